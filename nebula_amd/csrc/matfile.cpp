// matfile.cpp -- libnbl_mat.so: reading and writing the reference's HDF5
// material files (io/hdf5_file.cpp, common/units/*, and the file accesses of
// kieft_elastic/kieft_inelastic/full_penn::create and material::create).
//
// Host C++ over the HDF5 1.10 C and high-level (dimension scale, path) APIs.
// The unit system is restated here: a quantity is a double in the units
// eV, nm, ns, K, e with integer powers of those five dimensions.
#include <hdf5.h>
#include <hdf5_hl.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nbl_mat.h"

namespace {

thread_local std::string g_error;

struct mat_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template<typename F>
int guarded(F&& f)
{
    try {
        f();
        g_error.clear();
        return NBL_OK;
    } catch (const mat_error& e) {
        g_error = e.what();
        return NBL_ERR_INVALID;
    } catch (const std::bad_alloc&) {
        g_error = "allocation failed (bad_alloc)";
        return NBL_ERR_ALLOC;
    } catch (const std::exception& e) {
        g_error = e.what();
        return NBL_ERR_RUNTIME;
    }
}

void require(bool c, const std::string& msg)
{
    if (!c) throw mat_error(msg);
}

// ---- units (common/units/dimension.h, quantity.h, unit_parser.h, unit_system.h)
struct quantity {
    double v = 1;
    int d[5] = { 0, 0, 0, 0, 0 };   // energy, length, time, temperature, charge
    bool is(int e, int l, int t = 0, int k = 0, int c = 0) const
    {
        return d[0] == e && d[1] == l && d[2] == t && d[3] == k && d[4] == c;
    }
};

quantity q_pow(quantity a, int p)
{
    quantity r;
    r.v = std::pow(a.v, p);
    for (int i = 0; i < 5; ++i) r.d[i] = a.d[i] * p;
    return r;
}

quantity q_mul(quantity a, const quantity& b)
{
    a.v *= b.v;
    for (int i = 0; i < 5; ++i) a.d[i] += b.d[i];
    return a;
}

// default_unit_parser (unit_system.h:33-47)
const std::map<std::string, quantity>& unit_map()
{
    static const std::map<std::string, quantity> m = {
        { "eV", { 1, { 1, 0, 0, 0, 0 } } },
        { "nm", { 1, { 0, 1, 0, 0, 0 } } },
        { "K", { 1, { 0, 0, 0, 1, 0 } } },
        { "s", { 1e9, { 0, 0, 1, 0, 0 } } },
        { "m", { 1e9, { 0, 1, 0, 0, 0 } } },
        { "cm", { 1e7, { 0, 1, 0, 0, 0 } } },
        { "g", { 6.2415e15, { 1, -2, 2, 0, 0 } } },
        { "C", { 6.2415e18, { 0, 0, 0, 0, 1 } } },
        { "radian", { 1, { 0, 0, 0, 0, 0 } } },
    };
    return m;
}

// unit_parser::parse_unit (unit_parser.h:46-92): tokens "name[^power]"
// joined by '*' or '/', an optional leading '*' or '/'.
quantity parse_unit(const std::string& text)
{
    quantity out;
    size_t i = 0;
    const size_t n = text.size();
    auto new_unit = [](char c) { return c == '*' || c == '/'; };
    while (i < n) {
        int power = text[i] == '/' ? -1 : 1;
        if (new_unit(text[i])) ++i;
        std::string name;
        for (; i < n && text[i] != '^' && !new_unit(text[i]); ++i) name.push_back(text[i]);
        auto it = unit_map().find(name);
        require(it != unit_map().end(), "Unknown unit '" + name + "'.");
        if (i < n && text[i] == '^') {
            std::string p;
            for (++i; i < n && !new_unit(text[i]); ++i) p.push_back(text[i]);
            size_t used = 0;
            int pw = 0;
            try {
                pw = std::stoi(p, &used);
            } catch (const std::exception&) {
                throw mat_error("bad power '" + p + "' in unit '" + text + "'");
            }
            power *= pw;
        }
        out = q_mul(out, q_pow(it->second, power));
    }
    return out;
}

// ---- HDF5 handles -------------------------------------------------------
struct h5 {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    h5() = default;
    h5(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    h5(const h5&) = delete;
    h5& operator=(const h5&) = delete;
    h5(h5&& o) noexcept : id(o.id), close(o.close) { o.id = -1; }
    ~h5() { if (id >= 0 && close) close(id); }
    operator hid_t() const { return id; }
};

h5 vlen_string_type()
{
    h5 t(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(t, H5T_VARIABLE);
    return t;
}

// read_attribute_string (hdf5_file.cpp:55-68); fixed-length strings accepted too.
std::string read_string_attr(hid_t attr)
{
    h5 ft(H5Aget_type(attr), H5Tclose);
    require(ft >= 0 && H5Tget_class(ft) == H5T_STRING, "attribute is not a string");
    if (H5Tis_variable_str(ft) > 0) {
        h5 mt = vlen_string_type();
        char* buf = nullptr;
        require(H5Aread(attr, mt, &buf) >= 0 && buf, "cannot read string attribute");
        std::string s(buf);
        H5free_memory(buf);
        return s;
    }
    const size_t sz = H5Tget_size(ft);
    std::vector<char> buf(sz + 1, 0);
    h5 mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt, sz + 1);
    require(H5Aread(attr, mt, buf.data()) >= 0, "cannot read string attribute");
    return std::string(buf.data());
}

// read_attribute_quantity (hdf5_file.cpp:73-104): compound {value f64, unit string}
quantity read_quantity_attr(hid_t attr)
{
    struct rec { double value; char* unit; };
    h5 st = vlen_string_type();
    h5 ct(H5Tcreate(H5T_COMPOUND, sizeof(rec)), H5Tclose);
    H5Tinsert(ct, "value", HOFFSET(rec, value), H5T_NATIVE_DOUBLE);
    H5Tinsert(ct, "unit", HOFFSET(rec, unit), st);
    rec r{ 0, nullptr };
    require(H5Aread(attr, ct, &r) >= 0 && r.unit, "cannot read quantity attribute");
    std::string u(r.unit);
    H5free_memory(r.unit);
    quantity q = parse_unit(u);
    q.v = r.value * q.v;
    return q;
}

struct table {
    std::vector<hsize_t> dims;
    std::vector<float> data;
    quantity unit;
};

struct axis {
    int n = 0;
    double front = 0, back = 0;   // in the unit system
};

}  // namespace

struct nbl_matfile_s {
    std::string filename;
    hid_t file = -1;
    // owned arrays handed out through the raw structs
    std::vector<float> k_el_imfp, k_el_cos, k_in_imfp, k_in_w0, k_binding;
    std::vector<double> k_outer;
    std::vector<float> p_imfp, p_omega, p_q, p_binding;

    ~nbl_matfile_s() { if (file >= 0) H5Fclose(file); }

    h5 open_dataset(const std::string& name) const
    {
        h5 d(H5Dopen2(file, name.c_str(), H5P_DEFAULT), H5Dclose);
        require(d >= 0, "Could not read table '" + name + "' in " + filename);
        return d;
    }

    // fill_tableND (hdf5_file.cpp:234-289) + get_unit (:350-375)
    table read_table(const std::string& name, int rank, bool with_unit) const
    {
        h5 d = open_dataset(name);
        h5 sp(H5Dget_space(d), H5Sclose);
        require(H5Sget_simple_extent_ndims(sp) == rank, "Dataspace of '" + name + "' has unexpected dimension");
        table t;
        t.dims.resize(rank);
        H5Sget_simple_extent_dims(sp, t.dims.data(), nullptr);
        size_t total = 1;
        for (hsize_t v : t.dims) {
            require(v >= 1 && v < (hsize_t)std::numeric_limits<int>::max(), "table '" + name + "' too large");
            total *= (size_t)v;
        }
        t.data.resize(total);
        require(H5Dread(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, t.data.data()) >= 0,
                "cannot read table '" + name + "'");
        if (with_unit) {
            h5 a(H5Aopen(d, "units", H5P_DEFAULT), H5Aclose);
            require(a >= 0, "table '" + name + "' has no units attribute");
            t.unit = parse_unit(read_string_attr(a));
        }
        return t;
    }

    // get_dimscale (hdf5_file.cpp:111-153): the first scale attached to dim
    bool read_scale(const std::string& name, int dim, std::vector<double>& vals, quantity& unit) const
    {
        h5 d = open_dataset(name);
        struct ctx { std::vector<double>* v; quantity* u; std::string err; bool found; } c{ &vals, &unit, "", false };
        if (H5DSget_num_scales(d, (unsigned)dim) <= 0) return false;
        H5DSiterate_scales(d, (unsigned)dim, nullptr,
            [](hid_t, unsigned, hid_t dsid, void* data) -> herr_t {
                ctx& c = *static_cast<ctx*>(data);
                try {
                    h5 a(H5Aopen(dsid, "units", H5P_DEFAULT), H5Aclose);
                    require(a >= 0, "dimension scale has no units attribute");
                    *c.u = parse_unit(read_string_attr(a));
                    h5 sp(H5Dget_space(dsid), H5Sclose);
                    require(H5Sget_simple_extent_ndims(sp) == 1, "dimension scale is not 1D");
                    hsize_t n = 0;
                    H5Sget_simple_extent_dims(sp, &n, nullptr);
                    c.v->resize((size_t)n);
                    require(H5Dread(dsid, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, c.v->data()) >= 0,
                            "cannot read dimension scale");
                    c.found = true;
                } catch (const std::exception& e) {
                    c.err = e.what();
                }
                return 1;   // first scale only
            }, &c);
        require(c.err.empty(), name + ": " + c.err);
        return c.found;
    }

    // get_lin_dimscale / get_log_dimscale (hdf5_file.cpp:291-348): spacing
    // checked one-sidedly against EPSILON = 10 * FLT_EPSILON as the reference
    // does; a linear axis without a scale is 0..1.
    axis read_axis(const std::string& name, int dim, int n_expected, bool logarithmic, quantity* unit_out) const
    {
        std::vector<double> v;
        quantity unit;
        if (!read_scale(name, dim, v, unit)) {
            require(!logarithmic, "Dataset " + name + ": no dimension scale for a logarithmic axis");
            if (unit_out) *unit_out = quantity{};
            return { n_expected, 0.0, 1.0 };
        }
        require(v.size() == (size_t)n_expected, "Dimension scale has unexpected size.");
        const double eps = 10 * (double)std::numeric_limits<float>::epsilon();
        const double a = v.front(), b = v.back();
        for (int i = 0; i < n_expected; ++i) {
            double t;
            if (logarithmic) t = a * std::exp(i * (std::log(b / a) / (n_expected - 1)));
            else t = a + i * ((b - a) / (n_expected - 1));
            require(!(t - v[i] > eps), "Dataset " + name + (logarithmic ? ": Logarithmic" : ": Linear") +
                                           " dimension scale expected");
        }
        if (unit_out) *unit_out = unit;
        return { n_expected, a * unit.v, b * unit.v };
    }

    quantity property(const std::string& name) const
    {
        h5 a(H5Aopen(file, name.c_str(), H5P_DEFAULT), H5Aclose);
        require(a >= 0, "Could not read quantity property '" + name + "'");
        return read_quantity_attr(a);
    }
    bool has_property(const std::string& name) const { return H5Aexists(file, name.c_str()) > 0; }
};

namespace {

float energy_ev(const quantity& q, const char* what)
{
    require(q.is(1, 0), std::string(what) + " is not an energy");
    return (float)q.v;      // static_cast<real>(quantity / units::eV)
}

void same_axis(const axis& a, const axis& b, const std::string& what)
{
    require(a.n == b.n && a.front == b.front && a.back == b.back,
            "tables on different " + what + " axes are not supported (" + std::to_string(a.n) + " vs " +
                std::to_string(b.n) + " points)");
}

// table values times their unit, in float, as the reference's mem_scope loops do
void scale_values(std::vector<float>& v, const quantity& unit, double extra = 1.0)
{
    const float u = (float)(unit.v * extra);
    if (u == 1.0f) return;
    for (float& x : v) x = x * u;
}

// ---- writer ---------------------------------------------------------------
struct writer {
    h5 file, lcpl;
    explicit writer(const char* path)
        : file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose),
          lcpl(H5Pcreate(H5P_LINK_CREATE), H5Pclose)
    {
        require(file >= 0, std::string("cannot create ") + path);
        H5Pset_create_intermediate_group(lcpl, 1);
    }
    void string_attr(hid_t loc, const char* name, const std::string& value)
    {
        h5 t = vlen_string_type();
        h5 sp(H5Screate(H5S_SCALAR), H5Sclose);
        h5 a(H5Acreate2(loc, name, t, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
        const char* p = value.c_str();
        require(a >= 0 && H5Awrite(a, t, &p) >= 0, std::string("cannot write attribute ") + name);
    }
    void quantity_attr(const char* name, double value, const char* unit)
    {
        struct rec { double value; const char* unit; };
        h5 st = vlen_string_type();
        h5 ct(H5Tcreate(H5T_COMPOUND, sizeof(rec)), H5Tclose);
        H5Tinsert(ct, "value", HOFFSET(rec, value), H5T_NATIVE_DOUBLE);
        H5Tinsert(ct, "unit", HOFFSET(rec, unit), st);
        h5 sp(H5Screate(H5S_SCALAR), H5Sclose);
        h5 a(H5Acreate2(file, name, ct, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
        rec r{ value, unit };
        require(a >= 0 && H5Awrite(a, ct, &r) >= 0, std::string("cannot write attribute ") + name);
    }
    template<typename T>
    h5 dataset(const std::string& path, const std::vector<hsize_t>& dims, const T* data, const char* unit)
    {
        const bool dbl = sizeof(T) == 8;
        h5 sp(H5Screate_simple((int)dims.size(), dims.data(), nullptr), H5Sclose);
        h5 d(H5Dcreate2(file, path.c_str(), dbl ? H5T_IEEE_F64LE : H5T_IEEE_F32LE, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT),
             H5Dclose);
        require(d >= 0, "cannot create dataset " + path);
        require(H5Dwrite(d, dbl ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0,
                "cannot write dataset " + path);
        if (unit) string_attr(d, "units", unit);
        return d;
    }
    // a dimension scale dataset attached to `table`'s dimension `dim`
    void scale(hid_t table, const std::string& path, unsigned dim, double lo, double hi, int n, bool logarithmic,
               const char* unit, const char* label)
    {
        std::vector<double> v((size_t)n);
        for (int i = 0; i < n; ++i)
            v[(size_t)i] = logarithmic ? lo * std::exp(i * (std::log(hi / lo) / (n - 1))) : lo + i * ((hi - lo) / (n - 1));
        v.front() = lo;      // exact ends, so that a reader's geomspace/linspace
        v.back() = hi;       // over (front, back) reproduces the writer's axis
        h5 s = dataset(path, { (hsize_t)n }, v.data(), unit);
        require(H5DSset_scale(s, label) >= 0 && H5DSattach_scale(table, s, dim) >= 0, "cannot attach scale " + path);
    }
};

}  // namespace

extern "C" {

const char* nbl_mat_last_error(void) { return g_error.c_str(); }

int nbl_mat_parse_unit(const char* text, double* value, int32_t dims[5])
{
    return guarded([&] {
        require(text && value && dims, "nbl_mat_parse_unit: NULL argument");
        const quantity q = parse_unit(text);
        *value = q.v;
        for (int i = 0; i < 5; ++i) dims[i] = q.d[i];
    });
}

int nbl_mat_open(const char* path, nbl_matfile* out)
{
    return guarded([&] {
        require(path && out, "nbl_mat_open: NULL argument");
        *out = nullptr;
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        auto f = std::make_unique<nbl_matfile_s>();
        f->filename = path;
        f->file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
        require(f->file >= 0, std::string("Unable to open file ") + path);
        *out = f.release();
    });
}

int nbl_mat_close(nbl_matfile f)
{
    delete f;
    return NBL_OK;
}

int nbl_mat_exists(nbl_matfile f, const char* path, int32_t* out)
{
    return guarded([&] {
        require(f && path && out, "nbl_mat_exists: NULL argument");
        *out = H5LTpath_valid(f->file, path, 0) == 1 ? 1 : 0;
    });
}

int nbl_mat_get_string(nbl_matfile f, const char* name, char* buf, uint64_t cap, uint64_t* len)
{
    return guarded([&] {
        require(f && name && len, "nbl_mat_get_string: NULL argument");
        h5 a(H5Aopen(f->file, name, H5P_DEFAULT), H5Aclose);
        require(a >= 0, std::string("Could not read string property '") + name + "'");
        const std::string s = read_string_attr(a);
        *len = s.size();
        if (buf && cap) {
            const size_t k = std::min<size_t>(s.size(), (size_t)cap - 1);
            std::memcpy(buf, s.data(), k);
            buf[k] = 0;
        }
    });
}

int nbl_mat_get_quantity(nbl_matfile f, const char* name, double* value, int32_t dims[5])
{
    return guarded([&] {
        require(f && name && value && dims, "nbl_mat_get_quantity: NULL argument");
        const quantity q = f->property(name);
        *value = q.v;
        for (int i = 0; i < 5; ++i) dims[i] = q.d[i];
    });
}

int nbl_mat_kieft_raw(nbl_matfile f, nbl_kieft_raw* out)
{
    return guarded([&] {
        require(f && out, "nbl_mat_kieft_raw: NULL argument");
        require(H5LTpath_valid(f->file, "/kieft/elastic", 0) == 1,
                "Kieft elastic model not found in material file " + f->filename);
        require(H5LTpath_valid(f->file, "kieft/inelastic", 0) == 1,
                "Kieft inelastic model not found in material file " + f->filename);
        nbl_kieft_raw r{};
        // material::create, kieft_elastic::create, kieft_inelastic::create: properties
        r.barrier = energy_ev(f->property("barrier"), "barrier");
        r.fermi = energy_ev(f->property("fermi"), "fermi");
        r.band_gap = f->has_property("band_gap") ? energy_ev(f->property("band_gap"), "band_gap") : -1.0f;
        r.phonon_loss = energy_ev(f->property("phonon_loss"), "phonon_loss");
        const quantity a = f->property("effective_A");
        require(a.is(1, -2, 2), "effective_A is not a mass");
        r.effective_a_grams = a.v / 6.2415e15;
        // tables (elastic.h:128-157, inelastic.h:223-323)
        table el_imfp = f->read_table("/kieft/elastic/imfp", 1, true);
        table el_cos = f->read_table("/kieft/elastic/costheta_icdf", 2, false);
        table in_imfp = f->read_table("/kieft/inelastic/imfp", 1, true);
        table in_w0 = f->read_table("/kieft/inelastic/w0_icdf", 2, true);
        table bind = f->read_table("/kieft/ionization/binding_icdf", 2, true);
        require(el_imfp.unit.is(0, -1) && in_imfp.unit.is(0, -1), "imfp tables must have units of length^-1");
        require(in_w0.unit.is(1, 0) && bind.unit.is(1, 0), "w0_icdf and binding_icdf must have units of energy");
        const int nk = (int)el_imfp.dims[0], np = (int)el_cos.dims[1];
        quantity ku;
        const axis K = f->read_axis("/kieft/elastic/imfp", 0, nk, true, &ku);
        require(ku.is(1, 0), "energy axis is not an energy");
        same_axis(K, f->read_axis("/kieft/elastic/costheta_icdf", 0, (int)el_cos.dims[0], true, nullptr), "energy");
        same_axis(K, f->read_axis("/kieft/inelastic/imfp", 0, (int)in_imfp.dims[0], true, nullptr), "energy");
        same_axis(K, f->read_axis("/kieft/inelastic/w0_icdf", 0, (int)in_w0.dims[0], true, nullptr), "energy");
        same_axis(K, f->read_axis("/kieft/ionization/binding_icdf", 0, (int)bind.dims[0], true, nullptr), "energy");
        quantity pu;
        const axis P = f->read_axis("/kieft/elastic/costheta_icdf", 1, np, false, &pu);
        require(pu.is(0, 0), "probability axis is not dimensionless");
        same_axis(P, f->read_axis("/kieft/inelastic/w0_icdf", 1, (int)in_w0.dims[1], false, nullptr), "probability");
        same_axis(P, f->read_axis("/kieft/ionization/binding_icdf", 1, (int)bind.dims[1], false, nullptr), "probability");
        scale_values(el_imfp.data, el_imfp.unit);      // imfp * real(unit * nm)
        scale_values(in_imfp.data, in_imfp.unit);
        scale_values(in_w0.data, in_w0.unit);          // w0 * real(unit / eV)
        scale_values(bind.data, bind.unit);
        // outer shells (inelastic.h:277-289), in double
        table outer;
        {
            h5 d = f->open_dataset("/kieft/ionization/outer_shells");
            h5 sp(H5Dget_space(d), H5Sclose);
            require(H5Sget_simple_extent_ndims(sp) == 1, "outer_shells is not 1D");
            hsize_t n = 0;
            H5Sget_simple_extent_dims(sp, &n, nullptr);
            f->k_outer.assign((size_t)n, 0.0);
            if (n) require(H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, f->k_outer.data()) >= 0,
                           "cannot read outer_shells");
            h5 at(H5Aopen(d, "units", H5P_DEFAULT), H5Aclose);
            require(at >= 0, "outer_shells has no units attribute");
            const quantity u = parse_unit(read_string_attr(at));
            require(u.is(1, 0), "outer_shells must have units of energy");
            for (double& v : f->k_outer) v = v * u.v;
        }
        f->k_el_imfp = std::move(el_imfp.data); f->k_el_cos = std::move(el_cos.data);
        f->k_in_imfp = std::move(in_imfp.data); f->k_in_w0 = std::move(in_w0.data);
        f->k_binding = std::move(bind.data);
        r.n_k = nk; r.k_min = K.front; r.k_max = K.back;
        r.n_p = np; r.p_min = P.front; r.p_max = P.back;
        r.elastic_imfp = f->k_el_imfp.data(); r.elastic_costheta = f->k_el_cos.data();
        r.inelastic_imfp = f->k_in_imfp.data(); r.inelastic_w0 = f->k_in_w0.data();
        r.binding_icdf = f->k_binding.data();
        r.n_outer_shells = (int32_t)f->k_outer.size(); r.outer_shells = f->k_outer.data();
        *out = r;
    });
}

int nbl_mat_penn_raw(nbl_matfile f, nbl_penn_raw* out)
{
    return guarded([&] {
        require(f && out, "nbl_mat_penn_raw: NULL argument");
        require(H5LTpath_valid(f->file, "/full_penn", 0) == 1,
                "Full Penn model not found in material file " + f->filename);
        nbl_penn_raw r{};
        r.fermi = energy_ev(f->property("fermi"), "fermi");
        r.band_gap = f->has_property("band_gap") ? energy_ev(f->property("band_gap"), "band_gap") : -1.0f;
        table imfp = f->read_table("/full_penn/imfp", 1, true);
        table om = f->read_table("/full_penn/omega_icdf", 2, true);
        table q = f->read_table("/full_penn/q_icdf", 3, true);
        table b = f->read_table("/ionization/binding_icdf", 3, true);
        require(imfp.unit.is(0, -1) && q.unit.is(0, -1), "imfp and q_icdf must have units of length^-1");
        require(om.unit.is(1, 0) && b.unit.is(1, 0), "omega_icdf and binding_icdf must have units of energy");
        quantity u;
        const axis K = f->read_axis("/full_penn/imfp", 0, (int)imfp.dims[0], true, &u);
        require(u.is(1, 0), "energy axis is not an energy");
        same_axis(K, f->read_axis("/full_penn/omega_icdf", 0, (int)om.dims[0], true, nullptr), "energy");
        const axis P = f->read_axis("/full_penn/omega_icdf", 1, (int)om.dims[1], false, nullptr);
        const axis qK = f->read_axis("/full_penn/q_icdf", 0, (int)q.dims[0], true, nullptr);
        const axis qW = f->read_axis("/full_penn/q_icdf", 1, (int)q.dims[1], false, nullptr);
        const axis qP = f->read_axis("/full_penn/q_icdf", 2, (int)q.dims[2], false, nullptr);
        const axis iK = f->read_axis("/ionization/binding_icdf", 0, (int)b.dims[0], true, nullptr);
        const axis iW = f->read_axis("/ionization/binding_icdf", 1, (int)b.dims[1], true, &u);
        require(u.is(0, 0), "the binding table's omega/K axis is not dimensionless");
        const axis iP = f->read_axis("/ionization/binding_icdf", 2, (int)b.dims[2], false, nullptr);
        scale_values(imfp.data, imfp.unit);            // imfp * real(unit * nm)
        scale_values(om.data, om.unit);                // omega * real(unit / eV)
        scale_values(q.data, q.unit);                  // q in nm^-1 (the 0.19519 factor is the prep's)
        scale_values(b.data, b.unit);
        f->p_imfp = std::move(imfp.data); f->p_omega = std::move(om.data);
        f->p_q = std::move(q.data); f->p_binding = std::move(b.data);
        r.n_k = K.n; r.k_min = K.front; r.k_max = K.back;
        r.n_p = P.n; r.p_min = P.front; r.p_max = P.back;
        r.imfp = f->p_imfp.data(); r.omega_icdf = f->p_omega.data();
        r.q_nk = qK.n; r.q_nw = qW.n; r.q_np = qP.n;
        r.q_k_min = qK.front; r.q_k_max = qK.back; r.q_w_min = qW.front; r.q_w_max = qW.back;
        r.q_p_min = qP.front; r.q_p_max = qP.back; r.q_icdf = f->p_q.data();
        r.i_nk = iK.n; r.i_nw = iW.n; r.i_np = iP.n;
        r.i_k_min = iK.front; r.i_k_max = iK.back; r.i_w_min = iW.front; r.i_w_max = iW.back;
        r.i_p_min = iP.front; r.i_p_max = iP.back; r.binding_icdf = f->p_binding.data();
        *out = r;
    });
}

int nbl_mat_write(const char* path, const char* name, const nbl_kieft_raw* k, const nbl_penn_raw* p)
{
    return guarded([&] {
        require(path && name && k, "nbl_mat_write: NULL argument");
        require(k->n_k >= 2 && k->n_p >= 2 && k->k_min > 0 && k->k_max > k->k_min && k->p_max > k->p_min,
                "nbl_mat_write: bad Kieft axes");
        require(k->elastic_imfp && k->elastic_costheta && k->inelastic_imfp && k->inelastic_w0 && k->binding_icdf &&
                (k->n_outer_shells == 0 || k->outer_shells), "nbl_mat_write: NULL Kieft table");
        if (p) {
            require(p->fermi == k->fermi && p->band_gap == k->band_gap, "nbl_mat_write: Penn and Kieft properties differ");
            require(p->imfp && p->omega_icdf && p->q_icdf && p->binding_icdf, "nbl_mat_write: NULL Penn table");
            require(p->n_k >= 2 && p->n_p >= 2 && p->q_nk >= 2 && p->q_nw >= 2 && p->q_np >= 2 && p->i_nk >= 2 &&
                    p->i_nw >= 2 && p->i_np >= 2 && p->k_min > 0 && p->q_k_min > 0 && p->i_k_min > 0 && p->i_w_min > 0,
                    "nbl_mat_write: bad Penn axes");
        }
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        writer w(path);
        w.string_attr(w.file, "name", name);
        w.string_attr(w.file, "cstool_version", "nebula_amd nbl_mat_write 1");
        w.quantity_attr("barrier", k->barrier, "eV");
        w.quantity_attr("fermi", k->fermi, "eV");
        w.quantity_attr("band_gap", k->band_gap, "eV");
        w.quantity_attr("phonon_loss", k->phonon_loss, "eV");
        w.quantity_attr("effective_A", k->effective_a_grams, "g");
        const hsize_t nk = (hsize_t)k->n_k, np = (hsize_t)k->n_p;
        struct t2 { const char* path; const float* data; const char* unit; bool two_d; };
        const t2 kt[] = {
            { "/kieft/elastic/imfp", k->elastic_imfp, "nm^-1", false },
            { "/kieft/elastic/costheta_icdf", k->elastic_costheta, "", true },
            { "/kieft/inelastic/imfp", k->inelastic_imfp, "nm^-1", false },
            { "/kieft/inelastic/w0_icdf", k->inelastic_w0, "eV", true },
            { "/kieft/ionization/binding_icdf", k->binding_icdf, "eV", true },
        };
        for (const t2& t : kt) {
            h5 d = t.two_d ? w.dataset(t.path, { nk, np }, t.data, t.unit) : w.dataset(t.path, { nk }, t.data, t.unit);
            w.scale(d, std::string(t.path) + "_energy", 0, k->k_min, k->k_max, k->n_k, true, "eV", "energy");
            if (t.two_d)
                w.scale(d, std::string(t.path) + "_probability", 1, k->p_min, k->p_max, k->n_p, false, "", "probability");
        }
        static const double no_shell = 0.0;
        w.dataset("/kieft/ionization/outer_shells", { (hsize_t)k->n_outer_shells },
                  k->n_outer_shells ? k->outer_shells : &no_shell, "eV");
        if (p) {
            const hsize_t pk = (hsize_t)p->n_k, pp = (hsize_t)p->n_p;
            h5 a = w.dataset("/full_penn/imfp", { pk }, p->imfp, "nm^-1");
            w.scale(a, "/full_penn/imfp_energy", 0, p->k_min, p->k_max, p->n_k, true, "eV", "energy");
            h5 o = w.dataset("/full_penn/omega_icdf", { pk, pp }, p->omega_icdf, "eV");
            w.scale(o, "/full_penn/omega_icdf_energy", 0, p->k_min, p->k_max, p->n_k, true, "eV", "energy");
            w.scale(o, "/full_penn/omega_icdf_probability", 1, p->p_min, p->p_max, p->n_p, false, "", "probability");
            h5 q = w.dataset("/full_penn/q_icdf", { (hsize_t)p->q_nk, (hsize_t)p->q_nw, (hsize_t)p->q_np }, p->q_icdf,
                             "nm^-1");
            w.scale(q, "/full_penn/q_icdf_energy", 0, p->q_k_min, p->q_k_max, p->q_nk, true, "eV", "energy");
            w.scale(q, "/full_penn/q_icdf_omega_over_K", 1, p->q_w_min, p->q_w_max, p->q_nw, false, "", "omega/K");
            w.scale(q, "/full_penn/q_icdf_probability", 2, p->q_p_min, p->q_p_max, p->q_np, false, "", "probability");
            h5 b = w.dataset("/ionization/binding_icdf", { (hsize_t)p->i_nk, (hsize_t)p->i_nw, (hsize_t)p->i_np },
                             p->binding_icdf, "eV");
            w.scale(b, "/ionization/binding_icdf_energy", 0, p->i_k_min, p->i_k_max, p->i_nk, true, "eV", "energy");
            w.scale(b, "/ionization/binding_icdf_omega_over_K", 1, p->i_w_min, p->i_w_max, p->i_nw, true, "", "omega/K");
            w.scale(b, "/ionization/binding_icdf_probability", 2, p->i_p_min, p->i_p_max, p->i_np, false, "",
                    "probability");
        }
    });
}

}  // extern "C"
