// materials.hip -- nbl_materials_*, nbl_kieft_prepare, batched table lookups
// and the single-event physics entry point (injected uniforms).
#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>

#include "nbl_host.h"
#include "nbl_device.h"

using namespace nbl;

struct nbl_materials_s {
    std::vector<dmat> views;     // device pointers inside
    dbuf<float> data;            // every table of every material
    dbuf<dmat> dmats;
    float max_energy = 0;
};

namespace nbl {
const dmat* materials_device(nbl_materials m) { return m->dmats.p; }
const dmat& materials_view(nbl_materials m, int i) { return m->views[i]; }
int materials_count(nbl_materials m) { return (int)m->views.size(); }
float materials_max_energy(nbl_materials m) { return m->max_energy; }
const tri_rec* geometry_host_recs(nbl_geometry g);
const dgeom& geometry_view(nbl_geometry g);
}

namespace {

void check_t1(const nbl_table1d_desc& t, const char* what)
{
    if (t.n < 2 || !t.data || !std::isfinite(t.x_min) || !std::isfinite(t.x_max) || t.x_max == t.x_min)
        throw invalid_arg(std::string("nbl_materials_create: invalid 1D table ") + what);
}
void check_t2(const nbl_table2d_desc& t, const char* what)
{
    if (t.width < 2 || t.height < 2 || !t.data || !std::isfinite(t.x_min) || !std::isfinite(t.x_max) ||
        !std::isfinite(t.y_min) || !std::isfinite(t.y_max) || t.x_max == t.x_min || t.y_max == t.y_min)
        throw invalid_arg(std::string("nbl_materials_create: invalid 2D table ") + what);
}
// table_1D::create / table_2D::create axis steps (table_1D.inl:14-15, table_2D.inl:16-19)
dtab1 make_t1(const nbl_table1d_desc& t, const float* dev)
{
    return { dev, t.n, t.x_min, (float)(t.n - 1) / (t.x_max - t.x_min) };
}
dtab2 make_t2(const nbl_table2d_desc& t, const float* dev)
{
    return { dev, t.width, t.height, t.x_min, (float)(t.width - 1) / (t.x_max - t.x_min),
             t.y_min, (float)(t.height - 1) / (t.y_max - t.y_min) };
}
void check_t3(const nbl_table3d_desc& t, const char* what)
{
    if (t.width < 2 || t.height < 2 || t.depth < 2 || !t.data || !std::isfinite(t.x_min) || !std::isfinite(t.x_max) ||
        !std::isfinite(t.y_min) || !std::isfinite(t.y_max) || !std::isfinite(t.z_min) || !std::isfinite(t.z_max) ||
        t.x_max == t.x_min || t.y_max == t.y_min || t.z_max == t.z_min)
        throw invalid_arg(std::string("nbl_materials_create: invalid 3D table ") + what);
}
// table_3D::create axis steps (table_3D.inl:16-22)
dtab3 make_t3(const nbl_table3d_desc& t, const float* dev)
{
    return { dev, t.width, t.height, t.depth, t.x_min, (float)(t.width - 1) / (t.x_max - t.x_min),
             t.y_min, (float)(t.height - 1) / (t.y_max - t.y_min), t.z_min, (float)(t.depth - 1) / (t.z_max - t.z_min) };
}
size_t t3_size(const nbl_table3d_desc& t) { return (size_t)t.width * t.height * t.depth; }

__global__ __launch_bounds__(256) void k_tab1(dtab1 t, const float* __restrict__ x, uint32_t n, float* __restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = tab1_get(t, x[i]);
}
__global__ __launch_bounds__(256) void k_tab3(dtab3 t, int rounddown, const float* __restrict__ x,
    const float* __restrict__ y, const float* __restrict__ z, uint32_t n, float* __restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = rounddown ? tab3_rounddown(t, x[i], y[i], z[i]) : tab3_get(t, x[i], y[i], z[i]);
}
__global__ __launch_bounds__(256) void k_tab2(dtab2 t, int rounddown, const float* __restrict__ x,
    const float* __restrict__ y, uint32_t n, float* __restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = rounddown ? tab2_rounddown(t, x[i], y[i]) : tab2_get(t, x[i], y[i]);
}

struct phys_args {
    const dmat* mats;
    const tri_rec* tris;      // triangle records indexed by input id (intersect events)
    int event;
    const int32_t* material; const uint32_t* last_tri; const float* uniforms;
    uint32_t n;
    dparticle* io; int32_t* out_material; uint8_t* out_status;
    dparticle* out_sec; uint8_t* out_has_sec; float* out_scalar; uint8_t* out_type;
};

__global__ __launch_bounds__(256) void k_physics(phys_args a)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        rng_inject rng{ a.uniforms + (size_t)i * NBL_NUM_UNIFORMS, 0 };
        dparticle p = a.io[i];
        int mat = a.material[i];
        uint8_t st = OUT_ALIVE, has = 0, ty = 0;
        float scalar = 0;
        dparticle sec{};
        if (a.event == NBL_EVENT_SAMPLE_PATH) {
            scalar = sample_path(a.mats[mat], p.e, rng, &ty);
        } else if (a.event == NBL_EVENT_ELASTIC) {
            elastic_execute(a.mats[mat], p, rng);
        } else if (a.event == NBL_EVENT_INELASTIC) {
            has = inelastic_event(a.mats[mat], p, rng, &sec) ? 1 : 0;
        } else {
            st = (uint8_t)intersect_execute(a.mats, a.tris[a.last_tri[i]], p, mat, rng);
        }
        a.io[i] = p;
        a.out_material[i] = mat;
        a.out_status[i] = st;
        a.out_sec[i] = sec;
        a.out_has_sec[i] = has;
        a.out_scalar[i] = scalar;
        a.out_type[i] = ty;
    }
}

}  // namespace

extern "C" {

// kieft_elastic::create (elastic.h:120-166) and kieft_inelastic::create
// (inelastic.h:212-327) applied to tables on log-spaced energy and linear
// probability grids (the dimension scales of common/util/range.h:204-312).
int nbl_kieft_prepare(const nbl_kieft_raw* r, float* el_imfp, float* el_icdf, float* in_imfp,
    float* in_w0, float* ion, nbl_material_desc* d)
{
    return guarded([&] {
        require(r && el_imfp && el_icdf && in_imfp && in_w0 && ion && d, "nbl_kieft_prepare: NULL argument");
        require(r->n_k >= 2 && r->n_p >= 2, "nbl_kieft_prepare: need at least 2 points per axis");
        require(r->k_min > 0 && r->k_max > r->k_min && r->p_max > r->p_min, "nbl_kieft_prepare: bad axis range");
        require(r->elastic_imfp && r->elastic_costheta && r->inelastic_imfp && r->inelastic_w0 && r->binding_icdf,
            "nbl_kieft_prepare: NULL table");
        require(r->n_outer_shells == 0 || r->outer_shells, "nbl_kieft_prepare: NULL outer shells");
        const int nk = r->n_k, np = r->n_p;
        const size_t n2 = (size_t)nk * np;
        // geomspace: K_i = K_min * exp(i * log(K_max/K_min)/(n-1)); linspace for P
        std::vector<double> K(nk), P(np);
        const double ks = std::log(r->k_max / r->k_min) / (nk - 1), ps = (r->p_max - r->p_min) / (np - 1);
        for (int i = 0; i < nk; ++i) K[i] = r->k_min * std::exp(i * ks);
        for (int j = 0; j < np; ++j) P[j] = r->p_min + j * ps;
        const float lk_lo = (float)std::log(K[0]), lk_hi = (float)std::log(K[nk - 1]);
        const float p_lo = (float)P[0], p_hi = (float)P[np - 1];

        for (int i = 0; i < nk; ++i) el_imfp[i] = std::log(r->elastic_imfp[i] * 1.0f);
        for (size_t k = 0; k < n2; ++k) el_icdf[k] = std::max(-1.0f, std::min(1.0f, r->elastic_costheta[k]));
        for (int i = 0; i < nk; ++i) in_imfp[i] = std::log(r->inelastic_imfp[i] * 1.0f);
        const double fermi = (double)r->fermi;
        for (int i = 0; i < nk; ++i)
            for (int j = 0; j < np; ++j) {
                const size_t k = (size_t)i * np + j;
                in_w0[k] = (float)std::log(std::max(0.0, std::min<double>(K[i] - fermi, (double)(r->inelastic_w0[k] * 1.0f))));
            }

        // Binding table: outer shells below 100 eV, sorted; entries for K > 100 eV
        // take the rounddown value at log(K + 10 eV) of the table as rewritten
        // so far (the reference rewrites in place), -1 below 50 eV or if not
        // finite; otherwise the largest outer shell <= K.
        std::vector<double> shells;
        for (int s = 0; s < r->n_outer_shells; ++s)
            if (r->outer_shells[s] < 100.0) shells.push_back(r->outer_shells[s]);
        std::sort(shells.begin(), shells.end());
        std::copy(r->binding_icdf, r->binding_icdf + n2, ion);
        const float x_step = (float)(nk - 1) / (lk_hi - lk_lo), y_step = (float)(np - 1) / (p_hi - p_lo);
        auto rounddown = [&](float x, float y) {
            const float xi = (x - lk_lo) * x_step, yi = (y - p_lo) * y_step;
            const int lx = clampi((int)std::floor(xi), 0, nk - 1), ly = clampi((int)std::floor(yi), 0, np - 1);
            return ion[(size_t)lx * np + ly];
        };
        for (int i = 0; i < nk; ++i)
            for (int j = 0; j < np; ++j) {
                double b = -1;
                if (K[i] > 100.0) {
                    b = (double)rounddown((float)std::log(K[i] + 10.0), (float)P[j]);
                    if (b < 50.0 || !std::isfinite(b)) b = -1;
                }
                if (b < 0) {
                    // largest shell <= K: last element not greater than K
                    auto it = std::upper_bound(shells.begin(), shells.end(), K[i]);
                    if (it != shells.begin()) b = *(it - 1);
                }
                ion[(size_t)i * np + j] = (float)b;
            }
        d->barrier = r->barrier; d->fermi = r->fermi; d->band_gap = r->band_gap; d->phonon_loss = r->phonon_loss;
        // 2 * (2 m_e) / A with the reference's gram (6.2415e15 in eV ns^2 nm^-2)
        d->recoil_const = 2 * (float)((2 * (9.109383e-28 * 6.2415e15)) / (r->effective_a_grams * 6.2415e15));
        d->elastic_log_imfp = { nk, lk_lo, lk_hi, el_imfp };
        d->elastic_costheta_icdf = { nk, np, lk_lo, lk_hi, p_lo, p_hi, el_icdf };
        d->inelastic_log_imfp = { nk, lk_lo, lk_hi, in_imfp };
        d->inelastic_log_w0_icdf = { nk, np, lk_lo, lk_hi, p_lo, p_hi, in_w0 };
        d->ionisation = { nk, np, lk_lo, lk_hi, p_lo, p_hi, ion };
        d->inelastic_model = NBL_INELASTIC_KIEFT;
        d->penn_q_icdf = nbl_table3d_desc{};
        d->penn_ionisation = nbl_table3d_desc{};
    });
}

int nbl_penn_prepare(const nbl_penn_raw* r, float* log_imfp, float* log_omega_icdf, float* q_icdf, float* ionisation,
    nbl_material_desc* d)
{
    return guarded([&] {
        require(r && log_imfp && log_omega_icdf && q_icdf && ionisation && d, "nbl_penn_prepare: NULL argument");
        require(r->imfp && r->omega_icdf && r->q_icdf && r->binding_icdf, "nbl_penn_prepare: NULL table");
        require(r->n_k >= 2 && r->n_p >= 2 && r->q_nk >= 2 && r->q_nw >= 2 && r->q_np >= 2 && r->i_nk >= 2 &&
                r->i_nw >= 2 && r->i_np >= 2, "nbl_penn_prepare: need at least 2 points per axis");
        require(r->k_min > 0 && r->k_max > r->k_min && r->p_max > r->p_min && r->q_k_min > 0 &&
                r->q_k_max > r->q_k_min && r->q_w_max > r->q_w_min && r->q_p_max > r->q_p_min && r->i_k_min > 0 &&
                r->i_k_max > r->i_k_min && r->i_w_min > 0 && r->i_w_max > r->i_w_min && r->i_p_max > r->i_p_min,
            "nbl_penn_prepare: bad axis range");
        // full_penn.h:304-318: log of the IMFP in nm^-1
        const int nk = r->n_k, np = r->n_p;
        std::vector<double> K(nk);
        const double ks = std::log(r->k_max / r->k_min) / (nk - 1);
        for (int i = 0; i < nk; ++i) K[i] = r->k_min * std::exp(i * ks);
        const float lk_lo = (float)std::log(K[0]), lk_hi = (float)std::log(K[nk - 1]);
        const float p_lo = (float)r->p_min, p_hi = (float)(r->p_min + (np - 1) * ((r->p_max - r->p_min) / (np - 1)));
        for (int i = 0; i < nk; ++i) log_imfp[i] = std::log(r->imfp[i] * 1.0f);
        // full_penn.h:320-341: log(omega), omega clamped to [0, K - E_F]
        const double fermi = (double)r->fermi;
        for (int i = 0; i < nk; ++i)
            for (int j = 0; j < np; ++j) {
                const size_t k = (size_t)i * np + j;
                log_omega_icdf[k] = (float)std::log(std::max(0.0, std::min<double>(K[i] - fermi, (double)(r->omega_icdf[k] * 1.0f))));
            }
        // full_penn.h:344-366: q in units of hbar / sqrt(2 m) (0.19519 nm eV^1/2)
        const size_t nq = (size_t)r->q_nk * r->q_nw * r->q_np;
        const float q_unit = (float)(0.19519 * 1.0);
        for (size_t k = 0; k < nq; ++k) q_icdf[k] = r->q_icdf[k] * q_unit;
        // full_penn.h:369-396: binding energies below 50 eV (or not finite) -> -1
        const size_t ni = (size_t)r->i_nk * r->i_nw * r->i_np;
        for (size_t k = 0; k < ni; ++k) {
            float b = r->binding_icdf[k] * 1.0f;
            if (b < 50 || !std::isfinite(b)) b = -1;
            ionisation[k] = b;
        }
        auto lin_hi = [](double lo, double hi, int n) { return (float)(lo + (n - 1) * ((hi - lo) / (n - 1))); };
        auto log_end = [](double lo, double hi, int n, int i) { return lo * std::exp(i * (std::log(hi / lo) / (n - 1))); };
        d->fermi = r->fermi;
        d->band_gap = r->band_gap;
        d->inelastic_log_imfp = { nk, lk_lo, lk_hi, log_imfp };
        d->inelastic_log_w0_icdf = { nk, np, lk_lo, lk_hi, p_lo, p_hi, log_omega_icdf };
        d->ionisation = nbl_table2d_desc{};
        d->penn_q_icdf = { r->q_nk, r->q_nw, r->q_np,
            (float)std::log(log_end(r->q_k_min, r->q_k_max, r->q_nk, 0)),
            (float)std::log(log_end(r->q_k_min, r->q_k_max, r->q_nk, r->q_nk - 1)),
            (float)r->q_w_min, lin_hi(r->q_w_min, r->q_w_max, r->q_nw),
            (float)r->q_p_min, lin_hi(r->q_p_min, r->q_p_max, r->q_np), q_icdf };
        d->penn_ionisation = { r->i_nk, r->i_nw, r->i_np,
            (float)std::log(log_end(r->i_k_min, r->i_k_max, r->i_nk, 0)),
            (float)std::log(log_end(r->i_k_min, r->i_k_max, r->i_nk, r->i_nk - 1)),
            (float)std::log(log_end(r->i_w_min, r->i_w_max, r->i_nw, 0)),
            (float)std::log(log_end(r->i_w_min, r->i_w_max, r->i_nw, r->i_nw - 1)),
            (float)r->i_p_min, lin_hi(r->i_p_min, r->i_p_max, r->i_np), ionisation };
        d->inelastic_model = NBL_INELASTIC_PENN;
    });
}

int nbl_materials_create(const nbl_material_desc* mats, int32_t n, nbl_materials* out)
{
    return guarded([&] {
        require(out != nullptr, "nbl_materials_create: out is NULL");
        require(mats != nullptr && n > 0, "nbl_materials_create: no materials");
        require(n <= 32767, "nbl_materials_create: too many materials");
        *out = nullptr;
        ensure_device();
        auto m = std::make_unique<nbl_materials_s>();
        size_t total = 0;
        for (int i = 0; i < n; ++i) {
            const nbl_material_desc& d = mats[i];
            check_t1(d.elastic_log_imfp, "elastic_log_imfp");
            check_t1(d.inelastic_log_imfp, "inelastic_log_imfp");
            check_t2(d.elastic_costheta_icdf, "elastic_costheta_icdf");
            check_t2(d.inelastic_log_w0_icdf, "inelastic_log_w0_icdf");
            require(d.inelastic_model == NBL_INELASTIC_KIEFT || d.inelastic_model == NBL_INELASTIC_PENN,
                "nbl_materials_create: unknown inelastic model");
            require(d.inelastic_model == mats[0].inelastic_model,
                "nbl_materials_create: all materials of a set must use the same inelastic model");
            total += d.elastic_log_imfp.n + d.inelastic_log_imfp.n +
                     (size_t)d.elastic_costheta_icdf.width * d.elastic_costheta_icdf.height +
                     (size_t)d.inelastic_log_w0_icdf.width * d.inelastic_log_w0_icdf.height;
            if (d.inelastic_model == NBL_INELASTIC_KIEFT) {
                check_t2(d.ionisation, "ionisation");
                total += (size_t)d.ionisation.width * d.ionisation.height;
            } else {
                check_t3(d.penn_q_icdf, "penn_q_icdf");
                check_t3(d.penn_ionisation, "penn_ionisation");
                total += t3_size(d.penn_q_icdf) + t3_size(d.penn_ionisation);
            }
        }
        // one allocation; every table starts on a 256-byte boundary
        std::vector<float> host;
        host.reserve(total + 64 * 7 * n);
        auto put = [&](const float* src, size_t cnt) {
            while (host.size() % 64) host.push_back(0.f);
            const size_t off = host.size();
            host.insert(host.end(), src, src + cnt);
            return off;
        };
        struct offs { size_t a, b, c, d, e, f, g; };
        std::vector<offs> o(n);
        for (int i = 0; i < n; ++i) {
            const nbl_material_desc& d = mats[i];
            o[i].a = put(d.elastic_log_imfp.data, d.elastic_log_imfp.n);
            o[i].b = put(d.inelastic_log_imfp.data, d.inelastic_log_imfp.n);
            o[i].c = put(d.elastic_costheta_icdf.data, (size_t)d.elastic_costheta_icdf.width * d.elastic_costheta_icdf.height);
            o[i].d = put(d.inelastic_log_w0_icdf.data, (size_t)d.inelastic_log_w0_icdf.width * d.inelastic_log_w0_icdf.height);
            o[i].e = o[i].f = o[i].g = 0;
            if (d.inelastic_model == NBL_INELASTIC_KIEFT) {
                o[i].e = put(d.ionisation.data, (size_t)d.ionisation.width * d.ionisation.height);
            } else {
                o[i].f = put(d.penn_q_icdf.data, t3_size(d.penn_q_icdf));
                o[i].g = put(d.penn_ionisation.data, t3_size(d.penn_ionisation));
            }
        }
        m->data.alloc(host.size());
        m->data.upload(host.data(), host.size());
        float emax = __builtin_inff();
        for (int i = 0; i < n; ++i) {
            const nbl_material_desc& d = mats[i];
            dmat v;
            v.barrier = d.barrier; v.fermi = d.fermi; v.band_gap = d.band_gap;
            v.phonon_loss = d.phonon_loss; v.recoil_const = d.recoil_const;
            v.el_imfp = make_t1(d.elastic_log_imfp, m->data.p + o[i].a);
            v.in_imfp = make_t1(d.inelastic_log_imfp, m->data.p + o[i].b);
            v.el_icdf = make_t2(d.elastic_costheta_icdf, m->data.p + o[i].c);
            v.in_icdf = make_t2(d.inelastic_log_w0_icdf, m->data.p + o[i].d);
            v.model = d.inelastic_model;
            v.ion = dtab2{};
            v.pq = dtab3{};
            v.pion = dtab3{};
            if (v.model == INEL_KIEFT) {
                v.ion = make_t2(d.ionisation, m->data.p + o[i].e);
            } else {
                v.pq = make_t3(d.penn_q_icdf, m->data.p + o[i].f);
                v.pion = make_t3(d.penn_ionisation, m->data.p + o[i].g);
            }
            m->views.push_back(v);
            // get_max_energy: exp(scalemax) of both IMFP tables, minimum over
            // mechanisms and materials (elastic.h:46-49, cpu_material_manager.inl:42-55)
            for (const dtab1* t : { &v.el_imfp, &v.in_imfp }) {
                const float e = std::exp(t->x_min + (float)(t->n - 1) / t->x_step);
                if (e < emax) emax = e;
            }
        }
        m->max_energy = emax;
        m->dmats.alloc(n);
        m->dmats.upload(m->views.data(), n);
        NBL_HIP(hipDeviceSynchronize());
        *out = m.release();
    });
}

int nbl_materials_destroy(nbl_materials m)
{
    return guarded([&] { delete m; });
}

int nbl_materials_get_max_energy(nbl_materials m, float* out)
{
    return guarded([&] {
        require(m && out, "nbl_materials_get_max_energy: NULL argument");
        *out = m->max_energy;
    });
}

int nbl_table1d_get_batch(nbl_materials m, int32_t material, int32_t table, const float* x, uint32_t n, float* out)
{
    return guarded([&] {
        require(m != nullptr, "nbl_table1d_get_batch: NULL materials");
        require(material >= 0 && material < (int)m->views.size(), "nbl_table1d_get_batch: bad material");
        require(table == NBL_TAB_ELASTIC_IMFP || table == NBL_TAB_INELASTIC_IMFP, "nbl_table1d_get_batch: bad table");
        require(n == 0 || (x && out), "nbl_table1d_get_batch: NULL array");
        if (n == 0) return;
        const dmat& v = m->views[material];
        dbuf<float> dx(n), dout(n);
        dx.upload(x, n);
        hipLaunchKernelGGL(k_tab1, dim3(grid_for(n, 256)), dim3(256), 0, 0,
            table == NBL_TAB_ELASTIC_IMFP ? v.el_imfp : v.in_imfp, dx.p, n, dout.p);
        NBL_HIP(hipGetLastError());
        dout.download(out, n);
        NBL_HIP(hipDeviceSynchronize());
    });
}

int nbl_table2d_get_batch(nbl_materials m, int32_t material, int32_t table, int32_t rounddown,
    const float* x, const float* y, uint32_t n, float* out)
{
    return guarded([&] {
        require(m != nullptr, "nbl_table2d_get_batch: NULL materials");
        require(material >= 0 && material < (int)m->views.size(), "nbl_table2d_get_batch: bad material");
        require(table == NBL_TAB_ELASTIC_ICDF || table == NBL_TAB_INELASTIC_W0_ICDF || table == NBL_TAB_IONISATION,
            "nbl_table2d_get_batch: bad table");
        require(n == 0 || (x && y && out), "nbl_table2d_get_batch: NULL array");
        if (n == 0) return;
        const dmat& v = m->views[material];
        require(table != NBL_TAB_IONISATION || v.model == INEL_KIEFT,
            "nbl_table2d_get_batch: a full Penn material has no 2D ionisation table");
        const dtab2& t = table == NBL_TAB_ELASTIC_ICDF ? v.el_icdf : table == NBL_TAB_INELASTIC_W0_ICDF ? v.in_icdf : v.ion;
        dbuf<float> dx(n), dy(n), dout(n);
        dx.upload(x, n); dy.upload(y, n);
        hipLaunchKernelGGL(k_tab2, dim3(grid_for(n, 256)), dim3(256), 0, 0, t, rounddown ? 1 : 0, dx.p, dy.p, n, dout.p);
        NBL_HIP(hipGetLastError());
        dout.download(out, n);
        NBL_HIP(hipDeviceSynchronize());
    });
}

int nbl_table3d_get_batch(nbl_materials m, int32_t material, int32_t table, int32_t rounddown,
    const float* x, const float* y, const float* z, uint32_t n, float* out)
{
    return guarded([&] {
        require(m != nullptr, "nbl_table3d_get_batch: NULL materials");
        require(material >= 0 && material < (int)m->views.size(), "nbl_table3d_get_batch: bad material");
        require(table == NBL_TAB_PENN_Q_ICDF || table == NBL_TAB_PENN_IONISATION, "nbl_table3d_get_batch: bad table");
        const dmat& v = m->views[material];
        require(v.model == INEL_PENN, "nbl_table3d_get_batch: not a full Penn material");
        require(n == 0 || (x && y && z && out), "nbl_table3d_get_batch: NULL array");
        if (n == 0) return;
        const dtab3& t = table == NBL_TAB_PENN_Q_ICDF ? v.pq : v.pion;
        dbuf<float> dx(n), dy(n), dz(n), dout(n);
        dx.upload(x, n); dy.upload(y, n); dz.upload(z, n);
        hipLaunchKernelGGL(k_tab3, dim3(grid_for(n, 256)), dim3(256), 0, 0, t, rounddown ? 1 : 0, dx.p, dy.p, dz.p, n, dout.p);
        NBL_HIP(hipGetLastError());
        dout.download(out, n);
        NBL_HIP(hipDeviceSynchronize());
    });
}

int nbl_physics_batch(nbl_materials m, nbl_geometry g, int32_t event, const int32_t* material,
    const uint32_t* last_tri, const float* uniforms, uint32_t n, nbl_particle* io, int32_t* out_material,
    uint8_t* out_status, nbl_particle* out_secondary, uint8_t* out_has_secondary, float* out_scalar,
    uint8_t* out_type)
{
    return guarded([&] {
        require(m != nullptr, "nbl_physics_batch: NULL materials");
        require(event >= NBL_EVENT_SAMPLE_PATH && event <= NBL_EVENT_INTERSECT, "nbl_physics_batch: bad event");
        require(event != NBL_EVENT_INTERSECT || (g && last_tri), "nbl_physics_batch: intersect needs geometry and last_tri");
        require(n == 0 || (material && uniforms && io), "nbl_physics_batch: NULL array");
        if (n == 0) return;
        const int nm = (int)m->views.size();
        for (uint32_t i = 0; i < n; ++i) {
            if (event == NBL_EVENT_INTERSECT) {
                require(material[i] < nm, "nbl_physics_batch: material index out of range");
            } else {
                require(material[i] >= 0 && material[i] < nm, "nbl_physics_batch: material index out of range");
            }
        }
        // triangle records by input id, for intersect events
        dbuf<tri_rec> tris;
        dbuf<uint32_t> dlast;
        if (event == NBL_EVENT_INTERSECT) {
            const tri_rec* recs = geometry_host_recs(g);
            nbl_geometry_info info;
            if (nbl_geometry_get_info(g, &info) != NBL_OK) throw invalid_arg("nbl_physics_batch: bad geometry");
            std::vector<tri_rec> by_id(info.n_triangles_in);
            std::vector<uint8_t> have(info.n_triangles_in, 0);
            for (uint32_t k = 0; k < info.n_leaf_refs; ++k) { by_id[recs[k].id] = recs[k]; have[recs[k].id] = 1; }
            for (uint32_t i = 0; i < n; ++i) {
                require(last_tri[i] < info.n_triangles_in && have[last_tri[i]], "nbl_physics_batch: bad last_tri");
                const tri_rec& t = by_id[last_tri[i]];
                const int a = (int16_t)(t.mats & 0xFFFFu), b = (int16_t)(t.mats >> 16);
                require(a < nm && b < nm, "nbl_physics_batch: triangle material index beyond the material count");
            }
            tris.alloc(by_id.size());
            tris.upload(by_id.data(), by_id.size());
            dlast.alloc(n);
            dlast.upload(last_tri, n);
        }
        dbuf<int32_t> dmatl(n), dout_mat(n);
        dbuf<float> du((size_t)n * NBL_NUM_UNIFORMS), dscal(n);
        dbuf<dparticle> dio(n), dsec(n);
        dbuf<uint8_t> dst(n), dhas(n), dty(n);
        dmatl.upload(material, n);
        du.upload(uniforms, (size_t)n * NBL_NUM_UNIFORMS);
        static_assert(sizeof(dparticle) == sizeof(nbl_particle), "particle layout");
        dio.upload(reinterpret_cast<const dparticle*>(io), n);
        phys_args a{ m->dmats.p, tris.p, event, dmatl.p, dlast.p, du.p, n, dio.p, dout_mat.p, dst.p,
                     dsec.p, dhas.p, dscal.p, dty.p };
        hipLaunchKernelGGL(k_physics, dim3(grid_for(n, 256)), dim3(256), 0, 0, a);
        NBL_HIP(hipGetLastError());
        dio.download(reinterpret_cast<dparticle*>(io), n);
        if (out_material) dout_mat.download(out_material, n);
        if (out_status) dst.download(out_status, n);
        if (out_secondary) dsec.download(reinterpret_cast<dparticle*>(out_secondary), n);
        if (out_has_secondary) dhas.download(out_has_secondary, n);
        if (out_scalar) dscal.download(out_scalar, n);
        if (out_type) dty.download(out_type, n);
        NBL_HIP(hipDeviceSynchronize());
    });
}

}  // extern "C"
