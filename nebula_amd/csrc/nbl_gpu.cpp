// nbl_gpu.cpp -- command-line front end with the options, positional
// arguments, prescan and frame loop of the reference's nebula_gpu
// (gpu_main.cu:221-420 and its worker thread, gpu_main.cu:77-219), on one
// GPU, through libnbl_hip.so's C ABI.
//
//   nbl_gpu [options] <geometry.tri> <primaries.pri> <material0> .. <materialN>
//
// Options (same keys and defaults as the reference; --key=value or --key value):
//   --energy-threshold  lowest energy to simulate, eV          [0]
//   --capacity          electron capacity on the GPU           [1000000]
//   --prescan-size      electrons used for the prescan         [1000]
//   --batch-factor      multiplication factor for batch size   [0.9]
//   --seed              random seed                            [2026365851 = 0x78c7e39b]
//   --sort-primaries    sort primaries by position first       [0]
//   --inelastic         kieft | penn | auto                    [auto]
//   --persistent        0 | 1: run the primaries through the persistent
//                       transport kernel (one launch per 2^22 primaries)
//                       instead of the prescan and frame loop      [0]
//       (the reference fixes its inelastic model at compile time, full Penn by
//       default; auto = penn when every material file holds full Penn tables)
//
// Materials: the reference's HDF5 material files (cstool output), read
// through libnbl_mat.so (loaded from this program's directory on first use,
// include/nbl_mat.h), or binary .nblmat files holding the Kieft-Bosch raw
// tables (the nbl_kieft_raw fields), little-endian:
//   char magic[8] = "NBLMAT1\0"; uint32 n_k, n_p, n_outer_shells, 0;
//   float64 k_min, k_max, p_min, p_max, effective_a_grams;
//   float32 barrier, fermi, band_gap, phonon_loss;
//   float32 elastic_imfp[n_k], elastic_costheta[n_k][n_p],
//           inelastic_imfp[n_k], inelastic_w0[n_k][n_p], binding_icdf[n_k][n_p];
//   float64 outer_shells[n_outer_shells]
// (nebula_amd.synth.write_material_file writes it).
//
// Output: every detected electron as a 36-byte record on stdout, 7 float32
// (position, direction, energy) + 2 int32 (the primary's pixel), as the
// reference writes them (gpu_main.cu:199-206).  Messages go to stderr.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <dlfcn.h>
#include <unistd.h>

#include "../../include/nbl_hip.h"
#include "../../include/nbl_mat.h"

namespace {

void check(int rc, const char* what)
{
    if (rc != NBL_OK) throw std::runtime_error(std::string(what) + ": " + nbl_last_error());
}

// --key=value / --key value options, as the reference's cli_params (common/cli_params.h:54-90)
struct options {
    std::map<std::string, std::pair<std::string, std::string>> opt;   // key -> (description, value)
    std::vector<std::string> positional;
    void add(const std::string& k, const std::string& d, const std::string& v) { opt[k] = { d, v }; }
    void parse(int argc, char** argv)
    {
        for (int i = 1; i < argc; ++i) {
            const std::string item = argv[i];
            if (item.size() >= 2 && item[0] == '-' && item[1] == '-') {
                std::string key, value;
                const size_t eq = item.find('=', 2);
                if (eq == std::string::npos) {
                    key = item.substr(2);
                    value = ++i < argc ? argv[i] : "";
                    if (value == "=") value = ++i < argc ? argv[i] : "";
                } else {
                    key = item.substr(2, eq - 2);
                    value = item.substr(eq + 1);
                }
                auto it = opt.find(key);
                if (it == opt.end()) throw std::runtime_error("Unexpected key '" + key + "'");
                it->second.second = value;
            } else {
                positional.push_back(item);
            }
        }
    }
    template<typename T> T get(const std::string& k) const
    {
        std::istringstream ss(opt.at(k).second);
        T v{};
        ss >> v;
        return v;
    }
    void usage(std::ostream& o, const char* prog) const
    {
        o << "Usage: " << prog << " [options] <geometry.tri> <primaries.pri> [material0.nblmat] .. [materialN.nblmat]\n"
          << "Options:\n";
        for (auto& kv : opt) o << "\t--" << kv.first << ": " << kv.second.first << " [" << kv.second.second << "]\n";
    }
};

// libnbl_mat.so, opened on first use from the directory of this program
struct mat_api {
    void* so = nullptr;
    decltype(&nbl_mat_open) open = nullptr;
    decltype(&nbl_mat_close) close = nullptr;
    decltype(&nbl_mat_exists) exists = nullptr;
    decltype(&nbl_mat_kieft_raw) kieft_raw = nullptr;
    decltype(&nbl_mat_penn_raw) penn_raw = nullptr;
    decltype(&nbl_mat_last_error) last_error = nullptr;
    static mat_api& get()
    {
        static mat_api a;
        if (a.so) return a;
        char exe[4096] = {};
        const ssize_t n = readlink("/proc/self/exe", exe, sizeof exe - 1);
        std::string dir = n > 0 ? std::string(exe, (size_t)n) : std::string();
        dir = dir.substr(0, dir.find_last_of('/') + 1);
        a.so = dlopen((dir + "libnbl_mat.so").c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!a.so) throw std::runtime_error(std::string("HDF5 material files need libnbl_mat.so: ") + dlerror());
        auto sym = [&](auto& f, const char* name) {
            f = reinterpret_cast<std::remove_reference_t<decltype(f)>>(dlsym(a.so, name));
            if (!f) throw std::runtime_error(std::string("libnbl_mat.so lacks ") + name);
        };
        sym(a.open, "nbl_mat_open"); sym(a.close, "nbl_mat_close"); sym(a.exists, "nbl_mat_exists");
        sym(a.kieft_raw, "nbl_mat_kieft_raw"); sym(a.penn_raw, "nbl_mat_penn_raw");
        sym(a.last_error, "nbl_mat_last_error");
        return a;
    }
    void check(int rc, const char* what) const
    {
        if (rc != NBL_OK) throw std::runtime_error(std::string(what) + ": " + last_error());
    }
};

bool is_hdf5(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    char sig[8] = {};
    f.read(sig, 8);
    return f && std::memcmp(sig, "\x89HDF\r\n\x1a\n", 8) == 0;
}

// One material file (HDF5 or .nblmat) with its raw tables kept alive for the
// table preparation.
struct material_file {
    std::vector<float> el_imfp, el_cos, in_imfp, in_w0, binding;
    std::vector<double> outer;
    nbl_kieft_raw raw{};
    nbl_matfile h5 = nullptr;          // HDF5 file: owns the raw arrays
    bool has_penn = false;
    nbl_penn_raw penn{};
    material_file(const material_file&) = delete;
    material_file(material_file&& o) noexcept
        : el_imfp(std::move(o.el_imfp)), el_cos(std::move(o.el_cos)), in_imfp(std::move(o.in_imfp)),
          in_w0(std::move(o.in_w0)), binding(std::move(o.binding)), outer(std::move(o.outer)), raw(o.raw), h5(o.h5),
          has_penn(o.has_penn), penn(o.penn)
    {
        o.h5 = nullptr;
    }
    ~material_file() { if (h5) mat_api::get().close(h5); }
    explicit material_file(const std::string& path)
    {
        if (is_hdf5(path)) {
            const mat_api& m = mat_api::get();
            m.check(m.open(path.c_str(), &h5), "nbl_mat_open");
            m.check(m.kieft_raw(h5, &raw), "nbl_mat_kieft_raw");
            int32_t e = 0;
            m.check(m.exists(h5, "/full_penn", &e), "nbl_mat_exists");
            if (e) {
                m.check(m.penn_raw(h5, &penn), "nbl_mat_penn_raw");
                has_penn = true;
            }
            return;
        }
        std::ifstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("cannot open material file '" + path + "'");
        char magic[8];
        uint32_t hdr[4];
        double d[5];
        float fl[4];
        f.read(magic, 8);
        f.read(reinterpret_cast<char*>(hdr), sizeof hdr);
        f.read(reinterpret_cast<char*>(d), sizeof d);
        f.read(reinterpret_cast<char*>(fl), sizeof fl);
        if (!f || std::memcmp(magic, "NBLMAT1", 8) != 0) throw std::runtime_error("'" + path + "' is not an NBLMAT1 file");
        const uint32_t nk = hdr[0], np = hdr[1], no = hdr[2];
        if (nk < 2 || np < 2 || nk > (1u << 16) || np > (1u << 16) || no > 64)
            throw std::runtime_error("'" + path + "': implausible table sizes");
        auto rd = [&](auto& v, size_t n) {
            v.resize(n);
            f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(v[0])));
        };
        rd(el_imfp, nk); rd(el_cos, (size_t)nk * np); rd(in_imfp, nk); rd(in_w0, (size_t)nk * np);
        rd(binding, (size_t)nk * np); rd(outer, no);
        if (!f) throw std::runtime_error("'" + path + "': file too short");
        raw.barrier = fl[0]; raw.fermi = fl[1]; raw.band_gap = fl[2]; raw.phonon_loss = fl[3];
        raw.effective_a_grams = d[4];
        raw.n_k = (int32_t)nk; raw.k_min = d[0]; raw.k_max = d[1];
        raw.n_p = (int32_t)np; raw.p_min = d[2]; raw.p_max = d[3];
        raw.elastic_imfp = el_imfp.data(); raw.elastic_costheta = el_cos.data();
        raw.inelastic_imfp = in_imfp.data(); raw.inelastic_w0 = in_w0.data(); raw.binding_icdf = binding.data();
        raw.n_outer_shells = (int32_t)no; raw.outer_shells = outer.data();
    }
};

// Prepared tables of one material (nbl_kieft_prepare, then nbl_penn_prepare
// for full Penn) and its descriptor.
struct prepared_material {
    std::vector<float> el_imfp, el_icdf, in_imfp, in_w0, ion, p_imfp, p_omega, p_q, p_ion;
    nbl_material_desc desc{};
    prepared_material(const material_file& m, bool penn)
    {
        const size_t nk = m.raw.n_k, np = m.raw.n_p;
        el_imfp.resize(nk); in_imfp.resize(nk);
        el_icdf.resize(nk * np); in_w0.resize(nk * np); ion.resize(nk * np);
        check(nbl_kieft_prepare(&m.raw, el_imfp.data(), el_icdf.data(), in_imfp.data(), in_w0.data(), ion.data(), &desc),
              "nbl_kieft_prepare");
        if (!penn) return;
        const nbl_penn_raw& r = m.penn;
        p_imfp.resize((size_t)r.n_k); p_omega.resize((size_t)r.n_k * r.n_p);
        p_q.resize((size_t)r.q_nk * r.q_nw * r.q_np); p_ion.resize((size_t)r.i_nk * r.i_nw * r.i_np);
        check(nbl_penn_prepare(&r, p_imfp.data(), p_omega.data(), p_q.data(), p_ion.data(), &desc), "nbl_penn_prepare");
    }
};

// work_pool (common/work_pool.h:14-123) over the loaded primaries
struct work_pool {
    const nbl_particle* p;
    const uint32_t* tags;
    uint64_t n, next = 0;
    static void fill(uint8_t* data, nbl_particle* parts, uint32_t* t, uint32_t cap, void* user)
    {
        work_pool& w = *static_cast<work_pool*>(user);
        for (uint32_t i = 0; i < cap && w.next < w.n; ++i) {
            if (data[i] != 0) continue;
            parts[i] = w.p[w.next];
            t[i] = w.tags[w.next];
            data[i] = 1;
            ++w.next;
        }
    }
    bool done() const { return next >= n; }
};

struct record_writer {
    const std::vector<nbl_int2>* pixels;
    std::vector<char> buf;
    uint64_t count = 0;
    static void on_detect(const nbl_particle* p, uint32_t tag, void* user)
    {
        record_writer& w = *static_cast<record_writer*>(user);
        const float f[7] = { p->pos.x, p->pos.y, p->pos.z, p->dir.x, p->dir.y, p->dir.z, p->kin_energy };
        const nbl_int2 px = (*w.pixels)[tag];
        const int32_t i[2] = { px.x, px.y };
        const size_t o = w.buf.size();
        w.buf.resize(o + sizeof f + sizeof i);
        std::memcpy(w.buf.data() + o, f, sizeof f);
        std::memcpy(w.buf.data() + o + sizeof f, i, sizeof i);
        ++w.count;
        if (w.buf.size() >= (1u << 20)) w.flush();
    }
    void flush()
    {
        if (!buf.empty()) std::fwrite(buf.data(), 1, buf.size(), stdout);
        buf.clear();
    }
};

int run(int argc, char** argv)
{
    int maj = 0, min = 0, pat = 0;
    nbl_version(&maj, &min, &pat);
    std::clog << "This is nbl_gpu (libnbl_hip " << maj << "." << min << "." << pat << ")\n\n";
    options p;
    p.add("energy-threshold", "Lowest energy to simulate", "0");
    p.add("capacity", "Electron capacity on the GPU", "1000000");
    p.add("prescan-size", "Number of electrons to use for prescan", "1000");
    p.add("batch-factor", "Multiplication factor for electron batch size", "0.9");
    p.add("seed", "Random seed", std::to_string(0x78c7e39bu));
    p.add("sort-primaries", "Sort primary electrons before simulation", "0");
    p.add("inelastic", "Inelastic model: kieft, penn or auto", "auto");
    p.add("persistent", "Use the persistent transport kernel instead of the frame loop", "0");
    p.parse(argc, argv);
    const float energy_threshold = p.get<float>("energy-threshold");
    const uint32_t capacity = p.get<uint32_t>("capacity");
    const uint32_t prescan_size = p.get<uint32_t>("prescan-size");
    const float batch_factor = p.get<float>("batch-factor");
    const uint32_t seed = p.get<uint32_t>("seed");
    const bool sort_primaries = p.get<bool>("sort-primaries");
    const std::string inelastic = p.get<std::string>("inelastic");
    const bool persistent = p.get<bool>("persistent");
    if (inelastic != "kieft" && inelastic != "penn" && inelastic != "auto") {
        std::clog << "Error: --inelastic must be kieft, penn or auto\n";
        return 1;
    }
    const auto& pos = p.positional;
    if (pos.size() < 3 || capacity == 0 || prescan_size == 0 || !(batch_factor > 0)) {
        p.usage(std::clog, argv[0]);
        return 1;
    }

    // geometry
    std::clog << "Loading geometry..." << std::endl;
    uint32_t n_tri = 0;
    check(nbl_tri_file_load(pos[0].c_str(), nullptr, 0, &n_tri), "nbl_tri_file_load");
    std::vector<nbl_triangle> tris(n_tri);
    if (n_tri) check(nbl_tri_file_load(pos[0].c_str(), tris.data(), n_tri, &n_tri), "nbl_tri_file_load");
    if (tris.empty()) {
        std::clog << "Error: could not load triangles!\n";
        p.usage(std::clog, argv[0]);
        return 1;
    }
    {
        int max_material = -1;
        for (const auto& t : tris) max_material = std::max({ max_material, t.material_in, t.material_out });
        const int given = (int)pos.size() - 2;
        if (max_material > given - 1) {
            std::clog << "Error: not enough materials provided for this geometry!\n  Expected " << (max_material + 1)
                      << " materials, " << given << " provided.\n";
            return 1;
        }
        if (max_material < given - 1)
            std::clog << "Warning: too many materials provided for this geometry!\n  Expected " << (max_material + 1)
                      << " materials, " << given << " provided.\n";
    }
    nbl_geometry g = nullptr;
    check(nbl_geometry_create(tris.data(), n_tri, &g), "nbl_geometry_create");
    nbl_geometry_info gi{};
    check(nbl_geometry_get_info(g, &gi), "nbl_geometry_get_info");

    // materials
    std::clog << "Loading materials..." << std::endl;
    std::vector<material_file> files;
    std::vector<prepared_material> prep;
    files.reserve(pos.size() - 2);
    prep.reserve(pos.size() - 2);
    bool all_penn = true;
    for (size_t k = 2; k < pos.size(); ++k) {
        files.emplace_back(pos[k]);
        all_penn = all_penn && files.back().has_penn;
        if (inelastic == "penn" && !files.back().has_penn)
            throw std::runtime_error("Full Penn model not found in material file " + pos[k]);
    }
    const bool penn = inelastic == "penn" || (inelastic == "auto" && all_penn);
    for (size_t k = 2; k < pos.size(); ++k) {
        prep.emplace_back(files[k - 2], penn);
        std::clog << "  Material " << (k - 2) << ": " << pos[k] << "\n";
    }
    std::clog << "Physics models:\n  Kieft-Bosch elastic (Mott + phonon), "
              << (penn ? "full Penn inelastic" : "Kieft-Bosch inelastic")
              << ", boundary with quantum-mechanical transmission and refraction\n";
    std::vector<nbl_material_desc> descs;
    for (auto& m : prep) descs.push_back(m.desc);
    nbl_materials mats = nullptr;
    check(nbl_materials_create(descs.data(), (int32_t)descs.size(), &mats), "nbl_materials_create");
    float max_energy = 0;
    check(nbl_materials_get_max_energy(mats, &max_energy), "nbl_materials_get_max_energy");

    // primaries
    std::clog << "Loading primary electrons..." << std::endl;
    uint64_t n_pri = 0;
    check(nbl_pri_file_load(pos[1].c_str(), gi.aabb_min, gi.aabb_max, max_energy, nullptr, nullptr, 0, &n_pri, nullptr),
          "nbl_pri_file_load");
    std::vector<nbl_particle> prim(n_pri);
    std::vector<nbl_int2> pixels(n_pri);
    nbl_pri_warnings w{};
    if (n_pri)
        check(nbl_pri_file_load(pos[1].c_str(), gi.aabb_min, gi.aabb_max, max_energy, prim.data(), pixels.data(), n_pri,
                                &n_pri, &w), "nbl_pri_file_load");
    if (w.outside_geometry) std::clog << "WARNING: " << w.outside_geometry << " primary electrons starting outside geometry.\n";
    if (w.low_energy) std::clog << "WARNING: " << w.low_energy << " primary electrons starting with low kinetic energy.\n";
    if (w.high_energy)
        std::clog << "WARNING: " << w.high_energy << " primary electrons starting with kinetic energy higher than what"
                  << " is available in the material files. Limit is " << (max_energy / 1000) << " keV.\n";
    if (w.bad_direction) std::clog << "WARNING: " << w.bad_direction << " primary electrons start with unphysical direction vector.\n";
    if (sort_primaries) check(nbl_pri_sort_morton(prim.data(), pixels.data(), n_pri), "nbl_pri_sort_morton");
    check(nbl_pri_prescan_shuffle(prim.data(), pixels.data(), n_pri, prescan_size), "nbl_pri_prescan_shuffle");
    if (prim.empty()) {
        std::clog << "Error: could not load primary electrons!\n";
        p.usage(std::clog, argv[0]);
        return 1;
    }
    std::vector<uint32_t> tags(n_pri);
    std::iota(tags.begin(), tags.end(), 0u);     // GPU tag = index of the primary (gpu_main.cu:376-380)
    work_pool pool{ prim.data(), tags.data(), n_pri };
    std::clog << "\nLoaded " << n_tri << " triangles.\n  min = {" << gi.aabb_min.x << ", " << gi.aabb_min.y << ", "
              << gi.aabb_min.z << "}\n  max = {" << gi.aabb_max.x << ", " << gi.aabb_max.y << ", " << gi.aabb_max.z
              << "}\nLoaded " << n_pri << " primaries.\nLoaded " << descs.size() << " materials.\n\n" << std::flush;

    nbl_driver d = nullptr;
    check(nbl_driver_create(capacity, g, mats, energy_threshold, max_energy, seed, nullptr, &d), "nbl_driver_create");
    if (persistent) {
        // whole cascades in persistent launches; detected electrons come back
        // as records with their primary's tag (= index), in detection order
        const auto t0 = std::chrono::steady_clock::now();
        record_writer out{ &pixels, {}, 0 };
        constexpr uint64_t kChunk = 1ull << 22;
        uint64_t launches = 0;
        for (uint64_t o = 0; o < n_pri; o += kChunk) {
            const uint64_t k = std::min<uint64_t>(kChunk, n_pri - o);
            uint64_t cap_records = 4 * k;
            for (;;) {
                std::vector<nbl_particle> det((size_t)cap_records);
                std::vector<uint32_t> det_tags((size_t)cap_records);
                nbl_detection_buffer db{ det.data(), det_tags.data(), cap_records, 0 };
                nbl_tallies t{};
                t.detected = &db;
                check(nbl_driver_run_persistent_list(d, prim.data() + o, tags.data() + o, k, nullptr, &t),
                      "nbl_driver_run_persistent_list");
                ++launches;
                if (db.count > db.capacity) {     // more detections than room: run the chunk again
                    cap_records = 2 * db.count;
                    continue;
                }
                for (uint64_t i = 0; i < db.count; ++i) record_writer::on_detect(&det[i], det_tags[i], &out);
                break;
            }
            std::clog << " \rProgress " << std::fixed << (100.0 * (double)(o + k) / (double)n_pri) << "%" << std::flush;
        }
        out.flush();
        std::fflush(stdout);
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::clog << "\nSimulation (persistent, " << launches << " launches): " << n_pri << " primaries, " << out.count
                  << " detected electrons, " << sec << " s\n";
        nbl_driver_destroy(d);
        nbl_materials_destroy(mats);
        nbl_geometry_destroy(g);
        return 0;
    }
    const auto t0 = std::chrono::steady_clock::now();

    // prescan (gpu_main.cu:108-166): frame_size = 1 + iteration of the largest
    // running count; batch_size from the running/detected counts
    std::vector<std::pair<uint32_t, uint32_t>> stats;
    {
        const uint32_t k = (uint32_t)std::min<uint64_t>(prescan_size, n_pri - pool.next);
        uint32_t pushed = 0;
        check(nbl_driver_push(d, prim.data() + pool.next, tags.data() + pool.next, k, &pushed), "nbl_driver_push");
        pool.next += pushed;
        stats.push_back({ pushed, 0 });
    }
    while (stats.back().first > 0) {
        check(nbl_driver_do_iteration(d), "nbl_driver_do_iteration");
        uint32_t r = 0, det = 0;
        check(nbl_driver_get_running_count(d, &r), "get_running_count");
        check(nbl_driver_get_detected_count(d, &det), "get_detected_count");
        stats.push_back({ r, det });
        std::clog << " \rExecuting pre-scan | running: " << r << " | detected: " << det;
    }
    const uint32_t frame_size = 1 + (uint32_t)std::distance(stats.begin(),
        std::max_element(stats.begin(), stats.end(), [](auto a, auto b) { return a.first < b.first; }));
    uint32_t batch_size;
    {
        float acc = 0;
        for (size_t i = 2 * (size_t)frame_size; i < stats.size(); i += frame_size) acc += stats[i].first / float(prescan_size);
        const size_t f = std::min<size_t>(frame_size, stats.size() - 1);
        acc += 2 * stats[f].first / float(prescan_size);
        acc += 2 * stats[f].second / float(prescan_size);
        batch_size = acc > 0 ? (uint32_t)(batch_factor * capacity / acc) : capacity;
        batch_size = std::max<uint32_t>(1, std::min(batch_size, capacity));
    }
    std::clog << "\nframe_size = " << frame_size << " | batch_size = " << batch_size << std::endl;

    // frame loop (gpu_main.cu:179-213)
    record_writer out{ &pixels, {}, 0 };
    check(nbl_driver_allocate_input_buffers(d, batch_size), "allocate_input_buffers");
    check(nbl_driver_push_to_buffer(d, &work_pool::fill, &pool), "push_to_buffer");
    auto last_print = std::chrono::steady_clock::now();
    for (;;) {
        check(nbl_driver_push_to_simulation(d), "push_to_simulation");
        check(nbl_driver_buffer_detected(d), "buffer_detected");
        for (uint32_t i = 0; i < frame_size; ++i) check(nbl_driver_do_iteration(d), "do_iteration");
        check(nbl_driver_push_to_buffer(d, &work_pool::fill, &pool), "push_to_buffer");
        uint32_t running = 0;
        check(nbl_driver_flush_buffered(d, &record_writer::on_detect, &out, &running), "flush_buffered");
        const auto now = std::chrono::steady_clock::now();
        if (now - last_print > std::chrono::seconds(1)) {
            std::clog << " \rProgress " << std::fixed << (100.0 * (double)pool.next / (double)n_pri)
                      << "% | Running: " << running << std::flush;
            last_print = now;
        }
        if (running == 0 && pool.done()) break;
    }
    out.flush();
    std::fflush(stdout);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::clog << "\nSimulation: " << n_pri << " primaries, " << out.count << " detected electrons, " << s << " s\n";
    nbl_driver_destroy(d);
    nbl_materials_destroy(mats);
    nbl_geometry_destroy(g);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::clog << "Error: " << e.what() << "\n";
        return 1;
    }
}
