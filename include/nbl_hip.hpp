// nbl_hip.hpp -- header-only C++ facade over the C ABI of libnbl_hip.so with
// the method set of the reference's hot-path classes, so that host code
// written against them switches by changing types:
//
//   nbl::geometry::octree<true>              -> nbl_hip::octree
//       (geometry/octree.h:42-97)
//   nbl::gpu_material_manager<material_t>    -> nbl_hip::material_manager
//       (core/gpu_material_manager.h:37-74)
//   nbl::drivers::gpu_driver<sl, is, geom>   -> nbl_hip::gpu_driver
//       (drivers/gpu/gpu_driver.h:58-182)
//
// Error behaviour follows the reference: allocation failures throw
// std::bad_alloc (common/cuda/cuda_new.h:22-26), everything else
// std::runtime_error with the library's message.  Unlike the reference,
// objects own their device memory (RAII); create()/destroy() pairs are kept
// for source compatibility.
#ifndef NBL_HIP_HPP
#define NBL_HIP_HPP

#include <cstdint>
#include <functional>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "nbl_hip.h"

namespace nbl_hip {

inline void check(int rc)
{
    if (rc == NBL_OK) return;
    if (rc == NBL_ERR_ALLOC) throw std::bad_alloc();
    throw std::runtime_error(nbl_last_error());
}

using particle = nbl_particle;
using triangle = nbl_triangle;

class octree {
public:
    octree() = default;
    explicit octree(const std::vector<triangle>& tris) { check(nbl_geometry_create(tris.data(), (uint32_t)tris.size(), &h_)); }
    octree(octree&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    octree& operator=(octree&& o) noexcept { std::swap(h_, o.h_); return *this; }
    octree(const octree&) = delete;
    ~octree() { if (h_) nbl_geometry_destroy(h_); }

    static octree create(const std::vector<triangle>& tris) { return octree(tris); }
    static void destroy(octree& g) { octree tmp(std::move(g)); }

    nbl_geometry_info info() const { nbl_geometry_info i; check(nbl_geometry_get_info(h_, &i)); return i; }
    nbl_vec3 AABB_min() const { return info().aabb_min; }
    nbl_vec3 AABB_max() const { return info().aabb_max; }
    float get_max_extent() const { return info().max_extent; }

    // Batched forms of in_domain and propagate (host arrays).
    std::vector<uint8_t> in_domain(const std::vector<nbl_vec3>& pos) const
    {
        std::vector<uint8_t> out(pos.size());
        check(nbl_geometry_in_domain_batch(h_, pos.data(), (uint32_t)pos.size(), out.data()));
        return out;
    }
    void propagate(const std::vector<nbl_vec3>& pos, const std::vector<nbl_vec3>& dir, const std::vector<float>& dist,
        const std::vector<uint32_t>& ignore_tri, const std::vector<int32_t>& ignore_mat,
        std::vector<float>& out_dist, std::vector<uint32_t>& out_tri) const
    {
        const uint32_t n = (uint32_t)pos.size();
        out_dist.resize(n); out_tri.resize(n);
        check(nbl_geometry_propagate_batch(h_, pos.data(), dir.data(), dist.data(), ignore_tri.data(),
            ignore_mat.data(), n, out_dist.data(), out_tri.data()));
    }
    nbl_geometry handle() const { return h_; }

private:
    nbl_geometry h_ = nullptr;
};

class material_manager {
public:
    material_manager() = default;
    explicit material_manager(const std::vector<nbl_material_desc>& m)
    {
        check(nbl_materials_create(m.data(), (int32_t)m.size(), &h_));
    }
    material_manager(material_manager&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    material_manager(const material_manager&) = delete;
    ~material_manager() { if (h_) nbl_materials_destroy(h_); }

    static material_manager create(const std::vector<nbl_material_desc>& m) { return material_manager(m); }
    static void destroy(material_manager& m) { material_manager tmp(std::move(m)); }
    static bool is_physical(int32_t material_idx) { return material_idx >= 0; }
    float get_max_energy() const { float e; check(nbl_materials_get_max_energy(h_, &e)); return e; }
    nbl_materials handle() const { return h_; }

private:
    nbl_materials h_ = nullptr;
};

class gpu_driver {
public:
    using particle_index_t = uint32_t;
    using seed_t = uint64_t;

    gpu_driver(particle_index_t capacity, const material_manager& materials, const octree& geometry,
        float min_energy, float max_energy, seed_t seed = 0)
        : _min_energy(min_energy), _max_energy(max_energy)
    {
        check(nbl_driver_create(capacity, geometry.handle(), materials.handle(), min_energy, max_energy, seed, nullptr, &h_));
    }
    gpu_driver(const gpu_driver&) = delete;
    gpu_driver& operator=(const gpu_driver&) = delete;
    ~gpu_driver() { if (h_) nbl_driver_destroy(h_); }

    particle_index_t push(particle* particles, uint32_t* tags, particle_index_t N)
    {
        uint32_t pushed = 0;
        check(nbl_driver_push(h_, particles, tags, N, &pushed));
        return pushed;
    }
    template<typename detect_function>
    void flush_detected(detect_function function)
    {
        check(nbl_driver_flush_detected(h_, &trampoline<detect_function>, &function));
    }
    particle_index_t get_running_count() const { uint32_t n; check(nbl_driver_get_running_count(h_, &n)); return n; }
    particle_index_t get_detected_count() const { uint32_t n; check(nbl_driver_get_detected_count(h_, &n)); return n; }
    void allocate_input_buffers(particle_index_t N) { check(nbl_driver_allocate_input_buffers(h_, N)); }
    // pool must offer get_work(bool-like* data, particle*, uint32_t* tags, uint32_t N)
    template<typename work_pool>
    void push_to_buffer(work_pool& pool)
    {
        check(nbl_driver_push_to_buffer(h_, &fill_trampoline<work_pool>, &pool));
    }
    void push_to_simulation() { check(nbl_driver_push_to_simulation(h_)); }
    void buffer_detected() { check(nbl_driver_buffer_detected(h_)); }
    template<typename detect_function>
    particle_index_t flush_buffered(detect_function function)
    {
        uint32_t running = 0;
        check(nbl_driver_flush_buffered(h_, &trampoline<detect_function>, &function, &running));
        return running;
    }
    void do_iteration() { check(nbl_driver_do_iteration(h_)); }
    nbl_driver handle() const { return h_; }

    // Every primary run to the end of its cascade in persistent launches
    // (no reference counterpart: replaces the prescan + frame loop of
    // gpu_main.cu:108-213); function(particle, tag) is called for each
    // detected electron, in detection order.
    template<typename detect_function>
    nbl_tallies run_to_end(const particle* primaries, const uint32_t* tags, uint64_t n, detect_function function,
        const nbl_persistent_opts* opts = nullptr)
    {
        uint64_t cap = 4 * n + 64;
        for (;;) {
            std::vector<particle> det(cap);
            std::vector<uint32_t> det_tags(cap);
            nbl_detection_buffer db{ det.data(), det_tags.data(), cap, 0 };
            nbl_tallies t{};
            t.detected = &db;
            check(nbl_driver_run_persistent_list(h_, primaries, tags, n, opts, &t));
            if (db.count > cap) { cap = 2 * db.count; continue; }
            for (uint64_t i = 0; i < db.count; ++i) function(det[i], det_tags[i]);
            t.detected = nullptr;
            return t;
        }
    }

    float _min_energy, _max_energy;

private:
    template<typename F>
    static void trampoline(const nbl_particle* p, uint32_t tag, void* user) { (*static_cast<F*>(user))(*p, tag); }
    template<typename P>
    static void fill_trampoline(uint8_t* data, nbl_particle* parts, uint32_t* tags, uint32_t n, void* user)
    {
        static_cast<P*>(user)->get_work(data, parts, tags, n);
    }
    nbl_driver h_ = nullptr;
};

}  // namespace nbl_hip

#endif
