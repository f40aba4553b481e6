// transport_device.h -- one transport step of one electron, shared by the
// iteration driver and the persistent kernel.
#pragma once

#include "nbl_device.h"

namespace nbl {

// Outcome of the init step (the event the particle will execute next).
enum : uint8_t { EV_NONE = 0, EV_INELASTIC = 1, EV_ELASTIC = 2, EV_INTERSECT = 3, EV_TERMINATED = 4 };

struct lane_counters {
    uint32_t events = 0, lookups_1d = 0, lookups_2d = 0, rounddowns = 0;
    uint32_t node_steps = 0, tri_tests = 0, boundary = 0, secondaries = 0, pending = 0;
};

// kernels::init body (drivers/gpu/gpu_driver.inl:403-481): domain check,
// energy thresholds w.r.t. the vacuum level, free-path sampling in a
// physical material, normalisation, octree propagation and the move.
// p.dir is stored normalised (gpu_driver.inl:458, 480).
//
// init_sample is the part before the walk: false if the particle
// terminates, else the distance to travel and the sampled event type.
//
// init_sample_in is the same with the material given: m is the particle's
// material when physical (is_physical(material), material.h), else unused.
template<bool COUNT, typename R>
__device__ __forceinline__ bool init_sample_in(const dgeom& g, const dmat& m, bool physical, float min_energy,
    float max_energy, dparticle& p, R& rng, lane_counters& c, float& dist, uint8_t& type)
{
    if (COUNT) ++c.events;
    if (!in_domain(g, p.pos)) return false;
    type = 0;
    dist = g.max_extent;
    if (physical) {
        if (p.e < m.barrier + min_energy || p.e > m.barrier + max_energy) return false;
        dist = sample_path(m, p.e, rng, &type);
        if (COUNT) c.lookups_1d += 2;
    }
    p.dir = normalised_fast(p.dir);      // rnorm3df in the reference's GPU build (gpu_driver.inl:458)
    return true;
}
template<bool COUNT, typename R>
__device__ __forceinline__ bool init_sample(const dgeom& g, const dmat* __restrict__ mats, float min_energy,
    float max_energy, dparticle& p, int mat, R& rng, lane_counters& c, float& dist, uint8_t& type)
{
    // The fields path sampling reads, loaded together before the energy
    // test (the compiler would issue the tables' descriptors only after it).
    const dmat& src = mats[mat >= 0 ? mat : 0];
    dmat m;
    m.barrier = src.barrier;
    m.el_imfp = src.el_imfp;
    m.in_imfp = src.in_imfp;
    asm volatile("" ::"v"(m.barrier), "v"(m.el_imfp.d), "v"(m.el_imfp.n), "v"(m.el_imfp.x_min), "v"(m.el_imfp.x_step),
                 "v"(m.in_imfp.d), "v"(m.in_imfp.n), "v"(m.in_imfp.x_min), "v"(m.in_imfp.x_step));
    return init_sample_in<COUNT>(g, m, mat >= 0, min_energy, max_energy, p, rng, c, dist, type);
}

// The move after the walk: to the sampled event (EV_NONE, EV_INELASTIC or
// EV_ELASTIC) or to the hit triangle (EV_INTERSECT).
__device__ __forceinline__ uint8_t init_move(dparticle& p, float dist, float travelled, uint32_t hit, uint8_t type,
    uint32_t& last_tri)
{
    if (hit == NO_TRI) {
        p.pos = p.pos + p.dir * dist;
        return type;
    }
    p.pos = p.pos + p.dir * travelled;
    last_tri = hit;
    return EV_INTERSECT;
}

template<bool COUNT, typename R, typename CACHE = no_path_cache>
__device__ __forceinline__ uint8_t init_step(const dgeom& g, const dmat* __restrict__ mats, float min_energy,
    float max_energy, dparticle& p, int mat, uint32_t& last_tri, R& rng, lane_counters& c, CACHE* pc = nullptr)
{
    float dist;
    uint8_t type;
    if (!init_sample<COUNT>(g, mats, min_energy, max_energy, p, mat, rng, c, dist, type)) return EV_TERMINATED;
    uint32_t hit;
    prop_counters cnt{ 0, 0 };
    const float d = propagate<COUNT, CACHE>(g, p.pos, p.dir, dist, last_tri, mat, &hit, &cnt, pc);
    if (COUNT) { c.node_steps += cnt.node_steps; c.tri_tests += cnt.tri_tests; }
    return init_move(p, dist, d, hit, type, last_tri);
}

// Triangle record of an input triangle id.
__device__ __forceinline__ tri_rec load_tri(const dgeom& g, uint32_t id)
{
    const float4* rp = reinterpret_cast<const float4*>(g.recs + g.id_rec[id]);
    const float4 a = rp[0], b = rp[1], c = rp[2];
    tri_rec r;
    r.r0x = a.x; r.r0y = a.y; r.r0z = a.z; r.e1x = a.w;
    r.e1y = b.x; r.e1z = b.y; r.e2x = b.z; r.e2y = b.w;
    r.e2z = c.x; r.mats = __float_as_uint(c.y); r.id = __float_as_uint(c.z); r.pad = 0;
    return r;
}

// Synchronise the lanes of one wavefront around LDS and global-memory
// traffic between lanes (workgroup-scope fence: waits for outstanding
// loads and stores before any later access is issued).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
// number of set bits of mask below this lane
__device__ __forceinline__ uint32_t rank_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// The material all lanes with `need` are in, when they share one (in most
// loop iterations they do: one material under the beam, or one region of a
// scene); -1 when they differ or no lane needs one.  The result is
// wave-uniform, so a descriptor indexed by it is read with scalar loads.
__device__ __forceinline__ int wave_material(bool need, int mat)
{
    const uint64_t m = __ballot(need);
    if (m == 0) return -1;
    const int mu = __builtin_amdgcn_readlane(mat, (int)__builtin_ctzll(m));
    return __ballot(need && mat != mu) == 0 ? mu : -1;
}

// The material array as constant memory (address space 4): it is not
// written while a kernel runs, and uniform loads from this space are scalar
// loads (a generic pointer would give vector loads, one per lane).
typedef __attribute__((address_space(4))) const dmat cdmat;
__device__ __forceinline__ cdmat* const_mats(const dmat* m) { return (cdmat*)m; }

// A copy of material i of a constant material array, field by field (the
// fields a kernel does not use are never loaded).
__device__ __forceinline__ dmat load_material(cdmat* mats, int i)
{
    cdmat* c = mats + i;
    dmat m;
    m.barrier = c->barrier; m.fermi = c->fermi; m.band_gap = c->band_gap;
    m.phonon_loss = c->phonon_loss; m.recoil_const = c->recoil_const;
    m.el_imfp.d = c->el_imfp.d; m.el_imfp.n = c->el_imfp.n;
    m.el_imfp.x_min = c->el_imfp.x_min; m.el_imfp.x_step = c->el_imfp.x_step;
    m.in_imfp.d = c->in_imfp.d; m.in_imfp.n = c->in_imfp.n;
    m.in_imfp.x_min = c->in_imfp.x_min; m.in_imfp.x_step = c->in_imfp.x_step;
#define NBL_COPY_TAB2(f) \
    m.f.d = c->f.d; m.f.w = c->f.w; m.f.h = c->f.h; m.f.x_min = c->f.x_min; m.f.x_step = c->f.x_step; \
    m.f.y_min = c->f.y_min; m.f.y_step = c->f.y_step;
    NBL_COPY_TAB2(el_icdf) NBL_COPY_TAB2(in_icdf) NBL_COPY_TAB2(ion)
#undef NBL_COPY_TAB2
    m.model = c->model;
#define NBL_COPY_TAB3(f) \
    m.f.d = c->f.d; m.f.w = c->f.w; m.f.h = c->f.h; m.f.dp = c->f.dp; m.f.x_min = c->f.x_min; \
    m.f.x_step = c->f.x_step; m.f.y_min = c->f.y_min; m.f.y_step = c->f.y_step; m.f.z_min = c->f.z_min; \
    m.f.z_step = c->f.z_step;
    NBL_COPY_TAB3(pq) NBL_COPY_TAB3(pion)
#undef NBL_COPY_TAB3
    return m;
}

}  // namespace nbl
