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
template<bool COUNT, typename R>
__device__ __forceinline__ bool init_sample(const dgeom& g, const dmat* __restrict__ mats, float min_energy,
    float max_energy, dparticle& p, int mat, R& rng, lane_counters& c, float& dist, uint8_t& type)
{
    if (COUNT) ++c.events;
    if (!in_domain(g, p.pos)) return false;
    type = 0;
    dist = g.max_extent;
    if (mat >= 0) {
        const dmat& m = mats[mat];
        if (p.e < m.barrier + min_energy || p.e > m.barrier + max_energy) return false;
        dist = sample_path(m, p.e, rng, &type);
        if (COUNT) c.lookups_1d += 2;
    }
    p.dir = normalised_fast(p.dir);      // rnorm3df in the reference's GPU build (gpu_driver.inl:458)
    return true;
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

}  // namespace nbl
