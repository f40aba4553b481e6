// wave_propagate.h -- octree::propagate for all 64 lanes of a wavefront at
// once, with the leaf triangle tests shared across the wavefront.
//
// The scalar walk (nbl_device.h:propagate) loops over its leaf's triangles;
// in a wavefront that costs the longest leaf list of any lane, which near a
// rough surface is ~15x the mean.  Here every lane still walks its own cells
// (descent, exit plane, neighbour step), but the candidates of all lanes'
// leaves are pooled and tested 64 at a time, one candidate per lane.  Each
// lane then replays the reference's sequential rule "t > 0 && t <= best +
// EPSILON" over its own candidates' distances in leaf order, so results are
// identical to the scalar walk's (and the reference's), bit for bit.
#pragma once

#include "transport_device.h"

namespace nbl {

// Exclusive prefix sum over the 64 lanes (all lanes active); *total = sum.
// Inclusive scan with DPP lane moves (no LDS, no address registers): row_shr
// 1, 2, 4, 8 within each row of 16 lanes (a lane with no source adds 0),
// then row_bcast:15 adds each row's last lane to the next row (rows 1 and
// 3) and row_bcast:31 adds lane 31 to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_exclusive_sum(uint32_t v, uint32_t& total)
{
    uint32_t x = v;
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);   // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);   // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);   // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);   // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);   // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);   // row_bcast:31
    total = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    return x - v;
}

// Inclusive maximum over lanes 0..lane (all lanes active), with the lane
// moves of wave_exclusive_sum; a lane with no source keeps its value.
__device__ __forceinline__ uint32_t wave_inclusive_max(uint32_t x)
{
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false));
    return x;
}

// Owner of candidate base + lane: the last lane whose offset is <= it (the
// lane found has count > 0 whenever base + lane < total).  Every lane sends
// its index to the lane at (off - base), or to lane 0 when that is outside
// the round; a forward permute keeps the highest sender per lane, and lanes
// with count 0 share their offset with a higher lane, so lane s receives the
// last lane whose range starts at base + s (0 if none).  Lane 0 takes the
// owner of item base from a ballot (the last started range), and a running
// maximum gives every lane its owner.
__device__ __forceinline__ uint32_t pooled_owner(uint32_t off, uint32_t count, uint32_t base)
{
    const uint32_t lane = lane_id();
    const uint32_t slot = off - base < 64u ? off - base : 0u;
    uint32_t mark = (uint32_t)__builtin_amdgcn_ds_permute((int)(slot * 4u), (int)lane);
    const uint64_t started = __ballot(count > 0 && off <= base);
    if (lane == 0) mark = 63u - (uint32_t)__builtin_clzll(started | 1ull);
    return wave_inclusive_max(mark);
}

// pooled_owner carrying a payload < 2^26 of the owner lane along: the
// permute and the running maximum act on (lane << 26 | payload), whose
// order is the lanes' order, so every lane receives its owner's payload
// with its owner (key >> 26) and a round needs no further lane move.
constexpr uint32_t OWNER_PAYLOAD_BITS = 26;
__device__ __forceinline__ uint32_t pooled_owner_key(uint32_t off, uint32_t count, uint32_t base, uint32_t payload)
{
    const uint32_t lane = lane_id();
    const uint32_t key = (lane << OWNER_PAYLOAD_BITS) | payload;
    const uint32_t slot = off - base < 64u ? off - base : 0u;
    uint32_t mark = (uint32_t)__builtin_amdgcn_ds_permute((int)(slot * 4u), (int)key);
    const uint64_t started = __ballot(count > 0 && off <= base);
    const uint32_t open = 63u - (uint32_t)__builtin_clzll(started | 1ull);
    const uint32_t open_key = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)open);
    if (lane == 0) mark = open_key;
    return wave_inclusive_max(mark);
}

// A lane's walk between cells (the state octree::propagate keeps across
// iterations of its cell loop, octree.inl:32-190).
struct walk_state {
    f3 pos;              // current point
    float to_go;         // distance left
    uint64_t location;   // code of the next cell
    uint32_t cells;      // cells visited
    uint32_t rope;       // rope of the next cell (leaf_rope index), NO_ROPE for the first
};

__device__ __forceinline__ walk_state walk_begin(f3 start, float distance)
{
    return walk_state{ start, distance, 1ull, 0u, NO_ROPE };
}

// One cell of the walk of every lane with run == true; must be called by all
// 64 lanes of the wavefront (uniform control flow); lanes not running take
// part in the shared triangle tests only.  Returns true for lanes whose walk
// ended in this cell, with *hit (a triangle's input index or NO_TRI) and
// *travelled (propagate's return value for a walk of `distance`).
template<bool COUNT, typename CACHE = no_path_cache>
__device__ __forceinline__ bool walk_cell(bool run, const dgeom& g, walk_state& w, f3 dir, float distance,
    uint32_t ignore_tri, int ignore_mat, uint32_t* hit, float* travelled, prop_counters* cnt, CACHE* pc = nullptr)
{
    const uint32_t lane = lane_id();
    uint32_t first = 0, count = 0, leaf_id = 0;
    int axis = 0;
    float best = 0;
    if (run) {
        f3 c, s;
        const uint64_t child = descend<COUNT>(g, w.pos, w.location, c, s, cnt, pc, w.rope);
        leaf_id = (uint32_t)child;
        count = (uint32_t)((child >> 32) & 0x7FFFFFFFu);
        best = leaf_exit(c, s, w.pos, dir, axis);
        if (count > 0) {
            const leaf_info li = load_leaf(g, (uint32_t)child);
            if (COUNT) ++cnt->node_steps;        // the leaf record
            first = li.first;
            bool slab_culled;
            if (leaf_dropped(li, s, w.pos, dir, best, w.to_go, ignore_mat, slab_culled)) count = 0;
        }
    }
    uint32_t total;
    const uint32_t off = wave_exclusive_sum(count, total);
    uint32_t best_id = NO_TRI;
    // candidates [off, off + count) of every lane, 64 per round
    // a candidate's record is (first - off) + item of its owner: carried with
    // the owner search where record indices fit its payload
    const uint32_t rel = first - off;
    for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t item = base + lane;
        uint32_t owner, k;
        if (g.owner_payload) {
            constexpr uint32_t mask = (1u << OWNER_PAYLOAD_BITS) - 1;
            const uint32_t key = pooled_owner_key(off, count, base, rel & mask);
            owner = key >> OWNER_PAYLOAD_BITS;
            k = (key + item) & mask;
        } else {
            owner = pooled_owner(off, count, base);
            k = __shfl(rel, (int)owner) + item;
        }
        const f3 op{ __shfl(w.pos.x, (int)owner), __shfl(w.pos.y, (int)owner), __shfl(w.pos.z, (int)owner) };
        const f3 od{ __shfl(dir.x, (int)owner), __shfl(dir.y, (int)owner), __shfl(dir.z, (int)owner) };
        const uint32_t oit = __shfl(ignore_tri, (int)owner);
        const int oim = __shfl(ignore_mat, (int)owner);
        float t = -1;
        uint32_t id = NO_TRI;
        if (item < total) t = leaf_candidate<COUNT>(g.recs, k, op, od, oit, oim, id, cnt);
        // Only candidates with t > 0 can replace best.  They are few; take
        // them in item order (= each owner's leaf order) and let the owner
        // apply the reference's rule to each of its own.
        uint64_t m = __ballot(t > 0);
        while (m) {
            const uint32_t b = (uint32_t)__builtin_ctzll(m);
            m &= m - 1;
            const float tq = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t), b));
            const uint32_t idq = (uint32_t)__builtin_amdgcn_readlane((int)id, b);
            const uint32_t it = base + b;
            if (it >= off && it < off + count && (tq <= best + EPSILON)) { best = tq; axis = 0; best_id = idq; }
        }
    }
    if (!run) return false;
    if (best >= w.to_go) {
        *hit = NO_TRI; *travelled = distance;
        return true;
    }
    if (axis == 0) {
        *hit = best_id; *travelled = distance - w.to_go + best;
        return true;
    }
    w.to_go -= best;
    w.pos = w.pos + dir * best;
    next_cell(w.location, axis, dir);
    {   // the face crossed: axis -1, -2, -4 and the sign rule of next_cell
        const uint32_t ax = axis == -1 ? 0u : (axis == -2 ? 1u : 2u);
        const float da = ax == 0 ? dir.x : (ax == 1 ? dir.y : dir.z);
        w.rope = leaf_id * 6u + 2u * ax + (da >= 0 ? 0u : 1u);
    }
    if (++w.cells >= MAX_CELLS || w.location <= 1) {
        *hit = NO_TRI; *travelled = distance;
        return true;
    }
    return false;
}

// octree::propagate for all 64 lanes at once (all must call it); lanes with
// active == false take part in the shared tests only.
template<bool COUNT, typename CACHE = no_path_cache>
__device__ __forceinline__ float propagate_wave(bool active, const dgeom& g, f3 start, f3 dir, float distance,
    uint32_t ignore_tri, int ignore_mat, uint32_t* hit, prop_counters* cnt, CACHE* pc = nullptr)
{
    walk_state w = walk_begin(start, distance);
    float result = distance;
    *hit = NO_TRI;
    bool run = active;
    while (__ballot(run)) {
        if (walk_cell<COUNT, CACHE>(run, g, w, dir, distance, ignore_tri, ignore_mat, hit, &result, cnt, pc))
            run = false;
    }
    return result;
}

}  // namespace nbl
