// nbl_device.h -- device-side data layout and inline functions of the
// MI355X electron-transport path: octree traversal, Moller-Trumbore tests,
// table interpolation, XORWOW streams and the Kieft-Bosch / boundary physics.
//
// Numeric contract (shared with every kernel of the library): IEEE float32,
// correctly rounded division and square root, no fused multiply-add (the
// library is built with -ffp-contract=off), and the operation order of the
// reference, so that traversal and table interpolation are bit-identical to
// the reference's CPU code.  Transcendentals (log, exp, expm1, sin, cos,
// atan2) come from the ROCm device library and may differ from the host libm
// by an ulp or two; parts of the physics use the hardware square root,
// reciprocal and reciprocal square root (sqrt_fast, rcp_fast,
// normalised_fast below), about 1 ulp each.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nbl {

#define NBL_HD __host__ __device__ __forceinline__

constexpr float EPSILON = 1.1920928955078125e-06f;   // 10*FLT_EPSILON, config/data_types.h:94
constexpr float PI_F = 3.1415926535897932f;            // config/config.h:32
constexpr uint32_t NO_TRI = 0xFFFFFFFFu;
constexpr uint32_t MAX_CELLS = 1u << 20;               // cell-walk bound (never reached by a terminating walk)

// Special materials, core/cpu_material_manager.h:82-91
constexpr int MAT_NOP = -128, MAT_TERMINATOR = -127, MAT_DETECTOR = -126,
              MAT_DETECTOR_LT50 = -125, MAT_DETECTOR_GE50 = -124, MAT_VACUUM = -123,
              MAT_MIRROR = -122;

struct f3 { float x, y, z; };
NBL_HD f3 operator+(f3 a, f3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
NBL_HD f3 operator-(f3 a, f3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
NBL_HD f3 operator*(f3 a, float b) { return { a.x * b, a.y * b, a.z * b }; }
NBL_HD f3 operator*(float a, f3 b) { return { a * b.x, a * b.y, a * b.z }; }
NBL_HD f3 cross(f3 a, f3 b) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
NBL_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// normalise / normalised in their CPU form, vector_math.h:128-158: a * (1 / |a|)
NBL_HD f3 normalised(f3 a) { const float r = 1 / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); return a * r; }

// ---------------------------------------------------------------------------
// Geometry layout in HBM
//   nodes : 8 x uint64 child words per internal octree node (one 64-B line);
//           node 0 is the root.  A child word is either an internal node
//           index (bit 63 clear) or a leaf (bit 63 set) with its record
//           count in bits [32,63) and its ordinal in bits [0,32): the index
//           of its leaf_info (first record, cull bounds) and of its ropes.
//   recs  : 48-B triangle records, stored once per leaf reference, in the
//           leaf's triangle order, so a leaf is one contiguous run.
// ---------------------------------------------------------------------------
constexpr uint64_t LEAF_BIT = 1ull << 63;

struct __attribute__((aligned(16))) tri_rec {
    float r0x, r0y, r0z, e1x;
    float e1y, e1z, e2x, e2y;
    float e2z; uint32_t mats;  // int16 material_in | int16 material_out << 16
    uint32_t id;               // index of the triangle in the create() input
    uint32_t pad;
};
static_assert(sizeof(tri_rec) == 48, "tri_rec is 48 bytes");

// Cull record of a leaf (leaf_info), in a dense array indexed by the leaf's
// child word; it also holds the leaf's first record and its materials (every
// triangle's, when they agree).  Empty leaves have t = 2 and no slab.
//
// Side cone (ax, ay, az, t): if every triangle of the leaf has the same
// materials and the float normals cross(e1, e2) all lie within angle theta
// of the unit axis a, then for a direction d with |d|^2 within 1e-4 of 1,
// dot(a, d) > t implies that the reference's side test
// fl(dot(cross(e1, e2), d)) < 0 is false for every triangle of the leaf,
// and dot(a, d) < -t that it is true for every one: t = sin(theta + 1e-5)
// (1 + 1e-4) + 1e-6 covers the rounding of both dot products (< 2e-7 |n| |d|
// and < 4e-7) and the length of d.  t = 2 for leaves where this does not
// hold (mixed materials, a near-degenerate triangle, a cone wider than
// acos(0.2)).  Used to skip a leaf all of whose triangles the walk would
// skip as facing the ignored material (octree.inl:128-138).
//
// Slab (lo, hi) along a: a walk whose segment [pos, pos + T d] in the leaf,
// T = min(exit, distance left) + EPSILON <= 2.03 (sx + sy + sz) + 1e-3 (s =
// the cell's half sizes), lies above hi or below lo (heights a.x in float)
// and whose direction has |dot(a, d)| >= t + 0.0202 cannot get a distance
// in (0, T] from any of the leaf's triangles in the reference's float
// intersect_ray: every triangle lies in [lo + m, hi - m] along a, and m
// bounds how far the float hit point p + t d of intersect_ray can lie from
// the point of the triangle given by its float barycentrics (its rounding
// error analysis, with the triangles' shapes and sizes, the cell's size and
// position, and |d . n| >= 0.02 |n|, which |dot(a, d)| >= t + 0.0202
// guarantees through the cone since sin(theta + asin 0.02) <= t + 0.0201).
// lo = -inf, hi = +inf where no slab is kept.
struct __attribute__((aligned(32))) leaf_info {
    float ax, ay, az, t;       // side cone
    float lo, hi;              // slab
    uint32_t first;            // the leaf's first record
    uint32_t mats;             // its triangles' materials (as tri_rec::mats) when t < 2
};
static_assert(sizeof(leaf_info) == 32, "leaf_info is 32 bytes");

// Rope of a leaf face: entry 6 * leaf + 2 * axis + (direction component < 0)
// holds the child word of the node that the walk reaches by descending the
// neighbour's location code (next_cell, then octree.inl:58-69) from the
// root, that node's centre as the descent computes it (the same float sums
// of +-h), and the code's length; word 0 where the neighbour code leaves the
// root (the walk ends there and never reads it).
struct __attribute__((aligned(32))) leaf_rope { uint64_t word; float cx, cy, cz; uint32_t depth; };
static_assert(sizeof(leaf_rope) == 32, "leaf_rope is 32 bytes");
constexpr uint32_t NO_ROPE = 0xFFFFFFFFu;

struct dgeom {
    const uint64_t* nodes;
    const tri_rec* recs;
    const leaf_info* leaves;   // per leaf, indexed by its child word
    const leaf_rope* ropes;    // 6 per leaf
    uint32_t n_ropes;
    uint32_t owner_payload;    // 1: record indices fit 26 bits (the wavefront walk's owner search carries them)
    const uint32_t* id_rec;    // input triangle id -> index of one of its records
    float cx, cy, cz;          // root centre
    float hx, hy, hz;          // root half size
    float max_extent;          // 2*|halfsize|, octree.inl:300
    float mx, my, mz;          // (|centre| + half size) * 2^-16 per axis: more than 128 ulp
                               // of any cell centre, the leaf-box margin of descend()
    // derived once on the host (kernel arguments, so that the device does not
    // keep them in vector registers): fl(centre -+ half size), the domain
    // bounds of in_domain, and 0.5 * half size, the first level's offset
    float lox, loy, loz, hix, hiy, hiz;
    float h1x, h1y, h1z;
};

// Work counters of a walk: the loads it issues (node_steps: child words,
// leaf records, rope records) and the triangle records it tests (a leaf
// dropped by a cull costs none); the oracle counts the reference's own work.
struct prop_counters { uint32_t node_steps, tri_tests; };

// Fills the derived fields of a geometry view from its centre and half size
// (the same float operations the device would do).
NBL_HD void set_derived(dgeom& g)
{
    g.lox = g.cx - g.hx; g.loy = g.cy - g.hy; g.loz = g.cz - g.hz;
    g.hix = g.cx + g.hx; g.hiy = g.cy + g.hy; g.hiz = g.cz + g.hz;
    g.h1x = 0.5f * g.hx; g.h1y = 0.5f * g.hy; g.h1z = 0.5f * g.hz;
}

NBL_HD bool in_domain(const dgeom& g, f3 p)   // octree.inl:24-30
{
    return (p.x > g.lox) && (p.x < g.hix) && (p.y > g.loy) && (p.y < g.hiy) && (p.z > g.loz) && (p.z < g.hiz);
}

// triangle::intersect_ray, core/triangle.inl:7-28.
// The range tests on the barycentrics u = fl(u_num / det), v = fl(v_num / det)
// are decided first on estimates with the hardware reciprocal (relative
// error < 2^-21), against the thresholds widened or narrowed by a relative
// margin of 2^-12 of the estimate:
//   * clearly outside (most candidates): rejected;
//   * clearly inside: accepted without the divisions for u and v;
//   * otherwise, NaN included: the reference's divisions and tests.
// The margin exceeds the estimate's error (against the exact quotient, and so
// against the rounded u and v and their rounded sum) by a factor of 250 and
// the rounding of the shifted thresholds by far more, so every decision is
// the reference's, and the returned distance is its correctly rounded
// quotient: results stay bit-identical.
__device__ __forceinline__ float intersect_ray(f3 r0, f3 e1, f3 e2, f3 start, f3 dir)
{
    const f3 pvec = cross(dir, e2);
    const float det = dot(e1, pvec);
    if (fabsf(det) < EPSILON) return -1;
    const f3 tvec = start - r0;
    const float un = dot(tvec, pvec);
    const f3 qvec = cross(tvec, e1);
    const float vn = dot(dir, qvec);
    bool inside;
    {
        const float r = __builtin_amdgcn_rcpf(det);
        const float ua = un * r, va = vn * r;
        const float mu = fabsf(ua) * 0x1p-12f, mv = fabsf(va) * 0x1p-12f;
        if ((ua < -EPSILON - mu) || (ua > (1 + EPSILON) + mu) || (va < -EPSILON - mv) ||
            (ua + va > (1 + EPSILON) + (mu + mv)))
            return -1;
        inside = (ua > -EPSILON + mu) && (ua < (1 + EPSILON) - mu) && (va > -EPSILON + mv) &&
                 (ua + va < (1 + EPSILON) - (mu + mv));
    }
    if (!inside) {
        const float u = un / det;
        if ((u < -EPSILON) || (u > 1 + EPSILON)) return -1;
        const float v = vn / det;
        if ((v < -EPSILON) || (u + v > 1 + EPSILON)) return -1;
    }
    const float t = dot(e2, qvec) / det;
    if (t < 0) return -1;
    return t;
}

// Bit helpers of the descent, usable in host code as well (the descent is
// checked on the CPU against the reference-order descent).  clz64 needs x != 0.
NBL_HD int clz64(uint64_t x) { return __builtin_clzll(x); }
NBL_HD uint32_t as_u32(float f) { return __builtin_bit_cast(uint32_t, f); }

// One level of the descent: the child cell's centre along each axis moves by
// +-h, and h halves.  The reference's term AABB.w * halfsize * (+-0.5)
// (octree.inl:64-66, 78-80) is exactly +-h (products by powers of two), and
// it is added to the centre in the same order, so centres are bit-identical.
NBL_HD void cell_step(unsigned oct, f3& a, f3& h)
{
    a.x += (oct & 1u) ? h.x : -h.x;
    a.y += (oct & 2u) ? h.y : -h.y;
    a.z += (oct & 4u) ? h.z : -h.z;
    h.x *= 0.5f; h.y *= 0.5f; h.z *= 0.5f;
}
NBL_HD unsigned position_octant(f3 pos, f3 a)
{
    unsigned oct = 0;
    oct ^= (pos.x > a.x) ? 1u : 0u;
    oct ^= (pos.y > a.y) ? 2u : 0u;
    oct ^= (pos.z > a.z) ? 4u : 0u;
    return oct;
}

// Descent from the root to the leaf cell of the walk (octree.inl:47-84):
// along the location code first (it never passes through a leaf before its
// last digit), then by position, extending the code.  Returns the leaf's
// child word; c = leaf centre, s = leaf half size (AABB.w * halfsize).
struct no_path_cache {};
template<bool COUNT>
NBL_HD uint64_t descend(const dgeom& g, f3 pos, uint64_t& location, f3& c, f3& s,
    prop_counters* cnt, no_path_cache*, uint32_t = NO_ROPE)
{
    f3 a{ g.cx, g.cy, g.cz }, h{ g.h1x, g.h1y, g.h1z };
    uint32_t node = 0;
    uint64_t child = 0;
    bool leaf = false;
    for (int i = 60 - clz64(location); i >= 0; i -= 3) {
        const unsigned oct = (unsigned)(location >> i) & 7u;
        child = g.nodes[(size_t)node * 8 + oct];
        cell_step(oct, a, h);
        if (COUNT) ++cnt->node_steps;
        leaf = (child & LEAF_BIT) != 0;
        node = (uint32_t)child;
        if (leaf) break;
    }
    while (!leaf) {
        const unsigned oct = position_octant(pos, a);
        location = (location << 3) | oct;
        child = g.nodes[(size_t)node * 8 + oct];
        cell_step(oct, a, h);
        if (COUNT) ++cnt->node_steps;
        leaf = (child & LEAF_BIT) != 0;
        node = (uint32_t)child;
    }
    c = a;
    s = { h.x + h.x, h.y + h.y, h.z + h.z };
    return child;
}

// Memo of the lane's last descended path (the walk re-descends from the
// root for every cell, octree.inl:47-84).  Levels where the new path equals
// the memo's cost centre arithmetic only; the node where the paths part is
// read once from LDS, and only the levels below it are loaded from HBM.
// Internal nodes of levels 1..kLevels-1 are kept in LDS, the leaf word in
// registers, and the leaf's centre in LDS after the node words (the leaf-box
// test of a new walk below).  Centres are otherwise always recomputed.
struct lds_path_cache {
    static constexpr uint32_t kLevels = 12;
    static constexpr uint32_t kWords = kLevels + 3;   // per lane: node words, then the leaf centre
    uint32_t* node;      // this lane's column: node[l * 64] = internal node at depth l
    uint64_t loc;        // location code of the memo's leaf (1 = empty)
    uint64_t leaf;       // the leaf's child word
    uint32_t depth;      // the leaf's depth (0 = empty)
    NBL_HD unsigned digit(uint32_t l) const { return (unsigned)(loc >> (3 * (depth - l))) & 7u; }
    // A leaf's word is stored too: it lands on the level of the memo's own
    // leaf, which node_at never reads (it reads levels below the leaf).
    NBL_HD void store(uint32_t l, uint64_t child)
    {
        if (l < kLevels) node[l * 64] = (uint32_t)child;
    }
    NBL_HD void store_centre(f3 c)
    {
        float* f = reinterpret_cast<float*>(node + kLevels * 64);
        f[0] = c.x; f[64] = c.y; f[128] = c.z;
    }
    NBL_HD f3 centre() const
    {
        const float* f = reinterpret_cast<const float*>(node + kLevels * 64);
        return { f[0], f[64], f[128] };
    }
    // internal node at depth l < depth of the memo's path
    NBL_HD uint32_t node_at(const dgeom& g, uint32_t l) const
    {
        const uint32_t k = l < kLevels ? l : kLevels - 1;
        uint32_t n = k ? node[k * 64] : 0u;
        for (uint32_t j = k + 1; j <= l; ++j) n = (uint32_t)g.nodes[(size_t)n * 8 + digit(j)];
        return n;
    }
};
// rope entry k as two 16-B loads (k clamped to the table)
NBL_HD leaf_rope load_rope(const dgeom& g, uint32_t k)
{
    const float4* p = reinterpret_cast<const float4*>(g.ropes + min(k, g.n_ropes - 1));
    const float4 a = p[0], b = p[1];
    const uint64_t w = (uint64_t)as_u32(a.x) | ((uint64_t)as_u32(a.y) << 32);
    return leaf_rope{ w, a.z, a.w, b.x, as_u32(b.y) };
}
template<bool COUNT>
NBL_HD uint64_t descend(const dgeom& g, f3 pos, uint64_t& location, f3& c, f3& s,
    prop_counters* cnt, lds_path_cache* pc, uint32_t rope = NO_ROPE)
{
    f3 a{ g.cx, g.cy, g.cz }, h{ g.h1x, g.h1y, g.h1z };
    const uint32_t D = (uint32_t)(63 - clz64(location)) / 3;   // digits of the code
    const uint32_t M = pc->depth;
    uint32_t level = 0, node = 0;
    uint64_t child = 0;
    bool leaf = false, matched = false;
    if (rope != NO_ROPE) {
        // The next cell of a walk: the node its code leads to and that node's
        // centre come from the previous leaf's rope; the memo's levels below
        // D are that leaf's path, which the neighbour code shares up to its
        // last digit.  The position part below is shared with new walks.
        const leaf_rope r = load_rope(g, rope);
        child = r.word;
        a = { r.cx, r.cy, r.cz };
        h = { ldexpf(g.hx, -(int)(D + 1)), ldexpf(g.hy, -(int)(D + 1)), ldexpf(g.hz, -(int)(D + 1)) };
        level = D;
        if (COUNT) ++cnt->node_steps;            // the rope record (no descent loads)
        pc->store(level, child);
        node = (uint32_t)child;
        leaf = (child & LEAF_BIT) != 0;
    } else if (D == 0) {
        // A new walk (code 1).  It usually starts where the lane's last walk
        // ended, inside the memo's leaf.  The faces of a leaf (centre c, half
        // size s) are, in exact arithmetic, centre planes of its ancestors;
        // the float centres the descent compares with differ from the exact
        // ones by less than 11 ulp of the largest centre (21 levels of one
        // rounded addition each), so a point inside [c - s + m, c + s - m]
        // with m = 128 such ulp (the geometry's margin) is on the leaf's side
        // of every plane on its path, and the descent by position would
        // return this same leaf: take it without descending.
        if (M > 0) {
            const f3 cm = pc->centre();
            const float sx = ldexpf(g.hx, -(int)M), sy = ldexpf(g.hy, -(int)M), sz = ldexpf(g.hz, -(int)M);
            if (fabsf(pos.x - cm.x) < sx - g.mx && fabsf(pos.y - cm.y) < sy - g.my &&
                fabsf(pos.z - cm.z) < sz - g.mz) {
                location = pc->loc;
                c = cm;
                s = { sx, sy, sz };
                return pc->leaf;
            }
        }
        // otherwise by position from the root, along the memo while the
        // octants agree (what the general case below gives for D = 0)
        matched = true;
    } else {
        // leading digits the code shares with the memo's leaf
        uint32_t common;
        {
            const uint32_t md = min(D, M);
            const uint64_t x = (location >> (3 * (D - md))) ^ (pc->loc >> (3 * (M - md)));
            common = x == 0 ? md : md - 1 - (uint32_t)(63 - clz64(x)) / 3;
        }
        const bool same = M > 0 && common == D && D == M;
        if (!same && M > 0 && common == M) common = M - 1;   // a leaf is never inside the code's path
        const uint32_t shared = same ? D : common;
        while (level < shared) {
            ++level;
            cell_step((unsigned)(location >> (3 * (D - level))) & 7u, a, h);
        }
        if (same) {
            child = pc->leaf;
            leaf = true;
        } else if (common == D) {
            matched = true;                        // the code's node lies on the memo's path
        } else {
            node = pc->node_at(g, common);
            while (level < D) {
                ++level;
                const unsigned oct = (unsigned)(location >> (3 * (D - level))) & 7u;
                child = g.nodes[(size_t)node * 8 + oct];
                pc->store(level, child);
                cell_step(oct, a, h);
                if (COUNT) ++cnt->node_steps;
                node = (uint32_t)child;
                leaf = (child & LEAF_BIT) != 0;
                if (leaf) break;
            }
        }
    }
    const uint32_t position_from = level;
    if (matched) {
        // Along the memo while the position's octants agree with its digits:
        // centre arithmetic only.  A matched lane starts above the memo's
        // leaf (level < M) or with an empty memo (M = 0), so the loop ends
        // within M - level steps and the digit shift sh stays in [0, 3(M-1)].
        // Its code is the memo's code cut to its level, and stays so.
        uint32_t sh = 3 * (M - level);
        while (level < M) {
            const unsigned oct = position_octant(pos, a);
            sh -= 3;
            if (((unsigned)(pc->loc >> sh) & 7u) != oct) break;
            ++level;
            cell_step(oct, a, h);
        }
        if (level > position_from) location = pc->loc >> (3 * (M - level));
        if (M > 0 && level == M) { child = pc->leaf; leaf = true; }
        else node = pc->node_at(g, level);
    }
    const uint32_t loads_from = level;
    while (!leaf) {
        const unsigned oct = position_octant(pos, a);
        location = (location << 3) | oct;
        ++level;
        child = g.nodes[(size_t)node * 8 + oct];
        pc->store(level, child);
        node = (uint32_t)child;
        leaf = (child & LEAF_BIT) != 0;
        cell_step(oct, a, h);
    }
    if (COUNT) cnt->node_steps += level - loads_from;      // child words loaded
    pc->loc = location;
    pc->depth = level;
    pc->leaf = child;
    pc->store_centre(a);
    c = a;
    s = { h.x + h.x, h.y + h.y, h.z + h.z };
    return child;
}

// Exit plane of a leaf cell: AABB_intersect (octree.inl:211-220) and the
// choice of the nearest plane (octree.inl:97-112).  axis = -1, -2, -4.
__device__ __forceinline__ float leaf_exit(f3 c, f3 s, f3 pos, f3 dir, int& axis)
{
    const float tx = (c.x + copysignf(s.x + EPSILON, dir.x) - pos.x) / dir.x;
    const float ty = (c.y + copysignf(s.y + EPSILON, dir.y) - pos.y) / dir.y;
    const float tz = (c.z + copysignf(s.z + EPSILON, dir.z) - pos.z) / dir.z;
    if ((tx < ty) && (tx < tz)) { axis = -1; return tx; }
    if ((ty < tx) && (ty < tz)) { axis = -2; return ty; }
    axis = -4;
    return tz;
}

// One leaf candidate (octree.inl:117-147): skipped (-1) when it is the
// ignored triangle or its material on the ray's side is the ignored one,
// else the intersect_ray distance.
template<bool COUNT>
__device__ __forceinline__ float leaf_candidate(const tri_rec* __restrict__ recs, uint32_t k, f3 pos, f3 dir,
    uint32_t ignore_tri, int ignore_mat, uint32_t& id, prop_counters* cnt)
{
    const float4* rp = reinterpret_cast<const float4*>(recs + k);
    const float4 a = rp[0], b = rp[1], c = rp[2];
    // The record's three loads go out together: left to itself, the compiler
    // issues the id's load alone and the others only after the id test,
    // two round trips to L2 for every candidate.
    asm volatile("" ::"v"(a.x), "v"(a.y), "v"(a.z), "v"(a.w), "v"(b.x), "v"(b.y), "v"(b.z), "v"(b.w), "v"(c.x),
                 "v"(c.y), "v"(c.z));
    id = __float_as_uint(c.z);
    if (id == ignore_tri) return -1;
    const f3 r0{ a.x, a.y, a.z }, e1{ a.w, b.x, b.y }, e2{ b.z, b.w, c.x };
    const uint32_t mats = __float_as_uint(c.y);
    const int mat_in = (int)(int16_t)(mats & 0xFFFFu);
    const int mat_out = (int)(int16_t)(mats >> 16);
    const int mo = dot(cross(e1, e2), dir) < 0 ? mat_in : mat_out;
    if (mo == ignore_mat) return -1;
    if (COUNT) ++cnt->tri_tests;
    return intersect_ray(r0, e1, e2, pos, dir);
}

// Step to the neighbouring cell across the exit plane (octree.inl:165-185):
// flip the axis bit at the deepest level where that stays inside the parent.
// The reference climbs the code one digit at a time until the digit's axis
// bit equals `value` (then flips it), or the code is down to 1.  In closed
// form: the axis bits of the n digits sit at bits axis + 3k (k = 0 is the
// last digit); the lowest k whose bit allows the move is the lowest set bit
// of those bits (inverted when value is 0), and the code becomes the code
// shifted right by 3k with that bit flipped; none such: the code is 1.
NBL_HD void next_cell(uint64_t& location, int axis, f3 dir)
{
    const unsigned mask = (unsigned)(-axis);
    const unsigned a = mask == 4u ? 2u : mask >> 1;
    const bool neg = !((mask == 1u ? dir.x : (mask == 2u ? dir.y : dir.z)) >= 0);   // as the reference's >= 0 test
    const uint32_t n = location > 1 ? (uint32_t)(63 - clz64(location)) / 3 : 0u;   // digits, <= 21
    const uint64_t column = 0x1249249249249249ull & ((1ull << (3 * n)) - 1);       // bits 0, 3, .., 3(n-1)
    const uint64_t bits = location >> a;
    const uint64_t movable = (neg ? bits : ~bits) & column;
    location = movable ? (location >> __builtin_ctzll(movable)) ^ mask : 1ull;
}

// True when no triangle of the leaf (its leaf_info li, cell half sizes s)
// can take part in the walk of this lane, so that its candidates need not
// enter the shared tests; the lane's direction has |d|^2 within 1e-4 of 1,
// and either
//   * side cull: the side cone puts d clearly on one side of all the leaf's
//     normals and the material on that side is the ignored one, so the side
//     test (octree.inl:128-138) skips every triangle; or
//   * slab cull: the segment [pos, pos + T d], T = min(exit, to_go) +
//     EPSILON, lies outside the leaf's slab, so intersect_ray returns no
//     distance in (0, T] for any of them.  The walk keeps a distance only if
//     it is in (0, exit + EPSILON], and when to_go <= exit one beyond
//     to_go + EPSILON cannot change its result (the walk ends without a hit
//     whenever such a distance would be kept), so dropping the leaf leaves
//     the walk's result unchanged.
// slab_culled: the slab cull applied; then all the leaf's candidates pass
// the side test (|dot(a, d)| > t: the side is decided and is not the
// ignored one), and the reference tests each of them.
__device__ __forceinline__ bool leaf_dropped(const leaf_info& li, f3 s, f3 pos, f3 dir, float exit, float to_go,
    int ignore_mat, bool& slab_culled)
{
    slab_culled = false;
    if (!(fabsf(dot(dir, dir) - 1) < 1e-4f)) return false;
    const float c = li.ax * dir.x + li.ay * dir.y + li.az * dir.z;
    const int mat_in = (int)(int16_t)(li.mats & 0xFFFFu), mat_out = (int)(int16_t)(li.mats >> 16);
    if ((c > li.t && mat_out == ignore_mat) || (c < -li.t && mat_in == ignore_mat)) return true;
    if (!(fabsf(c) >= li.t + 0.0202f)) return false;
    const float T = fminf(exit, to_go) + EPSILON;
    if (!(T <= 2.03f * (s.x + s.y + s.z) + 1e-3f)) return false;
    const float h0 = li.ax * pos.x + li.ay * pos.y + li.az * pos.z;
    const float h1 = h0 + T * c;
    slab_culled = (h0 > li.hi && h1 > li.hi) || (h0 < li.lo && h1 < li.lo);
    return slab_culled;
}

// leaf_info k as two 16-B loads
__device__ __forceinline__ leaf_info load_leaf(const dgeom& g, uint32_t k)
{
    const float4* p = reinterpret_cast<const float4*>(g.leaves + k);
    const float4 a = p[0], b = p[1];
    return leaf_info{ a.x, a.y, a.z, a.w, b.x, b.y, __float_as_uint(b.z), __float_as_uint(b.w) };
}

// octree::propagate, geometry/octree.inl:32-192.  Same cell walk (location
// codes, re-descent from the root, neighbour by bit flip), same AABB and
// intersection arithmetic and the same "t <= best + EPSILON" replacement
// rule over each leaf's triangles in leaf order.  Returns the travelled
// distance; *hit gets the triangle's input index or NO_TRI.
template<bool COUNT, typename CACHE = no_path_cache>
__device__ __forceinline__ float propagate(const dgeom& g, f3 start, f3 dir, float distance,
    uint32_t ignore_tri, int ignore_mat, uint32_t* hit, prop_counters* cnt, CACHE* pc = nullptr)
{
    *hit = NO_TRI;
    f3 pos = start;
    float to_go = distance;
    uint64_t location = 1;
    uint32_t cells = 0;
    do {
        f3 c, s;
        const uint64_t child = descend<COUNT>(g, pos, location, c, s, cnt, pc);
        uint32_t count = (uint32_t)((child >> 32) & 0x7FFFFFFFu), first = 0;
        int axis;          // -1, -2, -4 for x, y, z; 0 once a triangle is kept
        float best = leaf_exit(c, s, pos, dir, axis);
        uint32_t best_id = NO_TRI;
        if (count > 0) {
            const leaf_info li = load_leaf(g, (uint32_t)child);
            if (COUNT) ++cnt->node_steps;            // the leaf record
            first = li.first;
            bool slab_culled;
            if (leaf_dropped(li, s, pos, dir, best, to_go, ignore_mat, slab_culled)) count = 0;
        }
        for (uint32_t k = first; k < first + count; ++k) {
            uint32_t id;
            const float t = leaf_candidate<COUNT>(g.recs, k, pos, dir, ignore_tri, ignore_mat, id, cnt);
            if ((t > 0) && (t <= best + EPSILON)) { best = t; axis = 0; best_id = id; }
        }
        if (best >= to_go) return distance;
        if (axis == 0) { *hit = best_id; return distance - to_go + best; }
        to_go -= best;
        pos = pos + dir * best;
        next_cell(location, axis, dir);
        if (++cells >= MAX_CELLS) return distance;
    } while (location > 1);
    return distance;
}

// ---------------------------------------------------------------------------
// Tables: table_1D::get, table_2D::get / get_rounddown (table_1D.inl:65-81,
// table_2D.inl:73-104).  2D storage (i, j) -> d[i*h + j], x = energy axis.
// ---------------------------------------------------------------------------
struct dtab1 { const float* d; int n; float x_min, x_step; };
// Table data lives in global memory; loads through this type are global
// loads (a generic pointer read from a descriptor would give flat loads,
// which also wait on the LDS and scalar-load counter).
typedef __attribute__((address_space(1))) const float gfloat;
__device__ __forceinline__ gfloat* gptr(const float* p) { return (gfloat*)p; }
struct dtab2 { const float* d; int w, h; float x_min, x_step, y_min, y_step; };

NBL_HD int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }
NBL_HD float clampr(float i, float lo, float hi) { return i < lo ? lo : (i > hi ? hi : i); }
// std::min / std::max argument conventions of the reference's CPU build
NBL_HD float minr(float a, float b) { return (b < a) ? b : a; }
NBL_HD float maxr(float a, float b) { return (a < b) ? b : a; }

// table_1D::get in two parts, so that independent lookups can have their
// loads in flight together: the fetch (index, the two values, the fraction)
// and the interpolation.
struct tab1_fetch { float v0, v1, f; };
__device__ __forceinline__ tab1_fetch tab1_load(const dtab1& t, float x)
{
    const float xi = (x - t.x_min) * t.x_step;
    const int lo = clampi((int)floorf(xi), 0, t.n - 2);
    return { gptr(t.d)[lo], gptr(t.d)[lo + 1], xi - (float)lo };
}
__device__ __forceinline__ float tab1_finish(const tab1_fetch& q) { return (1 - q.f) * q.v0 + q.f * q.v1; }
__device__ __forceinline__ float tab1_get(const dtab1& t, float x) { return tab1_finish(tab1_load(t, x)); }
__device__ __forceinline__ float tab2_get(const dtab2& t, float x, float y)
{
    const float xi = (x - t.x_min) * t.x_step;
    const float yi = (y - t.y_min) * t.y_step;
    const int lx = clampi((int)floorf(xi), 0, t.w - 2);
    const int ly = clampi((int)floorf(yi), 0, t.h - 2);
    gfloat* r0 = gptr(t.d) + (size_t)lx * t.h + ly;
    gfloat* r1 = r0 + t.h;
    const float v00 = r0[0], v01 = r0[1], v10 = r1[0], v11 = r1[1];
    const float fx = xi - (float)lx, fy = yi - (float)ly;
    return (1 - fx) * (1 - fy) * v00 + fx * (1 - fy) * v10 + (1 - fx) * fy * v01 + fx * fy * v11;
}
__device__ __forceinline__ float tab2_rounddown(const dtab2& t, float x, float y)
{
    const float xi = (x - t.x_min) * t.x_step;
    const float yi = (y - t.y_min) * t.y_step;
    const int lx = clampi((int)floorf(xi), 0, t.w - 1);
    const int ly = clampi((int)floorf(yi), 0, t.h - 1);
    return gptr(t.d)[(size_t)lx * t.h + ly];
}

// table_3D::get / get_rounddown (table_3D.inl:78-120); storage (i, j, k) ->
// d[(i*h + j)*dp + k], x = energy axis.
struct dtab3 { const float* d; int w, h, dp; float x_min, x_step, y_min, y_step, z_min, z_step; };

__device__ __forceinline__ float tab3_at(const dtab3& t, int i, int j, int k)
{
    return gptr(t.d)[((size_t)i * t.h + j) * t.dp + k];
}
__device__ __forceinline__ float tab3_get(const dtab3& t, float x, float y, float z)
{
    const float xi = (x - t.x_min) * t.x_step;
    const float yi = (y - t.y_min) * t.y_step;
    const float zi = (z - t.z_min) * t.z_step;
    const int lx = clampi((int)floorf(xi), 0, t.w - 2);
    const int ly = clampi((int)floorf(yi), 0, t.h - 2);
    const int lz = clampi((int)floorf(zi), 0, t.dp - 2);
    const float v000 = tab3_at(t, lx, ly, lz), v100 = tab3_at(t, lx + 1, ly, lz);
    const float v010 = tab3_at(t, lx, ly + 1, lz), v110 = tab3_at(t, lx + 1, ly + 1, lz);
    const float v001 = tab3_at(t, lx, ly, lz + 1), v101 = tab3_at(t, lx + 1, ly, lz + 1);
    const float v011 = tab3_at(t, lx, ly + 1, lz + 1), v111 = tab3_at(t, lx + 1, ly + 1, lz + 1);
    const float fx = xi - (float)lx, fy = yi - (float)ly, fz = zi - (float)lz;
    return (1 - fx) * (1 - fy) * (1 - fz) * v000
        + fx * (1 - fy) * (1 - fz) * v100
        + (1 - fx) * fy * (1 - fz) * v010
        + fx * fy * (1 - fz) * v110
        + (1 - fx) * (1 - fy) * fz * v001
        + fx * (1 - fy) * fz * v101
        + (1 - fx) * fy * fz * v011
        + fx * fy * fz * v111;
}
__device__ __forceinline__ float tab3_rounddown(const dtab3& t, float x, float y, float z)
{
    const float xi = (x - t.x_min) * t.x_step;
    const float yi = (y - t.y_min) * t.y_step;
    const float zi = (z - t.z_min) * t.z_step;
    return tab3_at(t, clampi((int)floorf(xi), 0, t.w - 1), clampi((int)floorf(yi), 0, t.h - 1),
                   clampi((int)floorf(zi), 0, t.dp - 1));
}

// Inelastic mechanism of a material set
constexpr int INEL_KIEFT = 0, INEL_PENN = 1;

// Material as held on the device: Kieft elastic plus Kieft inelastic
// (scatter_list<kieft_inelastic, kieft_elastic>) or full Penn inelastic
// (the reference's default, physics_config.h:21-24).  For full Penn,
// in_imfp and in_icdf hold its log IMFP and log omega ICDF (the same roles
// and the same table prep as Kieft's), pq its q ICDF and pion its 3D
// binding ICDF; ion is unused.
struct dmat {
    float barrier, fermi, band_gap, phonon_loss, recoil_const;
    dtab1 el_imfp, in_imfp;
    dtab2 el_icdf, in_icdf, ion;
    int model;
    dtab3 pq, pion;
};

// ---------------------------------------------------------------------------
// XORWOW (Marsaglia, J. Stat. Softw. 8(14), 2003, sec. 3.1) with the
// project's seeding: splitmix64 (Steele, Lea & Flood 2014) of the seed and
// the sequence number fills the five xorshift words and the Weyl counter.
// Uniforms: ((r >> 8) + 1) * 2^-24, in (0, 1].
// ---------------------------------------------------------------------------
struct rng_state { uint32_t x, y, z, w, v, d; };

NBL_HD uint64_t splitmix64_next(uint64_t& s)
{
    uint64_t r = (s += 0x9E3779B97F4A7C15ull);
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
}
NBL_HD rng_state rng_seed(uint64_t seed, uint64_t sequence)
{
    uint64_t s = seed ^ (sequence * 0xD1B54A32D192ED03ull);
    const uint64_t a = splitmix64_next(s), b = splitmix64_next(s), c = splitmix64_next(s);
    rng_state r{ (uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32), (uint32_t)c, (uint32_t)(c >> 32) };
    if ((r.x | r.y | r.z | r.w | r.v) == 0) r.x = 0x6C078965u;
    return r;
}
NBL_HD float rng_unit(rng_state& s)
{
    const uint32_t t = s.x ^ (s.x >> 2);
    s.x = s.y; s.y = s.z; s.z = s.w; s.w = s.v;
    s.v = (s.v ^ (s.v << 4)) ^ (t ^ (t << 1));
    s.d += 362437u;
    const uint32_t r = s.d + s.v;
    return (float)((r >> 8) + 1u) * 5.9604644775390625e-8f;
}

// A draw source: XORWOW stream, or injected uniforms for unit-parity runs.
struct rng_xorwow {
    rng_state s;
    __device__ __forceinline__ float unit() { return rng_unit(s); }
};
struct rng_inject {
    const float* u; int k;
    __device__ __forceinline__ float unit() { return k < 8 ? u[k++] : 0.5f; }
};

// log and exp where an error of a few ulp is not amplified (table
// coordinates log E, the imfp from its logarithm, the exponential draws):
// v_log_f32 / v_exp_f32 (log2, exp2: 1 ulp for normal arguments) and one
// multiplication, for normal positive inputs (energies, uniform draws
// >= 2^-24) and moderate arguments, against the 13-14 instructions of the
// device library's logf / expf.
__device__ __forceinline__ float log_fast(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945f; }
__device__ __forceinline__ float exp_fast(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896f); }

// random_generator draws, common/util/random.h:64-83.  uniform_vector takes
// cos(theta) first, then phi (the reference leaves the order to the compiler).
template<typename R> __device__ __forceinline__ float rng_phi(R& r) { return 2 * PI_F * r.unit(); }
template<typename R> __device__ __forceinline__ float rng_exponential(R& r, float tau) { return -tau * log_fast(r.unit()); }

// Square root and normalisation in the physics: one hardware instruction
// each (v_sqrt_f32, v_rsq_f32, about 1 ulp), as the reference's GPU build
// computes them (-prec-sqrt=false, rnorm3df; vector_math.h:110-158).  The
// same for reciprocals whose error no later step amplifies (v_rcp_f32,
// about 1 ulp).  The traversal, the table interpolation and the boundary
// crossing keep correctly rounded operations.
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ f3 normalised_fast(f3 a) { return a * __builtin_amdgcn_rsqf(dot(a, a)); }

// make_normal_vec (vector_math.h:171-186) for the azimuth phi = 2 pi u
// (random_generator::phi, random.h:64-67), without trigonometric range
// reduction: sin and cos of atan2(y, x) are y/r and x/r with r = |(x, y)|
// (0 and 1 for r = 0, as atan2(0, 0) = 0), and sin, cos of 2 pi u come from
// sincospi(2u).  The same vector as the reference's up to rounding.
__device__ __forceinline__ f3 normal_vec_frame(f3 dir, float sin_phi, float cos_phi)
{
    const float r = sqrt_fast(dir.x * dir.x + dir.y * dir.y);
    float cos_az = 1, sin_az = 0;
    if (r > 0) {
        const float ir = __builtin_amdgcn_rcpf(r);
        cos_az = dir.x * ir;
        sin_az = dir.y * ir;
    }
    const f3 unit_v{ dir.z * cos_az, dir.z * sin_az, -r };
    const f3 unit_u = cross(unit_v, dir);
    return unit_u * cos_phi + unit_v * sin_phi;
}
__device__ __forceinline__ f3 make_normal_vec_u(f3 dir, float u)
{
    float sin_phi, cos_phi;
    sincospif(2 * u, &sin_phi, &cos_phi);
    return normal_vec_frame(dir, sin_phi, cos_phi);
}
// The same with sin and cos of 2 pi u from the hardware (v_sin_f32 and
// v_cos_f32 take their argument in turns: one instruction each against the
// 42 of sincospif), for the Kieft scattering frame, whose outputs keep the
// physics tolerance with it.
__device__ __forceinline__ f3 make_normal_vec_turn(f3 dir, float u)
{
    return normal_vec_frame(dir, __builtin_amdgcn_sinf(u), __builtin_amdgcn_cosf(u));
}
// make_unit_vec (vector_math.h:194-206) for phi = 2 pi u.
__device__ __forceinline__ f3 make_unit_vec_u(float cos_theta, float u)
{
    float sin_phi, cos_phi;
    sincospif(2 * u, &sin_phi, &cos_phi);
    cos_theta = clampr(cos_theta, -1, 1);
    const float sin_theta = sqrt_fast(1 - cos_theta * cos_theta);
    return { sin_theta * cos_phi, sin_theta * sin_phi, cos_theta };
}
// random_generator::uniform_vector (random.h:72-75): cos(theta) first, then phi.
template<typename R> __device__ __forceinline__ f3 rng_uniform_vector(R& r)
{
    const float c = 2 * r.unit() - 1;
    const float u = r.unit();
    return make_unit_vec_u(c, u);
}

struct dparticle { f3 pos, dir; float e; };

// scatter_list::sample_path (scatter_list.h:106-112, 148-162): the elastic
// mechanism (type 2) is sampled first and wins ties; inelastic is type 1.
template<typename R>
__device__ __forceinline__ float sample_path(const dmat& m, float kin_energy, R& rng, uint8_t* type)
{
    const float le = log_fast(kin_energy);
    // both IMFP lookups' loads in flight before either is used (the
    // compiler otherwise issues the second only after the first's draw)
    const tab1_fetch q_el = tab1_load(m.el_imfp, le), q_in = tab1_load(m.in_imfp, le);
    asm volatile("" ::"v"(q_el.v0), "v"(q_el.v1), "v"(q_in.v0), "v"(q_in.v1));
    float best = __builtin_inff();
    uint8_t ty = 0;
    {
        const float imfp = exp_fast(tab1_finish(q_el));
        const float path = rng_exponential(rng, rcp_fast(imfp));
        if (path < best) { best = path; ty = 2; }
    }
    {
        const float imfp = exp_fast(tab1_finish(q_in));
        const float path = rng_exponential(rng, rcp_fast(imfp));
        if (path < best) { best = path; ty = 1; }
    }
    *type = ty;
    return best;
}

// Kieft elastic and inelastic scattering, split in three parts so that a
// wavefront with lanes of both kinds runs the shared middle part (the new
// direction frame) once:
//   scatter_begin  the ICDF table draw and everything before the frame;
//                  false when no rotation follows (inelastic phonon loss);
//   scatter_frame  p.dir normalised and the unit normal nd at a random
//                  azimuth, make_normal_vec (vector_math.h:171-186);
//   scatter_end    the rotation, energy loss and (inelastic) the secondary.
// Each lane draws its uniforms in the reference's order:
//   elastic    P(cos theta), phi
//   inelastic  P(omega'), P(binding), U(omega), phi, secondary (cos, phi).
struct scatter_state {
    float cos_theta;           // elastic
    float omega, binding;      // inelastic
    bool elastic;
};

// kieft_elastic::execute, physics/kieft/elastic.h:67-87;
// kieft_inelastic::execute, physics/kieft/inelastic.h:68-150.
// scatter_begin_v takes the ICDF value v of the first draw (cos theta or
// log omega') from the caller: the table, or the learned sampler.
template<typename R>
__device__ __forceinline__ bool scatter_begin_v(const dmat& m, dparticle& p, R& rng, bool elastic, float v,
    scatter_state& s)
{
    s.elastic = elastic;
    if (elastic) {
        s.cos_theta = clampr(v, -1, 1);
        return true;
    }
    const float omega0 = expf(v);
    float binding = tab2_rounddown(m.ion, logf(omega0), rng.unit());
    float omega;
    {
        const float omega_max = 0.5f * (p.e + omega0 - m.fermi);
        float omega_min = omega0;
        float w0 = minr(omega0 - 1, maxr(0, binding) - m.fermi);
        if (p.e > 2 * omega0) {
            omega_min = 0.5f * (p.e + omega0 - sqrtf(p.e * (p.e - 2 * omega0)));
            w0 = omega0;
        }
        const float U = rng.unit();
        if ((w0 > 0) && (omega_min > w0) && (omega_min < omega_max)) {
            // the quotients inside the logs stay correctly rounded: near 1,
            // the log amplifies their error by omega_max / w0
            const float iw0 = rcp_fast(w0);
            const float f_min = iw0 * logf((omega_min - w0) / omega_min);
            const float f_max = iw0 * logf((omega_max - w0) / omega_max);
            omega = -w0 * rcp_fast(expm1f(w0 * (f_min * (1 - U) + f_max * U)));
        } else {
            omega = omega0 * omega_max * rcp_fast(omega0 * (1 - U) + omega_max * U);
        }
    }
    if (binding < 0) {
        if (m.band_gap < 0) {
            // metals: nothing (inelastic.h:132-134)
        } else if (omega0 > m.band_gap) {
            binding = m.band_gap;
        } else {
            p.e -= omega0;               // optical phonon loss, inelastic.h:140-149
            return false;
        }
    }
    s.omega = omega;
    s.binding = maxr(0, binding);
    return true;
}
template<typename R>
__device__ __forceinline__ bool scatter_begin(const dmat& m, dparticle& p, R& rng, bool elastic, scatter_state& s)
{
    const float y = rng.unit();
    const float v = tab2_get(elastic ? m.el_icdf : m.in_icdf, log_fast(p.e), y);
    return scatter_begin_v(m, p, rng, elastic, v, s);
}

template<typename R>
__device__ __forceinline__ f3 scatter_frame(dparticle& p, R& rng)
{
    p.dir = normalised_fast(p.dir);
    return normalised_fast(make_normal_vec_turn(p.dir, rng.unit()));
}

// elastic.h:88-115 (phonon + recoil loss on); inelastic.h:151-206.
// Returns true and fills *sec when a secondary electron is created.
template<typename R>
__device__ __forceinline__ bool scatter_end(const dmat& m, dparticle& p, R& rng, const scatter_state& s, f3 nd,
    dparticle* sec)
{
    if (s.elastic) {
        const float cos_theta = s.cos_theta;
        const float sin_theta = sqrt_fast(1 - cos_theta * cos_theta);
        p.dir = p.dir * cos_theta + nd * sin_theta;
        if (p.e < 200) p.e -= minr(50e-3f, m.phonon_loss);
        else p.e -= m.recoil_const * (1 - cos_theta) * p.e;
        return false;
    }
    const float binding = s.binding, omega = s.omega;
    const float K = p.e - m.fermi + 2 * binding;
    const float dK = binding + omega;
    const float cos_theta = sqrt_fast(dK * rcp_fast(K));
    const float sin_theta = sqrt_fast(1 - clampr(cos_theta * cos_theta, 0, 1));
    f3 sdir = p.dir * cos_theta + nd * sin_theta;
    sdir = normalised_fast(sdir);
    const float sb = sqrt_fast(binding * rcp_fast(dK));
    sdir = sdir + sb * rng_uniform_vector(rng);
    sdir = normalised_fast(sdir);
    sec->e = m.fermi + omega - binding;
    sec->pos = p.pos;
    sec->dir = sdir;
    p.e -= omega;
    p.dir = p.dir - cos_theta * sdir;
    return true;
}

// kieft_elastic::execute, physics/kieft/elastic.h:67-115
template<typename R>
__device__ __forceinline__ void elastic_execute(const dmat& m, dparticle& p, R& rng)
{
    scatter_state s;
    scatter_begin(m, p, rng, true, s);
    const f3 nd = scatter_frame(p, rng);
    scatter_end(m, p, rng, s, nd, nullptr);
}

// kieft_inelastic::execute, physics/kieft/inelastic.h:68-206 (all options on).
// Returns true and fills *sec when a secondary electron is created.
template<typename R>
__device__ __forceinline__ bool inelastic_execute(const dmat& m, dparticle& p, R& rng, dparticle* sec)
{
    scatter_state s;
    if (!scatter_begin(m, p, rng, false, s)) return false;
    const f3 nd = scatter_frame(p, rng);
    return scatter_end(m, p, rng, s, nd, sec);
}

// Rejection sampling in full Penn's plasmon decay stops after this many
// proposals (the reference loops until acceptance; an accepted proposal
// comes with probability > 0.6 per round for random uniforms, so the cap is
// never reached by a random stream; it bounds the loop for fixed, injected
// uniforms).  The last proposal is kept.
constexpr int PENN_MAX_PROPOSALS = 1000;

// full_penn::execute (physics/full_penn.h:66-288), secondaries on.
// Returns true and fills *sec when a secondary electron is created.
template<typename R>
__device__ __forceinline__ bool penn_execute(const dmat& m, dparticle& p, R& rng, dparticle* sec)
{
    float omega;
    {
        const float x = logf(p.e);
        const float y = rng.unit();
        omega = expf(tab2_get(m.in_icdf, x, y));
    }
    float binding;
    {
        const float x = logf(p.e);
        const float y = logf(omega / p.e);
        const float z = rng.unit();
        binding = tab3_rounddown(m.pion, x, y, z);
    }
    p.dir = normalised_fast(p.dir);
    if (binding <= 0) {
        if (!(omega > m.band_gap)) {        // sub-band-gap loss (optical phonon)
            p.e -= omega;
            return false;
        }
        // Fermi-sea excitation
        float q;
        {
            const float x = logf(p.e);
            const float y = omega / p.e;
            const float z = rng.unit();
            q = tab3_get(m.pq, x, y, z);
        }
        const f3 prim_normal_dir = normalised_fast(make_normal_vec_u(p.dir, rng.unit()));
        f3 q_dir, q_normal_dir;
        {
            // correctly rounded: c near 1 amplifies its error in sqrt(1 - c^2)
            const float c = clampr((omega + q * q) / (2 * sqrtf(p.e) * q), -1, 1);
            const float s = sqrt_fast(1 - c * c);
            q_dir = normalised_fast(p.dir * c + prim_normal_dir * s);
            q_normal_dir = normalised_fast(make_normal_vec_u(q_dir, rng.unit()));
        }
        {
            const float c = clampr((2 * p.e - omega - q * q) / (2 * sqrtf(p.e * (p.e - omega))), -1, 1);
            const float s = sqrt_fast(1 - c * c);
            p.e -= omega;
            p.dir = p.dir * c + prim_normal_dir * s;
        }
        const float k_fermi = sqrt_fast(m.fermi);
        const float k_fermi_omega = sqrt_fast(m.fermi + omega);
        if (q < k_fermi_omega - k_fermi) {
            // plasmon decay: x = E_i / omega with p(x) ~ sqrt(x (x + 1)) on [max(0, f-1), f]
            const float f = m.fermi / omega;
            float x;
            if (f < 0.1f) {
                const float U = cbrtf(rng.unit());
                x = f * U * U;
            } else {
                const float lower = maxr(0, f - 1);
                const float upper = f;
                const float d1 = upper - lower;
                const float d2 = upper * upper - lower * lower;
                const float alpha = d2 / (d1 + d2);
                int tries = 0;
                do {
                    const float U1 = rng.unit();
                    const float U2 = rng.unit();
                    x = (U1 < alpha) ? sqrtf(d2 * U2 + lower * lower) : d1 * U2 + lower;
                } while (rng.unit() > sqrtf(x * x + x) / (x + 0.5f) && ++tries < PENN_MAX_PROPOSALS);
            }
            sec->e = x * omega + omega;
            sec->dir = rng_uniform_vector(rng);
        } else if (q < k_fermi_omega + k_fermi) {
            // single excitation from the Fermi sphere
            const float ki_parr = (omega - q * q) / (2 * q);
            const float kf_parr = ki_parr + q;
            const float ki_perp_min_sq = maxr(0, m.fermi - kf_parr * kf_parr);
            const float ki_perp_max_sq = m.fermi - ki_parr * ki_parr;
            const float ki_perp_sq = maxr(0, ki_perp_min_sq + (ki_perp_max_sq - ki_perp_min_sq) * rng.unit());
            const f3 ki = ki_parr * q_dir + sqrt_fast(ki_perp_sq) * q_normal_dir;
            sec->e = dot(ki, ki) + omega;
            sec->dir = ki + q * q_dir;
        } else {
            return false;                   // q too large for either: no secondary
        }
        sec->pos = p.pos;
        return true;
    }
    // inner-shell excitation (as Kieft, inelastic.h:151-206)
    const f3 normal_dir = normalised_fast(make_normal_vec_u(p.dir, rng.unit()));
    const float K = p.e - m.fermi + 2 * binding;
    const float dK = binding + omega;
    const float cos_theta = sqrt_fast(dK / K);
    const float sin_theta = sqrt_fast(1 - clampr(cos_theta * cos_theta, 0, 1));
    f3 sdir = p.dir * cos_theta + normal_dir * sin_theta;
    sdir = normalised_fast(sdir);
    sdir = sdir + sqrt_fast(binding / dK) * rng_uniform_vector(rng);
    sdir = normalised_fast(sdir);
    sec->e = m.fermi + omega - binding;
    sec->pos = p.pos;
    sec->dir = sdir;
    p.e -= omega;
    p.dir = p.dir - cos_theta * sdir;
    return true;
}

// The material's inelastic event: Kieft or full Penn.
template<typename R>
__device__ __forceinline__ bool inelastic_event(const dmat& m, dparticle& p, R& rng, dparticle* sec)
{
    return m.model == INEL_PENN ? penn_execute(m, p, rng, sec) : inelastic_execute(m, p, rng, sec);
}

constexpr int OUT_ALIVE = 0, OUT_DETECTED = 1, OUT_TERMINATED = 2;

// boundary_intersect<true,true,false>::execute, physics/boundary_intersect.h:39-156.
template<typename R>
__device__ __forceinline__ int intersect_execute(const dmat* mats, const tri_rec& tri, dparticle& p, int& mat, R& rng)
{
    const int t_in = (int)(int16_t)(tri.mats & 0xFFFFu), t_out = (int)(int16_t)(tri.mats >> 16);
    // both sides' barriers are read first, so that their loads overlap the
    // normal's arithmetic (the compiler would otherwise issue them after the
    // special-material cases, which need the normal)
    float bar_in = 0, bar_out = 0;
    if (t_in >= 0) bar_in = mats[t_in].barrier;
    if (t_out >= 0) bar_out = mats[t_out].barrier;
    const f3 nd = normalised(p.dir);
    const f3 e1{ tri.e1x, tri.e1y, tri.e1z }, e2{ tri.e2x, tri.e2y, tri.e2z };
    const f3 n = normalised(cross(e1, e2));
    const float cos_theta = dot(n, nd);
    asm volatile("" ::"v"(bar_in), "v"(bar_out));
    int mat_in, mat_out;
    float b_in, b_out;
    if (cos_theta > 0) { mat_in = t_in; mat_out = t_out; b_in = bar_in; b_out = bar_out; }
    else { mat_in = t_out; mat_out = t_in; b_in = bar_out; b_out = bar_in; }
    switch (mat_out) {
    case MAT_DETECTOR: return OUT_DETECTED;
    case MAT_DETECTOR_LT50: return p.e < 50 ? OUT_DETECTED : OUT_ALIVE;
    case MAT_DETECTOR_GE50: return p.e >= 50 ? OUT_DETECTED : OUT_ALIVE;
    case MAT_TERMINATOR: return OUT_TERMINATED;
    case MAT_MIRROR: p.dir = nd - 2 * n * cos_theta; return OUT_ALIVE;
    case MAT_NOP: return OUT_ALIVE;
    default: break;
    }
    float dU = 0;
    if (mat_in >= 0) dU -= b_in;
    if (mat_out >= 0) dU += b_out;
    if (p.e * cos_theta * cos_theta + dU > 0) {
        const float s = sqrtf(1 + dU / (p.e * cos_theta * cos_theta));
        const float T = 4 * s / ((1 + s) * (1 + s));
        if (rng.unit() < T) {
            p.dir = (nd - n * cos_theta) + s * n * cos_theta;
            p.e += dU;
            mat = mat_out;
            return OUT_ALIVE;
        }
    }
    p.dir = nd - 2 * n * cos_theta;
    return OUT_ALIVE;
}

}  // namespace nbl
