/*
 * nbl_oracle.cpp -- CPU oracle for the Nebula electron-transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain C++ restatement of the reference's
 * algorithm, used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker and the timed CPU baseline.  The HIP
 * library never links, loads or calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the reference's source/ directory).  Arithmetic is float32 with the
 * reference's operation order; build with -ffp-contract=off so that no
 * multiply-add is fused (the reference CPU build on x86-64 has no FMA).
 *
 * Parity pins (what checks each part) are listed in tests/test_oracle_*.py:
 *   - triangle::intersect_ray, octree builder, octree::propagate, trilist,
 *     table_1D/2D: bit-exact against the reference's own code compiled from
 *     its sources (oracle/_ref), and against committed golden fixtures.
 *   - physics (kieft elastic/inelastic, boundary_intersect), table prep,
 *     cpu driver: closed forms and invariants (energy bookkeeping,
 *     T = 4s/(1+s)^2, reflection preserves |d|, special materials).
 *   - XORWOW: Marsaglia (2003) recurrence and statistical batteries.
 */
#include "nbl_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <queue>
#include <random>
#include <thread>
#include <vector>

namespace orc {

/* ---- numeric contract (config/data_types.h:47-94, config/config.h:32) ---- */
struct vec3 { float x, y, z; };
static constexpr float EPSILON = 10 * std::numeric_limits<float>::epsilon();
static constexpr float PI = 3.1415926535897932f;

/* vector_math.h:10-105 */
static inline vec3 operator+(vec3 a, vec3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
static inline vec3 operator-(vec3 a, vec3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
static inline vec3 operator*(vec3 a, float b) { return { a.x * b, a.y * b, a.z * b }; }
static inline vec3 operator*(float a, vec3 b) { return { a * b.x, a * b.y, a * b.z }; }
static inline void operator+=(vec3& a, vec3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
static inline void operator-=(vec3& a, vec3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; }
static inline void operator*=(vec3& a, float b) { a.x *= b; a.y *= b; a.z *= b; }
static inline bool operator==(vec3 a, vec3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
static inline vec3 cross(vec3 a, vec3 b)
{
    return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
static inline float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* CPU forms of magnitude / normalised / normalise (vector_math.h:108-158) */
static inline float magnitude(vec3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline vec3 normalised(vec3 a) { return a * (1 / std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z)); }
static inline void normalise(vec3& a) { a *= 1 / std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

/* scalar_math.h:19-83 (CPU forms) */
static inline int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }
static inline float clampr(float i, float lo, float hi) { return i < lo ? lo : (i > hi ? hi : i); }
static inline int floorr(float x) { return (int)std::floor(x); }
static inline float minr(float a, float b) { return std::min(a, b); }
static inline float maxr(float a, float b) { return std::max(a, b); }
static inline float saturater(float x) { return clampr(x, 0, 1); }

/* make_normal_vec (vector_math.h:171-186) */
static inline vec3 make_normal_vec(vec3 dir, float phi)
{
    const float az = std::atan2(dir.y, dir.x);
    const float sin_az = std::sin(az), cos_az = std::cos(az);
    const vec3 unit_v{ dir.z * cos_az, dir.z * sin_az, -std::sqrt(dir.x * dir.x + dir.y * dir.y) };
    const vec3 unit_u = cross(unit_v, dir);
    const float sin_phi = std::sin(phi), cos_phi = std::cos(phi);
    return unit_u * cos_phi + unit_v * sin_phi;
}
/* make_unit_vec (vector_math.h:194-206) */
static inline vec3 make_unit_vec(float cos_theta, float phi)
{
    const float sin_phi = std::sin(phi), cos_phi = std::cos(phi);
    cos_theta = clampr(cos_theta, -1, 1);
    const float sin_theta = std::sqrt(1 - cos_theta * cos_theta);
    return { sin_theta * cos_phi, sin_theta * sin_phi, cos_theta };
}

/* ---- random numbers --------------------------------------------------- */
/* XORWOW of Marsaglia, "Xorshift RNGs", J. Stat. Softw. 8(14), 2003, sec. 3:
 * five xorshift words plus a Weyl sequence d += 362437.  Seeding is the
 * project's own: splitmix64 (Steele, Lea & Flood 2014) of seed and sequence.
 * Uniforms are ((r >> 8) + 1) * 2^-24, in (0, 1]. */
struct xorwow { uint32_t s[6]; };
static inline uint64_t splitmix64(uint64_t& z)
{
    uint64_t r = (z += 0x9E3779B97F4A7C15ull);
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
}
static inline void xorwow_seed(xorwow& st, uint64_t seed, uint64_t sequence)
{
    uint64_t z = seed ^ (sequence * 0xD1B54A32D192ED03ull);
    const uint64_t a = splitmix64(z), b = splitmix64(z), c = splitmix64(z);
    st.s[0] = (uint32_t)a; st.s[1] = (uint32_t)(a >> 32);
    st.s[2] = (uint32_t)b; st.s[3] = (uint32_t)(b >> 32);
    st.s[4] = (uint32_t)c; st.s[5] = (uint32_t)(c >> 32);
    if ((st.s[0] | st.s[1] | st.s[2] | st.s[3] | st.s[4]) == 0) st.s[0] = 0x6C078965u;
}
static inline uint32_t xorwow_next(xorwow& st)
{
    uint32_t t = st.s[0] ^ (st.s[0] >> 2);
    st.s[0] = st.s[1]; st.s[1] = st.s[2]; st.s[2] = st.s[3]; st.s[3] = st.s[4];
    st.s[4] = (st.s[4] ^ (st.s[4] << 4)) ^ (t ^ (t << 1));
    st.s[5] += 362437u;
    return st.s[5] + st.s[4];
}
static inline float u32_to_unit(uint32_t r) { return (float)((r >> 8) + 1u) * 5.9604644775390625e-8f; }

/* random_generator<gpu_flag> draws (common/util/random.h:56-83).  The order of
 * the two draws inside uniform_vector() is unspecified in C++ (argument
 * evaluation order); both sides of this project take cos(theta) first. */
template<typename base>
struct rng_t : base {
    using base::base;
    float phi() { return 2 * PI * this->unit(); }
    vec3 uniform_vector() { const float c = 2 * this->unit() - 1; const float p = phi(); return make_unit_vec(c, p); }
    float exponential(float tau) { return -tau * std::log(this->unit()); }
};
struct xorwow_base {
    xorwow st;
    xorwow_base() = default;
    explicit xorwow_base(const xorwow& s) : st(s) {}
    float unit() { return u32_to_unit(xorwow_next(st)); }
};
/* random_state<false> (random.inl:7-30): mt19937 + generate_canonical<float,24>. */
struct mt_base {
    std::mt19937 gen;
    explicit mt_base(uint32_t seed) : gen(seed) {}
    float unit() { return std::generate_canonical<float, std::numeric_limits<float>::digits>(gen); }
};
/* Injected uniforms: the j-th draw returns u[j]. */
struct stream_base {
    const float* u; int k = 0;
    explicit stream_base(const float* p) : u(p) {}
    float unit() { return k < NBL_NUM_UNIFORMS ? u[k++] : 0.5f; }
};
using rng_xorwow = rng_t<xorwow_base>;
using rng_mt = rng_t<mt_base>;
using rng_stream = rng_t<stream_base>;

/* ---- triangle (core/triangle.h:8-77, core/triangle.inl:1-103) ------------ */
struct triangle {
    int material_in, material_out;
    vec3 r0, e1, e2;
    vec3 r1() const { return r0 + e1; }
    vec3 r2() const { return r0 + e2; }
    vec3 normal() const { return cross(e1, e2); }
    vec3 aabb_min() const
    {
        const vec3 A = r0, B = r1(), C = r2();
        return { minr(minr(A.x, B.x), C.x), minr(minr(A.y, B.y), C.y), minr(minr(A.z, B.z), C.z) };
    }
    vec3 aabb_max() const
    {
        const vec3 A = r0, B = r1(), C = r2();
        return { maxr(maxr(A.x, B.x), C.x), maxr(maxr(A.y, B.y), C.y), maxr(maxr(A.z, B.z), C.z) };
    }
    /* triangle.inl:7-28, Moller-Trumbore with EPSILON guards */
    float intersect_ray(vec3 start, vec3 dir) const
    {
        const vec3 pvec = cross(dir, e2);
        const float det = dot(e1, pvec);
        if (std::abs(det) < EPSILON) return -1;
        const float u = dot(start - r0, pvec) / det;
        if ((u < -EPSILON) || (u > 1 + EPSILON)) return -1;
        const vec3 qvec = cross(start - r0, e1);
        const float v = dot(dir, qvec) / det;
        if ((v < -EPSILON) || (u + v > 1 + EPSILON)) return -1;
        const float t = dot(e2, qvec) / det;
        if (t < 0) return -1;
        return t;
    }
    /* triangle.inl:85-103 */
    bool perfect_overlap(const triangle& o) const
    {
        const vec3 a0 = r0, a1 = r1(), a2 = r2(), b0 = o.r0, b1 = o.r1(), b2 = o.r2();
        return (a0 == b0 && a1 == b1 && a2 == b2) || (a0 == b1 && a1 == b2 && a2 == b0) ||
               (a0 == b2 && a1 == b0 && a2 == b1) || (a0 == b0 && a1 == b2 && a2 == b1) ||
               (a0 == b1 && a1 == b0 && a2 == b2) || (a0 == b2 && a1 == b1 && a2 == b0);
    }
};
static inline triangle from_nbl(const nbl_triangle& t)
{
    return { t.material_in, t.material_out, { t.r0.x, t.r0.y, t.r0.z },
             { t.e1.x, t.e1.y, t.e1.z }, { t.e2.x, t.e2.y, t.e2.z } };
}

/* ---- triangle/box overlap: separating-axis test of Akenine-Moller (2001),
 * evaluated in double with the projections, vertex pairs and comparisons of
 * geometry/octree/tribox.cc:38-184 (the decisions on borderline cases must
 * be identical for identical leaf sets). ------------------------------------ */
static bool tri_box_overlap(const double c[3], const double h[3], const double tv[3][3])
{
    double v[3][3], e[3][3];
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) v[k][a] = tv[k][a] - c[a];
    for (int a = 0; a < 3; ++a) {
        e[0][a] = v[1][a] - v[0][a];
        e[1][a] = v[2][a] - v[1][a];
        e[2][a] = v[0][a] - v[2][a];
    }
    /* Axis-aligned slabs (tribox.cc:141-151). */
    for (int a = 0; a < 3; ++a) {
        double mn = v[0][a], mx = v[0][a];
        if (v[1][a] < mn) mn = v[1][a];
        if (v[1][a] > mx) mx = v[1][a];
        if (v[2][a] < mn) mn = v[2][a];
        if (v[2][a] > mx) mx = v[2][a];
        if (mn > h[a] || mx < -h[a]) return false;
    }
    /* Plane of the triangle (tribox.cc:38-60, 156-158). */
    {
        const double n[3] = { e[0][1] * e[1][2] - e[0][2] * e[1][1],
                              e[0][2] * e[1][0] - e[0][0] * e[1][2],
                              e[0][0] * e[1][1] - e[0][1] * e[1][0] };
        double vmin[3], vmax[3];
        for (int q = 0; q < 3; ++q) {
            if (n[q] > 0) { vmin[q] = -h[q] - v[0][q]; vmax[q] = h[q] - v[0][q]; }
            else          { vmin[q] = h[q] - v[0][q];  vmax[q] = -h[q] - v[0][q]; }
        }
        if (n[0] * vmin[0] + n[1] * vmin[1] + n[2] * vmin[2] > 0) return false;
        if (!(n[0] * vmax[0] + n[1] * vmax[1] + n[2] * vmax[2] >= 0)) return false;
    }
    /* Nine edge x axis tests (tribox.cc:64-107, 162-181).  For edge k the
     * three projections use vertex pairs {x-axis, y-axis, z-axis}. */
    static const int pair[3][3][2] = {
        { { 0, 2 }, { 0, 2 }, { 1, 2 } },   /* e0: X01, Y02, Z12 */
        { { 0, 2 }, { 0, 2 }, { 0, 1 } },   /* e1: X01, Y02, Z0  */
        { { 0, 1 }, { 0, 1 }, { 1, 2 } } }; /* e2: X2,  Y1,  Z12 */
    for (int k = 0; k < 3; ++k) {
        const double ex = e[k][0], ey = e[k][1], ez = e[k][2];
        const double fx = std::fabs(ex), fy = std::fabs(ey), fz = std::fabs(ez);
        for (int ax = 0; ax < 3; ++ax) {
            const int i0 = pair[k][ax][0], i1 = pair[k][ax][1];
            double p0, p1, rad;
            if (ax == 0) {        /* a=ez, b=ey: p = a*v[Y] - b*v[Z] */
                p0 = ez * v[i0][1] - ey * v[i0][2];
                p1 = ez * v[i1][1] - ey * v[i1][2];
                rad = fz * h[1] + fy * h[2];
            } else if (ax == 1) { /* a=ez, b=ex: p = -a*v[X] + b*v[Z] */
                p0 = -ez * v[i0][0] + ex * v[i0][2];
                p1 = -ez * v[i1][0] + ex * v[i1][2];
                rad = fz * h[0] + fx * h[2];
            } else {              /* a=ey, b=ex: p = a*v[X] - b*v[Y] */
                p0 = ey * v[i0][0] - ex * v[i0][1];
                p1 = ey * v[i1][0] - ex * v[i1][1];
                rad = fy * h[0] + fx * h[1];
            }
            const double mn = p0 < p1 ? p0 : p1, mx = p0 < p1 ? p1 : p0;
            if (mn > rad || mx < -rad) return false;
        }
    }
    return true;
}

/* octree_node::overlap (octree_builder.cpp:30-39): box inflated by 1.01 in double. */
static bool node_overlap(const triangle& tri, vec3 cc, vec3 ch)
{
    const double center[3] = { cc.x, cc.y, cc.z };
    const double half[3] = { 1.01 * ch.x, 1.01 * ch.y, 1.01 * ch.z };
    const vec3 a = tri.r0, b = tri.r1(), c = tri.r2();
    const double tv[3][3] = { { a.x, a.y, a.z }, { b.x, b.y, b.z }, { c.x, c.y, c.z } };
    return tri_box_overlap(center, half, tv);
}
/* get_octant_center_halfsize (octree_builder.cpp:47-64) */
static inline void octant_ch(int octant, vec3 c, vec3 h, vec3& oc, vec3& oh)
{
    oh = 0.5f * h;
    oc = { c.x + (((octant & 1) << 1) - 1) * oh.x,
           c.y + (((octant & 2)) - 1) * oh.y,
           c.z + (((octant & 4) >> 1) - 1) * oh.z };
}

/* ---- octree builder (octree_builder.cpp:93-352) --------------------------- */
struct build_node {
    int depth = 0;
    int64_t first_child = -1;       /* 8 consecutive nodes, or -1 for a leaf */
    std::vector<uint32_t> tris;     /* leaf triangle indices into root list  */
};
struct builder {
    static constexpr int max_depth = 21, split_count = 16; /* octree_builder.h:157-158 */
    std::vector<triangle> root_tris;
    std::vector<build_node> nodes;  /* nodes[0] is the root */
    vec3 center{ 0, 0, 0 }, halfsize{ 0, 0, 0 };

    int64_t make_children(int depth)
    {
        const int64_t first = (int64_t)nodes.size();
        for (int o = 0; o < 8; ++o) { build_node n; n.depth = depth; nodes.push_back(std::move(n)); }
        return first;
    }
    /* octree_child::insert (octree_builder.cpp:93-126) */
    bool insert_child(int64_t ni, uint32_t tri_index, const triangle& tri, vec3 c, vec3 h)
    {
        if (!node_overlap(tri, c, h)) return false;
        if (nodes[ni].first_child < 0) {
            for (uint32_t t : nodes[ni].tris)
                if (tri.perfect_overlap(root_tris[t])) return false; /* duplicate dropped */
            nodes[ni].tris.push_back(tri_index);
            if ((int)nodes[ni].tris.size() > split_count && nodes[ni].depth < max_depth) split(ni, c, h);
        } else {
            for (int o = 0; o < 8; ++o) {
                vec3 oc, oh; octant_ch(o, c, h, oc, oh);
                insert_child(nodes[ni].first_child + o, tri_index, tri, oc, oh);
            }
        }
        return true;
    }
    /* octree_child::split (octree_builder.cpp:140-156) */
    void split(int64_t ni, vec3 c, vec3 h)
    {
        const int64_t first = make_children(nodes[ni].depth + 1);
        nodes[ni].first_child = first;
        const std::vector<uint32_t> mine = nodes[ni].tris;
        for (int o = 0; o < 8; ++o) {
            vec3 oc, oh; octant_ch(o, c, h, oc, oh);
            for (uint32_t t : mine) insert_child(first + o, t, root_tris[t], oc, oh);
        }
        std::vector<uint32_t>().swap(nodes[ni].tris);
    }
    /* octree_root::octree_root(triangles) + insert (octree_builder.cpp:171-241) */
    void build(const std::vector<triangle>& input)
    {
        nodes.clear(); root_tris.clear();
        nodes.push_back(build_node{});
        if (input.empty()) return;
        vec3 mn = input[0].aabb_min(), mx = input[0].aabb_max();
        for (const triangle& t : input) {
            const vec3 a = t.aabb_min(), b = t.aabb_max();
            mn = { std::min(mn.x, a.x), std::min(mn.y, a.y), std::min(mn.z, a.z) };
            mx = { std::max(mx.x, b.x), std::max(mx.y, b.y), std::max(mx.z, b.z) };
        }
        center = 0.5f * (mn + mx);
        halfsize = 0.5f * (mx - mn) + vec3{ 1, 1, 1 };
        nodes[0].first_child = make_children(1);
        for (const triangle& t : input) {
            if (!node_overlap(t, center, halfsize)) continue;
            const uint32_t idx = (uint32_t)root_tris.size();
            root_tris.push_back(t);
            for (int o = 0; o < 8; ++o) {
                vec3 oc, oh; octant_ch(o, center, halfsize, oc, oh);
                insert_child(nodes[0].first_child + o, idx, t, oc, oh);
            }
        }
    }
};

/* linearized_octree (octree_builder.cpp:252-352) plus the map from
 * triangle_data positions back to input indices. */
struct geometry {
    std::vector<triangle> input;        /* as given to create()               */
    std::vector<int32_t> octree_data;   /* reference layout                  */
    std::vector<triangle> triangle_data;/* Morton-ordered                    */
    std::vector<uint32_t> tri_order;    /* triangle_data[k] == input[tri_order[k]] */
    vec3 center{ 0, 0, 0 }, halfsize{ 0, 0, 0 };
    float max_extent = 0;
    uint32_t n_nodes = 0, n_leaves = 0, max_depth = 0;

    void create(const std::vector<triangle>& tris)
    {
        input = tris;
        builder b; b.build(tris);
        center = b.center; halfsize = b.halfsize;
        max_extent = 2 * magnitude(halfsize);          /* octree.inl:263 */
        /* BFS == Morton order (octree_builder.cpp:259-276) */
        std::vector<int64_t> order;
        {
            std::queue<int64_t> q; q.push(0);
            while (!q.empty()) {
                const int64_t n = q.front(); q.pop();
                order.push_back(n);
                if (b.nodes[n].first_child >= 0)
                    for (int o = 0; o < 8; ++o) q.push(b.nodes[n].first_child + o);
            }
        }
        /* triangle order by first leaf visit (octree_builder.cpp:283-308) */
        std::vector<int32_t> tmap(b.root_tris.size(), -1);
        for (int64_t n : order) {
            if (b.nodes[n].first_child >= 0 || n == 0) continue;
            for (uint32_t t : b.nodes[n].tris) {
                if (tmap[t] >= 0) continue;
                tmap[t] = (int32_t)triangle_data.size();
                triangle_data.push_back(b.root_tris[t]);
                tri_order.push_back(t);  /* root index == input index (root never rejects) */
            }
        }
        /* index table (octree_builder.cpp:311-347) */
        std::vector<int64_t> offset(b.nodes.size(), -1);
        for (int64_t n : order) {
            offset[n] = (int64_t)octree_data.size();
            const build_node& bn = b.nodes[n];
            if (bn.first_child < 0 && n != 0) {
                ++n_leaves;
                for (uint32_t t : bn.tris) octree_data.push_back(tmap[t]);
                octree_data.push_back(-1);
            } else {
                ++n_nodes;
                for (int o = 0; o < 8; ++o) octree_data.push_back(0);
            }
            if ((uint32_t)bn.depth > max_depth) max_depth = (uint32_t)bn.depth;
        }
        for (int64_t n : order) {
            const build_node& bn = b.nodes[n];
            if (bn.first_child < 0) continue;
            for (int o = 0; o < 8; ++o) {
                const int64_t c = bn.first_child + o;
                int32_t v = (int32_t)offset[c];
                if (b.nodes[c].first_child < 0) v = -v;
                octree_data[offset[n] + o] = v;
            }
        }
    }

    /* octree::in_domain (octree.inl:24-30) */
    bool in_domain(vec3 p) const
    {
        return (p.x > center.x - halfsize.x) && (p.x < center.x + halfsize.x) &&
               (p.y > center.y - halfsize.y) && (p.y < center.y + halfsize.y) &&
               (p.z > center.z - halfsize.z) && (p.z < center.z + halfsize.z);
    }

    /* octree::propagate (octree.inl:32-192).  Returns the hit triangle as an
     * input index (NBL_NO_TRIANGLE for none).  The cell walk is capped at
     * max_cells visits, a bound never reached by a terminating walk. */
    static constexpr uint32_t max_cells = 1u << 20;
    float propagate(vec3 start, vec3 direction, float distance, uint32_t ignore_tri,
        int ignore_material, uint32_t& hit, uint32_t* node_steps, uint32_t* tri_tests) const
    {
        hit = NBL_NO_TRIANGLE;
        if (octree_data.empty()) return distance;
        vec3 pos = start;
        float distance_to_go = distance;
        uint64_t location = 1;
        uint32_t cells = 0;
        do {
            float ax = center.x, ay = center.y, az = center.z, aw = 1;
            int64_t index = 0;
            for (int i = 60 - __builtin_clzll(location); i >= 0; i -= 3) {
                const unsigned octant = (unsigned)(location >> i) & 7;
                index = octree_data[index + octant];
                ax += aw * halfsize.x * ((octant & 1) - 0.5f);
                ay += aw * halfsize.y * (((octant & 2) >> 1) - 0.5f);
                az += aw * halfsize.z * (((octant & 4) >> 2) - 0.5f);
                aw *= 0.5f;
                if (node_steps) ++*node_steps;
            }
            while (index >= 0) {
                unsigned octant = 0;
                octant ^= (pos.x > ax) ? 1 : 0;
                octant ^= (pos.y > ay) ? 2 : 0;
                octant ^= (pos.z > az) ? 4 : 0;
                location = (location << 3) | octant;
                index = octree_data[index + octant];
                ax += aw * halfsize.x * ((octant & 1) - 0.5f);
                ay += aw * halfsize.y * (((octant & 2) >> 1) - 0.5f);
                az += aw * halfsize.z * (((octant & 4) >> 2) - 0.5f);
                aw *= 0.5f;
                if (node_steps) ++*node_steps;
            }
            index = -index;

            /* AABB_intersect (octree.inl:211-220) */
            float intersect; int target;
            {
                const vec3 hs = aw * halfsize;
                const float tx = (ax + std::copysign(hs.x + EPSILON, direction.x) - pos.x) / direction.x;
                const float ty = (ay + std::copysign(hs.y + EPSILON, direction.y) - pos.y) / direction.y;
                const float tz = (az + std::copysign(hs.z + EPSILON, direction.z) - pos.z) / direction.z;
                if ((tx < ty) && (tx < tz)) { intersect = tx; target = -1; }
                else if ((ty < tx) && (ty < tz)) { intersect = ty; target = -2; }
                else { intersect = tz; target = -4; }
            }
            while (true) {
                const int32_t tidx = octree_data[index++];
                if (node_steps) ++*node_steps;
                if (tidx < 0) break;
                if (tri_order[tidx] == ignore_tri) continue;
                const triangle& tri = triangle_data[tidx];
                const int mat_out = dot(tri.normal(), direction) < 0 ? tri.material_in : tri.material_out;
                if (mat_out == ignore_material) continue;
                if (tri_tests) ++*tri_tests;
                const float t = tri.intersect_ray(pos, direction);
                if ((t > 0) && (t <= intersect + EPSILON)) { intersect = t; target = tidx; }
            }
            if (intersect >= distance_to_go) return distance;
            if (target >= 0) { hit = tri_order[target]; return distance - distance_to_go + intersect; }
            distance_to_go -= intersect;
            pos += direction * intersect;
            const unsigned mask = (unsigned)(-target);
            unsigned value;
            if (mask == 1) value = (direction.x >= 0) ? 0 : 1;
            else if (mask == 2) value = (direction.y >= 0) ? 0 : 2;
            else value = (direction.z >= 0) ? 0 : 4;
            while (location > 1) {
                if ((location & mask) == value) { location ^= mask; break; }
                location >>= 3;
            }
            if (++cells >= max_cells) return distance;
        } while (location > 1);
        return distance;
    }
};

/* trilist::propagate (geometry/trilist.inl:52-85) */
static float trilist_propagate(const std::vector<triangle>& tris, vec3 start, vec3 dir,
    float distance, uint32_t ignore_tri, int ignore_material, uint32_t& hit)
{
    float best = distance; hit = NBL_NO_TRIANGLE;
    for (uint32_t i = 0; i < (uint32_t)tris.size(); ++i) {
        if (i == ignore_tri) continue;
        const triangle& tri = tris[i];
        const int mat_out = dot(tri.normal(), dir) < 0 ? tri.material_in : tri.material_out;
        if (mat_out == ignore_material) continue;
        const float t = tri.intersect_ray(start, dir);
        if (t > 0 && t < best) { best = t; hit = i; }
    }
    return best;
}

/* ---- tables (common/util/table_1D.inl, table_2D.inl) --------------------- */
struct table1d {
    int n = 0; float x_min = 0, x_step = 0; std::vector<float> d;
    void create(float xmin, float xmax, int n_, const float* data)
    {
        n = n_; d.assign(data, data + n); x_min = xmin;
        x_step = (n - 1) / (xmax - xmin);              /* table_1D.inl:14-15 */
    }
    float get(float x) const                            /* table_1D.inl:65-81 */
    {
        const float x_index = (x - x_min) * x_step;
        const int low = clampi(floorr(x_index), 0, n - 2);
        const float lo = d[low], hi = d[low + 1];
        const float frac = x_index - low;
        return (1 - frac) * lo + frac * hi;
    }
    float scalemax() const { return x_min + (n - 1) / x_step; } /* table_1D.inl:102-106 */
};
struct table2d {
    int w = 0, h = 0; float x_min = 0, x_step = 0, y_min = 0, y_step = 0; std::vector<float> d;
    void create(float xmin, float xmax, int w_, float ymin, float ymax, int h_, const float* data)
    {
        w = w_; h = h_; d.assign(data, data + (size_t)w * h);
        x_min = xmin; x_step = (w - 1) / (xmax - xmin);  /* table_2D.inl:16-19 */
        y_min = ymin; y_step = (h - 1) / (ymax - ymin);
    }
    float at(int i, int j) const { return d[(size_t)i * h + j]; }
    float get(float x, float y) const                   /* table_2D.inl:73-94 */
    {
        const float x_index = (x - x_min) * x_step;
        const float y_index = (y - y_min) * y_step;
        const int low_x = clampi(floorr(x_index), 0, w - 2);
        const int low_y = clampi(floorr(y_index), 0, h - 2);
        const float v00 = at(low_x, low_y), v10 = at(low_x + 1, low_y);
        const float v01 = at(low_x, low_y + 1), v11 = at(low_x + 1, low_y + 1);
        const float fx = x_index - low_x, fy = y_index - low_y;
        return (1 - fx) * (1 - fy) * v00 + fx * (1 - fy) * v10 + (1 - fx) * fy * v01 + fx * fy * v11;
    }
    float rounddown(float x, float y) const             /* table_2D.inl:96-104 */
    {
        const float x_index = (x - x_min) * x_step;
        const float y_index = (y - y_min) * y_step;
        const int low_x = clampi(floorr(x_index), 0, w - 1);
        const int low_y = clampi(floorr(y_index), 0, h - 1);
        return at(low_x, low_y);
    }
};

/* table_3D (common/util/table_3D.inl): value(i,j,k) at (i*height + j)*depth + k. */
struct table3d {
    int w = 0, h = 0, dp = 0;
    float x_min = 0, x_step = 0, y_min = 0, y_step = 0, z_min = 0, z_step = 0;
    std::vector<float> d;
    void create(float xmin, float xmax, int w_, float ymin, float ymax, int h_, float zmin, float zmax, int dp_,
        const float* data)                              /* table_3D.inl:5-25 */
    {
        w = w_; h = h_; dp = dp_; d.assign(data, data + (size_t)w * h * dp);
        x_min = xmin; x_step = (w - 1) / (xmax - xmin);
        y_min = ymin; y_step = (h - 1) / (ymax - ymin);
        z_min = zmin; z_step = (dp - 1) / (zmax - zmin);
    }
    float at(int i, int j, int k) const { return d[((size_t)i * h + j) * dp + k]; }   /* table_3D.inl:64-74 */
    float get(float x, float y, float z) const          /* table_3D.inl:76-106 */
    {
        const float xi = (x - x_min) * x_step, yi = (y - y_min) * y_step, zi = (z - z_min) * z_step;
        const int lx = clampi(floorr(xi), 0, w - 2), ly = clampi(floorr(yi), 0, h - 2), lz = clampi(floorr(zi), 0, dp - 2);
        const float fx = xi - lx, fy = yi - ly, fz = zi - lz;
        return (1 - fx) * (1 - fy) * (1 - fz) * at(lx, ly, lz)
            + fx * (1 - fy) * (1 - fz) * at(lx + 1, ly, lz)
            + (1 - fx) * fy * (1 - fz) * at(lx, ly + 1, lz)
            + fx * fy * (1 - fz) * at(lx + 1, ly + 1, lz)
            + (1 - fx) * (1 - fy) * fz * at(lx, ly, lz + 1)
            + fx * (1 - fy) * fz * at(lx + 1, ly, lz + 1)
            + (1 - fx) * fy * fz * at(lx, ly + 1, lz + 1)
            + fx * fy * fz * at(lx + 1, ly + 1, lz + 1);
    }
    float rounddown(float x, float y, float z) const    /* table_3D.inl:108-118 */
    {
        const float xi = (x - x_min) * x_step, yi = (y - y_min) * y_step, zi = (z - z_min) * z_step;
        return at(clampi(floorr(xi), 0, w - 1), clampi(floorr(yi), 0, h - 1), clampi(floorr(zi), 0, dp - 1));
    }
};

/* ---- material (core/material.h, physics/kieft/elastic.h, inelastic.h) ----- */
/* For full Penn (physics/full_penn.h) in_imfp / in_icdf hold the log IMFP and
 * log omega ICDF, pq the q ICDF and pion the 3D binding ICDF; ion is unused. */
struct material {
    float barrier = 0, fermi = 0, band_gap = -1, phonon_loss = 0, recoil_const = 0;
    table1d el_imfp; table2d el_icdf; table1d in_imfp; table2d in_icdf; table2d ion;
    int model = NBL_INELASTIC_KIEFT; table3d pq, pion;
    /* kieft_*::get_max_energy and the energy_visitor min (scatter_list.h:135-146):
     * for_each visits elastic (index 1) first, then inelastic. */
    float max_energy() const
    {
        float m = std::numeric_limits<float>::infinity();
        const float e_el = std::exp(el_imfp.scalemax());
        if (e_el < m) m = e_el;
        const float e_in = std::exp(in_imfp.scalemax());
        if (e_in < m) m = e_in;
        return m;
    }
};
static void table1d_from(table1d& t, const nbl_table1d_desc& d) { t.create(d.x_min, d.x_max, d.n, d.data); }
static void table2d_from(table2d& t, const nbl_table2d_desc& d)
{
    t.create(d.x_min, d.x_max, d.width, d.y_min, d.y_max, d.height, d.data);
}

struct particle { vec3 pos, dir; float kin_energy; };
static inline particle from_nbl(const nbl_particle& p)
{
    return { { p.pos.x, p.pos.y, p.pos.z }, { p.dir.x, p.dir.y, p.dir.z }, p.kin_energy };
}
static inline nbl_particle to_nbl(const particle& p)
{
    return { { p.pos.x, p.pos.y, p.pos.z }, { p.dir.x, p.dir.y, p.dir.z }, p.kin_energy };
}

struct counters {
    uint64_t events = 0, lookups_1d = 0, lookups_2d = 0, rounddowns = 0, node_steps = 0,
             tri_tests = 0, boundary_events = 0, secondaries = 0;
};

/* scatter_list::sample_path (scatter_list.h:106-112, 148-162) with the
 * kieft sample_path of elastic.h:54-62 and inelastic.h:56-63.  for_each
 * visits the last mechanism first: elastic (type 2), then inelastic (type 1);
 * the minimum is kept with a strict '<' so elastic wins ties. */
template<typename R>
static inline void sample_path(const material& m, const particle& p, R& rng, uint8_t& type, float& dist, counters* c)
{
    type = 0; dist = std::numeric_limits<float>::infinity();
    {
        const float imfp = std::exp(m.el_imfp.get(std::log(p.kin_energy)));
        const float path = rng.exponential(1 / imfp);
        if (path < dist) { type = 2; dist = path; }
    }
    {
        const float imfp = std::exp(m.in_imfp.get(std::log(p.kin_energy)));
        const float path = rng.exponential(1 / imfp);
        if (path < dist) { type = 1; dist = path; }
    }
    if (c) c->lookups_1d += 2;
}

/* kieft_elastic::execute (elastic.h:67-115), acoustic phonon and recoil loss on. */
template<typename R>
static inline void elastic_execute(const material& m, particle& p, R& rng, counters* c)
{
    float cos_theta, sin_theta;
    {
        const float x = std::log(p.kin_energy);
        const float y = rng.unit();
        cos_theta = clampr(m.el_icdf.get(x, y), -1, 1);
        sin_theta = std::sqrt(1 - cos_theta * cos_theta);
    }
    normalise(p.dir);
    const vec3 normal_dir = normalised(make_normal_vec(p.dir, rng.phi()));
    p.dir = p.dir * cos_theta + normal_dir * sin_theta;
    if (p.kin_energy < 200)
        p.kin_energy -= minr(50e-3f, m.phonon_loss);
    else
        p.kin_energy -= m.recoil_const * (1 - cos_theta) * p.kin_energy;
    if (c) c->lookups_2d += 1;
}

/* kieft_inelastic::execute (inelastic.h:68-206), all four options on.
 * Returns true and fills sec if a secondary electron is created. */
template<typename R>
static inline bool inelastic_execute(const material& m, particle& p, R& rng, particle& sec, counters* c)
{
    float omega0;
    {
        const float x = std::log(p.kin_energy);
        const float y = rng.unit();
        omega0 = std::exp(m.in_icdf.get(x, y));
    }
    float binding;
    {
        const float x = std::log(omega0);
        const float y = rng.unit();
        binding = m.ion.rounddown(x, y);
    }
    if (c) { c->lookups_2d += 1; c->rounddowns += 1; }
    float omega;
    {
        const float omega_max = 0.5f * (p.kin_energy + omega0 - m.fermi);
        float omega_min = omega0;
        float w0 = minr(omega0 - 1, maxr(0, binding) - m.fermi);
        if (p.kin_energy > 2 * omega0) {
            omega_min = 0.5f * (p.kin_energy + omega0 - std::sqrt(p.kin_energy * (p.kin_energy - 2 * omega0)));
            w0 = omega0;
        }
        const float U = rng.unit();
        if ((w0 > 0) && (omega_min > w0) && (omega_min < omega_max)) {
            const float f_min = 1 / w0 * std::log((omega_min - w0) / omega_min);
            const float f_max = 1 / w0 * std::log((omega_max - w0) / omega_max);
            omega = -w0 / std::expm1(w0 * (f_min * (1 - U) + f_max * U));
        } else {
            omega = omega0 * omega_max / (omega0 * (1 - U) + omega_max * U);
        }
    }
    if (binding < 0) {
        if (m.band_gap < 0) {
            /* metals: nothing (inelastic.h:132-134) */
        } else if (omega0 > m.band_gap) {
            binding = m.band_gap;
        } else {
            p.kin_energy -= omega0;   /* optical phonon loss, exit (inelastic.h:140-149) */
            return false;
        }
    }
    binding = maxr(0, binding);
    normalise(p.dir);
    const vec3 normal_dir = normalised(make_normal_vec(p.dir, rng.phi()));
    const float _K = p.kin_energy - m.fermi + 2 * binding;
    const float dK = binding + omega;
    const float cos_theta = std::sqrt(dK / _K);
    const float sin_theta = std::sqrt(1 - saturater(cos_theta * cos_theta));
    vec3 secondary_dir = p.dir * cos_theta + normal_dir * sin_theta;
    normalise(secondary_dir);
    secondary_dir += std::sqrt(binding / dK) * rng.uniform_vector();
    normalise(secondary_dir);
    sec.kin_energy = m.fermi + omega - binding;
    sec.pos = p.pos;
    sec.dir = secondary_dir;
    p.kin_energy -= omega;
    p.dir -= cos_theta * secondary_dir;
    if (c) c->secondaries += 1;
    return true;
}

/* The reference's plasmon-decay rejection loop (full_penn.h:173-184) runs
 * until acceptance.  Both sides of this project stop after this many
 * proposals and keep the last one; with random uniforms the acceptance
 * probability per proposal exceeds 0.6, so the cap only bounds the loop for
 * fixed injected uniforms. */
static constexpr int PENN_MAX_PROPOSALS = 1000;

/* Plasmon decay's x = E_i / omega (full_penn.h:141-187): p(x) ~ sqrt(x (x + 1))
 * on [max(0, f - 1), f], f = E_F / omega; for f < 0.1, p(x) ~ sqrt(x) on [0, f]. */
template<typename R>
static inline float plasmon_x(float f, R& rng)
{
    if (f < 0.1f) {
        const float U = std::cbrt(rng.unit());
        return f * U * U;
    }
    const float lower = maxr(0, f - 1);
    const float upper = f;
    const float d1 = upper - lower;
    const float d2 = upper * upper - lower * lower;
    const float alpha = d2 / (d1 + d2);
    float x;
    int tries = 0;
    for (;;) {
        const float U1 = rng.unit();
        const float U2 = rng.unit();
        x = (U1 < alpha) ? std::sqrt(d2 * U2 + lower * lower) : d1 * U2 + lower;
        if (!(rng.unit() > std::sqrt(x * x + x) / (x + 0.5f))) break;
        if (++tries >= PENN_MAX_PROPOSALS) break;
    }
    return x;
}

/* full_penn<false, true>::execute (physics/full_penn.h:61-288).
 * Returns true and fills sec if a secondary electron is created. */
template<typename R>
static inline bool penn_execute(const material& m, particle& p, R& rng, particle& sec, counters* c)
{
    float omega;
    {   /* full_penn.h:73-79 */
        const float x = std::log(p.kin_energy);
        const float y = rng.unit();
        omega = std::exp(m.in_icdf.get(x, y));
    }
    float binding;
    {   /* full_penn.h:81-94: rounddown at (log K, log(omega/K), P) */
        const float x = std::log(p.kin_energy);
        const float y = std::log(omega / p.kin_energy);
        const float z = rng.unit();
        binding = m.pion.rounddown(x, y, z);
    }
    if (c) { c->lookups_2d += 1; c->rounddowns += 1; }
    normalise(p.dir);
    if (binding <= 0) {
        if (!(omega > m.band_gap)) {      /* full_penn.h:226-234: sub-band-gap loss */
            p.kin_energy -= omega;
            return false;
        }
        float q;                           /* full_penn.h:104-113 */
        {
            const float x = std::log(p.kin_energy);
            const float y = omega / p.kin_energy;
            const float z = rng.unit();
            q = m.pq.get(x, y, z);
        }
        const vec3 prim_normal_dir = normalised(make_normal_vec(p.dir, rng.phi()));
        vec3 q_dir, q_normal_dir;
        {   /* full_penn.h:115-123 */
            const float ct = clampr((omega + q * q) / (2 * std::sqrt(p.kin_energy) * q), -1, 1);
            const float st = std::sqrt(1 - ct * ct);
            q_dir = normalised(p.dir * ct + prim_normal_dir * st);
            q_normal_dir = normalised(make_normal_vec(q_dir, rng.phi()));
        }
        {   /* full_penn.h:126-131: deflect the primary */
            const float ct = clampr((2 * p.kin_energy - omega - q * q) /
                (2 * std::sqrt(p.kin_energy * (p.kin_energy - omega))), -1, 1);
            const float st = std::sqrt(1 - ct * ct);
            p.kin_energy -= omega;
            p.dir = p.dir * ct + prim_normal_dir * st;
        }
        const float k_fermi = std::sqrt(m.fermi);
        const float k_fermi_omega = std::sqrt(m.fermi + omega);
        if (q < k_fermi_omega - k_fermi) {
            /* plasmon decay, full_penn.h:141-187 */
            const float x = plasmon_x(m.fermi / omega, rng);
            sec.kin_energy = x * omega + omega;
            sec.dir = rng.uniform_vector();
        } else if (q < k_fermi_omega + k_fermi) {
            /* single excitation, full_penn.h:188-214 */
            const float ki_parr = (omega - q * q) / (2 * q);
            const float kf_parr = ki_parr + q;
            const float ki_perp_min_sq = maxr(0, m.fermi - kf_parr * kf_parr);
            const float ki_perp_max_sq = m.fermi - ki_parr * ki_parr;
            const float ki_perp_sq = maxr(0, ki_perp_min_sq + (ki_perp_max_sq - ki_perp_min_sq) * rng.unit());
            const vec3 ki = ki_parr * q_dir + std::sqrt(ki_perp_sq) * q_normal_dir;
            sec.kin_energy = dot(ki, ki) + omega;
            sec.dir = ki + q * q_dir;
        } else {
            return false;                  /* full_penn.h:215-221: no secondary */
        }
        sec.pos = p.pos;
        if (c) c->secondaries += 1;
        return true;
    }
    /* inner-shell excitation, full_penn.h:237-285 */
    const vec3 normal_dir = normalised(make_normal_vec(p.dir, rng.phi()));
    const float _K = p.kin_energy - m.fermi + 2 * binding;
    const float dK = binding + omega;
    const float cos_theta = std::sqrt(dK / _K);
    const float sin_theta = std::sqrt(1 - saturater(cos_theta * cos_theta));
    vec3 secondary_dir = p.dir * cos_theta + normal_dir * sin_theta;
    normalise(secondary_dir);
    secondary_dir += std::sqrt(binding / dK) * rng.uniform_vector();
    normalise(secondary_dir);
    sec.kin_energy = m.fermi + omega - binding;
    sec.pos = p.pos;
    sec.dir = secondary_dir;
    p.kin_energy -= omega;
    p.dir -= cos_theta * secondary_dir;
    if (c) c->secondaries += 1;
    return true;
}

/* The material's inelastic mechanism (physics_config.h:21-24 selects it at
 * compile time in the reference; here it is a per-set property). */
template<typename R>
static inline bool inelastic_event(const material& m, particle& p, R& rng, particle& sec, counters* c)
{
    return m.model == NBL_INELASTIC_PENN ? penn_execute(m, p, rng, sec, c) : inelastic_execute(m, p, rng, sec, c);
}

/* boundary_intersect<true,true,false>::execute (boundary_intersect.h:39-156).
 * Returns NBL_OUT_*; *mat is the particle's material (updated on transmission). */
template<typename R>
static inline int intersect_execute(const std::vector<material>& mats, const triangle& tri,
    particle& p, int& mat, R& rng, counters* c)
{
    if (c) c->boundary_events += 1;
    const vec3 normalised_dir = normalised(p.dir);
    const vec3 n = normalised(tri.normal());
    const float cos_theta = dot(n, normalised_dir);
    int mat_in, mat_out;
    if (cos_theta > 0) { mat_in = tri.material_in; mat_out = tri.material_out; }
    else { mat_in = tri.material_out; mat_out = tri.material_in; }
    switch (mat_out) {
    case NBL_MAT_DETECTOR: return NBL_OUT_DETECTED;
    case NBL_MAT_DETECTOR_LT50: return p.kin_energy < 50 ? NBL_OUT_DETECTED : NBL_OUT_ALIVE;
    case NBL_MAT_DETECTOR_GE50: return p.kin_energy >= 50 ? NBL_OUT_DETECTED : NBL_OUT_ALIVE;
    case NBL_MAT_TERMINATOR: return NBL_OUT_TERMINATED;
    case NBL_MAT_MIRROR: p.dir = normalised_dir - 2 * n * cos_theta; return NBL_OUT_ALIVE;
    case NBL_MAT_NOP: return NBL_OUT_ALIVE;
    default: break;
    }
    float dU = 0;
    if (mat_in >= 0) dU -= mats[mat_in].barrier;
    if (mat_out >= 0) dU += mats[mat_out].barrier;
    if (p.kin_energy * cos_theta * cos_theta + dU > 0) {
        const float s = std::sqrt(1 + dU / (p.kin_energy * cos_theta * cos_theta));
        const float T = 4 * s / ((1 + s) * (1 + s));
        if (rng.unit() < T) {
            p.dir = (normalised_dir - n * cos_theta) + s * n * cos_theta;
            p.kin_energy += dU;
            mat = mat_out;
            return NBL_OUT_ALIVE;
        }
    }
    p.dir = normalised_dir - 2 * n * cos_theta;
    return NBL_OUT_ALIVE;
}

/* ---- kieft table preparation (elastic.h:120-166, inelastic.h:212-327) ----- */
static int kieft_prepare(const nbl_kieft_raw& r, float* el_imfp, float* el_icdf, float* in_imfp,
    float* in_w0, float* ion, nbl_material_desc& d)
{
    if (r.n_k < 2 || r.n_p < 2 || !(r.k_min > 0) || !(r.k_max > r.k_min) || !(r.p_max > r.p_min)) return -1;
    const int nk = r.n_k, np = r.n_p;
    /* geomspace / linspace dimension scales (common/util/range.h:204-312) */
    const double kstep = std::log(r.k_max / r.k_min) / (nk - 1);
    std::vector<double> K(nk);
    for (int i = 0; i < nk; ++i) K[i] = r.k_min * std::exp(i * kstep);
    const double pstep = (r.p_max - r.p_min) / (np - 1);
    std::vector<double> P(np);
    for (int j = 0; j < np; ++j) P[j] = r.p_min + j * pstep;
    const float lk0 = (float)std::log(K.front()), lk1 = (float)std::log(K.back());
    const float p0 = (float)P.front(), p1 = (float)P.back();
    /* elastic.h:129-141: log(imfp * unit), unit = 1 nm^-1 */
    for (int i = 0; i < nk; ++i) el_imfp[i] = std::log(r.elastic_imfp[i] * 1.0f);
    /* elastic.h:144-157: clamp to [-1, 1] */
    for (size_t i = 0; i < (size_t)nk * np; ++i) el_icdf[i] = std::max(-1.f, std::min(1.f, r.elastic_costheta[i]));
    /* inelastic.h:224-237 */
    for (int i = 0; i < nk; ++i) in_imfp[i] = std::log(r.inelastic_imfp[i] * 1.0f);
    /* inelastic.h:240-263: log(max(0, min(K - fermi, w0))) in double */
    const double fermi = (double)r.fermi;
    for (int i = 0; i < nk; ++i)
        for (int j = 0; j < np; ++j)
            in_w0[(size_t)i * np + j] = (float)std::log(std::max(0.0,
                std::min<double>(K[i] - fermi, (double)(r.inelastic_w0[(size_t)i * np + j] * 1.0f))));
    /* inelastic.h:266-323: the binding table is rewritten in place, reading
     * rounddown values of the partially rewritten table. */
    std::vector<double> outer;
    for (int s = 0; s < r.n_outer_shells; ++s)
        if (r.outer_shells[s] < 100.0) outer.push_back(r.outer_shells[s]);
    std::sort(outer.begin(), outer.end());
    table2d t; t.create(lk0, lk1, nk, p0, p1, np, r.binding_icdf);
    for (int i = 0; i < nk; ++i) {
        for (int j = 0; j < np; ++j) {
            const float Pj = (float)P[j];
            double binding = -1;
            if (K[i] > 100.0) {
                binding = (double)t.rounddown((float)std::log(K[i] + 10.0), Pj) * 1.0;
                if (binding < 50.0 || !std::isfinite(binding)) binding = -1;
            }
            if (binding < 0) {
                auto it = std::lower_bound(outer.rbegin(), outer.rend(), K[i], std::greater<double>{});
                if (it != outer.rend()) binding = *it;
            }
            t.d[(size_t)i * np + j] = (float)binding;
        }
    }
    std::memcpy(ion, t.d.data(), sizeof(float) * (size_t)nk * np);
    d.barrier = r.barrier; d.fermi = r.fermi; d.band_gap = r.band_gap;
    d.phonon_loss = r.phonon_loss;
    /* elastic.h:161-163, gram = 6.2415e15 in the reference's unit system */
    d.recoil_const = 2 * (float)((2 * (9.109383e-28 * 6.2415e15)) / (r.effective_a_grams * 6.2415e15));
    d.elastic_log_imfp = { nk, lk0, lk1, el_imfp };
    d.elastic_costheta_icdf = { nk, np, lk0, lk1, p0, p1, el_icdf };
    d.inelastic_log_imfp = { nk, lk0, lk1, in_imfp };
    d.inelastic_log_w0_icdf = { nk, np, lk0, lk1, p0, p1, in_w0 };
    d.ionisation = { nk, np, lk0, lk1, p0, p1, ion };
    return 0;
}

/* ---- full Penn table preparation (physics/full_penn.h:296-399) ------------ */
/* geomspace(lo, hi, n)[i] and linspace(lo, hi, n)[i] (common/util/range.h:204-312) */
static inline double geom_at(double lo, double hi, int n, int i) { return lo * std::exp(i * (std::log(hi / lo) / (n - 1))); }
static inline double lin_at(double lo, double hi, int n, int i) { return lo + i * ((hi - lo) / (n - 1)); }

static int penn_prepare(const nbl_penn_raw& r, float* log_imfp, float* log_omega, float* q_icdf, float* ion,
    nbl_material_desc& d)
{
    if (r.n_k < 2 || r.n_p < 2 || r.q_nk < 2 || r.q_nw < 2 || r.q_np < 2 || r.i_nk < 2 || r.i_nw < 2 || r.i_np < 2)
        return -1;
    if (!(r.k_min > 0) || !(r.k_max > r.k_min) || !(r.p_max > r.p_min) || !(r.q_k_min > 0) ||
        !(r.q_k_max > r.q_k_min) || !(r.q_w_max > r.q_w_min) || !(r.q_p_max > r.q_p_min) || !(r.i_k_min > 0) ||
        !(r.i_k_max > r.i_k_min) || !(r.i_w_min > 0) || !(r.i_w_max > r.i_w_min) || !(r.i_p_max > r.i_p_min))
        return -1;
    const int nk = r.n_k, np = r.n_p;
    /* full_penn.h:304-318: log(imfp * unit), unit = 1 nm^-1, in float */
    for (int i = 0; i < nk; ++i) log_imfp[i] = std::log(r.imfp[i] * 1.0f);
    /* full_penn.h:320-341: log(max(0, min(K - fermi, omega))) in double */
    for (int i = 0; i < nk; ++i) {
        const double K = geom_at(r.k_min, r.k_max, nk, i);
        for (int j = 0; j < np; ++j)
            log_omega[(size_t)i * np + j] = (float)std::log(std::max(0.0,
                std::min<double>(K - (double)r.fermi, (double)(r.omega_icdf[(size_t)i * np + j] * 1.0f))));
    }
    /* full_penn.h:344-366: q * (hbar / sqrt(2 m) = 0.19519 nm eV^1/2), the
     * product taken in float with the unit rounded to float */
    const float q_unit = (float)(1.0 * 0.19519 * 1.0);
    for (size_t k = 0; k < (size_t)r.q_nk * r.q_nw * r.q_np; ++k) q_icdf[k] = r.q_icdf[k] * q_unit;
    /* full_penn.h:369-396: bindings below 50 eV or not finite become -1 */
    for (size_t k = 0; k < (size_t)r.i_nk * r.i_nw * r.i_np; ++k) {
        const float b = r.binding_icdf[k] * 1.0f;
        ion[k] = (b < 50 || !std::isfinite(b)) ? -1.0f : b;
    }
    d.fermi = r.fermi;
    d.band_gap = r.band_gap;
    const float lk0 = (float)std::log(geom_at(r.k_min, r.k_max, nk, 0));
    const float lk1 = (float)std::log(geom_at(r.k_min, r.k_max, nk, nk - 1));
    d.inelastic_log_imfp = { nk, lk0, lk1, log_imfp };
    d.inelastic_log_w0_icdf = { nk, np, lk0, lk1, (float)lin_at(r.p_min, r.p_max, np, 0),
                                (float)lin_at(r.p_min, r.p_max, np, np - 1), log_omega };
    d.ionisation = nbl_table2d_desc{};
    d.penn_q_icdf = { r.q_nk, r.q_nw, r.q_np,
        (float)std::log(geom_at(r.q_k_min, r.q_k_max, r.q_nk, 0)),
        (float)std::log(geom_at(r.q_k_min, r.q_k_max, r.q_nk, r.q_nk - 1)),
        (float)lin_at(r.q_w_min, r.q_w_max, r.q_nw, 0), (float)lin_at(r.q_w_min, r.q_w_max, r.q_nw, r.q_nw - 1),
        (float)lin_at(r.q_p_min, r.q_p_max, r.q_np, 0), (float)lin_at(r.q_p_min, r.q_p_max, r.q_np, r.q_np - 1),
        q_icdf };
    d.penn_ionisation = { r.i_nk, r.i_nw, r.i_np,
        (float)std::log(geom_at(r.i_k_min, r.i_k_max, r.i_nk, 0)),
        (float)std::log(geom_at(r.i_k_min, r.i_k_max, r.i_nk, r.i_nk - 1)),
        (float)std::log(geom_at(r.i_w_min, r.i_w_max, r.i_nw, 0)),
        (float)std::log(geom_at(r.i_w_min, r.i_w_max, r.i_nw, r.i_nw - 1)),
        (float)lin_at(r.i_p_min, r.i_p_max, r.i_np, 0), (float)lin_at(r.i_p_min, r.i_p_max, r.i_np, r.i_np - 1),
        ion };
    d.inelastic_model = NBL_INELASTIC_PENN;
    return 0;
}

/* ---- GPU-driver slot semantics (gpu_particle_manager.h:93-103, gpu_driver.inl) */
struct slots {
    uint32_t cap;
    std::vector<uint8_t> status;
    std::vector<uint32_t> particle_idx;
    std::vector<particle> parts;
    std::vector<uint32_t> tags;
    std::vector<int32_t> material;
    std::vector<uint32_t> last_tri;
    std::vector<xorwow> rng;
};
static inline bool slot_active(uint8_t s)   /* gpu_particle_manager.inl:271-283 */
{
    return !(s == NBL_ST_TERMINATED || s == NBL_ST_DETECTED || s == NBL_ST_PENDING);
}
/* Stable sort on status bits [0,2) (gpu_particle_manager.inl:237-246). */
static void sort_status(const uint8_t* keys, uint32_t n, uint32_t* out)
{
    uint32_t count[4] = { 0, 0, 0, 0 };
    for (uint32_t i = 0; i < n; ++i) ++count[keys[i] & 3];
    uint32_t off[4] = { 0, count[0], count[0] + count[1], count[0] + count[1] + count[2] };
    for (uint32_t i = 0; i < n; ++i) out[off[keys[i] & 3]++] = i;
}

/* ---- CPU particle manager (drivers/cpu/cpu_particle_manager.{h,inl}) ------ */
enum cpu_status { CPU_SCATTER, CPU_INTERSECT, CPU_NO_EVENT, CPU_DETECTED, CPU_TERMINATED };
struct cpu_particle {
    cpu_status status; uint8_t next_scatter; int current_material;
    particle data; uint32_t primary_tag, secondary_tag; uint32_t last_triangle;
};
struct cascade { uint32_t next_secondary_tag, running_count; };

/* Energy-deposition tally of energydeposit_cpu_driver (drivers/cpu/
 * energydeposit_cpu_driver.h:44-94): a scatter event whose particle's
 * energy changes deposits before.E - after.E at after.pos; here summed
 * into a voxel grid instead of written as records (cpu_energydep.cpp:184-192). */
struct deposit_grid {
    int nx = 0, ny = 0, nz = 0;
    vec3 lo{}, inv{};                  /* voxel index = floor((p - lo) * inv) */
    std::vector<double> e;
    double total = 0;
    uint64_t count = 0;
    void add(vec3 p, float de)
    {
        total += de; ++count;
        const int ix = (int)std::floor((p.x - lo.x) * inv.x), iy = (int)std::floor((p.y - lo.y) * inv.y),
                  iz = (int)std::floor((p.z - lo.z) * inv.z);
        if (ix < 0 || iy < 0 || iz < 0 || ix >= nx || iy >= ny || iz >= nz) return;
        e[((size_t)iz * ny + iy) * nx + ix] += de;
    }
};

template<typename R>
struct cpu_driver {
    const geometry& geom; const std::vector<material>& mats;
    float min_energy, max_energy;
    R rng;
    std::vector<cpu_particle> particles;
    std::map<uint32_t, cascade> cascades;
    counters c;
    deposit_grid* dep = nullptr;         /* energy-deposition tally, or none */

    cpu_driver(const geometry& g, const std::vector<material>& m, float mn, float mx, R r)
        : geom(g), mats(m), min_energy(mn), max_energy(mx), rng(r) {}

    void push(const particle* p, const uint32_t* tags, size_t n)   /* cpu_particle_manager.inl:22-43 */
    {
        particles.reserve(particles.size() + n);
        for (size_t i = 0; i < n; ++i) {
            particles.push_back({ CPU_NO_EVENT, 0, NBL_MAT_VACUUM, p[i], tags[i], 0, NBL_NO_TRIANGLE });
            cascades.insert(std::make_pair(tags[i], cascade{ 1, 1 }));
        }
    }
    bool active(size_t i) const
    {
        return !(particles[i].status == CPU_TERMINATED || particles[i].status == CPU_DETECTED);
    }
    void terminate(size_t i) { particles[i].status = CPU_TERMINATED; --cascades[particles[i].primary_tag].running_count; }
    void detect(size_t i) { particles[i].status = CPU_DETECTED; --cascades[particles[i].primary_tag].running_count; }

    /* cpu_driver::init (cpu_driver.inl:73-130) */
    void init(size_t idx)
    {
        c.events += 1;
        particle this_particle = particles[idx].data;
        if (!geom.in_domain(this_particle.pos)) { terminate(idx); return; }
        uint8_t type = 0; float dist = geom.max_extent;
        const int mi = particles[idx].current_material;
        if (mi >= 0) {
            const material& m = mats[mi];
            if (this_particle.kin_energy < m.barrier + min_energy || this_particle.kin_energy > m.barrier + max_energy) {
                terminate(idx); return;
            }
            sample_path(m, this_particle, rng, type, dist, &c);
        }
        normalise(this_particle.dir);
        uint32_t hit; uint32_t ns = 0, tt = 0;
        const float d = geom.propagate(this_particle.pos, this_particle.dir, dist,
            particles[idx].last_triangle, mi, hit, &ns, &tt);
        c.node_steps += ns; c.tri_tests += tt;
        cpu_particle& cp = particles[idx];
        if (hit == NBL_NO_TRIANGLE) {   /* set_scatter_event (cpu_particle_manager.inl:213-227) */
            if (type != 0) { cp.status = CPU_SCATTER; cp.next_scatter = type; }
            else cp.status = CPU_NO_EVENT;
            cp.data.pos += normalised(cp.data.dir) * dist;
        } else {                        /* set_intersect_event (cpu_particle_manager.inl:228-235) */
            cp.status = CPU_INTERSECT;
            cp.last_triangle = hit;
            cp.data.pos += normalised(cp.data.dir) * d;
        }
    }
    /* cpu_driver::intersect (cpu_driver.inl:133-143) */
    void intersect(size_t idx)
    {
        if (particles[idx].status != CPU_INTERSECT) return;
        particle p = particles[idx].data;
        int mat = particles[idx].current_material;
        const int r = intersect_execute(mats, geom.input[particles[idx].last_triangle], p, mat, rng, &c);
        if (r == NBL_OUT_DETECTED) { detect(idx); return; }
        if (r == NBL_OUT_TERMINATED) { terminate(idx); return; }
        particles[idx].data = p;
        particles[idx].current_material = mat;
    }
    /* cpu_driver::scatter (cpu_driver.inl:145-160) */
    void scatter(size_t idx)
    {
        if (particles[idx].status != CPU_SCATTER) return;
        particles[idx].last_triangle = NBL_NO_TRIANGLE;
        const material& m = mats[particles[idx].current_material];
        particle p = particles[idx].data;
        if (particles[idx].next_scatter == 2) {
            elastic_execute(m, p, rng, &c);
            particles[idx].data = p;
        } else {
            particle sec;
            const bool made = inelastic_event(m, p, rng, sec, &c);
            if (made) {   /* create_secondary (cpu_particle_manager.inl:181-196) */
                const uint32_t tag = particles[idx].primary_tag;
                const int mat = particles[idx].current_material;
                particles.push_back({ CPU_NO_EVENT, 0, mat, sec, tag, cascades[tag].next_secondary_tag++, NBL_NO_TRIANGLE });
                ++cascades[tag].running_count;
            }
            particles[idx].data = p;
        }
    }
    /* simple_cpu_driver::simulate_to_end (simple_cpu_driver.h:54-66); with a
     * deposit tally, energydeposit_cpu_driver::simulate_to_end
     * (energydeposit_cpu_driver.h:72-94): track = next_scatter after init */
    void simulate_to_end()
    {
        for (size_t idx = 0; idx < particles.size(); ++idx)
            while (active(idx)) {
                init(idx);
                const bool track = particles[idx].status == CPU_SCATTER;
                const float before = particles[idx].data.kin_energy;
                intersect(idx); scatter(idx);
                if (dep && track && before != particles[idx].data.kin_energy)
                    dep->add(particles[idx].data.pos, before - particles[idx].data.kin_energy);
            }
    }
    /* flush_detected + flush_terminated (cpu_driver.inl:59-68, cpu_particle_manager.inl:45-78) */
    template<typename F> void flush_detected(F f)
    {
        for (auto& p : particles)
            if (p.status == CPU_DETECTED) { f(p.data, p.primary_tag); p.status = CPU_TERMINATED; }
        particles.erase(std::remove_if(particles.begin(), particles.end(),
            [](const cpu_particle& x) { return x.status == CPU_TERMINATED; }), particles.end());
        for (auto i = cascades.begin(); i != cascades.end();) {
            if (i->second.running_count == 0) i = cascades.erase(i); else ++i;
        }
    }
};

} /* namespace orc */

using namespace orc;

struct orc_geometry_s { geometry g; };
struct orc_materials_s { std::vector<material> m; };
struct orc_slots_s { slots s; };

extern "C" {

orc_geometry orc_geometry_create(const nbl_triangle* tris, uint32_t n)
{
    std::vector<triangle> v(n);
    for (uint32_t i = 0; i < n; ++i) v[i] = from_nbl(tris[i]);
    orc_geometry g = new orc_geometry_s;
    g->g.create(v);
    return g;
}
void orc_geometry_destroy(orc_geometry g) { delete g; }
void orc_geometry_info(orc_geometry g, float* c, float* h, float* me, uint64_t* ls, uint32_t* nu,
    uint32_t* nn, uint32_t* nl, uint32_t* md)
{
    if (c) { c[0] = g->g.center.x; c[1] = g->g.center.y; c[2] = g->g.center.z; }
    if (h) { h[0] = g->g.halfsize.x; h[1] = g->g.halfsize.y; h[2] = g->g.halfsize.z; }
    if (me) *me = g->g.max_extent;
    if (ls) *ls = g->g.octree_data.size();
    if (nu) *nu = (uint32_t)g->g.triangle_data.size();
    if (nn) *nn = g->g.n_nodes;
    if (nl) *nl = g->g.n_leaves;
    if (md) *md = g->g.max_depth;
}
void orc_geometry_export(orc_geometry g, int32_t* od, uint32_t* to)
{
    if (od) std::memcpy(od, g->g.octree_data.data(), g->g.octree_data.size() * sizeof(int32_t));
    if (to) std::memcpy(to, g->g.tri_order.data(), g->g.tri_order.size() * sizeof(uint32_t));
}
void orc_in_domain_batch(orc_geometry g, const nbl_vec3* pos, uint32_t n, uint8_t* out)
{
    for (uint32_t i = 0; i < n; ++i) out[i] = g->g.in_domain({ pos[i].x, pos[i].y, pos[i].z }) ? 1 : 0;
}
void orc_propagate_batch(orc_geometry g, const nbl_vec3* pos, const nbl_vec3* dir, const float* dist,
    const uint32_t* ignore_tri, const int32_t* ignore_mat, uint32_t n, float* out_dist, uint32_t* out_tri,
    uint32_t* node_steps, uint32_t* tri_tests)
{
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t hit, ns = 0, tt = 0;
        out_dist[i] = g->g.propagate({ pos[i].x, pos[i].y, pos[i].z }, { dir[i].x, dir[i].y, dir[i].z },
            dist[i], ignore_tri[i], ignore_mat[i], hit, &ns, &tt);
        out_tri[i] = hit;
        if (node_steps) node_steps[i] = ns;
        if (tri_tests) tri_tests[i] = tt;
    }
}
void orc_trilist_propagate_batch(const nbl_triangle* tris, uint32_t ntri, const nbl_vec3* pos,
    const nbl_vec3* dir, const float* dist, const uint32_t* ignore_tri, const int32_t* ignore_mat,
    uint32_t n, float* out_dist, uint32_t* out_tri)
{
    std::vector<triangle> v(ntri);
    for (uint32_t i = 0; i < ntri; ++i) v[i] = from_nbl(tris[i]);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t hit;
        out_dist[i] = trilist_propagate(v, { pos[i].x, pos[i].y, pos[i].z }, { dir[i].x, dir[i].y, dir[i].z },
            dist[i], ignore_tri[i], ignore_mat[i], hit);
        out_tri[i] = hit;
    }
}
void orc_intersect_ray_batch(const nbl_triangle* tris, const nbl_vec3* pos, const nbl_vec3* dir,
    uint32_t n, float* out_t)
{
    for (uint32_t i = 0; i < n; ++i)
        out_t[i] = from_nbl(tris[i]).intersect_ray({ pos[i].x, pos[i].y, pos[i].z }, { dir[i].x, dir[i].y, dir[i].z });
}

int orc_kieft_prepare(const nbl_kieft_raw* raw, float* a, float* b, float* c, float* d, float* e, nbl_material_desc* desc)
{
    return kieft_prepare(*raw, a, b, c, d, e, *desc);
}
orc_materials orc_materials_create(const nbl_material_desc* mats, int32_t n)
{
    orc_materials m = new orc_materials_s;
    m->m.resize(n);
    for (int32_t i = 0; i < n; ++i) {
        material& t = m->m[i];
        t.barrier = mats[i].barrier; t.fermi = mats[i].fermi; t.band_gap = mats[i].band_gap;
        t.phonon_loss = mats[i].phonon_loss; t.recoil_const = mats[i].recoil_const;
        table1d_from(t.el_imfp, mats[i].elastic_log_imfp);
        table2d_from(t.el_icdf, mats[i].elastic_costheta_icdf);
        table1d_from(t.in_imfp, mats[i].inelastic_log_imfp);
        table2d_from(t.in_icdf, mats[i].inelastic_log_w0_icdf);
        t.model = mats[i].inelastic_model;
        if (t.model == NBL_INELASTIC_PENN) {
            const nbl_table3d_desc& q = mats[i].penn_q_icdf;
            const nbl_table3d_desc& b = mats[i].penn_ionisation;
            t.pq.create(q.x_min, q.x_max, q.width, q.y_min, q.y_max, q.height, q.z_min, q.z_max, q.depth, q.data);
            t.pion.create(b.x_min, b.x_max, b.width, b.y_min, b.y_max, b.height, b.z_min, b.z_max, b.depth, b.data);
        } else {
            table2d_from(t.ion, mats[i].ionisation);
        }
    }
    return m;
}
void orc_penn_plasmon_x(float f, uint32_t n, uint64_t seed, float* out)
{
    xorwow st; xorwow_seed(st, seed, 0);
    rng_xorwow rng(st);
    for (uint32_t i = 0; i < n; ++i) out[i] = plasmon_x(f, rng);
}
int orc_penn_prepare(const nbl_penn_raw* raw, float* a, float* b, float* c, float* d, nbl_material_desc* desc)
{
    return penn_prepare(*raw, a, b, c, d, *desc);
}
void orc_table3d_get_batch(orc_materials m, int32_t mi, int32_t table, int32_t rd, const float* x,
    const float* y, const float* z, uint32_t n, float* out)
{
    const material& t = m->m[mi];
    const table3d& tab = table == NBL_TAB_PENN_Q_ICDF ? t.pq : t.pion;
    for (uint32_t i = 0; i < n; ++i) out[i] = rd ? tab.rounddown(x[i], y[i], z[i]) : tab.get(x[i], y[i], z[i]);
}
void orc_materials_destroy(orc_materials m) { delete m; }
float orc_materials_max_energy(orc_materials m)  /* cpu_material_manager.inl:42-55 */
{
    float e = std::numeric_limits<float>::infinity();
    for (const material& t : m->m) { const float x = t.max_energy(); if (x < e) e = x; }
    return e;
}
void orc_table1d_get_batch(orc_materials m, int32_t mi, int32_t table, const float* x, uint32_t n, float* out)
{
    const material& t = m->m[mi];
    const table1d& tab = table == NBL_TAB_ELASTIC_IMFP ? t.el_imfp : t.in_imfp;
    for (uint32_t i = 0; i < n; ++i) out[i] = tab.get(x[i]);
}
void orc_table2d_get_batch(orc_materials m, int32_t mi, int32_t table, int32_t rd, const float* x,
    const float* y, uint32_t n, float* out)
{
    const material& t = m->m[mi];
    const table2d& tab = table == NBL_TAB_ELASTIC_ICDF ? t.el_icdf : table == NBL_TAB_INELASTIC_W0_ICDF ? t.in_icdf : t.ion;
    for (uint32_t i = 0; i < n; ++i) out[i] = rd ? tab.rounddown(x[i], y[i]) : tab.get(x[i], y[i]);
}
int orc_physics_batch(orc_materials m, orc_geometry g, int32_t event, const int32_t* material_idx,
    const uint32_t* last_tri, const float* uniforms, uint32_t n, nbl_particle* io, int32_t* out_material,
    uint8_t* out_status, nbl_particle* out_secondary, uint8_t* out_has_secondary, float* out_scalar,
    uint8_t* out_type)
{
    for (uint32_t i = 0; i < n; ++i) {
        rng_stream rng(uniforms + (size_t)i * NBL_NUM_UNIFORMS);
        particle p = from_nbl(io[i]);
        int mat = material_idx[i];
        uint8_t st = NBL_OUT_ALIVE;
        if (out_has_secondary) out_has_secondary[i] = 0;
        if (event == NBL_EVENT_SAMPLE_PATH) {
            uint8_t type; float dist;
            sample_path(m->m[mat], p, rng, type, dist, nullptr);
            if (out_scalar) out_scalar[i] = dist;
            if (out_type) out_type[i] = type;
        } else if (event == NBL_EVENT_ELASTIC) {
            elastic_execute(m->m[mat], p, rng, nullptr);
        } else if (event == NBL_EVENT_INELASTIC) {
            particle sec;
            const bool made = inelastic_event(m->m[mat], p, rng, sec, nullptr);
            if (made && out_secondary) out_secondary[i] = to_nbl(sec);
            if (out_has_secondary) out_has_secondary[i] = made ? 1 : 0;
        } else if (event == NBL_EVENT_INTERSECT) {
            if (!g) return -1;
            st = (uint8_t)intersect_execute(m->m, g->g.input[last_tri[i]], p, mat, rng, nullptr);
        } else {
            return -1;
        }
        io[i] = to_nbl(p);
        if (out_material) out_material[i] = mat;
        if (out_status) out_status[i] = st;
    }
    return 0;
}

void orc_rng_seed(uint64_t seed, uint64_t first, uint32_t n, uint32_t* states)
{
    for (uint32_t i = 0; i < n; ++i) {
        xorwow st; xorwow_seed(st, seed, first + i);
        std::memcpy(states + 6 * (size_t)i, st.s, sizeof st.s);
    }
}
void orc_rng_uniform_batch(uint32_t* states, uint32_t n, uint32_t draws, float* out)
{
    for (uint32_t i = 0; i < n; ++i) {
        xorwow st; std::memcpy(st.s, states + 6 * (size_t)i, sizeof st.s);
        for (uint32_t k = 0; k < draws; ++k) out[(size_t)i * draws + k] = u32_to_unit(xorwow_next(st));
        std::memcpy(states + 6 * (size_t)i, st.s, sizeof st.s);
    }
}

orc_slots orc_slots_create(uint32_t cap, uint64_t seed)
{
    orc_slots o = new orc_slots_s;
    slots& s = o->s;
    s.cap = cap;
    s.status.assign(cap, NBL_ST_TERMINATED);                 /* gpu_particle_manager.inl:27-32 */
    s.particle_idx.resize(cap);
    for (uint32_t i = 0; i < cap; ++i) s.particle_idx[i] = i;
    s.parts.assign(cap, particle{ { 0, 0, 0 }, { 0, 0, 0 }, 0 });
    s.tags.assign(cap, 0);
    s.material.assign(cap, NBL_MAT_VACUUM);
    s.last_tri.assign(cap, NBL_NO_TRIANGLE);
    s.rng.resize(cap);
    for (uint32_t i = 0; i < cap; ++i) xorwow_seed(s.rng[i], seed, i);  /* init_random_states (gpu_driver.inl:348-358) */
    return o;
}
void orc_slots_destroy(orc_slots s) { delete s; }
void orc_slots_read(orc_slots o, const nbl_slot_view* v)
{
    slots& s = o->s;
    for (uint32_t i = 0; i < s.cap; ++i) {
        if (v->status) v->status[i] = s.status[i];
        if (v->particle_idx) v->particle_idx[i] = s.particle_idx[i];
        if (v->particles) v->particles[i] = to_nbl(s.parts[i]);
        if (v->tags) v->tags[i] = s.tags[i];
        if (v->material) v->material[i] = s.material[i];
        if (v->last_tri) v->last_tri[i] = s.last_tri[i];
        if (v->rng) std::memcpy(v->rng + 6 * (size_t)i, s.rng[i].s, sizeof s.rng[i].s);
    }
}
void orc_slots_write(orc_slots o, const nbl_slot_view* v)
{
    slots& s = o->s;
    for (uint32_t i = 0; i < s.cap; ++i) {
        if (v->status) s.status[i] = v->status[i];
        if (v->particle_idx) s.particle_idx[i] = v->particle_idx[i];
        if (v->particles) s.parts[i] = from_nbl(v->particles[i]);
        if (v->tags) s.tags[i] = v->tags[i];
        if (v->material) s.material[i] = v->material[i];
        if (v->last_tri) s.last_tri[i] = v->last_tri[i];
        if (v->rng) std::memcpy(s.rng[i].s, v->rng + 6 * (size_t)i, sizeof s.rng[i].s);
    }
}
/* gpu_particle_manager::push (gpu_particle_manager.inl:105-164) */
uint32_t orc_slots_push(orc_slots o, const nbl_particle* p, const uint32_t* tags, uint32_t n)
{
    slots& s = o->s;
    uint32_t k = 0;
    for (uint32_t i = 0; i < s.cap && k < n; ++i) {
        if (s.status[i] != NBL_ST_TERMINATED) continue;
        s.status[i] = NBL_ST_NO_EVENT; s.parts[i] = from_nbl(p[k]); s.tags[i] = tags[k];
        s.material[i] = NBL_MAT_VACUUM; s.last_tri[i] = NBL_NO_TRIANGLE;
        ++k;
    }
    return k;
}
void orc_sort_status(const uint8_t* keys, uint32_t n, uint32_t* out) { sort_status(keys, n, out); }

/* gpu_driver::do_iteration (gpu_driver.inl:312-346) with the kernels of
 * gpu_driver.inl:403-567, run sequentially over slots. */
void orc_slots_iteration(orc_slots o, orc_materials mm, orc_geometry gg, float min_energy, float max_energy)
{
    slots& s = o->s;
    const std::vector<material>& mats = mm->m;
    const geometry& geom = gg->g;
    /* init (gpu_driver.inl:403-481) */
    for (uint32_t idx = 0; idx < s.cap; ++idx) {
        if (!slot_active(s.status[idx])) continue;
        particle p = s.parts[idx];
        if (!geom.in_domain(p.pos)) { s.status[idx] = NBL_ST_TERMINATED; continue; }
        uint8_t type = 0; float dist = geom.max_extent;
        const int mi = s.material[idx];
        if (mi >= 0) {
            const material& m = mats[mi];
            if (p.kin_energy < m.barrier + min_energy || p.kin_energy > m.barrier + max_energy) {
                s.status[idx] = NBL_ST_TERMINATED; continue;
            }
            rng_xorwow r(s.rng[idx]);
            sample_path(m, p, r, type, dist, nullptr);
            s.rng[idx] = r.st;
        }
        normalise(p.dir);
        uint32_t hit;
        const float d = geom.propagate(p.pos, p.dir, dist, s.last_tri[idx], mi, hit, nullptr, nullptr);
        if (hit == NBL_NO_TRIANGLE) {
            p.pos += p.dir * dist;
            s.status[idx] = type == 1 ? NBL_ST_INELASTIC_EVENT : type == 2 ? NBL_ST_ELASTIC_EVENT : NBL_ST_NO_EVENT;
        } else {
            p.pos += p.dir * d;
            s.status[idx] = NBL_ST_INTERSECT_EVENT;
            s.last_tri[idx] = hit;
        }
        s.parts[idx] = p;
    }
    /* sort (gpu_particle_manager.inl:237-246) */
    sort_status(s.status.data(), s.cap, s.particle_idx.data());
    /* intersect (gpu_driver.inl:483-504) */
    for (uint32_t i = 0; i < s.cap; ++i) {
        const uint32_t idx = s.particle_idx[i];
        if (s.status[idx] != NBL_ST_INTERSECT_EVENT) continue;
        rng_xorwow r(s.rng[idx]);
        particle p = s.parts[idx]; int mat = s.material[idx];
        const int res = intersect_execute(mats, geom.input[s.last_tri[idx]], p, mat, r, nullptr);
        s.rng[idx] = r.st;
        if (res == NBL_OUT_DETECTED) s.status[idx] = NBL_ST_DETECTED;
        else if (res == NBL_OUT_TERMINATED) s.status[idx] = NBL_ST_TERMINATED;
        else { s.parts[idx] = p; s.material[idx] = mat; }
    }
    /* elastic (gpu_driver.inl:543-567) */
    for (uint32_t i = 0; i < s.cap; ++i) {
        const uint32_t idx = s.particle_idx[i];
        if (s.status[idx] != NBL_ST_ELASTIC_EVENT) continue;
        s.last_tri[idx] = NBL_NO_TRIANGLE;
        rng_xorwow r(s.rng[idx]);
        elastic_execute(mats[s.material[idx]], s.parts[idx], r, nullptr);
        s.rng[idx] = r.st;
    }
    /* inelastic (gpu_driver.inl:506-541) with the secondary slot of thread i
     * at particle_idx[cap-1-i] (gpu_particle_manager.inl:354-376) */
    for (uint32_t i = 0; i < s.cap; ++i) {
        const uint32_t idx = s.particle_idx[i];
        if (!(s.status[idx] == NBL_ST_PENDING || s.status[idx] == NBL_ST_INELASTIC_EVENT)) continue;
        const uint32_t sec_idx = s.particle_idx[s.cap - 1 - i];
        if (s.status[sec_idx] != NBL_ST_TERMINATED) { s.status[idx] = NBL_ST_PENDING; continue; }
        s.status[idx] = NBL_ST_INELASTIC_EVENT;
        s.last_tri[idx] = NBL_NO_TRIANGLE;
        rng_xorwow r(s.rng[idx]);
        particle sec;
        const bool made = inelastic_event(mats[s.material[idx]], s.parts[idx], r, sec, nullptr);
        s.rng[idx] = r.st;
        if (made) {
            s.status[sec_idx] = NBL_ST_NEW_SECONDARY;
            s.parts[sec_idx] = sec;
            s.tags[sec_idx] = s.tags[idx];
            s.material[sec_idx] = s.material[idx];
            s.last_tri[sec_idx] = NBL_NO_TRIANGLE;
        }
    }
}

static int mt_run(orc_geometry g, orc_materials m, const nbl_particle* primaries, const uint32_t* tags,
    uint64_t n, float min_energy, uint32_t seed, int32_t n_threads, int32_t rng_kind, double time_limit_s,
    nbl_particle* det, uint32_t* det_tags, uint64_t max_det, uint64_t* hist, orc_run_stats* stats,
    const deposit_grid* dep_spec, deposit_grid* dep_out);

/* cpu_mt_main.cpp:158-219 */
int orc_cpu_mt_run(orc_geometry g, orc_materials m, const nbl_particle* primaries, const uint32_t* tags,
    uint64_t n, float min_energy, uint32_t seed, int32_t n_threads, int32_t rng_kind, double time_limit_s,
    nbl_particle* det, uint32_t* det_tags, uint64_t max_det, uint64_t* hist, orc_run_stats* stats)
{
    return mt_run(g, m, primaries, tags, n, min_energy, seed, n_threads, rng_kind, time_limit_s, det, det_tags,
        max_det, hist, stats, nullptr, nullptr);
}

/* cpu_energydep.cpp:158-219: the same run with an energy-deposition tally */
int orc_cpu_mt_run_deposit(orc_geometry g, orc_materials m, const nbl_particle* primaries, const uint32_t* tags,
    uint64_t n, float min_energy, uint32_t seed, int32_t n_threads, uint64_t* hist, orc_run_stats* stats,
    int32_t nx, int32_t ny, int32_t nz, const float* lo, const float* hi, double* energy, double* total,
    uint64_t* count)
{
    deposit_grid spec;
    spec.nx = nx; spec.ny = ny; spec.nz = nz;
    spec.lo = { lo[0], lo[1], lo[2] };
    spec.inv = { (float)nx / (hi[0] - lo[0]), (float)ny / (hi[1] - lo[1]), (float)nz / (hi[2] - lo[2]) };
    deposit_grid out = spec;
    out.e.assign((size_t)nx * ny * nz, 0.0);
    const int r = mt_run(g, m, primaries, tags, n, min_energy, seed, n_threads, ORC_RNG_XORWOW, 0, nullptr, nullptr,
        0, hist, stats, &spec, &out);
    if (energy) std::copy(out.e.begin(), out.e.end(), energy);
    if (total) *total = out.total;
    if (count) *count = out.count;
    return r;
}

static int mt_run(orc_geometry g, orc_materials m, const nbl_particle* primaries, const uint32_t* tags,
    uint64_t n, float min_energy, uint32_t seed, int32_t n_threads, int32_t rng_kind, double time_limit_s,
    nbl_particle* det, uint32_t* det_tags, uint64_t max_det, uint64_t* hist, orc_run_stats* stats,
    const deposit_grid* dep_spec, deposit_grid* dep_out)
{
    if (n_threads <= 0) n_threads = (int32_t)std::max(1u, std::thread::hardware_concurrency());
    const float max_energy = orc_materials_max_energy(m);
    std::mt19937 master(seed);                                   /* cpu_mt_main.cpp:63 */
    std::mutex pool_mutex, out_mutex;
    uint64_t next = 0;
    std::atomic<bool> stop{ false };
    orc_run_stats tot{};
    uint64_t n_det = 0;
    std::vector<uint64_t> h(NBL_HIST_BINS, 0);
    const auto t0 = std::chrono::steady_clock::now();
    auto sim_loop = [&](uint32_t thread_seed, int tid) {
        counters c; uint64_t prim = 0, d_se = 0, d_bse = 0, d_all = 0;
        std::vector<uint64_t> hl(NBL_HIST_BINS, 0);
        std::vector<std::pair<nbl_particle, uint32_t>> local;
        deposit_grid dl;
        if (dep_spec) { dl = *dep_spec; dl.e.assign((size_t)dl.nx * dl.ny * dl.nz, 0.0); }
        auto body = [&](auto& drv) {
            if (dep_spec) drv.dep = &dl;
            for (;;) {
                uint64_t i;
                {   /* work_pool::get_work(1) (work_pool.h:45-56) */
                    std::lock_guard<std::mutex> lk(pool_mutex);
                    if (next >= n || stop.load()) break;
                    i = next++;
                }
                const particle p = from_nbl(primaries[i]);
                drv.push(&p, &tags[i], 1);
                drv.simulate_to_end();
                drv.flush_detected([&](const particle& q, uint32_t t) {
                    ++d_all;
                    if (q.kin_energy < 50) ++d_se; else ++d_bse;
                    const int bin = (int)q.kin_energy;
                    if (bin >= 0 && bin < NBL_HIST_BINS) ++hl[bin];
                    if (det) local.push_back({ to_nbl(q), t });
                });
                ++prim;
                if (time_limit_s > 0 && (prim & 63) == 0 && tid == 0) {
                    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                    if (el > time_limit_s) stop.store(true);
                }
            }
            c = drv.c;
        };
        if (rng_kind == ORC_RNG_XORWOW) {
            xorwow st; xorwow_seed(st, thread_seed, (uint64_t)tid);
            rng_xorwow r(st);
            cpu_driver<rng_xorwow> drv(g->g, m->m, min_energy, max_energy, r);
            body(drv);
        } else {
            cpu_driver<rng_mt> drv(g->g, m->m, min_energy, max_energy, rng_mt(thread_seed));
            body(drv);
        }
        std::lock_guard<std::mutex> lk(out_mutex);
        tot.primaries += prim; tot.events += c.events; tot.detected += d_all;
        tot.detected_se += d_se; tot.detected_bse += d_bse;
        tot.lookups_1d += c.lookups_1d; tot.lookups_2d += c.lookups_2d; tot.rounddowns += c.rounddowns;
        tot.node_steps += c.node_steps; tot.tri_tests += c.tri_tests;
        tot.boundary_events += c.boundary_events; tot.secondaries += c.secondaries;
        for (int b = 0; b < NBL_HIST_BINS; ++b) h[b] += hl[b];
        if (dep_out) {
            for (size_t k = 0; k < dl.e.size(); ++k) dep_out->e[k] += dl.e[k];
            dep_out->total += dl.total; dep_out->count += dl.count;
        }
        for (auto& e : local) {
            if (n_det < max_det) { det[n_det] = e.first; if (det_tags) det_tags[n_det] = e.second; }
            ++n_det;
        }
    };
    std::vector<std::thread> threads;
    for (int t = 0; t < n_threads; ++t) threads.emplace_back(sim_loop, (uint32_t)master(), t);
    for (auto& t : threads) t.join();
    const auto t1 = std::chrono::steady_clock::now();
    tot.seconds = std::chrono::duration<double>(t1 - t0).count();
    tot.threads = (uint32_t)n_threads;
    if (hist) for (int b = 0; b < NBL_HIST_BINS; ++b) hist[b] = h[b];
    if (stats) *stats = tot;
    return 0;
}

} /* extern "C" */
