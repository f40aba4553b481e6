// octree_build.cpp -- host-side octree construction (see octree_build.h).
//
// The leaf sets must equal the reference's exactly, because the traversal's
// "t <= best + EPSILON" rule makes the hit triangle depend on which
// triangles a leaf holds and in which order.  The build therefore follows
// geometry/octree/octree_builder.cpp:93-241 step for step (insertion order,
// split cascade, per-leaf duplicate rejection) and evaluates the overlap test
// of geometry/octree/tribox.cc:109-184 with the same double-precision
// expressions.  Output goes straight to the device layout: one 64-byte word
// block per internal node and contiguous 48-byte records per leaf.
#include "octree_build.h"

#include <algorithm>
#include <cmath>
#include <deque>
#include <stdexcept>
#include <string>

namespace nbl {
namespace {

struct fvec { float x, y, z; };
inline bool same(fvec a, fvec b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// A triangle as the reference stores it (r0, e1 = r1 - r0, e2 = r2 - r0)
// with its vertices reconstructed in float (triangle.inl:36-49).
struct tri_in {
    fvec r0, e1, e2;
    fvec v[3];
    int32_t mi, mo;
};

struct box { float c[3], h[3]; };

// Child box of octant o (octree_builder.cpp:47-64): half = 0.5*h, centre
// shifted by -half or +half per axis, all in float.
inline box child_box(const box& b, int o)
{
    box r;
    for (int a = 0; a < 3; ++a) r.h[a] = 0.5f * b.h[a];
    r.c[0] = b.c[0] + (float)(((o & 1) << 1) - 1) * r.h[0];
    r.c[1] = b.c[1] + (float)((o & 2) - 1) * r.h[1];
    r.c[2] = b.c[2] + (float)(((o & 4) >> 1) - 1) * r.h[2];
    return r;
}

// Separating-axis triangle/box test (Akenine-Moller 2001) in double, with the
// box half size inflated by 1.01 (octree_builder.cpp:30-39).
bool overlaps(const tri_in& t, const box& b)
{
    const double hx = 1.01 * b.h[0], hy = 1.01 * b.h[1], hz = 1.01 * b.h[2];
    const double cx = b.c[0], cy = b.c[1], cz = b.c[2];
    const double ax = (double)t.v[0].x - cx, ay = (double)t.v[0].y - cy, az = (double)t.v[0].z - cz;
    const double bx = (double)t.v[1].x - cx, by = (double)t.v[1].y - cy, bz = (double)t.v[1].z - cz;
    const double qx = (double)t.v[2].x - cx, qy = (double)t.v[2].y - cy, qz = (double)t.v[2].z - cz;
    // slabs
    auto slab = [](double p, double q, double r, double h) {
        double lo = p, hi = p;
        if (q < lo) lo = q;
        if (q > hi) hi = q;
        if (r < lo) lo = r;
        if (r > hi) hi = r;
        return !(lo > h || hi < -h);
    };
    if (!slab(ax, bx, qx, hx) || !slab(ay, by, qy, hy) || !slab(az, bz, qz, hz)) return false;
    // edges: f = b - a, g = q - b, k = a - q
    const double fx = bx - ax, fy = by - ay, fz = bz - az;
    const double gx = qx - bx, gy = qy - by, gz = qz - bz;
    const double kx = ax - qx, ky = ay - qy, kz = az - qz;
    // triangle plane through a with normal f x g
    {
        const double nx = fy * gz - fz * gy, ny = fz * gx - fx * gz, nz = fx * gy - fy * gx;
        const double lox = nx > 0 ? -hx - ax : hx - ax, hix = nx > 0 ? hx - ax : -hx - ax;
        const double loy = ny > 0 ? -hy - ay : hy - ay, hiy = ny > 0 ? hy - ay : -hy - ay;
        const double loz = nz > 0 ? -hz - az : hz - az, hiz = nz > 0 ? hz - az : -hz - az;
        if (nx * lox + ny * loy + nz * loz > 0) return false;
        if (!(nx * hix + ny * hiy + nz * hiz >= 0)) return false;
    }
    // edge x unit-axis tests; each projects two of the three vertices
    auto sep = [](double p0, double p1, double rad) {
        const double lo = p0 < p1 ? p0 : p1, hi = p0 < p1 ? p1 : p0;
        return lo > rad || hi < -rad;
    };
    {   // edge f: x-axis (a,q), y-axis (a,q), z-axis (b,q)
        const double ux = std::fabs(fx), uy = std::fabs(fy), uz = std::fabs(fz);
        if (sep(fz * ay - fy * az, fz * qy - fy * qz, uz * hy + uy * hz)) return false;
        if (sep(-fz * ax + fx * az, -fz * qx + fx * qz, uz * hx + ux * hz)) return false;
        if (sep(fy * bx - fx * by, fy * qx - fx * qy, uy * hx + ux * hy)) return false;
    }
    {   // edge g: x-axis (a,q), y-axis (a,q), z-axis (a,b)
        const double ux = std::fabs(gx), uy = std::fabs(gy), uz = std::fabs(gz);
        if (sep(gz * ay - gy * az, gz * qy - gy * qz, uz * hy + uy * hz)) return false;
        if (sep(-gz * ax + gx * az, -gz * qx + gx * qz, uz * hx + ux * hz)) return false;
        if (sep(gy * ax - gx * ay, gy * bx - gx * by, uy * hx + ux * hy)) return false;
    }
    {   // edge k: x-axis (a,b), y-axis (a,b), z-axis (b,q)
        const double ux = std::fabs(kx), uy = std::fabs(ky), uz = std::fabs(kz);
        if (sep(kz * ay - ky * az, kz * by - ky * bz, uz * hy + uy * hz)) return false;
        if (sep(-kz * ax + kx * az, -kz * bx + kx * bz, uz * hx + ux * hz)) return false;
        if (sep(ky * bx - kx * by, ky * qx - kx * qy, uy * hx + ux * hy)) return false;
    }
    return true;
}

// Same vertex set in any of the six orders (triangle.inl:85-103).
bool perfect_overlap(const tri_in& a, const tri_in& b)
{
    const fvec* p = a.v; const fvec* q = b.v;
    return (same(p[0], q[0]) && same(p[1], q[1]) && same(p[2], q[2])) ||
           (same(p[0], q[1]) && same(p[1], q[2]) && same(p[2], q[0])) ||
           (same(p[0], q[2]) && same(p[1], q[0]) && same(p[2], q[1])) ||
           (same(p[0], q[0]) && same(p[1], q[2]) && same(p[2], q[1])) ||
           (same(p[0], q[1]) && same(p[1], q[0]) && same(p[2], q[2])) ||
           (same(p[0], q[2]) && same(p[1], q[1]) && same(p[2], q[0]));
}

struct bnode {
    int64_t kids = -1;               // first of 8 consecutive children, -1 = leaf
    int depth = 0;
    std::vector<uint32_t> list;      // leaf contents (input indices), ascending
};

class tree {
public:
    static constexpr int kMaxDepth = 21, kSplitCount = 16;   // octree_builder.h:157-158
    std::vector<tri_in> tris;
    std::vector<bnode> nodes;

    int64_t add_children(int depth)
    {
        const int64_t k = (int64_t)nodes.size();
        nodes.resize(nodes.size() + 8);
        for (int o = 0; o < 8; ++o) nodes[k + o].depth = depth;
        return k;
    }
    // octree_child::insert (octree_builder.cpp:93-126)
    void insert(int64_t ni, uint32_t t, const box& b)
    {
        if (!overlaps(tris[t], b)) return;
        if (nodes[ni].kids >= 0) {
            const int64_t k = nodes[ni].kids;
            for (int o = 0; o < 8; ++o) insert(k + o, t, child_box(b, o));
            return;
        }
        for (uint32_t s : nodes[ni].list)
            if (perfect_overlap(tris[t], tris[s])) return;
        nodes[ni].list.push_back(t);
        if ((int)nodes[ni].list.size() > kSplitCount && nodes[ni].depth < kMaxDepth) split(ni, b);
    }
    // octree_child::split (octree_builder.cpp:140-156)
    void split(int64_t ni, const box& b)
    {
        std::vector<uint32_t> mine;
        mine.swap(nodes[ni].list);
        const int64_t k = add_children(nodes[ni].depth + 1);
        nodes[ni].kids = k;
        for (int o = 0; o < 8; ++o) {
            const box cb = child_box(b, o);
            for (uint32_t t : mine) insert(k + o, t, cb);
        }
    }
};

// Cull bounds of the leaf records r[0, n) in the cell bx: the side cone and
// the slab of its leaf_info (nbl_device.h); first and mats are the caller's.
//
// Cone: the normals are the float cross products of the side test, so the
// cone holds the vectors that test uses; angles are measured in double
// against the stored float axis a.
//
// Slab: with u = 2^-24, |d| = 1 and D = |det| = |n| g, g = |d . n|/|n|, the
// reference's float intersect_ray (triangle.inl:7-28, no fused operations)
// has |det_f - det| <= 7u |e1||e2|, and for tvec = pos - r0 (|tvec| <= R)
// |un_f - un|, |vn_f - vn|, |tn_f - tn| <= 8u R |e2|, 8u R |e1|,
// 8u R |e1||e2|; with 7u |e1||e2| <= D/20, |det_f| >= 0.75 D and
//   |t_f - t| + |u_f - u||e1| + |v_f - v||e2|
//     <= (K/g)(u/0.75)(24 R + 7.7 (T + |e1| + |e2|)) + 1.1u (T + |e1| + |e2|) = E,
// K = |e1||e2|/|n|, for any hit the float tests accept (|u|, |v| < 1.1,
// t <= 1.1 T).  The float hit point p + t_f d then lies within E of
// r0 + u_f e1 + v_f e2, a point of the triangle grown by the EPSILON
// tolerance (extent ext along a), so a segment beyond the slab grown by
// 2E + ext, plus the rounding of the device's own heights, meets none.
leaf_info leaf_bounds(const tri_rec* r, size_t n, const box& bx)
{
    const float inf = HUGE_VALF;
    const leaf_info none{ 0, 0, 0, 2, -inf, inf, 0, 0 };
    std::vector<double> u(3 * n), kf(n);
    double sx = 0, sy = 0, sz = 0;
    for (size_t k = 0; k < n; ++k) {
        if (r[k].mats != r[0].mats) return none;
        const f3 e1{ r[k].e1x, r[k].e1y, r[k].e1z }, e2{ r[k].e2x, r[k].e2y, r[k].e2z };
        const f3 nf = cross(e1, e2);
        const double x = nf.x, y = nf.y, z = nf.z;
        const double len = std::sqrt(x * x + y * y + z * z);
        const double l1 = std::sqrt((double)e1.x * e1.x + (double)e1.y * e1.y + (double)e1.z * e1.z);
        const double l2 = std::sqrt((double)e2.x * e2.x + (double)e2.y * e2.y + (double)e2.z * e2.z);
        if (!(len > 1e-3 * l1 * l2)) return none;   // near-degenerate (or not finite)
        u[3 * k] = x / len; u[3 * k + 1] = y / len; u[3 * k + 2] = z / len;
        kf[k] = l1 * l2 / len;
        sx += u[3 * k]; sy += u[3 * k + 1]; sz += u[3 * k + 2];
    }
    const double sl = std::sqrt(sx * sx + sy * sy + sz * sz);
    if (!(sl > 0.5 * (double)n)) return none;
    const float ax = (float)(sx / sl), ay = (float)(sy / sl), az = (float)(sz / sl);
    const double al = std::sqrt((double)ax * ax + (double)ay * ay + (double)az * az);
    double cmin = 1;
    for (size_t k = 0; k < n; ++k)
        cmin = std::min(cmin, (u[3 * k] * ax + u[3 * k + 1] * ay + u[3 * k + 2] * az) / al);
    if (!(cmin >= 0.2)) return none;
    const double theta = std::acos(std::min(cmin, 1.0)) + 1e-9;
    const double tc = std::sin(theta + 1e-5) * (1 + 1e-4) + 1e-6;
    const leaf_info cone{ ax, ay, az, std::nextafter((float)tc, 2.0f), -inf, inf, 0, 0 };

    // slab
    // the device's |fl(a . d)| >= t + 0.0202 gives a . d >= sin(theta + asin g_min)
    // for |d| within 5e-5 of 1, so every |d . n| >= g_min |n|
    const double U = 0x1p-24, g_min = 0.02;
    const double half = std::sqrt((double)bx.h[0] * bx.h[0] + (double)bx.h[1] * bx.h[1] + (double)bx.h[2] * bx.h[2]);
    const double cn = std::sqrt((double)bx.c[0] * bx.c[0] + (double)bx.c[1] * bx.c[1] + (double)bx.c[2] * bx.c[2]);
    // bounds the device's segment limit fl(2.03 fl(sx + sy + sz) + 1e-3)
    const double tmax = 2.04 * ((double)bx.h[0] + bx.h[1] + bx.h[2]) + 1.01e-3;
    double lo = HUGE_VAL, hi = -HUGE_VAL, err = 0, ext = 0;
    for (size_t k = 0; k < n; ++k) {
        const double r0[3] = { r[k].r0x, r[k].r0y, r[k].r0z };
        const double e1[3] = { r[k].e1x, r[k].e1y, r[k].e1z }, e2[3] = { r[k].e2x, r[k].e2y, r[k].e2z };
        const double l1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        const double l2 = std::sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        if (!(7 * U * kf[k] / (0.75 * g_min) <= 0.05)) return cone;
        const double R = std::sqrt((r0[0] - bx.c[0]) * (r0[0] - bx.c[0]) + (r0[1] - bx.c[1]) * (r0[1] - bx.c[1]) +
                                   (r0[2] - bx.c[2]) * (r0[2] - bx.c[2])) + half + 1e-3;
        const double L = l1 + l2;
        err = std::max(err, kf[k] / g_min * U / 0.75 * (24 * R + 7.7 * (tmax + L)) + 1.1 * U * (tmax + L));
        ext = std::max(ext, 3 * (EPSILON + 3e-7) * L);
        for (int v = 0; v < 3; ++v) {
            const double p[3] = { r0[0] + (v == 1 ? e1[0] : v == 2 ? e2[0] : 0), r0[1] + (v == 1 ? e1[1] : v == 2 ? e2[1] : 0),
                                  r0[2] + (v == 1 ? e1[2] : v == 2 ? e2[2] : 0) };
            const double h = p[0] * ax + p[1] * ay + p[2] * az;
            lo = std::min(lo, h); hi = std::max(hi, h);
        }
    }
    if (!(std::isfinite(lo) && std::isfinite(hi))) return cone;
    // the device's heights fl(a . pos) and fl(h0 + fl(T fl(a . d)))
    const double eh = 8 * U * (cn + half + tmax + 1);
    const double delta = 2 * err + ext + eh + 1e-6;
    leaf_info res = cone;
    res.lo = std::nextafter((float)(lo - delta), -inf);
    res.hi = std::nextafter((float)(hi + delta), inf);
    return res;
}

// Ropes of every leaf (leaf_rope in nbl_device.h): for each face, the
// neighbour code of octree::propagate (octree.inl:165-185), descended from the
// root along its digits (octree.inl:58-69) with the device's centre steps
// (c += +-h, h halves, in float from the root's centre and half size / 2).
std::vector<leaf_rope> leaf_ropes(const built_octree& o, const std::vector<uint64_t>& leaf_code, const box& root)
{
    std::vector<leaf_rope> ropes(leaf_code.size() * 6, leaf_rope{ 0, 0, 0, 0, 0 });
    for (size_t i = 0; i < leaf_code.size(); ++i)
        for (uint32_t a = 0; a < 3; ++a)
            for (uint32_t neg = 0; neg < 2; ++neg) {
                const uint64_t mask = 1ull << a, value = neg ? mask : 0;
                uint64_t loc = leaf_code[i];
                bool moved = false;
                while (loc > 1) {
                    if ((loc & mask) == value) { loc ^= mask; moved = true; break; }
                    loc >>= 3;
                }
                if (!moved) continue;                  // the walk leaves the root here
                const uint32_t D = (uint32_t)(63 - __builtin_clzll(loc)) / 3;
                float c[3] = { root.c[0], root.c[1], root.c[2] }, h[3];
                for (int k = 0; k < 3; ++k) h[k] = 0.5f * root.h[k];
                uint64_t w = 0;
                uint32_t node = 0;
                for (uint32_t l = 1; l <= D; ++l) {
                    const unsigned oct = (unsigned)(loc >> (3 * (D - l))) & 7u;
                    w = o.nodes[(size_t)node * 8 + oct];
                    for (int k = 0; k < 3; ++k) {
                        c[k] += ((oct >> k) & 1u) ? h[k] : -h[k];
                        h[k] *= 0.5f;
                    }
                    if ((w & LEAF_BIT) && l < D) throw std::runtime_error("neighbour code passes through a leaf");
                    node = (uint32_t)w;
                }
                ropes[i * 6 + 2 * a + neg] = leaf_rope{ w, c[0], c[1], c[2], D };
            }
    return ropes;
}

}  // namespace

built_octree build_octree(const nbl_triangle* in, uint32_t n)
{
    if (n == 0) throw std::runtime_error("no triangles given");
    tree tr;
    tr.tris.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        tri_in& t = tr.tris[i];
        t.r0 = { in[i].r0.x, in[i].r0.y, in[i].r0.z };
        t.e1 = { in[i].e1.x, in[i].e1.y, in[i].e1.z };
        t.e2 = { in[i].e2.x, in[i].e2.y, in[i].e2.z };
        t.v[0] = t.r0;
        t.v[1] = { t.r0.x + t.e1.x, t.r0.y + t.e1.y, t.r0.z + t.e1.z };
        t.v[2] = { t.r0.x + t.e2.x, t.r0.y + t.e2.y, t.r0.z + t.e2.z };
        t.mi = in[i].material_in; t.mo = in[i].material_out;
        if (t.mi < -32768 || t.mi > 32767 || t.mo < -32768 || t.mo > 32767)
            throw std::runtime_error("material index of triangle " + std::to_string(i) + " outside int16 range");
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(t.v[k].x) || !std::isfinite(t.v[k].y) || !std::isfinite(t.v[k].z))
                throw std::runtime_error("triangle " + std::to_string(i) + " has a non-finite vertex");
    }
    // root box: AABB of all vertices + 1 per side (octree_builder.cpp:176-205)
    fvec lo = tr.tris[0].v[0], hi = lo;
    for (const tri_in& t : tr.tris) {
        fvec tl{ std::min(std::min(t.v[0].x, t.v[1].x), t.v[2].x), std::min(std::min(t.v[0].y, t.v[1].y), t.v[2].y),
                 std::min(std::min(t.v[0].z, t.v[1].z), t.v[2].z) };
        fvec th{ std::max(std::max(t.v[0].x, t.v[1].x), t.v[2].x), std::max(std::max(t.v[0].y, t.v[1].y), t.v[2].y),
                 std::max(std::max(t.v[0].z, t.v[1].z), t.v[2].z) };
        lo = { std::min(lo.x, tl.x), std::min(lo.y, tl.y), std::min(lo.z, tl.z) };
        hi = { std::max(hi.x, th.x), std::max(hi.y, th.y), std::max(hi.z, th.z) };
    }
    box root;
    root.c[0] = 0.5f * (lo.x + hi.x); root.c[1] = 0.5f * (lo.y + hi.y); root.c[2] = 0.5f * (lo.z + hi.z);
    root.h[0] = 0.5f * (hi.x - lo.x) + 1; root.h[1] = 0.5f * (hi.y - lo.y) + 1; root.h[2] = 0.5f * (hi.z - lo.z) + 1;

    tr.nodes.emplace_back();
    tr.nodes[0].kids = tr.add_children(1);
    for (uint32_t t = 0; t < n; ++t) {
        if (!overlaps(tr.tris[t], root)) continue;   // never true: the root box holds every vertex
        const int64_t k = tr.nodes[0].kids;
        for (int o = 0; o < 8; ++o) tr.insert(k + o, t, child_box(root, o));
    }

    // breadth-first order (== Morton order of the reference's linearization)
    std::vector<int64_t> order;
    order.reserve(tr.nodes.size());
    order.push_back(0);
    std::vector<box> boxes(tr.nodes.size());          // cell of every node
    std::vector<uint64_t> code(tr.nodes.size());      // location code of every node
    boxes[0] = root;
    code[0] = 1;
    for (size_t q = 0; q < order.size(); ++q) {
        const int64_t k = tr.nodes[order[q]].kids;
        if (k >= 0)
            for (int o = 0; o < 8; ++o) {
                order.push_back(k + o);
                boxes[k + o] = child_box(boxes[order[q]], o);
                code[k + o] = (code[order[q]] << 3) | (uint64_t)o;
            }
    }
    std::vector<uint64_t> leaf_code;
    std::vector<uint64_t> word(tr.nodes.size(), 0);   // child word of every node
    built_octree out;
    uint32_t n_internal = 0;
    std::vector<uint8_t> seen(n, 0);
    for (int64_t id : order) {
        const bnode& b = tr.nodes[id];
        out.max_depth = std::max<uint32_t>(out.max_depth, (uint32_t)b.depth);
        if (b.kids >= 0) { word[id] = n_internal++; continue; }
        if (out.recs.size() + b.list.size() > 0xFFFFFFFFull)
            throw std::runtime_error("too many leaf triangle references");
        const size_t start = out.recs.size();
        word[id] = LEAF_BIT | ((uint64_t)b.list.size() << 32) | (uint64_t)out.leaves.size();
        leaf_code.push_back(code[id]);
        ++out.n_leaves;
        for (uint32_t t : b.list) {
            const tri_in& s = tr.tris[t];
            tri_rec r;
            r.r0x = s.r0.x; r.r0y = s.r0.y; r.r0z = s.r0.z;
            r.e1x = s.e1.x; r.e1y = s.e1.y; r.e1z = s.e1.z;
            r.e2x = s.e2.x; r.e2y = s.e2.y; r.e2z = s.e2.z;
            r.mats = (uint32_t)(uint16_t)(int16_t)s.mi | ((uint32_t)(uint16_t)(int16_t)s.mo << 16);
            r.id = t; r.pad = 0;
            out.recs.push_back(r);
            if (!seen[t]) { seen[t] = 1; ++out.n_unique; }
        }
        leaf_info li{ 0, 0, 0, 2, -HUGE_VALF, HUGE_VALF, 0, 0 };
        if (!b.list.empty()) {
            li = leaf_bounds(&out.recs[start], b.list.size(), boxes[id]);
            li.mats = out.recs[start].mats;
        }
        li.first = (uint32_t)start;
        out.leaves.push_back(li);
    }
    out.nodes.assign((size_t)n_internal * 8, 0);
    for (int64_t id : order) {
        const bnode& b = tr.nodes[id];
        if (b.kids < 0) continue;
        for (int o = 0; o < 8; ++o) out.nodes[word[id] * 8 + o] = word[b.kids + o];
    }
    out.ropes = leaf_ropes(out, leaf_code, root);
    for (int a = 0; a < 3; ++a) { out.center[a] = root.c[a]; out.halfsize[a] = root.h[a]; }
    // 2*magnitude(halfsize) in its CPU form (octree.inl:263, vector_math.h:117)
    out.max_extent = 2 * std::sqrt(root.h[0] * root.h[0] + root.h[1] * root.h[1] + root.h[2] * root.h[2]);
    return out;
}

void export_linearized(const built_octree& o, std::vector<int32_t>& data, std::vector<uint32_t>& order_out)
{
    data.clear(); order_out.clear();
    // breadth-first walk over child words; root is internal node 0
    struct item { uint64_t w; int64_t parent_slot; };
    std::deque<item> q;
    q.push_back({ 0, -1 });
    std::vector<item> bfs;
    while (!q.empty()) {
        item it = q.front(); q.pop_front();
        bfs.push_back(it);
        if (!(it.w & LEAF_BIT))
            for (int c = 0; c < 8; ++c) q.push_back({ o.nodes[(it.w & 0xFFFFFFFFull) * 8 + c], (int64_t)bfs.size() - 1 });
    }
    // triangle order: first visit over leaves in BFS order (octree_builder.cpp:283-308)
    std::vector<int32_t> tmap;
    for (const item& it : bfs) {
        if (!(it.w & LEAF_BIT)) continue;
        const uint32_t first = leaf_first(o, it.w), cnt = (uint32_t)((it.w >> 32) & 0x7FFFFFFFu);
        for (uint32_t k = first; k < first + cnt; ++k) {
            const uint32_t id = o.recs[k].id;
            if (id >= tmap.size()) tmap.resize(id + 1, -1);
            if (tmap[id] >= 0) continue;
            tmap[id] = (int32_t)order_out.size();
            order_out.push_back(id);
        }
    }
    std::vector<int64_t> offset(bfs.size());
    std::vector<int> octant(bfs.size(), 0);
    {   // octant of each item within its parent
        std::vector<int> next(bfs.size(), 0);
        for (size_t i = 1; i < bfs.size(); ++i) octant[i] = next[bfs[i].parent_slot]++;
    }
    for (size_t i = 0; i < bfs.size(); ++i) {
        offset[i] = (int64_t)data.size();
        const uint64_t w = bfs[i].w;
        if (w & LEAF_BIT) {
            const uint32_t first = leaf_first(o, w), cnt = (uint32_t)((w >> 32) & 0x7FFFFFFFu);
            for (uint32_t k = first; k < first + cnt; ++k) data.push_back(tmap[o.recs[k].id]);
            data.push_back(-1);
        } else {
            for (int c = 0; c < 8; ++c) data.push_back(0);
        }
        if (i > 0) {
            const int64_t v = offset[i];
            data[offset[bfs[i].parent_slot] + octant[i]] = (int32_t)((w & LEAF_BIT) ? -v : v);
        }
    }
}

}  // namespace nbl
