// leaf_cull_check.cpp -- soundness of the octree's leaf culls on one scene
// (built by tests/test_leaf_culls.py with hipcc, run on the CPU).
//
// The device walk drops a leaf's candidates (wave_propagate.h: leaf_dropped,
// with the leaf's leaf_info of nebula_amd/csrc/nbl_device.h)
//   * side cull: when dot(axis, d) > t (or < -t) and the material on that
//     side is the ignored one -- exact only if then every triangle's float
//     side test fl(dot(cross(e1, e2), d)) < 0 of the reference
//     (octree.inl:128-138) comes out false (true);
//   * slab cull: when the segment [pos, pos + T d] lies outside the leaf's
//     slab -- exact only if the reference's float intersect_ray
//     (triangle.inl:7-28) then returns no distance in (0, T] for any of the
//     leaf's triangles.
// For every leaf with a cone, directions are drawn just beyond the side
// threshold (the hardest case) and uniformly beyond it, on both sides, with
// |d|^2 anywhere in the accepted window 1 +- 1e-4.  For every leaf with a
// slab, segments are drawn from points of the cell pushed against the slab,
// with directions down to the smallest accepted |dot(axis, d)| and lengths
// that bring the far end up to the slab; only those the device's own float
// decision drops are kept, and every triangle of the leaf is tested with
// them.
//
// It also checks every leaf's ropes (leaf_rope): for each face, the
// neighbour code by the reference's rule (octree.inl:165-185), descended
// from the root with the reference's own centre formula (octree.inl:62-67,
// AABB += AABB.w * halfsize * (bit - 0.5)), must give the rope's child word
// and, bit for bit, its centre.
//
// usage: leaf_cull_check <triangles.bin> <draws per leaf> <cone scale> <slab scale> [swap]
// A scale below 1 plants a mistake (a smaller side threshold, a thinner
// slab), and "swap" one in the ropes (the faces' signs exchanged); the check
// must then find violations.  Exit status 1 on any.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "octree_build.h"

using namespace nbl;

namespace {

// triangle::intersect_ray of the reference, in its operation order
float reference_intersect(const tri_rec& r, f3 start, f3 dir)
{
    const f3 r0{ r.r0x, r.r0y, r.r0z }, e1{ r.e1x, r.e1y, r.e1z }, e2{ r.e2x, r.e2y, r.e2z };
    const f3 pvec = cross(dir, e2);
    const float det = dot(e1, pvec);
    if (std::fabs(det) < EPSILON) return -1;
    const float u = dot(start - r0, pvec) / det;
    if ((u < -EPSILON) || (u > 1 + EPSILON)) return -1;
    const f3 qvec = cross(start - r0, e1);
    const float v = dot(dir, qvec) / det;
    if ((v < -EPSILON) || (u + v > 1 + EPSILON)) return -1;
    const float t = dot(e2, qvec) / det;
    if (t < 0) return -1;
    return t;
}

struct cell { double c[3], h[3]; };

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 5 && argc != 6) { fprintf(stderr, "usage: %s triangles.bin draws cone_scale slab_scale [swap]\n", argv[0]); return 2; }
    const bool swap_faces = argc == 6;
    FILE* f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 2; }
    std::vector<nbl_triangle> tris;
    nbl_triangle x;
    while (fread(&x, sizeof x, 1, f) == 1) tris.push_back(x);
    fclose(f);
    const int draws = atoi(argv[2]);
    const float cone_scale = (float)atof(argv[3]), slab_scale = (float)atof(argv[4]);
    const built_octree o = build_octree(tris.data(), (uint32_t)tris.size());

    // leaf cells from the node words (children of a cell: centre +- h/2)
    std::vector<std::pair<uint64_t, cell>> leaves;
    std::vector<uint64_t> codes;                        // location code of each entry of leaves
    {
        struct item { uint32_t node; cell cl; uint64_t code; };
        std::vector<item> stack;
        cell root{ { o.center[0], o.center[1], o.center[2] }, { o.halfsize[0], o.halfsize[1], o.halfsize[2] } };
        stack.push_back({ 0u, root, 1ull });
        while (!stack.empty()) {
            const item it = stack.back();
            stack.pop_back();
            for (int oc = 0; oc < 8; ++oc) {
                cell ch;
                for (int a = 0; a < 3; ++a) {
                    ch.h[a] = 0.5 * it.cl.h[a];
                    ch.c[a] = it.cl.c[a] + (((oc >> a) & 1) ? ch.h[a] : -ch.h[a]);
                }
                const uint64_t w = o.nodes[(size_t)it.node * 8 + oc];
                const uint64_t code = (it.code << 3) | (uint64_t)oc;
                if (w & LEAF_BIT) { leaves.push_back({ w, ch }); codes.push_back(code); }
                else stack.push_back({ (uint32_t)w, ch, code });
            }
        }
    }
    // ropes
    size_t n_ropes = 0, rope_bad = 0;
    for (size_t i = 0; i < leaves.size(); ++i) {
        const uint32_t ord = (uint32_t)leaves[i].first;
        for (uint32_t a = 0; a < 3; ++a)
            for (uint32_t neg = 0; neg < 2; ++neg) {
                uint64_t loc = codes[i];
                const unsigned mask = 1u << a, value = neg ? mask : 0u;
                bool moved = false;
                while (loc > 1) {
                    if ((loc & mask) == value) { loc ^= mask; moved = true; break; }
                    loc >>= 3;
                }
                if (!moved) continue;
                ++n_ropes;
                int digits = 0;
                for (uint64_t t = loc; t > 1; t >>= 3) ++digits;
                float ax = o.center[0], ay = o.center[1], az = o.center[2], aw = 1;
                uint32_t node = 0;
                uint64_t w = 0;
                for (int i2 = 3 * (digits - 1); i2 >= 0; i2 -= 3) {
                    const unsigned oct = (unsigned)(loc >> i2) & 7u;
                    w = o.nodes[(size_t)node * 8 + oct];
                    ax += aw * o.halfsize[0] * ((float)(oct & 1) - 0.5f);
                    ay += aw * o.halfsize[1] * ((float)((oct & 2) >> 1) - 0.5f);
                    az += aw * o.halfsize[2] * ((float)((oct & 4) >> 2) - 0.5f);
                    aw *= 0.5f;
                    node = (uint32_t)w;
                }
                const leaf_rope& r = o.ropes[(size_t)ord * 6 + 2 * a + (swap_faces ? 1 - neg : neg)];
                const bool same = r.word == w && r.depth == (uint32_t)digits &&
                                  !memcmp(&r.cx, &ax, 4) && !memcmp(&r.cy, &ay, 4) && !memcmp(&r.cz, &az, 4);
                if (!same) ++rope_bad;
            }
    }

    std::mt19937_64 rng(1);
    std::normal_distribution<double> N;
    std::uniform_real_distribution<double> U(0, 1);
    auto orthogonal = [&](double ax, double ay, double az, double& px, double& py, double& pz) {
        px = N(rng); py = N(rng); pz = N(rng);
        const double dp = px * ax + py * ay + pz * az;
        px -= dp * ax; py -= dp * ay; pz -= dp * az;
        const double pl = std::sqrt(px * px + py * py + pz * pz);
        px /= pl; py /= pl; pz /= pl;
    };
    size_t cand = 0, cand_cone = 0, cand_slab = 0, checks = 0, bad = 0, seg = 0, seg_bad = 0;
    for (const auto& [w, cl] : leaves) {
        const uint32_t cnt = (uint32_t)((w >> 32) & 0x7FFFFFFFu);
        if (cnt == 0) continue;
        cand += cnt;
        const leaf_info& li = o.leaves[(uint32_t)w];
        const uint32_t first = li.first;
        if (!(li.t < 2)) continue;
        cand_cone += cnt;
        leaf_info c = li;
        c.t *= cone_scale;
        const double al = std::sqrt((double)c.ax * c.ax + (double)c.ay * c.ay + (double)c.az * c.az);
        const double ax = c.ax / al, ay = c.ay / al, az = c.az / al;
        for (int s = 0; s < draws; ++s) {
            double px, py, pz;
            orthogonal(ax, ay, az, px, py, pz);
            // cosine to the axis: just beyond t for 3 draws in 4, else uniform beyond it
            const double cc = (s % 4 != 3) ? c.t * (1 + 1e-7 * U(rng)) + 1e-8 : c.t + (1 - c.t) * U(rng);
            const double sg = (s & 1) ? 1 : -1;
            const double sn = std::sqrt(std::max(0.0, 1 - cc * cc));
            const double len = 1 + (2 * U(rng) - 1) * 4.99e-5;
            const f3 d{ (float)(len * (sg * cc * ax + sn * px)), (float)(len * (sg * cc * ay + sn * py)),
                        (float)(len * (sg * cc * az + sn * pz)) };
            if (!(std::fabs(dot(d, d) - 1) < 1e-4f)) continue;
            const float cf = c.ax * d.x + c.ay * d.y + c.az * d.z;
            const int side = cf > c.t ? 1 : (cf < -c.t ? -1 : 0);
            if (side == 0) continue;
            ++checks;
            for (uint32_t k = first; k < first + cnt; ++k) {
                const tri_rec& r = o.recs[k];
                const f3 e1{ r.e1x, r.e1y, r.e1z }, e2{ r.e2x, r.e2y, r.e2z };
                const bool neg = dot(cross(e1, e2), d) < 0;
                if ((side > 0) == neg) ++bad;
            }
        }

        if (!(li.lo > -HUGE_VALF)) continue;
        cand_slab += cnt;
        struct { float lo, hi, g, tmax; } sl;
        {
            const float mid = 0.5f * (li.lo + li.hi), hw = 0.5f * (li.hi - li.lo) * slab_scale;
            sl.lo = mid - hw; sl.hi = mid + hw;
            // the device's limits on direction and segment length
            sl.g = li.t + 0.0202f;
            const float sx = (float)cl.h[0], sy = (float)cl.h[1], sz = (float)cl.h[2];
            sl.tmax = 2.03f * (sx + sy + sz) + 1e-3f;
        }
        for (int s = 0; s < draws; ++s) {
            const bool above = s & 1;
            // a point of the cell, moved along the axis to just beyond the slab when it stays inside
            double p[3];
            for (int a = 0; a < 3; ++a) p[a] = cl.c[a] + (2 * U(rng) - 1) * cl.h[a];
            const double hp = p[0] * ax + p[1] * ay + p[2] * az;
            if (s % 3 != 2) {
                const double target = above ? sl.hi + 1e-6 * (1 + std::fabs(sl.hi)) * U(rng) : sl.lo - 1e-6 * (1 + std::fabs(sl.lo)) * U(rng);
                for (int a = 0; a < 3; ++a) p[a] += (target - hp) * (a == 0 ? ax : a == 1 ? ay : az);
            }
            bool inside = true;
            for (int a = 0; a < 3; ++a) inside &= std::fabs(p[a] - cl.c[a]) <= cl.h[a];
            if (!inside) continue;
            const f3 pos{ (float)p[0], (float)p[1], (float)p[2] };
            // direction: |cos| just above g for half the draws, toward or away from the slab
            double qx, qy, qz;
            orthogonal(ax, ay, az, qx, qy, qz);
            const double g = sl.g;
            const double cc = (s % 4 < 2) ? std::min(1.0, g * (1 + 1e-6 * U(rng)) + 1e-7) : g + (1 - g) * U(rng);
            const double sg = (U(rng) < 0.5) ? 1 : -1;
            const double sn = std::sqrt(std::max(0.0, 1 - cc * cc));
            const double len = 1 + (2 * U(rng) - 1) * 4.99e-5;
            const f3 d{ (float)(len * (sg * cc * ax + sn * qx)), (float)(len * (sg * cc * ay + sn * qy)),
                        (float)(len * (sg * cc * az + sn * qz)) };
            // the device's decision (wave_propagate.h: leaf_dropped)
            if (!(std::fabs(dot(d, d) - 1) < 1e-4f)) continue;
            const float cf = li.ax * d.x + li.ay * d.y + li.az * d.z;
            if (!(std::fabs(cf) >= sl.g)) continue;
            const float h0 = li.ax * pos.x + li.ay * pos.y + li.az * pos.z;
            // length: up to the slab when heading for it, else anything up to tmax
            float T = sl.tmax * (float)U(rng);
            if ((above && cf < 0) || (!above && cf > 0)) {
                const double reach = above ? (h0 - sl.hi) / -cf : (sl.lo - h0) / cf;
                if (reach > 0 && s % 2 == 0) T = (float)std::min<double>(sl.tmax, reach * (1 - 1e-6));
            }
            if (!(T <= sl.tmax) || !(T > 0)) continue;
            const float h1 = h0 + T * cf;
            if (!((h0 > sl.hi && h1 > sl.hi) || (h0 < sl.lo && h1 < sl.lo))) continue;
            ++seg;
            for (uint32_t k = first; k < first + cnt; ++k) {
                const float t = reference_intersect(o.recs[k], pos, d);
                if (t > 0 && t <= T) ++seg_bad;
            }
        }
    }
    printf("candidates %zu in_cone_leaves %zu directions %zu violations %zu in_slab_leaves %zu segments %zu "
           "slab_violations %zu ropes %zu rope_violations %zu\n", cand, cand_cone, checks, bad, cand_slab, seg, seg_bad,
           n_ropes, rope_bad);
    return (bad != 0 || seg_bad != 0 || rope_bad != 0) ? 1 : 0;
}
