// side_cone_check.cpp -- soundness of the octree's leaf side cones on one
// scene (built by tests/test_side_cones.py with hipcc, run on the CPU).
//
// A leaf's cone (leaf_cone, nebula_amd/csrc/nbl_device.h) lets the walk skip
// the leaf when dot(axis, d) > t (or < -t) and the material on that side is
// the ignored one; that is only exact if then every triangle's float side
// test fl(dot(cross(e1, e2), d)) < 0 of the reference (octree.inl:128-138)
// comes out false (true).  For every leaf with a cone, directions are drawn
// just beyond the threshold (the hardest case) and uniformly beyond it, on
// both sides, with |d|^2 anywhere in the accepted window 1 +- 1e-4; every
// triangle of the leaf is tested with them.
//
// usage: side_cone_check <triangles.bin> <directions per leaf> <threshold scale>
// A threshold scale below 1 plants a mistake (the cone claims more than it
// holds); the check must then find violations.  Exit status 1 on any.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "octree_build.h"

using namespace nbl;

int main(int argc, char** argv)
{
    if (argc != 4) { fprintf(stderr, "usage: %s triangles.bin per_leaf scale\n", argv[0]); return 2; }
    FILE* f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 2; }
    std::vector<nbl_triangle> tris;
    nbl_triangle x;
    while (fread(&x, sizeof x, 1, f) == 1) tris.push_back(x);
    fclose(f);
    const int per_leaf = atoi(argv[2]);
    const float scale = (float)atof(argv[3]);
    const built_octree o = build_octree(tris.data(), (uint32_t)tris.size());

    std::mt19937_64 rng(1);
    std::normal_distribution<double> N;
    std::uniform_real_distribution<double> U(0, 1);
    size_t cand = 0, cand_cone = 0, checks = 0, bad = 0;
    for (uint64_t w : o.nodes) {
        if (!(w & LEAF_BIT)) continue;
        const uint32_t first = (uint32_t)w, cnt = (uint32_t)((w >> 32) & 0x7FFFFFFFu);
        if (cnt == 0) continue;
        cand += cnt;
        leaf_cone c = o.cone[first];
        if (!(c.t < 2)) continue;
        cand_cone += cnt;
        c.t *= scale;
        const double al = std::sqrt((double)c.ax * c.ax + (double)c.ay * c.ay + (double)c.az * c.az);
        const double ax = c.ax / al, ay = c.ay / al, az = c.az / al;
        for (int s = 0; s < per_leaf; ++s) {
            // a unit vector p orthogonal to the axis
            double px = N(rng), py = N(rng), pz = N(rng);
            const double dp = px * ax + py * ay + pz * az;
            px -= dp * ax; py -= dp * ay; pz -= dp * az;
            const double pl = std::sqrt(px * px + py * py + pz * pz);
            // cosine to the axis: just beyond t for 3 draws in 4, else uniform beyond it
            const double cc = (s % 4 != 3) ? c.t * (1 + 1e-7 * U(rng)) + 1e-8 : c.t + (1 - c.t) * U(rng);
            const double sg = (s & 1) ? 1 : -1;
            const double sn = std::sqrt(std::max(0.0, 1 - cc * cc));
            const double len = 1 + (2 * U(rng) - 1) * 4.99e-5;
            const f3 d{ (float)(len * (sg * cc * ax + sn * px / pl)), (float)(len * (sg * cc * ay + sn * py / pl)),
                        (float)(len * (sg * cc * az + sn * pz / pl)) };
            // the device's decision (wave_propagate.h: leaf_culled)
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
    }
    printf("candidates %zu in_cone_leaves %zu directions %zu violations %zu\n", cand, cand_cone, checks, bad);
    return bad != 0;
}
