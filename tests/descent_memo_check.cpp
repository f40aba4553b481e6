// descent_memo_check.cpp -- the persistent kernel's octree descent with its
// per-lane path memo and leaf ropes (nbl_device.h: descend with an
// lds_path_cache) against the reference-order descent (descend with
// no_path_cache: the location code's digits from the root, then the position,
// octree.inl:47-84), which the propagate parity tests pin to the oracle.
// Built by tests/test_descent_memo.py with hipcc and run on the CPU.
//
// One simulated lane makes a long random sequence of descents on one scene,
// as the walk does: new steps (code 1) at positions near the last leaf, far
// away, or exactly on one of its centre planes; and next cells of a walk (the
// reference's neighbour code of the last leaf, octree.inl:158-181), with and
// without the leaf's rope.  next_cell's closed form is checked against the
// reference's loop on the walk and on random codes.  Every call must return the same leaf word,
// location code, centre and half size as the reference-order descent, the
// same node-step count, and leave the memo's LDS levels on its code's path.
//
// New steps inside the memo's leaf by more than the geometry's margin take
// the leaf without descending; new steps are also placed within a few
// hundred ulp of the last leaf's faces, across that margin.
//
// usage: descent_memo_check <scene.tri> <calls> <seed> [cluster x y z r] [nudge] [nomargin]
//   cluster:  half of the far new steps land in the cube of half size r there
//   nudge:    the reference-order side sees positions on a centre plane one
//             ulp higher (a planted mistake: the check must then fail)
//   nomargin: the leaf-box test without its margin (a planted mistake: the
//             check must then fail on positions next to a leaf's faces)
#include "octree_build.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

using namespace nbl;

// The reference's neighbour-code step (octree.inl:158-181), as a loop: climb
// until the digit's axis bit equals the direction's value, flip it.
static uint64_t ref_next_cell(uint64_t location, int axis, f3 dir)
{
    const unsigned mask = (unsigned)(-axis);
    unsigned value;
    if (mask == 1) value = (dir.x >= 0) ? 0u : 1u;
    else if (mask == 2) value = (dir.y >= 0) ? 0u : 2u;
    else value = (dir.z >= 0) ? 0u : 4u;
    while (location > 1) {
        if ((location & mask) == value) { location ^= mask; break; }
        location >>= 3;
    }
    return location;
}

// next_cell (nbl_device.h, closed form) against the loop on random codes of
// every length 0..21, all six faces, direction components of both signs,
// +-0 and NaN.
static bool check_next_cell(std::mt19937_64& rng, long n)
{
    const float comps[] = { 1.f, -1.f, 0.f, -0.f, NAN, 1e-30f, -1e-30f };
    for (long i = 0; i < n; ++i) {
        const uint32_t digits = (uint32_t)(rng() % 22);
        uint64_t code = 1;
        for (uint32_t k = 0; k < digits; ++k) code = (code << 3) | (rng() & 7u);
        if (rng() % 4 == 0 && digits) code |= (rng() % 2 ? 0x1249249249249249ull : 0x2492492492492492ull) & ((1ull << (3 * digits)) - 1);
        const int axis = -(1 << (int)(rng() % 3));
        const f3 d{ comps[rng() % 7], comps[rng() % 7], comps[rng() % 7] };
        uint64_t got = code;
        next_cell(got, axis, d);
        if (got != ref_next_cell(code, axis, d)) {
            printf("next_cell MISMATCH code %llx axis %d: %llx / %llx\n", (unsigned long long)code, axis,
                (unsigned long long)got, (unsigned long long)ref_next_cell(code, axis, d));
            return false;
        }
    }
    return true;
}

int main(int argc, char** argv)
{
    if (argc < 4) { fprintf(stderr, "usage: %s scene.tri calls seed [cluster x y z r] [nudge]\n", argv[0]); return 2; }
    std::vector<nbl_triangle> tris;
    {
        FILE* f = fopen(argv[1], "rb");
        if (!f) { fprintf(stderr, "cannot open %s\n", argv[1]); return 2; }
        nbl_triangle t;
        while (fread(&t, sizeof t, 1, f) == 1) tris.push_back(t);
        fclose(f);
    }
    const long calls = atol(argv[2]);
    std::mt19937_64 rng((uint64_t)atoll(argv[3]));
    bool cluster = false, nudge = false, nomargin = false;
    float cx = 0, cy = 0, cz = 0, cr = 0;
    for (int i = 4; i < argc; ++i) {
        if (!strcmp(argv[i], "cluster") && i + 4 < argc) {
            cluster = true;
            cx = strtof(argv[i + 1], nullptr); cy = strtof(argv[i + 2], nullptr);
            cz = strtof(argv[i + 3], nullptr); cr = strtof(argv[i + 4], nullptr);
            i += 4;
        } else if (!strcmp(argv[i], "nudge")) nudge = true;
        else if (!strcmp(argv[i], "nomargin")) nomargin = true;
    }

    if (!check_next_cell(rng, 2000000)) return 1;
    const built_octree o = build_octree(tris.data(), (uint32_t)tris.size());
    dgeom g{};
    g.nodes = o.nodes.data();
    g.ropes = o.ropes.data();
    g.n_ropes = (uint32_t)o.ropes.size();
    g.cx = o.center[0]; g.cy = o.center[1]; g.cz = o.center[2];
    g.hx = o.halfsize[0]; g.hy = o.halfsize[1]; g.hz = o.halfsize[2];
    // the leaf-box margin as the library sets it (geometry_view)
    g.mx = nomargin ? 0.f : ldexpf(fabsf(g.cx) + g.hx, -16);
    g.my = nomargin ? 0.f : ldexpf(fabsf(g.cy) + g.hy, -16);
    g.mz = nomargin ? 0.f : ldexpf(fabsf(g.cz) + g.hz, -16);
    set_derived(g);
    const size_t n_words = o.nodes.size();

    std::vector<uint32_t> lds(lds_path_cache::kWords * 64, 0xFFFFFFFFu);   // one lane's column
    lds_path_cache pc{ lds.data(), 1ull, 0ull, 0u };
    prop_counters c_ref{}, c_memo{};
    std::uniform_real_distribution<float> U(-1.f, 1.f);

    f3 pos{ g.cx, g.cy, g.cz }, ls{ g.hx, g.hy, g.hz }, lc = pos;
    std::vector<f3> anc{ pos };   // centres of the internal nodes on the last path
    uint64_t last_loc = 1, last_leaf = 0;
    long n_new = 0, n_rope = 0, n_code = 0, n_plane = 0, n_face = 0, n_box = 0, max_depth = 0;
    for (long it = 0; it < calls; ++it) {
        uint64_t loc = 1;
        uint32_t rope = NO_ROPE;
        bool on_plane = false;
        const int kind = last_loc == 1 ? 0 : (int)(rng() % 3);
        if (kind == 0) {
            const int m = (int)(rng() % 8);
            if (m == 0 && cluster && rng() % 2) pos = { cx + cr * U(rng), cy + cr * U(rng), cz + cr * U(rng) };
            else if (m == 0) pos = { g.cx + g.hx * U(rng), g.cy + g.hy * U(rng), g.cz + g.hz * U(rng) };
            else if (m == 7 && pc.depth > 0) {
                // next to a face of the last leaf: within +-400 ulp of the
                // largest coordinate, on either side of the margin
                const uint32_t a = (uint32_t)(rng() % 3);
                const float sgn = rng() % 2 ? 1.f : -1.f;
                const float cc = a == 0 ? lc.x : a == 1 ? lc.y : lc.z, hh = a == 0 ? ls.x : a == 1 ? ls.y : ls.z;
                const float big = (a == 0 ? fabsf(g.cx) + g.hx : a == 1 ? fabsf(g.cy) + g.hy : fabsf(g.cz) + g.hz);
                const float ulp = ldexpf(1.f, std::ilogb(big) - 23);
                const float v = cc + sgn * (hh - ulp * (float)((long)(rng() % 801) - 400));
                if (a == 0) pos.x = v; else if (a == 1) pos.y = v; else pos.z = v;
                ++n_face;
            } else {
                const float sc = m < 4 ? 0.1f : 2.f;
                pos = { pos.x + sc * ls.x * U(rng), pos.y + sc * ls.y * U(rng), pos.z + sc * ls.z * U(rng) };
            }
            // exactly on the centre plane of an internal node of the last path
            // (the planes the descent compares positions with)
            if (rng() % 4 == 0) { pos.x = anc[rng() % anc.size()].x; on_plane = true; }
            if (rng() % 4 == 0) { pos.y = anc[rng() % anc.size()].y; on_plane = true; }
            ++n_new;
        } else {
            const uint32_t a = (uint32_t)(rng() % 3), neg = (uint32_t)(rng() % 2);
            const f3 d{ a == 0 ? (neg ? -0.6f : 0.6f) : 0.5f, a == 1 ? (neg ? -0.6f : 0.6f) : 0.5f,
                        a == 2 ? (neg ? -0.6f : 0.6f) : 0.5f };
            loc = ref_next_cell(last_loc, -(1 << a), d);
            uint64_t closed = last_loc;
            next_cell(closed, -(1 << a), d);
            if (closed != loc) { printf("next_cell MISMATCH on the walk at call %ld\n", it); return 1; }
            if (loc <= 1) { last_loc = 1; --it; continue; }   // the walk leaves the root: no next cell
            const float half = a == 0 ? ls.x : a == 1 ? ls.y : ls.z;
            const float step = (neg ? -1.f : 1.f) * half * (0.5f + 0.6f * (U(rng) + 1.f));
            if (a == 0) pos.x += step; else if (a == 1) pos.y += step; else pos.z += step;
            if (kind == 1) { rope = (uint32_t)last_leaf * 6u + 2u * a + neg; ++n_rope; }
            else ++n_code;
        }
        n_plane += on_plane;
        f3 pref = pos;
        if (nudge && on_plane) pref.x = nextafterf(pref.x, INFINITY), pref.y = nextafterf(pref.y, INFINITY);
        // new steps the leaf-box test takes without descending (counted, not checked here)
        if (kind == 0 && pc.depth > 0 && fabsf(pos.x - lc.x) < ls.x - g.mx && fabsf(pos.y - lc.y) < ls.y - g.my &&
            fabsf(pos.z - lc.z) < ls.z - g.mz) ++n_box;
        uint64_t l_ref = loc, l_memo = loc;
        f3 c_a{}, s_a{}, c_b{}, s_b{};
        const uint64_t w_ref = descend<true>(g, pref, l_ref, c_a, s_a, &c_ref, (no_path_cache*)nullptr);
        const uint64_t w_memo = descend<true>(g, pos, l_memo, c_b, s_b, &c_memo, &pc, rope);
        // the memo's descent loads a subset of the reference-order descent's nodes
        if (w_ref != w_memo || l_ref != l_memo || memcmp(&c_a, &c_b, sizeof c_a) || memcmp(&s_a, &s_b, sizeof s_a) ||
            c_memo.node_steps > c_ref.node_steps) {
            printf("MISMATCH call %ld kind %d rope %u code %llx: leaf %llx / %llx, code %llx / %llx, steps %u / %u\n",
                it, kind, rope, (unsigned long long)loc, (unsigned long long)w_ref, (unsigned long long)w_memo,
                (unsigned long long)l_ref, (unsigned long long)l_memo, c_ref.node_steps, c_memo.node_steps);
            return 1;
        }
        if (!(w_memo & LEAF_BIT)) { printf("call %ld: descent ended on an internal node\n", it); return 1; }
        // the memo's LDS levels hold the internal nodes of its code's path
        for (uint32_t l = 1; l < pc.depth && l < lds_path_cache::kLevels; ++l) {
            size_t n = 0;
            for (uint32_t j = 1; j <= l; ++j) {
                const size_t k = n * 8 + pc.digit(j);
                if (k >= n_words) { printf("call %ld: memo path leaves the node array\n", it); return 1; }
                n = (uint32_t)o.nodes[k];
            }
            if (lds[l * 64] != (uint32_t)n) { printf("call %ld: memo level %u is off the path\n", it, l); return 1; }
        }
        if ((long)pc.depth > max_depth) max_depth = pc.depth;
        last_loc = l_memo; last_leaf = (uint32_t)w_memo; ls = s_b; lc = c_b;
        {
            const uint32_t depth = (uint32_t)(63 - clz64(l_memo)) / 3;
            f3 a{ g.cx, g.cy, g.cz }, h{ 0.5f * g.hx, 0.5f * g.hy, 0.5f * g.hz };
            anc.assign(1, a);
            for (uint32_t l = 1; l < depth; ++l) {
                cell_step((unsigned)(l_memo >> (3 * (depth - l))) & 7u, a, h);
                anc.push_back(a);
            }
        }
        if (rng() % 64 == 0) last_loc = 1;
    }
    printf("ok calls %ld new %ld rope %ld code %ld on_plane %ld face %ld leaf_box %ld node_steps %u ref_node_steps %u "
           "max_depth %ld tree_depth %u\n", calls, n_new, n_rope, n_code, n_plane, n_face, n_box, c_memo.node_steps,
           c_ref.node_steps, max_depth, o.max_depth);
    return 0;
}
