/*
 * ref_harness.cpp -- thin C API around the reference's OWN geometry and table
 * code, compiled from its sources under /root/reference by oracle/Makefile
 * into oracle/_ref/libnbl_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY.  It pins the oracle's restatement of
 *   core/triangle.inl (intersect_ray),
 *   geometry/octree/octree_builder.cpp + tribox.cc (leaf sets, Morton order),
 *   geometry/octree.inl (in_domain, propagate),
 *   geometry/trilist.inl (brute-force propagate),
 *   common/util/table_1D.inl, table_2D.inl, table_3D.inl (get, get_rounddown),
 *   io/load_tri_file.cpp, io/load_pri_file.cpp (file formats, primary order)
 * bit for bit.  The physics headers are not included: they pull in the HDF5
 * material loader (io/hdf5_file.h), an external library.
 *
 * Triangle identities: the reference returns pointers into its own triangle
 * arrays.  They are mapped back to indices of the create() input by content
 * (duplicates are removed by the builder, so content is unique), and the
 * array base is recovered from the first pointer a probe ray returns.
 */
#include <cstring>
#include <cstdint>
#include <iostream>
#include <unordered_map>
#include <vector>

#include "config/config.h"
#include "core/triangle.h"
#include "core/events.h"
#include "geometry/octree.h"
#include "geometry/trilist.h"
#include "common/util/table_1D.h"
#include "common/util/table_2D.h"
#include "common/util/table_3D.h"
#include "io/load_tri_file.h"
#include "io/load_pri_file.h"

#include "../include/nbl_hip.h"

namespace {

struct key_hash {
    size_t operator()(const std::vector<uint32_t>& k) const
    {
        size_t h = 1469598103934665603ull;
        for (uint32_t w : k) { h ^= w; h *= 1099511628211ull; }
        return h;
    }
};
std::vector<uint32_t> tri_key(const triangle& t)
{
    const vec3 a = t.r0(), b = t.r1(), c = t.r2();
    const float f[9] = { a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z };
    std::vector<uint32_t> k(11);
    std::memcpy(k.data(), f, sizeof f);
    k[9] = (uint32_t)t.material_in; k[10] = (uint32_t)t.material_out;
    return k;
}
triangle to_tri(const nbl_triangle& t)
{
    const vec3 r0{ t.r0.x, t.r0.y, t.r0.z };
    const vec3 r1 = r0 + vec3{ t.e1.x, t.e1.y, t.e1.z };
    const vec3 r2 = r0 + vec3{ t.e2.x, t.e2.y, t.e2.z };
    triangle tri(r0, r1, r2, t.material_in, t.material_out);
    return tri;
}

struct ref_geom {
    std::vector<triangle> input;
    nbl::geometry::octree_builder::linearized_octree lin;
    nbl::geometry::octree<false> oct;
    std::vector<uint32_t> morton_to_input;
    std::unordered_map<std::vector<uint32_t>, uint32_t, key_hash> key_to_morton;
    const triangle* base = nullptr;
};

}  // namespace

extern "C" {

void* ref_geometry_create(const nbl_triangle* tris, uint32_t n)
{
    ref_geom* g = new ref_geom;
    for (uint32_t i = 0; i < n; ++i) g->input.push_back(to_tri(tris[i]));
    /* NB: triangle's constructor stores e1 = r1 - r0; the input already holds
     * e1, so r0 + e1 - r0 may differ by rounding.  Build from the stored form. */
    for (uint32_t i = 0; i < n; ++i) {
        static_assert(sizeof(triangle) == sizeof(nbl_triangle), "layout");
        std::memcpy(&g->input[i], &tris[i], sizeof(nbl_triangle));
    }
    g->lin = nbl::geometry::octree_builder::linearized_octree(
        nbl::geometry::octree_builder::octree_root(g->input));
    g->oct = nbl::geometry::octree<false>::create(g->lin);
    std::unordered_map<std::vector<uint32_t>, uint32_t, key_hash> in_map;
    for (uint32_t i = n; i-- > 0;) in_map[tri_key(g->input[i])] = i;   /* first occurrence wins */
    for (size_t k = 0; k < g->lin.triangle_data.size(); ++k) {
        const auto key = tri_key(g->lin.triangle_data[k]);
        g->morton_to_input.push_back(in_map.at(key));
        g->key_to_morton[key] = (uint32_t)k;
    }
    /* recover the base of octree<false>'s private triangle copy */
    for (size_t k = 0; k < g->lin.triangle_data.size() && !g->base; ++k) {
        const triangle& t = g->lin.triangle_data[k];
        const vec3 c = (1.0f / 3) * (t.r0() + t.r1() + t.r2());
        vec3 nrm = t.get_normal();
        const float m = magnitude(nrm);
        if (!(m > 0)) continue;
        nrm = nrm / m;
        const vec3 start = c + nrm * 0.5f;
        intersect_event e = g->oct.propagate(start, -nrm, 1.0f, nullptr, 0x7fff0000);
        if (e.isect_triangle) {
            const uint32_t mk = g->key_to_morton.at(tri_key(*e.isect_triangle));
            g->base = e.isect_triangle - mk;
        }
    }
    return g;
}

void ref_geometry_destroy(void* p)
{
    ref_geom* g = (ref_geom*)p;
    nbl::geometry::octree<false>::destroy(g->oct);
    delete g;
}

int ref_geometry_has_base(void* p) { return ((ref_geom*)p)->base != nullptr; }

void ref_geometry_info(void* p, float* center, float* halfsize, float* max_extent,
    uint64_t* lin_size, uint32_t* n_unique)
{
    ref_geom* g = (ref_geom*)p;
    center[0] = g->lin.center.x; center[1] = g->lin.center.y; center[2] = g->lin.center.z;
    halfsize[0] = g->lin.halfsize.x; halfsize[1] = g->lin.halfsize.y; halfsize[2] = g->lin.halfsize.z;
    *max_extent = g->oct.get_max_extent();
    *lin_size = g->lin.octree_data.size();
    *n_unique = (uint32_t)g->lin.triangle_data.size();
}

void ref_geometry_export(void* p, int32_t* octree_data, uint32_t* tri_order)
{
    ref_geom* g = (ref_geom*)p;
    std::memcpy(octree_data, g->lin.octree_data.data(), g->lin.octree_data.size() * sizeof(int32_t));
    std::memcpy(tri_order, g->morton_to_input.data(), g->morton_to_input.size() * sizeof(uint32_t));
}

void ref_in_domain_batch(void* p, const nbl_vec3* pos, uint32_t n, uint8_t* out)
{
    ref_geom* g = (ref_geom*)p;
    for (uint32_t i = 0; i < n; ++i) out[i] = g->oct.in_domain({ pos[i].x, pos[i].y, pos[i].z }) ? 1 : 0;
}

/* Returns -1 if an ignore_tri is given but the array base is unknown. */
int ref_propagate_batch(void* p, const nbl_vec3* pos, const nbl_vec3* dir, const float* dist,
    const uint32_t* ignore_tri, const int32_t* ignore_mat, uint32_t n, float* out_dist, uint32_t* out_tri)
{
    ref_geom* g = (ref_geom*)p;
    std::vector<int64_t> input_to_morton(g->input.size(), -1);
    for (size_t k = 0; k < g->morton_to_input.size(); ++k) input_to_morton[g->morton_to_input[k]] = (int64_t)k;
    for (uint32_t i = 0; i < n; ++i) {
        const triangle* ign = nullptr;
        if (ignore_tri[i] != NBL_NO_TRIANGLE) {
            if (!g->base) return -1;
            const int64_t mk = ignore_tri[i] < input_to_morton.size() ? input_to_morton[ignore_tri[i]] : -1;
            if (mk >= 0) ign = g->base + mk;
        }
        intersect_event e = g->oct.propagate({ pos[i].x, pos[i].y, pos[i].z }, { dir[i].x, dir[i].y, dir[i].z },
            dist[i], ign, ignore_mat[i]);
        out_dist[i] = e.isect_distance;
        out_tri[i] = e.isect_triangle ? g->morton_to_input[g->key_to_morton.at(tri_key(*e.isect_triangle))]
                                      : NBL_NO_TRIANGLE;
    }
    return 0;
}

void ref_trilist_propagate_batch(const nbl_triangle* tris, uint32_t ntri, const nbl_vec3* pos,
    const nbl_vec3* dir, const float* dist, const uint32_t* ignore_tri, const int32_t* ignore_mat,
    uint32_t n, float* out_dist, uint32_t* out_tri)
{
    std::vector<triangle> v(ntri, triangle({ 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 }, 0, 0));
    for (uint32_t i = 0; i < ntri; ++i) std::memcpy(&v[i], &tris[i], sizeof(nbl_triangle));
    auto tl = nbl::geometry::trilist<false>::create(v);
    /* base: the trilist keeps input order, so content of any hit gives it */
    std::unordered_map<std::vector<uint32_t>, uint32_t, key_hash> in_map;
    for (uint32_t i = ntri; i-- > 0;) in_map[tri_key(v[i])] = i;
    const triangle* base = nullptr;
    for (uint32_t i = 0; i < n; ++i) {
        const triangle* ign = nullptr;
        if (ignore_tri[i] != NBL_NO_TRIANGLE && base) ign = base + ignore_tri[i];
        intersect_event e = tl.propagate({ pos[i].x, pos[i].y, pos[i].z }, { dir[i].x, dir[i].y, dir[i].z },
            dist[i], ign, ignore_mat[i]);
        if (ignore_tri[i] != NBL_NO_TRIANGLE && !base) {
            /* find the base with an unfiltered call first, then redo */
            intersect_event probe = tl.propagate({ pos[i].x, pos[i].y, pos[i].z }, { dir[i].x, dir[i].y, dir[i].z },
                dist[i], nullptr, 0x7fff0000);
            if (probe.isect_triangle) {
                base = probe.isect_triangle - in_map.at(tri_key(*probe.isect_triangle));
                ign = base + ignore_tri[i];
                e = tl.propagate({ pos[i].x, pos[i].y, pos[i].z }, { dir[i].x, dir[i].y, dir[i].z },
                    dist[i], ign, ignore_mat[i]);
            }
        }
        out_dist[i] = e.isect_distance;
        out_tri[i] = e.isect_triangle ? in_map.at(tri_key(*e.isect_triangle)) : NBL_NO_TRIANGLE;
    }
    /* trilist<false>::destroy does not compile in the reference
     * (trilist.inl:41 calls a factory member that does not exist); the
     * harness frees nothing here. */
}

void ref_intersect_ray_batch(const nbl_triangle* tris, const nbl_vec3* pos, const nbl_vec3* dir,
    uint32_t n, float* out_t)
{
    for (uint32_t i = 0; i < n; ++i) {
        triangle t({ 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 }, 0, 0);
        std::memcpy(&t, &tris[i], sizeof(nbl_triangle));
        out_t[i] = t.intersect_ray({ pos[i].x, pos[i].y, pos[i].z }, { dir[i].x, dir[i].y, dir[i].z });
    }
}

void ref_table1d_get_batch(int32_t nx, float x_min, float x_max, const float* data,
    const float* x, uint32_t n, float* out)
{
    std::vector<float> d(data, data + nx);
    auto t = nbl::util::table_1D<float, false>::create(x_min, x_max, nx, d.data());
    for (uint32_t i = 0; i < n; ++i) out[i] = t.get(x[i]);
    nbl::util::table_1D<float, false>::destroy(t);
}

void ref_table2d_get_batch(int32_t w, int32_t h, float x_min, float x_max, float y_min, float y_max,
    const float* data, int32_t rounddown, const float* x, const float* y, uint32_t n, float* out)
{
    std::vector<float> d(data, data + (size_t)w * h);
    auto t = nbl::util::table_2D<float, false>::create(x_min, x_max, w, y_min, y_max, h, d.data());
    for (uint32_t i = 0; i < n; ++i) out[i] = rounddown ? t.get_rounddown(x[i], y[i]) : t.get(x[i], y[i]);
    nbl::util::table_2D<float, false>::destroy(t);
}

void ref_table3d_get_batch(int32_t w, int32_t h, int32_t dp, float x_min, float x_max, float y_min, float y_max,
    float z_min, float z_max, const float* data, int32_t rounddown, const float* x, const float* y, const float* z,
    uint32_t n, float* out)
{
    std::vector<float> d(data, data + (size_t)w * h * dp);
    auto t = nbl::util::table_3D<float, false>::create(x_min, x_max, w, y_min, y_max, h, z_min, z_max, dp, d.data());
    for (uint32_t i = 0; i < n; ++i)
        out[i] = rounddown ? t.get_rounddown(x[i], y[i], z[i]) : t.get(x[i], y[i], z[i]);
    nbl::util::table_3D<float, false>::destroy(t);
}

/* ---- io/load_tri_file.cpp, io/load_pri_file.cpp ---- */
/* load_tri_file; returns the count, fills out (if non-NULL) with r0, e1, e2
 * as the reference's triangle stores them.  -1 if the reference throws. */
int64_t ref_tri_file_load(const char* path, nbl_triangle* out)
{
    try {
        const std::vector<triangle> t = nbl::load_tri_file(path);
        // the reference's triangle is layout-identical to nbl_triangle
        // (material_in, material_out, r0, e1, e2; core/triangle.h:69-76)
        static_assert(sizeof(triangle) == sizeof(nbl_triangle), "triangle layout");
        if (out && !t.empty()) std::memcpy(out, t.data(), t.size() * sizeof(nbl_triangle));
        return (int64_t)t.size();
    } catch (...) {
        return -1;
    }
}

/* load_pri_file (+ sort_pri_file if sort, + prescan_shuffle(prescan)); returns the count. */
int64_t ref_pri_file_load(const char* path, const float* lo, const float* hi, float max_e, int sort, uint64_t prescan,
    nbl_particle* out, int32_t* pixels)
{
    auto pr = nbl::load_pri_file(path, vec3{ lo[0], lo[1], lo[2] }, vec3{ hi[0], hi[1], hi[2] }, max_e);
    if (sort && !pr.first.empty()) nbl::sort_pri_file(pr.first, pr.second);
    if (prescan) nbl::prescan_shuffle(pr.first, pr.second, prescan);
    for (size_t i = 0; i < pr.first.size(); ++i) {
        const particle& p = pr.first[i];
        if (out) out[i] = { { p.pos.x, p.pos.y, p.pos.z }, { p.dir.x, p.dir.y, p.dir.z }, p.kin_energy };
        if (pixels) { pixels[2 * i] = pr.second[i].x; pixels[2 * i + 1] = pr.second[i].y; }
    }
    return (int64_t)pr.first.size();
}

}  /* extern "C" */
