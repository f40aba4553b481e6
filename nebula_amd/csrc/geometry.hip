// geometry.hip -- nbl_geometry_*: octree upload, batched in_domain and
// propagate kernels, and the host-only linearization entry point.
#include <memory>

#include "nbl_host.h"
#include "octree_build.h"
#include "wave_propagate.h"

using namespace nbl;

struct nbl_geometry_s {
    built_octree host;
    dbuf<uint64_t> nodes;
    dbuf<tri_rec> recs;
    dbuf<leaf_info> leaves;
    dbuf<leaf_rope> ropes;
    dbuf<uint32_t> id_rec;
    dgeom view{};
    uint32_t n_in = 0;
    std::vector<uint8_t> has_rec;    // lazily built: id has a record
};

namespace nbl {
const dgeom& geometry_view(nbl_geometry g) { return g->view; }
const tri_rec* geometry_host_recs(nbl_geometry g) { return g->host.recs.data(); }
uint32_t geometry_n_input(nbl_geometry g) { return g->n_in; }
// true if triangle id has a device record (not dropped as a duplicate)
bool geometry_has_record(nbl_geometry g, uint32_t id)
{
    if (g->has_rec.empty()) {
        g->has_rec.assign(g->n_in, 0);
        for (const tri_rec& r : g->host.recs) g->has_rec[r.id] = 1;
    }
    return id < g->n_in && g->has_rec[id];
}
// largest material index used by any kept triangle (-1 if none is physical)
int geometry_max_material(nbl_geometry g)
{
    int mx = -1;
    for (const tri_rec& r : g->host.recs) {
        mx = std::max(mx, (int)(int16_t)(r.mats & 0xFFFFu));
        mx = std::max(mx, (int)(int16_t)(r.mats >> 16));
    }
    return mx;
}
}

namespace {

// One ray per lane, grid-stride by wavefronts, with the wavefront walk of the
// transport kernel (shared leaf tests, per-lane LDS path memo).  Inputs and
// outputs are device arrays.
__global__ __launch_bounds__(256) void k_propagate(dgeom g, const f3* __restrict__ pos,
    const f3* __restrict__ dir, const float* __restrict__ dist, const uint32_t* __restrict__ ignore_tri,
    const int32_t* __restrict__ ignore_mat, uint32_t n, float* __restrict__ out_dist,
    uint32_t* __restrict__ out_tri)
{
    __shared__ uint32_t memo[4 * lds_path_cache::kWords * 64];
    lds_path_cache pc{ memo + (threadIdx.x >> 6) * lds_path_cache::kWords * 64 + (threadIdx.x & 63), 1ull, 0ull, 0u };
    // wave-uniform loop: every lane of a wavefront enters propagate_wave together
    for (uint32_t w0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); w0 < n; w0 += gridDim.x * blockDim.x) {
        const uint32_t i = w0 + (threadIdx.x & 63u);
        const bool on = i < n;
        const uint32_t j = on ? i : w0;
        uint32_t hit;
        const float d = propagate_wave<false>(on, g, pos[j], dir[j], dist[j], ignore_tri[j], ignore_mat[j], &hit,
                                              nullptr, &pc);
        if (on) { out_dist[i] = d; out_tri[i] = hit; }
    }
}

__global__ __launch_bounds__(256) void k_in_domain(dgeom g, const f3* __restrict__ pos, uint32_t n,
    uint8_t* __restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = in_domain(g, pos[i]) ? 1 : 0;
}

void launch_propagate(nbl_geometry g, const f3* pos, const f3* dir, const float* dist,
    const uint32_t* it, const int32_t* im, uint32_t n, float* od, uint32_t* ot, hipStream_t s)
{
    if (n == 0) return;
    // at most 4 resident blocks per CU-sized slice: lanes walk several rays each
    hipLaunchKernelGGL(k_propagate, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, g->view, pos, dir, dist, it, im, n, od, ot);
    NBL_HIP(hipGetLastError());
}

}  // namespace

extern "C" {

int nbl_geometry_create(const nbl_triangle* tris, uint32_t n, nbl_geometry* out)
{
    return guarded([&] {
        require(out != nullptr, "nbl_geometry_create: out is NULL");
        require(tris != nullptr && n > 0, "nbl_geometry_create: no triangles");
        *out = nullptr;
        ensure_device();
        auto g = std::make_unique<nbl_geometry_s>();
        g->host = build_octree(tris, n);
        g->n_in = n;
        g->nodes.alloc(g->host.nodes.size());
        g->nodes.upload(g->host.nodes.data(), g->host.nodes.size());
        g->recs.alloc(std::max<size_t>(g->host.recs.size(), 1));
        g->recs.upload(g->host.recs.data(), g->host.recs.size());
        g->leaves.alloc(std::max<size_t>(g->host.leaves.size(), 1));
        g->leaves.upload(g->host.leaves.data(), g->host.leaves.size());
        g->ropes.alloc(std::max<size_t>(g->host.ropes.size(), 6));
        g->ropes.upload(g->host.ropes.data(), g->host.ropes.size());
        // one record per input id (NO_TRI for triangles dropped as duplicates)
        std::vector<uint32_t> id_rec(n, NO_TRI);
        for (size_t k = g->host.recs.size(); k-- > 0;) id_rec[g->host.recs[k].id] = (uint32_t)k;
        g->id_rec.alloc(n);
        g->id_rec.upload(id_rec.data(), n);
        NBL_HIP(hipDeviceSynchronize());
        dgeom& v = g->view;
        v.nodes = g->nodes.p; v.recs = g->recs.p; v.leaves = g->leaves.p; v.ropes = g->ropes.p; v.n_ropes = (uint32_t)std::max<size_t>(g->host.ropes.size(), 6); v.id_rec = g->id_rec.p;
        v.owner_payload = g->host.recs.size() <= ((size_t)1 << OWNER_PAYLOAD_BITS) ? 1u : 0u;
        v.cx = g->host.center[0]; v.cy = g->host.center[1]; v.cz = g->host.center[2];
        v.hx = g->host.halfsize[0]; v.hy = g->host.halfsize[1]; v.hz = g->host.halfsize[2];
        v.mx = ldexpf(std::fabs(v.cx) + v.hx, -16);
        v.my = ldexpf(std::fabs(v.cy) + v.hy, -16);
        v.mz = ldexpf(std::fabs(v.cz) + v.hz, -16);
        set_derived(v);
        v.max_extent = g->host.max_extent;
        *out = g.release();
    });
}

int nbl_geometry_destroy(nbl_geometry g)
{
    return guarded([&] { delete g; });
}

int nbl_geometry_get_info(nbl_geometry g, nbl_geometry_info* info)
{
    return guarded([&] {
        require(g && info, "nbl_geometry_get_info: NULL argument");
        const built_octree& h = g->host;
        info->aabb_min = { h.center[0] - h.halfsize[0], h.center[1] - h.halfsize[1], h.center[2] - h.halfsize[2] };
        info->aabb_max = { h.center[0] + h.halfsize[0], h.center[1] + h.halfsize[1], h.center[2] + h.halfsize[2] };
        info->max_extent = h.max_extent;
        info->n_triangles_in = g->n_in;
        info->n_triangles_unique = h.n_unique;
        info->n_nodes = (uint32_t)(h.nodes.size() / 8);
        info->n_leaves = h.n_leaves;
        info->n_leaf_refs = (uint32_t)h.recs.size();
        info->max_depth = h.max_depth;
        info->linearized_size = (uint64_t)info->n_nodes * 8 + h.recs.size() + h.n_leaves;
        info->device_bytes = g->nodes.bytes() + g->recs.bytes() + g->leaves.bytes() + g->ropes.bytes();
    });
}

int nbl_geometry_export_linearized(nbl_geometry g, int32_t* octree_data, uint32_t* triangle_order)
{
    return guarded([&] {
        require(g != nullptr, "nbl_geometry_export_linearized: NULL geometry");
        std::vector<int32_t> d; std::vector<uint32_t> o;
        export_linearized(g->host, d, o);
        if (octree_data) std::memcpy(octree_data, d.data(), d.size() * sizeof(int32_t));
        if (triangle_order) std::memcpy(triangle_order, o.data(), o.size() * sizeof(uint32_t));
    });
}

int nbl_octree_linearize(const nbl_triangle* tris, uint32_t n, int32_t* octree_data, uint64_t* octree_size,
    uint32_t* triangle_order, uint32_t* n_unique, float* center3, float* halfsize3)
{
    return guarded([&] {
        require(tris != nullptr && n > 0, "nbl_octree_linearize: no triangles");
        const built_octree o = build_octree(tris, n);
        std::vector<int32_t> d; std::vector<uint32_t> ord;
        export_linearized(o, d, ord);
        if (octree_size) *octree_size = d.size();
        if (n_unique) *n_unique = (uint32_t)ord.size();
        if (octree_data) std::memcpy(octree_data, d.data(), d.size() * sizeof(int32_t));
        if (triangle_order) std::memcpy(triangle_order, ord.data(), ord.size() * sizeof(uint32_t));
        for (int a = 0; a < 3; ++a) {
            if (center3) center3[a] = o.center[a];
            if (halfsize3) halfsize3[a] = o.halfsize[a];
        }
    });
}

int nbl_geometry_in_domain_batch(nbl_geometry g, const nbl_vec3* pos, uint32_t n, uint8_t* out)
{
    return guarded([&] {
        require(g && (n == 0 || (pos && out)), "nbl_geometry_in_domain_batch: NULL argument");
        if (n == 0) return;
        dbuf<f3> dp(n); dbuf<uint8_t> dout(n);
        dp.upload(reinterpret_cast<const f3*>(pos), n);
        hipLaunchKernelGGL(k_in_domain, dim3(grid_for(n, 256)), dim3(256), 0, 0, g->view, dp.p, n, dout.p);
        NBL_HIP(hipGetLastError());
        dout.download(out, n);
        NBL_HIP(hipDeviceSynchronize());
    });
}

int nbl_geometry_propagate_batch(nbl_geometry g, const nbl_vec3* pos, const nbl_vec3* dir, const float* dist,
    const uint32_t* ignore_tri, const int32_t* ignore_mat, uint32_t n, float* out_dist, uint32_t* out_tri)
{
    return guarded([&] {
        require(g != nullptr, "nbl_geometry_propagate_batch: NULL geometry");
        require(n == 0 || (pos && dir && dist && ignore_tri && ignore_mat && out_dist && out_tri),
            "nbl_geometry_propagate_batch: NULL array");
        if (n == 0) return;
        dbuf<f3> dp(n), dd(n); dbuf<float> ds(n), od(n); dbuf<uint32_t> di(n), ot(n); dbuf<int32_t> dm(n);
        dp.upload(reinterpret_cast<const f3*>(pos), n);
        dd.upload(reinterpret_cast<const f3*>(dir), n);
        ds.upload(dist, n); di.upload(ignore_tri, n); dm.upload(ignore_mat, n);
        launch_propagate(g, dp.p, dd.p, ds.p, di.p, dm.p, n, od.p, ot.p, 0);
        od.download(out_dist, n); ot.download(out_tri, n);
        NBL_HIP(hipDeviceSynchronize());
    });
}

int nbl_geometry_propagate_batch_device(nbl_geometry g, const nbl_vec3* pos, const nbl_vec3* dir,
    const float* dist, const uint32_t* ignore_tri, const int32_t* ignore_mat, uint32_t n,
    float* out_dist, uint32_t* out_tri)
{
    return guarded([&] {
        require(g != nullptr, "nbl_geometry_propagate_batch_device: NULL geometry");
        launch_propagate(g, reinterpret_cast<const f3*>(pos), reinterpret_cast<const f3*>(dir), dist,
            ignore_tri, ignore_mat, n, out_dist, out_tri, 0);
    });
}

}  // extern "C"
