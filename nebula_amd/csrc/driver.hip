// driver.hip -- nbl_driver_*: the gpu_driver replacement (drivers/gpu/
// gpu_driver.inl), its particle slots (gpu_particle_manager.inl), the
// in-house stable status sort, and the XORWOW entry points.
//
// One iteration = init kernel (slot order) -> stable sort of status bits
// [0,2) -> intersect+elastic kernel -> inelastic kernel.  Event kernels run
// over the sorted index and touch only their own status bucket; free slots
// (TERMINATED, bucket 3) sit at the end of the index, so thread i of the
// inelastic kernel may put a secondary into slot pidx[cap-1-i] if that slot
// is free, and otherwise leaves its particle PENDING for the next iteration
// (gpu_driver.inl:506-541, gpu_particle_manager.inl:354-376).
#include <algorithm>
#include <memory>

#include "driver.h"

using namespace nbl;

namespace {

constexpr uint8_t ST_PENDING = NBL_ST_PENDING, ST_INELASTIC = NBL_ST_INELASTIC_EVENT,
                  ST_ELASTIC = NBL_ST_ELASTIC_EVENT, ST_INTERSECT = NBL_ST_INTERSECT_EVENT,
                  ST_NO_EVENT = NBL_ST_NO_EVENT, ST_DETECTED = NBL_ST_DETECTED,
                  ST_NEW_SECONDARY = NBL_ST_NEW_SECONDARY, ST_TERMINATED = NBL_ST_TERMINATED;

__device__ __forceinline__ bool active(uint8_t s)   // gpu_particle_manager.inl:271-283
{
    return !(s == ST_TERMINATED || s == ST_DETECTED || s == ST_PENDING);
}

// ---- sort ------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sort_count(const uint8_t* __restrict__ keys, uint32_t n,
    uint32_t n_blocks, uint32_t* __restrict__ block_counts)
{
    __shared__ uint32_t cnt[4];
    if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * (sorter::kBlock * sorter::kItems);
    uint32_t c[4] = { 0, 0, 0, 0 };
    for (uint32_t k = 0; k < sorter::kItems; ++k) {
        const uint32_t i = base + k * sorter::kBlock + threadIdx.x;
        if (i < n) {
            const uint32_t b = keys[i] & 3u;
            c[0] += b == 0; c[1] += b == 1; c[2] += b == 2; c[3] += b == 3;
        }
    }
    for (int b = 0; b < 4; ++b) if (c[b]) atomicAdd(&cnt[b], c[b]);
    __syncthreads();
    if (threadIdx.x < 4) block_counts[threadIdx.x * n_blocks + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of the bucket-major counts (one block, serial over chunks);
// bucket[b] receives the size of bucket b.
__global__ __launch_bounds__(1024) void k_sort_scan(uint32_t* __restrict__ block_counts, uint32_t n_blocks,
    uint32_t* __restrict__ bucket)
{
    __shared__ uint32_t tmp[1024];
    __shared__ uint32_t carry;
    const uint32_t total = 4 * n_blocks;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t start = 0; start < total; start += 1024) {
        const uint32_t i = start + threadIdx.x;
        const uint32_t v = i < total ? block_counts[i] : 0;
        tmp[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024; off <<= 1) {     // Hillis-Steele inclusive scan
            const uint32_t add = threadIdx.x >= off ? tmp[threadIdx.x - off] : 0;
            __syncthreads();
            tmp[threadIdx.x] += add;
            __syncthreads();
        }
        const uint32_t c = carry;
        if (i < total) block_counts[i] = c + tmp[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c + tmp[1023];
        __syncthreads();
    }
    if (threadIdx.x < 4) {
        // size of bucket b = start of bucket b+1 - start of bucket b
        const uint32_t s = block_counts[threadIdx.x * n_blocks];
        const uint32_t e = threadIdx.x < 3 ? block_counts[(threadIdx.x + 1) * n_blocks] : carry;
        bucket[threadIdx.x] = e - s;
    }
}

// Stable scatter: tiles of 256 consecutive keys; ranks from per-bucket
// wavefront ballots, wave totals combined in LDS, running offsets per bucket.
__global__ __launch_bounds__(256) void k_sort_scatter(const uint8_t* __restrict__ keys, uint32_t n,
    uint32_t n_blocks, const uint32_t* __restrict__ block_offsets, uint32_t* __restrict__ out)
{
    __shared__ uint32_t wave_cnt[4][4];     // [wave][bucket]
    __shared__ uint32_t running[4];
    const uint32_t wave = threadIdx.x >> 6;
    if (threadIdx.x < 4) running[threadIdx.x] = block_offsets[threadIdx.x * n_blocks + blockIdx.x];
    __syncthreads();
    const uint32_t base = blockIdx.x * (sorter::kBlock * sorter::kItems);
    for (uint32_t k = 0; k < sorter::kItems; ++k) {
        const uint32_t i = base + k * sorter::kBlock + threadIdx.x;
        const uint32_t b = i < n ? (keys[i] & 3u) : 4u;
        const uint64_t m0 = __ballot(b == 0), m1 = __ballot(b == 1), m2 = __ballot(b == 2), m3 = __ballot(b == 3);
        const uint64_t mine = b == 0 ? m0 : b == 1 ? m1 : b == 2 ? m2 : m3;
        const uint32_t l = lane_id();
        if (l < 4) wave_cnt[wave][l] = __popcll(l == 0 ? m0 : l == 1 ? m1 : l == 2 ? m2 : m3);
        __syncthreads();
        if (b < 4) {
            uint32_t pos = running[b] + rank_below(mine);
            for (uint32_t w = 0; w < wave; ++w) pos += wave_cnt[w][b];
            out[pos] = i;
        }
        __syncthreads();
        if (threadIdx.x < 4)
            running[threadIdx.x] += wave_cnt[0][threadIdx.x] + wave_cnt[1][threadIdx.x] +
                                    wave_cnt[2][threadIdx.x] + wave_cnt[3][threadIdx.x];
        __syncthreads();
    }
}

// ---- slot kernels ------------------------------------------------------------
__global__ void k_seed_rng(rng_state* __restrict__ rng, uint32_t n, uint64_t seed, uint64_t first)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        rng[i] = rng_seed(seed, first + i);
}

__global__ void k_rng_uniform(rng_state* __restrict__ st, uint32_t n, uint32_t draws, float* __restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        rng_state s = st[i];
        for (uint32_t k = 0; k < draws; ++k) out[(size_t)i * draws + k] = rng_unit(s);
        st[i] = s;
    }
}

__global__ void k_fill_slots(dslots s)   // gpu_particle_manager.inl:26-63
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.cap; i += gridDim.x * blockDim.x) {
        s.status[i] = ST_TERMINATED; s.pidx[i] = i; s.tags[i] = 0;
        s.mat[i] = MAT_VACUUM; s.last_tri[i] = NO_TRI;
        s.parts[i] = dparticle{ { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    }
}

// kernels::init (gpu_driver.inl:403-481), slot order
__global__ __launch_bounds__(256) void k_init(dslots s, dgeom g, const dmat* __restrict__ mats, float min_e, float max_e)
{
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= s.cap || !active(s.status[idx])) return;
    dparticle p = s.parts[idx];
    const int mat = s.mat[idx];
    uint32_t last = s.last_tri[idx];
    rng_xorwow r{ s.rng[idx] };
    lane_counters c;
    const uint8_t ev = init_step<false>(g, mats, min_e, max_e, p, mat, last, r, c);
    if (ev == EV_TERMINATED) { s.status[idx] = ST_TERMINATED; return; }
    if (mat >= 0) s.rng[idx] = r.s;
    s.parts[idx] = p;
    if (ev == EV_INTERSECT) { s.status[idx] = ST_INTERSECT; s.last_tri[idx] = last; }
    else s.status[idx] = ev == EV_INELASTIC ? ST_INELASTIC : ev == EV_ELASTIC ? ST_ELASTIC : ST_NO_EVENT;
}

// kernels::intersect + kernels::elastic (gpu_driver.inl:483-504, 543-567)
// over the elastic and intersect buckets of the sorted index
__global__ __launch_bounds__(256) void k_events_a(dslots s, dgeom g, const dmat* __restrict__ mats)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lo = s.bucket[0], hi = s.bucket[0] + s.bucket[1] + s.bucket[2];
    if (i < lo || i >= hi) return;
    const uint32_t idx = s.pidx[i];
    const uint8_t st = s.status[idx];
    if (st == ST_INTERSECT) {
        dparticle p = s.parts[idx];
        int mat = s.mat[idx];
        rng_xorwow r{ s.rng[idx] };
        const tri_rec t = load_tri(g, s.last_tri[idx]);
        const int res = intersect_execute(mats, t, p, mat, r);
        s.rng[idx] = r.s;
        if (res == OUT_DETECTED) s.status[idx] = ST_DETECTED;
        else if (res == OUT_TERMINATED) s.status[idx] = ST_TERMINATED;
        else { s.parts[idx] = p; s.mat[idx] = mat; }
    } else if (st == ST_ELASTIC) {
        s.last_tri[idx] = NO_TRI;
        dparticle p = s.parts[idx];
        rng_xorwow r{ s.rng[idx] };
        elastic_execute(mats[s.mat[idx]], p, r);
        s.rng[idx] = r.s;
        s.parts[idx] = p;
    }
}

// kernels::inelastic (gpu_driver.inl:506-541) over the inelastic bucket
__global__ __launch_bounds__(256) void k_events_b(dslots s, const dmat* __restrict__ mats)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.bucket[0]) return;
    const uint32_t idx = s.pidx[i];
    const uint32_t sec = s.pidx[s.cap - 1 - i];
    if (s.status[sec] != ST_TERMINATED) { s.status[idx] = ST_PENDING; return; }
    s.status[idx] = ST_INELASTIC;
    s.last_tri[idx] = NO_TRI;
    dparticle p = s.parts[idx];
    const int mat = s.mat[idx];
    rng_xorwow r{ s.rng[idx] };
    dparticle q;
    const bool made = inelastic_event(mats[mat], p, r, &q);
    s.rng[idx] = r.s;
    s.parts[idx] = p;
    if (made) {   // create_secondary (gpu_particle_manager.inl:359-369)
        s.status[sec] = ST_NEW_SECONDARY;
        s.parts[sec] = q;
        s.tags[sec] = s.tags[idx];
        s.mat[sec] = mat;
        s.last_tri[sec] = NO_TRI;
    }
}

// kernels::push_buffer_in (gpu_driver.inl:367-389)
__global__ void k_push_buffer_in(dslots s, uint8_t* __restrict__ data, const dparticle* __restrict__ parts,
    const uint32_t* __restrict__ tags, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || i >= s.cap) return;
    if (data[i] != 1) return;
    const uint32_t idx = s.pidx[s.cap - i - 1];
    if (s.status[idx] != ST_TERMINATED) return;
    s.status[idx] = ST_NO_EVENT; s.parts[idx] = parts[i]; s.tags[idx] = tags[i];
    s.mat[idx] = MAT_VACUUM; s.last_tri[idx] = NO_TRI;
    data[i] = 0;
}

// kernels::terminate_detected (gpu_driver.inl:391-401)
__global__ void k_terminate_detected(dslots s)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.cap && s.status[i] == ST_DETECTED) s.status[i] = ST_TERMINATED;
}

// writes pushed particles into the given free slots (gpu_particle_manager.inl:129-160)
__global__ void k_place(dslots s, const uint32_t* __restrict__ slots, const dparticle* __restrict__ parts,
    const uint32_t* __restrict__ tags, uint32_t n)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = slots[k];
    s.status[i] = ST_NO_EVENT; s.parts[i] = parts[k]; s.tags[i] = tags[k];
    s.mat[i] = MAT_VACUUM; s.last_tri[i] = NO_TRI;
}

template<typename T> void host_alloc(T** p, size_t n)
{
    NBL_HIP(hipHostMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault));
}
template<typename T> void host_free(T*& p)
{
    if (p) (void)hipHostFree(p);
    p = nullptr;
}

void launch_iteration(nbl_driver d)
{
    const dgeom& g = geometry_view(d->geom);
    const dmat* mats = materials_device(d->mats);
    const dim3 grid((d->cap + 255) / 256), block(256);
    hipLaunchKernelGGL(k_init, grid, block, 0, d->stream, d->view(), g, mats, d->min_energy, d->max_energy);
    d->sort.run(d->status.p, d->cap, d->pidx.p, d->stream);
    hipLaunchKernelGGL(k_events_a, grid, block, 0, d->stream, d->view(), g, mats);
    hipLaunchKernelGGL(k_events_b, grid, block, 0, d->stream, d->view(), mats);
    NBL_HIP(hipGetLastError());
}

nbl_driver checked(nbl_driver d, const char* fn)
{
    if (!d) throw invalid_arg(std::string(fn) + ": NULL driver");
    return d;
}

}  // namespace

namespace nbl {

void sorter::init(uint32_t n)
{
    n_blocks = std::max(1u, (unsigned)((n + kBlock * kItems - 1) / (kBlock * kItems)));
    block_counts.alloc((size_t)4 * n_blocks);
    bucket.alloc(4);
}

void sorter::run(const uint8_t* keys, uint32_t n, uint32_t* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sort_count, dim3(n_blocks), dim3(kBlock), 0, s, keys, n, n_blocks, block_counts.p);
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, s, block_counts.p, n_blocks, bucket.p);
    hipLaunchKernelGGL(k_sort_scatter, dim3(n_blocks), dim3(kBlock), 0, s, keys, n, n_blocks, block_counts.p, out);
    NBL_HIP(hipGetLastError());
}

}  // namespace nbl

nbl_driver_s::~nbl_driver_s()
{
    if (stream) (void)hipStreamSynchronize(stream);
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    host_free(hin_data); host_free(hin_parts); host_free(hin_tags);
    host_free(hout_status); host_free(hout_parts); host_free(hout_tags);
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    if (stream) (void)hipStreamDestroy(stream);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
}

extern "C" {

int nbl_driver_create(uint32_t capacity, nbl_geometry g, nbl_materials m, float min_energy, float max_energy,
    uint64_t seed, const nbl_driver_opts* opts, nbl_driver* out)
{
    return guarded([&] {
        require(out != nullptr, "nbl_driver_create: out is NULL");
        *out = nullptr;
        require(g && m, "nbl_driver_create: NULL geometry or materials");
        require(capacity > 0 && capacity <= (1u << 30), "nbl_driver_create: capacity must be in [1, 2^30]");
        // gpu_main.cu:300-322: every triangle material must exist
        require(geometry_max_material(g) < materials_count(m),
            "nbl_driver_create: geometry uses a material index beyond the material count");
        ensure_device();
        auto d = std::make_unique<nbl_driver_s>();
        d->cap = capacity; d->geom = g; d->mats = m;
        d->min_energy = min_energy; d->max_energy = max_energy; d->seed = seed;
        if (opts && opts->threads_per_block) {
            require(opts->threads_per_block % 64 == 0 && opts->threads_per_block <= 1024,
                "nbl_driver_create: threads_per_block must be a multiple of 64 up to 1024");
            d->tpb = opts->threads_per_block;
        }
        NBL_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
        NBL_HIP(hipStreamCreateWithFlags(&d->copy_stream, hipStreamNonBlocking));
        NBL_HIP(hipEventCreate(&d->ev_start));
        NBL_HIP(hipEventCreate(&d->ev_stop));
        d->status.alloc(capacity); d->pidx.alloc(capacity); d->parts.alloc(capacity);
        d->tags.alloc(capacity); d->mat.alloc(capacity); d->last_tri.alloc(capacity); d->rng.alloc(capacity);
        d->sort.init(capacity);
        d->dout_status.alloc(capacity); d->dout_parts.alloc(capacity); d->dout_tags.alloc(capacity);
        host_alloc(&d->hout_status, capacity); host_alloc(&d->hout_parts, capacity); host_alloc(&d->hout_tags, capacity);
        const unsigned grid = grid_for(capacity, 256);
        hipLaunchKernelGGL(k_fill_slots, dim3(grid), dim3(256), 0, d->stream, d->view());
        hipLaunchKernelGGL(k_seed_rng, dim3(grid), dim3(256), 0, d->stream, d->rng.p, capacity, seed, (uint64_t)0);
        NBL_HIP(hipGetLastError());
        NBL_HIP(hipMemsetAsync(d->sort.bucket.p, 0, 4 * sizeof(uint32_t), d->stream));
        NBL_HIP(hipStreamSynchronize(d->stream));
        nbl_driver h = d.release();
        if (nbl_driver_buffer_detected(h) != NBL_OK) { delete h; throw hip_error("nbl_driver_create: buffer_detected failed"); }
        *out = h;
    });
}

int nbl_driver_destroy(nbl_driver d)
{
    return guarded([&] { delete d; });
}

int nbl_driver_push(nbl_driver d, const nbl_particle* particles, const uint32_t* tags, uint32_t n, uint32_t* pushed)
{
    return guarded([&] {
        checked(d, "nbl_driver_push");
        require(n == 0 || (particles && tags), "nbl_driver_push: NULL array");
        std::vector<uint8_t> st(d->cap);
        d->status.download(st.data(), d->cap, d->stream);
        NBL_HIP(hipStreamSynchronize(d->stream));
        std::vector<uint32_t> slots;
        for (uint32_t i = 0; i < d->cap && slots.size() < n; ++i)
            if (st[i] == ST_TERMINATED) slots.push_back(i);
        const uint32_t k = (uint32_t)slots.size();
        if (k) {
            dbuf<uint32_t> ds(k), dt(k); dbuf<dparticle> dp(k);
            ds.upload(slots.data(), k, d->stream);
            dp.upload(reinterpret_cast<const dparticle*>(particles), k, d->stream);
            dt.upload(tags, k, d->stream);
            hipLaunchKernelGGL(k_place, dim3((k + 255) / 256), dim3(256), 0, d->stream, d->view(), ds.p, dp.p, dt.p, k);
            NBL_HIP(hipGetLastError());
            NBL_HIP(hipStreamSynchronize(d->stream));
        }
        if (pushed) *pushed = k;
    });
}

int nbl_driver_do_iteration(nbl_driver d)
{
    return guarded([&] {
        checked(d, "nbl_driver_do_iteration");
        launch_iteration(d);
        NBL_HIP(hipStreamSynchronize(d->stream));
    });
}

static uint32_t count_status(nbl_driver d, bool detected)
{
    std::vector<uint8_t> st(d->cap);
    d->status.download(st.data(), d->cap, d->stream);
    NBL_HIP(hipStreamSynchronize(d->stream));
    uint32_t n = 0;
    for (uint8_t s : st) n += detected ? (s == ST_DETECTED) : (s != ST_TERMINATED && s != ST_DETECTED);
    return n;
}

int nbl_driver_get_running_count(nbl_driver d, uint32_t* out)
{
    return guarded([&] {
        checked(d, "nbl_driver_get_running_count");
        require(out != nullptr, "nbl_driver_get_running_count: out is NULL");
        *out = count_status(d, false);
    });
}

int nbl_driver_get_detected_count(nbl_driver d, uint32_t* out)
{
    return guarded([&] {
        checked(d, "nbl_driver_get_detected_count");
        require(out != nullptr, "nbl_driver_get_detected_count: out is NULL");
        *out = count_status(d, true);
    });
}

int nbl_driver_flush_detected(nbl_driver d, nbl_detect_fn fn, void* user)
{
    return guarded([&] {
        checked(d, "nbl_driver_flush_detected");
        std::vector<uint8_t> st(d->cap);
        std::vector<dparticle> ps(d->cap);
        std::vector<uint32_t> tg(d->cap);
        d->status.download(st.data(), d->cap, d->stream);
        d->parts.download(ps.data(), d->cap, d->stream);
        d->tags.download(tg.data(), d->cap, d->stream);
        NBL_HIP(hipStreamSynchronize(d->stream));
        for (uint32_t i = 0; i < d->cap; ++i)
            if (st[i] == ST_DETECTED && fn) fn(reinterpret_cast<const nbl_particle*>(&ps[i]), tg[i], user);
        hipLaunchKernelGGL(k_terminate_detected, dim3((d->cap + 255) / 256), dim3(256), 0, d->stream, d->view());
        NBL_HIP(hipGetLastError());
        NBL_HIP(hipStreamSynchronize(d->stream));
    });
}

int nbl_driver_allocate_input_buffers(nbl_driver d, uint32_t n)
{
    return guarded([&] {
        checked(d, "nbl_driver_allocate_input_buffers");
        NBL_HIP(hipStreamSynchronize(d->stream));
        host_free(d->hin_data); host_free(d->hin_parts); host_free(d->hin_tags);
        d->in_size = n;
        d->din_data.alloc(n); d->din_parts.alloc(n); d->din_tags.alloc(n);
        host_alloc(&d->hin_data, n); host_alloc(&d->hin_parts, n); host_alloc(&d->hin_tags, n);
        if (n) NBL_HIP(hipMemsetAsync(d->din_data.p, 0, n, d->stream));   // init_buffer_data
        NBL_HIP(hipStreamSynchronize(d->stream));
    });
}

int nbl_driver_push_to_buffer(nbl_driver d, nbl_pool_fill_fn fill, void* user)
{
    return guarded([&] {
        checked(d, "nbl_driver_push_to_buffer");
        const uint32_t n = d->in_size;
        if (n == 0) return;
        hipStream_t s = d->copy_stream;
        NBL_HIP(hipMemcpyAsync(d->hin_data, d->din_data.p, n, hipMemcpyDeviceToHost, s));
        NBL_HIP(hipMemcpyAsync(d->hin_parts, d->din_parts.p, n * sizeof(dparticle), hipMemcpyDeviceToHost, s));
        NBL_HIP(hipMemcpyAsync(d->hin_tags, d->din_tags.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        NBL_HIP(hipStreamSynchronize(s));
        if (fill) fill(d->hin_data, reinterpret_cast<nbl_particle*>(d->hin_parts), d->hin_tags, n, user);
        NBL_HIP(hipMemcpyAsync(d->din_data.p, d->hin_data, n, hipMemcpyHostToDevice, s));
        NBL_HIP(hipMemcpyAsync(d->din_parts.p, d->hin_parts, n * sizeof(dparticle), hipMemcpyHostToDevice, s));
        NBL_HIP(hipMemcpyAsync(d->din_tags.p, d->hin_tags, n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        NBL_HIP(hipStreamSynchronize(s));
    });
}

int nbl_driver_push_to_simulation(nbl_driver d)
{
    return guarded([&] {
        checked(d, "nbl_driver_push_to_simulation");
        d->sort.run(d->status.p, d->cap, d->pidx.p, d->stream);
        if (d->in_size) {
            hipLaunchKernelGGL(k_push_buffer_in, dim3((d->in_size + 255) / 256), dim3(256), 0, d->stream,
                d->view(), d->din_data.p, d->din_parts.p, d->din_tags.p, d->in_size);
            NBL_HIP(hipGetLastError());
        }
        NBL_HIP(hipStreamSynchronize(d->stream));
    });
}

int nbl_driver_buffer_detected(nbl_driver d)
{
    return guarded([&] {
        checked(d, "nbl_driver_buffer_detected");
        NBL_HIP(hipMemcpyAsync(d->dout_status.p, d->status.p, d->cap, hipMemcpyDeviceToDevice, d->stream));
        NBL_HIP(hipMemcpyAsync(d->dout_parts.p, d->parts.p, d->cap * sizeof(dparticle), hipMemcpyDeviceToDevice, d->stream));
        NBL_HIP(hipMemcpyAsync(d->dout_tags.p, d->tags.p, d->cap * sizeof(uint32_t), hipMemcpyDeviceToDevice, d->stream));
        hipLaunchKernelGGL(k_terminate_detected, dim3((d->cap + 255) / 256), dim3(256), 0, d->stream, d->view());
        NBL_HIP(hipGetLastError());
        NBL_HIP(hipStreamSynchronize(d->stream));
    });
}

int nbl_driver_flush_buffered(nbl_driver d, nbl_detect_fn fn, void* user, uint32_t* running)
{
    return guarded([&] {
        checked(d, "nbl_driver_flush_buffered");
        hipStream_t s = d->copy_stream;
        const uint32_t n = d->cap;
        NBL_HIP(hipMemcpyAsync(d->hout_status, d->dout_status.p, n, hipMemcpyDeviceToHost, s));
        NBL_HIP(hipMemcpyAsync(d->hout_parts, d->dout_parts.p, n * sizeof(dparticle), hipMemcpyDeviceToHost, s));
        NBL_HIP(hipMemcpyAsync(d->hout_tags, d->dout_tags.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        NBL_HIP(hipStreamSynchronize(s));
        uint32_t run = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const uint8_t st = d->hout_status[i];
            if (st != ST_TERMINATED && st != ST_DETECTED) ++run;
            if (st == ST_DETECTED && fn) fn(reinterpret_cast<const nbl_particle*>(&d->hout_parts[i]), d->hout_tags[i], user);
        }
        if (running) *running = run;
    });
}

int nbl_driver_read_slots(nbl_driver d, const nbl_slot_view* v)
{
    return guarded([&] {
        checked(d, "nbl_driver_read_slots");
        require(v != nullptr, "nbl_driver_read_slots: NULL view");
        const uint32_t n = d->cap;
        if (v->status) d->status.download(v->status, n, d->stream);
        if (v->particle_idx) d->pidx.download(v->particle_idx, n, d->stream);
        if (v->particles) d->parts.download(reinterpret_cast<dparticle*>(v->particles), n, d->stream);
        if (v->tags) d->tags.download(v->tags, n, d->stream);
        if (v->material) d->mat.download(v->material, n, d->stream);
        if (v->last_tri) d->last_tri.download(v->last_tri, n, d->stream);
        if (v->rng) d->rng.download(reinterpret_cast<rng_state*>(v->rng), n, d->stream);
        NBL_HIP(hipStreamSynchronize(d->stream));
    });
}

int nbl_driver_write_slots(nbl_driver d, const nbl_slot_view* v)
{
    return guarded([&] {
        checked(d, "nbl_driver_write_slots");
        require(v != nullptr, "nbl_driver_write_slots: NULL view");
        const uint32_t n = d->cap;
        if (v->last_tri) {
            for (uint32_t i = 0; i < n; ++i)
                require(v->last_tri[i] == NO_TRI || geometry_has_record(d->geom, v->last_tri[i]),
                    "nbl_driver_write_slots: last_tri is not a triangle of the geometry");
        }
        if (v->material) {
            const int nm = materials_count(d->mats);
            for (uint32_t i = 0; i < n; ++i)
                require(v->material[i] < nm, "nbl_driver_write_slots: material index beyond the material count");
        }
        if (v->status) d->status.upload(v->status, n, d->stream);
        if (v->particle_idx) d->pidx.upload(v->particle_idx, n, d->stream);
        if (v->particles) d->parts.upload(reinterpret_cast<const dparticle*>(v->particles), n, d->stream);
        if (v->tags) d->tags.upload(v->tags, n, d->stream);
        if (v->material) d->mat.upload(v->material, n, d->stream);
        if (v->last_tri) d->last_tri.upload(v->last_tri, n, d->stream);
        if (v->rng) d->rng.upload(reinterpret_cast<const rng_state*>(v->rng), n, d->stream);
        NBL_HIP(hipStreamSynchronize(d->stream));
    });
}

int nbl_sort_status_batch(const uint8_t* keys, uint32_t n, uint32_t* out_index)
{
    return guarded([&] {
        require(n == 0 || (keys && out_index), "nbl_sort_status_batch: NULL array");
        if (n == 0) return;
        ensure_device();
        sorter s; s.init(n);
        dbuf<uint8_t> dk(n); dbuf<uint32_t> dout(n);
        dk.upload(keys, n);
        s.run(dk.p, n, dout.p, 0);
        dout.download(out_index, n);
        NBL_HIP(hipDeviceSynchronize());
    });
}

int nbl_rng_seed(uint64_t seed, uint64_t first_sequence, uint32_t n, uint32_t* states)
{
    return guarded([&] {
        require(n == 0 || states, "nbl_rng_seed: NULL states");
        for (uint32_t i = 0; i < n; ++i) {
            const rng_state r = rng_seed(seed, first_sequence + i);
            std::memcpy(states + 6 * (size_t)i, &r, sizeof r);
        }
    });
}

int nbl_rng_uniform_batch(const uint32_t* states, uint32_t n, uint32_t draws, float* out)
{
    return guarded([&] {
        require(n == 0 || (states && out), "nbl_rng_uniform_batch: NULL array");
        if (n == 0 || draws == 0) return;
        ensure_device();
        dbuf<rng_state> ds(n); dbuf<float> dout((size_t)n * draws);
        ds.upload(reinterpret_cast<const rng_state*>(states), n);
        hipLaunchKernelGGL(k_rng_uniform, dim3(grid_for(n, 256)), dim3(256), 0, 0, ds.p, n, draws, dout.p);
        NBL_HIP(hipGetLastError());
        dout.download(out, (size_t)n * draws);
        NBL_HIP(hipDeviceSynchronize());
    });
}

}  // extern "C"
