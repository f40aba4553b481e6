// sampler.hip -- nbl_sampler_*: the learned ICDF sampler.
//
// Two networks (elastic cos(theta)-ICDF, inelastic omega'-ICDF; the binding
// table stays a table) are trained on samples drawn from the material
// tables themselves, with hand-written FP32 kernels: sample generation
// (the transport's own tab2_get lookups), dense forward and backward passes
// in feature-major layout, embedding gradients and Adam.  The trained FP32
// weights are then rounded to BF16 and stored as MFMA A operands for
// mlp_eval_wave (mlp_device.h), which the persistent kernel and
// nbl_sampler_eval_batch use.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <vector>

#include "nbl_host.h"
#include "mlp_device.h"
#include "sampler.h"

using namespace nbl;

namespace nbl {
const dmat* materials_device(nbl_materials m);
int materials_count(nbl_materials m);
const dmat& materials_view(nbl_materials m, int i);
}

namespace {

// FP32 parameter block of one network (offsets in floats)
constexpr int P_W1 = 0;                          // [64][16]
constexpr int P_B1 = P_W1 + MLP_H * MLP_IN;      // [64]
constexpr int P_W2 = P_B1 + MLP_H;               // [64][64]
constexpr int P_B2 = P_W2 + MLP_H * MLP_H;
constexpr int P_W3 = P_B2 + MLP_H;
constexpr int P_B3 = P_W3 + MLP_H * MLP_H;
constexpr int P_W4 = P_B3 + MLP_H;               // [64]
constexpr int P_B4 = P_W4 + MLP_H;               // [1] (+3 pad)
constexpr int P_EMB = P_B4 + 4;                  // [MLP_MAX_MAT][8]
constexpr int P_TOTAL = P_EMB + MLP_MAX_MAT * MLP_EMB;
// BF16 A operands, in bf16 elements: layer 1 [m 4][lane 64][4], layers 2 and 3
// [m 4][q 2][lane 64][8], output [q 2][lane 64][8]
constexpr int PK_W1 = 0, PK_W2 = PK_W1 + 4 * 64 * 4, PK_W3 = PK_W2 + 8 * 64 * 8, PK_W4 = PK_W3 + 8 * 64 * 8;
constexpr int N_PACK = PK_W4 + 2 * 64 * 8;
static_assert(P_TOTAL == NBL_SAMPLER_PARAMS, "parameter block layout");
static_assert(MLP_KIND_ELASTIC == NBL_SAMPLER_ELASTIC && MLP_KIND_INELASTIC == NBL_SAMPLER_INELASTIC, "kinds");

__device__ __forceinline__ uint64_t smix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float unit24(uint32_t r) { return (float)((r >> 8) + 1u) * 5.9604644775390625e-8f; }

struct gen_args {
    const dmat* mats;
    int n_mat, kind;
    float le_lo, le_hi;        // log energy range
    float e_shift, e_scale;
    float y_shift, y_scale;
    uint64_t seed;
    uint32_t B;
};

// Training target in the network's output space (before standardisation).
__device__ __forceinline__ float target_of(int kind, float v)
{
    return kind == MLP_KIND_ELASTIC ? logf(1.000001f - fminf(fmaxf(v, -1.0f), 1.0f)) : v;
}

// Samples of one step: material, log E, u -> features X[16][B], target T[B], material M[B].
__global__ __launch_bounds__(256) void k_gen(gen_args a, const float* __restrict__ prm, uint64_t step,
    float* __restrict__ X, float* __restrict__ T, int* __restrict__ M)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.B) return;
    const uint64_t h1 = smix(a.seed ^ smix(step * 0x100000001B3ull + b));
    const uint64_t h2 = smix(h1);
    const int mat = (int)((h1 >> 40) % (uint64_t)a.n_mat);
    const float le = a.le_lo + (a.le_hi - a.le_lo) * unit24((uint32_t)h1);
    const float u = unit24((uint32_t)h2);
    const dmat& m = a.mats[mat];
    const float v = tab2_get(a.kind == MLP_KIND_ELASTIC ? m.el_icdf : m.in_icdf, le, u);
    T[b] = (target_of(a.kind, v) - a.y_shift) / a.y_scale;
    M[b] = mat;
    const size_t B = a.B;
    float sc[4];
    mlp_scalars(a.e_shift, a.e_scale, le, u, sc);
    for (int i = 0; i < 4; ++i) X[i * B + b] = sc[i];
    const float* e = prm + P_EMB + mat * MLP_EMB;
    for (int i = 0; i < MLP_EMB; ++i) X[(4 + i) * B + b] = e[i];
    for (int i = 12; i < MLP_IN; ++i) X[i * B + b] = 0.0f;
}

// Y[n][b] = act(bias[n] + sum_k W[n][k] X[k][b]), n < 64
template<int K, bool RELU>
__global__ __launch_bounds__(256) void k_fwd(const float* __restrict__ W, const float* __restrict__ bias,
    const float* __restrict__ X, float* __restrict__ Y, uint32_t B)
{
    __shared__ float w[MLP_H * K + MLP_H];
    for (int i = threadIdx.x; i < MLP_H * K; i += blockDim.x) w[i] = W[i];
    for (int i = threadIdx.x; i < MLP_H; i += blockDim.x) w[MLP_H * K + i] = bias[i];
    __syncthreads();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float x[K];
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] = X[(size_t)k * B + b];
    for (int n = 0; n < MLP_H; ++n) {
        float acc = w[MLP_H * K + n];
#pragma unroll
        for (int k = 0; k < K; ++k) acc += w[n * K + k] * x[k];
        Y[(size_t)n * B + b] = RELU ? fmaxf(acc, 0.0f) : acc;
    }
}

// Output layer and squared-error loss: dY[b] = 2 (y - T) / B; loss += (y - T)^2.
__global__ __launch_bounds__(256) void k_out(const float* __restrict__ prm, const float* __restrict__ H3,
    const float* __restrict__ T, float* __restrict__ dY, double* __restrict__ loss, uint32_t B)
{
    __shared__ float w[MLP_H + 1];
    __shared__ double red[256];
    for (int i = threadIdx.x; i <= MLP_H; i += blockDim.x) w[i] = prm[P_W4 + i];
    __syncthreads();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0;
    if (b < B) {
        float y = w[MLP_H];
        for (int n = 0; n < MLP_H; ++n) y += w[n] * H3[(size_t)n * B + b];
        // table cells with no finite value (e.g. log of 0 below the Fermi
        // level, energies no electron reaches) carry no loss
        const float t = T[b];
        const float d = isfinite(t) ? y - t : 0.0f;
        dY[b] = 2.0f * d / (float)B;
        l = (double)d * d;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// dH3[n][b] = dY[b] w4[n] (H3 > 0); gradient of w4, b4 accumulated per block.
__global__ __launch_bounds__(256) void k_out_bwd(const float* __restrict__ prm, const float* __restrict__ H3,
    const float* __restrict__ dY, float* __restrict__ dH3, float* __restrict__ grad, uint32_t B)
{
    __shared__ float gw[MLP_H + 1];
    for (int i = threadIdx.x; i <= MLP_H; i += blockDim.x) gw[i] = 0;
    __syncthreads();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) {
        const float d = dY[b];
        for (int n = 0; n < MLP_H; ++n) {
            const float h = H3[(size_t)n * B + b];
            dH3[(size_t)n * B + b] = h > 0 ? d * prm[P_W4 + n] : 0.0f;
            atomicAdd(&gw[n], d * h);
        }
        atomicAdd(&gw[MLP_H], d);
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= MLP_H; i += blockDim.x) atomicAdd(&grad[P_W4 + i], gw[i]);
}

// dX[k][b] = (sum_n W[n][k] dY[n][b]) * (Hprev[k][b] > 0 if MASK)
template<int K, bool MASK>
__global__ __launch_bounds__(256) void k_bwd_data(const float* __restrict__ W, const float* __restrict__ dY,
    const float* __restrict__ Hprev, float* __restrict__ dX, uint32_t B)
{
    __shared__ float w[MLP_H * K];
    for (int i = threadIdx.x; i < MLP_H * K; i += blockDim.x) w[i] = W[i];
    __syncthreads();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float d[MLP_H];
#pragma unroll
    for (int n = 0; n < MLP_H; ++n) d[n] = dY[(size_t)n * B + b];
    for (int k = 0; k < K; ++k) {
        float acc = 0;
#pragma unroll
        for (int n = 0; n < MLP_H; ++n) acc += w[n * K + k] * d[n];
        if (MASK && !(Hprev[(size_t)k * B + b] > 0)) acc = 0;
        dX[(size_t)k * B + b] = acc;
    }
}

// gW[n][k] += sum_b dY[n][b] X[k][b], gb[n] += sum_b dY[n][b]; one chunk of
// samples per block, staged through LDS 64 samples at a time.  Thread t owns
// row n = t / 4 and the K/4 columns of quarter t % 4.
template<int K>
__global__ __launch_bounds__(256) void k_bwd_weight(const float* __restrict__ dY, const float* __restrict__ X,
    float* __restrict__ gW, float* __restrict__ gb, uint32_t B, uint32_t chunk)
{
    constexpr int TS = 64;
    constexpr int PER = K / 4;
    __shared__ float sd[MLP_H][TS + 1];
    __shared__ float sx[K][TS + 1];
    const int n = threadIdx.x / 4, k0 = (threadIdx.x % 4) * PER;
    float acc[PER];
    float accb = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) acc[i] = 0;
    const uint32_t b0 = blockIdx.x * chunk, b1 = min(B, b0 + chunk);
    for (uint32_t t0 = b0; t0 < b1; t0 += TS) {
        for (int i = threadIdx.x; i < MLP_H * TS; i += blockDim.x) {
            const int r = i / TS, c = i % TS;
            sd[r][c] = (t0 + c < b1) ? dY[(size_t)r * B + t0 + c] : 0.0f;
        }
        for (int i = threadIdx.x; i < K * TS; i += blockDim.x) {
            const int r = i / TS, c = i % TS;
            sx[r][c] = (t0 + c < b1) ? X[(size_t)r * B + t0 + c] : 0.0f;
        }
        __syncthreads();
#pragma unroll 2
        for (int c = 0; c < TS; ++c) {
            const float d = sd[n][c];
#pragma unroll
            for (int i = 0; i < PER; ++i) acc[i] += d * sx[k0 + i][c];
            accb += d;
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) atomicAdd(&gW[n * K + k0 + i], acc[i]);
    if (threadIdx.x % 4 == 0) atomicAdd(&gb[n], accb);
}

// Embedding gradient: gE[mat][i] += dX[4 + i][b]
__global__ __launch_bounds__(256) void k_emb_grad(const float* __restrict__ dX, const int* __restrict__ M,
    float* __restrict__ grad, uint32_t B, int n_mat)
{
    __shared__ float ge[MLP_MAX_MAT * MLP_EMB];
    for (int i = threadIdx.x; i < n_mat * MLP_EMB; i += blockDim.x) ge[i] = 0;
    __syncthreads();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) {
        const int m = M[b];
        for (int i = 0; i < MLP_EMB; ++i) atomicAdd(&ge[m * MLP_EMB + i], dX[(size_t)(4 + i) * B + b]);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n_mat * MLP_EMB; i += blockDim.x) atomicAdd(&grad[P_EMB + i], ge[i]);
}

// Adam (Kingma & Ba 2015) with bias correction; clears the gradient.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ prm, float* __restrict__ grad,
    float* __restrict__ mom, float* __restrict__ vel, int n, float lr, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i];
    const float m = 0.9f * mom[i] + 0.1f * g;
    const float v = 0.999f * vel[i] + 0.001f * g * g;
    mom[i] = m; vel[i] = v;
    prm[i] -= lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
    grad[i] = 0;
}

// BF16 A operands in the order of mlp_device.h (one bf16 per thread).
__global__ void k_pack(const float* __restrict__ prm, uint16_t* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N_PACK) return;
    float v;
    if (i < PK_W2) {                                    // layer 1: [m][lane][e]
        const int m = i / 256, lane = (i / 4) % 64, e = i % 4;
        v = prm[P_W1 + (16 * m + lane % 16) * MLP_IN + 4 * (lane / 16) + e];
    } else if (i < PK_W4) {                             // layers 2, 3: [m][q][lane][e]
        const bool l2 = i < PK_W3;
        const int r = i - (l2 ? PK_W2 : PK_W3);
        const int m = r / 1024, q = (r / 512) % 2, lane = (r / 8) % 64, e = r % 8;
        v = prm[(l2 ? P_W2 : P_W3) + (16 * m + lane % 16) * MLP_H + mlp_pi(q, lane / 16, e)];
    } else {                                            // output: [q][lane][e], row 0 only
        const int r = i - PK_W4;
        const int q = r / 512, lane = (r / 8) % 64, e = r % 8;
        v = lane % 16 == 0 ? prm[P_W4 + mlp_pi(q, lane / 16, e)] : 0.0f;
    }
    out[i] = to_bf16(v);
}

// Batch evaluation through the transport's path (mlp_eval_needed): each
// wavefront evaluates its odd and its even lanes' samples in two compacted
// passes, so every value goes through the lane permutation the persistent
// kernel uses.
__global__ __launch_bounds__(256) void k_eval(dmlp net, const int* __restrict__ mat, const float* __restrict__ energy,
    const float* __restrict__ u, uint32_t n, float* __restrict__ out)
{
    for (uint32_t w0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); w0 < n; w0 += gridDim.x * blockDim.x) {
        const uint32_t i = w0 + (threadIdx.x & 63u);
        const bool on = i < n, odd = (threadIdx.x & 1u) != 0;
        const uint32_t k = on ? i : w0;
        const int m = mat[k];
        const float le = logf(energy[k]), uu = u[k];
        const float v1 = mlp_eval_needed(net, on && odd, m, le, uu);
        const float v0 = mlp_eval_needed(net, on && !odd, m, le, uu);
        if (on) out[i] = odd ? v1 : v0;
    }
}

// Known-answer tests of the matrix-core layouts: a 16x16x16 product (D1 =
// A[16][16] B[16][16]) and a 16x16x32 product (D2 = A[16][32] B[32][16]),
// operands loaded as mlp_device.h assumes.
__global__ void k_selftest(const float* __restrict__ A, const float* __restrict__ Bm, float* __restrict__ D1,
    float* __restrict__ D2)
{
    const int l = threadIdx.x, i = l % 16, g = l / 16;
    const bf16x4 a = pack4(A[i * 16 + 4 * g], A[i * 16 + 4 * g + 1], A[i * 16 + 4 * g + 2], A[i * 16 + 4 * g + 3]);
    const bf16x4 b = pack4(Bm[(4 * g) * 16 + i], Bm[(4 * g + 1) * 16 + i], Bm[(4 * g + 2) * 16 + i], Bm[(4 * g + 3) * 16 + i]);
    f32x4 acc = { 0, 0, 0, 0 };
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) D1[(4 * g + r) * 16 + i] = acc[r];
    // 16x16x32: A row i = l % 16, k = 8g + e; B col j = l % 16, k = 8g + e
    const float* A2 = A + 256;                 // [16][32]
    const float* B2 = Bm + 256;                // [32][16]
    bf16x4 alo = pack4(A2[i * 32 + 8 * g], A2[i * 32 + 8 * g + 1], A2[i * 32 + 8 * g + 2], A2[i * 32 + 8 * g + 3]);
    bf16x4 ahi = pack4(A2[i * 32 + 8 * g + 4], A2[i * 32 + 8 * g + 5], A2[i * 32 + 8 * g + 6], A2[i * 32 + 8 * g + 7]);
    bf16x4 blo = pack4(B2[(8 * g) * 16 + i], B2[(8 * g + 1) * 16 + i], B2[(8 * g + 2) * 16 + i], B2[(8 * g + 3) * 16 + i]);
    bf16x4 bhi = pack4(B2[(8 * g + 4) * 16 + i], B2[(8 * g + 5) * 16 + i], B2[(8 * g + 6) * 16 + i], B2[(8 * g + 7) * 16 + i]);
    f32x4 acc2 = { 0, 0, 0, 0 };
    acc2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(join8(alo, ahi), join8(blo, bhi), acc2, 0, 0, 0);
    for (int r = 0; r < 4; ++r) D2[(4 * g + r) * 16 + i] = acc2[r];
}

}  // namespace

struct nbl_sampler_s {
    nbl_materials mats = nullptr;
    int n_mat = 0;
    dbuf<float> prm[2];            // FP32 parameters per network
    dbuf<uint16_t> packed[2];      // BF16 A operands per network
    dmlp view[2]{};
    uint32_t kinds = 3;            // bit k: network k trained and used
    nbl_sampler_report report{};
};

namespace nbl {
const dmlp* sampler_views(nbl_sampler s) { return s ? s->view : nullptr; }
uint32_t sampler_kinds(nbl_sampler s) { return s ? s->kinds : 0u; }
nbl_materials sampler_materials(nbl_sampler s) { return s ? s->mats : nullptr; }
}

namespace {

void make_view(nbl_sampler_s& s, int k, float e_shift, float e_scale, float y_shift, float y_scale, float b4)
{
    dmlp& v = s.view[k];
    const float* p = s.prm[k].p;
    const uint16_t* pk = s.packed[k].p;
    v.w1 = reinterpret_cast<const bf16x4*>(pk + PK_W1);
    v.w2 = reinterpret_cast<const bf16x8*>(pk + PK_W2);
    v.w3 = reinterpret_cast<const bf16x8*>(pk + PK_W3);
    v.w4 = reinterpret_cast<const bf16x8*>(pk + PK_W4);
    v.b1 = p + P_B1; v.b2 = p + P_B2; v.b3 = p + P_B3; v.emb = p + P_EMB;
    v.b4 = b4; v.e_shift = e_shift; v.e_scale = e_scale; v.y_shift = y_shift; v.y_scale = y_scale;
    v.kind = k;
}

// He-uniform weights, zero biases, small random embeddings (host LCG, seeded).
std::vector<float> init_params(uint64_t seed)
{
    std::vector<float> p(P_TOTAL, 0.0f);
    uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() { s = s * 6364136223846793005ull + 1442695040888963407ull; return (float)((s >> 40) * (1.0 / 16777216.0)) * 2 - 1; };
    auto fill = [&](int off, int n, float a) { for (int i = 0; i < n; ++i) p[off + i] = a * uni(); };
    fill(P_W1, MLP_H * MLP_IN, std::sqrt(6.0f / 12));
    fill(P_W2, MLP_H * MLP_H, std::sqrt(6.0f / MLP_H));
    fill(P_W3, MLP_H * MLP_H, std::sqrt(6.0f / MLP_H));
    fill(P_W4, MLP_H, std::sqrt(3.0f / MLP_H));
    fill(P_EMB, MLP_MAX_MAT * MLP_EMB, 0.5f);
    return p;
}

void train_network(nbl_sampler_s& S, int kind, const nbl_sampler_opts& o, float le_lo, float le_hi, double& loss_out)
{
    const uint32_t B = o.batch;
    const float e_shift = 0.5f * (le_lo + le_hi), e_scale = 2.0f / (le_hi - le_lo);
    gen_args a{ materials_device(S.mats), S.n_mat, kind, le_lo, le_hi, e_shift, e_scale, 0.0f, 1.0f,
                o.seed * 2 + (uint64_t)kind, B };
    auto& prm = S.prm[kind];
    prm.alloc(P_TOTAL);
    const std::vector<float> init = init_params(o.seed * 2 + (uint64_t)kind + 17);
    prm.upload(init.data(), P_TOTAL);
    dbuf<float> grad(P_TOTAL), mom(P_TOTAL), vel(P_TOTAL);
    NBL_HIP(hipMemset(grad.p, 0, grad.bytes())); NBL_HIP(hipMemset(mom.p, 0, mom.bytes()));
    NBL_HIP(hipMemset(vel.p, 0, vel.bytes()));
    dbuf<float> X((size_t)MLP_IN * B), T(B), Y(B), dY(B), dX((size_t)MLP_IN * B);
    dbuf<float> H1((size_t)MLP_H * B), H2((size_t)MLP_H * B), H3((size_t)MLP_H * B);
    dbuf<float> D1((size_t)MLP_H * B), D2((size_t)MLP_H * B), D3((size_t)MLP_H * B);
    dbuf<int> M(B);
    dbuf<double> loss(1);
    const dim3 g((B + 255) / 256), blk(256);
    // target standardisation from one batch
    {
        hipLaunchKernelGGL(k_gen, g, blk, 0, 0, a, prm.p, (uint64_t)~0ull, X.p, T.p, M.p);
        NBL_HIP(hipGetLastError());
        std::vector<float> t(B);
        T.download(t.data(), B);
        double s1 = 0, s2 = 0;
        size_t cnt = 0;
        for (float x : t)
            if (std::isfinite(x)) { s1 += x; s2 += (double)x * x; ++cnt; }
        require(cnt > B / 2, "nbl_sampler_create: the tables have no finite values over most of the energy range");
        const double mean = s1 / cnt, var = std::max(s2 / cnt - mean * mean, 1e-12);
        a.y_shift = (float)mean;
        a.y_scale = (float)std::sqrt(var);
    }
    const uint64_t steps = std::max<uint64_t>(1, o.samples / B);
    // weight gradients: one block per 128 samples (at least 32 blocks for
    // the default batch; the step time is set by these reductions' depth)
    const uint32_t chunk = 128, nchunk = (B + chunk - 1) / chunk;
    double last = 0;
    for (uint64_t st = 0; st < steps; ++st) {
        const bool probe = st + 1 == steps || (st % 64) == 63;
        if (probe) NBL_HIP(hipMemsetAsync(loss.p, 0, sizeof(double), 0));
        hipLaunchKernelGGL(k_gen, g, blk, 0, 0, a, prm.p, st, X.p, T.p, M.p);
        hipLaunchKernelGGL((k_fwd<MLP_IN, true>), g, blk, 0, 0, prm.p + P_W1, prm.p + P_B1, X.p, H1.p, B);
        hipLaunchKernelGGL((k_fwd<MLP_H, true>), g, blk, 0, 0, prm.p + P_W2, prm.p + P_B2, H1.p, H2.p, B);
        hipLaunchKernelGGL((k_fwd<MLP_H, true>), g, blk, 0, 0, prm.p + P_W3, prm.p + P_B3, H2.p, H3.p, B);
        hipLaunchKernelGGL(k_out, g, blk, 0, 0, prm.p, H3.p, T.p, dY.p, loss.p, B);
        hipLaunchKernelGGL(k_out_bwd, g, blk, 0, 0, prm.p, H3.p, dY.p, D3.p, grad.p, B);
        hipLaunchKernelGGL((k_bwd_weight<MLP_H>), dim3(nchunk), blk, 0, 0, D3.p, H2.p, grad.p + P_W3, grad.p + P_B3, B, chunk);
        hipLaunchKernelGGL((k_bwd_data<MLP_H, true>), g, blk, 0, 0, prm.p + P_W3, D3.p, H2.p, D2.p, B);
        hipLaunchKernelGGL((k_bwd_weight<MLP_H>), dim3(nchunk), blk, 0, 0, D2.p, H1.p, grad.p + P_W2, grad.p + P_B2, B, chunk);
        hipLaunchKernelGGL((k_bwd_data<MLP_H, true>), g, blk, 0, 0, prm.p + P_W2, D2.p, H1.p, D1.p, B);
        hipLaunchKernelGGL((k_bwd_weight<MLP_IN>), dim3(nchunk), blk, 0, 0, D1.p, X.p, grad.p + P_W1, grad.p + P_B1, B, chunk);
        hipLaunchKernelGGL((k_bwd_data<MLP_IN, false>), g, blk, 0, 0, prm.p + P_W1, D1.p, (const float*)nullptr, dX.p, B);
        hipLaunchKernelGGL(k_emb_grad, g, blk, 0, 0, dX.p, M.p, grad.p, B, S.n_mat);
        // cosine decay of the learning rate to 1 % of its start
        const double frac = (double)st / (double)steps;
        const float lr = (float)(o.lr * (0.01 + 0.99 * 0.5 * (1 + std::cos(3.141592653589793 * frac))));
        const float c1 = 1.0f - std::pow(0.9f, (float)(st + 1)), c2 = 1.0f - std::pow(0.999f, (float)(st + 1));
        hipLaunchKernelGGL(k_adam, dim3((P_TOTAL + 255) / 256), blk, 0, 0, prm.p, grad.p, mom.p, vel.p, P_TOTAL, lr, c1, c2);
        NBL_HIP(hipGetLastError());
        if (probe) {
            double l = 0;
            loss.download(&l, 1);
            last = l / B;
            require(std::isfinite(last), "nbl_sampler_create: training diverged");
        }
    }
    loss_out = last;
    S.packed[kind].alloc(N_PACK);
    hipLaunchKernelGGL(k_pack, dim3((N_PACK + 255) / 256), blk, 0, 0, prm.p, S.packed[kind].p);
    NBL_HIP(hipGetLastError());
    float b4 = 0;
    NBL_HIP(hipMemcpy(&b4, prm.p + P_B4, sizeof(float), hipMemcpyDeviceToHost));
    make_view(S, kind, e_shift, e_scale, a.y_shift, a.y_scale, b4);
    NBL_HIP(hipDeviceSynchronize());
}

}  // namespace

extern "C" {

int nbl_sampler_create(nbl_materials m, const nbl_sampler_opts* opts, nbl_sampler* out)
{
    return guarded([&] {
        require(m && out, "nbl_sampler_create: NULL argument");
        *out = nullptr;
        nbl_sampler_opts o{};
        o.samples = 100000000ull; o.batch = 4096; o.lr = 4e-3f; o.e_min = 0; o.e_max = 0; o.seed = 1;
        if (opts) {
            if (opts->samples) o.samples = opts->samples;
            if (opts->batch) o.batch = opts->batch;
            if (opts->lr > 0) o.lr = opts->lr;
            o.e_min = opts->e_min; o.e_max = opts->e_max; o.seed = opts->seed;
            o.kinds = opts->kinds;
        }
        const uint32_t kinds = o.kinds ? o.kinds : 3u;
        require(kinds <= 3u, "nbl_sampler_create: kinds must be a mask of the elastic and inelastic bits");
        require(o.batch >= 256 && o.batch <= (1u << 22) && o.batch % 256 == 0,
            "nbl_sampler_create: batch must be a multiple of 256 in [256, 4194304]");
        const int n_mat = materials_count(m);
        require(n_mat >= 1 && n_mat <= MLP_MAX_MAT, "nbl_sampler_create: 1..64 materials");
        // training energy range: the caller's, else the tables' own (common to all materials)
        float lo = -1e30f, hi = 1e30f;
        for (int i = 0; i < n_mat; ++i) {
            for (const dtab2* t : { &materials_view(m, i).el_icdf, &materials_view(m, i).in_icdf }) {
                lo = std::max(lo, t->x_min);
                hi = std::min(hi, t->x_min + (float)(t->w - 1) / t->x_step);
            }
        }
        if (o.e_min > 0) lo = std::max(lo, std::log(o.e_min));
        if (o.e_max > 0) hi = std::min(hi, std::log(o.e_max));
        require(hi > lo, "nbl_sampler_create: empty energy range");
        auto s = std::make_unique<nbl_sampler_s>();
        s->mats = m;
        s->n_mat = n_mat;
        s->kinds = kinds;
        const auto t0 = std::chrono::steady_clock::now();
        if (kinds & (1u << MLP_KIND_ELASTIC)) train_network(*s, MLP_KIND_ELASTIC, o, lo, hi, s->report.loss_elastic);
        if (kinds & (1u << MLP_KIND_INELASTIC))
            train_network(*s, MLP_KIND_INELASTIC, o, lo, hi, s->report.loss_inelastic);
        s->report.train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        s->report.steps = std::max<uint64_t>(1, o.samples / o.batch);
        s->report.samples = s->report.steps * o.batch;
        s->report.e_min = std::exp(lo);
        s->report.e_max = std::exp(hi);
        *out = s.release();
    });
}

int nbl_sampler_get_report(nbl_sampler s, nbl_sampler_report* r)
{
    return guarded([&] {
        require(s && r, "nbl_sampler_get_report: NULL argument");
        *r = s->report;
    });
}

int nbl_sampler_get_params(nbl_sampler s, int kind, float* params, uint32_t n, float* consts)
{
    return guarded([&] {
        require(s != nullptr, "nbl_sampler_get_params: NULL sampler");
        require(kind == NBL_SAMPLER_ELASTIC || kind == NBL_SAMPLER_INELASTIC, "nbl_sampler_get_params: bad kind");
        require(s->kinds & (1u << kind), "nbl_sampler_get_params: this network was not trained");
        require(n == (uint32_t)P_TOTAL, "nbl_sampler_get_params: n must be NBL_SAMPLER_PARAMS");
        if (params) s->prm[kind].download(params, P_TOTAL);
        if (consts) {
            const dmlp& v = s->view[kind];
            consts[0] = v.e_shift; consts[1] = v.e_scale; consts[2] = v.y_shift; consts[3] = v.y_scale;
        }
    });
}

int nbl_sampler_eval_batch(nbl_sampler s, int kind, const int32_t* mat, const float* energy, const float* u,
    uint32_t n, float* out)
{
    return guarded([&] {
        require(s != nullptr, "nbl_sampler_eval_batch: NULL sampler");
        require(kind == NBL_SAMPLER_ELASTIC || kind == NBL_SAMPLER_INELASTIC, "nbl_sampler_eval_batch: bad kind");
        require(s->kinds & (1u << kind), "nbl_sampler_eval_batch: this network was not trained");
        if (n == 0) return;
        require(mat && energy && u && out, "nbl_sampler_eval_batch: NULL array");
        for (uint32_t i = 0; i < n; ++i) {
            require(mat[i] >= 0 && mat[i] < s->n_mat, "nbl_sampler_eval_batch: material index out of range");
            require(energy[i] > 0 && std::isfinite(energy[i]), "nbl_sampler_eval_batch: energy must be > 0");
            require(u[i] > 0 && u[i] <= 1, "nbl_sampler_eval_batch: u must be in (0, 1]");
        }
        dbuf<int> dm(n); dbuf<float> de(n), du(n), dout(n);
        dm.upload(mat, n); de.upload(energy, n); du.upload(u, n);
        hipLaunchKernelGGL(k_eval, dim3(grid_for(n, 256, 1024)), dim3(256), 0, 0, s->view[kind], dm.p, de.p, du.p, n, dout.p);
        NBL_HIP(hipGetLastError());
        dout.download(out, n);
    });
}

int nbl_sampler_destroy(nbl_sampler s)
{
    return guarded([&] { delete s; });
}

int nbl_sampler_selftest(double* max_abs_err)
{
    return guarded([&] {
        require(max_abs_err != nullptr, "nbl_sampler_selftest: NULL argument");
        ensure_device();
        // small integers and halves: exact in BF16 and in the FP32 sums
        // A: [16][16] then [16][32]; B: [16][16] then [32][16]
        std::vector<float> A(256 + 512), Bm(256 + 512), D1(256), D2(256);
        for (int i = 0; i < 768; ++i) {
            A[i] = (float)((i * 7) % 9 - 4) * 0.5f;
            Bm[i] = (float)((i * 5) % 11 - 5) * 0.25f;
        }
        dbuf<float> dA(768), dB(768), dD1(256), dD2(256);
        dA.upload(A.data(), 768); dB.upload(Bm.data(), 768);
        hipLaunchKernelGGL(k_selftest, dim3(1), dim3(64), 0, 0, dA.p, dB.p, dD1.p, dD2.p);
        NBL_HIP(hipGetLastError());
        dD1.download(D1.data(), 256);
        dD2.download(D2.data(), 256);
        NBL_HIP(hipDeviceSynchronize());
        double err = 0;
        for (int i = 0; i < 16; ++i)
            for (int j = 0; j < 16; ++j) {
                double r1 = 0, r2 = 0;
                for (int k = 0; k < 16; ++k) r1 += (double)A[i * 16 + k] * Bm[k * 16 + j];
                for (int k = 0; k < 32; ++k) r2 += (double)A[256 + i * 32 + k] * Bm[256 + k * 16 + j];
                err = std::max(err, std::max(std::fabs(r1 - D1[i * 16 + j]), std::fabs(r2 - D2[i * 16 + j])));
            }
        *max_abs_err = err;
    });
}

}  // extern "C"
