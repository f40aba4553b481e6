// mlp_device.h -- the learned ICDF sampler: a BF16 MLP evaluated by a whole
// wavefront on the matrix cores (v_mfma_f32_16x16x16_bf16, FP32 accumulation).
//
// Network (one per table kind, shared by all materials):
//   features f (16): (log E - shift) * scale, 2u - 1, log(u)/8, log(1-u)/8,
//                    the material's 8-dim embedding, 4 zeros
//   h1 = relu(W1 f + b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3)   (64 each)
//   y  = w4 . h3 + b4, then y * y_scale + y_shift and the kind's transform:
//        elastic    cos(theta) = 1 + 1e-6 - exp(y)     (trained on log(1 - cos + 1e-6))
//        inelastic  log(omega') = y                    (the table's own value)
//
// Matrix layout.  Samples are the N dimension: a wavefront's 64 lanes hold
// 64 samples, in 4 blocks of 16.  For v_mfma_f32_16x16x16_bf16 lane l holds
//   A (16x16, rows i, cols k):  i = l % 16, k = 4 (l / 16) + e, e = 0..3
//   B (16x16, rows k, cols j):  j = l % 16, k = 4 (l / 16) + e
//   D (16x16, rows i, cols j):  j = l % 16, i = 4 (l / 16) + r, r = 0..3
// so a layer's D block for outputs 16m..16m+15 is, lane by lane and element
// by element, the next layer's B block for inputs 16m..16m+15: activations
// stay in registers from layer to layer.  Weights are stored pre-swizzled as
// A operands: block (m, q) lane l = W[16m + l%16][16q + 4(l/16) + e].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nbl {

constexpr int MLP_IN = 16;       // input features (12 used, padded to one K block)
constexpr int MLP_H = 64;        // hidden width
constexpr int MLP_EMB = 8;       // embedding width
constexpr int MLP_MAX_MAT = 64;  // materials per sampler
constexpr int MLP_KIND_ELASTIC = 0, MLP_KIND_INELASTIC = 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Device view of one network.
struct dmlp {
    const bf16x4* w1;         // A operands, [m 4][q 1][lane 64]
    const bf16x4* w2;         // [m 4][q 4][lane 64]
    const bf16x4* w3;         // [m 4][q 4][lane 64]
    const float* b1;          // 64
    const float* b2;
    const float* b3;
    const float* w4;          // 64 (output layer, FP32 on the vector ALU)
    const float* emb;         // [material][8]
    float b4;
    float e_shift, e_scale;   // feature 0
    float y_scale, y_shift;   // output de-standardisation
    int kind;
};

// float -> bf16, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ uint16_t to_bf16(float x)
{
    uint32_t b = __float_as_uint(x);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d)
{
    bf16x4 v;
    v[0] = (short)to_bf16(a); v[1] = (short)to_bf16(b); v[2] = (short)to_bf16(c); v[3] = (short)to_bf16(d);
    return v;
}
__device__ __forceinline__ bf16x4 relu_pack(f32x4 d)
{
    return pack4(fmaxf(d[0], 0.0f), fmaxf(d[1], 0.0f), fmaxf(d[2], 0.0f), fmaxf(d[3], 0.0f));
}

// The four scalar input features of a sample (features 0..3).
__device__ __forceinline__ void mlp_scalars(const dmlp& n, float log_e, float u, float s[4])
{
    s[0] = (log_e - n.e_shift) * n.e_scale;
    s[1] = 2.0f * u - 1.0f;
    s[2] = logf(u) * 0.125f;
    s[3] = logf(fmaxf(1.0f - u, 5.9604644775390625e-8f)) * 0.125f;
}

// Output transform of the network's raw value.
__device__ __forceinline__ float mlp_value(const dmlp& n, float y)
{
    const float v = y * n.y_scale + n.y_shift;
    return n.kind == MLP_KIND_ELASTIC ? 1.000001f - expf(v) : v;
}

// Evaluate the network for the samples held by lanes 0 .. 16 nblk - 1 (one
// per lane).  Must be called by all 64 lanes with the same nblk (1..4).
// Returns this lane's sample's value (defined for lanes < 16 nblk).
__device__ __forceinline__ float mlp_eval_blocks(const dmlp& n, int mat, float log_e, float u, int nblk)
{
    const uint32_t lane = __lane_id();
    const uint32_t j = lane & 15u, g = lane >> 4;
    float sc[4];
    mlp_scalars(n, log_e, u, sc);
    const int mat0 = mat < 0 ? 0 : mat;
    float mine = 0;
    // One block of 16 samples at a time; compiler barriers keep weight and
    // bias loads next to their use (L1/L2 hits) instead of hoisting them all
    // into registers.
#pragma unroll 1
    for (int nb = 0; nb < nblk; ++nb) {
        asm volatile("" ::: "memory");
        // B operand of layer 1: features 4g..4g+3 of the sample in lane j + 16 nb
        // (g = 0: the scalars, g = 1, 2: embedding halves, g = 3: zeros)
        const int src = (int)(j + 16u * nb);
        const float s0 = __shfl(sc[0], src), s1 = __shfl(sc[1], src), s2 = __shfl(sc[2], src), s3 = __shfl(sc[3], src);
        const int sm = __shfl(mat0, src);
        float x0 = 0, x1 = 0, x2 = 0, x3 = 0;
        if (g == 0) {
            x0 = s0; x1 = s1; x2 = s2; x3 = s3;
        } else if (g < 3) {
            const float4 e = *reinterpret_cast<const float4*>(n.emb + (size_t)sm * MLP_EMB + 4 * (g - 1));
            x0 = e.x; x1 = e.y; x2 = e.z; x3 = e.w;
        }
        const bf16x4 bin = pack4(x0, x1, x2, x3);
        // layer 1
        bf16x4 bh[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const bf16x4* wm = n.w1 + m * 64;
            asm volatile("" : "+s"(wm));
            const int o = 16 * m + 4 * (int)g;
            const float4 bb = *reinterpret_cast<const float4*>(n.b1 + o);
            f32x4 acc = { bb.x, bb.y, bb.z, bb.w };
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wm[lane], bin, acc, 0, 0, 0);
            bh[m] = relu_pack(acc);
        }
        // layers 2 and 3
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {
            const bf16x4* w = layer == 0 ? n.w2 : n.w3;
            const float* b = layer == 0 ? n.b2 : n.b3;
            bf16x4 nh[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                // The block's base pointer passes through an empty asm statement,
                // so its four operand addresses are formed here (scalar base +
                // lane offset) and not hoisted out of the loop as 64-bit vector
                // addresses, which would hold ~64 registers.
                const bf16x4* wm = w + m * 4 * 64;
                asm volatile("" : "+s"(wm));
                const int o = 16 * m + 4 * (int)g;
                const float4 bb = *reinterpret_cast<const float4*>(b + o);
                f32x4 acc = { bb.x, bb.y, bb.z, bb.w };
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wm[q * 64 + lane], bh[q], acc, 0, 0, 0);
                nh[m] = relu_pack(acc);
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) bh[m] = nh[m];
        }
        // output layer: this lane's 16 hidden values of sample j + 16 nb, summed over the 4 groups
        float s = 0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const float4 ww = *reinterpret_cast<const float4*>(n.w4 + 16 * m + 4 * (int)g);
            s += ww.x * from_bf16((uint16_t)bh[m][0]);
            s += ww.y * from_bf16((uint16_t)bh[m][1]);
            s += ww.z * from_bf16((uint16_t)bh[m][2]);
            s += ww.w * from_bf16((uint16_t)bh[m][3]);
        }
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (g == (uint32_t)nb) mine = s;
    }
    return mlp_value(n, mine + n.b4);
}

// All 64 lanes' samples (lanes whose sample is not needed pass any finite input).
__device__ __forceinline__ float mlp_eval_wave(const dmlp& n, int mat, float log_e, float u)
{
    return mlp_eval_blocks(n, mat, log_e, u, 4);
}

// Only the samples of lanes with need == true: they are moved to the first
// lanes (a stable partition by one permute), the network runs over
// ceil(count / 16) blocks, and each lane reads its value back.  Must be
// called by all 64 lanes; returns the value for lanes with need == true.
__device__ __forceinline__ float mlp_eval_needed(const dmlp& n, bool need, int mat, float log_e, float u)
{
    const uint64_t m = __ballot(need);
    const uint32_t cnt = (uint32_t)__popcll(m);
    if (cnt == 0) return 0.0f;
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    const uint32_t lane = __lane_id();
    // slot of this lane's sample: needed samples first, in lane order
    const int slot = (int)(need ? below : cnt + (lane - below));
    const int addr = slot * 4;
    const float le = __int_as_float(__builtin_amdgcn_ds_permute(addr, __float_as_int(log_e)));
    const float uu = __int_as_float(__builtin_amdgcn_ds_permute(addr, __float_as_int(u)));
    const int mm = __builtin_amdgcn_ds_permute(addr, mat);
    const float v = mlp_eval_blocks(n, mm, le, uu, (int)((cnt + 15) / 16));
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

}  // namespace nbl
