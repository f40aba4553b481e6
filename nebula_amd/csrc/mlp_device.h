// mlp_device.h -- the learned ICDF sampler: a BF16 MLP evaluated by a whole
// wavefront on the matrix cores (v_mfma_f32_16x16x16_bf16 and _16x16x32_bf16,
// FP32 accumulation).
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
// 64 samples, in 4 blocks of 16.  Layer 1 (K = 16) uses
// v_mfma_f32_16x16x16_bf16, layers 2, 3 and the output layer (K = 64)
// v_mfma_f32_16x16x32_bf16.  With g = l / 16, lane l holds
//   16x16x16  A: row l%16, k = 4g + e (e < 4);  B: col l%16, k = 4g + e
//   16x16x32  A: row l%16, k = 8g + e (e < 8);  B: col l%16, k = 8g + e
//   D (both): col l%16, rows 4g + r (r < 4)
// so the D block of outputs 16m..16m+15 holds, lane by lane, hidden units
// 16m + 4g + r.  K order inside a product is free, so the 16x16x32 products
// of the next layer take K slot (q, g, e) as hidden unit
//   pi(q, g, e) = 16 (2q + e/4) + 4g + e%4,
// which makes their B operand the two D blocks 2q and 2q+1 of this lane
// side by side: activations stay in registers from layer to layer.  Weights
// are stored pre-swizzled as A operands in that order:
//   layer 1 block (m)     lane l: W1[16m + l%16][4g + e]
//   layer 2/3 block (m,q) lane l: W[16m + l%16][pi(q, g, e)]
//   output block (q)      lane l: w4[pi(q, g, e)] in row 0 (l%16 == 0), else 0
// and nbl_sampler_selftest checks both instructions' layouts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nbl {

constexpr int MLP_IN = 16;       // input features (12 used, padded to one K block)
constexpr int MLP_H = 64;        // hidden width
constexpr int MLP_EMB = 8;       // embedding width
constexpr int MLP_MAX_MAT = 64;  // materials per sampler
constexpr int MLP_KIND_ELASTIC = 0, MLP_KIND_INELASTIC = 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));     // 16x16x16 operand (4 x bf16 bits)
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));    // 16x16x32 operand
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// The networks' parameters as global memory (address space 1): loads
// through generic pointers would be flat loads, which also wait on the LDS
// and scalar-load counter the network's lane permutes use.
typedef __attribute__((address_space(1))) const bf16x4 gbf16x4;
typedef __attribute__((address_space(1))) const bf16x8 gbf16x8;
typedef __attribute__((address_space(1))) const f32x4 gf32x4;

// Device view of one network.
struct dmlp {
    const bf16x4* w1;         // A operands, [m 4][lane 64]
    const bf16x8* w2;         // [m 4][q 2][lane 64]
    const bf16x8* w3;         // [m 4][q 2][lane 64]
    const bf16x8* w4;         // [q 2][lane 64] (row 0 of a 16-row A operand)
    const float* b1;          // 64
    const float* b2;
    const float* b3;
    const float* emb;         // [material][8]
    float b4;
    float e_shift, e_scale;   // feature 0
    float y_scale, y_shift;   // output de-standardisation
    int kind;
};

// Hidden unit held by K slot (q, g, e) of a 16x16x32 product (see above).
__host__ __device__ constexpr int mlp_pi(int q, int g, int e) { return 16 * (2 * q + (e >> 2)) + 4 * g + (e & 3); }

// float -> bf16, round to nearest even (finite inputs); the device code uses
// the hardware conversion v_cvt_pk_bf16_f32, which rounds the same way.
__host__ __device__ __forceinline__ uint16_t to_bf16(float x)
{
    uint32_t b = __float_as_uint(x);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// two floats -> two bf16 in one word (low half first), one instruction on gfx950
__device__ __forceinline__ uint32_t cvt2(float a, float b)
{
    const f32x2 v = { a, b };
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}
__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d)
{
    const u32x2 w = { cvt2(a, b), cvt2(c, d) };
    return __builtin_bit_cast(bf16x4, w);
}
__device__ __forceinline__ bf16x4 relu_pack(f32x4 d)
{
    return pack4(fmaxf(d[0], 0.0f), fmaxf(d[1], 0.0f), fmaxf(d[2], 0.0f), fmaxf(d[3], 0.0f));
}
// D blocks 2q and 2q+1 side by side: the B operand of K slots q of a 16x16x32 product
__device__ __forceinline__ bf16x8 join8(bf16x4 lo, bf16x4 hi)
{
    const u32x2 a = __builtin_bit_cast(u32x2, lo), b = __builtin_bit_cast(u32x2, hi);
    const u32x4 w = { a[0], a[1], b[0], b[1] };
    return __builtin_bit_cast(bf16x8, w);
}

// The four scalar input features of a sample (features 0..3), with the
// hardware logarithm (training and inference use this same function).
__device__ __forceinline__ void mlp_scalars(float e_shift, float e_scale, float log_e, float u, float s[4])
{
    s[0] = (log_e - e_shift) * e_scale;
    s[1] = 2.0f * u - 1.0f;
    s[2] = __logf(u) * 0.125f;
    s[3] = __logf(fmaxf(1.0f - u, 5.9604644775390625e-8f)) * 0.125f;
}

// Output transform of the network's raw value.
__device__ __forceinline__ float mlp_value(const dmlp& n, float y)
{
    const float v = y * n.y_scale + n.y_shift;
    return n.kind == MLP_KIND_ELASTIC ? 1.000001f - __expf(v) : v;
}

// Evaluate the network for the samples held by lanes 0 .. 16 nblk - 1 (one
// per lane).  Must be called by all 64 lanes with the same nblk (1..4).
// Returns this lane's sample's value (defined for lanes < 16 nblk).
__device__ __forceinline__ float mlp_eval_blocks(const dmlp& n, int mat, float log_e, float u, int nblk)
{
    const uint32_t lane = __lane_id();
    const uint32_t j = lane & 15u, g = lane >> 4;
    float sc[4];
    mlp_scalars(n.e_shift, n.e_scale, log_e, u, sc);
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
            const f32x4 e = *(gf32x4*)(n.emb + (size_t)sm * MLP_EMB + 4 * (g - 1));
            x0 = e.x; x1 = e.y; x2 = e.z; x3 = e.w;
        }
        const bf16x4 bin = pack4(x0, x1, x2, x3);
        // layer 1 (K = 16)
        bf16x4 bh[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            gbf16x4* wm = (gbf16x4*)(n.w1 + m * 64);
            asm volatile("" : "+s"(wm));
            const int o = 16 * m + 4 * (int)g;
            const f32x4 bb = *(gf32x4*)(n.b1 + o);
            f32x4 acc = { bb.x, bb.y, bb.z, bb.w };
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wm[lane], bin, acc, 0, 0, 0);
            bh[m] = relu_pack(acc);
        }
        // layers 2 and 3 (K = 64: two 16x16x32 products per output block)
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {
            const bf16x8* w = layer == 0 ? n.w2 : n.w3;
            const float* b = layer == 0 ? n.b2 : n.b3;
            const bf16x8 in0 = join8(bh[0], bh[1]), in1 = join8(bh[2], bh[3]);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                // The block's base pointer passes through an empty asm statement,
                // so its operand addresses are formed here (scalar base + lane
                // offset) and not hoisted out of the loop as 64-bit vector
                // addresses, which would hold many registers.
                gbf16x8* wm = (gbf16x8*)(w + m * 2 * 64);
                asm volatile("" : "+s"(wm));
                const int o = 16 * m + 4 * (int)g;
                const f32x4 bb = *(gf32x4*)(b + o);
                f32x4 acc = { bb.x, bb.y, bb.z, bb.w };
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wm[lane], in0, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wm[64 + lane], in1, acc, 0, 0, 0);
                bh[m] = relu_pack(acc);
            }
        }
        // output layer: row 0 of a product whose other rows are zero; the
        // value of sample j + 16 nb lands in lane j, element 0
        {
            gbf16x8* wm = (gbf16x8*)n.w4;
            asm volatile("" : "+s"(wm));
            f32x4 acc = { 0, 0, 0, 0 };
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wm[lane], join8(bh[0], bh[1]), acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wm[64 + lane], join8(bh[2], bh[3]), acc, 0, 0, 0);
            const float v = __shfl(acc[0], (int)j);
            if (g == (uint32_t)nb) mine = v;
        }
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
