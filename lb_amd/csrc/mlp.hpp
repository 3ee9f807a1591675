// mlp.hpp -- the closure network on CDNA4 matrix cores (device code).
//
// Network: 12 -> 64 -> 64 -> 1, SiLU hidden activations, BF16 weights and
// activations on the matrix cores (v_mfma_f32_16x16x16_bf16 for layer 1,
// v_mfma_f32_16x16x32_bf16 for layer 2), FP32 accumulation; output head
// C = c_les * softplus(z) / ln 2 (z = 0 gives the Smagorinsky constant).
//
// Matrix-core layout (16x16x16, one wave): D[16x16] = A[16x16] B[16x16] + C;
// lane l = 16 q + s holds A[s][4q+j], B[4q+j][s] and D[4q+j][s], j = 0..3
// (16x16x32: A[s][8q+j], B[8q+j][s], j = 0..7; D as before).  Both layouts
// were checked against a host product on the GPU before use.
// Every layer is computed transposed -- hidden units along M, samples along
// N -- so that a layer's D fragment is the next layer's B fragment as is:
//   H1^T = W1 X^T,  H2^T = W2 A1^T,  z = w3 . A2 + b3.
// Inference runs the network in the scaled variable g = -log2(e) h, in
// which the logistic function is 1 / (1 + 2^g) and needs no multiply before
// the hardware exp2: the BF16 image of W1 and the FP32 b1, b2 carry the
// factor -log2(e), the activations are stored as g sigma = -log2(e) SiLU(h),
// so W2 needs no factor ((-log2 e)(-ln 2) = 1), and w3 carries -ln 2.  Up to
// the BF16 rounding of the scaled W1 and activations this is the same
// network; it saves one multiply per activation (128 per cell).
// A wave evaluates 64 samples as four tiles of 16; the 16 samples of tile t
// are the cells of lanes 16t .. 16t+15, whose features are staged in LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "features.hpp"

namespace lbm {

constexpr int NHID = 64;
constexpr int OFF_W1 = 0;                      // [64][12]
constexpr int OFF_B1 = OFF_W1 + NHID * NFEAT;  // [64]
constexpr int OFF_W2 = OFF_B1 + NHID;          // [64][64]
constexpr int OFF_B2 = OFF_W2 + NHID * NHID;   // [64]
constexpr int OFF_W3 = OFF_B2 + NHID;          // [64]
constexpr int OFF_B3 = OFF_W3 + NHID;          // [1]
constexpr int NPARAM = OFF_B3 + 1;             // 5057
constexpr float LN2F = 0.69314718055994531f;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits, round to nearest even.
__device__ __forceinline__ short f2bf(float x)
{
    uint32_t b = __float_as_uint(x);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (short)(b >> 16);
}

// Logistic function from the hardware exp2 and reciprocal.
__device__ __forceinline__ float sigmoid_hw(float h)
{
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * h));
}

__device__ __forceinline__ float silu_hw(float h) { return h * sigmoid_hw(h); }

// -log2(e) SiLU(h) in the scaled variable g = -log2(e) h: g / (1 + 2^g).
constexpr float NEG_LOG2E = -1.4426950408889634f;
constexpr float NEG_LN2 = -0.69314718055994531f;
__device__ __forceinline__ float gsilu_hw(float g)
{
    return g * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(g));
}

__device__ __forceinline__ float dsilu_hw(float h)
{
    const float s = sigmoid_hw(h);
    return s * (1.0f + h * (1.0f - s));
}

__device__ __forceinline__ f32x4 mfma_bf16(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Inference weights as the kernels read them.
struct MlpWeights {
    const uint16_t* w1;   // [64][16] BF16, columns 12..15 zero
    const uint16_t* w2;   // [64][64] BF16
    const float* b1;      // [64]
    const float* b2;      // [64]
    const float* w3;      // [64]
    const float* norm;    // [4][12]: mean, 1/standard deviation, envelope low, high (or null)
    float b3;
    float c;              // head scale (Smagorinsky constant)
};

typedef short s16x8 __attribute__((ext_vector_type(8)));

// Row stride (BF16 elements) of a wave's feature staging area: 18 rather
// than 16, so that the 16 lanes that read or write one column hit 16
// different bank pairs (9 dwords per row; 9 is odd).
constexpr int XS_STRIDE = 18;

// LDS image of the weights and the input normalisation, plus a per-wave
// feature staging area.  The weights are stored in matrix-core fragment
// order -- entry [..][lane] is what that lane feeds the instruction -- so a
// wave reads a fragment as 64 consecutive 8- or 16-byte words, free of bank
// conflicts.  (In row order the 16 lanes of a column read rows 128 bytes
// apart, all in one bank pair: a 16-way conflict on every fragment.)
// The normalisation sits here rather than in kernel arguments: its 48
// values held in scalar registers through the step made the compiler spill
// scalar registers.
struct MlpLds {
    bf16x4 w1f[4][64];         // [hidden block][lane]: W1[16 hb + s][4 q .. 4 q + 3] (scaled image)
    s16x8 w2f[4][2][64];       // [output block][32-block][lane]: W2[16 mb + s][32 kb + 4 q ..] and [.. + 16 ..]
    float b1[NHID], b2[NHID], w3[NHID];
    float norm[4][NFEAT];      // mean, 1/sd (no division to normalise), envelope low, high
                               // (the envelope in normalised units)
    uint16_t xs[4][64 * XS_STRIDE];   // per wave: [lane][feature], BF16
};

// Cooperative copy of the weights into LDS (fragment order) by a 256-thread
// workgroup.  Call in uniform control flow, followed by a workgroup barrier.
__device__ __forceinline__ void mlp_load_lds(MlpLds& L, const MlpWeights& W)
{
    const int t = threadIdx.x;
    {
        const int hb = t >> 6, lane = t & 63, q = lane >> 4, s = lane & 15;
        L.w1f[hb][lane] = *(const bf16x4*)&W.w1[(hb * 16 + s) * 16 + 4 * q];
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int f = t + 256 * r;   // 512 fragments: [mb][kb][lane]
        const int mb = f >> 7, kb = (f >> 6) & 1, lane = f & 63, q = lane >> 4, s = lane & 15;
        const bf16x4 w0 = *(const bf16x4*)&W.w2[(mb * 16 + s) * NHID + kb * 32 + 4 * q];
        const bf16x4 w1 = *(const bf16x4*)&W.w2[(mb * 16 + s) * NHID + kb * 32 + 16 + 4 * q];
        L.w2f[mb][kb][lane] = s16x8{w0[0], w0[1], w0[2], w0[3], w1[0], w1[1], w1[2], w1[3]};
    }
    if (t < NHID) {   // biases and output weights in the scaled variable
        L.b1[t] = NEG_LOG2E * W.b1[t];
        L.b2[t] = NEG_LOG2E * W.b2[t];
        L.w3[t] = NEG_LN2 * W.w3[t];
    }
    if (W.norm && t < 4 * NFEAT) (&L.norm[0][0])[t] = W.norm[t];
}

// Ordering of LDS accesses between lanes of one wave.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_hw __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8_hw __attribute__((ext_vector_type(8)));

// Four FP32 values to BF16 (round to nearest even, v_cvt_pk_bf16_f32).
__device__ __forceinline__ bf16x4 pack_bf16x4(float a, float b, float c, float d)
{
    const f32x2 lo = {a, b}, hi = {c, d};
    const bf16x2_hw l = __builtin_convertvector(lo, bf16x2_hw), h = __builtin_convertvector(hi, bf16x2_hw);
    bf16x4 r;
    __builtin_memcpy(&r, &l, 4);
    __builtin_memcpy((char*)&r + 4, &h, 4);
    return r;
}

// z of this lane's sample (normalised features xn).  Must be called by all
// 64 lanes of the wave in uniform control flow; lanes without a sample pass
// zeros and ignore the result.  Layer 2 contracts 32 hidden units per
// v_mfma_f32_16x16x32_bf16 in the order 4q..4q+3, 16+4q..16+4q+3 of each
// 32-block: that is the order in which a lane already holds the two
// layer-1 output fragments, so the fragments feed the instruction unchanged
// and only the weight fragment is read in that order.
// Every weight fragment, bias and output weight is read from LDS once per
// wave and applied to the four tiles in turn: four independent products per
// fragment, and a quarter of the LDS traffic of a tile-by-tile loop.  Each
// tile's arithmetic, and its order, is the same as evaluating it alone.
__device__ __forceinline__ float mlp_wave_forward(MlpLds& L, const float* xn, float b3)
{
    const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) & 3;
    const int q = lane >> 4, s = lane & 15;
    uint16_t* xs = L.xs[wave];
#pragma unroll
    for (int k = 0; k < 16; k += 4) {
        const bf16x4 v = pack_bf16x4(k < NFEAT ? xn[k] : 0.0f, k + 1 < NFEAT ? xn[k + 1] : 0.0f,
                                     k + 2 < NFEAT ? xn[k + 2] : 0.0f, k + 3 < NFEAT ? xn[k + 3] : 0.0f);
        *(bf16x4*)&xs[lane * XS_STRIDE + k] = v;
    }
    wave_lds_sync();
    bf16x4 xb[4];   // inputs of tile t: the cells of lanes 16t .. 16t+15
#pragma unroll
    for (int t = 0; t < 4; ++t) xb[t] = *(const bf16x4*)&xs[(16 * t + s) * XS_STRIDE + 4 * q];
    bf16x4 a1[4][4];   // layer-1 activations [tile][hidden block]
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
        const bf16x4 w = L.w1f[hb][lane];
        f32x4 bias;
#pragma unroll
        for (int j = 0; j < 4; ++j) bias[j] = L.b1[hb * 16 + 4 * q + j];
        // the four products first, then their activations: the matrix core
        // works on the next product while the activations run
        f32x4 c[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) c[t] = mfma_bf16(w, xb[t], bias);
#pragma unroll
        for (int t = 0; t < 4; ++t)
            a1[t][hb] = pack_bf16x4(gsilu_hw(c[t][0]), gsilu_hw(c[t][1]), gsilu_hw(c[t][2]), gsilu_hw(c[t][3]));
    }
    float zp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        f32x4 c[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) c[0][j] = L.b2[mb * 16 + 4 * q + j];
#pragma unroll
        for (int t = 1; t < 4; ++t) c[t] = c[0];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            const s16x8 wa = L.w2f[mb][kb][lane];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const s16x8 xa = {a1[t][2 * kb][0],     a1[t][2 * kb][1],     a1[t][2 * kb][2],
                                  a1[t][2 * kb][3],     a1[t][2 * kb + 1][0], a1[t][2 * kb + 1][1],
                                  a1[t][2 * kb + 1][2], a1[t][2 * kb + 1][3]};
                c[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_hw, wa),
                                                               __builtin_bit_cast(bf16x8_hw, xa), c[t], 0, 0, 0);
            }
        }
        float w3[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) w3[j] = L.w3[mb * 16 + 4 * q + j];
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) zp[t] += w3[j] * gsilu_hw(c[t][j]);
    }
    float zme = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        float v = zp[t];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (q == t) zme = v;
    }
    wave_lds_sync();   // the staging area is reused by the next call
    return zme + b3;
}

// Normalise features with the LDS image's mean and 1/sd; true when every
// one lies inside the training envelope.  The tests combine without
// short-circuit: a chain of per-feature branches made the compiler shuffle
// (and spill) scalar registers at every one.
__device__ __forceinline__ bool mlp_normalise(const MlpLds& L, const float* x, float* xn)
{
    unsigned ok = 1u;
#pragma unroll
    for (int k = 0; k < NFEAT; ++k) {
        xn[k] = (x[k] - L.norm[0][k]) * L.norm[1][k];
        // inside [lo, hi] exactly when the median of (lo, x, hi) is x (false for NaN)
        ok &= (unsigned)(__builtin_amdgcn_fmed3f(L.norm[2][k], xn[k], L.norm[3][k]) == xn[k]);
    }
    return ok != 0u;
}

// Output head C = c softplus(z) / ln 2 from the hardware exp2 and log2:
// softplus(z) / ln 2 = max(z, 0) / ln 2 + log2(1 + e), e = 2^(-|z| / ln 2).
// For e < 2^-12 (where 1 + e would round) log2(1 + e) = e (1 - e/2) / ln 2
// to within e^3.  Relative error a few 1e-7; the library log1p(exp(z)) it
// replaces took ~250 instructions per wave in the step kernel.
__device__ __forceinline__ float mlp_coefficient(float z, float c)
{
    constexpr float IL2 = 1.4426950408889634f;   // 1 / ln 2
    const float e = __builtin_amdgcn_exp2f(-IL2 * __builtin_fabsf(z));
    const float l = e < 0x1p-12f ? e * (IL2 - 0.5f * IL2 * e) : __builtin_amdgcn_logf(1.0f + e);
    return c * (__builtin_fmaxf(z, 0.0f) * IL2 + l);
}

}  // namespace lbm
