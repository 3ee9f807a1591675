// closure.hip -- training and stand-alone inference of the closure network
// (include/lbm_closure.h, lbm_mlp_*).
//
// Training step = k_mlp_grad (forward and backward of a batch; every
// workgroup writes its partial gradient) + k_mlp_reduce (deterministic FP64
// sum of the partials, then Adam).  Matrix products run on the matrix cores
// (BF16 operands, FP32 accumulation) in the transposed-layer layout of
// mlp.hpp; the weight gradients contract over the 16 samples of a tile, so
// their operands are transposed through LDS first.  With BF = false the same
// kernel forms every 16x16x16 product in FP32 through LDS instead, which
// checks the data flow to FP32 rounding.  The loss is accumulated in FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/lbm_closure.h"
#include "mlp.hpp"

using namespace lbm;

namespace {

constexpr int TPB_T = 256;

struct GradArgs {
    const float* x;       // [n][12]
    const float* y;       // [n]
    const float* w;       // [n] or null
    int64_t n;
    int64_t first;        // deterministic batch start
    int batch;
    int random;           // 1: hashed sample indices
    uint64_t seed;
    int64_t step;
    const float* p;       // parameters
    float c;
    float* part;          // [gridDim.x][NPARAM]
    double* part_loss;    // [gridDim.x]
};

struct TrainLds {
    float w1[NHID * 16];
    float w2[NHID * NHID];
    float b1[NHID], b2[NHID], w3[NHID];
    float ea[4][256], eb[4][256];
    union {
        struct {
            float t1[4][NHID * 16];
            float t2[4][NHID * 16];
        } tr;
        float acc[NPARAM];
    } u;
    double loss[4];
};

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ int64_t sample_index(const GradArgs& a, int64_t pos)
{
    if (!a.random) return a.first + pos;
    const uint64_t h = mix64(a.seed + 0x9E3779B97F4A7C15ull * (uint64_t)(a.step * (int64_t)a.batch + pos + 1));
    return (int64_t)(h % (uint64_t)a.n);
}

// D = A B + C for one 16x16x16 tile (fragments in the mlp.hpp layout).
template <bool BF>
__device__ __forceinline__ f32x4 mma(TrainLds& L, int wave, f32x4 a, f32x4 b, f32x4 c)
{
    if constexpr (BF) {
        bf16x4 ab, bb;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            ab[j] = f2bf(a[j]);
            bb[j] = f2bf(b[j]);
        }
        return mfma_bf16(ab, bb, c);
    } else {
        const int lane = threadIdx.x & 63, q = lane >> 4, s = lane & 15;
        float* EA = L.ea[wave];
        float* EB = L.eb[wave];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            EA[s * 16 + 4 * q + j] = a[j];
            EB[(4 * q + j) * 16 + s] = b[j];
        }
        wave_lds_sync();
        f32x4 d;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float acc = c[j];
            const int m = 4 * q + j;
#pragma unroll
            for (int k = 0; k < 16; ++k) acc = fmaf(EA[m * 16 + k], EB[k * 16 + s], acc);
            d[j] = acc;
        }
        wave_lds_sync();
        return d;
    }
}

template <bool BF>
__global__ __launch_bounds__(TPB_T) void k_mlp_grad(GradArgs a)
{
    __shared__ TrainLds L;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, q = lane >> 4, s = lane & 15;
    for (int i = tid; i < NHID * 16; i += TPB_T) {
        const int h = i >> 4, k = i & 15;
        L.w1[i] = k < NFEAT ? a.p[OFF_W1 + h * NFEAT + k] : 0.0f;
    }
    for (int i = tid; i < NHID * NHID; i += TPB_T) L.w2[i] = a.p[OFF_W2 + i];
    if (tid < NHID) {
        L.b1[tid] = a.p[OFF_B1 + tid];
        L.b2[tid] = a.p[OFF_B2 + tid];
        L.w3[tid] = a.p[OFF_W3 + tid];
    }
    const float b3 = a.p[OFF_B3];
    __syncthreads();

    f32x4 gW2[4][4], gW1[4];
    float gb1[4][4], gb2[4][4], gw3[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            gW1[u][j] = 0.0f;
            gb1[u][j] = gb2[u][j] = gw3[u][j] = 0.0f;
#pragma unroll
            for (int v = 0; v < 4; ++v) gW2[u][v][j] = 0.0f;
        }
    float gb3 = 0.0f;
    double loss = 0.0;
    const float inv_b = 1.0f / (float)a.batch;
    const int ntiles = a.batch / 16;
    float* T1 = L.u.tr.t1[wave];
    float* T2 = L.u.tr.t2[wave];

    for (int tile = blockIdx.x * 4 + wave; tile < ntiles; tile += gridDim.x * 4) {
        const int64_t i = sample_index(a, (int64_t)tile * 16 + s);
        f32x4 xb, xr;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = 4 * q + j;
            xb[j] = f < NFEAT ? a.x[i * NFEAT + f] : 0.0f;
            const int64_t ir = sample_index(a, (int64_t)tile * 16 + 4 * q + j);
            xr[j] = s < NFEAT ? a.x[ir * NFEAT + s] : 0.0f;
        }
        // forward: H1^T = W1 X^T + b1, H2^T = W2 A1^T + b2
        f32x4 h1[4], a1[4], h2[4];
#pragma unroll
        for (int hb = 0; hb < 4; ++hb) {
            f32x4 w, c;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                w[j] = L.w1[(hb * 16 + s) * 16 + 4 * q + j];
                c[j] = L.b1[hb * 16 + 4 * q + j];
            }
            h1[hb] = mma<BF>(L, wave, w, xb, c);
#pragma unroll
            for (int j = 0; j < 4; ++j) a1[hb][j] = silu_hw(h1[hb][j]);
        }
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            f32x4 c;
#pragma unroll
            for (int j = 0; j < 4; ++j) c[j] = L.b2[mb * 16 + 4 * q + j];
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                f32x4 w;
#pragma unroll
                for (int j = 0; j < 4; ++j) w[j] = L.w2[(mb * 16 + s) * NHID + kb * 16 + 4 * q + j];
                c = mma<BF>(L, wave, w, a1[kb], c);
            }
            h2[mb] = c;
        }
        float zp = 0.0f;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int j = 0; j < 4; ++j) zp += L.w3[mb * 16 + 4 * q + j] * silu_hw(h2[mb][j]);
        zp += __shfl_xor(zp, 16, 64);
        zp += __shfl_xor(zp, 32, 64);
        const float z = zp + b3;
        const float C = mlp_coefficient(z, a.c);
        const float wi = a.w ? a.w[i] : 1.0f;
        const float e = C - a.y[i];
        if (q == 0) loss += (double)wi * (double)e * (double)e;
        const float dz = 2.0f * wi * e * inv_b * a.c * sigmoid_hw(z) / LN2F;
        if (q == 0) gb3 += dz;

        // backward through layer 3 and the second activation
        f32x4 dh2[4], dh1[4];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int h = mb * 16 + 4 * q + j;
                gw3[mb][j] += silu_hw(h2[mb][j]) * dz;
                dh2[mb][j] = L.w3[h] * dz * dsilu_hw(h2[mb][j]);
                gb2[mb][j] += dh2[mb][j];
            }
        // dA1^T = W2^T dH2^T
#pragma unroll
        for (int hb = 0; hb < 4; ++hb) {
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                f32x4 w;
#pragma unroll
                for (int j = 0; j < 4; ++j) w[j] = L.w2[(kb * 16 + 4 * q + j) * NHID + hb * 16 + s];
                c = mma<BF>(L, wave, w, dh2[kb], c);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                dh1[hb][j] = c[j] * dsilu_hw(h1[hb][j]);
                gb1[hb][j] += dh1[hb][j];
            }
        }
        // dW2 = dH2^T A1 and dW1 = dH1^T X, contracting over the tile's samples
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                T1[(b * 16 + 4 * q + j) * 16 + s] = a1[b][j];
                T2[(b * 16 + 4 * q + j) * 16 + s] = dh2[b][j];
            }
        wave_lds_sync();
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            f32x4 A;
#pragma unroll
            for (int j = 0; j < 4; ++j) A[j] = T2[(mb * 16 + s) * 16 + 4 * q + j];
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                f32x4 B;
#pragma unroll
                for (int j = 0; j < 4; ++j) B[j] = T1[(nb * 16 + s) * 16 + 4 * q + j];
                gW2[mb][nb] = mma<BF>(L, wave, A, B, gW2[mb][nb]);
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) T2[(b * 16 + 4 * q + j) * 16 + s] = dh1[b][j];
        wave_lds_sync();
#pragma unroll
        for (int hb = 0; hb < 4; ++hb) {
            f32x4 A;
#pragma unroll
            for (int j = 0; j < 4; ++j) A[j] = T2[(hb * 16 + s) * 16 + 4 * q + j];
            gW1[hb] = mma<BF>(L, wave, A, xr, gW1[hb]);
        }
        wave_lds_sync();
    }

    // per-hidden-unit sums over the 16 sample lanes, scalars over the wave
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                gb1[b][j] += __shfl_xor(gb1[b][j], off, 64);
                gb2[b][j] += __shfl_xor(gb2[b][j], off, 64);
                gw3[b][j] += __shfl_xor(gw3[b][j], off, 64);
            }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        gb3 += __shfl_xor(gb3, off, 64);
        loss += __shfl_xor(loss, off, 64);
    }
    __syncthreads();
    for (int i = tid; i < NPARAM; i += TPB_T) L.u.acc[i] = 0.0f;
    __syncthreads();
    for (int w = 0; w < 4; ++w) {   // waves add in a fixed order
        if (wave == w) {
#pragma unroll
            for (int mb = 0; mb < 4; ++mb)
#pragma unroll
                for (int nb = 0; nb < 4; ++nb)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        L.u.acc[OFF_W2 + (mb * 16 + 4 * q + j) * NHID + nb * 16 + s] += gW2[mb][nb][j];
            if (s < NFEAT) {
#pragma unroll
                for (int hb = 0; hb < 4; ++hb)
#pragma unroll
                    for (int j = 0; j < 4; ++j) L.u.acc[OFF_W1 + (hb * 16 + 4 * q + j) * NFEAT + s] += gW1[hb][j];
            }
            if (s == 0) {
#pragma unroll
                for (int b = 0; b < 4; ++b)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int h = b * 16 + 4 * q + j;
                        L.u.acc[OFF_B1 + h] += gb1[b][j];
                        L.u.acc[OFF_B2 + h] += gb2[b][j];
                        L.u.acc[OFF_W3 + h] += gw3[b][j];
                    }
            }
            if (lane == 0) {
                L.u.acc[OFF_B3] += gb3;
                L.loss[w] = loss;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < NPARAM; i += TPB_T) a.part[(size_t)blockIdx.x * NPARAM + i] = L.u.acc[i];
    if (tid == 0) a.part_loss[blockIdx.x] = L.loss[0] + L.loss[1] + L.loss[2] + L.loss[3];
}

struct ReduceArgs {
    const float* part;
    const double* part_loss;
    int nparts;
    float* grad;          // gradient out, or null
    double* loss_out;     // sum of the partial losses, or null
    float *p, *m, *v;     // Adam state, used when adam != 0
    float lr, beta1, beta2, eps, bc1, bc2;
    int adam;
};

__global__ __launch_bounds__(TPB_T) void k_mlp_reduce(ReduceArgs a)
{
    const int i = blockIdx.x * TPB_T + threadIdx.x;
    if (i == 0 && a.loss_out) {
        double s = 0.0;
        for (int k = 0; k < a.nparts; ++k) s += a.part_loss[k];
        *a.loss_out = s;
    }
    if (i >= NPARAM) return;
    double g = 0.0;
    for (int k = 0; k < a.nparts; ++k) g += (double)a.part[(size_t)k * NPARAM + i];
    const float gf = (float)g;
    if (a.grad) a.grad[i] = gf;
    if (a.adam) {
        const float m = a.beta1 * a.m[i] + (1.0f - a.beta1) * gf;
        const float v = a.beta2 * a.v[i] + (1.0f - a.beta2) * gf * gf;
        a.m[i] = m;
        a.v[i] = v;
        a.p[i] -= a.lr * (m / a.bc1) / (sqrtf(v / a.bc2) + a.eps);
    }
}

__global__ __launch_bounds__(TPB_T) void k_mlp_forward(MlpWeights W, const float* x, int64_t n, float* out)
{
    __shared__ MlpLds L;
    mlp_load_lds(L, W);
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * TPB_T + threadIdx.x;
    const bool valid = i < n;
    float xn[NFEAT];
#pragma unroll
    for (int k = 0; k < NFEAT; ++k) xn[k] = valid ? x[i * NFEAT + k] : 0.0f;
    const float z = mlp_wave_forward(L, xn, W.b3);
    if (valid) out[i] = mlp_coefficient(z, W.c);
}

uint16_t host_bf16(float x)
{
    uint32_t b;
    std::memcpy(&b, &x, 4);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}

}  // namespace

// BF16 inference image of a parameter vector: -log2(e) W1 padded to
// [64][16] (inference works in the scaled variable g = -log2(e) h), W2.
void lbm_mlp_pack_bf16(const float* p, uint16_t* w1, uint16_t* w2)
{
    for (int h = 0; h < NHID; ++h)
        for (int k = 0; k < 16; ++k)
            w1[h * 16 + k] = k < NFEAT ? host_bf16(NEG_LOG2E * p[OFF_W1 + h * NFEAT + k]) : 0;
    for (int i = 0; i < NHID * NHID; ++i) w2[i] = host_bf16(p[OFF_W2 + i]);
}

struct lbm_mlp {
    hipStream_t stream = nullptr;
    float* p = nullptr;       // parameters
    float* m = nullptr;       // Adam moments
    float* v = nullptr;
    float* x = nullptr;       // training set
    float* y = nullptr;
    float* w = nullptr;
    int64_t n = 0;
    float* part = nullptr;
    double* part_loss = nullptr;
    int max_parts = 256;
    float* grad = nullptr;
    double* loss = nullptr;   // per-step loss history
    int loss_cap = 0;
    uint16_t* wb = nullptr;   // BF16 inference image: w1 [64*16], w2 [64*64]
    int64_t t = 0;            // Adam step count
};

#define HCHK(x)                                 \
    do {                                        \
        if ((x) != hipSuccess) {                \
            (void)hipGetLastError();            \
            return LBM_E_HIP;                   \
        }                                       \
    } while (0)

static uint64_t host_mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

extern "C" {

void lbm_mlp_destroy(lbm_mlp* m)
{
    if (!m) return;
    void* bufs[] = {m->p, m->m, m->v, m->x, m->y, m->w, m->part, m->part_loss, m->grad, m->loss, m->wb};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

lbm_status lbm_mlp_create(const float* params, uint64_t seed, lbm_mlp** out)
{
    if (!out) return LBM_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return LBM_E_HIP;
    }
    lbm_mlp* m = new (std::nothrow) lbm_mlp();
    if (!m) return LBM_E_OOM;
    std::vector<float> h(NPARAM, 0.0f);
    if (params) {
        std::memcpy(h.data(), params, sizeof(float) * NPARAM);
    } else {
        // N(0, 1/fan_in) by Box-Muller on a counter-based generator
        uint64_t ctr = seed;
        auto normal = [&]() {
            const double u1 = ((host_mix(ctr++) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
            const double u2 = (host_mix(ctr++) >> 11) * (1.0 / 9007199254740992.0);
            return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        };
        for (int i = 0; i < NHID * NFEAT; ++i) h[OFF_W1 + i] = (float)(normal() / std::sqrt((double)NFEAT));
        for (int i = 0; i < NHID * NHID; ++i) h[OFF_W2 + i] = (float)(normal() / std::sqrt((double)NHID));
    }
    bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&m->p, sizeof(float) * NPARAM) == hipSuccess;
    ok = ok && hipMalloc(&m->m, sizeof(float) * NPARAM) == hipSuccess;
    ok = ok && hipMalloc(&m->v, sizeof(float) * NPARAM) == hipSuccess;
    ok = ok && hipMalloc(&m->part, sizeof(float) * NPARAM * m->max_parts) == hipSuccess;
    ok = ok && hipMalloc(&m->part_loss, sizeof(double) * m->max_parts) == hipSuccess;
    ok = ok && hipMalloc(&m->grad, sizeof(float) * NPARAM) == hipSuccess;
    ok = ok && hipMalloc(&m->wb, sizeof(uint16_t) * (NHID * 16 + NHID * NHID)) == hipSuccess;
    ok = ok && hipMemcpy(m->p, h.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemset(m->m, 0, sizeof(float) * NPARAM) == hipSuccess;
    ok = ok && hipMemset(m->v, 0, sizeof(float) * NPARAM) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        lbm_mlp_destroy(m);
        return LBM_E_HIP;
    }
    *out = m;
    return LBM_OK;
}

lbm_status lbm_mlp_set_data(lbm_mlp* m, const float* x, const float* y, const float* w, int64_t n)
{
    if (!m || !x || !y || n <= 0) return LBM_E_ARG;
    for (float** b : {&m->x, &m->y, &m->w})
        if (*b) {
            (void)hipFree(*b);
            *b = nullptr;
        }
    m->n = 0;
    HCHK(hipMalloc(&m->x, sizeof(float) * NFEAT * n));
    HCHK(hipMalloc(&m->y, sizeof(float) * n));
    HCHK(hipMemcpy(m->x, x, sizeof(float) * NFEAT * n, hipMemcpyHostToDevice));
    HCHK(hipMemcpy(m->y, y, sizeof(float) * n, hipMemcpyHostToDevice));
    if (w) {
        HCHK(hipMalloc(&m->w, sizeof(float) * n));
        HCHK(hipMemcpy(m->w, w, sizeof(float) * n, hipMemcpyHostToDevice));
    }
    m->n = n;
    return LBM_OK;
}

static int grad_parts(const lbm_mlp* m, int batch)
{
    const int tiles = batch / 16;
    int g = (tiles + 3) / 4;
    return g < m->max_parts ? g : m->max_parts;
}

static GradArgs grad_args(lbm_mlp* m, int batch, float c)
{
    GradArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = m->x;
    a.y = m->y;
    a.w = m->w;
    a.n = m->n;
    a.batch = batch;
    a.p = m->p;
    a.c = c;
    a.part = m->part;
    a.part_loss = m->part_loss;
    return a;
}

lbm_status lbm_mlp_gradient(lbm_mlp* m, int64_t first, int32_t count, float c_les, int32_t use_bf16,
                            double* loss_sum, float* grad)
{
    if (!m || count <= 0 || count % 16 || first < 0 || first + count > m->n || !m->x) return LBM_E_ARG;
    GradArgs a = grad_args(m, count, c_les);
    a.first = first;
    const int parts = grad_parts(m, count);
    if (use_bf16) k_mlp_grad<true><<<parts, TPB_T, 0, m->stream>>>(a);
    else k_mlp_grad<false><<<parts, TPB_T, 0, m->stream>>>(a);
    HCHK(hipGetLastError());
    ReduceArgs r;
    std::memset(&r, 0, sizeof(r));
    r.part = m->part;
    r.part_loss = m->part_loss;
    r.nparts = parts;
    r.grad = m->grad;
    if (!m->loss || m->loss_cap < 1) {
        if (m->loss) (void)hipFree(m->loss);
        m->loss = nullptr;
        HCHK(hipMalloc(&m->loss, sizeof(double)));
        m->loss_cap = 1;
    }
    r.loss_out = m->loss;
    k_mlp_reduce<<<(NPARAM + TPB_T - 1) / TPB_T, TPB_T, 0, m->stream>>>(r);
    HCHK(hipGetLastError());
    if (grad) HCHK(hipMemcpyAsync(grad, m->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->stream));
    double ls = 0.0;
    HCHK(hipMemcpyAsync(&ls, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HCHK(hipStreamSynchronize(m->stream));
    if (loss_sum) *loss_sum = ls;
    return LBM_OK;
}

lbm_status lbm_mlp_train(lbm_mlp* m, int32_t steps, int32_t batch, float lr, float c_les, uint64_t seed,
                         double* loss)
{
    if (!m || steps < 0 || batch <= 0 || batch % 16 || !m->x) return LBM_E_ARG;
    if (steps == 0) return LBM_OK;
    if (m->loss_cap < steps) {
        if (m->loss) (void)hipFree(m->loss);
        m->loss = nullptr;
        HCHK(hipMalloc(&m->loss, sizeof(double) * steps));
        m->loss_cap = steps;
    }
    const int parts = grad_parts(m, batch);
    GradArgs a = grad_args(m, batch, c_les);
    a.random = 1;
    a.seed = seed;
    ReduceArgs r;
    std::memset(&r, 0, sizeof(r));
    r.part = m->part;
    r.part_loss = m->part_loss;
    r.nparts = parts;
    r.p = m->p;
    r.m = m->m;
    r.v = m->v;
    r.lr = lr;
    r.beta1 = 0.9f;
    r.beta2 = 0.999f;
    r.eps = 1e-8f;
    r.adam = 1;
    for (int k = 0; k < steps; ++k) {
        a.step = k;
        k_mlp_grad<true><<<parts, TPB_T, 0, m->stream>>>(a);
        m->t++;
        r.bc1 = (float)(1.0 - std::pow(0.9, (double)m->t));
        r.bc2 = (float)(1.0 - std::pow(0.999, (double)m->t));
        r.loss_out = m->loss + k;
        k_mlp_reduce<<<(NPARAM + TPB_T - 1) / TPB_T, TPB_T, 0, m->stream>>>(r);
    }
    HCHK(hipGetLastError());
    if (loss) {
        std::vector<double> h(steps);
        HCHK(hipMemcpyAsync(h.data(), m->loss, sizeof(double) * steps, hipMemcpyDeviceToHost, m->stream));
        HCHK(hipStreamSynchronize(m->stream));
        for (int k = 0; k < steps; ++k) loss[k] = h[k] / batch;
    }
    HCHK(hipStreamSynchronize(m->stream));
    return LBM_OK;
}

lbm_status lbm_mlp_get_params(lbm_mlp* m, float* params)
{
    if (!m || !params) return LBM_E_ARG;
    HCHK(hipStreamSynchronize(m->stream));
    HCHK(hipMemcpy(params, m->p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    return LBM_OK;
}

lbm_status lbm_mlp_set_params(lbm_mlp* m, const float* params)
{
    if (!m || !params) return LBM_E_ARG;
    HCHK(hipStreamSynchronize(m->stream));
    HCHK(hipMemcpy(m->p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    HCHK(hipMemset(m->m, 0, sizeof(float) * NPARAM));
    HCHK(hipMemset(m->v, 0, sizeof(float) * NPARAM));
    m->t = 0;
    return LBM_OK;
}

lbm_status lbm_mlp_forward(lbm_mlp* m, const float* x, int64_t n, float c_les, float* c_out)
{
    if (!m || !x || !c_out || n <= 0) return LBM_E_ARG;
    std::vector<float> p(NPARAM);
    lbm_status st = lbm_mlp_get_params(m, p.data());
    if (st != LBM_OK) return st;
    std::vector<uint16_t> wb(NHID * 16 + NHID * NHID);
    lbm_mlp_pack_bf16(p.data(), wb.data(), wb.data() + NHID * 16);
    HCHK(hipMemcpy(m->wb, wb.data(), sizeof(uint16_t) * wb.size(), hipMemcpyHostToDevice));
    float *dx = nullptr, *dout = nullptr;
    if (hipMalloc(&dx, sizeof(float) * NFEAT * n) != hipSuccess) {
        (void)hipGetLastError();
        return LBM_E_OOM;
    }
    if (hipMalloc(&dout, sizeof(float) * n) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(dx);
        return LBM_E_OOM;
    }
    lbm_status rs = LBM_OK;
    MlpWeights W;
    W.w1 = m->wb;
    W.w2 = m->wb + NHID * 16;
    W.b1 = m->p + OFF_B1;
    W.b2 = m->p + OFF_B2;
    W.w3 = m->p + OFF_W3;
    W.norm = nullptr;   // the caller passes normalised features
    W.b3 = p[OFF_B3];
    W.c = c_les;
    if (hipMemcpy(dx, x, sizeof(float) * NFEAT * n, hipMemcpyHostToDevice) != hipSuccess) rs = LBM_E_HIP;
    if (rs == LBM_OK) {
        k_mlp_forward<<<(unsigned)((n + TPB_T - 1) / TPB_T), TPB_T, 0, m->stream>>>(W, dx, n, dout);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(m->stream) != hipSuccess ||
            hipMemcpy(c_out, dout, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess)
            rs = LBM_E_HIP;
    }
    (void)hipFree(dx);
    (void)hipFree(dout);
    return rs;
}

}  // extern "C"
