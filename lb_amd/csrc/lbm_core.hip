// lbm_core.hip -- host side of the MI355X (gfx950) collide-stream core:
// flag-grid classification, device buffers, launches and the C ABI of
// include/lbm_core.h.  The device code is in lbm_kernels.hpp.
//
// Lattice in HBM: the caller's nx*ny*nz cells plus a one-cell solid halo on
// every axis without index wrap, so that populations streaming to or from
// outside the caller's lattice have a slot to live in (the reference reads
// and writes outside its arrays there, StreamingAB.h:51-56; the oracle
// leaves such populations untouched, which the halo reproduces).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

#include "../../include/lbm_closure.h"
#include "lbm_kernels.hpp"

using namespace lbm;

// closure.hip: BF16 inference image of a parameter vector
void lbm_mlp_pack_bf16(const float* p, uint16_t* w1, uint16_t* w2);

struct lbm_ctx {
    lbm_desc d;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    size_t N = 0;              // padded cells
    size_t NU = 0;             // caller's cells
    size_t elem = 0;           // bytes per stored population
    size_t rbytes = 0;         // bytes per arithmetic value
    Geo g{};
    bool flags_needed = false; // any non-fluid cell (halo included)
    int n_inlet = 0, n_eq = 0, n_outlet = 0, n_ghost = 0, n_pub = 0;
    int n_upd = 0;                // moving-wall and Zou-He inlet entries
    std::vector<int> ghost_off;  // launch boundaries of the ghost list (dependency levels)
    void* pops = nullptr;
    uint8_t* cls = nullptr;
    uint64_t* node = nullptr;
    int32_t* bcidx = nullptr;
    void* inl_state = nullptr;
    NodeDev* inl = nullptr;
    NodeDev* outl = nullptr;
    NodeDev* ghost = nullptr;
    void* pub_mac = nullptr;
    void* inl_post = nullptr;     // hand-over to the inlet rule (k_inlet_state, or deferred): f* and rho, u
    void* inl_mac = nullptr;
    void* pub_post = nullptr;
    void* macro = nullptr;     // 4 planes, arithmetic type, padded layout
    double* stage = nullptr;   // 5 planes FP64, caller layout
    double* partials = nullptr;
    std::vector<double> h_partials;
    int64_t steps = 0;
    bool initialized = false;
    bool first = false;
    bool nan_seen = false;
    bool diag_next = false;    // the next launched step reduces mass, momentum and wall force
    double last_ms = 0.0;
    size_t dev_bytes = 0;
    // learned closure
    bool mlp_ready = false;
    float* cl_params = nullptr;   // FP32 parameters
    uint16_t* cl_wb = nullptr;    // BF16 W1 [64][16] + W2 [64][64]
    float cl_b3 = 0.0f;
    float* cl_norm = nullptr;     // [4][12] mean, 1/sd, envelope low, high
    float* ccache = nullptr;      // coefficient per cell held between network evaluations
    int cl_every = 1;             // evaluate the network every cl_every steps
    int64_t cl_count = 0;         // steps since the last (re)start of the refresh cycle
    // time averaging (MomentTimeAvg)
    double* avg = nullptr;        // 4 FP64 planes, caller layout: sums of rho, ux, uy, uz
    int64_t avg_count = 0;        // states added
    bool avg_pending = false;     // the current state is counted but not yet added
    bool avg_each = false;        // launch_steps: count the state after every step
    // probe records taken without a synchronisation
    double* probe_buf = nullptr;  // [PROBE_CAP][4] on the device
    int probe_n = 0;              // records in probe_buf
    std::vector<double> probe_host;  // records drained from the device, oldest first
    // outlet omega damping
    int sp_axis = -1, sp_begin = 0, sp_interval = 1;
    // ErrorQuadratic on the device
    std::vector<uint8_t> h_flag1;   // flag == 1 per caller cell (the cells the norm sums over)
    uint8_t* flag1 = nullptr;       // the same on the device, uploaded by the first evaluation
    void* uerr = nullptr;           // previous velocity, 3 planes of the arithmetic type, caller layout
    // per-block partials of a diagnostics step go here instead of `partials` when set
    double* diag_dst = nullptr;
    // asynchronous snapshots: rho, u, omega into snap_dev on the context stream,
    // then a copy to pinned host memory on copy_stream while steps go on
    double* snap_dev = nullptr;       // 5 planes FP64, caller layout
    double* snap_host = nullptr;      // pinned, 5 planes
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_snap = nullptr, ev_snap_copy = nullptr;
    bool snap_pending = false;
    int64_t snap_step = 0;
    // Deferred interpolated outlet: the rule runs in the next reader of an
    // outlet cell, with the post-collision records of the step that produced
    // the state.  Records alternate between pub_post (even steps) and
    // pub_post2 (odd steps), so a step can read the previous step's while
    // writing its own.
    bool outl_deferred = false;
    bool outl_pending = false;      // the current state's outlet cells still need the rule
    void* pub_post2 = nullptr;
    // Deferred inlet: the moving-wall and Zou-He rules run in the next
    // step's kernel from the records (inl_post, inl_mac) of the step that
    // produced the state.
    bool inl_deferred = false;
    bool inl_pending = false;       // the inlet cells' stored state still needs its rule
};

static constexpr int PROBE_CAP = 4096;

// The interpolated outlet's weight c_s = 1/sqrt(3) (OutletDensityInterpolated.h:226)
// in the constants' precision.
template <class R>
static R outlet_cs(const lbm_ctx* c)
{
    return (c->d.compat & LBM_COMPAT_F32CONST) ? (R)(1.f / sqrtf(3.f)) : (R)(1.0 / std::sqrt(3.0));
}

// Post-collision records written by step t (deferred outlet: alternating buffers).
static void* pub_records(const lbm_ctx* c, int64_t t)
{
    return (c->outl_deferred && (t & 1)) ? c->pub_post2 : c->pub_post;
}

#define HIPCHK(x)                                   \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) return LBM_E_HIP;     \
    } while (0)

static lbm_status dev_alloc(lbm_ctx* c, void** p, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return LBM_E_OOM;
    }
    c->dev_bytes += bytes;
    return LBM_OK;
}

template <class ST>
static void fill_common(const lbm_ctx* c, typename ST::T** p, typename ST::R* rest, Consts<typename ST::R>& K)
{
    using R = typename ST::R;
    K = make_consts<R>(c->d.compat);
    for (int k = 0; k < 4; ++k) rest[k] = ST::dev ? (R)(K.w[k] * (R)c->d.rho0) : R(0);
    typename ST::T* base = (typename ST::T*)c->pops;
    for (int q = 0; q < Q; ++q) p[q] = base + (size_t)q * c->N;
}

static MlpWeights closure_weights(const lbm_ctx* c)
{
    MlpWeights W;
    W.w1 = c->cl_wb;
    W.w2 = c->cl_wb ? c->cl_wb + NHID * 16 : nullptr;
    W.b1 = c->cl_params ? c->cl_params + OFF_B1 : nullptr;
    W.b2 = c->cl_params ? c->cl_params + OFF_B2 : nullptr;
    W.w3 = c->cl_params ? c->cl_params + OFF_W3 : nullptr;
    W.norm = c->cl_norm;
    W.b3 = c->cl_b3;
    W.c = (float)c->d.c_les;
    return W;
}

template <class ST>
static AuxArgs<ST> aux_args(const lbm_ctx* c)
{
    using R = typename ST::R;
    AuxArgs<ST> a;
    std::memset(&a, 0, sizeof(a));
    fill_common<ST>(c, a.p, a.rest, a.K);
    a.cls = c->cls;
    a.bcidx = c->bcidx;
    a.inl_state = (R*)c->inl_state;
    a.macro = (R*)c->macro;
    a.g = c->g;
    a.tau = (R)c->d.tau;
    a.cles = (R)c->d.c_les;
    a.omega0 = R(1) / (R)c->d.tau;
    a.closure = c->d.closure;
    a.lit = c->d.arith == LBM_ARITH_LITERAL;
    a.odd = (int)(c->steps & 1);
    a.first = c->first ? 1 : 0;
    a.io = c->stage;
    a.sp_axis = c->sp_axis;
    a.sp_begin = c->sp_begin;
    a.sp_deriv = ((R)c->d.sponge_high - (R)c->d.sponge_low) / (R)c->sp_interval;
    a.outl_pending = c->outl_pending ? 1 : 0;
    a.outl = c->outl;
    a.pub_prev = (const R*)pub_records(c, c->steps - 1);
    a.cs = outlet_cs<R>(c);
    a.crev20 = (c->d.compat & LBM_COMPAT_CREV20) ? 1 : 0;
    a.inl_pending = c->inl_pending ? 1 : 0;
    a.inl = c->inl;
    a.inl_post = (const R*)c->inl_post;
    a.inl_mac = (const R*)c->inl_mac;
    return a;
}

static unsigned blocks_for(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }

template <class ST>
static PatchArgs<ST> patch_args(const lbm_ctx* c)
{
    using R = typename ST::R;
    PatchArgs<ST> a;
    std::memset(&a, 0, sizeof(a));
    fill_common<ST>(c, a.p, a.rest, a.K);
    a.cls = c->cls;
    a.bcidx = c->bcidx;
    a.inl_state = (R*)c->inl_state;
    a.pub_mac = (const R*)c->pub_mac;
    a.pub_post = (const R*)c->pub_post;
    a.g = c->g;
    a.nu = (R)c->d.nu;
    // the outlet's interpolation weight (OutletDensityInterpolated.h:226) is
    // 1/sqrt(3) in the constants' precision
    a.cs = outlet_cs<R>(c);
    a.inletrho = (c->d.compat & LBM_COMPAT_INLETRHO) ? 1 : 0;
    a.crev20 = (c->d.compat & LBM_COMPAT_CREV20) ? 1 : 0;
    return a;
}

// Boundary rules that read what other cells produced in the same step, in
// the reference's order: inlet, outlet, periodic.
template <class ST>
static lbm_status launch_patches(lbm_ctx* c, int odd_next)
{
    PatchArgs<ST> a = patch_args<ST>(c);
    a.odd = odd_next;
    a.pub_post = (const typename ST::R*)pub_records(c, c->steps);   // this step's records
    if (c->n_upd && !c->inl_deferred) {
        PatchArgs<ST> u = a;
        u.nodes = c->inl;
        u.n = c->n_inlet;
        u.pub_mac = (const typename ST::R*)c->inl_mac;
        u.pub_post = (const typename ST::R*)c->inl_post;
        k_inlet_state<ST><<<blocks_for(u.n), TPB, 0, c->stream>>>(u);
    }
    if (c->n_eq) {
        a.nodes = c->inl;
        a.n = c->n_inlet;
        k_inlet_eq<ST><<<blocks_for(a.n), TPB, 0, c->stream>>>(a);
    }
    if (c->n_outlet && !c->outl_deferred) {
        a.nodes = c->outl;
        a.n = c->n_outlet;
        k_outlet<ST><<<blocks_for(a.n), TPB, 0, c->stream>>>(a);
    }
    for (size_t l = 0; l + 1 < c->ghost_off.size(); ++l) {
        a.nodes = c->ghost + c->ghost_off[l];
        a.n = c->ghost_off[l + 1] - c->ghost_off[l];
        k_periodic<ST><<<blocks_for(a.n), TPB, 0, c->stream>>>(a);
    }
    HIPCHK(hipGetLastError());
    return LBM_OK;
}

// Add the pending state to the time-average sums with its own pass.
template <class ST>
static lbm_status avg_flush_st(lbm_ctx* c)
{
    if (!c->avg_pending) return LBM_OK;
    AuxArgs<ST> a = aux_args<ST>(c);
    a.io = c->avg;
    k_moment_add<ST><<<blocks_for(c->NU), TPB, 0, c->stream>>>(a);
    HIPCHK(hipGetLastError());
    c->avg_pending = false;
    return LBM_OK;
}

// Seed of the stochastic-rounding stores of step t (any fixed bijective mix
// of the step counter: runs with the same inputs give the same results).
static uint32_t step_seed(uint64_t t)
{
    uint64_t x = t * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    x ^= x >> 31;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    return (uint32_t)(x >> 32);
}

// The boundary-aware step: 16-bit storage takes the kernel held to 7 waves
// per SIMD, the other formats the plain entry point.
template <class ST, int COLL, bool LIT, bool ODD, bool INL>
static void launch_flagged(dim3 grid, dim3 block, hipStream_t s, const StepArgs<ST>& a)
{
    if constexpr (sizeof(typename ST::T) == 2)
        k_step_w7<ST, COLL, LIT, ODD, INL><<<grid, block, 0, s>>>(a);
    else
        k_step<ST, COLL, LIT, ODD, true, false, INL><<<grid, block, 0, s>>>(a);
}

template <class ST, int COLL, bool LIT>
static lbm_status launch_steps(lbm_ctx* c, int nsteps)
{
    using R = typename ST::R;
    StepArgs<ST> a;
    std::memset(&a, 0, sizeof(a));
    fill_common<ST>(c, a.p, a.rest, a.K);
    a.cls = c->cls;
    a.node = c->node;
    a.bcidx = c->bcidx;
    a.inl_state = (R*)c->inl_state;
    a.inl = c->inl;
    a.pub_mac = (R*)c->pub_mac;
    a.pub_post = (R*)c->pub_post;
    a.inl_post = (R*)c->inl_post;
    a.inl_mac = (R*)c->inl_mac;
    a.macro = (const R*)c->macro;
    a.partials = c->diag_dst ? c->diag_dst : c->partials;
    a.ccache = c->ccache;
    a.g = c->g;
    a.sp_axis = c->sp_axis;
    a.sp_begin = c->sp_begin;
    a.sp_deriv = ((R)c->d.sponge_high - (R)c->d.sponge_low) / (R)c->sp_interval;
    a.outl = c->outl;
    a.cs = outlet_cs<R>(c);
    a.omega0 = R(1) / (R)c->d.tau;   // 1.f / tau (OmegaNo.h:44, NonDim...h:114)
    a.tau = (R)c->d.tau;
    a.cles = (R)c->d.c_les;
    a.crev20 = (c->d.compat & LBM_COMPAT_CREV20) ? 1 : 0;
    a.closure = c->d.closure;
    const bool mlp = c->d.closure == LBM_CLOSURE_MLP;
    MlpStepArgs<ST> ma;
    if (mlp) {
        if (!c->mlp_ready) return LBM_E_STATE;
        ma.mlp = closure_weights(c);
    }
    const bool patches = (c->n_upd && !c->inl_deferred) || c->n_eq || (c->n_outlet && !c->outl_deferred) ||
                         c->n_ghost;
    const dim3 grid((unsigned)c->g.ntiles), block(TPB);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int s = 0; s < nsteps; ++s) {
        a.first = c->first ? 1 : 0;
        // a counted state is added by the step that reads it; the network step
        // does not fuse the add, so it gets its own pass first
        if (mlp && c->avg_pending && c->cl_count % c->cl_every == 0) {
            lbm_status st = avg_flush_st<ST>(c);
            if (st != LBM_OK) return st;
        }
        a.avg = c->avg_pending ? c->avg : nullptr;
        c->avg_pending = false;
        a.diag = (s == 0 && c->diag_next) ? 1 : 0;
        a.pub_post = (R*)pub_records(c, c->steps);
        a.pub_prev = (const R*)pub_records(c, c->steps - 1);
        a.outl_pending = c->outl_pending ? 1 : 0;
        a.inl_pending = c->inl_pending ? 1 : 0;
        if constexpr (ST::sr) a.seed = step_seed((uint64_t)c->steps);
        const bool odd = (c->steps & 1) != 0;
        if constexpr (sizeof(R) == 4 && !LIT && COLL <= LBM_COLLISION_CUM2015) {
            if (mlp && c->cl_count % c->cl_every == 0) {
                // network evaluated in this step; coefficients kept for the next ones
                ma.s = a;
                if (c->flags_needed) {
                    if (odd) k_step_mlp<ST, COLL, true, true><<<grid, block, 0, c->stream>>>(ma);
                    else k_step_mlp<ST, COLL, false, true><<<grid, block, 0, c->stream>>>(ma);
                } else {
                    if (odd) k_step_mlp<ST, COLL, true, false><<<grid, block, 0, c->stream>>>(ma);
                    else k_step_mlp<ST, COLL, false, false><<<grid, block, 0, c->stream>>>(ma);
                }
            } else if (mlp) {
                if (c->flags_needed) {
                    if (odd) k_step<ST, COLL, false, true, true, true><<<grid, block, 0, c->stream>>>(a);
                    else k_step<ST, COLL, false, false, true, true><<<grid, block, 0, c->stream>>>(a);
                } else {
                    if (odd) k_step<ST, COLL, false, true, false, true><<<grid, block, 0, c->stream>>>(a);
                    else k_step<ST, COLL, false, false, false, true><<<grid, block, 0, c->stream>>>(a);
                }
            }
            if (mlp) c->cl_count++;
        }
        if (!mlp) {
            if (c->flags_needed && c->inl_deferred) {
                if (odd) launch_flagged<ST, COLL, LIT, true, true>(grid, block, c->stream, a);
                else launch_flagged<ST, COLL, LIT, false, true>(grid, block, c->stream, a);
            } else if (c->flags_needed) {
                if (odd) launch_flagged<ST, COLL, LIT, true, false>(grid, block, c->stream, a);
                else launch_flagged<ST, COLL, LIT, false, false>(grid, block, c->stream, a);
            } else {
                if (odd) k_step<ST, COLL, LIT, true, false><<<grid, block, 0, c->stream>>>(a);
                else k_step<ST, COLL, LIT, false, false><<<grid, block, 0, c->stream>>>(a);
            }
        }
        HIPCHK(hipGetLastError());
        if (patches) {
            lbm_status st = launch_patches<ST>(c, odd ? 0 : 1);
            if (st != LBM_OK) return st;
        }
        c->first = false;
        c->steps++;
        c->outl_pending = c->outl_deferred;
        c->inl_pending = c->inl_deferred;
        if (c->avg_each) {
            c->avg_pending = true;
            c->avg_count++;
        }
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    return LBM_OK;
}

template <class ST, bool LIT>
static lbm_status launch_coll(lbm_ctx* c, int nsteps)
{
    switch (c->d.collision) {
    case LBM_COLLISION_CUM2015: return launch_steps<ST, LBM_COLLISION_CUM2015, LIT>(c, nsteps);
    case LBM_COLLISION_CUM2017: return launch_steps<ST, LBM_COLLISION_CUM2017, LIT>(c, nsteps);
    case LBM_COLLISION_COMBINED: return launch_steps<ST, LBM_COLLISION_COMBINED, LIT>(c, nsteps);
    default: return launch_steps<ST, LBM_COLLISION_SRT, LIT>(c, nsteps);
    }
}

template <class ST>
static lbm_status launch_by_coll(lbm_ctx* c, int nsteps)
{
    return c->d.arith == LBM_ARITH_LITERAL ? launch_coll<ST, true>(c, nsteps) : launch_coll<ST, false>(c, nsteps);
}

template <class ST>
static lbm_status launch_by_coll_fast(lbm_ctx* c, int nsteps)
{
    return launch_coll<ST, false>(c, nsteps);
}

// Dispatch a functor templated on the storage policy.
template <class F>
static lbm_status with_storage(const lbm_ctx* c, F&& fn)
{
    switch (c->d.storage) {
    case LBM_STORAGE_F64: return fn(StoreF64{});
    case LBM_STORAGE_F32: return fn(StoreF32{});
    // the stochastic-rounding formats share the layout; only the step's
    // streaming stores round differently
    case LBM_STORAGE_F16DEV:
    case LBM_STORAGE_F16DEV_SR: return fn(StoreF16Dev{});
    case LBM_STORAGE_BF16DEV:
    case LBM_STORAGE_BF16DEV_SR: return fn(StoreBF16Dev{});
    default: return LBM_E_ARG;
    }
}

// ---------------------------------------------------------------------------
// Flag-grid classification
// ---------------------------------------------------------------------------

namespace {

struct Layout {
    int ux, uy, uz;   // caller's lattice
    int px, py, pz;   // halo per axis
    int nx, ny, nz;   // padded
    bool w[3];
    size_t user(int x, int y, int z) const
    {
        return (size_t)x + (size_t)ux * ((size_t)y + (size_t)uy * (size_t)z);
    }
    size_t pad(int x, int y, int z) const
    {
        return (size_t)(x + px) + (size_t)nx * ((size_t)(y + py) + (size_t)ny * (size_t)(z + pz));
    }
    // neighbour (x+cx, y+cy, z+cz) on the caller's lattice, with index wrap
    bool nbr(int x, int y, int z, int cx, int cy, int cz, int& ox, int& oy, int& oz) const
    {
        ox = x + cx;
        oy = y + cy;
        oz = z + cz;
        if (ox < 0 || ox >= ux) { if (!w[0]) return false; ox = (ox + ux) % ux; }
        if (oy < 0 || oy >= uy) { if (!w[1]) return false; oy = (oy + uy) % uy; }
        if (oz < 0 || oz >= uz) { if (!w[2]) return false; oz = (oz + uz) % uz; }
        return true;
    }
};

enum Kind : uint8_t { K_SOLID, K_STREAM, K_INLET };

struct Plan {
    std::vector<uint8_t> cls;
    std::vector<uint64_t> node;
    std::vector<int32_t> bcidx;
    std::vector<NodeDev> inl, outl, ghost;
    std::vector<int> ghost_off;
    int n_eq = 0;
    int n_pub = 0;
    bool flags_needed = false;
    // interpolated outlet applied by the next reader of the cell (see lbm_ctx)
    bool defer_outlet = false;
    bool defer_inlet = false;
};

NodeDev node_dev(const Layout& L, const lbm_bc_node& b)
{
    NodeDev n;
    n.x = b.x + L.px;
    n.y = b.y + L.py;
    n.z = b.z + L.pz;
    n.kind = b.kind;
    n.nx = b.nx;
    n.ny = b.ny;
    n.nz = b.nz;
    n.ux = b.ux;
    n.uy = b.uy;
    n.uz = b.uz;
    n.rho = b.rho;
    n.regular = b.regular;
    n.s0 = n.s1 = -1;
    n.partner = -1;
    return n;
}

}  // namespace

static lbm_status classify(const lbm_desc* d, const Layout& L, Plan& P)
{
    const size_t NU = (size_t)L.ux * L.uy * L.uz;
    const size_t N = (size_t)L.nx * L.ny * L.nz;
    std::vector<uint8_t> kind(NU);
    for (size_t i = 0; i < NU; ++i) {
        const int F = d->flags[i];
        if (F == 0) kind[i] = K_SOLID;
        else if (F > 0 || F == -2 || F == -3) kind[i] = K_STREAM;   // StreamingAB.h:49
        else if (F == -1) kind[i] = K_INLET;
        else return LBM_E_UNSUPPORTED;   // other negative flags (symmetry planes) are not built here
    }
    std::vector<uint8_t> wall(NU, 0);
    std::vector<int32_t> inlet_of(NU, -1), pub_of(NU, -1);
    std::vector<uint8_t> has_outlet(NU, 0);
    std::vector<int32_t> outlet_entry(NU, -1);
    P.inl.clear();
    P.outl.clear();
    P.ghost.clear();
    P.n_eq = 0;
    P.n_pub = 0;
    auto pub_slot = [&](size_t iu) -> int32_t {
        if (pub_of[iu] < 0) pub_of[iu] = P.n_pub++;
        return pub_of[iu];
    };
    // slot of a neighbour whose rho/u/f* a patch reads: -1 for solid cells
    // (zero in the reference), error for inlet cells (not published here)
    auto ref_slot = [&](int x, int y, int z, int32_t& slot) -> lbm_status {
        const size_t j = L.user(x, y, z);
        if (kind[j] == K_SOLID) slot = -1;
        else if (kind[j] == K_INLET) return LBM_E_UNSUPPORTED;
        else slot = pub_slot(j);
        return LBM_OK;
    };

    if (d->wall_mesh != 0 && d->wall_mesh != 1) return LBM_E_ARG;
    // BounceBackWallHalfMesh (BounceBackWallHalfMesh.h:38-50): every cell with flag 2
    if (d->wall_mesh)
        for (size_t i = 0; i < NU; ++i) wall[i] = d->flags[i] == 2 ? 1 : 0;
    for (int n = 0; n < d->n_bc; ++n) {
        const lbm_bc_node& b = d->bc[n];
        if (b.x < 0 || b.y < 0 || b.z < 0 || b.x >= L.ux || b.y >= L.uy || b.z >= L.uz) return LBM_E_ARG;
        const size_t i = L.user(b.x, b.y, b.z);
        const int F = d->flags[i];
        switch (b.kind) {
        case LBM_BC_WALL:
            if (F <= 0) return LBM_E_ARG;  // wall list holds fluid nodes only (geometryMesherBoundary.h:540-542)
            if (!d->wall_mesh) wall[i] = 1;
            break;
        case LBM_BC_INLET_MOVING_WALL:
        case LBM_BC_INLET_ZOU_HE:
        case LBM_BC_INLET_EQUILIBRIUM: {
            if (F != -1 || inlet_of[i] >= 0) return LBM_E_ARG;
            NodeDev e = node_dev(L, b);
            if (b.kind == LBM_BC_INLET_EQUILIBRIUM) {
                // neighbours x - n and x - 2n along +-x (InletVelocityEquilibrium.h:58-97)
                const int sx = (b.nx == 1.0f && b.ny == 0.0f && b.nz == 0.0f) ? 1
                             : (b.nx == -1.0f && b.ny == 0.0f && b.nz == 0.0f) ? -1 : 0;
                int x1, y1, z1, x2, y2, z2;
                if (sx != 0 && L.nbr(b.x, b.y, b.z, -sx, 0, 0, x1, y1, z1) &&
                    L.nbr(b.x, b.y, b.z, -2 * sx, 0, 0, x2, y2, z2)) {
                    lbm_status st = ref_slot(x1, y1, z1, e.s0);
                    if (st == LBM_OK) st = ref_slot(x2, y2, z2, e.s1);
                    if (st != LBM_OK) return st;
                    P.n_eq++;
                } else {
                    e.kind = -1;   // the reference applies nothing to this node
                }
            }
            inlet_of[i] = (int32_t)P.inl.size();
            P.inl.push_back(e);
            break;
        }
        case LBM_BC_OUTLET_INTERPOLATED: {
            if (kind[i] == K_SOLID) return LBM_E_ARG;
            if (kind[i] == K_INLET || has_outlet[i]) return LBM_E_UNSUPPORTED;
            has_outlet[i] = 1;
            const int sx = (b.nx == 1.0f && b.ny == 0.0f && b.nz == 0.0f) ? 1
                         : (b.nx == -1.0f && b.ny == 0.0f && b.nz == 0.0f) ? -1 : 0;
            if (sx == 0) break;   // other normals: no action (OutletDensityInterpolated.h:249)
            NodeDev e = node_dev(L, b);
            e.s0 = pub_slot(i);
            int xu, yu, zu;
            if (L.nbr(b.x, b.y, b.z, -sx, 0, 0, xu, yu, zu)) {
                lbm_status st = ref_slot(xu, yu, zu, e.s1);
                if (st != LBM_OK) return st;
            } else {
                e.s1 = -2;   // only the irregular-node copy applies
            }
            // irregular +x nodes copy f_crev into f_q where (x, y - c_y, z - c_z) is solid
            // (OutletDensityInterpolated.h:178-193); cells off the lattice are skipped
            uint32_t mask = 0;
            if (sx == 1 && !b.regular)
                for (int q = 0; q < Q; ++q) {
                    int x1, y1, z1;
                    if (L.nbr(b.x, b.y, b.z, 0, -CY[q], -CZ[q], x1, y1, z1) && kind[L.user(x1, y1, z1)] == K_SOLID)
                        mask |= 1u << q;
                }
            e.partner = (int32_t)mask;
            outlet_entry[i] = (int32_t)P.outl.size();
            P.outl.push_back(e);
            break;
        }
        case LBM_BC_OUTLET_EQUILIBRIUM:
        case LBM_BC_OUTLET_NEIGHBOUR_EQ:
        case LBM_BC_OUTLET_NEIGHBOUR_EQ_RF: {
            if (kind[i] == K_SOLID) return LBM_E_ARG;
            if (kind[i] == K_INLET || has_outlet[i]) return LBM_E_UNSUPPORTED;
            has_outlet[i] = 1;
            NodeDev e = node_dev(L, b);
            e.s0 = pub_slot(i);   // the node's own rho, u
            uint32_t mask = 0;
            if (b.kind != LBM_BC_OUTLET_EQUILIBRIUM) {
                for (int q = 0; q < Q; ++q) {
                    const float nc = b.nx * (float)CX[q] + b.ny * (float)CY[q] + b.nz * (float)CZ[q];
                    if (!(nc < 0.0f)) continue;
                    int x1, y1, z1;
                    // the reference reads outside its arrays here (no bounds checks); skipped
                    if (!L.nbr(b.x, b.y, b.z, CX[q], CY[q], CZ[q], x1, y1, z1)) continue;
                    const size_t j = L.user(x1, y1, z1);
                    if (!b.regular && !(d->flags[j] > 0)) continue;
                    if (kind[j] == K_INLET) return LBM_E_UNSUPPORTED;
                    if (kind[j] != K_SOLID) pub_slot(j);
                    mask |= 1u << q;
                }
            }
            e.partner = (int32_t)mask;
            P.outl.push_back(e);
            break;
        }
        case LBM_BC_PERIODIC_GHOST: {
            if (kind[i] == K_SOLID) return LBM_E_ARG;
            if (kind[i] == K_INLET) return LBM_E_UNSUPPORTED;
            const float ax = std::fabs(b.nx), ay = std::fabs(b.ny), az = std::fabs(b.nz);
            const int axis = (ax == 1.0f && ay == 0.0f && az == 0.0f) ? 0
                           : (ax == 0.0f && ay == 1.0f && az == 0.0f) ? 1
                           : (ax == 0.0f && ay == 0.0f && az == 1.0f) ? 2 : -1;
            if (axis < 0) break;   // no action (Periodic.h:136)
            const int ext = axis == 0 ? L.ux : axis == 1 ? L.uy : L.uz;
            if (b.per_idx < 0 || b.per_idx >= ext) return LBM_E_ARG;
            NodeDev e = node_dev(L, b);
            const int qx = axis == 0 ? b.per_idx : b.x, qy = axis == 1 ? b.per_idx : b.y,
                      qz = axis == 2 ? b.per_idx : b.z;
            e.partner = (int32_t)L.pad(qx, qy, qz);
            e.s0 = b.regular ? -1 : pub_slot(i);
            P.ghost.push_back(e);
            break;
        }
        default:
            return LBM_E_ARG;
        }
    }
    // a neighbour-equilibrium entry must not read a cell that another outlet
    // entry of the same launch writes
    for (const NodeDev& e : P.outl) {
        if (e.kind != LBM_BC_OUTLET_NEIGHBOUR_EQ && e.kind != LBM_BC_OUTLET_NEIGHBOUR_EQ_RF) continue;
        for (int q = 0; q < Q; ++q) {
            if (!(((uint32_t)e.partner >> q) & 1u)) continue;
            int x1, y1, z1;
            L.nbr(e.x - L.px, e.y - L.py, e.z - L.pz, CX[q], CY[q], CZ[q], x1, y1, z1);
            if (has_outlet[L.user(x1, y1, z1)]) return LBM_E_UNSUPPORTED;
        }
    }
    // flag -1 cells without an inlet entry keep their state and stream it
    for (size_t i = 0; i < NU; ++i) {
        if (kind[i] == K_INLET && inlet_of[i] < 0) {
            NodeDev e{};
            const int x = (int)(i % L.ux), y = (int)((i / L.ux) % L.uy), z = (int)(i / ((size_t)L.ux * L.uy));
            e.x = x + L.px;
            e.y = y + L.py;
            e.z = z + L.pz;
            e.kind = -1;
            e.s0 = e.s1 = e.partner = -1;
            inlet_of[i] = (int32_t)P.inl.size();
            P.inl.push_back(e);
        }
    }

    // Ghost entries act in list order (Periodic.h:50-138): an entry must run
    // after every earlier entry that writes its cell or its partner, or reads
    // its cell as a partner.  Entries of one level are independent.
    if (!P.ghost.empty()) {
        std::unordered_map<int64_t, int> last_write, last_read;
        std::vector<int> level(P.ghost.size());
        int max_level = 0;
        for (size_t k = 0; k < P.ghost.size(); ++k) {
            const NodeDev& e = P.ghost[k];
            const int64_t self = (int64_t)e.x + (int64_t)L.nx * ((int64_t)e.y + (int64_t)L.ny * e.z);
            const int64_t part = e.partner;
            int lv = 0;
            auto bump = [&](std::unordered_map<int64_t, int>& m, int64_t key) {
                auto it = m.find(key);
                if (it != m.end()) lv = std::max(lv, it->second + 1);
            };
            bump(last_write, self);
            bump(last_write, part);
            bump(last_read, self);
            level[k] = lv;
            last_write[self] = std::max(last_write.count(self) ? last_write[self] : 0, lv);
            if (part != self) last_read[part] = std::max(last_read.count(part) ? last_read[part] : 0, lv);
            max_level = std::max(max_level, lv);
        }
        std::vector<NodeDev> sorted;
        sorted.reserve(P.ghost.size());
        P.ghost_off.assign(1, 0);
        for (int lv = 0; lv <= max_level; ++lv) {
            for (size_t k = 0; k < P.ghost.size(); ++k)
                if (level[k] == lv) sorted.push_back(P.ghost[k]);
            P.ghost_off.push_back((int)sorted.size());
        }
        P.ghost.swap(sorted);
    } else {
        P.ghost_off.clear();
    }

    // cell classes and stream masks on the padded lattice
    // An interpolated outlet reads only its own state and records the step
    // published, so its rule can run in whichever kernel next reads the cell
    // (the next step's load or a read-back) instead of a patch kernel over an
    // x-plane of cells, each of whose populations sits on its own cache line.
    // Ghost layers (which read outlet cells as partners or write them after the
    // outlet rule), the other outlet kinds and the learned closure keep the
    // patch kernel.  So do the 16-bit storage formats: with the rule in the
    // step kernel their obstacle-flow step took 1.22-1.24 ms against 1.00 ms
    // with the patch kernel (512x256x256 lattice, one MI355X), while FP32
    // gained 2 % and FP64 9 %.
    const bool wide = d->storage == LBM_STORAGE_F64 || d->storage == LBM_STORAGE_F32;
    P.defer_outlet = wide && !P.outl.empty() && P.ghost.empty() && d->closure != LBM_CLOSURE_MLP;
    // The moving-wall and Zou-He rules read only the records of the step
    // that produced the state (the cell's own post-collision populations and
    // rho, u), so they can run in the next step's kernel, whose inlet cells
    // read their state anyway.  That pays where the inlet cells fill whole
    // rows of a tile (an inlet plane normal to y or z: a cavity lid) -- one
    // launch per step instead of two, +10 % on a 64^3 cavity.  On a plane
    // normal to x each wave holds one inlet cell, whose rule holds up the
    // other 63 lanes: the obstacle workload's step took 0.7 % (FP32) to 8 %
    // (FP16) longer than with the patch kernel, so such inlets keep it (an
    // equilibrium inlet acts only on x normals, so it is never deferred).
    // Ghost layers (whose partners can be inlet cells) and the learned
    // closure keep the patch kernel too.
    bool rule = false;
    P.defer_inlet = P.ghost.empty() && d->closure != LBM_CLOSURE_MLP;
    for (const NodeDev& e : P.inl) {
        if (e.nx != 0.0f) P.defer_inlet = false;
        if (e.kind == LBM_BC_INLET_MOVING_WALL || e.kind == LBM_BC_INLET_ZOU_HE) rule = true;
    }
    P.defer_inlet = P.defer_inlet && rule;
    for (const NodeDev& e : P.outl)
        if (e.kind != LBM_BC_OUTLET_INTERPOLATED) P.defer_outlet = false;
    P.cls.assign(N, CL_SOLID);
    P.node.assign(N, 0);
    P.bcidx.assign(N, -1);
    P.flags_needed = N != NU;
    for (int z = 0; z < L.uz; ++z)
        for (int y = 0; y < L.uy; ++y)
            for (int x = 0; x < L.ux; ++x) {
                const size_t iu = L.user(x, y, z), ip = L.pad(x, y, z);
                if (kind[iu] == K_SOLID) {
                    P.flags_needed = true;
                    continue;
                }
                uint32_t S = 0, K = 0;
                for (int q = 1; q < Q; ++q) {
                    int ox, oy, oz;
                    // source x - c_q: solid -> bounce-back (wall nodes) or untouched;
                    // outside the lattice -> untouched
                    if (!L.nbr(x, y, z, -CX[q], -CY[q], -CZ[q], ox, oy, oz)) {
                        if (wall[iu]) return LBM_E_UNSUPPORTED;
                        S |= 1u << q;
                    } else if (kind[L.user(ox, oy, oz)] == K_SOLID) {
                        S |= 1u << q;
                    }
                    // destination x + c_q: solid, inlet or outside -> no streaming write
                    if (!L.nbr(x, y, z, CX[q], CY[q], CZ[q], ox, oy, oz) ||
                        kind[L.user(ox, oy, oz)] != K_STREAM)
                        K |= 1u << q;
                }
                // the Combined operator's 2017 branch: every non-solid cell whose flag
                // is not 1 (CollisionCumD3Q27TurbulentCombined.h:1071-1076)
                const bool c17 = d->collision == LBM_COLLISION_COMBINED && d->flags[iu] != 1;
                P.node[ip] = (uint64_t)S | ((uint64_t)K << 32) | ((uint64_t)c17 << NW_CUM2017) |
                             ((uint64_t)wall[iu] << 63);
                if (kind[iu] == K_INLET) {
                    P.cls[ip] = CL_INLET;
                    P.bcidx[ip] = inlet_of[iu];
                    P.flags_needed = true;
                } else if (P.defer_outlet && outlet_entry[iu] >= 0) {
                    P.cls[ip] = CL_OUTL;   // publishes through its entry's own slot
                    P.bcidx[ip] = outlet_entry[iu];
                    P.flags_needed = true;
                } else if (pub_of[iu] >= 0) {
                    P.cls[ip] = CL_PUB;
                    P.bcidx[ip] = pub_of[iu];
                    P.flags_needed = true;
                } else if (S != 0 || K != 0 || wall[iu] || c17) {
                    P.cls[ip] = CL_BND;
                    P.flags_needed = true;
                } else {
                    P.cls[ip] = CL_FLUID;
                }
            }
    return LBM_OK;
}

// Upload host FP64 planes (caller layout) into the padded macro buffer.
template <class R>
static lbm_status upload_macro(lbm_ctx* c, const double* rho, const double* ux, const double* uy, const double* uz)
{
    const size_t N = c->N;
    std::vector<R> h(4 * N, R(0));
    const Geo& g = c->g;
    size_t iu = 0;
    for (int z = 0; z < g.uz; ++z)
        for (int y = 0; y < g.uy; ++y)
            for (int x = 0; x < g.ux; ++x, ++iu) {
                const size_t i = (size_t)(x + g.px) + (size_t)g.nx * ((size_t)(y + g.py) + (size_t)g.ny * (z + g.pz));
                h[i] = rho ? (R)rho[iu] : (R)c->d.rho0;
                h[N + i] = ux ? (R)ux[iu] : R(0);
                h[2 * N + i] = uy ? (R)uy[iu] : R(0);
                h[3 * N + i] = uz ? (R)uz[iu] : R(0);
            }
    HIPCHK(hipMemcpyAsync(c->macro, h.data(), 4 * N * sizeof(R), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

// Tile width 2^bxl along x: the widest whose padding waste stays under 15%,
// else the one with the least waste.
static int pick_bx_log2(int nx)
{
    int best = 8;
    double best_util = 0.0;
    for (int b = 8; b >= 5; --b) {
        const int w = 1 << b;
        const double util = (double)nx / (double)(((nx + w - 1) / w) * w);
        if (util >= 0.85) return b;
        if (util > best_util) {
            best_util = util;
            best = b;
        }
    }
    return best;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char* lbm_status_str(lbm_status s)
{
    switch (s) {
    case LBM_OK: return "ok";
    case LBM_E_ARG: return "invalid argument";
    case LBM_E_HIP: return "HIP runtime error";
    case LBM_E_OOM: return "device out of memory";
    case LBM_E_NAN: return "non-finite mass or momentum";
    case LBM_E_STATE: return "call out of order";
    case LBM_E_UNSUPPORTED: return "not supported by this build";
    default: return "unknown status";
    }
}

void lbm_destroy(lbm_ctx* c)
{
    if (!c) return;
    void* bufs[] = {c->pops,     c->cls,      c->node,  c->bcidx, c->inl_state, c->inl,      c->outl,
                    c->ghost,    c->pub_mac,  c->pub_post, c->macro, c->stage, c->partials, c->cl_params, c->cl_norm,
                    c->cl_wb,    c->ccache,   c->avg,   c->probe_buf, c->inl_post, c->inl_mac,
                    c->flag1,    c->uerr,     c->snap_dev, c->pub_post2};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->snap_pending && c->ev_snap_copy) (void)hipEventSynchronize(c->ev_snap_copy);
    if (c->snap_host) (void)hipHostFree(c->snap_host);
    if (c->ev_snap) (void)hipEventDestroy(c->ev_snap);
    if (c->ev_snap_copy) (void)hipEventDestroy(c->ev_snap_copy);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

lbm_status lbm_create(const lbm_desc* d, lbm_ctx** out)
{
    if (!d || !out) return LBM_E_ARG;
    *out = nullptr;
    if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || !d->flags) return LBM_E_ARG;
    if (d->n_bc < 0 || (d->n_bc > 0 && !d->bc)) return LBM_E_ARG;
    if (d->collision < LBM_COLLISION_SRT || d->collision > LBM_COLLISION_COMBINED) return LBM_E_ARG;
    if (d->closure != LBM_CLOSURE_NONE && d->closure != LBM_CLOSURE_SMAGORINSKY && d->closure != LBM_CLOSURE_MLP)
        return LBM_E_ARG;
    // the learned closure runs in the fast FP32 arithmetic
    if (d->closure == LBM_CLOSURE_MLP && d->arith != LBM_ARITH_FAST) return LBM_E_ARG;
    if (d->closure == LBM_CLOSURE_MLP && d->storage == LBM_STORAGE_F64) return LBM_E_UNSUPPORTED;
    if (d->closure == LBM_CLOSURE_MLP && d->collision > LBM_COLLISION_CUM2015) return LBM_E_UNSUPPORTED;
    if (d->closure == LBM_CLOSURE_MLP && d->outlet_sponge) return LBM_E_UNSUPPORTED;
    if (d->storage < LBM_STORAGE_F64 || d->storage > LBM_STORAGE_BF16DEV_SR) return LBM_E_ARG;
    if (d->arith != LBM_ARITH_LITERAL && d->arith != LBM_ARITH_FAST) return LBM_E_ARG;
    if (d->arith == LBM_ARITH_LITERAL && d->storage >= LBM_STORAGE_F16DEV) return LBM_E_ARG;
    // tau > 1/2 keeps omega in (0, 2); NonDim...h:79 asserts this (compiled out by -DNDEBUG)
    if (!(d->tau > 0.5) || !std::isfinite(d->tau)) return LBM_E_ARG;
    if (d->closure != LBM_CLOSURE_NONE && !(d->c_les >= 0.0)) return LBM_E_ARG;
    if (d->periodic[0] && d->nx < 2) return LBM_E_ARG;
    if (d->periodic[1] && d->ny < 2) return LBM_E_ARG;
    if (d->periodic[2] && d->nz < 2) return LBM_E_ARG;

    Layout L;
    L.ux = d->nx;
    L.uy = d->ny;
    L.uz = d->nz;
    L.w[0] = d->periodic[0] != 0;
    L.w[1] = d->periodic[1] != 0;
    L.w[2] = d->periodic[2] != 0;
    const size_t elem = d->storage == LBM_STORAGE_F64 ? 8 : d->storage == LBM_STORAGE_F32 ? 4 : 2;
    const int line = (int)(128 / elem);   // populations per 128-byte cache line
    // A halo along x is a whole line of solid cells on each side, so that the
    // caller's rows start on a line and the step's line-aligned accesses are
    // aligned in fact; only its innermost cells are ever read.
    L.px = L.w[0] ? 0 : line;
    L.py = L.w[1] ? 0 : 1;
    L.pz = L.w[2] ? 0 : 1;
    L.nx = L.w[0] ? L.ux : (L.px + L.ux + 1 + line - 1) / line * line;
    L.ny = L.uy + 2 * L.py;
    L.nz = L.uz + 2 * L.pz;
    const size_t N = (size_t)L.nx * L.ny * L.nz;
    const size_t NU = (size_t)L.ux * L.uy * L.uz;
    if (N * elem > 0xFFFFFFFFull || N >= 0x7FFFFFFFull) return LBM_E_ARG;  // 32-bit plane offsets

    Plan P;
    lbm_status st = classify(d, L, P);
    if (st != LBM_OK) return st;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return LBM_E_HIP;
    }

    lbm_ctx* c = new (std::nothrow) lbm_ctx();
    if (!c) return LBM_E_OOM;
    c->d = *d;
    c->d.flags = nullptr;
    c->d.bc = nullptr;
    c->N = N;
    c->NU = NU;
    c->elem = elem;
    c->rbytes = d->storage == LBM_STORAGE_F64 ? 8 : 4;
    c->flags_needed = P.flags_needed;
    c->n_inlet = (int)P.inl.size();
    c->n_eq = P.n_eq;
    for (const NodeDev& e : P.inl)
        if (e.kind == LBM_BC_INLET_MOVING_WALL || e.kind == LBM_BC_INLET_ZOU_HE) c->n_upd++;
    c->n_outlet = (int)P.outl.size();
    if (d->outlet_sponge) {
        // OutletNeighbourEquilibriumOmega::outletOmega (:18-79): the first outlet
        // entry's normal picks the axis; the last ceil(0.1 n) cells of it
        const lbm_bc_node* b0 = nullptr;
        for (int n = 0; n < d->n_bc && !b0; ++n)
            if (d->bc[n].kind == LBM_BC_OUTLET_INTERPOLATED || d->bc[n].kind == LBM_BC_OUTLET_EQUILIBRIUM ||
                d->bc[n].kind == LBM_BC_OUTLET_NEIGHBOUR_EQ || d->bc[n].kind == LBM_BC_OUTLET_NEIGHBOUR_EQ_RF)
                b0 = &d->bc[n];
        if (!b0) {
            lbm_destroy(c);
            return LBM_E_ARG;
        }
        const float an[3] = {std::fabs(b0->nx), std::fabs(b0->ny), std::fabs(b0->nz)};
        int axis = -1;
        for (int k = 0; k < 3; ++k)
            if (an[k] == 1.0f && an[(k + 1) % 3] == 0.0f && an[(k + 2) % 3] == 0.0f) axis = k;
        if (axis < 0) {
            lbm_destroy(c);
            return LBM_E_UNSUPPORTED;
        }
        const int dim = axis == 0 ? d->nx : axis == 1 ? d->ny : d->nz;
        c->sp_interval = (int)std::ceil(dim * 0.1f);
        c->sp_begin = dim - c->sp_interval;
        c->sp_axis = axis;
    }
    c->h_flag1.resize(NU);
    for (size_t i = 0; i < NU; ++i) c->h_flag1[i] = d->flags[i] == 1 ? 1 : 0;
    c->n_ghost = (int)P.ghost.size();
    c->n_pub = P.n_pub;
    c->outl_deferred = P.defer_outlet;
    c->inl_deferred = P.defer_inlet;
    c->ghost_off = P.ghost_off;
    (void)hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        lbm_destroy(c);
        return LBM_E_HIP;
    }
    // tile geometry: 256 threads = (1 << bx_log2) x by cells
    const int bxl = pick_bx_log2(L.ux);
    c->g.nx = L.nx;
    c->g.ny = L.ny;
    c->g.nz = L.nz;
    c->g.px = L.px;
    c->g.py = L.py;
    c->g.pz = L.pz;
    c->g.ux = L.ux;
    c->g.uy = L.uy;
    c->g.uz = L.uz;
    c->g.bx_log2 = bxl;
    c->g.by = TPB >> bxl;
    c->g.ntx = (L.ux + (1 << bxl) - 1) >> bxl;
    c->g.nty = (L.uy + c->g.by - 1) / c->g.by;
    c->g.ntiles = c->g.ntx * c->g.nty * L.uz;
    c->g.wx = L.w[0];
    c->g.wy = L.w[1];
    c->g.wz = L.w[2];

    const size_t rb = c->rbytes;
    st = dev_alloc(c, &c->pops, N * elem * Q);
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->cls, N);
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->node, N * sizeof(uint64_t));
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->bcidx, N * sizeof(int32_t));
    if (st == LBM_OK) st = dev_alloc(c, &c->inl_state, (size_t)c->n_inlet * Q * rb);
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->inl, P.inl.size() * sizeof(NodeDev));
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->outl, P.outl.size() * sizeof(NodeDev));
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->ghost, P.ghost.size() * sizeof(NodeDev));
    if (st == LBM_OK) st = dev_alloc(c, &c->pub_mac, (size_t)c->n_pub * PUB_MAC * rb);
    if (st == LBM_OK) st = dev_alloc(c, &c->pub_post, (size_t)c->n_pub * Q * rb);
    if (st == LBM_OK && c->outl_deferred) st = dev_alloc(c, &c->pub_post2, (size_t)c->n_pub * Q * rb);
    if (st == LBM_OK) st = dev_alloc(c, &c->inl_post, c->n_upd ? (size_t)c->n_inlet * Q * rb : 0);
    if (st == LBM_OK) st = dev_alloc(c, &c->inl_mac, c->n_upd ? (size_t)c->n_inlet * PUB_MAC * rb : 0);
    if (st == LBM_OK) st = dev_alloc(c, &c->macro, 4 * N * rb);
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->stage, 5 * NU * sizeof(double));
    const size_t nparts = std::max((size_t)c->g.ntiles, (size_t)blocks_for(NU));
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->partials, nparts * 8 * sizeof(double));
    if (st != LBM_OK) {
        lbm_destroy(c);
        return st;
    }
    c->h_partials.resize(nparts * 8);
    auto up = [&](void* dst, const void* src, size_t bytes) {
        return bytes == 0 || hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream) == hipSuccess;
    };
    bool ok = up(c->cls, P.cls.data(), N) && up(c->node, P.node.data(), N * sizeof(uint64_t)) &&
              up(c->bcidx, P.bcidx.data(), N * sizeof(int32_t)) &&
              up(c->inl, P.inl.data(), P.inl.size() * sizeof(NodeDev)) &&
              up(c->outl, P.outl.data(), P.outl.size() * sizeof(NodeDev)) &&
              up(c->ghost, P.ghost.data(), P.ghost.size() * sizeof(NodeDev));
    ok = ok && hipMemsetAsync(c->pops, 0, N * elem * Q, c->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(c->inl_state, 0, std::max<size_t>(16, (size_t)c->n_inlet * Q * rb), c->stream) ==
                   hipSuccess;
    ok = ok && hipMemsetAsync(c->pub_mac, 0, std::max<size_t>(16, (size_t)c->n_pub * PUB_MAC * rb), c->stream) ==
                   hipSuccess;
    ok = ok && hipMemsetAsync(c->pub_post, 0, std::max<size_t>(16, (size_t)c->n_pub * Q * rb), c->stream) ==
                   hipSuccess;
    if (c->pub_post2)
        ok = ok && hipMemsetAsync(c->pub_post2, 0, std::max<size_t>(16, (size_t)c->n_pub * Q * rb), c->stream) ==
                       hipSuccess;
    ok = ok && hipStreamSynchronize(c->stream) == hipSuccess;
    if (!ok) {
        lbm_destroy(c);
        return LBM_E_HIP;
    }
    *out = c;
    return LBM_OK;
}

// Add a pending time-average state before the state changes or the sums are read.
static lbm_status avg_flush(lbm_ctx* c)
{
    if (!c->avg_pending) return LBM_OK;
    return with_storage(c, [&](auto stv) -> lbm_status { return avg_flush_st<decltype(stv)>(c); });
}

lbm_status lbm_init_equilibrium(lbm_ctx* c, const double* rho, const double* ux, const double* uy, const double* uz)
{
    if (!c) return LBM_E_ARG;
    lbm_status st = avg_flush(c);
    if (st != LBM_OK) return st;
    st = c->rbytes == 8 ? upload_macro<double>(c, rho, ux, uy, uz)
                                   : upload_macro<float>(c, rho, ux, uy, uz);
    if (st != LBM_OK) return st;
    // solid cells carry rho = u = 0 in the reference (InitializationEquilibriumConstVector.h:84-89)
    c->steps = 0;
    st = with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        AuxArgs<ST> a = aux_args<ST>(c);
        a.odd = 0;
        k_init<ST><<<blocks_for(c->N), TPB, 0, c->stream>>>(a);
        HIPCHK(hipGetLastError());
        return LBM_OK;
    });
    if (st != LBM_OK) return st;
    // the convergence norm's previous velocity starts again from zero
    if (c->uerr) HIPCHK(hipMemsetAsync(c->uerr, 0, 3 * c->NU * c->rbytes, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->initialized = true;
    c->first = true;
    c->outl_pending = false;
    c->inl_pending = false;
    c->nan_seen = false;
    c->cl_count = 0;   // the next step evaluates the closure network
    return LBM_OK;
}

lbm_status lbm_set_populations(lbm_ctx* c, const double* f)
{
    if (!c || !f) return LBM_E_ARG;
    {
        lbm_status st = avg_flush(c);
        if (st != LBM_OK) return st;
    }
    for (int q = 0; q < Q; ++q) {
        HIPCHK(hipMemcpyAsync(c->stage, f + (size_t)q * c->NU, c->NU * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        lbm_status st = with_storage(c, [&](auto stv) -> lbm_status {
            using ST = decltype(stv);
            AuxArgs<ST> a = aux_args<ST>(c);
            a.q = q;
            k_scatter<ST><<<blocks_for(c->NU), TPB, 0, c->stream>>>(a);
            HIPCHK(hipGetLastError());
            return LBM_OK;
        });
        if (st != LBM_OK) return st;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->initialized = true;
    c->first = false;
    c->outl_pending = false;   // the given populations are the state, outlet rule included
    c->inl_pending = false;    // and inlet rule
    c->cl_count = 0;   // the next step evaluates the closure network
    return LBM_OK;
}

lbm_status lbm_step(lbm_ctx* c, int32_t nsteps)
{
    if (!c || nsteps < 0) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    if (nsteps == 0) return LBM_OK;
    lbm_status st;
    switch (c->d.storage) {
    case LBM_STORAGE_F64: st = launch_by_coll<StoreF64>(c, nsteps); break;
    case LBM_STORAGE_F32: st = launch_by_coll<StoreF32>(c, nsteps); break;
    case LBM_STORAGE_F16DEV: st = launch_by_coll_fast<StoreF16Dev>(c, nsteps); break;
    case LBM_STORAGE_BF16DEV: st = launch_by_coll_fast<StoreBF16Dev>(c, nsteps); break;
    case LBM_STORAGE_F16DEV_SR: st = launch_by_coll_fast<StoreF16DevSR>(c, nsteps); break;
    case LBM_STORAGE_BF16DEV_SR: st = launch_by_coll_fast<StoreBF16DevSR>(c, nsteps); break;
    default: return LBM_E_ARG;
    }
    if (st != LBM_OK) return st;
    c->last_ms = -(double)nsteps;  // resolved at the next sync
    return LBM_OK;
}

static lbm_status avg_alloc(lbm_ctx* c)
{
    if (c->avg) return LBM_OK;
    lbm_status st = dev_alloc(c, (void**)&c->avg, 4 * c->NU * sizeof(double));
    if (st != LBM_OK) return st;
    HIPCHK(hipMemsetAsync(c->avg, 0, 4 * c->NU * sizeof(double), c->stream));
    c->avg_count = 0;
    c->avg_pending = false;
    return LBM_OK;
}

lbm_status lbm_moment_add(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = avg_alloc(c);
    if (st == LBM_OK) st = avg_flush(c);   // the same state counted twice
    if (st != LBM_OK) return st;
    c->avg_pending = true;
    c->avg_count++;
    return LBM_OK;
}

lbm_status lbm_step_moment_add(lbm_ctx* c, int32_t nsteps)
{
    if (!c || nsteps < 0) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = avg_alloc(c);
    if (st != LBM_OK) return st;
    c->avg_each = true;
    st = lbm_step(c, nsteps);
    c->avg_each = false;
    return st;
}

lbm_status lbm_moment_avg(lbm_ctx* c, double* rho, double* ux, double* uy, double* uz, int64_t* count)
{
    if (!c) return LBM_E_ARG;
    if (count) *count = c->avg ? c->avg_count : 0;
    if (!c->avg || c->avg_count == 0) return LBM_E_STATE;
    lbm_status st = avg_flush(c);
    if (st != LBM_OK) return st;
    double* outs[4] = {rho, ux, uy, uz};
    for (int k = 0; k < 4; ++k)
        if (outs[k])
            HIPCHK(hipMemcpyAsync(outs[k], c->avg + k * c->NU, c->NU * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // momentAvg (MomentTimeAvg.h:47-77): the sums divided by the counter
    const double n = (double)c->avg_count;
    for (int k = 0; k < 4; ++k)
        if (outs[k])
            for (size_t i = 0; i < c->NU; ++i) outs[k][i] /= n;
    return LBM_OK;
}

lbm_status lbm_moment_avg_reset(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (!c->avg) return LBM_OK;
    HIPCHK(hipMemsetAsync(c->avg, 0, 4 * c->NU * sizeof(double), c->stream));
    c->avg_count = 0;
    c->avg_pending = false;
    return LBM_OK;
}

lbm_status lbm_sync(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->last_ms < 0) {
        const double n = -c->last_ms;
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_ms = (double)ms / n;
    }
    return c->nan_seen ? LBM_E_NAN : LBM_OK;
}

// rho, u, omega of the current state into 5 FP64 planes (the staging
// planes unless dst is given).
static lbm_status run_macro(lbm_ctx* c, double* dst = nullptr)
{
    return with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        AuxArgs<ST> a = aux_args<ST>(c);
        if (dst) a.io = dst;
        k_macro<ST><<<blocks_for(c->NU), TPB, 0, c->stream>>>(a);
        HIPCHK(hipGetLastError());
        return LBM_OK;
    });
}

// Closure read-back into device buffers (any may be null).
static lbm_status run_closure_eval(lbm_ctx* c, float* feats, int32_t* env, double* coef, double* omega)
{
    return with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        if constexpr (sizeof(typename ST::R) != 4) {
            return LBM_E_UNSUPPORTED;
        } else {
            ClosureEvalArgs<ST> e;
            std::memset(&e, 0, sizeof(e));
            e.a = aux_args<ST>(c);
            e.mlp = closure_weights(c);
            e.feats = feats;
            e.env = env;
            e.coef = coef;
            e.omega = omega;
            e.ccache = (c->cl_count % c->cl_every == 0) ? nullptr : c->ccache;
            k_closure_eval<ST><<<blocks_for(c->NU), TPB, 0, c->stream>>>(e);
            HIPCHK(hipGetLastError());
            return LBM_OK;
        }
    });
}

lbm_status lbm_get_macroscopic(lbm_ctx* c, double* rho, double* ux, double* uy, double* uz)
{
    if (!c) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    st = run_macro(c);
    if (st != LBM_OK) return st;
    double* outs[4] = {rho, ux, uy, uz};
    for (int k = 0; k < 4; ++k)
        if (outs[k])
            HIPCHK(hipMemcpyAsync(outs[k], c->stage + k * c->NU, c->NU * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_probe(lbm_ctx* c, int32_t x, int32_t y, int32_t z, double out[4])
{
    if (!c || !out) return LBM_E_ARG;
    if (x < 0 || y < 0 || z < 0 || x >= c->g.ux || y >= c->g.uy || z >= c->g.uz) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    st = with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        AuxArgs<ST> a = aux_args<ST>(c);
        k_probe<ST><<<1, 64, 0, c->stream>>>(a, x, y, z);
        HIPCHK(hipGetLastError());
        return LBM_OK;
    });
    if (st != LBM_OK) return st;
    HIPCHK(hipMemcpyAsync(out, c->stage, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

// Move the device probe records to the host list (synchronises).
static lbm_status probe_drain(lbm_ctx* c)
{
    if (c->probe_n == 0) return LBM_OK;
    const size_t old = c->probe_host.size();
    c->probe_host.resize(old + 4 * (size_t)c->probe_n);
    HIPCHK(hipMemcpyAsync(c->probe_host.data() + old, c->probe_buf, 4 * (size_t)c->probe_n * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->probe_n = 0;
    return LBM_OK;
}

lbm_status lbm_probe_record(lbm_ctx* c, int32_t x, int32_t y, int32_t z)
{
    if (!c) return LBM_E_ARG;
    if (x < 0 || y < 0 || z < 0 || x >= c->g.ux || y >= c->g.uy || z >= c->g.uz) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    if (!c->probe_buf) {
        lbm_status st = dev_alloc(c, (void**)&c->probe_buf, (size_t)PROBE_CAP * 4 * sizeof(double));
        if (st != LBM_OK) return st;
    }
    if (c->probe_n == PROBE_CAP) {
        lbm_status st = probe_drain(c);
        if (st != LBM_OK) return st;
    }
    double* slot = c->probe_buf + 4 * (size_t)c->probe_n;
    lbm_status st = with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        AuxArgs<ST> a = aux_args<ST>(c);
        a.io = slot;
        k_probe<ST><<<1, 64, 0, c->stream>>>(a, x, y, z);
        HIPCHK(hipGetLastError());
        return LBM_OK;
    });
    if (st != LBM_OK) return st;
    c->probe_n++;
    return LBM_OK;
}

lbm_status lbm_probe_fetch(lbm_ctx* c, double* out, int64_t max_records, int64_t* n)
{
    if (!c || !n) return LBM_E_ARG;
    lbm_status st = probe_drain(c);
    if (st != LBM_OK) return st;
    const int64_t have = (int64_t)(c->probe_host.size() / 4);
    *n = have;
    if (!out) return LBM_OK;
    if (max_records < have) return LBM_E_ARG;
    std::copy(c->probe_host.begin(), c->probe_host.end(), out);
    c->probe_host.clear();
    return LBM_OK;
}

lbm_status lbm_get_omega(lbm_ctx* c, double* omega)
{
    if (!c || !omega) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    if (c->d.closure == LBM_CLOSURE_MLP) {
        if (!c->mlp_ready) return LBM_E_STATE;
        st = run_closure_eval(c, nullptr, nullptr, nullptr, c->stage + 4 * c->NU);
    } else {
        st = run_macro(c);
    }
    if (st != LBM_OK) return st;
    HIPCHK(hipMemcpyAsync(omega, c->stage + 4 * c->NU, c->NU * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_snapshot_begin(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    if (c->snap_pending) return LBM_E_STATE;
    if (c->d.closure == LBM_CLOSURE_MLP && !c->mlp_ready) return LBM_E_STATE;
    if (!c->snap_dev) {
        lbm_status st = dev_alloc(c, (void**)&c->snap_dev, 5 * c->NU * sizeof(double));
        if (st != LBM_OK) return st;
        if (hipHostMalloc((void**)&c->snap_host, 5 * c->NU * sizeof(double), hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            c->snap_host = nullptr;
            return LBM_E_OOM;
        }
        if (hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_snap, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_snap_copy, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            return LBM_E_HIP;
        }
    }
    // the previous snapshot's copy has finished (it was collected), so
    // snap_dev is free; the planes are formed behind the queued steps
    lbm_status st = run_macro(c, c->snap_dev);
    if (st != LBM_OK) return st;
    if (c->d.closure == LBM_CLOSURE_MLP) {
        st = run_closure_eval(c, nullptr, nullptr, nullptr, c->snap_dev + 4 * c->NU);
        if (st != LBM_OK) return st;
    }
    HIPCHK(hipEventRecord(c->ev_snap, c->stream));
    HIPCHK(hipStreamWaitEvent(c->copy_stream, c->ev_snap, 0));
    HIPCHK(hipMemcpyAsync(c->snap_host, c->snap_dev, 5 * c->NU * sizeof(double), hipMemcpyDeviceToHost,
                          c->copy_stream));
    HIPCHK(hipEventRecord(c->ev_snap_copy, c->copy_stream));
    c->snap_pending = true;
    c->snap_step = c->steps;
    return LBM_OK;
}

lbm_status lbm_snapshot_wait(lbm_ctx* c, double* rho, double* ux, double* uy, double* uz, double* omega,
                             int64_t* step)
{
    if (!c) return LBM_E_ARG;
    if (!c->snap_pending) return LBM_E_STATE;
    HIPCHK(hipEventSynchronize(c->ev_snap_copy));
    c->snap_pending = false;
    double* outs[5] = {rho, ux, uy, uz, omega};
    for (int k = 0; k < 5; ++k)
        if (outs[k]) std::memcpy(outs[k], c->snap_host + (size_t)k * c->NU, c->NU * sizeof(double));
    if (step) *step = c->snap_step;
    return LBM_OK;
}

lbm_status lbm_get_population(lbm_ctx* c, int32_t q, double* f)
{
    if (!c || !f || q < 0 || q >= Q) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    st = with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        AuxArgs<ST> a = aux_args<ST>(c);
        a.q = q;
        k_gather<ST><<<blocks_for(c->NU), TPB, 0, c->stream>>>(a);
        HIPCHK(hipGetLastError());
        return LBM_OK;
    });
    if (st != LBM_OK) return st;
    HIPCHK(hipMemcpyAsync(f, c->stage, c->NU * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_get_populations(lbm_ctx* c, double* f)
{
    if (!c || !f) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    for (int q = 0; q < Q; ++q) {
        st = with_storage(c, [&](auto stv) -> lbm_status {
            using ST = decltype(stv);
            AuxArgs<ST> a = aux_args<ST>(c);
            a.q = q;
            k_gather<ST><<<blocks_for(c->NU), TPB, 0, c->stream>>>(a);
            HIPCHK(hipGetLastError());
            return LBM_OK;
        });
        if (st != LBM_OK) return st;
        HIPCHK(hipMemcpyAsync(f + (size_t)q * c->NU, c->stage, c->NU * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_conservation(lbm_ctx* c, double out[4])
{
    if (!c || !out) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    const unsigned nb = blocks_for(c->NU);
    st = with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        AuxArgs<ST> a = aux_args<ST>(c);
        a.io = c->partials;
        k_sums<ST><<<nb, TPB, 0, c->stream>>>(a);
        HIPCHK(hipGetLastError());
        return LBM_OK;
    });
    if (st != LBM_OK) return st;
    HIPCHK(hipMemcpyAsync(c->h_partials.data(), c->partials, (size_t)nb * 4 * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s[4] = {0, 0, 0, 0};
    for (unsigned b = 0; b < nb; ++b)
        for (int k = 0; k < 4; ++k) s[k] += c->h_partials[(size_t)b * 4 + k];
    bool finite = true;
    for (int k = 0; k < 4; ++k) {
        out[k] = s[k];
        finite = finite && std::isfinite(s[k]);
    }
    if (!finite) {
        c->nan_seen = true;
        return LBM_E_NAN;
    }
    return LBM_OK;
}

lbm_status lbm_step_count(lbm_ctx* c, int64_t* steps)
{
    if (!c || !steps) return LBM_E_ARG;
    *steps = c->steps;
    return LBM_OK;
}

lbm_status lbm_last_kernel_ms(lbm_ctx* c, double* ms)
{
    if (!c || !ms) return LBM_E_ARG;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    *ms = c->last_ms;
    return LBM_OK;
}

lbm_status lbm_device_info(int32_t* count, char* name, size_t name_len)
{
    if (!count) return LBM_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        *count = 0;
        return LBM_E_HIP;
    }
    *count = n;
    if (name && name_len > 0) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return LBM_E_HIP;
        std::snprintf(name, name_len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    }
    return LBM_OK;
}

lbm_status lbm_bandwidth_probe(size_t bytes, int32_t iters, double* gbps, double* ms_per_iter)
{
    if (bytes < 16 || iters <= 0) return LBM_E_ARG;
    const size_t n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) { (void)hipGetLastError(); return LBM_E_OOM; }
    if (hipMalloc(&b, n * 16) != hipSuccess) { (void)hipGetLastError(); (void)hipFree(a); return LBM_E_OOM; }
    hipStream_t s;
    hipEvent_t e0, e1;
    lbm_status st = LBM_OK;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) { (void)hipFree(a); (void)hipFree(b); return LBM_E_HIP; }
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const unsigned nb = (unsigned)((n + TPB - 1) / TPB);
    (void)hipMemsetAsync(a, 0, n * 16, s);
    k_copy<<<nb, TPB, 0, s>>>(a, b, n);  // warm-up
    (void)hipEventRecord(e0, s);
    for (int k = 0; k < iters; ++k) k_copy<<<nb, TPB, 0, s>>>(a, b, n);
    (void)hipEventRecord(e1, s);
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) st = LBM_E_HIP;
    float ms = 0.f;
    if (st == LBM_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = LBM_E_HIP;
    if (st == LBM_OK) {
        const double per = (double)ms / iters;
        if (ms_per_iter) *ms_per_iter = per;
        if (gbps) *gbps = 2.0 * (double)(n * 16) / (per * 1e-3) / 1e9;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return st;
}

lbm_status lbm_device_bytes(lbm_ctx* c, size_t* bytes)
{
    if (!c || !bytes) return LBM_E_ARG;
    *bytes = c->dev_bytes;
    return LBM_OK;
}

lbm_status lbm_step_diagnostics(lbm_ctx* c, double out[7])
{
    if (!c || !out) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    if (c->d.closure == LBM_CLOSURE_MLP) return LBM_E_UNSUPPORTED;
    c->diag_next = true;
    lbm_status st = lbm_step(c, 1);
    c->diag_next = false;
    if (st != LBM_OK) return st;
    st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    const size_t nb = (size_t)c->g.ntiles;
    HIPCHK(hipMemcpy(c->h_partials.data(), c->partials, nb * 8 * sizeof(double), hipMemcpyDeviceToHost));
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (size_t b = 0; b < nb; ++b)
        for (int k = 0; k < 7; ++k) s[k] += c->h_partials[b * 8 + k];
    bool finite = true;
    for (int k = 0; k < 7; ++k) {
        out[k] = s[k];
        finite = finite && std::isfinite(s[k]);
    }
    if (!finite) {
        c->nan_seen = true;
        return LBM_E_NAN;
    }
    return LBM_OK;
}

}  // extern "C"

// Internal hooks of noalloc.hip (not part of the C ABI).
//
// Geometry of the diagnostics step: tiles (one partial record each), the
// device and stream of the context.
lbm_status lbm_internal_diag_info(lbm_ctx* c, int* ntiles, int* device, hipStream_t* stream)
{
    if (!c) return LBM_E_ARG;
    *ntiles = c->g.ntiles;
    *device = c->device;
    *stream = c->stream;
    return LBM_OK;
}

// lbm_step_diagnostics' step with its per-tile partial records (8 doubles
// each) written to dst instead of the context's buffer; asynchronous.
lbm_status lbm_internal_diag_step(lbm_ctx* c, double* dst)
{
    if (!c || !dst) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    if (c->d.closure == LBM_CLOSURE_MLP) return LBM_E_UNSUPPORTED;
    c->diag_dst = dst;
    c->diag_next = true;
    lbm_status st = lbm_step(c, 1);
    c->diag_next = false;
    c->diag_dst = nullptr;
    return st;
}

void lbm_internal_mark_nan(lbm_ctx* c)
{
    if (c) c->nan_seen = true;
}

// out[0 .. ncomp) = sum over nb records of parts (stride doubles apart), on
// the device in a fixed order; asynchronous on `stream`.
lbm_status lbm_internal_reduce(const double* parts, int nb, int stride, int ncomp, double* out, hipStream_t stream)
{
    if (nb < 1 || ncomp < 1 || ncomp > 8 || stride < ncomp) return LBM_E_ARG;
    k_reduce_partials<<<1, TPB, 0, stream>>>(parts, nb, stride, ncomp, out);
    HIPCHK(hipGetLastError());
    return LBM_OK;
}

extern "C" {

lbm_status lbm_error_quadratic(lbm_ctx* c, double* err)
{
    if (!c || !err) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    if (!c->uerr) {
        // first evaluation: the previous velocity starts at zero, as the
        // reference's u_error field and the host loop's copy of it do
        st = dev_alloc(c, &c->uerr, 3 * c->NU * c->rbytes);
        if (st == LBM_OK) st = dev_alloc(c, (void**)&c->flag1, c->NU);
        if (st != LBM_OK) return st;
        HIPCHK(hipMemsetAsync(c->uerr, 0, 3 * c->NU * c->rbytes, c->stream));
        HIPCHK(hipMemcpyAsync(c->flag1, c->h_flag1.data(), c->NU, hipMemcpyHostToDevice, c->stream));
    }
    const unsigned nb = blocks_for(c->NU);
    double* res = c->partials + (size_t)nb * 2;   // the partials buffer holds >= 8 doubles per block
    st = with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        AuxArgs<ST> a = aux_args<ST>(c);
        a.io = c->partials;
        k_error<ST><<<nb, TPB, 0, c->stream>>>(a, (typename ST::R*)c->uerr, c->flag1);
        HIPCHK(hipGetLastError());
        return LBM_OK;
    });
    if (st != LBM_OK) return st;
    k_reduce_partials<<<1, TPB, 0, c->stream>>>(c->partials, (int)nb, 2, 2, res);
    HIPCHK(hipGetLastError());
    double sums[2];
    HIPCHK(hipMemcpyAsync(sums, res, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *err = std::sqrt(sums[0]) / std::sqrt(sums[1]);
    if (std::isnan(sums[0]) || std::isnan(sums[1])) {
        c->nan_seen = true;
        return LBM_E_NAN;
    }
    return LBM_OK;
}

lbm_status lbm_set_closure(lbm_ctx* c, const float* params, const float* norm)
{
    if (!c || !params || !norm) return LBM_E_ARG;
    if (c->d.closure != LBM_CLOSURE_MLP) return LBM_E_STATE;
    float N[4 * NFEAT];   // mean, 1/sd, envelope low, high
    for (int k = 0; k < NFEAT; ++k) {
        const float sd = norm[NFEAT + k];
        N[k] = norm[k];
        N[NFEAT + k] = 1.0f / sd;
        N[2 * NFEAT + k] = norm[2 * NFEAT + k];
        N[3 * NFEAT + k] = norm[3 * NFEAT + k];
        if (!(sd > 0.0f) || !std::isfinite(sd) || !std::isfinite(N[k]) || !std::isfinite(N[NFEAT + k]))
            return LBM_E_ARG;
    }
    for (int i = 0; i < NPARAM; ++i)
        if (!std::isfinite(params[i])) return LBM_E_ARG;
    if (!c->cl_params) {
        lbm_status st = dev_alloc(c, (void**)&c->cl_params, sizeof(float) * NPARAM);
        if (st != LBM_OK) return st;
    }
    if (!c->cl_wb) {
        lbm_status st = dev_alloc(c, (void**)&c->cl_wb, sizeof(uint16_t) * (NHID * 16 + NHID * NHID));
        if (st != LBM_OK) return st;
    }
    if (!c->cl_norm) {
        lbm_status st = dev_alloc(c, (void**)&c->cl_norm, sizeof(N));
        if (st != LBM_OK) return st;
    }
    std::vector<uint16_t> wb(NHID * 16 + NHID * NHID);
    lbm_mlp_pack_bf16(params, wb.data(), wb.data() + NHID * 16);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->cl_params, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->cl_wb, wb.data(), sizeof(uint16_t) * wb.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->cl_norm, N, sizeof(N), hipMemcpyHostToDevice));
    if (!c->ccache) {
        lbm_status st = dev_alloc(c, (void**)&c->ccache, sizeof(float) * c->N);
        if (st != LBM_OK) return st;
    }
    c->cl_b3 = params[OFF_B3];
    c->cl_count = 0;
    c->mlp_ready = true;
    return LBM_OK;
}

lbm_status lbm_set_closure_refresh(lbm_ctx* c, int32_t every)
{
    if (!c || every < 1) return LBM_E_ARG;
    if (c->d.closure != LBM_CLOSURE_MLP) return LBM_E_STATE;
    c->cl_every = every;
    c->cl_count = 0;
    return LBM_OK;
}

lbm_status lbm_get_closure_features(lbm_ctx* c, float* features, int32_t* in_envelope, double* coefficient)
{
    if (!c) return LBM_E_ARG;
    if (c->d.closure != LBM_CLOSURE_MLP || !c->mlp_ready || !c->initialized) return LBM_E_STATE;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    const size_t NU = c->NU;
    float* df = nullptr;
    int32_t* de = nullptr;
    double* dc = nullptr;
    st = LBM_OK;
    if (features && hipMalloc(&df, sizeof(float) * NFEAT * NU) != hipSuccess) st = LBM_E_OOM;
    if (st == LBM_OK && in_envelope && hipMalloc(&de, sizeof(int32_t) * NU) != hipSuccess) st = LBM_E_OOM;
    if (st == LBM_OK && coefficient) dc = c->stage;
    if (st == LBM_OK) st = run_closure_eval(c, df, de, dc, nullptr);
    if (st == LBM_OK && df &&
        hipMemcpyAsync(features, df, sizeof(float) * NFEAT * NU, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        st = LBM_E_HIP;
    if (st == LBM_OK && de &&
        hipMemcpyAsync(in_envelope, de, sizeof(int32_t) * NU, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        st = LBM_E_HIP;
    if (st == LBM_OK && dc &&
        hipMemcpyAsync(coefficient, dc, sizeof(double) * NU, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        st = LBM_E_HIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess && st == LBM_OK) st = LBM_E_HIP;
    (void)hipGetLastError();
    if (df) (void)hipFree(df);
    if (de) (void)hipFree(de);
    return st;
}

lbm_status lbm_filter(lbm_ctx* c, int32_t factor, double* out)
{
    if (!c || !out || factor < 2) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    const int fx = c->g.ux, fy = c->g.uy, fz = c->g.uz;
    if (fx % factor || fy % factor || fz % factor) return LBM_E_ARG;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    const int cx = fx / factor, cy = fy / factor, cz = fz / factor;
    const size_t NC = (size_t)cx * cy * cz;
    st = with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        AuxArgs<ST> a = aux_args<ST>(c);
        k_filter<ST><<<blocks_for(NC), TPB, 0, c->stream>>>(a, factor, cx, cy, cz);
        HIPCHK(hipGetLastError());
        return LBM_OK;
    });
    if (st != LBM_OK) return st;
    HIPCHK(hipMemcpyAsync(out, c->stage, 10 * NC * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

}  // extern "C"
