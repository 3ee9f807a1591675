// lbm_core.hip -- MI355X (gfx950) lattice-Boltzmann collide-stream core.
//
// One fused kernel per time step: load the 27 populations of a cell, derive
// rho/u, the relaxation rate (constant or Smagorinsky), collide (cumulant
// 2015 or BGK) and stream in place with the A-A ("esoteric twist") pattern,
// applying bounce-back and moving-wall boundaries from the flag grid in the
// same pass.  It replaces the reference's per-step operator chain
// OmegaLES -> Collision -> StreamingAB -> BounceBackWallHalf -> Inlet ->
// MomentDensityVelocityN27 (SolverTurbulentLES.h:187-214), which reads and
// writes two full lattices plus rho/u/omega every step.
//
// A-A layout: a single population array P[q][cell] (structure of arrays,
// x fastest).  With t the step counter:
//   even t: cell x reads f_q from P[q][x] and writes post-collision f*_q to
//           P[opp q][x];
//   odd t:  cell x reads f_q from P[opp q][x - c_q] and writes f*_q to
//           P[q][x + c_q].
// Every slot is read and written by the same thread within a step, so the
// scheme is race-free without a second lattice.
//
// Boundaries (reference semantics, see the oracle for the AB restatement):
//  * source x - c_q solid (flag 0): the cell itself writes the slot its
//    neighbour would have written -- f*_{crev q} for wall-list nodes
//    (half-way bounce-back, BounceBackWallHalf.h:46-47) or its own
//    pre-collision f_q otherwise (no bounce-back: StreamingAB.h:58-65 leaves
//    the slot untouched);
//  * destination x + c_q solid or inlet: the streaming write is skipped;
//  * inlet cells (flag -1) keep their own state in a side buffer, stream
//    their pre-collision populations (StreamingAB.h:62-65) and update it with
//    the moving-wall rule (InletVelocityMovingWall.h:43-49).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/lbm_core.h"
#include "collide.hpp"
#include "storage.hpp"

namespace lbm {

enum : uint8_t { CL_SOLID = 0, CL_FLUID = 1, CL_BND = 2, CL_INLET_MW = 3 };

constexpr int TPB = 256;
constexpr uint64_t MASK27 = (1ull << 27) - 1;

struct Geo {
    int nx, ny, nz;
    int bx_log2, by;   // tile: (1 << bx_log2) x by cells, 256 threads
    int ntx, nty;      // tiles along x and y
    int ntiles;
    int wx, wy, wz;    // index wrap per axis
};

struct InletDev {
    float nx, ny, nz;
    float ux, uy, uz;
    int32_t kind;
    int32_t pad;
};

template <class ST>
struct StepArgs {
    typename ST::T* p[Q];
    const uint8_t* cls;
    const uint64_t* node;
    const int32_t* bcidx;
    typename ST::R* inl_state;      // [n_inlet][27]
    const InletDev* inl;
    const typename ST::R* macro;    // rho, ux, uy, uz planes for the first step
    double* partials;               // [ntiles][4] when diag != 0
    Geo g;
    Consts<typename ST::R> K;
    typename ST::R rest[4];         // w*rho0 per weight class (deviation storage)
    typename ST::R omega0, tau, cles;
    int crev20;
    int first;
    int diag;
    int closure;
};

// Workgroups are dispatched round-robin over the 8 XCDs; renumber so that
// consecutive tiles (neighbouring x segments, which share partial cache
// lines on odd steps) land on the same XCD and its L2.
__device__ __forceinline__ int xcd_tile(int bid, int G)
{
    const int k = G >> 3, r = G & 7;
    const int x = bid & 7, y = bid >> 3;
    return (x < r) ? x * (k + 1) + y : r * (k + 1) + (x - r) * k + y;
}

struct Cell {
    int x, y, z;
    uint32_t i;
    int dxm, dxp, dym, dyp, dzm, dzp;
    bool inside;
};

__device__ __forceinline__ Cell cell_of(const Geo& g)
{
    Cell c;
    const int t = xcd_tile((int)blockIdx.x, g.ntiles);
    const int tx = t % g.ntx;
    const int rem = t / g.ntx;
    const int ty = rem % g.nty;
    c.z = rem / g.nty;
    c.x = (tx << g.bx_log2) + ((int)threadIdx.x & ((1 << g.bx_log2) - 1));
    c.y = ty * g.by + ((int)threadIdx.x >> g.bx_log2);
    c.inside = c.x < g.nx && c.y < g.ny;
    const int sx = 1, sy = g.nx, sz = g.nx * g.ny;
    c.i = (uint32_t)(c.x + g.nx * (c.y + g.ny * c.z));
    c.dxm = (c.x > 0 || !g.wx) ? -sx : (g.nx - 1) * sx;
    c.dxp = (c.x < g.nx - 1 || !g.wx) ? sx : -(g.nx - 1) * sx;
    c.dym = (c.y > 0 || !g.wy) ? -sy : (g.ny - 1) * sy;
    c.dyp = (c.y < g.ny - 1 || !g.wy) ? sy : -(g.ny - 1) * sy;
    c.dzm = (c.z > 0 || !g.wz) ? -sz : (g.nz - 1) * sz;
    c.dzp = (c.z < g.nz - 1 || !g.wz) ? sz : -(g.nz - 1) * sz;
    return c;
}

// Index of x + c_q.
__device__ __forceinline__ uint32_t nb(const Cell& c, int q)
{
    int d = 0;
    d += CX[q] > 0 ? c.dxp : (CX[q] < 0 ? c.dxm : 0);
    d += CY[q] > 0 ? c.dyp : (CY[q] < 0 ? c.dym : 0);
    d += CZ[q] > 0 ? c.dzp : (CZ[q] < 0 ? c.dzm : 0);
    return (uint32_t)((int)c.i + d);
}

// Read the pre-collision populations of a streamed cell at step parity ODD.
template <class ST, bool ODD>
__device__ __forceinline__ void load_pop(typename ST::T* const* p, const typename ST::R* rest, const Cell& c,
                                         typename ST::R* f)
{
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if (ODD)
            f[q] = ST::load(p[OPP[q]], nb(c, OPP[q]), rest[WCLASS[q]]);
        else
            f[q] = ST::load(p[q], c.i, rest[WCLASS[q]]);
    }
}

// Streaming write of post-collision population q (goes to x + c_q).
template <class ST, bool ODD>
__device__ __forceinline__ void store_out(typename ST::T* const* p, const typename ST::R* rest, const Cell& c, int q,
                                          typename ST::R v)
{
    if (ODD)
        ST::store(p[q], nb(c, q), v, rest[WCLASS[q]]);
    else
        ST::store(p[OPP[q]], c.i, v, rest[WCLASS[q]]);
}

// Write the value cell x will read as incoming population q next step
// (used when x - c_q is solid and nobody else writes that slot).
template <class ST, bool ODD>
__device__ __forceinline__ void store_in_next(typename ST::T* const* p, const typename ST::R* rest, const Cell& c,
                                              int q, typename ST::R v)
{
    if (ODD)
        ST::store(p[q], c.i, v, rest[WCLASS[q]]);
    else
        ST::store(p[OPP[q]], nb(c, OPP[q]), v, rest[WCLASS[q]]);
}

// Moving-wall update of an inlet cell's own state
// (InletVelocityMovingWall.h:43-49): for n.c_q > 0,
// f(x, c_rev[q]) = f*(x, q) - 6 w_q rho (c_q . u_w).  The dot products are
// float as in the reference (its normals and velocities are float vectors).
template <class R>
__device__ __forceinline__ void moving_wall(const Consts<R>& K, const InletDev& b, bool crev20, const R* post, R rho,
                                            R* state)
{
    LBM_FP_STRICT
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const float nc = b.nx * (float)CX[q] + b.ny * (float)CY[q] + b.nz * (float)CZ[q];
        if (nc > 0.0f) {
            const float cu = (float)CX[q] * b.ux + (float)CY[q] * b.uy + (float)CZ[q] * b.uz;
            const R v = post[q] - R(6) * K.w[WCLASS[q]] * rho * (R)cu;
            if (q == 20) {
                if (crev20) state[21] = v; else state[19] = v;
            } else {
                state[OPP[q]] = v;
            }
        }
    }
}

template <class R>
__device__ __forceinline__ R fast_rcp(R x)
{
    return R(1) / x;
}

template <class ST, int COLL, bool LIT, bool ODD, bool FLAGS>
__global__ __launch_bounds__(TPB) void k_step(StepArgs<ST> a)
{
    using R = typename ST::R;
    const Cell c = cell_of(a.g);
    bool active = c.inside;
    uint8_t cls = CL_FLUID;
    if (FLAGS && active) cls = a.cls[c.i];
    if (cls == CL_SOLID) active = false;
    double dm[4] = {0.0, 0.0, 0.0, 0.0};

    if (active) {
        R f[Q];
        int bi = -1;
        if (FLAGS && cls == CL_INLET_MW) {
            bi = a.bcidx[c.i];
#pragma unroll
            for (int q = 0; q < Q; ++q) f[q] = a.inl_state[(size_t)bi * Q + q];
        } else {
            load_pop<ST, ODD>(a.p, a.rest, c, f);
        }

        R rho, ux, uy, uz, omega;
        R P[Q];
        if (LIT) {
            if (a.first) {
                const size_t N = (size_t)a.g.nx * a.g.ny * a.g.nz;
                rho = a.macro[c.i];
                ux = a.macro[N + c.i];
                uy = a.macro[2 * N + c.i];
                uz = a.macro[3 * N + c.i];
            } else {
                moments_literal(f, rho, ux, uy, uz);
            }
            omega = a.closure == LBM_CLOSURE_SMAGORINSKY ? omega_les_literal(a.K, f, rho, ux, uy, uz, a.tau, a.cles)
                                                         : a.omega0;
#pragma unroll
            for (int q = 0; q < Q; ++q) P[q] = f[q];
            if (COLL == LBM_COLLISION_CUM2015)
                cum2015_literal(a.K, P, rho, ux, uy, uz, omega);
            else
                srt_literal(a.K, P, rho, ux, uy, uz, omega);
            if (a.diag) {
                dm[0] = (double)rho;
                dm[1] = (double)(rho * ux);
                dm[2] = (double)(rho * uy);
                dm[3] = (double)(rho * uz);
            }
        } else {
            const Raw2<R> r = raw_moments(f);
            if (a.first) {
                const size_t N = (size_t)a.g.nx * a.g.ny * a.g.nz;
                rho = a.macro[c.i];
                ux = a.macro[N + c.i];
                uy = a.macro[2 * N + c.i];
                uz = a.macro[3 * N + c.i];
            } else {
                rho = r.m0;
                const R ir = fast_rcp(rho);
                ux = r.jx * ir;
                uy = r.jy * ir;
                uz = r.jz * ir;
            }
            const Cen2<R> ce = central2(r, ux, uy, uz);
            omega = a.closure == LBM_CLOSURE_SMAGORINSKY ? omega_les_fast(ce, rho, a.tau, a.cles) : a.omega0;
            if (COLL == LBM_COLLISION_CUM2015) {
                cum2015_fast(a.K, ce, r.m0, rho, ux, uy, uz, omega, P);
            } else {
#pragma unroll
                for (int q = 0; q < Q; ++q) P[q] = f[q];
                srt_fast(a.K, P, rho, ux, uy, uz, omega);
            }
            if (a.diag) {
                dm[0] = (double)r.m0;
                dm[1] = (double)r.jx;
                dm[2] = (double)r.jy;
                dm[3] = (double)r.jz;
            }
        }

        if (!FLAGS || cls == CL_FLUID) {
#pragma unroll
            for (int q = 0; q < Q; ++q) store_out<ST, ODD>(a.p, a.rest, c, q, P[q]);
        } else {
            const uint64_t nw = a.node[c.i];
            const uint32_t S = (uint32_t)(nw & MASK27);
            const uint32_t Kk = (uint32_t)((nw >> 32) & MASK27);
            const bool wall = (nw >> 63) != 0;
            if (cls == CL_INLET_MW) {
                // stream the pre-collision state, keep the updated one
#pragma unroll
                for (int q = 0; q < Q; ++q)
                    if (!((Kk >> q) & 1u)) store_out<ST, ODD>(a.p, a.rest, c, q, f[q]);
                R st[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) st[q] = f[q];
                moving_wall(a.K, a.inl[bi], a.crev20 != 0, P, rho, st);
#pragma unroll
                for (int q = 0; q < Q; ++q) a.inl_state[(size_t)bi * Q + q] = st[q];
            } else {
#pragma unroll
                for (int q = 0; q < Q; ++q)
                    if (!((Kk >> q) & 1u)) store_out<ST, ODD>(a.p, a.rest, c, q, P[q]);
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    if ((S >> q) & 1u) {
                        R v;
                        if (!wall)
                            v = f[q];
                        else if (q == 20)
                            v = a.crev20 ? P[21] : P[19];
                        else
                            v = P[OPP[q]];
                        store_in_next<ST, ODD>(a.p, a.rest, c, q, v);
                    }
                }
            }
        }
    }

    if (a.diag) {
        // wave-level FP64 reduction, then across the 4 waves of the block
        __shared__ double red[TPB / 64][4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double v = dm[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            dm[k] = v;
        }
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) red[wv][k] = dm[k];
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < TPB / 64; ++w) s += red[w][threadIdx.x];
            a.partials[(size_t)blockIdx.x * 4 + threadIdx.x] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// Auxiliary kernels: initialisation, gather/scatter, moments, omega, sums.
// ---------------------------------------------------------------------------

template <class ST>
struct AuxArgs {
    typename ST::T* p[Q];
    const uint8_t* cls;
    const int32_t* bcidx;
    typename ST::R* inl_state;
    typename ST::R* macro;   // rho, ux, uy, uz planes
    Geo g;
    Consts<typename ST::R> K;
    typename ST::R rest[4];
    typename ST::R tau, cles, omega0;
    int closure;
    int lit;
    int odd;        // parity of the stored state
    int q;          // plane for gather/scatter
    double* io;     // staging buffer (FP64)
};

__device__ __forceinline__ uint32_t aux_cell(const Geo& g, int& x, int& y, int& z, bool& ok)
{
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    ok = i < N;
    const uint32_t ii = (uint32_t)(ok ? i : 0);
    x = (int)(ii % (uint32_t)g.nx);
    const uint32_t r = ii / (uint32_t)g.nx;
    y = (int)(r % (uint32_t)g.ny);
    z = (int)(r / (uint32_t)g.ny);
    return ii;
}

__device__ __forceinline__ Cell aux_neigh(const Geo& g, int x, int y, int z, uint32_t i)
{
    Cell c;
    c.x = x; c.y = y; c.z = z; c.i = i; c.inside = true;
    const int sx = 1, sy = g.nx, sz = g.nx * g.ny;
    c.dxm = (x > 0 || !g.wx) ? -sx : (g.nx - 1) * sx;
    c.dxp = (x < g.nx - 1 || !g.wx) ? sx : -(g.nx - 1) * sx;
    c.dym = (y > 0 || !g.wy) ? -sy : (g.ny - 1) * sy;
    c.dyp = (y < g.ny - 1 || !g.wy) ? sy : -(g.ny - 1) * sy;
    c.dzm = (z > 0 || !g.wz) ? -sz : (g.nz - 1) * sz;
    c.dzp = (z < g.nz - 1 || !g.wz) ? sz : -(g.nz - 1) * sz;
    return c;
}

// Current populations of a non-solid cell.
template <class ST>
__device__ __forceinline__ void cell_state(const AuxArgs<ST>& a, const Cell& c, uint8_t cls,
                                           typename ST::R* f)
{
    if (cls == CL_INLET_MW) {
        const int bi = a.bcidx[c.i];
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = a.inl_state[(size_t)bi * Q + q];
    } else if (a.odd) {
        load_pop<ST, true>(a.p, a.rest, c, f);
    } else {
        load_pop<ST, false>(a.p, a.rest, c, f);
    }
}

// f = f_eq(rho, u) from the macro planes (InitializationEquilibriumConstVector.h:111-126).
template <class ST>
__global__ __launch_bounds__(TPB) void k_init(AuxArgs<ST> a)
{
    using R = typename ST::R;
    int x, y, z;
    bool ok;
    const uint32_t i = aux_cell(a.g, x, y, z, ok);
    if (!ok) return;
    const size_t N = (size_t)a.g.nx * a.g.ny * a.g.nz;
    const uint8_t cls = a.cls[i];
    const R rho = a.macro[i], ux = a.macro[N + i], uy = a.macro[2 * N + i], uz = a.macro[3 * N + i];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const R v = cls == CL_SOLID ? R(0) : feq_q(a.K, q, ux, uy, uz, rho);
        ST::store(a.p[q], i, v, a.rest[WCLASS[q]]);
        if (cls == CL_INLET_MW) a.inl_state[(size_t)a.bcidx[i] * Q + q] = v;
    }
}

// Write plane q of a host-provided state (FP64 staging) at the current parity.
template <class ST>
__global__ __launch_bounds__(TPB) void k_scatter(AuxArgs<ST> a)
{
    using R = typename ST::R;
    int x, y, z;
    bool ok;
    const uint32_t i = aux_cell(a.g, x, y, z, ok);
    if (!ok) return;
    const uint8_t cls = a.cls[i];
    if (cls == CL_SOLID) return;
    const int q = a.q;
    const R v = (R)a.io[i];
    if (cls == CL_INLET_MW) {
        a.inl_state[(size_t)a.bcidx[i] * Q + q] = v;
        return;
    }
    const Cell c = aux_neigh(a.g, x, y, z, i);
    // same slot as load_pop reads
    uint32_t j = i;
    int plane = q;
    if (a.odd) {
        plane = OPP[q];
        int d = 0;
        const int o = OPP[q];
        d += CX[o] > 0 ? c.dxp : (CX[o] < 0 ? c.dxm : 0);
        d += CY[o] > 0 ? c.dyp : (CY[o] < 0 ? c.dym : 0);
        d += CZ[o] > 0 ? c.dzp : (CZ[o] < 0 ? c.dzm : 0);
        j = (uint32_t)((int)i + d);
    }
    ST::store(a.p[plane], j, v, a.rest[WCLASS[q]]);
}

// Gather plane q of the current state into FP64 staging (0 on solids).
template <class ST>
__global__ __launch_bounds__(TPB) void k_gather(AuxArgs<ST> a)
{
    using R = typename ST::R;
    int x, y, z;
    bool ok;
    const uint32_t i = aux_cell(a.g, x, y, z, ok);
    if (!ok) return;
    const uint8_t cls = a.cls[i];
    if (cls == CL_SOLID) {
        a.io[i] = 0.0;
        return;
    }
    const Cell c = aux_neigh(a.g, x, y, z, i);
    R f[Q];
    cell_state(a, c, cls, f);
    R v = f[0];
#pragma unroll
    for (int q = 1; q < Q; ++q)
        if (q == a.q) v = f[q];
    a.io[i] = (double)v;
}

// rho, u (and omega) of the current state into 4 (5) FP64 staging planes.
template <class ST>
__global__ __launch_bounds__(TPB) void k_macro(AuxArgs<ST> a)
{
    using R = typename ST::R;
    int x, y, z;
    bool ok;
    const uint32_t i = aux_cell(a.g, x, y, z, ok);
    if (!ok) return;
    const size_t N = (size_t)a.g.nx * a.g.ny * a.g.nz;
    const uint8_t cls = a.cls[i];
    if (cls == CL_SOLID) {
        for (int k = 0; k < 5; ++k) a.io[k * N + i] = 0.0;
        return;
    }
    const Cell c = aux_neigh(a.g, x, y, z, i);
    R f[Q];
    cell_state(a, c, cls, f);
    R rho, ux, uy, uz, om;
    if (a.lit) {
        moments_literal(f, rho, ux, uy, uz);
        om = a.closure == LBM_CLOSURE_SMAGORINSKY ? omega_les_literal(a.K, f, rho, ux, uy, uz, a.tau, a.cles)
                                                  : a.omega0;
    } else {
        const Raw2<R> r = raw_moments(f);
        rho = r.m0;
        ux = r.jx / rho;
        uy = r.jy / rho;
        uz = r.jz / rho;
        const Cen2<R> ce = central2(r, ux, uy, uz);
        om = a.closure == LBM_CLOSURE_SMAGORINSKY ? omega_les_fast(ce, rho, a.tau, a.cles) : a.omega0;
    }
    a.io[i] = (double)rho;
    a.io[N + i] = (double)ux;
    a.io[2 * N + i] = (double)uy;
    a.io[3 * N + i] = (double)uz;
    a.io[4 * N + i] = (double)om;
}

// rho, u of a single cell (probe).
template <class ST>
__global__ void k_probe(AuxArgs<ST> a, int x, int y, int z)
{
    using R = typename ST::R;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t i = (uint32_t)(x + a.g.nx * (y + a.g.ny * z));
    const uint8_t cls = a.cls[i];
    if (cls == CL_SOLID) {
        for (int k = 0; k < 4; ++k) a.io[k] = 0.0;
        return;
    }
    const Cell c = aux_neigh(a.g, x, y, z, i);
    R f[Q];
    cell_state(a, c, cls, f);
    R rho, ux, uy, uz;
    moments_literal(f, rho, ux, uy, uz);
    a.io[0] = (double)rho;
    a.io[1] = (double)ux;
    a.io[2] = (double)uy;
    a.io[3] = (double)uz;
}

// Per-block FP64 sums of mass and momentum over non-solid cells.
template <class ST>
__global__ __launch_bounds__(TPB) void k_sums(AuxArgs<ST> a)
{
    using R = typename ST::R;
    int x, y, z;
    bool ok;
    const uint32_t i = aux_cell(a.g, x, y, z, ok);
    double dm[4] = {0.0, 0.0, 0.0, 0.0};
    if (ok && a.cls[i] != CL_SOLID) {
        const Cell c = aux_neigh(a.g, x, y, z, i);
        R f[Q];
        cell_state(a, c, a.cls[i], f);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double v = (double)f[q];
            dm[0] += v;
            dm[1] += CX[q] * v;
            dm[2] += CY[q] * v;
            dm[3] += CZ[q] * v;
        }
    }
    __shared__ double red[TPB / 64][4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double v = dm[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        dm[k] = v;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 4; ++k) red[wv][k] = dm[k];
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int w = 0; w < TPB / 64; ++w) s += red[w][threadIdx.x];
        a.io[(size_t)blockIdx.x * 4 + threadIdx.x] = s;
    }
}

// Plain stream copy used as the bandwidth reference.
__global__ __launch_bounds__(TPB) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace lbm

// ===========================================================================
// Host side
// ===========================================================================

using namespace lbm;

struct lbm_ctx {
    lbm_desc d;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    size_t N = 0;
    size_t elem = 0;           // bytes per stored population
    size_t rbytes = 0;         // bytes per arithmetic value
    Geo g{};
    bool flags_needed = false; // any non-fluid cell
    int n_inlet = 0;
    void* pops = nullptr;
    uint8_t* cls = nullptr;
    uint64_t* node = nullptr;
    int32_t* bcidx = nullptr;
    void* inl_state = nullptr;
    InletDev* inl = nullptr;
    void* macro = nullptr;     // 4 planes, arithmetic type
    double* stage = nullptr;   // 5 planes FP64
    double* partials = nullptr;
    std::vector<double> h_partials;
    int64_t steps = 0;
    bool initialized = false;
    bool first = false;
    bool nan_seen = false;
    double last_ms = 0.0;
    size_t dev_bytes = 0;
};

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

template <class R>
static Consts<R> make_consts(uint32_t compat)
{
    Consts<R> K;
    if (compat & LBM_COMPAT_F32CONST) {
        K.w[0] = (R)(8.f / 27.f);
        K.w[1] = (R)(2.f / 27.f);
        K.w[2] = (R)(1.f / 54.f);
        K.w[3] = (R)(1.f / 216.f);
        K.c13 = (R)(1.f / 3.f);
    } else {
        K.w[0] = (R)(8.0 / 27.0);
        K.w[1] = (R)(2.0 / 27.0);
        K.w[2] = (R)(1.0 / 54.0);
        K.w[3] = (R)(1.0 / 216.0);
        K.c13 = (R)(1.0 / 3.0);
    }
    return K;
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
    a.io = c->stage;
    return a;
}

static unsigned aux_blocks(const lbm_ctx* c) { return (unsigned)((c->N + TPB - 1) / TPB); }

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
    a.macro = (const R*)c->macro;
    a.partials = c->partials;
    a.g = c->g;
    a.omega0 = R(1) / (R)c->d.tau;   // 1.f / tau (OmegaNo.h:44, NonDim...h:114)
    a.tau = (R)c->d.tau;
    a.cles = (R)c->d.c_les;
    a.crev20 = (c->d.compat & LBM_COMPAT_CREV20) ? 1 : 0;
    a.closure = c->d.closure;
    const dim3 grid((unsigned)c->g.ntiles), block(TPB);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int s = 0; s < nsteps; ++s) {
        a.first = c->first ? 1 : 0;
        a.diag = 0;
        const bool odd = (c->steps & 1) != 0;
        if (c->flags_needed) {
            if (odd) k_step<ST, COLL, LIT, true, true><<<grid, block, 0, c->stream>>>(a);
            else k_step<ST, COLL, LIT, false, true><<<grid, block, 0, c->stream>>>(a);
        } else {
            if (odd) k_step<ST, COLL, LIT, true, false><<<grid, block, 0, c->stream>>>(a);
            else k_step<ST, COLL, LIT, false, false><<<grid, block, 0, c->stream>>>(a);
        }
        HIPCHK(hipGetLastError());
        c->first = false;
        c->steps++;
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    return LBM_OK;
}

template <class ST>
static lbm_status launch_by_coll(lbm_ctx* c, int nsteps)
{
    const bool lit = c->d.arith == LBM_ARITH_LITERAL;
    if (c->d.collision == LBM_COLLISION_CUM2015)
        return lit ? launch_steps<ST, LBM_COLLISION_CUM2015, true>(c, nsteps)
                   : launch_steps<ST, LBM_COLLISION_CUM2015, false>(c, nsteps);
    return lit ? launch_steps<ST, LBM_COLLISION_SRT, true>(c, nsteps)
               : launch_steps<ST, LBM_COLLISION_SRT, false>(c, nsteps);
}

template <class ST>
static lbm_status launch_by_coll_fast(lbm_ctx* c, int nsteps)
{
    if (c->d.collision == LBM_COLLISION_CUM2015) return launch_steps<ST, LBM_COLLISION_CUM2015, false>(c, nsteps);
    return launch_steps<ST, LBM_COLLISION_SRT, false>(c, nsteps);
}

// Dispatch a functor templated on the storage policy.
template <class F>
static lbm_status with_storage(const lbm_ctx* c, F&& fn)
{
    switch (c->d.storage) {
    case LBM_STORAGE_F64: return fn(StoreF64{});
    case LBM_STORAGE_F32: return fn(StoreF32{});
    case LBM_STORAGE_F16DEV: return fn(StoreF16Dev{});
    case LBM_STORAGE_BF16DEV: return fn(StoreBF16Dev{});
    default: return LBM_E_ARG;
    }
}

// ---------------------------------------------------------------------------
// Flag-grid classification
// ---------------------------------------------------------------------------

static bool neighbour(const lbm_desc* d, int x, int y, int z, int cx, int cy, int cz, size_t* out)
{
    int xx = x + cx, yy = y + cy, zz = z + cz;
    if (xx < 0 || xx >= d->nx) { if (!d->periodic[0]) return false; xx = (xx + d->nx) % d->nx; }
    if (yy < 0 || yy >= d->ny) { if (!d->periodic[1]) return false; yy = (yy + d->ny) % d->ny; }
    if (zz < 0 || zz >= d->nz) { if (!d->periodic[2]) return false; zz = (zz + d->nz) % d->nz; }
    *out = (size_t)xx + (size_t)d->nx * ((size_t)yy + (size_t)d->ny * (size_t)zz);
    return true;
}

static lbm_status classify(const lbm_desc* d, std::vector<uint8_t>& cls, std::vector<uint64_t>& node,
                           std::vector<int32_t>& bcidx, std::vector<InletDev>& inl, bool& flags_needed)
{
    const size_t N = (size_t)d->nx * d->ny * d->nz;
    cls.assign(N, CL_SOLID);
    node.assign(N, 0);
    bcidx.assign(N, -1);
    inl.clear();
    std::vector<uint8_t> wall(N, 0);
    for (int n = 0; n < d->n_bc; ++n) {
        const lbm_bc_node& b = d->bc[n];
        if (b.x < 0 || b.y < 0 || b.z < 0 || b.x >= d->nx || b.y >= d->ny || b.z >= d->nz) return LBM_E_ARG;
        const size_t i = (size_t)b.x + (size_t)d->nx * ((size_t)b.y + (size_t)d->ny * (size_t)b.z);
        const int F = d->flags[i];
        switch (b.kind) {
        case LBM_BC_WALL:
            if (F <= 0) return LBM_E_ARG;  // wall list holds fluid nodes only (geometryMesherBoundary.h:540-542)
            wall[i] = 1;
            break;
        case LBM_BC_INLET_MOVING_WALL: {
            if (F != -1 || bcidx[i] >= 0) return LBM_E_ARG;
            InletDev e{b.nx, b.ny, b.nz, b.ux, b.uy, b.uz, b.kind, 0};
            bcidx[i] = (int32_t)inl.size();
            inl.push_back(e);
            break;
        }
        case LBM_BC_INLET_EQUILIBRIUM:
        case LBM_BC_OUTLET_INTERPOLATED:
        case LBM_BC_PERIODIC_GHOST:
            return LBM_E_UNSUPPORTED;
        default:
            return LBM_E_ARG;
        }
    }
    // cell classes
    for (size_t i = 0; i < N; ++i) {
        const int F = d->flags[i];
        if (F == 0) cls[i] = CL_SOLID;
        else if (F > 0) cls[i] = CL_FLUID;
        else if (F == -1) {
            if (bcidx[i] < 0) return LBM_E_ARG;  // inlet flag without an inlet entry
            cls[i] = CL_INLET_MW;
        } else {
            return LBM_E_UNSUPPORTED;
        }
    }
    flags_needed = false;
    for (int z = 0; z < d->nz; ++z)
        for (int y = 0; y < d->ny; ++y)
            for (int x = 0; x < d->nx; ++x) {
                const size_t i = (size_t)x + (size_t)d->nx * ((size_t)y + (size_t)d->ny * (size_t)z);
                if (cls[i] == CL_SOLID) { flags_needed = true; continue; }
                uint32_t S = 0, K = 0;
                for (int q = 1; q < Q; ++q) {
                    size_t j;
                    if (!neighbour(d, x, y, z, -CX[q], -CY[q], -CZ[q], &j)) return LBM_E_ARG;
                    if (cls[j] == CL_SOLID) S |= 1u << q;
                    if (!neighbour(d, x, y, z, CX[q], CY[q], CZ[q], &j)) return LBM_E_ARG;
                    if (cls[j] == CL_SOLID || cls[j] == CL_INLET_MW) K |= 1u << q;
                }
                uint64_t w = (uint64_t)S | ((uint64_t)K << 32) | ((uint64_t)wall[i] << 63);
                node[i] = w;
                if (cls[i] == CL_INLET_MW) { flags_needed = true; continue; }
                if (S != 0 || K != 0) {
                    cls[i] = CL_BND;
                    flags_needed = true;
                }
            }
    return LBM_OK;
}

// Upload host FP64 planes into the arithmetic-type macro buffer.
template <class R>
static lbm_status upload_macro(lbm_ctx* c, const double* rho, const double* ux, const double* uy, const double* uz)
{
    const size_t N = c->N;
    std::vector<R> h(4 * N);
    for (size_t i = 0; i < N; ++i) {
        h[i] = rho ? (R)rho[i] : (R)c->d.rho0;
        h[N + i] = ux ? (R)ux[i] : R(0);
        h[2 * N + i] = uy ? (R)uy[i] : R(0);
        h[3 * N + i] = uz ? (R)uz[i] : R(0);
    }
    HIPCHK(hipMemcpyAsync(c->macro, h.data(), 4 * N * sizeof(R), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
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
    void* bufs[] = {c->pops, c->cls, c->node, c->bcidx, c->inl_state, c->inl, c->macro, c->stage, c->partials};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
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
    if (d->collision != LBM_COLLISION_SRT && d->collision != LBM_COLLISION_CUM2015) return LBM_E_ARG;
    if (d->closure == LBM_CLOSURE_MLP) return LBM_E_UNSUPPORTED;
    if (d->closure != LBM_CLOSURE_NONE && d->closure != LBM_CLOSURE_SMAGORINSKY) return LBM_E_ARG;
    if (d->storage < LBM_STORAGE_F64 || d->storage > LBM_STORAGE_BF16DEV) return LBM_E_ARG;
    if (d->arith != LBM_ARITH_LITERAL && d->arith != LBM_ARITH_FAST) return LBM_E_ARG;
    if (d->arith == LBM_ARITH_LITERAL && d->storage >= LBM_STORAGE_F16DEV) return LBM_E_ARG;
    // tau > 1/2 keeps omega in (0, 2); NonDim...h:79 asserts this (compiled out by -DNDEBUG)
    if (!(d->tau > 0.5) || !std::isfinite(d->tau)) return LBM_E_ARG;
    if (d->closure == LBM_CLOSURE_SMAGORINSKY && !(d->c_les >= 0.0)) return LBM_E_ARG;
    const size_t N = (size_t)d->nx * d->ny * d->nz;
    const size_t elem = d->storage == LBM_STORAGE_F64 ? 8 : d->storage == LBM_STORAGE_F32 ? 4 : 2;
    if (N * elem > 0xFFFFFFFFull || N >= 0x7FFFFFFFull) return LBM_E_ARG;  // 32-bit plane offsets
    if (d->periodic[0] && d->nx < 2) return LBM_E_ARG;
    if (d->periodic[1] && d->ny < 2) return LBM_E_ARG;
    if (d->periodic[2] && d->nz < 2) return LBM_E_ARG;

    std::vector<uint8_t> cls;
    std::vector<uint64_t> node;
    std::vector<int32_t> bcidx;
    std::vector<InletDev> inl;
    bool flags_needed = false;
    lbm_status st = classify(d, cls, node, bcidx, inl, flags_needed);
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
    c->elem = elem;
    c->rbytes = d->storage == LBM_STORAGE_F64 ? 8 : 4;
    c->flags_needed = flags_needed;
    c->n_inlet = (int)inl.size();
    (void)hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        lbm_destroy(c);
        return LBM_E_HIP;
    }
    // tile geometry: 256 threads = (1 << bx_log2) x by cells
    int bxl = 6;
    if (d->nx > 64) bxl = 7;
    if (d->nx > 128) bxl = 8;
    c->g.nx = d->nx;
    c->g.ny = d->ny;
    c->g.nz = d->nz;
    c->g.bx_log2 = bxl;
    c->g.by = TPB >> bxl;
    c->g.ntx = (d->nx + (1 << bxl) - 1) >> bxl;
    c->g.nty = (d->ny + c->g.by - 1) / c->g.by;
    c->g.ntiles = c->g.ntx * c->g.nty * d->nz;
    c->g.wx = d->periodic[0] != 0;
    c->g.wy = d->periodic[1] != 0;
    c->g.wz = d->periodic[2] != 0;

    st = dev_alloc(c, &c->pops, N * elem * Q);
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->cls, N);
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->node, N * sizeof(uint64_t));
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->bcidx, N * sizeof(int32_t));
    if (st == LBM_OK) st = dev_alloc(c, &c->inl_state, (size_t)c->n_inlet * Q * c->rbytes);
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->inl, inl.size() * sizeof(InletDev));
    if (st == LBM_OK) st = dev_alloc(c, &c->macro, 4 * N * c->rbytes);
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->stage, 5 * N * sizeof(double));
    const size_t nparts = std::max((size_t)c->g.ntiles, (size_t)aux_blocks(c));
    if (st == LBM_OK) st = dev_alloc(c, (void**)&c->partials, nparts * 4 * sizeof(double));
    if (st != LBM_OK) {
        lbm_destroy(c);
        return st;
    }
    c->h_partials.resize(nparts * 4);
    if (hipMemcpyAsync(c->cls, cls.data(), N, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->node, node.data(), N * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->bcidx, bcidx.data(), N * sizeof(int32_t), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        (!inl.empty() && hipMemcpyAsync(c->inl, inl.data(), inl.size() * sizeof(InletDev), hipMemcpyHostToDevice,
                                        c->stream) != hipSuccess) ||
        hipMemsetAsync(c->pops, 0, N * elem * Q, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        lbm_destroy(c);
        return LBM_E_HIP;
    }
    *out = c;
    return LBM_OK;
}

lbm_status lbm_init_equilibrium(lbm_ctx* c, const double* rho, const double* ux, const double* uy, const double* uz)
{
    if (!c) return LBM_E_ARG;
    lbm_status st = c->rbytes == 8 ? upload_macro<double>(c, rho, ux, uy, uz)
                                   : upload_macro<float>(c, rho, ux, uy, uz);
    if (st != LBM_OK) return st;
    // solid cells carry rho = u = 0 in the reference (InitializationEquilibriumConstVector.h:84-89)
    c->steps = 0;
    st = with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        AuxArgs<ST> a = aux_args<ST>(c);
        a.odd = 0;
        k_init<ST><<<aux_blocks(c), TPB, 0, c->stream>>>(a);
        HIPCHK(hipGetLastError());
        return LBM_OK;
    });
    if (st != LBM_OK) return st;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->initialized = true;
    c->first = true;
    c->nan_seen = false;
    return LBM_OK;
}

lbm_status lbm_set_populations(lbm_ctx* c, const double* f)
{
    if (!c || !f) return LBM_E_ARG;
    for (int q = 0; q < Q; ++q) {
        HIPCHK(hipMemcpyAsync(c->stage, f + (size_t)q * c->N, c->N * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        lbm_status st = with_storage(c, [&](auto stv) -> lbm_status {
            using ST = decltype(stv);
            AuxArgs<ST> a = aux_args<ST>(c);
            a.q = q;
            k_scatter<ST><<<aux_blocks(c), TPB, 0, c->stream>>>(a);
            HIPCHK(hipGetLastError());
            return LBM_OK;
        });
        if (st != LBM_OK) return st;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->initialized = true;
    c->first = false;
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
    default: return LBM_E_ARG;
    }
    if (st != LBM_OK) return st;
    c->last_ms = -(double)nsteps;  // resolved at the next sync
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

lbm_status lbm_get_macroscopic(lbm_ctx* c, double* rho, double* ux, double* uy, double* uz)
{
    if (!c) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    const size_t N = c->N;
    double* outs[4] = {rho, ux, uy, uz};
    if (c->first) {
        // before the first step the reference's rho/u arrays hold the initial values
        std::vector<uint8_t> cls(N);
        HIPCHK(hipMemcpy(cls.data(), c->cls, N, hipMemcpyDeviceToHost));
        if (c->rbytes == 8) {
            std::vector<double> h(4 * N);
            HIPCHK(hipMemcpy(h.data(), c->macro, 4 * N * sizeof(double), hipMemcpyDeviceToHost));
            for (int k = 0; k < 4; ++k)
                if (outs[k])
                    for (size_t i = 0; i < N; ++i) outs[k][i] = cls[i] == CL_SOLID ? 0.0 : h[k * N + i];
        } else {
            std::vector<float> h(4 * N);
            HIPCHK(hipMemcpy(h.data(), c->macro, 4 * N * sizeof(float), hipMemcpyDeviceToHost));
            for (int k = 0; k < 4; ++k)
                if (outs[k])
                    for (size_t i = 0; i < N; ++i) outs[k][i] = cls[i] == CL_SOLID ? 0.0 : (double)h[k * N + i];
        }
        return LBM_OK;
    }
    st = with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        AuxArgs<ST> a = aux_args<ST>(c);
        k_macro<ST><<<aux_blocks(c), TPB, 0, c->stream>>>(a);
        HIPCHK(hipGetLastError());
        return LBM_OK;
    });
    if (st != LBM_OK) return st;
    for (int k = 0; k < 4; ++k)
        if (outs[k])
            HIPCHK(hipMemcpyAsync(outs[k], c->stage + k * N, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_probe(lbm_ctx* c, int32_t x, int32_t y, int32_t z, double out[4])
{
    if (!c || !out) return LBM_E_ARG;
    if (x < 0 || y < 0 || z < 0 || x >= c->g.nx || y >= c->g.ny || z >= c->g.nz) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    if (c->first) {
        double tmp[4];
        const size_t i = (size_t)x + (size_t)c->g.nx * ((size_t)y + (size_t)c->g.ny * (size_t)z);
        uint8_t cls = 0;
        HIPCHK(hipMemcpy(&cls, c->cls + i, 1, hipMemcpyDeviceToHost));
        for (int k = 0; k < 4; ++k) {
            if (c->rbytes == 8) {
                HIPCHK(hipMemcpy(&tmp[k], (double*)c->macro + k * c->N + i, 8, hipMemcpyDeviceToHost));
            } else {
                float v;
                HIPCHK(hipMemcpy(&v, (float*)c->macro + k * c->N + i, 4, hipMemcpyDeviceToHost));
                tmp[k] = v;
            }
            out[k] = cls == CL_SOLID ? 0.0 : tmp[k];
        }
        return LBM_OK;
    }
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

lbm_status lbm_get_omega(lbm_ctx* c, double* omega)
{
    if (!c || !omega) return LBM_E_ARG;
    if (!c->initialized) return LBM_E_STATE;
    lbm_status st = lbm_sync(c);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    st = with_storage(c, [&](auto stv) -> lbm_status {
        using ST = decltype(stv);
        AuxArgs<ST> a = aux_args<ST>(c);
        k_macro<ST><<<aux_blocks(c), TPB, 0, c->stream>>>(a);
        HIPCHK(hipGetLastError());
        return LBM_OK;
    });
    if (st != LBM_OK) return st;
    HIPCHK(hipMemcpyAsync(omega, c->stage + 4 * c->N, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
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
            k_gather<ST><<<aux_blocks(c), TPB, 0, c->stream>>>(a);
            HIPCHK(hipGetLastError());
            return LBM_OK;
        });
        if (st != LBM_OK) return st;
        HIPCHK(hipMemcpyAsync(f + (size_t)q * c->N, c->stage, c->N * sizeof(double), hipMemcpyDeviceToHost,
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
    const unsigned nb = aux_blocks(c);
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

}  // extern "C"
