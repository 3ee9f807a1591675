// lbm_kernels.hpp -- device code of the collide-stream core (gfx950).
//
// One fused kernel per time step (k_step): load the 27 populations of a cell,
// derive rho/u, the relaxation rate (constant or Smagorinsky), collide
// (cumulant 2015 or BGK) and stream in place with the A-A pattern, applying
// bounce-back and the moving-wall inlet from the flag grid in the same pass.
// It replaces the reference's per-step operator chain
// OmegaLES -> Collision -> StreamingAB -> BounceBackWallHalf -> Inlet ->
// Outlet -> Periodic -> MomentDensityVelocityN27 (SolverTurbulentLES.h:187-214),
// which reads and writes two full lattices plus rho/u/omega every step.
//
// A-A layout: one population array P[q][cell] (structure of arrays, x
// fastest, a solid halo on every non-periodic axis: one cell in y and z, a
// whole cache line of cells in x so that rows start on a line).  With t the
// number of completed steps:
//   even t: cell x reads f_q from P[q][x] and writes post-collision f*_q to
//           P[opp q][x];
//   odd t:  cell x reads f_q from P[opp q][x - c_q] and writes f*_q to
//           P[q][x + c_q].
// Every slot is read and written by the same thread within a step, so the
// scheme is race-free without a second lattice.
//
// Boundaries (reference semantics; the oracle restates the AB step):
//  * source x - c_q solid: the cell writes the slot that source would have
//    written -- f*_{crev q} for wall-list nodes (half-way bounce-back,
//    BounceBackWallHalf.h:46-47) or its own pre-collision f_q otherwise
//    (StreamingAB.h:58-65 leaves the slot untouched);
//  * destination x + c_q solid or inlet: the streaming write is skipped;
//  * inlet cells keep their own state in a side buffer and stream their
//    pre-collision populations (StreamingAB.h:62-65); the moving-wall rule
//    (InletVelocityMovingWall.h:43-49) updates that state in the same pass;
//  * the equilibrium inlet, interpolated outlet and ghost-layer periodic
//    rules read values that other cells produce in the same step, so the
//    fused kernel publishes them (neighbour rho/u, post-collision populations)
//    into small side buffers and per-list patch kernels apply the rules after
//    the step, in the reference's order (inlet, outlet, periodic).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/lbm_core.h"
#include "collide.hpp"
#include "mlp.hpp"
#include "storage.hpp"

namespace lbm {

// per-cell class
enum : uint8_t {
    CL_SOLID = 0,
    CL_FLUID = 1,  // streamed, every neighbour streamed too
    CL_BND = 2,    // streamed, some neighbour solid, inlet or outside (masks in the node word),
                   // or a cell the Combined operator collides with its 2017 branch
    CL_INLET = 3,  // inlet node (flag -1): own state in a side buffer
    CL_PUB = 4,    // streamed like CL_BND; also publishes rho, u_x and f* for a patch kernel
    CL_OUTL = 5    // interpolated outlet node whose rule its next reader applies; publishes
                   // like CL_PUB through its entry's slot (bcidx = outlet entry)
};

constexpr int TPB = 256;
// node word of a non-CL_FLUID cell: bits 0-26 source mask S (populations that
// do not stream in), bits 32-58 destination mask K (streaming writes dropped),
// bit 62 the Combined operator's 2017 branch (flag != 1), bit 63 wall-list node
constexpr uint64_t MASK27 = (1ull << 27) - 1;
constexpr int NW_CUM2017 = 62;
constexpr int PUB_MAC = 4;   // published rho, ux, uy, uz per publish slot

struct Geo {
    int nx, ny, nz;     // padded lattice
    int px, py, pz;     // offset of the caller's lattice: halo width per axis (0 on
                        // index-wrap axes; along x a whole cache line of solid cells)
    int ux, uy, uz;     // caller's lattice
    int bx_log2, by;    // tile: (1 << bx_log2) x by cells, 256 threads
    int ntx, nty;       // tiles along x and y
    int ntiles;
    int wx, wy, wz;     // index wrap per axis
};

// Boundary-node record on the device (padded coordinates).  Publish slots
// index the per-step side buffers written by CL_PUB cells; -1 stands for a
// solid cell, whose rho, u and f* the reference keeps at zero.
struct NodeDev {
    int32_t x, y, z;
    int32_t kind;      // LBM_BC_*
    float nx, ny, nz;
    float ux, uy, uz;
    float rho;
    int32_t regular;
    int32_t s0;        // equilibrium inlet: slot of x - n; outlet and ghost: own slot
    int32_t s1;        // equilibrium inlet: slot of x - 2n; outlet: upstream slot, -2 when outside the lattice
    int32_t partner;   // ghost: padded index of the partner cell
};

template <class ST>
struct StepArgs {
    typename ST::T* p[Q];
    const uint8_t* cls;
    const uint64_t* node;
    const int32_t* bcidx;          // inlet: state index; CL_PUB: publish slot
    typename ST::R* inl_state;     // [n_inlet][27]
    typename ST::R* inl_post;      // [n_inlet][27]: moving-wall / Zou-He cells' f* of this step
    typename ST::R* inl_mac;       // [n_inlet][4]: their rho, ux, uy, uz
    const NodeDev* inl;            // [n_inlet]
    typename ST::R* pub_mac;       // [n_pub][PUB_MAC]: rho, u_x, u_y, u_z
    typename ST::R* pub_post;      // [n_pub][27]: post-collision populations
    const typename ST::R* pub_prev;  // the previous step's records (deferred outlet)
    const NodeDev* outl;           // outlet entries (deferred outlet)
    const typename ST::R* macro;   // rho, ux, uy, uz planes for the first step
    double* partials;              // [ntiles][8] when diag != 0: mass, momentum, wall force
    float* ccache;                 // learned closure: coefficient per cell, held between refreshes
    double* avg;                   // non-null: add rho, u of the state this step reads to the
                                   // time-average sums (4 FP64 planes, caller layout)
    Geo g;
    Consts<typename ST::R> K;
    typename ST::R rest[4];        // w*rho0 per weight class (deviation storage)
    typename ST::R omega0, tau, cles;
    int crev20;
    int first;
    int diag;
    int closure;
    uint32_t seed;                 // per-step seed of the stochastic-rounding stores
    int sp_axis, sp_begin;         // outlet omega damping: axis (-1 off) and first cell
    typename ST::R sp_deriv;       // omega decrement per cell past sp_begin
    int outl_pending;              // CL_OUTL cells' loaded state still needs the outlet rule
    typename ST::R cs;             // the outlet's interpolation weight
    int inl_pending;               // inlet cells' stored state still needs its rule (deferred inlet)
};

// Outlet omega damping (OutletNeighbourEquilibriumOmega.h:18-79) of a cell at
// padded coordinates: omega - deriv * (coordinate - begin) past begin.
template <class R>
__device__ __forceinline__ R sponge(R omega, int axis, int begin, R deriv, const Geo& g, int x, int y, int z)
{
    LBM_FP_STRICT
    if (axis < 0) return omega;
    const int v = axis == 0 ? x - g.px : axis == 1 ? y - g.py : z - g.pz;
    if (v < begin) return omega;
    return omega - deriv * (R)(v - begin);
}

// OutletDensityInterpolated::outlet (OutletDensityInterpolated.h:170-252) on
// one node: interpolated incoming populations, then the density reset.
// The interpolated rule itself on the node's current populations d (the
// state after streaming), with pub the post-collision records of the step
// that produced that state.  b.partner holds the irregular +x node's copy
// mask (host-computed: directions whose cell (x, y - c_y, z - c_z) is solid),
// applied in q order as the reference's loop does.
template <class R>
__device__ __forceinline__ bool outlet_blend(R* d, const NodeDev& b, const R* pub, R cs, bool crev20, R& rho, R& ux,
                                             R& uy, R& uz)
{
    LBM_FP_STRICT
    const uint32_t mask = (uint32_t)b.partner;
#pragma unroll
    for (int vel = 0; vel < Q; ++vel)
        if ((mask >> vel) & 1u) d[vel] = d[crev_of(vel, crev20)];
    if (b.s1 == -2) return false;
    const R* self = pub + (size_t)b.s0 * Q;
#pragma unroll
    for (int vel = 0; vel < Q; ++vel) {
        const float nc = b.nx * (float)CX[vel] + b.ny * (float)CY[vel] + b.nz * (float)CZ[vel];
        if (nc < 0.0f) {
            const R up = b.s1 >= 0 ? pub[(size_t)b.s1 * Q + vel] : R(0);
            d[vel] = cs * up + (R(1) - cs) * self[vel];
        }
    }
    const R density = b.rho;
    rho = d[0] + d[1] + d[2] + d[3] + d[4] + d[5] + d[6] + d[7] + d[8] + d[9] + d[10] + d[11] + d[12] + d[13] +
          d[14] + d[15] + d[16] + d[17] + d[18] + d[19] + d[20] + d[21] + d[22] + d[23] + d[24] + d[25] + d[26];
    ux = (d[1] - d[2] + d[7] - d[8] + d[9] - d[10] - d[11] + d[12] - d[15] + d[16] - d[19] + d[20] - d[21] + d[22] +
          d[23] - d[24] - d[25] + d[26]) / density;
    uy = (-d[5] + d[6] - d[11] + d[12] + d[13] - d[14] + d[15] - d[16] + d[17] - d[18] + d[19] - d[20] - d[21] +
          d[22] - d[23] + d[24] - d[25] + d[26]) / density;
    uz = (-d[3] + d[4] - d[7] + d[8] + d[9] - d[10] - d[13] + d[14] + d[17] - d[18] - d[19] + d[20] + d[21] - d[22] -
          d[23] + d[24] - d[25] + d[26]) / density;
    return true;
}

// The density reset of population vel: f - f_eq(u, rho) + f_eq(u, rho_b = 1).
template <class R>
__device__ __forceinline__ R outlet_reset(const Consts<R>& K, int vel, R v, R ux, R uy, R uz, R rho)
{
    LBM_FP_STRICT
    return v - feq_q(K, vel, ux, uy, uz, rho) + feq_q(K, vel, ux, uy, uz, R(1));
}

template <class R>
__device__ __forceinline__ void interp_outlet(R* d, const NodeDev& b, const R* pub, const Consts<R>& K, R cs,
                                              bool crev20)
{
    R rho, ux, uy, uz;
    if (outlet_blend(d, b, pub, cs, crev20, rho, ux, uy, uz)) {
#pragma unroll
        for (int vel = 0; vel < Q; ++vel) d[vel] = outlet_reset(K, vel, d[vel], ux, uy, uz, rho);
    }
}

// The rule as the next reader of a deferred outlet cell applies it: to the
// populations just loaded, rounded through the storage format as the patch
// kernel's store and the next load would round them.
template <class ST>
__device__ __forceinline__ void deferred_outlet(typename ST::R* f, const NodeDev& b, const typename ST::R* pub,
                                                const Consts<typename ST::R>& K, typename ST::R cs, bool crev20,
                                                const typename ST::R* rest)
{
    interp_outlet(f, b, pub, K, cs, crev20);
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = ST::round_trip(f[q], rest[WCLASS[q]]);
}

// Arguments of the collide-stream step with the learned closure.
template <class ST>
struct MlpStepArgs {
    StepArgs<ST> s;
    MlpWeights mlp;
};

struct Cell {
    int x, y, z;
    uint32_t i;
    int dxm, dxp, dym, dyp, dzm, dzp;
    bool inside;
};

__device__ __forceinline__ void set_deltas(const Geo& g, Cell& c)
{
    const int sx = 1, sy = g.nx, sz = g.nx * g.ny;
    c.dxm = (c.x > 0 || !g.wx) ? -sx : (g.nx - 1) * sx;
    c.dxp = (c.x < g.nx - 1 || !g.wx) ? sx : -(g.nx - 1) * sx;
    c.dym = (c.y > 0 || !g.wy) ? -sy : (g.ny - 1) * sy;
    c.dyp = (c.y < g.ny - 1 || !g.wy) ? sy : -(g.ny - 1) * sy;
    c.dzm = (c.z > 0 || !g.wz) ? -sz : (g.nz - 1) * sz;
    c.dzp = (c.z < g.nz - 1 || !g.wz) ? sz : -(g.nz - 1) * sz;
}

__device__ __forceinline__ Cell cell_at(const Geo& g, int x, int y, int z)
{
    Cell c;
    c.x = x;
    c.y = y;
    c.z = z;
    c.inside = true;
    c.i = (uint32_t)(x + g.nx * (y + g.ny * z));
    set_deltas(g, c);
    return c;
}

// Tiles cover the caller's cells only (halo cells are solid and never
// collide).  Workgroups go to the XCDs in the hardware's round-robin order
// (workgroup b on XCD b mod 8), which keeps all of them on one narrow band of
// each plane.  Where a row of tiles holds several tiles (ntx > 1), the tiles of
// one row group are given workgroup ids that are equal mod 8: one XCD then
// handles the whole row group, and the odd step's x-shifted reads across a
// tile edge find the neighbour tile's cache lines in the same L2.
__device__ __forceinline__ int tile_of_block(const Geo& g, int b)
{
    if (g.ntx == 1) return b;
    const int nrg = g.ntiles / g.ntx;   // row groups: (y tile, z) pairs
    const int full = nrg >> 3;          // complete groups of 8 row groups
    const int grp = 8 * g.ntx;
    const int gi = b / grp;
    int R, k;
    if (gi < full) {
        const int rem = b - gi * grp;
        k = rem >> 3;
        R = gi * 8 + (rem & 7);
    } else {
        const int rem = b - full * grp;
        R = full * 8 + rem / g.ntx;
        k = rem % g.ntx;
    }
    return R * g.ntx + k;
}

__device__ __forceinline__ Cell cell_of(const Geo& g)
{
    Cell c;
    const int t = tile_of_block(g, (int)blockIdx.x);
    const int tx = t % g.ntx;
    const int rem = t / g.ntx;
    const int ty = rem % g.nty;
    c.z = g.pz + rem / g.nty;
    const int ux = (tx << g.bx_log2) + ((int)threadIdx.x & ((1 << g.bx_log2) - 1));
    const int uy = ty * g.by + ((int)threadIdx.x >> g.bx_log2);
    c.x = g.px + ux;
    c.y = g.py + uy;
    c.inside = ux < g.ux && uy < g.uy;
    c.i = (uint32_t)(c.x + g.nx * (c.y + g.ny * c.z));
    set_deltas(g, c);
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

// Slot holding the pre-collision population q of cell c when the next step
// has parity odd (the read location of the A-A scheme).
__device__ __forceinline__ void read_slot(const Cell& c, int q, bool odd, int& plane, uint32_t& idx)
{
    if (odd) {
        plane = OPP[q];
        idx = nb(c, OPP[q]);
    } else {
        plane = q;
        idx = c.i;
    }
}

// A wave's access to population q covers whole cache lines when it is not
// shifted along x: every access of the even step, and on the odd step the
// directions with c_x = 0 (tiles are whole rows or power-of-two segments).
__host__ __device__ constexpr bool line_aligned(bool odd, int q) { return !odd || CX[q] == 0; }
__host__ __device__ constexpr int read_plane(bool odd, int q) { return odd ? OPP[q] : q; }
__device__ __forceinline__ uint32_t read_index(const Cell& c, bool odd, int q) { return odd ? nb(c, OPP[q]) : c.i; }

// The boundary path needs a few pre-collision values after the collision; it
// reads them again instead of keeping all 27 in registers through it.  Here:
// incoming population q whose source x - c_q is solid or outside.  Nobody
// writes its read slot this step (the cell's own write there is masked, since
// x - c_q is no streaming destination, and solid cells do not run), so the
// slot still holds the value read at the start.
template <class ST, bool ODD>
__device__ __forceinline__ typename ST::R reload_pop(typename ST::T* const* p, const typename ST::R* rest,
                                                     const Cell& c, int q)
{
    return ST::template get<false>(p[read_plane(ODD, q)], read_index(c, ODD, q), rest[WCLASS[q]]);
}

template <class ST, bool ODD>
__device__ __forceinline__ void load_pop(typename ST::T* const* p, const typename ST::R* rest, const Cell& c,
                                         typename ST::R* f)
{
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (line_aligned(ODD, q))
            f[q] = ST::template get<true>(p[read_plane(ODD, q)], read_index(c, ODD, q), rest[WCLASS[q]]);
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (!line_aligned(ODD, q))
            f[q] = ST::template get<false>(p[read_plane(ODD, q)], read_index(c, ODD, q), rest[WCLASS[q]]);
}

// Streaming write of post-collision population q (goes to x + c_q).
// h: the cell's random bits for stochastic-rounding storage (else unused);
// NT: the streaming hint, for the line-aligned directions.
template <class ST, bool ODD, bool NT>
__device__ __forceinline__ void store_out(typename ST::T* const* p, const typename ST::R* rest, const Cell& c, int q,
                                          typename ST::R v, uint32_t h)
{
    if (ODD)
        ST::template put<NT>(p[q], nb(c, q), v, rest[WCLASS[q]], sr_bits(h, q));
    else
        ST::template put<NT>(p[OPP[q]], c.i, v, rest[WCLASS[q]], sr_bits(h, q));
}

// Write the value cell x will read as incoming population q next step
// (used when x - c_q is solid and nobody else writes that slot).
template <class ST, bool ODD>
__device__ __forceinline__ void store_in_next(typename ST::T* const* p, const typename ST::R* rest, const Cell& c,
                                              int q, typename ST::R v, uint32_t h)
{
    if (ODD)
        ST::template put<false>(p[q], c.i, v, rest[WCLASS[q]], sr_bits(h, 27 + q));
    else
        ST::template put<false>(p[OPP[q]], nb(c, OPP[q]), v, rest[WCLASS[q]], sr_bits(h, 27 + q));
}

// Moving-wall update of an inlet cell's own state
// (InletVelocityMovingWall.h:43-49): for n.c_q > 0,
// f(x, c_rev[q]) = f*(x, q) - 6 w_q rho (c_q . u_w).  The dot products are
// float as in the reference (its normals and velocities are float vectors).
template <class R>
__device__ __forceinline__ void moving_wall(const Consts<R>& K, const NodeDev& b, bool crev20, const R* post, R rho,
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

// InletVelocityZouHe (InletVelocityZouHe.h:80-101): for every outgoing
// direction q (n.c_q > 0), state[c_rev q] = (f*(q) - f_eq(q; rho, u)) +
// f_eq(c_rev q; rho, u_inlet).  The inlet velocity is a float vector in the
// reference, so that equilibrium's polynomial is float.
template <class R>
__device__ __forceinline__ void zou_he(const Consts<R>& K, const NodeDev& b, bool crev20, const R* post, R rho, R ux,
                                       R uy, R uz, R* state)
{
    LBM_FP_STRICT
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const float nc = b.nx * (float)CX[q] + b.ny * (float)CY[q] + b.nz * (float)CZ[q];
        if (nc > 0.0f) {
            const R fneq = post[q] - feq_q(K, q, ux, uy, uz, rho);
            const int r = crev_of(q, crev20);
            const float uc = (float)CX[r] * b.ux + (float)CY[r] * b.uy + (float)CZ[r] * b.uz;
            const float u2 = b.ux * b.ux + b.uy * b.uy + b.uz * b.uz;
            state[r] = fneq + K.w[WCLASS[r]] * rho * (R)(1.f + 3.f * uc + 4.5f * uc * uc - 1.5f * u2);
        }
    }
}

// The update a moving-wall or Zou-He inlet cell's rule makes to its stored
// state st after a step, from that step's records: the cell's own
// post-collision populations (post) and its rho, u (own).
template <class R>
__device__ __forceinline__ void inlet_rule(const Consts<R>& K, const NodeDev& b, const R* post, const R* own,
                                           bool crev20, R* st)
{
    if (b.kind == LBM_BC_INLET_MOVING_WALL)
        moving_wall(K, b, crev20, post, own[0], st);
    else if (b.kind == LBM_BC_INLET_ZOU_HE)
        zou_he(K, b, crev20, post, own[0], own[1], own[2], own[3], st);
}

template <class R>
__device__ __forceinline__ R fast_rcp(R x)
{
    return R(1) / x;
}

// Caller-layout index of a padded cell.
__device__ __forceinline__ size_t user_index(const Geo& g, const Cell& c)
{
    return (size_t)(c.x - g.px) + (size_t)g.ux * ((size_t)(c.y - g.py) + (size_t)g.uy * (size_t)(c.z - g.pz));
}

// MomentTimeAvg::momentAdd of one cell (MomentTimeAvg.h:19-45): running FP64
// sums of rho and u, four planes of nu cells.
__device__ __forceinline__ void moment_add(double* avg, size_t nu, size_t iu, double rho, double ux, double uy,
                                           double uz)
{
    avg[iu] += rho;
    avg[nu + iu] += ux;
    avg[2 * nu + iu] += uy;
    avg[3 * nu + iu] += uz;
}

// Streaming, bounce-back and boundary bookkeeping of one cell after its
// collision: f = pre-collision populations, P = post-collision populations.
// (k_step keeps its own inline copy of this code: passing the population
// arrays through a helper raised its register count from 48-64 to > 100.)
template <class ST, bool ODD, bool FLAGS>
__device__ __forceinline__ void stream_cell(const StepArgs<ST>& a, const Cell& c, uint8_t cls, int bi, bool inlet,
                                            typename ST::R (&P)[Q],
                                            typename ST::R rho, typename ST::R ux, typename ST::R uy,
                                            typename ST::R uz)
{
    using R = typename ST::R;
    const uint32_t hsr = ST::sr ? sr_hash(c.i, a.seed) : 0u;
    // node word of a boundary cell (0 for plain fluid): skip masks and walls
    const uint64_t nw = (FLAGS && cls != CL_FLUID) ? a.node[c.i] : 0ull;
    const uint32_t S = (uint32_t)(nw & MASK27);
    const uint32_t Kk = (uint32_t)((nw >> 32) & MASK27);
    const bool wall = (nw >> 63) != 0;
    // boundary work first, as in k_step
    if (FLAGS && cls != CL_FLUID && inlet) {
        const int kind = a.inl[bi].kind;
        if (kind == LBM_BC_INLET_MOVING_WALL || kind == LBM_BC_INLET_ZOU_HE) {
            R* m = a.inl_mac + (size_t)bi * PUB_MAC;
            m[0] = rho;
            m[1] = ux;
            m[2] = uy;
            m[3] = uz;
            R* dst = a.inl_post + (size_t)bi * Q;
#pragma unroll
            for (int q = 0; q < Q; ++q) dst[q] = P[q];
        }
        // an inlet cell streams its pre-collision state (StreamingAB.h:62-65)
#pragma unroll
        for (int q = 0; q < Q; ++q) P[q] = a.inl_state[(size_t)bi * Q + q];
    }
    if (FLAGS && cls != CL_FLUID && !inlet) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if ((S >> q) & 1u) {
                R v;
                if (!wall)
                    v = reload_pop<ST, ODD>(a.p, a.rest, c, q);
                else if (q == 20)
                    v = a.crev20 ? P[21] : P[19];
                else
                    v = P[OPP[q]];
                store_in_next<ST, ODD>(a.p, a.rest, c, q, v, hsr);
            }
        }
        // data the boundary patch kernels of this step read
        if (cls == CL_PUB) {
            a.pub_mac[(size_t)bi * PUB_MAC] = rho;
            a.pub_mac[(size_t)bi * PUB_MAC + 1] = ux;
            a.pub_mac[(size_t)bi * PUB_MAC + 2] = uy;
            a.pub_mac[(size_t)bi * PUB_MAC + 3] = uz;
            R* dst = a.pub_post + (size_t)bi * Q;
#pragma unroll
            for (int q = 0; q < Q; ++q) dst[q] = P[q];
        }
    }
    // streaming writes, shared by every lane of the wave (boundary lanes drop
    // the writes their mask skips)
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (line_aligned(ODD, q) && !((Kk >> q) & 1u))
            store_out<ST, ODD, true>(a.p, a.rest, c, q, P[q], hsr);
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (!line_aligned(ODD, q) && !((Kk >> q) & 1u))
            store_out<ST, ODD, false>(a.p, a.rest, c, q, P[q], hsr);
}

// INL: inlet rules deferred into this kernel (a separate instantiation: the
// rule's code, even untaken, cost the obstacle workload's flagged step 1-2 %
// in FP32/FP64 and 8 % in FP16 storage).
template <class ST, int COLL, bool LIT, bool ODD, bool FLAGS, bool CC, bool INL>
__device__ __forceinline__ void step_cells(const StepArgs<ST>& a)
{
#include "step_body.inc"
}

// The step kernel.  For FP32 storage it runs the body through step_cells;
// for the 2- and 8-byte formats the body is compiled in the kernel itself.
// The two forms are the same source but compile to different schedules: the
// even FP32 step through step_cells computes all 27 post-collision values
// before its stores (107 VGPRs) and ran 2.9 % faster on the periodic 256^3
// workload (cumulant 2017: 3.2 %), while the 16-bit steps ran 2.1-2.6 %
// slower that way and FP64 0.5 %.
template <class ST, int COLL, bool LIT, bool ODD, bool FLAGS, bool CC = false, bool INL = false>
__global__ __launch_bounds__(TPB) void k_step(StepArgs<ST> a)
{
    if constexpr (sizeof(typename ST::T) == 4) {
        step_cells<ST, COLL, LIT, ODD, FLAGS, CC, INL>(a);
    } else {
#include "step_body.inc"
    }
}

// The boundary-aware step for 16-bit storage, held to 7 waves per SIMD (72
// VGPRs instead of 75-76, which allow 6).  A 16-bit wave moves half the bytes
// of an FP32 one, so it needs more waves in flight: on the obstacle workload
// the step took 2 % less time.  FP32 was unchanged, and 8 waves spilled and
// ran 7 % slower.
template <class ST, int COLL, bool LIT, bool ODD, bool INL = false>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(7, 7))) void k_step_w7(StepArgs<ST> a)
{
    step_cells<ST, COLL, LIT, ODD, true, false, INL>(a);
}

// Moments, Smagorinsky rate and closure features of one cell, shared by the
// collide-stream kernel and the closure read-back kernel.
template <class ST>
struct ClosurePrep {
    typename ST::R rho, ux, uy, uz, m0, pin, om_s;
    Cen2<typename ST::R> ce;
    float x[NFEAT];   // raw features
    float xn[NFEAT];  // normalised features
    bool env;         // inside the training envelope
};

template <class ST>
__device__ __forceinline__ void closure_prep(const typename ST::R* f, bool first, const typename ST::R* macro,
                                             size_t N, uint32_t i, typename ST::R tau, typename ST::R cles,
                                             const MlpLds& L, ClosurePrep<ST>& o)
{
    using R = typename ST::R;
    const Raw2<R> r = raw_moments(f);
    o.m0 = r.m0;
    if (first) {
        o.rho = macro[i];
        o.ux = macro[N + i];
        o.uy = macro[2 * N + i];
        o.uz = macro[3 * N + i];
    } else {
        o.rho = r.m0;
        const R ir = fast_rcp(o.rho);
        o.ux = r.jx * ir;
        o.uy = r.jy * ir;
        o.uz = r.jz * ir;
    }
    o.ce = central2(r, o.ux, o.uy, o.uz);
    o.om_s = omega_les_fast(o.ce, o.rho, tau, cles);
    R om_unused;
    cell_features(o.ce, o.rho, tau, cles, o.x, o.pin, om_unused);
    o.env = mlp_normalise(L, o.x, o.xn);
}

// Collide-stream step with the learned closure (fast arithmetic, FP32 math).
// Phase 1 loads a cell and forms its features; phase 2 evaluates the network
// for the wave's 64 cells on the matrix cores (uniform control flow: every
// lane takes part, cells without a collision feed zeros); phase 3 collides
// with omega from the network's coefficient -- or from the Smagorinsky
// formula when a feature leaves the training envelope -- and streams.
template <class ST, int COLL, bool ODD, bool FLAGS>
__global__ __launch_bounds__(TPB) void k_step_mlp(MlpStepArgs<ST> ma)
{
    const StepArgs<ST>& a = ma.s;
    using R = typename ST::R;
    static_assert(sizeof(R) == 4, "the learned closure runs in FP32 arithmetic");
    __shared__ MlpLds L;
    mlp_load_lds(L, ma.mlp);
    __syncthreads();

    const Cell c = cell_of(a.g);
    bool active = c.inside;
    uint8_t cls = CL_FLUID;
    if (FLAGS && active) cls = a.cls[c.i];
    if (cls == CL_SOLID) active = false;
    const bool inlet = FLAGS && cls == CL_INLET;
    int bi = -1;
    R f[Q];
    ClosurePrep<ST> cp;
    float xn[NFEAT];
#pragma unroll
    for (int k = 0; k < NFEAT; ++k) xn[k] = 0.0f;
    cp.env = false;
    if (active) {
        if (FLAGS && cls >= CL_INLET) bi = a.bcidx[c.i];
        if (inlet) {
#pragma unroll
            for (int q = 0; q < Q; ++q) f[q] = a.inl_state[(size_t)bi * Q + q];
        } else {
            load_pop<ST, ODD>(a.p, a.rest, c, f);
        }
        closure_prep<ST>(f, a.first != 0, a.macro, (size_t)a.g.nx * a.g.ny * a.g.nz, c.i, a.tau, a.cles, L, cp);
        if (cp.env) {
#pragma unroll
            for (int k = 0; k < NFEAT; ++k) xn[k] = cp.xn[k];
        }
    }
    const float z = mlp_wave_forward(L, xn, ma.mlp.b3);
    if (active) {
        const R omega = cp.env ? omega_from_coefficient(mlp_coefficient(z, ma.mlp.c), cp.pin, a.tau) : cp.om_s;
        if (a.ccache) a.ccache[c.i] = cp.env ? mlp_coefficient(z, ma.mlp.c) : (float)a.cles;
        R P[Q];
        if (COLL == LBM_COLLISION_CUM2015) {
            cum2015_fast(a.K, cp.ce, cp.m0, cp.rho, cp.ux, cp.uy, cp.uz, omega, P);
        } else {
#pragma unroll
            for (int q = 0; q < Q; ++q) P[q] = f[q];
            srt_fast(a.K, P, cp.rho, cp.ux, cp.uy, cp.uz, omega);
        }
        // the streaming offsets are formed again here: kept from the loads of
        // an odd step, the 27 neighbour offsets held 24 registers through the
        // network
        Cell cs = c;
        __asm__ volatile("" : "+v"(cs.i), "+v"(cs.dxm), "+v"(cs.dxp), "+v"(cs.dym), "+v"(cs.dyp), "+v"(cs.dzm),
                         "+v"(cs.dzp));
        stream_cell<ST, ODD, FLAGS>(a, cs, cls, bi, inlet, P, cp.rho, cp.ux, cp.uy, cp.uz);
    }
}

// ---------------------------------------------------------------------------
// Boundary patch kernels (one thread per list entry), run after k_step on the
// state the next step will read.
// ---------------------------------------------------------------------------

template <class ST>
struct PatchArgs {
    typename ST::T* p[Q];
    const uint8_t* cls;
    const int32_t* bcidx;
    const NodeDev* nodes;          // this launch's entries
    int n;
    typename ST::R* inl_state;     // indexed like the inlet list
    const typename ST::R* pub_mac;
    const typename ST::R* pub_post;
    Geo g;
    Consts<typename ST::R> K;
    typename ST::R rest[4];
    typename ST::R nu, cs;
    int crev20;
    int inletrho;   // LBM_COMPAT_INLETRHO: the reference's nu/3 inlet extrapolation
    int odd;   // parity of the step that will read the state
};

// True when padded coordinate (x, y, z) lies on the caller's lattice
// (halo cells do not; the reference skips neighbours outside its arrays).
__device__ __forceinline__ bool on_lattice(const Geo& g, int x, int y, int z)
{
    return x >= g.px && x < g.px + g.ux && y >= g.py && y < g.py + g.uy && z >= g.pz && z < g.pz + g.uz;
}

__device__ __forceinline__ int wrap(int v, int n, int w) { return w ? (v + n) % n : v; }

template <class ST>
__device__ __forceinline__ void load_node(const PatchArgs<ST>& a, const Cell& c, typename ST::R* d, int* pl,
                                          uint32_t* ix)
{
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        read_slot(c, q, a.odd != 0, pl[q], ix[q]);
        d[q] = ST::load(a.p[pl[q]], ix[q], a.rest[WCLASS[q]]);
    }
}

template <class ST>
__device__ __forceinline__ void store_node(const PatchArgs<ST>& a, const typename ST::R* d, const int* pl,
                                           const uint32_t* ix)
{
#pragma unroll
    for (int q = 0; q < Q; ++q) ST::store(a.p[pl[q]], ix[q], d[q], a.rest[WCLASS[q]]);
}

// InletVelocityEquilibrium::inlet (InletVelocityEquilibrium.h:51-103):
// rho_b is extrapolated from rho and u_x of the first two interior
// neighbours as they entered this step's collision, and the inlet state
// becomes f_eq(u_b, rho_b).  Only +-x normals act, as in the reference.
// With duxb, dux1 the differences each branch of the reference forms,
// rho_b = rho_1 + 3 nu (duxb - dux1) = rho_1 + 3 nu (x_b - x_1) d2u/dx2 (the
// steady x-momentum balance d(rho/3)/dx = nu d2u/dx2); the reference's
// nu/3, and its sign on the +x branch (:80), under LBM_COMPAT_INLETRHO.
template <class ST>
__global__ __launch_bounds__(TPB) void k_inlet_eq(PatchArgs<ST> a)
{
    LBM_FP_STRICT
    using R = typename ST::R;
    const int j = blockIdx.x * TPB + threadIdx.x;
    if (j >= a.n) return;
    const NodeDev b = a.nodes[j];
    if (b.kind != LBM_BC_INLET_EQUILIBRIUM) return;
    const R uxb = b.ux, uyb = b.uy, uzb = b.uz;
    const R rho1 = b.s0 >= 0 ? a.pub_mac[(size_t)b.s0 * PUB_MAC] : R(0);
    const R ux1 = b.s0 >= 0 ? a.pub_mac[(size_t)b.s0 * PUB_MAC + 1] : R(0);
    const R ux2 = b.s1 >= 0 ? a.pub_mac[(size_t)b.s1 * PUB_MAC + 1] : R(0);
    R rhob;
    if (b.nx == 1.0f && b.ny == 0.0f && b.nz == 0.0f) {
        const R duxb = (uxb - ux1);
        const R dux1 = (ux1 - ux2);
        rhob = a.inletrho ? rho1 + a.nu / 3 * (dux1 - duxb) : rho1 + R(3) * a.nu * (duxb - dux1);
    } else if (b.nx == -1.0f && b.ny == 0.0f && b.nz == 0.0f) {
        const R duxb = (ux1 - uxb);
        const R dux1 = (ux2 - ux1);
        rhob = a.inletrho ? rho1 + a.nu / 3 * (duxb - dux1) : rho1 + R(3) * a.nu * (duxb - dux1);
    } else {
        return;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) a.inl_state[(size_t)j * Q + q] = feq_q(a.K, q, uxb, uyb, uzb, rhob);
}

// State update of moving-wall and Zou-He inlet cells (the rules of
// InletVelocityMovingWall.h:43-49 and InletVelocityZouHe.h:80-101, applied by
// moving_wall / zou_he) from the collision the step kernel handed over: the
// cell's post-collision populations (pub_post here) and its rho, u (pub_mac).
// Runs after the step kernel, which streamed the previous state.
template <class ST>
__global__ __launch_bounds__(TPB) void k_inlet_state(PatchArgs<ST> a)
{
    using R = typename ST::R;
    const int j = blockIdx.x * TPB + threadIdx.x;
    if (j >= a.n) return;
    const NodeDev b = a.nodes[j];
    if (b.kind != LBM_BC_INLET_MOVING_WALL && b.kind != LBM_BC_INLET_ZOU_HE) return;
    R P[Q], st[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        P[q] = a.pub_post[(size_t)j * Q + q];
        st[q] = a.inl_state[(size_t)j * Q + q];
    }
    const R* m = a.pub_mac + (size_t)j * PUB_MAC;
    if (b.kind == LBM_BC_INLET_MOVING_WALL)
        moving_wall(a.K, b, a.crev20 != 0, P, m[0], st);
    else
        zou_he(a.K, b, a.crev20 != 0, P, m[0], m[1], m[2], m[3], st);
#pragma unroll
    for (int q = 0; q < Q; ++q) a.inl_state[(size_t)j * Q + q] = st[q];
}

// OutletDensityInterpolated::outlet (OutletDensityInterpolated.h:170-252):
// irregular +x nodes first copy f_crev into f_q where the cell at
// (x, y - c_y, z - c_z) is solid (sequentially in q), then incoming
// populations are interpolated between the node's and the upstream cell's
// post-collision values and the density is reset to the outlet density.
// OutletDensityEquilibrium (OutletDensityEquilibrium.h:55-76) and the
// neighbour-equilibrium outlets (OutletNeighbourEquilibrium.h:78-126,
// ...OmegaRF.h:143-198) on one list entry.  rho, u of a cell are the moments
// its collision used in this step (published by the step kernel), which is
// what the reference's rho/u fields hold at this point.  b.partner holds the
// mask of incoming directions the rule acts on (host-computed: n.c < 0, the
// neighbour x + c inside the lattice and, at irregular nodes, flagged > 0).
template <class ST>
__device__ __forceinline__ void outlet_equilibrium(const PatchArgs<ST>& a, const NodeDev& b)
{
    LBM_FP_STRICT
    using R = typename ST::R;
    const Cell c = cell_at(a.g, b.x, b.y, b.z);
    R d[Q];
    int pl[Q];
    uint32_t ix[Q];
    load_node(a, c, d, pl, ix);
    const R* mx = a.pub_mac + (size_t)b.s0 * PUB_MAC;
    const R ux = mx[1], uy = mx[2], uz = mx[3];
    R density = b.rho;
    if (b.kind == LBM_BC_OUTLET_EQUILIBRIUM) {
#pragma unroll
        for (int vel = 0; vel < Q; ++vel) {
            const float nc = b.nx * (float)CX[vel] + b.ny * (float)CY[vel] + b.nz * (float)CZ[vel];
            if (nc > 0.0f) d[crev_of(vel, a.crev20 != 0)] = feq_q(a.K, vel, ux, uy, uz, density);
        }
    } else {
        // the RF variant's backflow correction is written with a double literal 0.5
        if (b.kind == LBM_BC_OUTLET_NEIGHBOUR_EQ_RF && ux < R(0))
            density = (R)((double)density - 0.5 * (double)ux * (double)ux * (double)density);
        const uint32_t mask = (uint32_t)b.partner;
#pragma unroll
        for (int vel = 0; vel < Q; ++vel) {
            if (!((mask >> vel) & 1u)) continue;
            const int xn = wrap(b.x + CX[vel], a.g.nx, a.g.wx), yn = wrap(b.y + CY[vel], a.g.ny, a.g.wy),
                      zn = wrap(b.z + CZ[vel], a.g.nz, a.g.wz);
            const Cell nc = cell_at(a.g, xn, yn, zn);
            R fEqNeigh = R(0), fNeigh = R(0);   // a solid neighbour: rho = 0, f = 0 in the reference
            if (a.cls[nc.i] != CL_SOLID) {
                const R* mn = a.pub_mac + (size_t)a.bcidx[nc.i] * PUB_MAC;
                fEqNeigh = feq_q(a.K, vel, mn[1], mn[2], mn[3], mn[0]);
                int pn;
                uint32_t in;
                read_slot(nc, vel, a.odd != 0, pn, in);
                fNeigh = ST::load(a.p[pn], in, a.rest[WCLASS[vel]]);
            }
            const R fEqBC = feq_q(a.K, vel, ux, uy, uz, density);
            d[vel] = fEqBC + (fEqNeigh - fNeigh);
        }
    }
    store_node(a, d, pl, ix);
}

template <class ST>
__global__ __launch_bounds__(TPB) void k_outlet(PatchArgs<ST> a)
{
    LBM_FP_STRICT
    using R = typename ST::R;
    const int j = blockIdx.x * TPB + threadIdx.x;
    if (j >= a.n) return;
    const NodeDev b = a.nodes[j];
    if (b.kind != LBM_BC_OUTLET_INTERPOLATED) {
        outlet_equilibrium(a, b);
        return;
    }
    const Cell c = cell_at(a.g, b.x, b.y, b.z);
    R d[Q];
    int pl[Q];
    uint32_t ix[Q];
    load_node(a, c, d, pl, ix);
    interp_outlet(d, b, a.pub_post, a.K, a.cs, a.crev20 != 0);
    store_node(a, d, pl, ix);
}

// Periodic::periodic (Periodic.h:43-138): populations with n.c < 0 are
// copied from the partner cell (the node with its normal-axis index replaced
// by the periodic index); irregular nodes also bounce their own
// post-collision populations back off solid neighbours, sequentially in q.
// Entries that depend on each other run in separate launches, in list order.
template <class ST>
__global__ __launch_bounds__(TPB) void k_periodic(PatchArgs<ST> a)
{
    using R = typename ST::R;
    const int j = blockIdx.x * TPB + threadIdx.x;
    if (j >= a.n) return;
    const NodeDev b = a.nodes[j];
    const Cell c = cell_at(a.g, b.x, b.y, b.z);
    R d[Q];
    int pl[Q];
    uint32_t ix[Q];
    load_node(a, c, d, pl, ix);
    const uint32_t pi = (uint32_t)b.partner;
    const bool self_partner = pi == c.i;
    const uint8_t pcls = a.cls[pi];
    const int px = (int)(pi % (uint32_t)a.g.nx);
    const int py = (int)((pi / (uint32_t)a.g.nx) % (uint32_t)a.g.ny);
    const int pz = (int)(pi / ((uint32_t)a.g.nx * (uint32_t)a.g.ny));
    const Cell pc = cell_at(a.g, px, py, pz);
    const R* post = a.pub_post + (size_t)b.s0 * Q;
#pragma unroll
    for (int vel = 0; vel < Q; ++vel) {
        const float nc = b.nx * (float)CX[vel] + b.ny * (float)CY[vel] + b.nz * (float)CZ[vel];
        if (nc < 0.0f && !self_partner) {
            R v;
            if (pcls == CL_SOLID) {
                v = R(0);
            } else if (pcls == CL_INLET) {
                v = a.inl_state[(size_t)a.bcidx[pi] * Q + vel];
            } else {
                int ppl;
                uint32_t pix;
                read_slot(pc, vel, a.odd != 0, ppl, pix);
                v = ST::load(a.p[ppl], pix, a.rest[WCLASS[vel]]);
            }
            d[vel] = v;
        }
        if (!b.regular && nc <= 0.0f) {
            const int xx = wrap(b.x + CX[vel], a.g.nx, a.g.wx);
            const int yy = wrap(b.y + CY[vel], a.g.ny, a.g.wy);
            const int zz = wrap(b.z + CZ[vel], a.g.nz, a.g.wz);
            if (on_lattice(a.g, xx, yy, zz) && a.cls[(uint32_t)(xx + a.g.nx * (yy + a.g.ny * zz))] == CL_SOLID)
                d[crev_of(vel, a.crev20 != 0)] = post[vel];
        }
    }
    store_node(a, d, pl, ix);
}

// ---------------------------------------------------------------------------
// Auxiliary kernels: initialisation, gather/scatter, moments, omega, sums.
// These iterate over the caller's cells (no halo); io is in caller layout.
// ---------------------------------------------------------------------------

template <class ST>
struct AuxArgs {
    typename ST::T* p[Q];
    const uint8_t* cls;
    const int32_t* bcidx;
    typename ST::R* inl_state;
    typename ST::R* macro;   // rho, ux, uy, uz planes (padded layout)
    Geo g;
    Consts<typename ST::R> K;
    typename ST::R rest[4];
    typename ST::R tau, cles, omega0;
    int closure;
    int lit;
    int odd;        // parity of the stored state
    int first;      // no step taken since lbm_init_equilibrium: rho, u are the macro planes
    int q;          // plane for gather/scatter
    double* io;     // staging buffer (FP64, caller layout)
    int sp_axis, sp_begin;      // outlet omega damping (as StepArgs)
    typename ST::R sp_deriv;
    int outl_pending;           // deferred outlet: CL_OUTL cells still need the rule
    const NodeDev* outl;
    const typename ST::R* pub_prev;   // records of the step that produced the state
    typename ST::R cs;
    int crev20;
    int inl_pending;            // deferred inlet: inlet cells' stored state still needs its rule
    const NodeDev* inl;
    const typename ST::R* inl_post;
    const typename ST::R* inl_mac;
};

// Thread -> caller cell -> padded cell.
__device__ __forceinline__ bool user_cell(const Geo& g, uint32_t& iu, Cell& c)
{
    const size_t NU = (size_t)g.ux * g.uy * g.uz;
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= NU) return false;
    iu = (uint32_t)i;
    const int x = (int)(iu % (uint32_t)g.ux);
    const uint32_t r = iu / (uint32_t)g.ux;
    const int y = (int)(r % (uint32_t)g.uy);
    const int z = (int)(r / (uint32_t)g.uy);
    c = cell_at(g, x + g.px, y + g.py, z + g.pz);
    return true;
}

// Current populations of a non-solid cell (a deferred outlet cell's with its
// pending rule applied).
template <class ST>
__device__ __forceinline__ void cell_state(const AuxArgs<ST>& a, const Cell& c, uint8_t cls, typename ST::R* f)
{
    if (cls == CL_INLET) {
        const int bi = a.bcidx[c.i];
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = a.inl_state[(size_t)bi * Q + q];
        if (a.inl_pending)
            inlet_rule(a.K, a.inl[bi], a.inl_post + (size_t)bi * Q, a.inl_mac + (size_t)bi * PUB_MAC,
                       a.crev20 != 0, f);
        return;
    }
    if (a.odd)
        load_pop<ST, true>(a.p, a.rest, c, f);
    else
        load_pop<ST, false>(a.p, a.rest, c, f);
    if (cls == CL_OUTL && a.outl_pending)
        deferred_outlet<ST>(f, a.outl[a.bcidx[c.i]], a.pub_prev, a.K, a.cs, a.crev20 != 0, a.rest);
}

// f = f_eq(rho, u) from the macro planes (InitializationEquilibriumConstVector.h:111-126);
// runs over the padded lattice.
template <class ST>
__global__ __launch_bounds__(TPB) void k_init(AuxArgs<ST> a)
{
    using R = typename ST::R;
    const size_t N = (size_t)a.g.nx * a.g.ny * a.g.nz;
    const size_t ii = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (ii >= N) return;
    const uint32_t i = (uint32_t)ii;
    const uint8_t cls = a.cls[i];
    const R rho = a.macro[i], ux = a.macro[N + i], uy = a.macro[2 * N + i], uz = a.macro[3 * N + i];
    const bool inlet = cls == CL_INLET;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const R v = cls == CL_SOLID ? R(0) : feq_q(a.K, q, ux, uy, uz, rho);
        ST::store(a.p[q], i, v, a.rest[WCLASS[q]]);
        if (inlet) a.inl_state[(size_t)a.bcidx[i] * Q + q] = v;
    }
}

// Write plane q of a host-provided state (FP64 staging) at the current parity.
template <class ST>
__global__ __launch_bounds__(TPB) void k_scatter(AuxArgs<ST> a)
{
    using R = typename ST::R;
    uint32_t iu;
    Cell c;
    if (!user_cell(a.g, iu, c)) return;
    const uint8_t cls = a.cls[c.i];
    if (cls == CL_SOLID) return;
    const int q = a.q;
    const R v = (R)a.io[iu];
    if (cls == CL_INLET) {
        a.inl_state[(size_t)a.bcidx[c.i] * Q + q] = v;
        return;
    }
    int pl;
    uint32_t ix;
    read_slot(c, q, a.odd != 0, pl, ix);
    ST::store(a.p[pl], ix, v, a.rest[WCLASS[q]]);
}

// Gather plane q of the current state into FP64 staging (0 on solids).
template <class ST>
__global__ __launch_bounds__(TPB) void k_gather(AuxArgs<ST> a)
{
    uint32_t iu;
    Cell c;
    if (!user_cell(a.g, iu, c)) return;
    const uint8_t cls = a.cls[c.i];
    if (cls == CL_SOLID) {
        a.io[iu] = 0.0;
        return;
    }
    const int q = a.q;
    if ((cls == CL_INLET && a.inl_pending) || (cls == CL_OUTL && a.outl_pending)) {
        typename ST::R f[Q];
        cell_state(a, c, cls, f);
        a.io[iu] = (double)f[q];
        return;
    }
    if (cls == CL_INLET) {
        a.io[iu] = (double)a.inl_state[(size_t)a.bcidx[c.i] * Q + q];
        return;
    }
    int pl;
    uint32_t ix;
    read_slot(c, q, a.odd != 0, pl, ix);
    a.io[iu] = (double)ST::load(a.p[pl], ix, a.rest[WCLASS[q]]);
}

// rho, u (and omega when om != nullptr) of a non-solid cell of the current state.
template <class ST>
__device__ __forceinline__ void cell_macro(const AuxArgs<ST>& a, const Cell& c, uint8_t cls, typename ST::R& rho,
                                           typename ST::R& ux, typename ST::R& uy, typename ST::R& uz,
                                           typename ST::R* om)
{
    using R = typename ST::R;
    R f[Q];
    cell_state(a, c, cls, f);
    const size_t N = (size_t)a.g.nx * a.g.ny * a.g.nz;
    if (a.lit) {
        if (a.first) {
            rho = a.macro[c.i];
            ux = a.macro[N + c.i];
            uy = a.macro[2 * N + c.i];
            uz = a.macro[3 * N + c.i];
        } else {
            moments_literal(f, rho, ux, uy, uz);
        }
        if (om)
            *om = sponge(a.closure == LBM_CLOSURE_SMAGORINSKY ? omega_les_literal(a.K, f, rho, ux, uy, uz, a.tau, a.cles)
                                                              : a.omega0,
                         a.sp_axis, a.sp_begin, a.sp_deriv, a.g, c.x, c.y, c.z);
    } else {
        const Raw2<R> r = raw_moments(f);
        if (a.first) {
            rho = a.macro[c.i];
            ux = a.macro[N + c.i];
            uy = a.macro[2 * N + c.i];
            uz = a.macro[3 * N + c.i];
        } else {
            rho = r.m0;
            ux = r.jx / rho;
            uy = r.jy / rho;
            uz = r.jz / rho;
        }
        if (om) {
            // the rate from the velocity as the step forms it (reciprocal, then products)
            R vx = ux, vy = uy, vz = uz;
            if (!a.first) {
                const R ir = fast_rcp(rho);
                vx = r.jx * ir;
                vy = r.jy * ir;
                vz = r.jz * ir;
            }
            const Cen2<R> ce = central2(r, vx, vy, vz);
            *om = sponge(a.closure == LBM_CLOSURE_SMAGORINSKY ? omega_les_fast(ce, rho, a.tau, a.cles) : a.omega0,
                         a.sp_axis, a.sp_begin, a.sp_deriv, a.g, c.x, c.y, c.z);
        }
    }
}

// rho, u and omega of the current state into 5 FP64 staging planes.
template <class ST>
__global__ __launch_bounds__(TPB) void k_macro(AuxArgs<ST> a)
{
    using R = typename ST::R;
    uint32_t iu;
    Cell c;
    if (!user_cell(a.g, iu, c)) return;
    const size_t NU = (size_t)a.g.ux * a.g.uy * a.g.uz;
    const uint8_t cls = a.cls[c.i];
    if (cls == CL_SOLID) {
        for (int k = 0; k < 5; ++k) a.io[k * NU + iu] = 0.0;
        return;
    }
    R rho, ux, uy, uz, om;
    cell_macro(a, c, cls, rho, ux, uy, uz, &om);
    a.io[iu] = (double)rho;
    a.io[NU + iu] = (double)ux;
    a.io[2 * NU + iu] = (double)uy;
    a.io[3 * NU + iu] = (double)uz;
    a.io[4 * NU + iu] = (double)om;
}

// MomentTimeAvg::momentAdd as its own pass: rho, u of the current state (the
// values k_macro returns) added to the sums in a.io.  A step launched with
// StepArgs::avg does the same inside the step instead.
template <class ST>
__global__ __launch_bounds__(TPB) void k_moment_add(AuxArgs<ST> a)
{
    using R = typename ST::R;
    uint32_t iu;
    Cell c;
    if (!user_cell(a.g, iu, c)) return;
    const uint8_t cls = a.cls[c.i];
    if (cls == CL_SOLID) return;
    R rho, ux, uy, uz;
    cell_macro(a, c, cls, rho, ux, uy, uz, (R*)nullptr);
    moment_add(a.io, (size_t)a.g.ux * a.g.uy * a.g.uz, iu, (double)rho, (double)ux, (double)uy, (double)uz);
}

// rho, u of a single cell (probe), caller coordinates, as k_macro forms them;
// written to a.io[0..3].
template <class ST>
__global__ void k_probe(AuxArgs<ST> a, int x, int y, int z)
{
    using R = typename ST::R;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const Cell c = cell_at(a.g, x + a.g.px, y + a.g.py, z + a.g.pz);
    const uint8_t cls = a.cls[c.i];
    if (cls == CL_SOLID) {
        for (int k = 0; k < 4; ++k) a.io[k] = 0.0;
        return;
    }
    R rho, ux, uy, uz;
    cell_macro(a, c, cls, rho, ux, uy, uz, (R*)nullptr);
    a.io[0] = (double)rho;
    a.io[1] = (double)ux;
    a.io[2] = (double)uy;
    a.io[3] = (double)uz;
}

// ErrorQuadratic::errorEvaluation (ErrorQuadratic.h:33-90) of the current
// state against the velocity uerr (3 planes, caller layout, the arithmetic
// type like the reference's u_error field): per block, the FP64 sums of
// sqrt(|u - u_prev|^2) and sqrt(|u|^2) over cells whose flag is 1 (the third
// difference is u_y - u_prev,z, ErrorQuadratic.h:43-44), written to
// a.io[2 block + {0, 1}]; then u_prev = u on every cell (0 on solid cells).
// u is what lbm_get_macroscopic returns.
template <class ST>
__global__ __launch_bounds__(TPB) void k_error(AuxArgs<ST> a, typename ST::R* uerr, const uint8_t* flag1)
{
    LBM_FP_STRICT
    using R = typename ST::R;
    uint32_t iu;
    Cell c;
    double sm[2] = {0.0, 0.0};
    if (user_cell(a.g, iu, c)) {
        const size_t NU = (size_t)a.g.ux * a.g.uy * a.g.uz;
        const uint8_t cls = a.cls[c.i];
        R rho = R(0), ux = R(0), uy = R(0), uz = R(0);
        if (cls != CL_SOLID) cell_macro(a, c, cls, rho, ux, uy, uz, (R*)nullptr);
        if (flag1[iu]) {
            const R ex = uerr[iu], ey = uerr[NU + iu], ez = uerr[2 * NU + iu];
            const R add1 = (ux - ex) * (ux - ex) + (uy - ey) * (uy - ey) + (uy - ez) * (uy - ez);
            const R add2 = ux * ux + uy * uy + uz * uz;
            sm[0] = (double)sqrt(add1);
            sm[1] = (double)sqrt(add2);
        }
        uerr[iu] = ux;
        uerr[NU + iu] = uy;
        uerr[2 * NU + iu] = uz;
    }
    __shared__ double red[TPB / 64][2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        double v = sm[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        sm[k] = v;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[wv][0] = sm[0];
        red[wv][1] = sm[1];
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double s = 0.0;
        for (int w = 0; w < TPB / 64; ++w) s += red[w][threadIdx.x];
        a.io[(size_t)blockIdx.x * 2 + threadIdx.x] = s;
    }
}

// Sum of nb per-block partial records (stride doubles apart, ncomp <= 8
// components each) into out[0 .. ncomp), by one workgroup in a fixed order:
// thread t adds records t, t + 256, ... in turn, then a tree over the
// threads.  Deterministic for a given nb.
__global__ __launch_bounds__(TPB) void k_reduce_partials(const double* parts, int nb, int stride, int ncomp,
                                                         double* out)
{
    __shared__ double red[TPB][8];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = (int)threadIdx.x; b < nb; b += TPB)
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < ncomp) acc[k] += parts[(size_t)b * stride + k];
#pragma unroll
    for (int k = 0; k < 8; ++k) red[threadIdx.x][k] = acc[k];
    __syncthreads();
    for (int h = TPB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int k = 0; k < 8; ++k) red[threadIdx.x][k] += red[threadIdx.x + h][k];
        __syncthreads();
    }
    if ((int)threadIdx.x < ncomp) out[threadIdx.x] = red[0][threadIdx.x];
}

// Per-block FP64 sums of mass and momentum over non-solid cells.
template <class ST>
__global__ __launch_bounds__(TPB) void k_sums(AuxArgs<ST> a)
{
    using R = typename ST::R;
    uint32_t iu;
    Cell c;
    double dm[4] = {0.0, 0.0, 0.0, 0.0};
    if (user_cell(a.g, iu, c) && a.cls[c.i] != CL_SOLID) {
        R f[Q];
        cell_state(a, c, a.cls[c.i], f);
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

template <class ST>
struct ClosureEvalArgs {
    AuxArgs<ST> a;
    MlpWeights mlp;
    float* feats;     // [12][NU] normalised features, or null
    int32_t* env;     // [NU], or null
    double* coef;     // [NU] coefficient the next step uses, or null
    double* omega;    // [NU] rate the next step uses, or null
    const float* ccache;  // held coefficients when the next step does not refresh, or null
};

// Closure inputs and outputs of the current state, per caller cell; the
// same per-cell code as k_step_mlp.
template <class ST>
__global__ __launch_bounds__(TPB) void k_closure_eval(ClosureEvalArgs<ST> e)
{
    using R = typename ST::R;
    __shared__ MlpLds L;
    mlp_load_lds(L, e.mlp);
    __syncthreads();
    const AuxArgs<ST>& a = e.a;
    uint32_t iu = 0;
    Cell c;
    const bool have = user_cell(a.g, iu, c);
    const bool active = have && a.cls[c.i] != CL_SOLID;
    ClosurePrep<ST> cp;
    cp.env = false;
    float xn[NFEAT];
#pragma unroll
    for (int k = 0; k < NFEAT; ++k) xn[k] = 0.0f;
    if (active) {
        R f[Q];
        cell_state(a, c, a.cls[c.i], f);
        closure_prep<ST>(f, a.first != 0, a.macro, (size_t)a.g.nx * a.g.ny * a.g.nz, c.i, a.tau, a.cles, L, cp);
        if (cp.env) {
#pragma unroll
            for (int k = 0; k < NFEAT; ++k) xn[k] = cp.xn[k];
        }
    }
    const float z = mlp_wave_forward(L, xn, e.mlp.b3);
    if (!have) return;
    const size_t NU = (size_t)a.g.ux * a.g.uy * a.g.uz;
    if (e.feats)
        for (int k = 0; k < NFEAT; ++k) e.feats[k * NU + iu] = active ? cp.xn[k] : 0.0f;
    if (e.env) e.env[iu] = (active && cp.env) ? 1 : 0;
    const float C = e.ccache ? e.ccache[c.i] : (cp.env ? mlp_coefficient(z, e.mlp.c) : (float)a.cles);
    if (e.coef) e.coef[iu] = active ? (double)C : 0.0;
    if (e.omega) {
        const R om = e.ccache ? omega_from_coefficient((R)C, pin_over_rho(cp.ce, cp.rho), a.tau)
                              : (cp.env ? omega_from_coefficient(C, cp.pin, a.tau) : cp.om_s);
        e.omega[iu] = active ? (double)om : 0.0;
    }
}

// Box filter over fct^3 blocks of caller cells: averages of rho, u and u u
// over the non-solid cells of each block, 10 FP64 planes in a.io.
template <class ST>
__global__ __launch_bounds__(TPB) void k_filter(AuxArgs<ST> a, int fct, int cx, int cy, int cz)
{
    using R = typename ST::R;
    const size_t NC = (size_t)cx * cy * cz;
    const size_t ic = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (ic >= NC) return;
    const int X = (int)(ic % (size_t)cx), Y = (int)((ic / (size_t)cx) % (size_t)cy), Z = (int)(ic / ((size_t)cx * cy));
    double acc[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) acc[k] = 0.0;
    int cnt = 0;
    for (int dz = 0; dz < fct; ++dz)
        for (int dy = 0; dy < fct; ++dy)
            for (int dx = 0; dx < fct; ++dx) {
                const Cell c = cell_at(a.g, X * fct + dx + a.g.px, Y * fct + dy + a.g.py, Z * fct + dz + a.g.pz);
                const uint8_t cls = a.cls[c.i];
                if (cls == CL_SOLID) continue;
                R rho, ux, uy, uz;
                if (a.first) {
                    const size_t N = (size_t)a.g.nx * a.g.ny * a.g.nz;
                    rho = a.macro[c.i];
                    ux = a.macro[N + c.i];
                    uy = a.macro[2 * N + c.i];
                    uz = a.macro[3 * N + c.i];
                } else {
                    R f[Q];
                    cell_state(a, c, cls, f);
                    moments_literal(f, rho, ux, uy, uz);
                }
                const double u = ux, v = uy, w = uz;
                acc[0] += (double)rho;
                acc[1] += u;
                acc[2] += v;
                acc[3] += w;
                acc[4] += u * u;
                acc[5] += v * v;
                acc[6] += w * w;
                acc[7] += u * v;
                acc[8] += u * w;
                acc[9] += v * w;
                ++cnt;
            }
    const double inv = cnt ? 1.0 / cnt : 0.0;
#pragma unroll
    for (int k = 0; k < 10; ++k) a.io[k * NC + ic] = acc[k] * inv;
}

// Plain stream copy used as the bandwidth reference.
__global__ __launch_bounds__(TPB) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace lbm
