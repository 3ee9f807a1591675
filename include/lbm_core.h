/*
 * lbm_core.h -- C ABI of the MI355X lattice-Boltzmann collide-stream core.
 *
 * One context owns one D3Q27 lattice in HBM (single in-place A-A population
 * array, structure-of-arrays [q][z][y][x]) plus one HIP stream.  The caller
 * owns every host pointer; the library copies what it needs.  No exception
 * crosses this boundary; every entry point returns an lbm_status.  A context
 * is not thread-safe: drive it from one host thread.
 *
 * The per-step operator sequence replaced by lbm_step() is the reference's
 *   Turbulence::omega -> Collision -> StreamingAB -> BounceBackWallHalf
 *   -> Inlet -> Outlet -> Periodic -> MomentDensityVelocityN27
 * as called from SolverTurbulentLES::runSimulation
 * (3D/simulation/src/solvers/SolverTurbulentLES.h:184-214) and
 * SolverLaminar::runSimulation (3D/simulation/src/solvers/SolverLaminar.h:132-157).
 * Each reference policy is a `static void op(LBMDataPointer&, LBMConstantsPointer&)`
 * (e.g. CollisionCumD3Q27Turbulent2015.h:19, StreamingAB.h:19,
 * BounceBackWallHalf.h:19, OmegaLES.h:25, MomentDensityVelocityN27.h:19); here
 * the whole sequence is one fused kernel per time step.
 */
#ifndef LBM_CORE_H
#define LBM_CORE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_Q 27

typedef struct lbm_ctx lbm_ctx;

typedef enum lbm_status {
    LBM_OK = 0,
    LBM_E_ARG = 1,          /* invalid argument / inconsistent flag grid */
    LBM_E_HIP = 2,          /* HIP runtime error (no device, launch failure) */
    LBM_E_OOM = 3,          /* device allocation failed */
    LBM_E_NAN = 4,          /* non-finite mass or momentum detected */
    LBM_E_STATE = 5,        /* call out of order (e.g. step before init) */
    LBM_E_UNSUPPORTED = 6   /* valid request that this build does not implement */
} lbm_status;

/* Collision operator (reference: collisions/CollisionSRT*.h,
 * CollisionCumD3Q27Turbulent2015.h, CollisionCumD3Q27Turbulent2017.h,
 * CollisionCumD3Q27TurbulentCombined.h).  CUM2017 is the parametrized cumulant
 * (omega_3..5 from omega_1 with magnitude limiters, fourth-order A/B terms);
 * COMBINED collides cells whose flag is 1 (fluid, no solid neighbour) with the
 * 2015 operator and every other non-solid cell with the 2017 one
 * (Combined.h:1066-1078).  The learned closure runs with SRT and CUM2015. */
enum { LBM_COLLISION_SRT = 0, LBM_COLLISION_CUM2015 = 1, LBM_COLLISION_CUM2017 = 2, LBM_COLLISION_COMBINED = 3 };

/* Relaxation-rate closure (reference: turbulence/OmegaNo.h, turbulence/OmegaLES.h).
 * LBM_CLOSURE_MLP is the learned sub-grid closure (no reference counterpart). */
enum { LBM_CLOSURE_NONE = 0, LBM_CLOSURE_SMAGORINSKY = 1, LBM_CLOSURE_MLP = 2 };

/* Population storage in HBM.  F16DEV/BF16DEV store g_i = f_i - w_i*rho0 and do
 * all collision arithmetic in FP32, rounding g to nearest.  The *_SR formats
 * have the same layout; the collide-stream step stores g with stochastic
 * rounding (unbiased: E[stored] = g; random bits from a hash of the cell and
 * the step counter, so a run is reproducible).  Other writers (initialisation,
 * lbm_set_populations, boundary patches) round to nearest. */
enum {
    LBM_STORAGE_F64 = 0,
    LBM_STORAGE_F32 = 1,
    LBM_STORAGE_F16DEV = 2,
    LBM_STORAGE_BF16DEV = 3,
    LBM_STORAGE_F16DEV_SR = 4,
    LBM_STORAGE_BF16DEV_SR = 5
};

/* Collision arithmetic.  LITERAL follows the reference's operation order
 * (three-pass central-moment transform, full cumulant set); FAST uses the
 * algebraically equivalent reduced form (omega_2..omega_10 = 1 makes every
 * post-collision cumulant of order >= 3 vanish). */
enum { LBM_ARITH_LITERAL = 0, LBM_ARITH_FAST = 1 };

/* Compatibility switches reproducing reference quirks. */
#define LBM_COMPAT_CREV20    0x1u  /* D3Q27.h:120-124 c_rev[20] = 21 (correct: 19)          */
#define LBM_COMPAT_F32CONST  0x2u  /* D3Q27.h:113-118 float weights; Cum2015.h:280-282 1.f/3.f;
                                      Cum2017.h:244-246, 349-362 0.01f, 2.f/3.f, 4.f/3.f    */
#define LBM_COMPAT_LIM210    0x4u  /* Cum2017.h:254-255, 264-265 the 210/012 limiters read
                                      |C_120 +- C_102| (correct: |C_210 +- C_012|)          */
#define LBM_COMPAT_TRACE17   0x8u  /* Cum2017.h:295 subtracts (rho - 1), Combined.h:798 (rho + 1)
                                      from the trace in D_x u (correct: k_000, as Cum2015.h:253) */
#define LBM_COMPAT_INLETRHO  0x10u /* InletVelocityEquilibrium.h:80, 95 extrapolate the inlet density
                                      with nu/3 where the steady x-momentum balance
                                      d(rho/3)/dx = nu d2u/dx2 gives 3 nu, and the +x branch
                                      subtracts the second difference its -x branch adds
                                      (correct: rho_b = rho_1 + 3 nu (x_b - x_1) d2u/dx2)      */
#define LBM_COMPAT_EQ44      0x20u /* Cum2017.h:351, Combined.h:351 pair C_220 + C_202 - 2 C_022 with
                                      D_x u + D_y v - 2 D_z w; the axis convention of Eq. 43 (and
                                      symmetry under x <-> z) gives -2 D_x u + D_y v + D_z w.  The
                                      reference's form exchanges the A terms of C_220 and C_022 */
#define LBM_COMPAT_OMEGA4    0x40u /* Cum2017.h:259-260, Combined.h:259-260 omega_4's denominator
                                      omega_2 (56 - 42 omega_1 + 8 omega_1^2) - 8 omega_1, where
                                      the published rate (Geier, Pasquali, Schoenherr 2017) has
                                      9 omega_1^2: with 8 the 2017 operator's viscosity is
                                      second-order accurate, with 9 fourth-order */
#define LBM_COMPAT_REFERENCE                                                                       \
    (LBM_COMPAT_CREV20 | LBM_COMPAT_F32CONST | LBM_COMPAT_LIM210 | LBM_COMPAT_TRACE17 | LBM_COMPAT_INLETRHO | \
     LBM_COMPAT_EQ44 | LBM_COMPAT_OMEGA4)

/* Boundary-node kinds.  The flag grid keeps the reference encoding
 * (geometryMesherBoundary.h): 0 solid, >0 fluid (1 + number of solid
 * neighbours), -1 inlet, -2 outlet, -3 periodic ghost. */
enum {
    LBM_BC_WALL = 0,                 /* BounceBackWallHalf.h:29-52, wall list        */
    LBM_BC_INLET_MOVING_WALL = 1,    /* InletVelocityMovingWall.h:31-53              */
    LBM_BC_INLET_EQUILIBRIUM = 2,    /* InletVelocityEquilibrium.h:51-103            */
    LBM_BC_OUTLET_INTERPOLATED = 3,  /* OutletDensityInterpolated.h:170-252          */
    LBM_BC_PERIODIC_GHOST = 4,       /* Periodic.h:43-138                            */
    LBM_BC_OUTLET_EQUILIBRIUM = 5,   /* OutletDensityEquilibrium.h:55-76             */
    LBM_BC_OUTLET_NEIGHBOUR_EQ = 6,  /* OutletNeighbourEquilibrium.h:78-126          */
    LBM_BC_OUTLET_NEIGHBOUR_EQ_RF = 7, /* OutletNeighbourEquilibriumOmegaRF.h:143-198 */
    LBM_BC_INLET_ZOU_HE = 8          /* InletVelocityZouHe.h:80-101                  */
};

/* One boundary-list entry; mirrors the reference PODs boundaryConditionWall /
 * Inlet / Outlet / Periodic (traits/LBMTraits.h:17-61).  Normals, velocities
 * and densities are float there, so they are float here. */
typedef struct lbm_bc_node {
    int32_t x, y, z;
    int32_t kind;          /* LBM_BC_* */
    float nx, ny, nz;      /* outward normal */
    float ux, uy, uz;      /* inlet velocity, lattice units */
    float rho;             /* outlet target density, lattice units */
    int32_t per_idx;       /* periodic partner index along the normal axis */
    int32_t regular;       /* reference 'regular' bit */
} lbm_bc_node;

typedef struct lbm_desc {
    int32_t nx, ny, nz;        /* lattice size including any solid halo */
    int32_t periodic[3];       /* 1 = index-wrap periodicity along x/y/z */
    int32_t collision;         /* LBM_COLLISION_* */
    int32_t closure;           /* LBM_CLOSURE_* */
    int32_t storage;           /* LBM_STORAGE_* */
    int32_t arith;             /* LBM_ARITH_* */
    uint32_t compat;           /* LBM_COMPAT_* bits */
    int32_t n_bc;              /* entries in bc[] */
    double tau;                /* molecular relaxation time, omega = 1/tau (LBMConstants::tau) */
    double nu;                 /* lattice viscosity used by the equilibrium inlet (LBMConstants::ny) */
    double c_les;              /* Smagorinsky constant in the reference formula (LBMConstants::CLES) */
    double rho0;               /* reference density of the deviation storage */
    const int32_t* flags;      /* host, reference encoding, x fastest, nx*ny*nz; copied */
    const lbm_bc_node* bc;     /* host, copied; may be NULL when n_bc == 0 */
    /* Outlet omega damping (OutletNeighbourEquilibriumOmega.h:18-79, used by the
     * ...Omega and ...OmegaRF outlets): omega -= (high - low) / m * p over the last
     * m = ceil(0.1 n) cells of the axis of the first outlet entry's normal, p = 0..m-1,
     * at the high end of that axis.  Needs an axis-aligned normal. */
    int32_t outlet_sponge;     /* 1 = on */
    float sponge_low, sponge_high;   /* LBMConstants::omegaDumpingLow / High */
    /* Wall policy.  0: BounceBackWallHalf (Wall/BounceBackWallHalf.h:29-52), half-way
     * bounce-back on the nodes of the wall list (LBM_BC_WALL entries; the mesher
     * lists cells with flag > 2).  1: BounceBackWallHalfMesh
     * (Wall/BounceBackWallHalfMesh.h:38-50), on every cell whose flag is exactly 2
     * (one solid neighbour), whatever the wall list holds.  Other fluid cells keep
     * the populations their solid neighbours would send (StreamingAB.h:58-65). */
    int32_t wall_mesh;
} lbm_desc;

/* Create a context on the current HIP device.  Validates the flag grid and the
 * boundary lists.  Neighbours outside the array on an axis without index wrap
 * are treated as the reference's out-of-array reads are restated: populations
 * streaming from there are left untouched, streaming there is dropped (the
 * core pads a solid halo for their slots; StreamingAB.h:51-56 has no
 * bounds checks).  A wall-list node with such a neighbour is
 * LBM_E_UNSUPPORTED, as are flags below -3, outlet or ghost entries on inlet
 * cells, and inlet/outlet rules that would read an inlet cell as a neighbour. */
lbm_status lbm_create(const lbm_desc* desc, lbm_ctx** out);

/* Equilibrium initialisation f = f_eq(rho, u) on every non-solid cell
 * (InitializationEquilibriumConstVector.h:74-126).  NULL rho means rho0
 * everywhere; NULL velocity components mean 0.  Arrays are nx*ny*nz, x fastest.
 * The next lbm_step uses these rho/u for its first collision, as the reference
 * does (its moments are only recomputed at the end of a step). */
lbm_status lbm_init_equilibrium(lbm_ctx* ctx, const double* rho, const double* ux,
                                const double* uy, const double* uz);

/* Overwrite the population state (layout [q][z][y][x], FP64).  The next step
 * derives rho/u from these populations. */
lbm_status lbm_set_populations(lbm_ctx* ctx, const double* f);

/* Advance nsteps time steps; asynchronous on the context stream. */
lbm_status lbm_step(lbm_ctx* ctx, int32_t nsteps);

/* Wait for queued work; returns LBM_E_NAN if a diagnostic found NaN/Inf. */
lbm_status lbm_sync(lbm_ctx* ctx);

/* rho and u of every non-solid cell exactly as MomentDensityVelocityN27
 * (moments/MomentDensityVelocityN27.h:65-81) computes them from the current
 * populations; 0 on solid cells.  Any pointer may be NULL.  Synchronises. */
lbm_status lbm_get_macroscopic(lbm_ctx* ctx, double* rho, double* ux, double* uy, double* uz);

/* Current (post-stream, post-boundary) populations, [q][z][y][x], FP64;
 * 0 on solid cells.  Synchronises. */
lbm_status lbm_get_populations(lbm_ctx* ctx, double* f);

/* Population q of every cell (as lbm_get_populations, one plane): what
 * outputerVTK::distributionFunctionVTK writes (postprocesors/outputerVTK.h:
 * 660-740).  Synchronises. */
lbm_status lbm_get_population(lbm_ctx* ctx, int32_t q, double* f);

/* Output snapshots that do not stall the stepping.  The reference writes VTK
 * output at its plot cadence by copying rho and u to the host and waiting
 * (SolverTurbulentLES.h:243-249, outputerVTK.h:59-120).  lbm_snapshot_begin
 * queues, behind the steps already queued, a kernel that forms rho, u (the
 * values of lbm_get_macroscopic) and omega (lbm_get_omega) of that state into
 * a device buffer, and a copy of it into pinned host memory on a second
 * stream; it returns at once, and the steps queued after it run while the copy
 * goes on.  lbm_snapshot_wait waits for that copy only and hands out the
 * planes (any pointer may be NULL) and the step count of the state.  One
 * snapshot at a time: begin while another is pending is LBM_E_STATE.  The
 * first begin allocates 5 values per cell on the device and in pinned host
 * memory. */
lbm_status lbm_snapshot_begin(lbm_ctx* ctx);
lbm_status lbm_snapshot_wait(lbm_ctx* ctx, double* rho, double* ux, double* uy, double* uz, double* omega,
                             int64_t* step);

/* rho, ux, uy, uz of one cell (as lbm_get_macroscopic) without copying the
 * whole field: the reference's probeWrite reads four values
 * (SolverTurbulentLES.h:150-177).  Synchronises. */
lbm_status lbm_probe(lbm_ctx* ctx, int32_t x, int32_t y, int32_t z, double out[4]);

/* The same four values recorded on the device without synchronising, for a
 * probe written every step (probe_every_it = 1 makes the reference copy to
 * the host and wait every step, SolverTurbulentLES.h:150-177, 223-228).
 * lbm_probe_fetch returns the records made since the last fetch, oldest
 * first, as {rho, ux, uy, uz} quadruples: *n = their number; out may be NULL
 * to ask for the count only (nothing is cleared); otherwise max_records must
 * be >= *n and the records are cleared.  Synchronises. */
lbm_status lbm_probe_record(lbm_ctx* ctx, int32_t x, int32_t y, int32_t z);
lbm_status lbm_probe_fetch(lbm_ctx* ctx, double* out, int64_t max_records, int64_t* n);

/* Relaxation rate omega that the next step will use on every cell, computed
 * from the current state as Turbulence::omega does at the start of a step
 * (OmegaLES.h:69-131 / OmegaNo.h:34-47); 0 on solid cells.  Synchronises. */
lbm_status lbm_get_omega(lbm_ctx* ctx, double* omega);

/* FP64 reductions over non-solid cells of the current state:
 * out = {sum rho, sum rho*ux, sum rho*uy, sum rho*uz}.  Synchronises;
 * returns LBM_E_NAN if any is non-finite. */
lbm_status lbm_conservation(lbm_ctx* ctx, double out[4]);

/* Advance one step and reduce, in FP64 over that step's cells:
 * out[0..3] = mass and momentum of the state the step starts from (the
 * moments its collision uses), out[4..6] = momentum the fluid loses across
 * the half-way bounce-back links of wall-list nodes in this step, i.e. the
 * force on the walls (drag and lift on an obstacle): for every wall node x
 * and direction q whose source x - c_q is solid,
 * F -= c_q (f_bounced(x, q) + f*(x, opp q)).  With the true opposite table
 * this is the momentum-exchange sum 2 f*(x, k) c_k over links k into the
 * solid.  (No reference counterpart; the learned closure is excluded.) */
lbm_status lbm_step_diagnostics(lbm_ctx* ctx, double out[7]);

/* lbm_step_diagnostics without allocating: no counterpart in the reference.
 * For callers that cannot call hipMalloc (code under stream capture, tight
 * stepping loops, harnesses that forbid allocation).  Same step and the same
 * seven values as lbm_step_diagnostics, whose per-tile FP64 records sum on the
 * host; here they go to a fixed device array of the library (one per GPU,
 * 16 MiB, held until the process ends) and are summed on the device, so the
 * results agree to FP64 summation rounding.  A lattice needs at most
 * LBM_NOALLOC_MAX_TILES tiles of 256 cells (a step's workgroups; 67.1 M cells
 * when rows are multiples of 256), else LBM_E_UNSUPPORTED before any step is
 * taken.  Calls that overlap on one GPU (other threads or streams) race on the
 * shared array.  Synchronises. */
#define LBM_NOALLOC_MAX_TILES 262144
lbm_status lbm_step_diagnostics_noalloc(lbm_ctx* ctx, double out[7]);

/* ErrorQuadratic::errorEvaluation (errorEvaluations/ErrorQuadratic.h:33-90),
 * the solver loop's convergence norm, on the device: with u the velocity
 * lbm_get_macroscopic returns and u_prev the velocity of the previous call
 * (zero before the first),
 *   err = sqrt(sum sqrt((ux - ux_prev)^2 + (uy - uy_prev)^2 + (uy - uz_prev)^2))
 *       / sqrt(sum sqrt(ux^2 + uy^2 + uz^2))
 * over cells whose flag is 1 (the third difference keeps the reference's
 * uy - uz_prev, :43-44); then u_prev = u.  lbm_init_equilibrium sets u_prev
 * back to zero.  Each cell's terms are formed in the
 * arithmetic type (FP64 or FP32, the reference's RealType) and summed in
 * FP64; only the two sums cross to the host.  The first call allocates
 * u_prev (3 values per cell).  Synchronises.  Returns LBM_E_NAN (and
 * lbm_sync keeps returning it) when a sum is not a number, which is the
 * reference loop's break condition (SolverTurbulentLES.h:238-241). */
lbm_status lbm_error_quadratic(lbm_ctx* ctx, double* err);

/* Time averaging, MomentTimeAvg (moments/MomentTimeAvg.h:19-80) as the
 * solver uses it (SolverTurbulentLES.h:216-221, 255-264).  Running sums of
 * rho, ux, uy, uz per cell, in FP64, on the device (4 * 8 bytes per cell,
 * allocated by the first call).  The values added are those
 * lbm_get_macroscopic would return for the state (0 on solid cells).
 *  - lbm_moment_add: momentAdd for the current state.  The add itself is done
 *    by the next lbm_step's kernel, which reads that state anyway (no extra
 *    pass over the populations); a call that changes or reads the state first
 *    does it with a pass of its own.
 *  - lbm_step_moment_add: nsteps steps, counting the state after each one
 *    (the reference's momentAdd after every momentUpdate of the window).
 *  - lbm_moment_avg: momentAvg, sums / count into caller arrays (any may be
 *    NULL; the sums are kept, so averaging can continue); *count = states
 *    added.  LBM_E_STATE when nothing was added.  Synchronises.
 *  - lbm_moment_avg_reset: zero the sums and the count. */
lbm_status lbm_moment_add(lbm_ctx* ctx);
lbm_status lbm_step_moment_add(lbm_ctx* ctx, int32_t nsteps);
lbm_status lbm_moment_avg(lbm_ctx* ctx, double* rho, double* ux, double* uy, double* uz, int64_t* count);
lbm_status lbm_moment_avg_reset(lbm_ctx* ctx);

/* Steps completed so far. */
lbm_status lbm_step_count(lbm_ctx* ctx, int64_t* steps);

/* Average device time of one step over the last lbm_step call, in
 * milliseconds (HIP events on the context stream): the collide-stream kernel
 * plus any boundary patch kernels (equilibrium inlet, interpolated outlet,
 * ghost-layer periodic). */
lbm_status lbm_last_kernel_ms(lbm_ctx* ctx, double* ms);

/* Device-to-device copy of `bytes` bytes (a plain float4 stream kernel) run
 * `iters` times on the current device: the practical HBM bandwidth this
 * core's roofline fractions are reported next to.  gbps = 2*bytes/time. */
lbm_status lbm_bandwidth_probe(size_t bytes, int32_t iters, double* gbps, double* ms_per_iter);

/* Bytes of device memory held by the context. */
lbm_status lbm_device_bytes(lbm_ctx* ctx, size_t* bytes);

/* Number of visible HIP devices, and the name / gfx target of the current
 * one (name may be NULL).  Returns LBM_E_HIP when no device is visible. */
lbm_status lbm_device_info(int32_t* count, char* name, size_t name_len);

/* One node of a bounding-volume hierarchy over a triangle mesh: an inner node
 * has children left/right (count = 0); a leaf holds triangles
 * order[first .. first + count) (left = right = -1). */
typedef struct lbm_bvh_node {
    double lo[3], hi[3];
    int32_t left, right, first, count;
} lbm_bvh_node;

/* The OFF voxelizer's inside test on the current device: for every cell
 * (i, j, k) of a dims[0] x dims[1] x dims[2] grid (x fastest) the query point
 * ((i + 0.5 - additional) / resolution + bbmin), held in a float as the
 * reference's RealType is (geometryMesherOFF.h:90-103), is inside when the
 * segment to `outside` meets the triangles in an odd number of distinct
 * points, points closer than 1e-5 counted once (geometryHandlerOFF.h:135-182).
 * grid[cell] = 1 inside, 0 outside.  nodes/order are a hierarchy over the
 * triangles (the host voxelizer's), traversed in the same order so that hits
 * are merged in the same order.  Synchronises.  LBM_E_UNSUPPORTED if a
 * segment needs more than 64 hits or a hierarchy deeper than 64 levels
 * (none computed). (No reference counterpart as a GPU routine.) */
lbm_status lbm_voxelize(const double* vertices, int32_t n_vertices, const int32_t* triangles, int32_t n_triangles,
                        const lbm_bvh_node* nodes, int32_t n_nodes, const int32_t* order, const int32_t dims[3],
                        double additional, double resolution, const double bbmin[3], const double outside[3],
                        int32_t* grid);

const char* lbm_status_str(lbm_status s);
void lbm_destroy(lbm_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* LBM_CORE_H */
