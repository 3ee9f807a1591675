/*
 * lbm_host.h -- C ABI of the host-side setup library (liblbm_host.so).
 *
 * Thin C entry points over the C++ classes in lb_amd/csrc/host/lbm_host.hpp,
 * which keep the reference's setup API:
 *   - OFF voxelization: geometryHandlerOFF + geometryMesherOFF
 *     (3D/preprocess/src/geometry/geometryHandlerOFF.h:63-182,
 *      geometryMesherOFF.h:45-106, driven as in 3D/preprocess/preprocess.cpp:59-82);
 *   - flag-grid files: outputerMesh::MeshMatrixOut / MeshMatrixIn
 *     (3D/simulation/src/postprocesors/outputerMesh.h:19-211);
 *   - boundary meshing: geometryMesherBoundary
 *     (3D/simulation/src/geometry/geometryMesherBoundary.h:47-617);
 *   - a whole case (unit conversion, solver time loop, output) as
 *     3D/simulation/main.cu:26-243 drives it, on top of include/lbm_core.h.
 * Status codes: 0 ok, nonzero failure (documented per function).
 */
#ifndef LBM_HOST_H
#define LBM_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "lbm_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Voxelize an OFF/NOFF/COFF surface: dims = ceil((BBmax-BBmin)*res + 2*add)
 * in float, query point of cell (i,j,k) = ((i+0.5-add)/res + BBmin) held in
 * float, fluid (1) when the segment to `outside` crosses the surface an odd
 * number of times.  grid may be NULL (sizes only); otherwise it receives
 * dims[0]*dims[1]*dims[2] ints, x fastest.  threads > 0: that many CPU
 * threads; 0: all cores; < 0: the inside tests run on the current GPU
 * (lbm_voxelize, same hierarchy and arithmetic, same grid).
 * Returns 0, 1 (empty polyhedron), 2 (unreadable / malformed file) or, for a
 * failed GPU run, 10 + its lbm_status. */
int lbmh_voxelize_off(const char* path, float resolution, float additional, const double outside[3], int threads,
                      int32_t dims[3], float bbmin[3], float bbmax[3], int32_t* grid);

/* Write / read the flag-grid text format.  Read with grid == NULL returns the
 * header only.  Returns 0, or 1 on I/O error / truncated or malformed file. */
int lbmh_mesh_write(const char* path, const int32_t dims[3], float resolution, float additional,
                    const float bbmin[3], const float bbmax[3], const int32_t* grid);
int lbmh_mesh_read(const char* path, int32_t dims[3], float* resolution, float* additional, float bbmin[3],
                   float bbmax[3], int32_t* grid);

/* Boundary mesher over a 0/1 grid (x fastest). */
typedef struct lbmh_mesher lbmh_mesher;
int lbmh_mesher_create(const int32_t dims[3], const int32_t* grid01, float resolution, float additional,
                       const float bbmin[3], lbmh_mesher** out);
/* meshingBoundaryWall (:439-484) */
int lbmh_mesher_wall(lbmh_mesher* m);
/* meshingBoundaryConditionInletUniform (:47-108); kind = LBM_BC_INLET_*.
 * Returns 1 where the reference throws (all velocity components equal). */
int lbmh_mesher_inlet(lbmh_mesher* m, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                      const float normal[3], const float velocity[3], int32_t kind);
/* meshingBoundaryConditionInletParaboloidCircle (:111-178) and
 * ...ParaboloidRectangle (:180-274): the cuboid's fluid cells become inlet
 * nodes (kind = LBM_BC_INLET_*), then every inlet entry whose centre lies in
 * the cuboid gets velocity -v n with, for the circle, v = 3 u_mean (1 - (r/R)^2)
 * within radius R of `center` (0 outside; its mean over the disc is 1.5 u_mean),
 * and for the rectangle v = 2 u_mean prod_i d_i (L_i - d_i) / prod_i (L_i/2)^2
 * over the axes across the normal (d_i from the rectangle's low corner, L_i =
 * dims[i]; peak 2 u_mean).  The rectangle version sets dims[] along the normal
 * to 1, as the reference's reference parameters are.  The reference's two
 * functions start from the uniform inlet with velocity (0, 0, 0), which that
 * function rejects (:52-54), so as written they throw; here the cells are
 * meshed without that check.  Returns 0, or 1 for a NULL argument. */
int lbmh_mesher_inlet_paraboloid_circle(lbmh_mesher* m, const double p1[3], const double p2[3], const double p3[3],
                                        int32_t id, const float normal[3], const double center[3], float radius,
                                        float mean_velocity, int32_t kind);
int lbmh_mesher_inlet_paraboloid_rectangle(lbmh_mesher* m, const double p1[3], const double p2[3],
                                           const double p3[3], int32_t id, const float normal[3],
                                           const double center[3], float dims[3], float mean_velocity, int32_t kind);
/* meshingBoundaryConditionOutlet (:276-325) */
int lbmh_mesher_outlet(lbmh_mesher* m, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                       const float normal[3], float density);
/* meshingBoundaryConditionPeriodic (:379-437) */
int lbmh_mesher_periodic(lbmh_mesher* m, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                         const float normal[3], int32_t periodic_index);
/* compileBoundaryArray* (:486-600) + arrayTransfer: flags (nx*ny*nz) and the
 * boundary list (walls, inlets, outlets, periodic, in that order).  bc may be
 * NULL to query the count; *n_bc is in/out (capacity / count). */
int lbmh_mesher_compile(lbmh_mesher* m, int32_t* flags, lbm_bc_node* bc, int32_t* n_bc);
void lbmh_mesher_destroy(lbmh_mesher* m);

/* A whole case as main.cu drives it: constants by LBMConstants field name,
 * geometry, unit conversion, initialisation and the time loop. */
typedef struct lbmh_case lbmh_case;
int lbmh_case_create(lbmh_case** out);
/* Set/get an LBMConstants field by its reference name (e.g. "tau",
 * "U_lb", "dimX_int", "CLES", "collision", "storage", "periodic_y").
 * Returns 1 for an unknown name. */
int lbmh_case_set(lbmh_case* c, const char* name, double value);
int lbmh_case_get(lbmh_case* c, const char* name, double* value);
int lbmh_case_set_results_dir(lbmh_case* c, const char* dir);
/* Geometry from a flag-grid file (MeshMatrixIn) or from a 0/1 grid. */
int lbmh_case_load_mesh(lbmh_case* c, const char* path);
int lbmh_case_set_grid(lbmh_case* c, const int32_t dims[3], const int32_t* grid01, float resolution,
                       float additional, const float bbmin[3], const float bbmax[3]);
/* Boundary meshing in main.cu order (wall first, then cuboids, then compile). */
int lbmh_case_mesh_wall(lbmh_case* c);
int lbmh_case_inlet(lbmh_case* c, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                    const float normal[3], const float velocity[3], int32_t kind);
int lbmh_case_inlet_paraboloid_circle(lbmh_case* c, const double p1[3], const double p2[3], const double p3[3],
                                      int32_t id, const float normal[3], const double center[3], float radius,
                                      float mean_velocity, int32_t kind);
int lbmh_case_inlet_paraboloid_rectangle(lbmh_case* c, const double p1[3], const double p2[3], const double p3[3],
                                         int32_t id, const float normal[3], const double center[3], float dims[3],
                                         float mean_velocity, int32_t kind);
int lbmh_case_outlet(lbmh_case* c, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                     const float normal[3], float density);
int lbmh_case_periodic(lbmh_case* c, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                       const float normal[3], int32_t periodic_index);
int lbmh_case_compile(lbmh_case* c);
/* SolverTurbulentLES::convertToLattice / initializeSimulation / runSimulation;
 * the latter two return lbm_status values. */
int lbmh_case_convert_to_lattice(lbmh_case* c);
int lbmh_case_initialize(lbmh_case* c);
int lbmh_case_run(lbmh_case* c);
/* Flags of the case and its device lattice (for lbm_get_* reads). */
int lbmh_case_flags(lbmh_case* c, int32_t* flags);
lbm_ctx* lbmh_case_lattice(lbmh_case* c);
/* InitializationEquilibriumVariables (InitializationEquilibriumVariables.h:24-178):
 * lbmh_read_vtk_fields reads rho (after the first "LOOKUP_TABLE default" line) and u
 * (after the first "VECTORS U double" line) for dims[0]*dims[1]*dims[2] cells, x
 * fastest, each value parsed into float, lattice units as written by
 * variablesLatticeVTK; any output may be NULL.  Returns 0, 1 (cannot open), 2 (rho
 * short), 3 (u short), 4 (bad arguments).  lbmh_case_set_init_file makes
 * lbmh_case_initialize start from such a file (name relative to the results
 * directory) instead of the uniform state; inlet cells still get the inlet
 * velocity. */
int lbmh_read_vtk_fields(const char* path, const int32_t dims[3], double* rho, double* ux, double* uy, double* uz);
int lbmh_case_set_init_file(lbmh_case* c, const char* name);

/* outputerVTK writers (postprocesors/outputerVTK.h:59-822) on the case's
 * current state, into the results directory: which = "variables" (physical
 * units, file variables.<index>.vtk), "variablesLattice" (variablesLattice<index>.vtk),
 * "omega", "tau", "distributionFunction" (population <index>,
 * distributionFunction<index>.vtk), "variablesTimeAvg" or
 * "variablesTimeAvgLattice" (the averages of the last run; both write
 * variablesTimeAvg.vtk, as the reference's do).  Returns 0, 1 (unknown name or
 * no averages) or 2 (read or write failed). */
int lbmh_case_write_vtk(lbmh_case* c, const char* which, int32_t index);

/* Error norms recorded by the loop (ErrorQuadratic), oldest first. */
int lbmh_case_errors(lbmh_case* c, double* out, int32_t* n);
/* Time averages of rho, ux, uy, uz (lattice units, x fastest) that the run
 * formed over its last iterationsMomentAvg - 1 steps (MomentTimeAvg,
 * SolverTurbulentLES.h:216-264; also written to variablesTimeAvg.vtk when
 * output is on); *count = states averaged.  Non-zero when the run averaged
 * nothing. */
int lbmh_case_time_avg(lbmh_case* c, double* rho, double* ux, double* uy, double* uz, int64_t* count);
void lbmh_case_destroy(lbmh_case* c);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HOST_H */
