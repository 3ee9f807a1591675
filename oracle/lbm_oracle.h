/*
 * lbm_oracle.h -- CPU oracle of the reference collide-stream path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of stloufra/LB's
 * unfused two-lattice ("AB") time step, used as the checker for the HIP core
 * and as the CPU baseline of bench.py.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * It keeps the reference's arrays (df, df_post, rho, u, omega, flags, BC
 * lists), its operator order (SolverTurbulentLES.h:184-214) and the operation
 * order of every arithmetic expression, so that an FP64 run reproduces the
 * reference's FP64 fields.  Layout: df[q][z][y][x], x fastest.
 *
 * Extensions beyond the reference, each documented where implemented:
 *   - index-wrap periodicity (desc.periodic), equivalent to the reference's
 *     ghost-layer Periodic policy for interior cells;
 *   - corrected tables/literals when desc.compat clears the quirk bits.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>
#include "../include/lbm_core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

/* Same descriptor as the HIP core.  storage F64 -> double arithmetic,
 * anything else -> float arithmetic (the reference's RealType switch,
 * traits/LBMTraits.h:72).  Returns an lbm_status code. */
int orc_create(const lbm_desc* d, orc** out);
int orc_init_equilibrium(orc* o, const double* rho, const double* ux, const double* uy, const double* uz);
int orc_set_populations(orc* o, const double* f);
int orc_step(orc* o, int nsteps);
int orc_get_macroscopic(orc* o, double* rho, double* ux, double* uy, double* uz);
/* MomentTimeAvg: add the current rho, u to FP64 sums; sums / count. */
int orc_moment_add(orc* o);
int orc_moment_avg(orc* o, double* rho, double* ux, double* uy, double* uz, long long* count);
int orc_get_populations(orc* o, double* f);
int orc_get_post_populations(orc* o, double* f);
int orc_get_omega(orc* o, double* omega);
/* ErrorQuadratic::errorEvaluation (errorEvaluations/ErrorQuadratic.h:33-90). */
int orc_error_quadratic(orc* o, double* err);
void orc_destroy(orc* o);
void orc_set_threads(int n);
int orc_get_threads(void);

/* Deliberate mistakes, for tests that show that a pin fails on them (the
 * oracle is otherwise always built and run with none):
 *   NO_A          the 2017 operator's fourth-order coefficient A = 0
 *   OUTLET_SWAP   the interpolated outlet weights c_s and 1 - c_s exchanged
 *   INLET_NU3     the corrected equilibrium inlet uses nu/3 instead of 3 nu
 *   BACK_221      the back-transform's 221 moment with 2 S110 S111 for 4 S110 S111
 *   INLET_SIGN    the corrected equilibrium inlet's +x branch with the sign flipped
 *   A_NEG         the 2017 operator's A with its sign flipped
 *   NO_B          the 2017 operator's fourth-order coefficient B = 0 */
enum { ORC_PLANT_NO_A = 1, ORC_PLANT_OUTLET_SWAP = 2, ORC_PLANT_INLET_NU3 = 4, ORC_PLANT_BACK_221 = 8,
       ORC_PLANT_INLET_SIGN = 16, ORC_PLANT_A_NEG = 32, ORC_PLANT_NO_B = 64 };
void orc_plant(int bits);

/* Single-cell operators, for unit tests against closed forms.
 * f and out hold 27 values in reference direction order.  flag is the
 * cell's flag-grid value; only LBM_COLLISION_COMBINED reads it (1 selects the
 * 2015 branch, anything else the 2017 branch). */
void orc_cell_collide(int collision, uint32_t compat, int use_double, const double* f, double rho,
                      double ux, double uy, double uz, double omega, int flag, double* out);
double orc_cell_omega_les(uint32_t compat, int use_double, const double* f, double rho, double ux,
                          double uy, double uz, double tau, double c_les);
void orc_cell_feq(uint32_t compat, int use_double, double rho, double ux, double uy, double uz, double* out);
void orc_tables(uint32_t compat, int* c /*27*3*/, int* c_rev /*27*/, double* w /*27*/);
/* The three-pass central-moment transform of the cumulant operator and its
 * inverse (FP64): k[9*i + 3*j + l] = central moment of orders (i, j, l). */
void orc_cell_central_moments(const double* f, double ux, double uy, double uz, double* k);
void orc_cell_from_central_moments(const double* k, double ux, double uy, double uz, double* f);

#ifdef __cplusplus
}
#endif

#endif
