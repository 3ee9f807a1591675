/*
 * lbm_oracle.c -- CPU oracle of the reference collide-stream path.
 * TEST INFRASTRUCTURE ONLY (see lbm_oracle.h): the checker for the HIP core
 * and the CPU baseline of bench.py; never linked into the product.
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* D3Q27 velocity set in the reference's slot order
 * (3D/simulation/src/solvers/models/D3Q27/D3Q27.h:81-111).  The order defines
 * the population layout, so it is kept. */
static const int ORC_C[27][3] = {
    {0, 0, 0},
    {1, 0, 0}, {-1, 0, 0}, {0, 0, -1}, {0, 0, 1}, {0, -1, 0}, {0, 1, 0},
    {1, 0, -1}, {-1, 0, 1}, {1, 0, 1}, {-1, 0, -1}, {-1, -1, 0}, {1, 1, 0},
    {0, 1, -1}, {0, -1, 1}, {-1, 1, 0}, {1, -1, 0}, {0, 1, 1}, {0, -1, -1},
    {-1, 1, -1}, {1, -1, 1}, {-1, -1, 1}, {1, 1, -1}, {1, -1, -1}, {-1, 1, 1}, {-1, -1, -1}, {1, 1, 1}};

/* Opposite direction, computed from ORC_C (the corrected table). */
static int orc_opposite(int q)
{
    for (int p = 0; p < 27; ++p)
        if (ORC_C[p][0] == -ORC_C[q][0] && ORC_C[p][1] == -ORC_C[q][1] && ORC_C[p][2] == -ORC_C[q][2]) return p;
    return -1;
}

/* Reverse table: the reference's c_rev (D3Q27.h:120-124) differs from the
 * true opposite only at slot 20, where it holds 21 (quirk Q1). */
static void orc_crev(uint32_t compat, int* crev)
{
    for (int q = 0; q < 27; ++q) crev[q] = orc_opposite(q);
    if (compat & LBM_COMPAT_CREV20) crev[20] = 21;
}

/* Weight class of a direction: number of non-zero components. */
static int orc_nnz(int q) { return (ORC_C[q][0] != 0) + (ORC_C[q][1] != 0) + (ORC_C[q][2] != 0); }

/* Deliberate mistakes (ORC_PLANT_* in lbm_oracle.h), set only by tests that
 * show a pin rejects them; 0 in every other use. */
static int g_plant = 0;

void orc_plant(int bits) { g_plant = bits; }

/* ---------------- double instantiation ---------------- */
#define REAL double
#define FN(x) x##_d
#define ORC_SQRT sqrt
#define ORC_FABS fabs
#include "lbm_oracle_impl.h"
#undef REAL
#undef FN
#undef ORC_SQRT
#undef ORC_FABS

/* ---------------- float instantiation ---------------- */
#define REAL float
#define FN(x) x##_f
#define ORC_SQRT sqrtf
#define ORC_FABS fabsf
#include "lbm_oracle_impl.h"
#undef REAL
#undef FN
#undef ORC_SQRT
#undef ORC_FABS

struct orc {
    int dbl;
    double rho0;
    state_d d;
    state_f f;
    double* avg;          /* time-average sums: rho, ux, uy, uz planes */
    long long avg_count;
};

static int g_threads = 0;

void orc_set_threads(int n)
{
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Cumulant constants: 1.f/3.f, 2.f/3.f, 4.f/3.f and lambda = 0.01f
 * (CollisionCumD3Q27Turbulent2015.h:280-306, Turbulent2017.h:244-246, 349-362),
 * and omega_4's 8 omega_1^2 (Turbulent2017.h:260; the published rate has 9). */
#define ORC_FILL_CK(K, T)                                                                  \
    do {                                                                                   \
        const int f32c_ = (compat & LBM_COMPAT_F32CONST) != 0;                             \
        (K)->c13 = f32c_ ? (T)(1.f / 3.f) : (T)(1.0 / 3.0);                                \
        (K)->c23 = f32c_ ? (T)(2.f / 3.f) : (T)(2.0 / 3.0);                                \
        (K)->c43 = f32c_ ? (T)(4.f / 3.f) : (T)(4.0 / 3.0);                                \
        (K)->lam = f32c_ ? (T)0.01f : (T)0.01;                                             \
        (K)->w4q = (compat & LBM_COMPAT_OMEGA4) ? (T)8 : (T)9;                             \
    } while (0)

/* Weights: D3Q27.h:113-118 writes them as float quotients (8.f/27.f ...),
 * which survive RealType = double (quirk Q2). */
#define ORC_FILL_CONSTS(S, T)                                                              \
    do {                                                                                   \
        for (int q = 0; q < 27; ++q) {                                                     \
            int k = orc_nnz(q);                                                            \
            if (compat & LBM_COMPAT_F32CONST)                                              \
                (S)->w[q] = (T)(k == 0 ? 8.f / 27.f : k == 1 ? 2.f / 27.f : k == 2 ? 1.f / 54.f : 1.f / 216.f); \
            else                                                                           \
                (S)->w[q] = (T)(k == 0 ? 8.0 / 27.0 : k == 1 ? 2.0 / 27.0 : k == 2 ? 1.0 / 54.0 : 1.0 / 216.0); \
        }                                                                                  \
        orc_crev(compat, (S)->crev);                                                       \
        (S)->c13 = (compat & LBM_COMPAT_F32CONST) ? (T)(1.f / 3.f) : (T)(1.0 / 3.0);       \
        ORC_FILL_CK(&(S)->ck, T);                                                          \
        (S)->cs = (compat & LBM_COMPAT_F32CONST) ? (T)(1.f / sqrtf(3.f)) : (T)(1.0 / sqrt(3.0)); \
    } while (0)

#define ORC_ALLOC_STATE(S, T)                                                               \
    do {                                                                                    \
        (S)->nx = d->nx; (S)->ny = d->ny; (S)->nz = d->nz; (S)->N = N;                      \
        (S)->per[0] = d->periodic[0] != 0; (S)->per[1] = d->periodic[1] != 0;               \
        (S)->per[2] = d->periodic[2] != 0;                                                  \
        (S)->collision = d->collision; (S)->closure = d->closure; (S)->compat = compat;     \
        (S)->tau = (T)d->tau; (S)->nu = (T)d->nu; (S)->CLES = (T)d->c_les;                  \
        (S)->sponge = d->outlet_sponge != 0;                                                \
        (S)->wall_mesh = d->wall_mesh != 0;                                                 \
        (S)->sponge_low = (T)d->sponge_low; (S)->sponge_high = (T)d->sponge_high;           \
        ORC_FILL_CONSTS(S, T);                                                              \
        (S)->flags = (int*)malloc(sizeof(int) * N);                                         \
        (S)->df = (T*)calloc((size_t)27 * N, sizeof(T));                                    \
        (S)->df_post = (T*)calloc((size_t)27 * N, sizeof(T));                               \
        (S)->rho = (T*)calloc((size_t)N, sizeof(T));                                        \
        (S)->u = (T*)calloc((size_t)3 * N, sizeof(T));                                      \
        (S)->u_err = (T*)calloc((size_t)3 * N, sizeof(T));                                  \
        (S)->omega = (T*)calloc((size_t)N, sizeof(T));                                      \
        (S)->n_bc = d->n_bc;                                                                \
        (S)->bc = (lbm_bc_node*)malloc(sizeof(lbm_bc_node) * (d->n_bc > 0 ? d->n_bc : 1));  \
        if (!(S)->flags || !(S)->df || !(S)->df_post || !(S)->rho || !(S)->u || !(S)->u_err || \
            !(S)->omega || !(S)->bc) {                                                      \
            orc_destroy(o);                                                                 \
            return LBM_E_OOM;                                                               \
        }                                                                                   \
        memcpy((S)->flags, d->flags, sizeof(int) * N);                                      \
        if (d->n_bc > 0) memcpy((S)->bc, d->bc, sizeof(lbm_bc_node) * d->n_bc);             \
    } while (0)

int orc_create(const lbm_desc* d, orc** out)
{
    if (!d || !out || !d->flags || d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return LBM_E_ARG;
    if (d->n_bc > 0 && !d->bc) return LBM_E_ARG;
    if (d->collision < LBM_COLLISION_SRT || d->collision > LBM_COLLISION_COMBINED) return LBM_E_ARG;
    if (d->closure != LBM_CLOSURE_NONE && d->closure != LBM_CLOSURE_SMAGORINSKY) return LBM_E_UNSUPPORTED;
    orc* o = (orc*)calloc(1, sizeof(orc));
    if (!o) return LBM_E_OOM;
    long N = (long)d->nx * d->ny * d->nz;
    uint32_t compat = d->compat;
    o->dbl = d->storage == LBM_STORAGE_F64;
    o->rho0 = d->rho0;
    if (o->dbl)
        ORC_ALLOC_STATE(&o->d, double);
    else
        ORC_ALLOC_STATE(&o->f, float);
    *out = o;
    return LBM_OK;
}

#define ORC_FREE_STATE(S)                                                                   \
    do {                                                                                    \
        free((S)->flags); free((S)->df); free((S)->df_post); free((S)->rho); free((S)->u);   \
        free((S)->u_err); free((S)->omega); free((S)->bc);                                  \
    } while (0)

void orc_destroy(orc* o)
{
    if (!o) return;
    ORC_FREE_STATE(&o->d);
    ORC_FREE_STATE(&o->f);
    free(o->avg);
    free(o);
}

int orc_init_equilibrium(orc* o, const double* rho, const double* ux, const double* uy, const double* uz)
{
    if (!o) return LBM_E_ARG;
    if (o->dbl)
        init_eq_d(&o->d, rho, ux, uy, uz, o->rho0);
    else
        init_eq_f(&o->f, rho, ux, uy, uz, o->rho0);
    return LBM_OK;
}

int orc_set_populations(orc* o, const double* f)
{
    if (!o || !f) return LBM_E_ARG;
    if (o->dbl)
        set_pop_d(&o->d, f);
    else
        set_pop_f(&o->f, f);
    return LBM_OK;
}

int orc_step(orc* o, int nsteps)
{
    if (!o || nsteps < 0) return LBM_E_ARG;
    for (int k = 0; k < nsteps; ++k) {
        if (o->dbl)
            step_d(&o->d);
        else
            step_f(&o->f);
    }
    return LBM_OK;
}

#define ORC_COPY_OUT(S)                                                                     \
    do {                                                                                    \
        for (long i = 0; i < (S)->N; ++i) {                                                 \
            if (rho) rho[i] = (S)->rho[i];                                                  \
            if (ux) ux[i] = (S)->u[3 * i];                                                  \
            if (uy) uy[i] = (S)->u[3 * i + 1];                                              \
            if (uz) uz[i] = (S)->u[3 * i + 2];                                              \
        }                                                                                   \
    } while (0)

int orc_get_macroscopic(orc* o, double* rho, double* ux, double* uy, double* uz)
{
    if (!o) return LBM_E_ARG;
    if (o->dbl)
        ORC_COPY_OUT(&o->d);
    else
        ORC_COPY_OUT(&o->f);
    return LBM_OK;
}

/* MomentTimeAvg::momentAdd (MomentTimeAvg.h:19-45): every cell's rho and u
 * (the fields of the last moment update) added to the sums; the sums are
 * double here whatever the state's precision. */
#define ORC_MOMENT_ADD(S)                                                                   \
    do {                                                                                    \
        const long N_ = (S)->N;                                                             \
        for (long i = 0; i < N_; ++i) {                                                     \
            o->avg[i] += (double)(S)->rho[i];                                               \
            o->avg[N_ + i] += (double)(S)->u[3 * i];                                        \
            o->avg[2 * N_ + i] += (double)(S)->u[3 * i + 1];                                \
            o->avg[3 * N_ + i] += (double)(S)->u[3 * i + 2];                                \
        }                                                                                   \
    } while (0)

int orc_moment_add(orc* o)
{
    if (!o) return LBM_E_ARG;
    const long N = o->dbl ? o->d.N : o->f.N;
    if (!o->avg) {
        o->avg = (double*)calloc((size_t)4 * N, sizeof(double));
        if (!o->avg) return LBM_E_OOM;
        o->avg_count = 0;
    }
    if (o->dbl)
        ORC_MOMENT_ADD(&o->d);
    else
        ORC_MOMENT_ADD(&o->f);
    o->avg_count++;
    return LBM_OK;
}

/* MomentTimeAvg::momentAvg (MomentTimeAvg.h:47-77): sums / counter. */
int orc_moment_avg(orc* o, double* rho, double* ux, double* uy, double* uz, long long* count)
{
    if (!o) return LBM_E_ARG;
    if (count) *count = o->avg ? o->avg_count : 0;
    if (!o->avg || o->avg_count == 0) return LBM_E_STATE;
    const long N = o->dbl ? o->d.N : o->f.N;
    double* outs[4] = {rho, ux, uy, uz};
    for (int k = 0; k < 4; ++k)
        if (outs[k])
            for (long i = 0; i < N; ++i) outs[k][i] = o->avg[k * N + i] / (double)o->avg_count;
    return LBM_OK;
}

int orc_get_populations(orc* o, double* f)
{
    if (!o || !f) return LBM_E_ARG;
    if (o->dbl) {
        for (long i = 0; i < 27 * o->d.N; ++i) f[i] = o->d.df[i];
    } else {
        for (long i = 0; i < 27 * o->f.N; ++i) f[i] = o->f.df[i];
    }
    return LBM_OK;
}

int orc_get_post_populations(orc* o, double* f)
{
    if (!o || !f) return LBM_E_ARG;
    if (o->dbl) {
        for (long i = 0; i < 27 * o->d.N; ++i) f[i] = o->d.df_post[i];
    } else {
        for (long i = 0; i < 27 * o->f.N; ++i) f[i] = o->f.df_post[i];
    }
    return LBM_OK;
}

int orc_get_omega(orc* o, double* w)
{
    if (!o || !w) return LBM_E_ARG;
    /* the rate the next step would use: Turbulence::omega on the current state,
     * then the outlet's omega damping (SolverTurbulentLES.h:188-193) */
    if (o->dbl) {
        op_omega_d(&o->d);
        op_outlet_omega_d(&o->d);
        for (long i = 0; i < o->d.N; ++i) w[i] = o->d.flags[i] != 0 ? o->d.omega[i] : 0.0;
    } else {
        op_omega_f(&o->f);
        op_outlet_omega_f(&o->f);
        for (long i = 0; i < o->f.N; ++i) w[i] = o->f.flags[i] != 0 ? o->f.omega[i] : 0.0;
    }
    return LBM_OK;
}

int orc_error_quadratic(orc* o, double* err)
{
    if (!o || !err) return LBM_E_ARG;
    *err = o->dbl ? error_quadratic_d(&o->d) : error_quadratic_f(&o->f);
    return LBM_OK;
}

/* ---- single-cell entry points for unit tests ---- */

#define ORC_CELL_CONSTS(T)                                                                  \
    T w[27]; int crev[27]; T c13; struct { T c13, c23, c43, lam, w4q; } ck;                     \
    for (int q = 0; q < 27; ++q) {                                                          \
        int k = orc_nnz(q);                                                                 \
        w[q] = (compat & LBM_COMPAT_F32CONST)                                               \
                   ? (T)(k == 0 ? 8.f / 27.f : k == 1 ? 2.f / 27.f : k == 2 ? 1.f / 54.f : 1.f / 216.f) \
                   : (T)(k == 0 ? 8.0 / 27.0 : k == 1 ? 2.0 / 27.0 : k == 2 ? 1.0 / 54.0 : 1.0 / 216.0); \
    }                                                                                       \
    orc_crev(compat, crev); (void)crev;                                                     \
    c13 = (compat & LBM_COMPAT_F32CONST) ? (T)(1.f / 3.f) : (T)(1.0 / 3.0);                 \
    ORC_FILL_CK(&ck, T);

void orc_tables(uint32_t compat, int* c, int* c_rev, double* wout)
{
    ORC_CELL_CONSTS(double)
    (void)c13; (void)ck;
    for (int q = 0; q < 27; ++q) {
        if (c) { c[3 * q] = ORC_C[q][0]; c[3 * q + 1] = ORC_C[q][1]; c[3 * q + 2] = ORC_C[q][2]; }
        if (c_rev) c_rev[q] = crev[q];
        if (wout) wout[q] = w[q];
    }
}

/* One cell of collision operator `collision`; for LBM_COLLISION_COMBINED,
 * `flag` is the cell's flag-grid value (1 selects the 2015 branch). */
void orc_cell_collide(int collision, uint32_t compat, int use_double, const double* f, double rho, double ux,
                      double uy, double uz, double omega, int flag, double* out)
{
    if (use_double) {
        ORC_CELL_CONSTS(double)
        (void)c13;
        ck_d k = {ck.c13, ck.c23, ck.c43, ck.lam, ck.w4q};
        double o[27];
        collide_one_d(collision, compat, w, &k, flag, f, rho, ux, uy, uz, omega, o);
        for (int q = 0; q < 27; ++q) out[q] = o[q];
    } else {
        ORC_CELL_CONSTS(float)
        (void)c13;
        ck_f k = {ck.c13, ck.c23, ck.c43, ck.lam, ck.w4q};
        float fi[27], o[27];
        for (int q = 0; q < 27; ++q) fi[q] = (float)f[q];
        collide_one_f(collision, compat, w, &k, flag, fi, (float)rho, (float)ux, (float)uy, (float)uz,
                      (float)omega, o);
        for (int q = 0; q < 27; ++q) out[q] = o[q];
    }
}

double orc_cell_omega_les(uint32_t compat, int use_double, const double* f, double rho, double ux, double uy,
                          double uz, double tau, double c_les)
{
    if (use_double) {
        ORC_CELL_CONSTS(double)
        (void)c13; (void)ck;
        return omega_les_d(w, f, rho, ux, uy, uz, tau, c_les);
    } else {
        ORC_CELL_CONSTS(float)
        (void)c13; (void)ck;
        float fi[27];
        for (int q = 0; q < 27; ++q) fi[q] = (float)f[q];
        return omega_les_f(w, fi, (float)rho, (float)ux, (float)uy, (float)uz, (float)tau, (float)c_les);
    }
}

void orc_cell_feq(uint32_t compat, int use_double, double rho, double ux, double uy, double uz, double* out)
{
    if (use_double) {
        ORC_CELL_CONSTS(double)
        (void)c13; (void)ck;
        for (int q = 0; q < 27; ++q) out[q] = feq_d(w, q, ux, uy, uz, rho);
    } else {
        ORC_CELL_CONSTS(float)
        (void)c13; (void)ck;
        for (int q = 0; q < 27; ++q) out[q] = feq_f(w, q, (float)ux, (float)uy, (float)uz, (float)rho);
    }
}

void orc_cell_central_moments(const double* f, double ux, double uy, double uz, double* kout)
{
    double F[3][3][3], A[3][3][3], B[3][3][3], k[3][3][3];
    for (int q = 0; q < 27; ++q) F[ORC_C[q][0] + 1][ORC_C[q][1] + 1][ORC_C[q][2] + 1] = f[q];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) fwd_d(F[a][b][0], F[a][b][1], F[a][b][2], uz, &A[a][b][0], &A[a][b][1], &A[a][b][2]);
    for (int a = 0; a < 3; ++a)
        for (int r = 0; r < 3; ++r) fwd_d(A[a][0][r], A[a][1][r], A[a][2][r], uy, &B[a][0][r], &B[a][1][r], &B[a][2][r]);
    for (int b = 0; b < 3; ++b)
        for (int r = 0; r < 3; ++r) fwd_d(B[0][b][r], B[1][b][r], B[2][b][r], ux, &k[0][b][r], &k[1][b][r], &k[2][b][r]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int l = 0; l < 3; ++l) kout[9 * i + 3 * j + l] = k[i][j][l];
}

void orc_cell_from_central_moments(const double* kin, double ux, double uy, double uz, double* f)
{
    double s[3][3][3], A[3][3][3], B[3][3][3], F[3][3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int l = 0; l < 3; ++l) s[i][j][l] = kin[9 * i + 3 * j + l];
    for (int b = 0; b < 3; ++b)
        for (int r = 0; r < 3; ++r) inv_d(s[0][b][r], s[1][b][r], s[2][b][r], ux, &A[1][b][r], &A[0][b][r], &A[2][b][r]);
    for (int a = 0; a < 3; ++a)
        for (int r = 0; r < 3; ++r) inv_d(A[a][0][r], A[a][1][r], A[a][2][r], uy, &B[a][1][r], &B[a][0][r], &B[a][2][r]);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) inv_d(B[a][b][0], B[a][b][1], B[a][b][2], uz, &F[a][b][1], &F[a][b][0], &F[a][b][2]);
    for (int q = 0; q < 27; ++q) f[q] = F[ORC_C[q][0] + 1][ORC_C[q][1] + 1][ORC_C[q][2] + 1];
}
