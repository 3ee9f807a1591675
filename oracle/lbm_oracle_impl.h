/*
 * lbm_oracle_impl.h -- precision-generic body of the CPU oracle.
 * TEST INFRASTRUCTURE ONLY (see lbm_oracle.h).
 *
 * Included twice by lbm_oracle.c with REAL = double / float and FN(x) giving
 * the suffixed name.  Every expression keeps the operand order of the
 * reference line it cites, and float literals stay float literals, so that
 * C's promotion rules reproduce the reference's mixed float/RealType
 * arithmetic (quirk Q2).  Build with -ffp-contract=off.
 */

/* Constants of the cumulant relaxation: 1/3, 2/3, 4/3 and the limiter
 * parameter lambda = 0.01 (float literals in the reference: 1.f/3.f, 2.f/3.f,
 * -4.f/3.f, 0.01f; exact values when the F32CONST quirk is off), and the
 * omega_1^2 coefficient of omega_4's denominator (reference 8, published 9;
 * LBM_COMPAT_OMEGA4). */
typedef struct FN(ck) {
    REAL c13, c23, c43, lam, w4q;
} FN(ck);

typedef struct FN(state) {
    int nx, ny, nz;
    long N;
    int per[3];
    int collision, closure;
    uint32_t compat;
    REAL tau, nu, CLES, c13, cs;
    FN(ck) ck;
    int sponge;             /* outlet omega damping on */
    int wall_mesh;          /* BounceBackWallHalfMesh instead of the wall list */
    REAL sponge_low, sponge_high;
    REAL w[27];
    int crev[27];
    int* flags;
    REAL* df;       /* [q][N] */
    REAL* df_post;  /* [q][N] */
    REAL* rho;      /* [N] */
    REAL* u;        /* [3N], AoS as the reference's StaticVector field */
    REAL* omega;    /* [N] */
    REAL* u_err;    /* [3N] previous velocity for ErrorQuadratic */
    lbm_bc_node* bc;
    int n_bc;
} FN(state);

/* f_equilibrium (OmegaLES.h:47-66, CollisionSRTTurbulent.h:35-54,
 * InitializationEquilibriumConstVector.h:53-71, InletVelocityEquilibrium.h:34-49):
 *   uc = c0*ux + c1*uy + c2*uz ; u2 = ux*ux + uy*uy + uz*uz
 *   w * rho * (1.f + 3.f*uc + 4.5f*uc*uc - 1.5f*u2)                               */
static inline REAL FN(feq)(const REAL* w, int q, REAL ux, REAL uy, REAL uz, REAL rho)
{
    REAL uc = ORC_C[q][0] * ux + ORC_C[q][1] * uy + ORC_C[q][2] * uz;
    REAL u2 = ux * ux + uy * uy + uz * uz;
    return w[q] * rho * (1.f + 3.f * uc + 4.5f * uc * uc - 1.5f * u2);
}

/* One step of the Geier-2015 central-moment transform along one axis for the
 * (m, z, p) triple of a line (CollisionCumD3Q27Turbulent2015.h:86-114 z-pass,
 * :118-146 y-pass, :150-178 x-pass; all three passes use this form). */
static inline void FN(fwd)(REAL m, REAL z, REAL p, REAL u, REAL* o0, REAL* o1, REAL* o2)
{
    REAL k0 = (p + m) + z;
    *o0 = k0;
    *o1 = (p - m) - u * k0;
    *o2 = (p + m) - 2.f * u * (p - m) + u * u * k0;
}

/* Inverse transform of one (order 0,1,2) triple back to (z, m, p)
 * (CollisionCumD3Q27Turbulent2015.h:389-419 x-pass, :422-452 y-pass,
 * :455-485 z-pass). */
static inline void FN(inv)(REAL a0, REAL a1, REAL a2, REAL u, REAL* vz, REAL* vm, REAL* vp)
{
    *vz = a0 * (1.f - u * u) - 2.f * u * a1 - a2;
    *vm = (a0 * (u * u - u) + a1 * (2.f * u - 1.f) + a2) * 0.5f;
    *vp = (a0 * (u * u + u) + a1 * (2.f * u + 1.f) + a2) * 0.5f;
}

/* Variants of the cumulant operator.  All share the central-moment and
 * cumulant transforms (identical text in the three reference files) and
 * differ in the relaxation block. */
enum {
    FN(V2015) = 0,      /* CollisionCumD3Q27Turbulent2015.h:32-516                    */
    FN(V2017) = 1,      /* CollisionCumD3Q27Turbulent2017.h:32-570                    */
    FN(V2017_COMB) = 2  /* CollisionCumD3Q27TurbulentCombined.h:32-573 (cum2017 lambda) */
};

/* Cumulant collision of one cell in the reference's operation order.
 * 2015: omega1 = local omega, omega2..omega10 = 1 (Turbulent2015.h:240-249).
 * 2017: omega1 = local omega, omega2 = 1, omega3..5 from omega1/omega2 with
 * the magnitude limiters, omega6..10 = 1, and the A/B fourth-order terms
 * (Turbulent2017.h:238-370).  Quirks (applied when the flag is set):
 *  - lim210: the 210+012 and 210-012 limiters read |C_120 +- C_102|
 *    (Turbulent2017.h:254-255, 264-265); corrected: |C_210 +- C_012|;
 *  - trace17: the bulk term of D_x u subtracts (rho - 1) in Turbulent2017.h:295
 *    and (rho + 1) in Combined.h:798 instead of the zeroth moment k_000
 *    that Turbulent2015.h:253 subtracts (the trace of the raw central second
 *    moments is k_000 at equilibrium, so only k_000 leaves the rest state
 *    fixed).
 * f/out in reference direction order. */
static void FN(cumulant)(int variant, int lim210, int trace17, int eq44, const FN(ck) * K_, const REAL* f, REAL rho, REAL ux,
                         REAL uy, REAL uz, REAL omg, REAL* out)
{
    const REAL c13 = K_->c13;
    REAL F[3][3][3], A[3][3][3], B[3][3][3], k[3][3][3];
    int a, b, r, q;
    for (q = 0; q < 27; ++q) F[ORC_C[q][0] + 1][ORC_C[q][1] + 1][ORC_C[q][2] + 1] = f[q];

    /* central moments: z, then y, then x pass (Eq. G2015 43-45) */
    for (a = 0; a < 3; ++a)
        for (b = 0; b < 3; ++b)
            FN(fwd)(F[a][b][0], F[a][b][1], F[a][b][2], uz, &A[a][b][0], &A[a][b][1], &A[a][b][2]);
    for (a = 0; a < 3; ++a)
        for (r = 0; r < 3; ++r)
            FN(fwd)(A[a][0][r], A[a][1][r], A[a][2][r], uy, &B[a][0][r], &B[a][1][r], &B[a][2][r]);
    for (b = 0; b < 3; ++b)
        for (r = 0; r < 3; ++r)
            FN(fwd)(B[0][b][r], B[1][b][r], B[2][b][r], ux, &k[0][b][r], &k[1][b][r], &k[2][b][r]);

#define K(i, j, l) k[i][j][l]
    const REAL k_000 = K(0,0,0);
    /* cumulants, Eq. G2015 47-54 (:188-232) */
    const REAL C_110 = K(1,1,0), C_101 = K(1,0,1), C_011 = K(0,1,1);
    const REAL C_200 = K(2,0,0), C_020 = K(0,2,0), C_002 = K(0,0,2);
    const REAL C_120 = K(1,2,0), C_012 = K(0,1,2), C_201 = K(2,0,1);
    const REAL C_210 = K(2,1,0), C_021 = K(0,2,1), C_102 = K(1,0,2);
    const REAL C_111 = K(1,1,1);
    const REAL C_211 = K(2,1,1) - (K(2,0,0) * K(0,1,1) + 2.f * K(1,0,1) * K(1,1,0)) / rho;
    const REAL C_121 = K(1,2,1) - (K(0,2,0) * K(1,0,1) + 2.f * K(1,1,0) * K(0,1,1)) / rho;
    const REAL C_112 = K(1,1,2) - (K(0,0,2) * K(1,1,0) + 2.f * K(0,1,1) * K(1,0,1)) / rho;
    const REAL C_220 = K(2,2,0) - (K(0,2,0) * K(2,0,0) + 2.f * K(1,1,0) * K(1,1,0)) / rho;
    const REAL C_022 = K(0,2,2) - (K(0,0,2) * K(0,2,0) + 2.f * K(0,1,1) * K(0,1,1)) / rho;
    const REAL C_202 = K(2,0,2) - (K(2,0,0) * K(0,0,2) + 2.f * K(1,0,1) * K(1,0,1)) / rho;
    const REAL C_122 = K(1,2,2) - (K(0,2,0) * K(1,0,2) + K(0,0,2) * K(1,2,0) + 4.f * K(0,1,1) * K(1,1,1) +
                                   2.f * (K(1,1,0) * K(0,1,2) + K(1,0,1) * K(0,2,1))) / rho;
    const REAL C_212 = K(2,1,2) - (K(0,0,2) * K(2,1,0) + K(2,0,0) * K(0,1,2) + 4.f * K(1,0,1) * K(1,1,1) +
                                   2.f * (K(0,1,1) * K(2,0,1) + K(1,1,0) * K(1,0,2))) / rho;
    const REAL C_221 = K(2,2,1) - (K(2,0,0) * K(0,2,1) + K(0,2,0) * K(2,0,1) + 4.f * K(1,1,0) * K(1,1,1) +
                                   2.f * (K(1,0,1) * K(1,2,0) + K(0,1,1) * K(2,1,0))) / rho;
    const REAL C_222 = K(2,2,2) - (4.f * K(1,1,1) * K(1,1,1) + K(2,0,0) * K(0,2,2) + K(0,2,0) * K(2,0,2) +
                                   K(0,0,2) * K(2,2,0) +
                                   4.f * (K(0,1,1) * K(2,1,1) + K(1,0,1) * K(1,2,1) + K(1,1,0) * K(1,1,2)) +
                                   2.f * (K(1,2,0) * K(1,0,2) + K(2,1,0) * K(0,1,2) + K(2,0,1) * K(0,2,1))) / rho +
                       (16.f * K(1,1,0) * K(1,0,1) * K(0,1,1) +
                        4.f * (K(1,0,1) * K(1,0,1) * K(0,2,0) + K(0,1,1) * K(0,1,1) * K(2,0,0) +
                               K(1,1,0) * K(1,1,0) * K(0,0,2)) +
                        2.f * K(2,0,0) * K(0,2,0) * K(0,0,2)) / rho / rho;

    REAL Dxu, Dyv, Dzw, Cs_110, Cs_101, Cs_011, Cs_200, Cs_020, Cs_002;
    REAL Cs_120, Cs_102, Cs_210, Cs_012, Cs_021, Cs_201, Cs_111;
    REAL Cs_220, Cs_202, Cs_022, Cs_211, Cs_121, Cs_112, Cs_221, Cs_212, Cs_122, Cs_222;
    if (variant == FN(V2015)) {
        /* relaxation rates (:240-249) */
        const REAL omega1 = omg;
        const REAL omega2 = 1.f, omega3 = 1.f, omega4 = 1.f, omega5 = 1.f, omega6 = 1.f;
        const REAL omega7 = 1.f, omega8 = 1.f, omega9 = 1.f, omega10 = 1.f;

        /* Eq. G2015 58-60 (:252-259) */
        Dxu = -omega1 * 0.5f / rho * (2.f * C_200 - C_020 - C_002) -
              omega2 * 0.5f / rho * (C_200 + C_020 + C_002 - k_000);
        Dyv = Dxu + 3.f * omega1 * 0.5f / rho * (C_200 - C_020);
        Dzw = Dxu + 3.f * omega1 * 0.5f / rho * (C_200 - C_002);

        /* Eq. G2015 55-57, 61-73 (:264-316) */
        Cs_110 = (1.f - omega1) * C_110;
        Cs_101 = (1.f - omega1) * C_101;
        Cs_011 = (1.f - omega1) * C_011;
        const REAL Eq61 = (1.f - omega1) * (C_200 - C_020) -
                          3.f * rho * (1.f - omega1 * 0.5f) * (ux * ux * Dxu - uy * uy * Dyv);
        const REAL Eq64 = (1.f - omega1) * (C_200 - C_002) -
                          3.f * rho * (1.f - omega1 * 0.5f) * (ux * ux * Dxu - uz * uz * Dzw);
        const REAL Eq65 = k_000 * omega2 + (1.f - omega2) * (C_200 + C_020 + C_002) -
                          3.f * rho * (1.f - omega2 / 2.f) * (ux * ux * Dxu + uy * uy * Dyv + uz * uz * Dzw);
        Cs_200 = c13 * (Eq61 + Eq64 + Eq65);
        Cs_020 = c13 * (Eq64 - 2.f * Eq61 + Eq65);
        Cs_002 = c13 * (Eq61 - 2.f * Eq64 + Eq65);
        Cs_120 = (-C_102 - C_120) * omega3 * 0.5f + (C_102 - C_120) * omega4 * 0.5f + C_120;
        Cs_102 = (-C_102 - C_120) * omega3 * 0.5f + (-C_102 + C_120) * omega4 * 0.5f + C_102;
        Cs_210 = (-C_012 - C_210) * omega3 * 0.5f + (C_012 - C_210) * omega4 * 0.5f + C_210;
        Cs_012 = (-C_012 - C_210) * omega3 * 0.5f + (-C_012 + C_210) * omega4 * 0.5f + C_012;
        Cs_021 = (-C_021 - C_201) * omega3 * 0.5f + (-C_021 + C_201) * omega4 * 0.5f + C_021;
        Cs_201 = (-C_021 - C_201) * omega3 * 0.5f + (C_021 - C_201) * omega4 * 0.5f + C_201;
        Cs_111 = (1.f - omega5) * C_111;
        const REAL Eq71 = (1.f - omega6) * (C_220 - 2.f * C_202 + C_022);
        const REAL Eq72 = (1.f - omega6) * (C_220 + C_202 - 2.f * C_022);
        const REAL Eq73 = (1.f - omega7) * (C_220 + C_202 + C_022);
        Cs_220 = c13 * (Eq71 + Eq72 + Eq73);
        Cs_202 = c13 * (-Eq71 + Eq73);
        Cs_022 = c13 * (-Eq72 + Eq73);
        Cs_211 = (1.f - omega8) * C_211;
        Cs_121 = (1.f - omega8) * C_121;
        Cs_112 = (1.f - omega8) * C_112;
        Cs_221 = (1.f - omega9) * C_221;
        Cs_212 = (1.f - omega9) * C_212;
        Cs_122 = (1.f - omega9) * C_122;
        Cs_222 = (1.f - omega10) * C_222;
    } else {
        /* relaxation rates, Geier 2017 (Turbulent2017.h:240-291) */
        const REAL omega1 = omg;
        const REAL omega2 = 1.f;
        const REAL lambda3 = K_->lam, lambda4 = K_->lam, lambda5 = K_->lam;
        const REAL omega3 = 8.f * (omega1 - 2.f) * (omega2 * (3.f * omega1 - 1.f) - 5.f * omega1) /
                            (8.f * (5.f - 2.f * omega1) * omega1 + omega2 * (8.f + omega1 * (9.f * omega1 - 26.f)));
        /* limiters: omega + (1 - omega) |C| / (rho lambda + |C|) */
        const REAL a210p = lim210 ? C_120 + C_102 : C_210 + C_012;
        const REAL a210m = lim210 ? C_120 - C_102 : C_210 - C_012;
        const REAL omega120p102 =
            omega3 + (1.f - omega3) * ORC_FABS(C_120 + C_102) / (rho * lambda3 + ORC_FABS(C_120 + C_102));
        const REAL omega210p012 = omega3 + (1.f - omega3) * ORC_FABS(a210p) / (rho * lambda3 + ORC_FABS(a210p));
        const REAL omega201p021 =
            omega3 + (1.f - omega3) * ORC_FABS(C_201 + C_021) / (rho * lambda3 + ORC_FABS(C_201 + C_021));
        const REAL omega4 = 8.f * (omega1 - 2.f) * (omega1 + omega2 * (3.f * omega1 - 7.f)) /
                            (omega2 * (56.f - 42.f * omega1 + K_->w4q * omega1 * omega1) - 8.f * omega1);
        const REAL omega120m102 =
            omega4 + (1.f - omega4) * ORC_FABS(C_120 - C_102) / (rho * lambda4 + ORC_FABS(C_120 - C_102));
        const REAL omega210m012 = omega4 + (1.f - omega4) * ORC_FABS(a210m) / (rho * lambda4 + ORC_FABS(a210m));
        const REAL omega201m021 =
            omega4 + (1.f - omega4) * ORC_FABS(C_201 - C_021) / (rho * lambda4 + ORC_FABS(C_201 - C_021));
        const REAL omega5 = 24.f * (omega1 - 2.f) *
                            (4.f * omega1 * omega1 + omega1 * omega2 * (18.f - 13.f * omega1) +
                             omega2 * omega2 * (2.f + omega1 * (6.f * omega1 - 11.f))) /
                            (16.f * omega1 * omega1 * (omega1 - 6.f) -
                             2.f * omega1 * omega2 * (216.f + 5.f * omega1 * (9.f * omega1 - 46.f)) +
                             omega2 * omega2 * (omega1 * (3.f * omega1 - 10.f) * (15.f * omega1 - 28.f) - 48.f));
        const REAL omega111 = omega5 + (1.f - omega5) * ORC_FABS(C_111) / (rho * lambda5 + ORC_FABS(C_111));
        const REAL omega6 = 1.f, omega7 = 1.f, omega8 = 1.f, omega9 = 1.f, omega10 = 1.f;
        const REAL A = (g_plant & ORC_PLANT_NO_A) ? (REAL)0 :
                       ((g_plant & ORC_PLANT_A_NEG) ? -1 : 1) *
                       (4.f * omega1 * omega1 + 2.f * omega1 * omega2 * (omega1 - 6.f) +
                        omega2 * omega2 * (omega1 * (10.f - 3.f * omega1) - 4.f)) /
                       (omega1 - omega2) / (omega2 * (2.f + 3.f * omega1) - 8.f * omega1);
        const REAL B = (g_plant & ORC_PLANT_NO_B) ? (REAL)0 :
                       (4.f * omega1 * omega2 * (9.f * omega1 - 16.f) - 4.f * omega1 * omega1 -
                        2.f * omega2 * omega2 * (2.f + 9.f * omega1 * (omega1 - 2.f))) /
                       3.f / (omega1 - omega2) / (omega2 * (2.f + 3.f * omega1) - 8.f * omega1);

        /* Eq. G2015 58-60 with the 2017 bulk term (Turbulent2017.h:294-301,
         * Combined.h:797-798) and the off-diagonal rates (:304-306) */
        REAL tr;
        if (!trace17) tr = k_000;
        else if (variant == FN(V2017)) tr = (-1 + rho);
        else tr = -(-1 - rho);
        Dxu = -omega1 * 0.5f / rho * (2.f * C_200 - C_020 - C_002) - omega2 * 0.5f / rho * (C_200 + C_020 + C_002 - tr);
        Dyv = Dxu + 3.f * omega1 * 0.5f / rho * (C_200 - C_020);
        Dzw = Dxu + 3.f * omega1 * 0.5f / rho * (C_200 - C_002);
        const REAL DxvDyu = -3.f * omega1 / rho * C_110;
        const REAL DxwDzu = -3.f * omega1 / rho * C_101;
        const REAL DywDzv = -3.f * omega1 / rho * C_011;

        /* second order as in 2015 (:310-329) */
        Cs_110 = (1.f - omega1) * C_110;
        Cs_101 = (1.f - omega1) * C_101;
        Cs_011 = (1.f - omega1) * C_011;
        const REAL Eq61 = (1.f - omega1) * (C_200 - C_020) -
                          3.f * rho * (1.f - omega1 * 0.5f) * (ux * ux * Dxu - uy * uy * Dyv);
        const REAL Eq64 = (1.f - omega1) * (C_200 - C_002) -
                          3.f * rho * (1.f - omega1 * 0.5f) * (ux * ux * Dxu - uz * uz * Dzw);
        const REAL Eq65 = k_000 * omega2 + (1.f - omega2) * (C_200 + C_020 + C_002) -
                          3.f * rho * (1.f - omega2 / 2.f) * (ux * ux * Dxu + uy * uy * Dyv + uz * uz * Dzw);
        Cs_200 = c13 * (Eq61 + Eq64 + Eq65);
        Cs_020 = c13 * (Eq64 - 2.f * Eq61 + Eq65);
        Cs_002 = c13 * (Eq61 - 2.f * Eq64 + Eq65);

        /* third order, Eq. G2017 117-122 and 42 (:331-346) */
        const REAL Eq117 = (1.f - omega120p102) * (C_120 + C_102);
        const REAL Eq118 = (1.f - omega210p012) * (C_210 + C_012);
        const REAL Eq119 = (1.f - omega201p021) * (C_201 + C_021);
        const REAL Eq120 = (1.f - omega120m102) * (C_120 - C_102);
        const REAL Eq121 = (1.f - omega210m012) * (C_210 - C_012);
        const REAL Eq122 = (1.f - omega201m021) * (C_201 - C_021);
        Cs_120 = 0.5f * (Eq120 + Eq117);
        Cs_102 = 0.5f * (-Eq120 + Eq117);
        Cs_210 = 0.5f * (Eq121 + Eq118);
        Cs_012 = 0.5f * (-Eq121 + Eq118);
        Cs_021 = 0.5f * (-Eq122 + Eq119);
        Cs_201 = 0.5f * (Eq122 + Eq119);
        Cs_111 = (1.f - omega111) * C_111;

        /* fourth order, Eq. G2017 43-48 (:349-365) */
        const REAL Eq43 = K_->c23 * (1.f / omega1 - 0.5f) * omega6 * A * rho * (Dxu - 2.f * Dyv + Dzw) +
                          (1.f - omega6) * (C_220 - 2.f * C_202 + C_022);
        /* Eq. 44 as the reference writes it (LBM_COMPAT_EQ44), or with the strain
         * combination that Eq. 43's axis convention gives (-2 Dxu + Dyv + Dzw) */
        const REAL Eq44 = K_->c23 * (1.f / omega1 - 0.5f) * omega6 * A * rho *
                              (eq44 ? (Dxu + Dyv - 2.f * Dzw) : (Dyv + Dzw - 2.f * Dxu)) +
                          (1.f - omega6) * (C_220 + C_202 - 2.f * C_022);
        const REAL Eq45 = -K_->c43 * (1.f / omega1 - 0.5f) * omega7 * A * rho * (Dxu + Dyv + Dzw) +
                          (1.f - omega7) * (C_220 + C_202 + C_022);
        Cs_220 = c13 * (Eq43 + Eq44 + Eq45);
        Cs_202 = c13 * (-Eq43 + Eq45);
        Cs_022 = c13 * (-Eq44 + Eq45);
        Cs_211 = -c13 * (1.f / omega1 - 0.5f) * omega8 * B * rho * DywDzv + (1.f - omega8) * C_211;
        Cs_121 = -c13 * (1.f / omega1 - 0.5f) * omega8 * B * rho * DxwDzu + (1.f - omega8) * C_121;
        Cs_112 = -c13 * (1.f / omega1 - 0.5f) * omega8 * B * rho * DxvDyu + (1.f - omega8) * C_112;

        /* fifth and sixth order (:367-370) */
        Cs_221 = (1.f - omega9) * C_221;
        Cs_212 = (1.f - omega9) * C_212;
        Cs_122 = (1.f - omega9) * C_122;
        Cs_222 = (1.f - omega10) * C_222;
    }
#undef K

    /* back to central moments, Eq. G2015 81-87 (:322-380) */
    REAL s[3][3][3];
#define S(i, j, l) s[i][j][l]
    S(0,0,0) = k_000;
    S(1,1,0) = Cs_110; S(1,0,1) = Cs_101; S(0,1,1) = Cs_011;
    S(2,0,0) = Cs_200; S(0,2,0) = Cs_020; S(0,0,2) = Cs_002;
    S(1,2,0) = Cs_120; S(0,1,2) = Cs_012; S(2,0,1) = Cs_201;
    S(2,1,0) = Cs_210; S(0,2,1) = Cs_021; S(1,0,2) = Cs_102;
    S(1,1,1) = Cs_111;
    S(1,0,0) = -k[1][0][0]; S(0,1,0) = -k[0][1][0]; S(0,0,1) = -k[0][0][1];
    S(2,1,1) = Cs_211 + (S(2,0,0) * S(0,1,1) + 2.f * S(1,0,1) * S(1,1,0)) / rho;
    S(1,2,1) = Cs_121 + (S(0,2,0) * S(1,0,1) + 2.f * S(1,1,0) * S(0,1,1)) / rho;
    S(1,1,2) = Cs_112 + (S(0,0,2) * S(1,1,0) + 2.f * S(0,1,1) * S(1,0,1)) / rho;
    S(2,2,0) = Cs_220 + (S(0,2,0) * S(2,0,0) + 2.f * S(1,1,0) * S(1,1,0)) / rho;
    S(0,2,2) = Cs_022 + (S(0,0,2) * S(0,2,0) + 2.f * S(0,1,1) * S(0,1,1)) / rho;
    S(2,0,2) = Cs_202 + (S(2,0,0) * S(0,0,2) + 2.f * S(1,0,1) * S(1,0,1)) / rho;
    S(1,2,2) = Cs_122 + (S(0,2,0) * S(1,0,2) + S(0,0,2) * S(1,2,0) + 4.f * S(0,1,1) * S(1,1,1) +
                         2.f * (S(1,1,0) * S(0,1,2) + S(1,0,1) * S(0,2,1))) / rho;
    S(2,1,2) = Cs_212 + (S(0,0,2) * S(2,1,0) + S(2,0,0) * S(0,1,2) + 4.f * S(1,0,1) * S(1,1,1) +
                         2.f * (S(0,1,1) * S(2,0,1) + S(1,1,0) * S(1,0,2))) / rho;
    S(2,2,1) = Cs_221 + (S(2,0,0) * S(0,2,1) + S(0,2,0) * S(2,0,1) +
                         ((g_plant & ORC_PLANT_BACK_221) ? 2.f : 4.f) * S(1,1,0) * S(1,1,1) +
                         2.f * (S(1,0,1) * S(1,2,0) + S(0,1,1) * S(2,1,0))) / rho;
    S(2,2,2) = Cs_222 +
               (4.f * S(1,1,1) * S(1,1,1) + S(2,0,0) * S(0,2,2) + S(0,2,0) * S(2,0,2) + S(0,0,2) * S(2,2,0) +
                4.f * (S(0,1,1) * S(2,1,1) + S(1,0,1) * S(1,2,1) + S(1,1,0) * S(1,1,2)) +
                2.f * (S(1,2,0) * S(1,0,2) + S(2,1,0) * S(0,1,2) + S(2,0,1) * S(0,2,1))) / rho -
               (16.f * S(1,1,0) * S(1,0,1) * S(0,1,1) +
                4.f * (S(1,0,1) * S(1,0,1) * S(0,2,0) + S(0,1,1) * S(0,1,1) * S(2,0,0) +
                       S(1,1,0) * S(1,1,0) * S(0,0,2)) +
                2.f * S(2,0,0) * S(0,2,0) * S(0,0,2)) / rho / rho;
#undef S

    /* back to populations: x, y, z pass (Eq. G2015 88-96) */
    for (b = 0; b < 3; ++b)
        for (r = 0; r < 3; ++r)
            FN(inv)(s[0][b][r], s[1][b][r], s[2][b][r], ux, &A[1][b][r], &A[0][b][r], &A[2][b][r]);
    for (a = 0; a < 3; ++a)
        for (r = 0; r < 3; ++r)
            FN(inv)(A[a][0][r], A[a][1][r], A[a][2][r], uy, &B[a][1][r], &B[a][0][r], &B[a][2][r]);
    for (a = 0; a < 3; ++a)
        for (b = 0; b < 3; ++b)
            FN(inv)(B[a][b][0], B[a][b][1], B[a][b][2], uz, &F[a][b][1], &F[a][b][0], &F[a][b][2]);
    for (q = 0; q < 27; ++q) out[q] = F[ORC_C[q][0] + 1][ORC_C[q][1] + 1][ORC_C[q][2] + 1];
}

/* BGK: f - (f - f_eq)*omega (CollisionSRTTurbulent.h:56-75, CollisionSRTLaminar.h) */
static void FN(srt)(const REAL* w, const REAL* f, REAL rho, REAL ux, REAL uy, REAL uz, REAL omg, REAL* out)
{
    for (int q = 0; q < 27; ++q) out[q] = f[q] - (f[q] - FN(feq)(w, q, ux, uy, uz, rho)) * omg;
}

/* Smagorinsky relaxation rate (OmegaLES.h:69-131). */
static REAL FN(omega_les)(const REAL* w, const REAL* f, REAL rho, REAL u_x, REAL u_y, REAL u_z, REAL tau,
                          REAL CLES)
{
    REAL pi[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int vel = 0; vel < 27; vel++) {
        REAL feq = FN(feq)(w, vel, u_x, u_y, u_z, rho);
        REAL df = f[vel];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) pi[a][b] += ORC_C[vel][a] * ORC_C[vel][b] * (df - feq);
    }
    REAL PI = pi[0][0] * pi[0][0];
    PI += pi[0][1] * pi[0][1];
    PI += pi[0][2] * pi[0][2];
    PI += pi[1][0] * pi[1][0];
    PI += pi[1][1] * pi[1][1];
    PI += pi[1][2] * pi[1][2];
    PI += pi[2][0] * pi[2][0];
    PI += pi[2][1] * pi[2][1];
    PI += pi[2][2] * pi[2][2];
    REAL rho_inv = 1.0f / rho;
    REAL tau_sq = tau * tau;
    REAL CLES_term = 18.f * CLES * rho_inv;
    PI = ORC_SQRT(PI);
    REAL tauLES = tau * 0.5f + 0.5f * ORC_SQRT(tau_sq + CLES_term * PI);
    return 1.f / tauLES;
}

static inline long FN(idx)(const FN(state) * s, int x, int y, int z)
{
    return (long)x + (long)s->nx * ((long)y + (long)s->ny * (long)z);
}

/* Neighbour coordinate with optional index wrap; returns 0 when outside the
 * array along a non-wrapped axis. */
static inline int FN(nb)(const FN(state) * s, int x, int y, int z, long* out)
{
    if (x < 0 || x >= s->nx) { if (!s->per[0]) return 0; x = (x + s->nx) % s->nx; }
    if (y < 0 || y >= s->ny) { if (!s->per[1]) return 0; y = (y + s->ny) % s->ny; }
    if (z < 0 || z >= s->nz) { if (!s->per[2]) return 0; z = (z + s->nz) % s->nz; }
    *out = FN(idx)(s, x, y, z);
    return 1;
}

#define DF(s, q, i) ((s)->df[(long)(q) * (s)->N + (i)])
#define DP(s, q, i) ((s)->df_post[(long)(q) * (s)->N + (i)])

static void FN(load)(const FN(state) * s, const REAL* arr, long i, REAL* f)
{
    for (int q = 0; q < 27; ++q) f[q] = arr[(long)q * s->N + i];
}

/* --- operators, one per reference policy --- */

/* OmegaLES::omega / OmegaNo::omega (OmegaLES.h:134-136, OmegaNo.h:34-47) */
static void FN(op_omega)(FN(state) * s)
{
    long N = s->N;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
        if (s->flags[i] == 0) continue;
        if (s->closure == LBM_CLOSURE_SMAGORINSKY) {
            REAL f[27];
            FN(load)(s, s->df, i, f);
            s->omega[i] = FN(omega_les)(s->w, f, s->rho[i], s->u[3 * i], s->u[3 * i + 1], s->u[3 * i + 2], s->tau,
                                        s->CLES);
        } else {
            s->omega[i] = 1.f / s->tau;
        }
    }
}

/* One cell of the configured operator.  The Combined operator collides
 * cells with flag 1 (fluid, no solid neighbour) with the 2015 lambda and every
 * other non-solid cell with its 2017 lambda
 * (CollisionCumD3Q27TurbulentCombined.h:1066-1078). */
static void FN(collide_one)(int collision, uint32_t compat, const REAL* w, const FN(ck) * ck, int flag,
                            const REAL* f, REAL rho, REAL ux, REAL uy, REAL uz, REAL omg, REAL* o)
{
    const int lim = (compat & LBM_COMPAT_LIM210) != 0, tr = (compat & LBM_COMPAT_TRACE17) != 0;
    const int e44 = (compat & LBM_COMPAT_EQ44) != 0;
    if (collision == LBM_COLLISION_SRT)
        FN(srt)(w, f, rho, ux, uy, uz, omg, o);
    else if (collision == LBM_COLLISION_CUM2015 || (collision == LBM_COLLISION_COMBINED && flag == 1))
        FN(cumulant)(FN(V2015), lim, tr, e44, ck, f, rho, ux, uy, uz, omg, o);
    else if (collision == LBM_COLLISION_CUM2017)
        FN(cumulant)(FN(V2017), lim, tr, e44, ck, f, rho, ux, uy, uz, omg, o);
    else
        FN(cumulant)(FN(V2017_COMB), lim, tr, e44, ck, f, rho, ux, uy, uz, omg, o);
}

/* Collision*::collision (Cum2015.h:521-536, Cum2017.h:575-590, Combined.h:1066-1086,
 * CollisionSRTTurbulent.h:56-80) */
static void FN(op_collision)(FN(state) * s)
{
    long N = s->N;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
        if (s->flags[i] == 0) continue;
        REAL f[27], o[27];
        FN(load)(s, s->df, i, f);
        FN(collide_one)(s->collision, s->compat, s->w, &s->ck, s->flags[i], f, s->rho[i], s->u[3 * i],
                        s->u[3 * i + 1], s->u[3 * i + 2], s->omega[i], o);
        for (int q = 0; q < 27; ++q) DP(s, q, i) = o[q];
    }
}

/* StreamingAB::streaming (StreamingAB.h:38-68): pull from x - c_i. */
static void FN(op_streaming)(FN(state) * s)
{
#pragma omp parallel for schedule(static)
    for (int z = 0; z < s->nz; ++z)
        for (int y = 0; y < s->ny; ++y)
            for (int x = 0; x < s->nx; ++x) {
                long i = FN(idx)(s, x, y, z);
                int mv = s->flags[i];
                if (!(mv > 0 || mv == -2 || mv == -3)) continue;
                for (int vel = 0; vel < 27; ++vel) {
                    long j;
                    if (!FN(nb)(s, x - ORC_C[vel][0], y - ORC_C[vel][1], z - ORC_C[vel][2], &j)) continue;
                    int nv = s->flags[j];
                    if (nv > 0 || nv == -2 || nv == -3)
                        DF(s, vel, i) = DP(s, vel, j);
                    else if (nv == -1)
                        DF(s, vel, i) = DF(s, vel, j);
                }
            }
}

/* BounceBackWallHalf::bounceBackWall (BounceBackWallHalf.h:29-52) */
static void FN(op_bounce_back)(FN(state) * s)
{
    if (s->wall_mesh) {
        /* BounceBackWallHalfMesh::bounceBackWall (BounceBackWallHalfMesh.h:38-50):
         * every cell whose flag is 2 */
        for (int z = 0; z < s->nz; ++z)
            for (int y = 0; y < s->ny; ++y)
                for (int x = 0; x < s->nx; ++x) {
                    long i = FN(idx)(s, x, y, z);
                    if (s->flags[i] != 2) continue;
                    for (int vel = 0; vel < 27; ++vel) {
                        long j;
                        if (!FN(nb)(s, x - ORC_C[vel][0], y - ORC_C[vel][1], z - ORC_C[vel][2], &j)) continue;
                        if (s->flags[j] == 0) DF(s, vel, i) = DP(s, s->crev[vel], i);
                    }
                }
        return;
    }
    for (int n = 0; n < s->n_bc; ++n) {
        const lbm_bc_node* b = &s->bc[n];
        if (b->kind != LBM_BC_WALL) continue;
        long i = FN(idx)(s, b->x, b->y, b->z);
        for (int vel = 0; vel < 27; ++vel) {
            long j;
            if (!FN(nb)(s, b->x - ORC_C[vel][0], b->y - ORC_C[vel][1], b->z - ORC_C[vel][2], &j)) continue;
            if (s->flags[j] == 0) DF(s, vel, i) = DP(s, s->crev[vel], i);
        }
    }
}

/* InletVelocityMovingWall::inlet (InletVelocityMovingWall.h:31-53) and
 * InletVelocityEquilibrium::inlet (InletVelocityEquilibrium.h:51-103). */
static void FN(op_inlet)(FN(state) * s)
{
    for (int n = 0; n < s->n_bc; ++n) {
        const lbm_bc_node* b = &s->bc[n];
        long i = FN(idx)(s, b->x, b->y, b->z);
        if (b->kind == LBM_BC_INLET_MOVING_WALL) {
            for (int vel = 0; vel < 27; ++vel) {
                if (b->nx * ORC_C[vel][0] + b->ny * ORC_C[vel][1] + b->nz * ORC_C[vel][2] > 0) {
                    DF(s, s->crev[vel], i) =
                        DP(s, vel, i) - 6 * s->w[vel] * s->rho[i] *
                                            (ORC_C[vel][0] * b->ux + ORC_C[vel][1] * b->uy + ORC_C[vel][2] * b->uz);
                }
            }
        } else if (b->kind == LBM_BC_INLET_ZOU_HE) {
            /* InletVelocityZouHe::inlet (InletVelocityZouHe.h:80-101): for every
             * outgoing v (n.c > 0), slot c_rev[v] = (f*(v) - f_eq(v; rho, u)) +
             * f_eq(c_rev[v]; rho, u_inlet), the inlet velocity's polynomial in float */
            for (int vel = 0; vel < 27; ++vel) {
                if (b->nx * ORC_C[vel][0] + b->ny * ORC_C[vel][1] + b->nz * ORC_C[vel][2] > 0) {
                    REAL fNeqRev = DP(s, vel, i) - FN(feq)(s->w, vel, s->u[3 * i], s->u[3 * i + 1],
                                                          s->u[3 * i + 2], s->rho[i]);
                    const int r = s->crev[vel];
                    float uc = ORC_C[r][0] * b->ux + ORC_C[r][1] * b->uy + ORC_C[r][2] * b->uz;
                    float u2 = b->ux * b->ux + b->uy * b->uy + b->uz * b->uz;
                    DF(s, r, i) = fNeqRev + s->w[r] * s->rho[i] * (1.f + 3.f * uc + 4.5f * uc * uc - 1.5f * u2);
                }
            }
        } else if (b->kind == LBM_BC_INLET_EQUILIBRIUM) {
            /* Reference (LBM_COMPAT_INLETRHO): rho_b = rho_1 + nu/3 (dux1 - duxb) on a +x
             * normal (:80), rho_1 + nu/3 (duxb - dux1) on a -x normal (:95).  Corrected:
             * both branches rho_1 + 3 nu (duxb - dux1), i.e. rho_b - rho_1 =
             * 3 nu (x_b - x_1) d2u/dx2 from d(rho/3)/dx = nu d2u/dx2. */
            const int ref = (s->compat & LBM_COMPAT_INLETRHO) != 0;
            const REAL cnu = (g_plant & ORC_PLANT_INLET_NU3) ? s->nu / 3 : 3 * s->nu;
            const REAL uxb = b->ux, uyb = b->uy, uzb = b->uz;
            long i1, i2;
            if (b->nx == 1.0f && b->ny == 0.0f && b->nz == 0.0f) {
                if (FN(nb)(s, b->x - 1, b->y, b->z, &i1) && FN(nb)(s, b->x - 2, b->y, b->z, &i2)) {
                    REAL ux1 = s->u[3 * i1], ux2 = s->u[3 * i2];
                    REAL duxb = (uxb - ux1);
                    REAL dux1 = (ux1 - ux2);
                    REAL rhob = ref ? s->rho[i1] + s->nu / 3 * (dux1 - duxb)
                                    : (g_plant & ORC_PLANT_INLET_SIGN) ? s->rho[i1] + cnu * (dux1 - duxb)
                                                                       : s->rho[i1] + cnu * (duxb - dux1);
                    for (int v = 0; v < 27; ++v) DF(s, v, i) = FN(feq)(s->w, v, uxb, uyb, uzb, rhob);
                }
            }
            if (b->nx == -1.0f && b->ny == 0.0f && b->nz == 0.0f) {
                if (FN(nb)(s, b->x + 1, b->y, b->z, &i1) && FN(nb)(s, b->x + 2, b->y, b->z, &i2)) {
                    REAL ux1 = s->u[3 * i1], ux2 = s->u[3 * i2];
                    REAL duxb = (ux1 - uxb);
                    REAL dux1 = (ux2 - ux1);
                    REAL rhob = ref ? s->rho[i1] + s->nu / 3 * (duxb - dux1) : s->rho[i1] + cnu * (duxb - dux1);
                    for (int v = 0; v < 27; ++v) DF(s, v, i) = FN(feq)(s->w, v, uxb, uyb, uzb, rhob);
                }
            }
        }
    }
}

/* OutletDensityInterpolated::outlet (OutletDensityInterpolated.h:170-252) */
/* f_equilibrium_outlet / f_equilibrium of the outlet variants
 * (OutletNeighbourEquilibrium.h:35-74): u of cell i, the given density. */
static inline REAL FN(feq_at)(const FN(state) * s, int vel, long i, REAL density)
{
    return FN(feq)(s->w, vel, s->u[3 * i], s->u[3 * i + 1], s->u[3 * i + 2], density);
}

/* OutletDensityEquilibrium::outlet (OutletDensityEquilibrium.h:55-76): for
 * every outgoing direction (n.c > 0) the slot c_rev[vel] receives the
 * equilibrium of vel at the target density and the node's velocity.
 * OutletNeighbourEquilibrium::outlet (OutletNeighbourEquilibrium.h:78-126),
 * ...Omega (:139-186), ...OmegaRF (:143-198): for every incoming direction
 * (n.c < 0), f = f_eq(rho_target, u(x)) + (f_eq(rho, u)(x + c) - f(x + c)),
 * at irregular nodes only where the neighbour's flag is > 0; the RF variant
 * lowers the target density by rho u_x^2 / 2 where u_x(x) < 0. */
static void FN(op_outlet_eq)(FN(state) * s, const lbm_bc_node* b)
{
    long i = FN(idx)(s, b->x, b->y, b->z);
    REAL density = b->rho;
    if (b->kind == LBM_BC_OUTLET_EQUILIBRIUM) {
        for (int vel = 0; vel < 27; ++vel)
            if (b->nx * ORC_C[vel][0] + b->ny * ORC_C[vel][1] + b->nz * ORC_C[vel][2] > 0)
                DF(s, s->crev[vel], i) = FN(feq_at)(s, vel, i, density);
        return;
    }
    if (b->kind == LBM_BC_OUTLET_NEIGHBOUR_EQ_RF && s->u[3 * i] < 0)
        density = density - 0.5 * s->u[3 * i] * s->u[3 * i] * density;
    for (int vel = 0; vel < 27; ++vel) {
        if (!(b->nx * ORC_C[vel][0] + b->ny * ORC_C[vel][1] + b->nz * ORC_C[vel][2] < 0)) continue;
        long j;
        /* the reference has no bounds checks; a neighbour outside the array is skipped */
        if (!FN(nb)(s, b->x + ORC_C[vel][0], b->y + ORC_C[vel][1], b->z + ORC_C[vel][2], &j)) continue;
        if (!b->regular && !(s->flags[j] > 0)) continue;
        REAL fEqNeigh = FN(feq_at)(s, vel, j, s->rho[j]);
        REAL fNeqNeigh = DF(s, vel, j);
        REAL fEqBC = FN(feq_at)(s, vel, i, density);
        DF(s, vel, i) = fEqBC + (fEqNeigh - fNeqNeigh);
    }
}

/* OutletNeighbourEquilibriumOmega::outletOmega (:18-79): omega decreases
 * linearly over the last 10 % of the lattice along the first outlet entry's
 * normal axis (at the high end of that axis whatever the normal's sign):
 * omega -= (high - low) / interval * position, position = 0 .. interval - 1. */
static void FN(op_outlet_omega)(FN(state) * s)
{
    if (!s->sponge) return;
    const lbm_bc_node* b0 = 0;
    for (int n = 0; n < s->n_bc && !b0; ++n)
        if (s->bc[n].kind >= LBM_BC_OUTLET_INTERPOLATED && s->bc[n].kind != LBM_BC_PERIODIC_GHOST) b0 = &s->bc[n];
    if (!b0) return;
    const float nx = b0->nx, ny = b0->ny, nz = b0->nz;
    const float px = fabsf(fabsf(nx) - 1), py = fabsf(fabsf(ny) - 1), pz = fabsf(fabsf(nz) - 1);
    const int dx01 = (int)ceilf(s->nx * 0.1f), dy01 = (int)ceilf(s->ny * 0.1f), dz01 = (int)ceilf(s->nz * 0.1f);
    const int dX = (int)(dx01 * fabsf(nx) + s->nx * fabsf(px));
    const int dY = (int)(dy01 * fabsf(ny) + s->ny * fabsf(py));
    const int dZ = (int)(dz01 * fabsf(nz) + s->nz * fabsf(pz));
    const int bX = s->nx - dX, bY = s->ny - dY, bZ = s->nz - dZ;
    const int interval = (int)fabsf(dx01 * nx + dy01 * ny + dz01 * nz);
    const REAL lo = s->sponge_low, hi = s->sponge_high;
    const REAL derivation = (hi - lo) / interval;
    for (int z = bZ; z < s->nz; ++z)
        for (int y = bY; y < s->ny; ++y)
            for (int x = bX; x < s->nx; ++x) {
                const int pos = (int)((x - bX) * nx * nx + (y - bY) * ny * ny + (z - bZ) * nz * nz);
                const REAL damp = derivation * pos;
                long i = FN(idx)(s, x, y, z);
                s->omega[i] = s->omega[i] - damp;
            }
}

static void FN(op_outlet)(FN(state) * s)
{
    for (int n = 0; n < s->n_bc; ++n) {
        const lbm_bc_node* b = &s->bc[n];
        if (b->kind == LBM_BC_OUTLET_EQUILIBRIUM || b->kind == LBM_BC_OUTLET_NEIGHBOUR_EQ ||
            b->kind == LBM_BC_OUTLET_NEIGHBOUR_EQ_RF) {
            FN(op_outlet_eq)(s, b);
            continue;
        }
        if (b->kind != LBM_BC_OUTLET_INTERPOLATED) continue;
        const REAL density = b->rho;
        long i = FN(idx)(s, b->x, b->y, b->z);
        int xp = (b->nx == 1.0f && b->ny == 0.0f && b->nz == 0.0f);
        int xm = (b->nx == -1.0f && b->ny == 0.0f && b->nz == 0.0f);
        if (!xp && !xm) continue;
        if (xp && !b->regular) {
            for (int vel = 0; vel < 27; ++vel) {
                long j;
                if (!FN(nb)(s, b->x, b->y - ORC_C[vel][1], b->z - ORC_C[vel][2], &j)) continue;
                if (s->flags[j] == 0) DF(s, vel, i) = DF(s, s->crev[vel], i);
            }
        }
        long iu;
        if (!FN(nb)(s, b->x + (xp ? -1 : 1), b->y, b->z, &iu)) continue;
        for (int vel = 0; vel < 27; ++vel) {
            if (b->nx * ORC_C[vel][0] + b->ny * ORC_C[vel][1] + b->nz * ORC_C[vel][2] < 0) {
                if (g_plant & ORC_PLANT_OUTLET_SWAP)
                    DF(s, vel, i) = (1 - s->cs) * DP(s, vel, iu) + s->cs * DP(s, vel, i);
                else
                    DF(s, vel, i) = s->cs * DP(s, vel, iu) + (1 - s->cs) * DP(s, vel, i);
            }
        }
        REAL d[27];
        FN(load)(s, s->df, i, d);
        REAL rho = d[0] + d[1] + d[2] + d[3] + d[4] + d[5] + d[6] + d[7] + d[8] + d[9] + d[10] + d[11] + d[12] +
                   d[13] + d[14] + d[15] + d[16] + d[17] + d[18] + d[19] + d[20] + d[21] + d[22] + d[23] + d[24] +
                   d[25] + d[26];
        REAL ux = (d[1] - d[2] + d[7] - d[8] + d[9] - d[10] - d[11] + d[12] - d[15] + d[16] - d[19] + d[20] -
                   d[21] + d[22] + d[23] - d[24] - d[25] + d[26]) / density;
        REAL uy = (-d[5] + d[6] - d[11] + d[12] + d[13] - d[14] + d[15] - d[16] + d[17] - d[18] + d[19] -
                   d[20] - d[21] + d[22] - d[23] + d[24] - d[25] + d[26]) / density;
        REAL uz = (-d[3] + d[4] - d[7] + d[8] + d[9] - d[10] - d[13] + d[14] + d[17] - d[18] - d[19] + d[20] +
                   d[21] - d[22] - d[23] + d[24] - d[25] + d[26]) / density;
        for (int vel = 0; vel < 27; ++vel)
            DF(s, vel, i) = DF(s, vel, i) - FN(feq)(s->w, vel, ux, uy, uz, rho) + FN(feq)(s->w, vel, ux, uy, uz, 1.f);
    }
}

/* Periodic::periodic (Periodic.h:43-138), ghost-layer periodicity. */
static void FN(op_periodic)(FN(state) * s)
{
    for (int n = 0; n < s->n_bc; ++n) {
        const lbm_bc_node* b = &s->bc[n];
        if (b->kind != LBM_BC_PERIODIC_GHOST) continue;
        long i = FN(idx)(s, b->x, b->y, b->z);
        float ax = fabsf(b->nx), ay = fabsf(b->ny), az = fabsf(b->nz);
        int axis = (ax == 1.0f && ay == 0.0f && az == 0.0f) ? 0
                 : (ax == 0.0f && ay == 1.0f && az == 0.0f) ? 1
                 : (ax == 0.0f && ay == 0.0f && az == 1.0f) ? 2 : -1;
        if (axis < 0) continue;
        long ip = axis == 0 ? FN(idx)(s, b->per_idx, b->y, b->z)
                : axis == 1 ? FN(idx)(s, b->x, b->per_idx, b->z)
                            : FN(idx)(s, b->x, b->y, b->per_idx);
        for (int vel = 0; vel < 27; ++vel) {
            float nc = b->nx * ORC_C[vel][0] + b->ny * ORC_C[vel][1] + b->nz * ORC_C[vel][2];
            if (nc < 0) DF(s, vel, i) = DF(s, vel, ip);
            if (!b->regular && nc <= 0) {
                long j;
                if (FN(nb)(s, b->x + ORC_C[vel][0], b->y + ORC_C[vel][1], b->z + ORC_C[vel][2], &j) &&
                    s->flags[j] == 0)
                    DF(s, s->crev[vel], i) = DP(s, vel, i);
            }
        }
    }
}

/* MomentDensityVelocityN27::momentUpdate (MomentDensityVelocityN27.h:27-83) */
static void FN(moments_cell)(const REAL* d, REAL* rho, REAL* ux, REAL* uy, REAL* uz)
{
    REAL r = d[0] + d[1] + d[2] + d[3] + d[4] + d[5] + d[6] + d[7] + d[8] + d[9] + d[10] + d[11] + d[12] + d[13] +
             d[14] + d[15] + d[16] + d[17] + d[18] + d[19] + d[20] + d[21] + d[22] + d[23] + d[24] + d[25] + d[26];
    *rho = r;
    *ux = (d[1] - d[2] + d[7] - d[8] + d[9] - d[10] - d[11] + d[12] - d[15] + d[16] - d[19] + d[20] - d[21] +
           d[22] + d[23] - d[24] - d[25] + d[26]) / r;
    *uy = (-d[5] + d[6] - d[11] + d[12] + d[13] - d[14] + d[15] - d[16] + d[17] - d[18] + d[19] - d[20] - d[21] +
           d[22] - d[23] + d[24] - d[25] + d[26]) / r;
    *uz = (-d[3] + d[4] - d[7] + d[8] + d[9] - d[10] - d[13] + d[14] + d[17] - d[18] - d[19] + d[20] + d[21] -
           d[22] - d[23] + d[24] - d[25] + d[26]) / r;
}

static void FN(op_moments)(FN(state) * s)
{
    long N = s->N;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
        if (s->flags[i] == 0) continue;
        REAL d[27];
        FN(load)(s, s->df, i, d);
        FN(moments_cell)(d, &s->rho[i], &s->u[3 * i], &s->u[3 * i + 1], &s->u[3 * i + 2]);
    }
}

/* SolverTurbulentLES::runSimulation body (SolverTurbulentLES.h:187-214). */
static void FN(step)(FN(state) * s)
{
    FN(op_omega)(s);
    FN(op_outlet_omega)(s);
    FN(op_collision)(s);
    FN(op_streaming)(s);
    FN(op_bounce_back)(s);
    FN(op_inlet)(s);
    FN(op_outlet)(s);
    FN(op_periodic)(s);
    FN(op_moments)(s);
}

/* InitializationEquilibriumConstVector (:74-126), with per-cell rho/u. */
static void FN(init_eq)(FN(state) * s, const double* rho, const double* ux, const double* uy, const double* uz,
                        double rho0)
{
    long N = s->N;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
        if (s->flags[i] != 0) {
            s->rho[i] = rho ? (REAL)rho[i] : (REAL)rho0;
            s->u[3 * i] = ux ? (REAL)ux[i] : 0;
            s->u[3 * i + 1] = uy ? (REAL)uy[i] : 0;
            s->u[3 * i + 2] = uz ? (REAL)uz[i] : 0;
        } else {
            s->rho[i] = 0.f;
            s->u[3 * i] = s->u[3 * i + 1] = s->u[3 * i + 2] = 0.f;
        }
        s->u_err[3 * i] = s->u_err[3 * i + 1] = s->u_err[3 * i + 2] = 0;
        for (int vel = 0; vel < 27; ++vel) {
            REAL v = 0.f;
            if (s->flags[i] != 0) v = FN(feq)(s->w, vel, s->u[3 * i], s->u[3 * i + 1], s->u[3 * i + 2], s->rho[i]);
            DF(s, vel, i) = v;
            DP(s, vel, i) = v;
        }
    }
}

static void FN(set_pop)(FN(state) * s, const double* f)
{
    long N = s->N;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i)
        for (int q = 0; q < 27; ++q) DF(s, q, i) = s->flags[i] != 0 ? (REAL)f[(long)q * N + i] : 0;
    FN(op_moments)(s);
}

/* ErrorQuadratic::errorEvaluation (ErrorQuadratic.h:33-90), including the
 * y-z component typo of :43-44.  The TNL reduce order is not reproduced
 * (summation here is sequential in double). */
static double FN(error_quadratic)(FN(state) * s)
{
    double e1 = 0.0, e2 = 0.0;
    for (long i = 0; i < s->N; ++i) {
        if (s->flags[i] != 1) continue;
        const REAL* u = &s->u[3 * i];
        const REAL* e = &s->u_err[3 * i];
        REAL add1 = (u[0] - e[0]) * (u[0] - e[0]) + (u[1] - e[1]) * (u[1] - e[1]) + (u[1] - e[2]) * (u[1] - e[2]);
        REAL add2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        e1 += ORC_SQRT(add1);
        e2 += ORC_SQRT(add2);
    }
    for (long i = 0; i < 3 * s->N; ++i) s->u_err[i] = s->u[i];
    return sqrt(e1) / sqrt(e2);
}

#undef DF
#undef DP
