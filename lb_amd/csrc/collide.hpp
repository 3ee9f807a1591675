// collide.hpp -- per-cell collision arithmetic of the collide-stream core.
//
// Header-only, compiled by hipcc for the device and by g++ for a host-side
// unit test that checks it bit-for-bit against the CPU oracle.
//
// Two arithmetic variants of the Geier-2015 cumulant operator:
//  * literal: the operation order of the reference
//    (3D/simulation/src/solvers/collisions/CollisionCumD3Q27Turbulent2015.h:84-485),
//    FP contraction off, so FP64 results match the reference's FP64 fields;
//  * fast: the same operator reduced algebraically.  The reference fixes
//    omega_2..omega_10 = 1 (:240-249); then every post-collision cumulant of
//    order >= 3 is zero and the post-collision central moments of order 4 and
//    6 are products of order-2 ones (Eq. G2015 81-84).  Only raw moments up to
//    order 2 are needed on the way in.
// Populations are passed in reference slot order (d3q27.hpp).
#pragma once

#include <cmath>

#include "../../include/lbm_core.h"
#include "d3q27.hpp"

namespace lbm {

#if defined(__clang__)
#define LBM_FP_STRICT _Pragma("clang fp contract(off)")
#define LBM_FP_FAST _Pragma("clang fp contract(fast)")
#else
#define LBM_FP_STRICT
#define LBM_FP_FAST
#endif

// Lattice constants selected by the compatibility bits.
template <class R>
struct Consts {
    R w[4];     // weight per class: rest, axis, edge, corner
    R c13;      // the 1/3 factor of Eq. G2015 61-73 (reference: 1.f/3.f)
    R c23, c43; // 2/3 and 4/3 of Eq. G2017 43-45 (reference: 2.f/3.f, 4.f/3.f)
    R lam;      // limiter parameter lambda_3..5 of Geier 2017 (reference: 0.01f)
    int lim210;   // the 210/012 limiters read |C_120 +- C_102| (reference quirk)
    int trace17;  // 2017 bulk term subtracts rho -+ 1 instead of k_000 (reference quirk)
    int eq44;     // 2017 Eq. 44 with D_x u + D_y v - 2 D_z w (reference quirk)
    R w4q;        // omega_1^2 coefficient of omega_4's denominator (reference 8, published 9)
};

// Constants for a set of LBM_COMPAT_* bits.  Float literals of the reference
// (D3Q27.h:113-118, CollisionCumD3Q27Turbulent2015.h:280-306,
// CollisionCumD3Q27Turbulent2017.h:244-246, 349-362) keep their float
// rounding under LBM_COMPAT_F32CONST.
template <class R>
inline Consts<R> make_consts(uint32_t compat)
{
    Consts<R> K;
    const bool f32 = (compat & LBM_COMPAT_F32CONST) != 0;
    K.w[0] = f32 ? (R)(8.f / 27.f) : (R)(8.0 / 27.0);
    K.w[1] = f32 ? (R)(2.f / 27.f) : (R)(2.0 / 27.0);
    K.w[2] = f32 ? (R)(1.f / 54.f) : (R)(1.0 / 54.0);
    K.w[3] = f32 ? (R)(1.f / 216.f) : (R)(1.0 / 216.0);
    K.c13 = f32 ? (R)(1.f / 3.f) : (R)(1.0 / 3.0);
    K.c23 = f32 ? (R)(2.f / 3.f) : (R)(2.0 / 3.0);
    K.c43 = f32 ? (R)(4.f / 3.f) : (R)(4.0 / 3.0);
    K.lam = f32 ? (R)0.01f : (R)0.01;
    K.lim210 = (compat & LBM_COMPAT_LIM210) != 0;
    K.trace17 = (compat & LBM_COMPAT_TRACE17) != 0;
    K.eq44 = (compat & LBM_COMPAT_EQ44) != 0;
    K.w4q = (compat & LBM_COMPAT_OMEGA4) ? R(8) : R(9);
    return K;
}

// Variants of the cumulant operator: the Geier-2015 operator, the 2017
// operator of CollisionCumD3Q27Turbulent2017.h and the 2017 branch of
// CollisionCumD3Q27TurbulentCombined.h (they differ only in the bulk-term quirk).
enum { CV_2015 = 0, CV_2017 = 1, CV_2017C = 2 };

template <class R>
LBM_HD R sqrt_r(R x)
{
    return std::sqrt(x);
}

// Reciprocal for the fast arithmetic: the hardware FP32 reciprocal (within
// 1 ulp) on the device, a correctly rounded quotient otherwise.
template <class R>
LBM_HD R frcp(R x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(R) == 4) return __builtin_amdgcn_rcpf(x);
#endif
    return R(1) / x;
}

// Base-10 logarithm for the closure features: the hardware log2 (within
// 1 ulp for normal inputs) times log10(2) on the device in FP32, the library
// function otherwise.
template <class R>
LBM_HD R log10_r(R x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(R) == 4) return __builtin_amdgcn_logf(x) * 0.30102999566398120f;
#endif
    return std::log10(x);
}

// ---------------------------------------------------------------------------
// Equilibrium and moments, literal operation order.
// ---------------------------------------------------------------------------

// f_eq of direction q: w*rho*(1 + 3uc + 4.5uc^2 - 1.5u^2) with
// uc = cx*ux + cy*uy + cz*uz, u2 = ux*ux + uy*uy + uz*uz (OmegaLES.h:47-66).
template <class R>
LBM_HD R feq_q(const Consts<R>& K, int q, R ux, R uy, R uz, R rho)
{
    LBM_FP_STRICT
    R uc = R(CX[q]) * ux + R(CY[q]) * uy + R(CZ[q]) * uz;
    R u2 = ux * ux + uy * uy + uz * uz;
    return K.w[WCLASS[q]] * rho * (R(1) + R(3) * uc + R(4.5) * uc * uc - R(1.5) * u2);
}

// rho = sum f, u = (sum c f)/rho with the term order of
// MomentDensityVelocityN27.h:65-81.
template <class R>
LBM_HD void moments_literal(const R* d, R& rho, R& ux, R& uy, R& uz)
{
    LBM_FP_STRICT
    R r = d[0] + d[1] + d[2] + d[3] + d[4] + d[5] + d[6] + d[7] + d[8] + d[9] + d[10] + d[11] + d[12] + d[13] +
          d[14] + d[15] + d[16] + d[17] + d[18] + d[19] + d[20] + d[21] + d[22] + d[23] + d[24] + d[25] + d[26];
    rho = r;
    ux = (d[1] - d[2] + d[7] - d[8] + d[9] - d[10] - d[11] + d[12] - d[15] + d[16] - d[19] + d[20] - d[21] +
          d[22] + d[23] - d[24] - d[25] + d[26]) / r;
    uy = (-d[5] + d[6] - d[11] + d[12] + d[13] - d[14] + d[15] - d[16] + d[17] - d[18] + d[19] - d[20] - d[21] +
          d[22] - d[23] + d[24] - d[25] + d[26]) / r;
    uz = (-d[3] + d[4] - d[7] + d[8] + d[9] - d[10] - d[13] + d[14] + d[17] - d[18] - d[19] + d[20] + d[21] -
          d[22] - d[23] + d[24] - d[25] + d[26]) / r;
}

// BGK collision f - (f - f_eq)*omega (CollisionSRTTurbulent.h:56-75).
template <class R>
LBM_HD void srt_literal(const Consts<R>& K, R* f, R rho, R ux, R uy, R uz, R omega)
{
    LBM_FP_STRICT
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = f[q] - (f[q] - feq_q(K, q, ux, uy, uz, rho)) * omega;
}

// Smagorinsky rate from Pi = sum c c (f - f_eq), literal loop (OmegaLES.h:84-129).
template <class R>
LBM_HD R omega_les_literal(const Consts<R>& K, const R* f, R rho, R ux, R uy, R uz, R tau, R cles)
{
    LBM_FP_STRICT
    R p[9] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const R neq = f[q] - feq_q(K, q, ux, uy, uz, rho);
        const int c[3] = {CX[q], CY[q], CZ[q]};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) p[3 * a + b] += R(c[a] * c[b]) * neq;
    }
    R pi = p[0] * p[0];
#pragma unroll
    for (int k = 1; k < 9; ++k) pi += p[k] * p[k];
    const R rho_inv = R(1) / rho;
    const R cterm = R(18) * cles * rho_inv;
    pi = sqrt_r(pi);
    const R tau_les = tau * R(0.5) + R(0.5) * sqrt_r(tau * tau + cterm * pi);
    return R(1) / tau_les;
}

// ---------------------------------------------------------------------------
// Literal cumulant operator.
// ---------------------------------------------------------------------------

// One line of the forward central-moment transform, in place:
// (f_m, f_z, f_p) -> (k_0, k_1, k_2)   (Eq. G2015 43-45).
template <class R>
LBM_HD void fwd_line(R& m, R& z, R& p, R u)
{
    LBM_FP_STRICT
    const R s = p + m;
    const R d = p - m;
    const R k0 = s + z;
    const R k1 = d - u * k0;
    const R k2 = s - R(2) * u * d + u * u * k0;
    m = k0;
    z = k1;
    p = k2;
}

// One line of the backward transform, in place:
// (k_0, k_1, k_2) -> (f_m, f_z, f_p)   (Eq. G2015 88-96).
template <class R>
LBM_HD void inv_line(R& v0, R& v1, R& v2, R u)
{
    LBM_FP_STRICT
    const R a0 = v0, a1 = v1, a2 = v2;
    const R fz = a0 * (R(1) - u * u) - R(2) * u * a1 - a2;
    const R fm = (a0 * (u * u - u) + a1 * (R(2) * u - R(1)) + a2) * R(0.5);
    const R fp = (a0 * (u * u + u) + a1 * (R(2) * u + R(1)) + a2) * R(0.5);
    v0 = fm;
    v1 = fz;
    v2 = fp;
}

// v[9*a + 3*b + c]: a, b, c index x, y, z (velocity class or moment order).
template <class R>
LBM_HD void fwd_all(R* v, R ux, R uy, R uz)
{
#pragma unroll
    for (int ab = 0; ab < 9; ++ab) fwd_line(v[3 * ab], v[3 * ab + 1], v[3 * ab + 2], uz);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) fwd_line(v[9 * a + c], v[9 * a + 3 + c], v[9 * a + 6 + c], uy);
#pragma unroll
    for (int bc = 0; bc < 9; ++bc) fwd_line(v[bc], v[9 + bc], v[18 + bc], ux);
}

template <class R>
LBM_HD void inv_all(R* v, R ux, R uy, R uz)
{
#pragma unroll
    for (int bc = 0; bc < 9; ++bc) inv_line(v[bc], v[9 + bc], v[18 + bc], ux);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) inv_line(v[9 * a + c], v[9 * a + 3 + c], v[9 * a + 6 + c], uy);
#pragma unroll
    for (int ab = 0; ab < 9; ++ab) inv_line(v[3 * ab], v[3 * ab + 1], v[3 * ab + 2], uz);
}

// b*c + 2*d*e : the order-4 correction of Eq. G2015 51-52 / 81-82.
template <class R>
LBM_HD R pair4(R b, R c, R d, R e)
{
    LBM_FP_STRICT
    return b * c + R(2) * d * e;
}

// a*b + c*d + 4*e*f + 2*(g*h + i*j) : the order-5 correction of Eq. 53 / 83.
template <class R>
LBM_HD R pair5(R a, R b, R c, R d, R e, R f, R g, R h, R i, R j)
{
    LBM_FP_STRICT
    return a * b + c * d + R(4) * e * f + R(2) * (g * h + i * j);
}

// Order-6 corrections of Eq. G2015 54 / 84, split as A/rho and B/rho/rho.
template <class R>
LBM_HD R six_a(const R* k)
{
    LBM_FP_STRICT
#define M_(i, j, l) k[9 * (i) + 3 * (j) + (l)]
    return R(4) * M_(1, 1, 1) * M_(1, 1, 1) + M_(2, 0, 0) * M_(0, 2, 2) + M_(0, 2, 0) * M_(2, 0, 2) +
           M_(0, 0, 2) * M_(2, 2, 0) +
           R(4) * (M_(0, 1, 1) * M_(2, 1, 1) + M_(1, 0, 1) * M_(1, 2, 1) + M_(1, 1, 0) * M_(1, 1, 2)) +
           R(2) * (M_(1, 2, 0) * M_(1, 0, 2) + M_(2, 1, 0) * M_(0, 1, 2) + M_(2, 0, 1) * M_(0, 2, 1));
}

template <class R>
LBM_HD R six_b(const R* k)
{
    LBM_FP_STRICT
    return R(16) * M_(1, 1, 0) * M_(1, 0, 1) * M_(0, 1, 1) +
           R(4) * (M_(1, 0, 1) * M_(1, 0, 1) * M_(0, 2, 0) + M_(0, 1, 1) * M_(0, 1, 1) * M_(2, 0, 0) +
                   M_(1, 1, 0) * M_(1, 1, 0) * M_(0, 0, 2)) +
           R(2) * M_(2, 0, 0) * M_(0, 2, 0) * M_(0, 0, 2);
}

// The four order-4/5 correction sets shared by the forward (cumulant) and
// backward (central-moment) directions.  sign = -1 forward, +1 backward.
template <class R>
LBM_HD void corr45(const R* k, R rho, R* out /* 211 121 112 220 022 202 122 212 221 */)
{
    LBM_FP_STRICT
    out[0] = pair4(M_(2, 0, 0), M_(0, 1, 1), M_(1, 0, 1), M_(1, 1, 0)) / rho;
    out[1] = pair4(M_(0, 2, 0), M_(1, 0, 1), M_(1, 1, 0), M_(0, 1, 1)) / rho;
    out[2] = pair4(M_(0, 0, 2), M_(1, 1, 0), M_(0, 1, 1), M_(1, 0, 1)) / rho;
    out[3] = pair4(M_(0, 2, 0), M_(2, 0, 0), M_(1, 1, 0), M_(1, 1, 0)) / rho;
    out[4] = pair4(M_(0, 0, 2), M_(0, 2, 0), M_(0, 1, 1), M_(0, 1, 1)) / rho;
    out[5] = pair4(M_(2, 0, 0), M_(0, 0, 2), M_(1, 0, 1), M_(1, 0, 1)) / rho;
    out[6] = pair5(M_(0, 2, 0), M_(1, 0, 2), M_(0, 0, 2), M_(1, 2, 0), M_(0, 1, 1), M_(1, 1, 1), M_(1, 1, 0),
                   M_(0, 1, 2), M_(1, 0, 1), M_(0, 2, 1)) / rho;
    out[7] = pair5(M_(0, 0, 2), M_(2, 1, 0), M_(2, 0, 0), M_(0, 1, 2), M_(1, 0, 1), M_(1, 1, 1), M_(0, 1, 1),
                   M_(2, 0, 1), M_(1, 1, 0), M_(1, 0, 2)) / rho;
    out[8] = pair5(M_(2, 0, 0), M_(0, 2, 1), M_(0, 2, 0), M_(2, 0, 1), M_(1, 1, 0), M_(1, 1, 1), M_(1, 0, 1),
                   M_(1, 2, 0), M_(0, 1, 1), M_(2, 1, 0)) / rho;
}

// Cumulant collision in the reference's operation order: Geier 2015
// (CollisionCumD3Q27Turbulent2015.h:84-485) or Geier 2017
// (CollisionCumD3Q27Turbulent2017.h:84-539; Combined.h:84-542 for CV_2017C).
// f: in/out populations; rho, u: the macroscopic fields the reference passes
// in (moments of the previous step); omg: omega_1.
template <int V, class R>
LBM_HD void cum_literal(const Consts<R>& K, R* f, R rho, R ux, R uy, R uz, R omg)
{
    LBM_FP_STRICT
    R k[27];
#pragma unroll
    for (int q = 0; q < Q; ++q) k[CELL3(q)] = f[q];
    fwd_all(k, ux, uy, uz);

    // cumulants, Eq. G2015 47-54
    R c45[9];
    corr45(k, rho, c45);
    const R k000 = M_(0, 0, 0);
    const R C200 = M_(2, 0, 0), C020 = M_(0, 2, 0), C002 = M_(0, 0, 2);
    const R C110 = M_(1, 1, 0), C101 = M_(1, 0, 1), C011 = M_(0, 1, 1);
    const R C120 = M_(1, 2, 0), C012 = M_(0, 1, 2), C201 = M_(2, 0, 1);
    const R C210 = M_(2, 1, 0), C021 = M_(0, 2, 1), C102 = M_(1, 0, 2), C111 = M_(1, 1, 1);
    const R C211 = M_(2, 1, 1) - c45[0], C121 = M_(1, 2, 1) - c45[1], C112 = M_(1, 1, 2) - c45[2];
    const R C220 = M_(2, 2, 0) - c45[3], C022 = M_(0, 2, 2) - c45[4], C202 = M_(2, 0, 2) - c45[5];
    const R C122 = M_(1, 2, 2) - c45[6], C212 = M_(2, 1, 2) - c45[7], C221 = M_(2, 2, 1) - c45[8];
    const R C222 = M_(2, 2, 2) - six_a(k) / rho + six_b(k) / rho / rho;

    R s[27];
#define S_(i, j, l) s[9 * (i) + 3 * (j) + (l)]
    R cs220, cs202, cs022, cs211, cs121, cs112, cs221, cs212, cs122, cs222;
    S_(0, 0, 0) = k000;
    S_(1, 0, 0) = -M_(1, 0, 0);
    S_(0, 1, 0) = -M_(0, 1, 0);
    S_(0, 0, 1) = -M_(0, 0, 1);
    if constexpr (V == CV_2015) {
        // relaxation (omega_2..omega_10 = 1 as configured by the reference)
        const R w1 = omg, w2 = R(1), w3 = R(1), w4 = R(1), w5 = R(1), w6 = R(1), w7 = R(1), w8 = R(1), w9 = R(1),
                w10 = R(1);
        const R Dxu =
            -w1 * R(0.5) / rho * (R(2) * C200 - C020 - C002) - w2 * R(0.5) / rho * (C200 + C020 + C002 - k000);
        const R Dyv = Dxu + R(3) * w1 * R(0.5) / rho * (C200 - C020);
        const R Dzw = Dxu + R(3) * w1 * R(0.5) / rho * (C200 - C002);
        const R e61 =
            (R(1) - w1) * (C200 - C020) - R(3) * rho * (R(1) - w1 * R(0.5)) * (ux * ux * Dxu - uy * uy * Dyv);
        const R e64 =
            (R(1) - w1) * (C200 - C002) - R(3) * rho * (R(1) - w1 * R(0.5)) * (ux * ux * Dxu - uz * uz * Dzw);
        const R e65 = k000 * w2 + (R(1) - w2) * (C200 + C020 + C002) -
                      R(3) * rho * (R(1) - w2 / R(2)) * (ux * ux * Dxu + uy * uy * Dyv + uz * uz * Dzw);
        const R e71 = (R(1) - w6) * (C220 - R(2) * C202 + C022);
        const R e72 = (R(1) - w6) * (C220 + C202 - R(2) * C022);
        const R e73 = (R(1) - w7) * (C220 + C202 + C022);
        S_(1, 1, 0) = (R(1) - w1) * C110;
        S_(1, 0, 1) = (R(1) - w1) * C101;
        S_(0, 1, 1) = (R(1) - w1) * C011;
        S_(2, 0, 0) = K.c13 * (e61 + e64 + e65);
        S_(0, 2, 0) = K.c13 * (e64 - R(2) * e61 + e65);
        S_(0, 0, 2) = K.c13 * (e61 - R(2) * e64 + e65);
        S_(1, 2, 0) = (-C102 - C120) * w3 * R(0.5) + (C102 - C120) * w4 * R(0.5) + C120;
        S_(1, 0, 2) = (-C102 - C120) * w3 * R(0.5) + (-C102 + C120) * w4 * R(0.5) + C102;
        S_(2, 1, 0) = (-C012 - C210) * w3 * R(0.5) + (C012 - C210) * w4 * R(0.5) + C210;
        S_(0, 1, 2) = (-C012 - C210) * w3 * R(0.5) + (-C012 + C210) * w4 * R(0.5) + C012;
        S_(0, 2, 1) = (-C021 - C201) * w3 * R(0.5) + (-C021 + C201) * w4 * R(0.5) + C021;
        S_(2, 0, 1) = (-C021 - C201) * w3 * R(0.5) + (C021 - C201) * w4 * R(0.5) + C201;
        S_(1, 1, 1) = (R(1) - w5) * C111;
        cs220 = K.c13 * (e71 + e72 + e73);
        cs202 = K.c13 * (-e71 + e73);
        cs022 = K.c13 * (-e72 + e73);
        cs211 = (R(1) - w8) * C211, cs121 = (R(1) - w8) * C121, cs112 = (R(1) - w8) * C112;
        cs221 = (R(1) - w9) * C221, cs212 = (R(1) - w9) * C212, cs122 = (R(1) - w9) * C122;
        cs222 = (R(1) - w10) * C222;
    } else {
        // relaxation rates, Geier 2017 (Turbulent2017.h:240-291): omega_2 = 1,
        // omega_3..5 from omega_1 with magnitude limiters, omega_6..10 = 1
        const R w1 = omg, w2 = R(1);
        const R lam = K.lam;
        const R w3 = R(8) * (w1 - R(2)) * (w2 * (R(3) * w1 - R(1)) - R(5) * w1) /
                     (R(8) * (R(5) - R(2) * w1) * w1 + w2 * (R(8) + w1 * (R(9) * w1 - R(26))));
        const R a210p = K.lim210 ? C120 + C102 : C210 + C012;
        const R a210m = K.lim210 ? C120 - C102 : C210 - C012;
        const R w120p102 = w3 + (R(1) - w3) * std::fabs(C120 + C102) / (rho * lam + std::fabs(C120 + C102));
        const R w210p012 = w3 + (R(1) - w3) * std::fabs(a210p) / (rho * lam + std::fabs(a210p));
        const R w201p021 = w3 + (R(1) - w3) * std::fabs(C201 + C021) / (rho * lam + std::fabs(C201 + C021));
        const R w4 = R(8) * (w1 - R(2)) * (w1 + w2 * (R(3) * w1 - R(7))) /
                     (w2 * (R(56) - R(42) * w1 + K.w4q * w1 * w1) - R(8) * w1);
        const R w120m102 = w4 + (R(1) - w4) * std::fabs(C120 - C102) / (rho * lam + std::fabs(C120 - C102));
        const R w210m012 = w4 + (R(1) - w4) * std::fabs(a210m) / (rho * lam + std::fabs(a210m));
        const R w201m021 = w4 + (R(1) - w4) * std::fabs(C201 - C021) / (rho * lam + std::fabs(C201 - C021));
        const R w5 = R(24) * (w1 - R(2)) *
                     (R(4) * w1 * w1 + w1 * w2 * (R(18) - R(13) * w1) + w2 * w2 * (R(2) + w1 * (R(6) * w1 - R(11)))) /
                     (R(16) * w1 * w1 * (w1 - R(6)) - R(2) * w1 * w2 * (R(216) + R(5) * w1 * (R(9) * w1 - R(46))) +
                      w2 * w2 * (w1 * (R(3) * w1 - R(10)) * (R(15) * w1 - R(28)) - R(48)));
        const R w111 = w5 + (R(1) - w5) * std::fabs(C111) / (rho * lam + std::fabs(C111));
        const R w6 = R(1), w7 = R(1), w8 = R(1), w9 = R(1), w10 = R(1);
        const R A = (R(4) * w1 * w1 + R(2) * w1 * w2 * (w1 - R(6)) + w2 * w2 * (w1 * (R(10) - R(3) * w1) - R(4))) /
                    (w1 - w2) / (w2 * (R(2) + R(3) * w1) - R(8) * w1);
        const R B = (R(4) * w1 * w2 * (R(9) * w1 - R(16)) - R(4) * w1 * w1 -
                     R(2) * w2 * w2 * (R(2) + R(9) * w1 * (w1 - R(2)))) /
                    R(3) / (w1 - w2) / (w2 * (R(2) + R(3) * w1) - R(8) * w1);
        // bulk term: k_000 (corrected) or the reference's rho - 1 / rho + 1
        R tr = k000;
        if (K.trace17) tr = V == CV_2017 ? R(-1) + rho : -(R(-1) - rho);
        const R Dxu = -w1 * R(0.5) / rho * (R(2) * C200 - C020 - C002) - w2 * R(0.5) / rho * (C200 + C020 + C002 - tr);
        const R Dyv = Dxu + R(3) * w1 * R(0.5) / rho * (C200 - C020);
        const R Dzw = Dxu + R(3) * w1 * R(0.5) / rho * (C200 - C002);
        const R DxvDyu = -R(3) * w1 / rho * C110;
        const R DxwDzu = -R(3) * w1 / rho * C101;
        const R DywDzv = -R(3) * w1 / rho * C011;
        const R e61 =
            (R(1) - w1) * (C200 - C020) - R(3) * rho * (R(1) - w1 * R(0.5)) * (ux * ux * Dxu - uy * uy * Dyv);
        const R e64 =
            (R(1) - w1) * (C200 - C002) - R(3) * rho * (R(1) - w1 * R(0.5)) * (ux * ux * Dxu - uz * uz * Dzw);
        const R e65 = k000 * w2 + (R(1) - w2) * (C200 + C020 + C002) -
                      R(3) * rho * (R(1) - w2 / R(2)) * (ux * ux * Dxu + uy * uy * Dyv + uz * uz * Dzw);
        S_(1, 1, 0) = (R(1) - w1) * C110;
        S_(1, 0, 1) = (R(1) - w1) * C101;
        S_(0, 1, 1) = (R(1) - w1) * C011;
        S_(2, 0, 0) = K.c13 * (e61 + e64 + e65);
        S_(0, 2, 0) = K.c13 * (e64 - R(2) * e61 + e65);
        S_(0, 0, 2) = K.c13 * (e61 - R(2) * e64 + e65);
        // third order, Eq. G2017 117-122 and 42 (:331-346)
        const R e117 = (R(1) - w120p102) * (C120 + C102);
        const R e118 = (R(1) - w210p012) * (C210 + C012);
        const R e119 = (R(1) - w201p021) * (C201 + C021);
        const R e120 = (R(1) - w120m102) * (C120 - C102);
        const R e121 = (R(1) - w210m012) * (C210 - C012);
        const R e122 = (R(1) - w201m021) * (C201 - C021);
        S_(1, 2, 0) = R(0.5) * (e120 + e117);
        S_(1, 0, 2) = R(0.5) * (-e120 + e117);
        S_(2, 1, 0) = R(0.5) * (e121 + e118);
        S_(0, 1, 2) = R(0.5) * (-e121 + e118);
        S_(0, 2, 1) = R(0.5) * (-e122 + e119);
        S_(2, 0, 1) = R(0.5) * (e122 + e119);
        S_(1, 1, 1) = (R(1) - w111) * C111;
        // fourth order, Eq. G2017 43-48 (:349-365)
        const R e43 = K.c23 * (R(1) / w1 - R(0.5)) * w6 * A * rho * (Dxu - R(2) * Dyv + Dzw) +
                      (R(1) - w6) * (C220 - R(2) * C202 + C022);
        const R e44 = K.c23 * (R(1) / w1 - R(0.5)) * w6 * A * rho *
                          (K.eq44 ? (Dxu + Dyv - R(2) * Dzw) : (Dyv + Dzw - R(2) * Dxu)) +
                      (R(1) - w6) * (C220 + C202 - R(2) * C022);
        const R e45 = -K.c43 * (R(1) / w1 - R(0.5)) * w7 * A * rho * (Dxu + Dyv + Dzw) +
                      (R(1) - w7) * (C220 + C202 + C022);
        cs220 = K.c13 * (e43 + e44 + e45);
        cs202 = K.c13 * (-e43 + e45);
        cs022 = K.c13 * (-e44 + e45);
        cs211 = -K.c13 * (R(1) / w1 - R(0.5)) * w8 * B * rho * DywDzv + (R(1) - w8) * C211;
        cs121 = -K.c13 * (R(1) / w1 - R(0.5)) * w8 * B * rho * DxwDzu + (R(1) - w8) * C121;
        cs112 = -K.c13 * (R(1) / w1 - R(0.5)) * w8 * B * rho * DxvDyu + (R(1) - w8) * C112;
        cs221 = (R(1) - w9) * C221, cs212 = (R(1) - w9) * C212, cs122 = (R(1) - w9) * C122;
        cs222 = (R(1) - w10) * C222;
    }

    // back to central moments, Eq. G2015 81-84
    R b45[9];
    corr45(s, rho, b45);
    S_(2, 1, 1) = cs211 + b45[0];
    S_(1, 2, 1) = cs121 + b45[1];
    S_(1, 1, 2) = cs112 + b45[2];
    S_(2, 2, 0) = cs220 + b45[3];
    S_(0, 2, 2) = cs022 + b45[4];
    S_(2, 0, 2) = cs202 + b45[5];
    // the order-5 corrections read the order-4 values just set
    corr45(s, rho, b45);
    S_(1, 2, 2) = cs122 + b45[6];
    S_(2, 1, 2) = cs212 + b45[7];
    S_(2, 2, 1) = cs221 + b45[8];
    S_(2, 2, 2) = cs222 + six_a(s) / rho - six_b(s) / rho / rho;
#undef S_

    inv_all(s, ux, uy, uz);
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = s[CELL3(q)];
}
#undef M_

// ---------------------------------------------------------------------------
// Fast variant (FP32 / FP16 / BF16 paths).
// ---------------------------------------------------------------------------

// Raw moments up to order 2, summed along z, y, x lines.
template <class R>
struct Raw2 {
    R m0, jx, jy, jz, mxx, myy, mzz, mxy, mxz, myz;
};

template <class R>
LBM_HD Raw2<R> raw_moments(const R* f)
{
    // z-lines: sum, first moment, second moment for each (cx, cy)
    R s0[9], s1[9], s2[9];
#pragma unroll
    for (int ab = 0; ab < 9; ++ab) {
        R m = R(0), z = R(0), p = R(0);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if (3 * (CX[q] + 1) + (CY[q] + 1) != ab) continue;
            if (CZ[q] < 0) m = f[q];
            else if (CZ[q] == 0) z = f[q];
            else p = f[q];
        }
        s2[ab] = p + m;
        s1[ab] = p - m;
        s0[ab] = s2[ab] + z;
    }
    // y-lines for each cx: (order_y, order_z) in {(0,0),(1,0),(0,1),(2,0),(1,1),(0,2)}
    R t00[3], t10[3], t01[3], t20[3], t11[3], t02[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int m = 3 * a, z = 3 * a + 1, p = 3 * a + 2;
        t20[a] = s0[p] + s0[m];
        t10[a] = s0[p] - s0[m];
        t00[a] = t20[a] + s0[z];
        t11[a] = s1[p] - s1[m];
        t01[a] = s1[p] + s1[m] + s1[z];
        t02[a] = s2[p] + s2[m] + s2[z];
    }
    Raw2<R> r;
    r.mxx = t00[2] + t00[0];
    r.jx = t00[2] - t00[0];
    r.m0 = r.mxx + t00[1];
    r.mxy = t10[2] - t10[0];
    r.mxz = t01[2] - t01[0];
    r.jy = t10[2] + t10[0] + t10[1];
    r.jz = t01[2] + t01[0] + t01[1];
    r.myy = t20[2] + t20[0] + t20[1];
    r.myz = t11[2] + t11[0] + t11[1];
    r.mzz = t02[2] + t02[0] + t02[1];
    return r;
}

// Central moments of order 1 and 2 about (ux, uy, uz).
template <class R>
struct Cen2 {
    R k100, k010, k001, kxx, kyy, kzz, kxy, kxz, kyz;
};

template <class R>
LBM_HD Cen2<R> central2(const Raw2<R>& r, R ux, R uy, R uz)
{
    LBM_FP_STRICT  // same bits in every kernel that forms the Smagorinsky rate
    Cen2<R> c;
    c.k100 = r.jx - ux * r.m0;
    c.k010 = r.jy - uy * r.m0;
    c.k001 = r.jz - uz * r.m0;
    c.kxx = r.mxx - R(2) * ux * r.jx + ux * ux * r.m0;
    c.kyy = r.myy - R(2) * uy * r.jy + uy * uy * r.m0;
    c.kzz = r.mzz - R(2) * uz * r.jz + uz * uz * r.m0;
    c.kxy = r.mxy - ux * r.jy - uy * r.jx + ux * uy * r.m0;
    c.kxz = r.mxz - ux * r.jz - uz * r.jx + ux * uz * r.m0;
    c.kyz = r.myz - uy * r.jz - uz * r.jy + uy * uz * r.m0;
    return c;
}

// Smagorinsky rate from the non-equilibrium stress.  For D3Q27 the
// second-order equilibrium moments are rho*(u u + I/3), so
// Pi_neq = central second moments - rho/3 I (the reference builds the same
// tensor with a 27-term loop, OmegaLES.h:84-101).
template <class R>
LBM_HD R omega_les_fast(const Cen2<R>& c, R rho, R tau, R cles)
{
    // contraction off: the learned closure's fallback evaluates this same
    // formula in another kernel and must give the same bits
    LBM_FP_STRICT
    const R third = rho * R(1.0 / 3.0);
    const R pxx = c.kxx - third, pyy = c.kyy - third, pzz = c.kzz - third;
    const R n2 = pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (c.kxy * c.kxy + c.kxz * c.kxz + c.kyz * c.kyz);
    const R tau_les = R(0.5) * tau + R(0.5) * sqrt_r(tau * tau + R(18) * cles * sqrt_r(n2) / rho);
    return R(1) / tau_les;
}

// Inverse line transform with order-1 term zero: (a0, 0, a2).
template <class R>
LBM_HD void inv_even(R a0, R a2, R uu, R uum, R uup, R& fm, R& fz, R& fp)
{
    fz = a0 * (R(1) - uu) - a2;
    fm = R(0.5) * (a0 * uum + a2);
    fp = R(0.5) * (a0 * uup + a2);
}

template <class R>
LBM_HD void inv_gen(R a0, R a1, R a2, R u, R uu, R uum, R uup, R& fm, R& fz, R& fp)
{
    fz = a0 * (R(1) - uu) - R(2) * u * a1 - a2;
    fm = R(0.5) * (a0 * uum + a1 * (R(2) * u - R(1)) + a2);
    fp = R(0.5) * (a0 * uup + a1 * (R(2) * u + R(1)) + a2);
}

// Reduced cumulant collision.  m0 is the zeroth moment of f, rho/u the
// fields the step uses (equal to the moments of f except on the first step
// after an equilibrium initialisation).  Writes post-collision populations.
template <class R>
LBM_HD void cum2015_fast(const Consts<R>& K, const Cen2<R>& c, R m0, R rho, R ux, R uy, R uz, R omg, R* f)
{
    const R ir = R(1) / rho;
    const R C200 = c.kxx, C020 = c.kyy, C002 = c.kzz;
    const R hw = R(0.5) * omg * ir;
    const R Dxu = -hw * (R(2) * C200 - C020 - C002) - R(0.5) * ir * (C200 + C020 + C002 - m0);
    const R Dyv = Dxu + R(3) * hw * (C200 - C020);
    const R Dzw = Dxu + R(3) * hw * (C200 - C002);
    const R om1 = R(1) - omg;
    const R A = R(3) * rho * (R(1) - R(0.5) * omg);
    const R xx = ux * ux * Dxu, yy = uy * uy * Dyv, zz = uz * uz * Dzw;
    const R e61 = om1 * (C200 - C020) - A * (xx - yy);
    const R e64 = om1 * (C200 - C002) - A * (xx - zz);
    const R e65 = m0 - R(1.5) * rho * (xx + yy + zz);
    const R s200 = K.c13 * (e61 + e64 + e65);
    const R s020 = K.c13 * (e64 - R(2) * e61 + e65);
    const R s002 = K.c13 * (e61 - R(2) * e64 + e65);
    const R s110 = om1 * c.kxy, s101 = om1 * c.kxz, s011 = om1 * c.kyz;
    // order-4 and order-6 central moments of the post-collision state
    const R s211 = (s200 * s011 + R(2) * s101 * s110) * ir;
    const R s121 = (s020 * s101 + R(2) * s110 * s011) * ir;
    const R s112 = (s002 * s110 + R(2) * s011 * s101) * ir;
    const R s220 = (s020 * s200 + R(2) * s110 * s110) * ir;
    const R s022 = (s002 * s020 + R(2) * s011 * s011) * ir;
    const R s202 = (s200 * s002 + R(2) * s101 * s101) * ir;
    const R s222 = (s200 * s022 + s020 * s202 + s002 * s220 + R(4) * (s011 * s211 + s101 * s121 + s110 * s112)) * ir -
                   (R(16) * s110 * s101 * s011 +
                    R(4) * (s101 * s101 * s020 + s011 * s011 * s200 + s110 * s110 * s002) +
                    R(2) * s200 * s020 * s002) * ir * ir;

    // x pass: (order_x 0,1,2) -> (m, z, p) for each (order_y, order_z)
    const R uxx = ux * ux, uxm = uxx - ux, uxp = uxx + ux;
    R X[27];  // X[9*a + 3*oy + oz], a = x class
#define XS(a, oy, oz) X[9 * (a) + 3 * (oy) + (oz)]
    inv_gen(m0, -c.k100, s200, ux, uxx, uxm, uxp, XS(0, 0, 0), XS(1, 0, 0), XS(2, 0, 0));
    inv_gen(-c.k001, s101, R(0), ux, uxx, uxm, uxp, XS(0, 0, 1), XS(1, 0, 1), XS(2, 0, 1));
    inv_even(s002, s202, uxx, uxm, uxp, XS(0, 0, 2), XS(1, 0, 2), XS(2, 0, 2));
    inv_gen(-c.k010, s110, R(0), ux, uxx, uxm, uxp, XS(0, 1, 0), XS(1, 1, 0), XS(2, 1, 0));
    inv_even(s011, s211, uxx, uxm, uxp, XS(0, 1, 1), XS(1, 1, 1), XS(2, 1, 1));
    {
        const R a1 = s112;  // (0, s112, 0)
        XS(1, 1, 2) = -R(2) * ux * a1;
        XS(0, 1, 2) = a1 * (ux - R(0.5));
        XS(2, 1, 2) = a1 * (ux + R(0.5));
    }
    inv_even(s020, s220, uxx, uxm, uxp, XS(0, 2, 0), XS(1, 2, 0), XS(2, 2, 0));
    {
        const R a1 = s121;  // (0, s121, 0)
        XS(1, 2, 1) = -R(2) * ux * a1;
        XS(0, 2, 1) = a1 * (ux - R(0.5));
        XS(2, 2, 1) = a1 * (ux + R(0.5));
    }
    inv_even(s022, s222, uxx, uxm, uxp, XS(0, 2, 2), XS(1, 2, 2), XS(2, 2, 2));

    // y pass
    const R uyy = uy * uy, uym = uyy - uy, uyp = uyy + uy;
    R Y[27];  // Y[9*a + 3*b + oz]
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int oz = 0; oz < 3; ++oz)
            inv_gen(XS(a, 0, oz), XS(a, 1, oz), XS(a, 2, oz), uy, uyy, uym, uyp, Y[9 * a + oz], Y[9 * a + 3 + oz],
                    Y[9 * a + 6 + oz]);
#undef XS
    // z pass
    const R uzz = uz * uz, uzm = uzz - uz, uzp = uzz + uz;
    R F[27];
#pragma unroll
    for (int ab = 0; ab < 9; ++ab)
        inv_gen(Y[3 * ab], Y[3 * ab + 1], Y[3 * ab + 2], uz, uzz, uzm, uzp, F[3 * ab], F[3 * ab + 1], F[3 * ab + 2]);
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = F[CELL3(q)];
}

// Reduced Geier-2017 collision (fast variant).  omega_6..omega_10 = 1 as in
// the reference (Turbulent2017.h:279-283), so the post-collision cumulants of
// order 5 and 6 vanish and those of order 4 are the A/B gradient terms alone;
// order 3 relaxes with the limited omega_3..5 and so needs the pre-collision
// third-order central moments, which come from the forward transform of f
// (the compiler keeps only the passes they need).  Orders 0..2 come from c,
// the same central moments the Smagorinsky rate was formed from.
template <int V, class R>
LBM_HD void cum2017_fast(const Consts<R>& K, const R* fin, const Cen2<R>& c, R m0, R rho, R ux, R uy, R uz, R omg,
                         R* f)
{
    R k[27];
#pragma unroll
    for (int q = 0; q < Q; ++q) k[CELL3(q)] = fin[q];
    fwd_all(k, ux, uy, uz);
#define M_(i, j, l) k[9 * (i) + 3 * (j) + (l)]
    const R C120 = M_(1, 2, 0), C102 = M_(1, 0, 2), C210 = M_(2, 1, 0), C012 = M_(0, 1, 2);
    const R C201 = M_(2, 0, 1), C021 = M_(0, 2, 1), C111 = M_(1, 1, 1);
#undef M_
    const R ir = R(1) / rho;
    const R C200 = c.kxx, C020 = c.kyy, C002 = c.kzz;

    // rates and limiters (Turbulent2017.h:240-291) with omega_2 = 1 substituted
    const R w1 = omg;
    const R w3 = -R(8) * (w1 - R(2)) * (R(2) * w1 + R(1)) * frcp(R(8) + w1 * (R(14) - R(7) * w1));
    const R w4 = R(8) * (w1 - R(2)) * (R(4) * w1 - R(7)) * frcp(R(56) - R(50) * w1 + K.w4q * w1 * w1);
    const R w5 = R(24) * (w1 - R(2)) * (R(2) + w1 * (R(7) - R(3) * w1)) *
                 frcp(R(16) * w1 * w1 * (w1 - R(6)) - R(2) * w1 * (R(216) + R(5) * w1 * (R(9) * w1 - R(46))) +
                      (w1 * (R(3) * w1 - R(10)) * (R(15) * w1 - R(28)) - R(48)));
    const R iden = frcp((w1 - R(1)) * (R(2) - R(5) * w1));
    const R A = (R(4) * w1 * w1 + R(2) * w1 * (w1 - R(6)) + w1 * (R(10) - R(3) * w1) - R(4)) * iden;
    const R B = (R(4) * w1 * (R(9) * w1 - R(16)) - R(4) * w1 * w1 - R(2) * (R(2) + R(9) * w1 * (w1 - R(2)))) *
                (iden * R(1.0 / 3.0));
    const R rl = rho * K.lam;
    auto lim = [rl](R w, R a) {
        const R m = std::fabs(a);
        return (R(1) - w) * rl * frcp(rl + m);   // 1 - (w + (1 - w) |a| / (rl + |a|))
    };
    const R p120 = C120 + C102, m120 = C120 - C102;
    const R p210 = C210 + C012, m210 = C210 - C012;
    const R p201 = C201 + C021, m201 = C201 - C021;
    const R e117 = lim(w3, p120) * p120;
    const R e118 = lim(w3, K.lim210 ? p120 : p210) * p210;
    const R e119 = lim(w3, p201) * p201;
    const R e120 = lim(w4, m120) * m120;
    const R e121 = lim(w4, K.lim210 ? m120 : m210) * m210;
    const R e122 = lim(w4, m201) * m201;
    const R s120 = R(0.5) * (e120 + e117), s102 = R(0.5) * (e117 - e120);
    const R s210 = R(0.5) * (e121 + e118), s012 = R(0.5) * (e118 - e121);
    const R s201 = R(0.5) * (e122 + e119), s021 = R(0.5) * (e119 - e122);
    const R s111 = lim(w5, C111) * C111;

    // second order as in the 2015 operator, with the 2017 bulk term
    R tr = m0;
    if (K.trace17) tr = V == CV_2017 ? rho - R(1) : rho + R(1);
    const R hw = R(0.5) * omg * ir;
    const R Dxu = -hw * (R(2) * C200 - C020 - C002) - R(0.5) * ir * (C200 + C020 + C002 - tr);
    const R Dyv = Dxu + R(3) * hw * (C200 - C020);
    const R Dzw = Dxu + R(3) * hw * (C200 - C002);
    const R om1 = R(1) - omg;
    const R A3 = R(3) * rho * (R(1) - R(0.5) * omg);
    const R xx = ux * ux * Dxu, yy = uy * uy * Dyv, zz = uz * uz * Dzw;
    const R e61 = om1 * (C200 - C020) - A3 * (xx - yy);
    const R e64 = om1 * (C200 - C002) - A3 * (xx - zz);
    const R e65 = m0 - R(1.5) * rho * (xx + yy + zz);
    const R s200 = K.c13 * (e61 + e64 + e65);
    const R s020 = K.c13 * (e64 - R(2) * e61 + e65);
    const R s002 = K.c13 * (e61 - R(2) * e64 + e65);
    const R s110 = om1 * c.kxy, s101 = om1 * c.kxz, s011 = om1 * c.kyz;

    // fourth-order cumulants, Eq. G2017 43-48 with omega_6..8 = 1: (1/omega_1 - 1/2) rho A, B
    const R g = (frcp(w1) - R(0.5)) * rho;
    const R ga = g * A, gb = -K.c13 * g * B * (-R(3) * w1 * ir);
    const R e43 = K.c23 * ga * (Dxu - R(2) * Dyv + Dzw);
    const R e44 = K.c23 * ga * (K.eq44 ? (Dxu + Dyv - R(2) * Dzw) : (Dyv + Dzw - R(2) * Dxu));
    const R e45 = -K.c43 * ga * (Dxu + Dyv + Dzw);
    const R c220 = K.c13 * (e43 + e44 + e45), c202 = K.c13 * (e45 - e43), c022 = K.c13 * (e45 - e44);
    const R c211 = gb * c.kyz, c121 = gb * c.kxz, c112 = gb * c.kxy;

    // back to central moments, Eq. G2015 81-84 (orders 5 and 6 from products)
    R s[27];
#define S_(i, j, l) s[9 * (i) + 3 * (j) + (l)]
    S_(0, 0, 0) = m0;
    S_(1, 0, 0) = -c.k100;
    S_(0, 1, 0) = -c.k010;
    S_(0, 0, 1) = -c.k001;
    S_(2, 0, 0) = s200, S_(0, 2, 0) = s020, S_(0, 0, 2) = s002;
    S_(1, 1, 0) = s110, S_(1, 0, 1) = s101, S_(0, 1, 1) = s011;
    S_(1, 2, 0) = s120, S_(1, 0, 2) = s102, S_(2, 1, 0) = s210;
    S_(0, 1, 2) = s012, S_(2, 0, 1) = s201, S_(0, 2, 1) = s021, S_(1, 1, 1) = s111;
    S_(2, 1, 1) = c211 + (s200 * s011 + R(2) * s101 * s110) * ir;
    S_(1, 2, 1) = c121 + (s020 * s101 + R(2) * s110 * s011) * ir;
    S_(1, 1, 2) = c112 + (s002 * s110 + R(2) * s011 * s101) * ir;
    S_(2, 2, 0) = c220 + (s020 * s200 + R(2) * s110 * s110) * ir;
    S_(0, 2, 2) = c022 + (s002 * s020 + R(2) * s011 * s011) * ir;
    S_(2, 0, 2) = c202 + (s200 * s002 + R(2) * s101 * s101) * ir;
    S_(1, 2, 2) = (s020 * s102 + s002 * s120 + R(4) * s011 * s111 + R(2) * (s110 * s012 + s101 * s021)) * ir;
    S_(2, 1, 2) = (s002 * s210 + s200 * s012 + R(4) * s101 * s111 + R(2) * (s011 * s201 + s110 * s102)) * ir;
    S_(2, 2, 1) = (s200 * s021 + s020 * s201 + R(4) * s110 * s111 + R(2) * (s101 * s120 + s011 * s210)) * ir;
    S_(2, 2, 2) = (R(4) * s111 * s111 + s200 * S_(0, 2, 2) + s020 * S_(2, 0, 2) + s002 * S_(2, 2, 0) +
                   R(4) * (s011 * S_(2, 1, 1) + s101 * S_(1, 2, 1) + s110 * S_(1, 1, 2)) +
                   R(2) * (s120 * s102 + s210 * s012 + s201 * s021)) * ir -
                  (R(16) * s110 * s101 * s011 + R(4) * (s101 * s101 * s020 + s011 * s011 * s200 + s110 * s110 * s002) +
                   R(2) * s200 * s020 * s002) * ir * ir;
#undef S_
    // x, y, z passes of the backward transform
    const R uxx = ux * ux, uxm = uxx - ux, uxp = uxx + ux;
    R X[27];
#pragma unroll
    for (int bc = 0; bc < 9; ++bc) inv_gen(s[bc], s[9 + bc], s[18 + bc], ux, uxx, uxm, uxp, X[bc], X[9 + bc], X[18 + bc]);
    const R uyy = uy * uy, uym = uyy - uy, uyp = uyy + uy;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int oz = 0; oz < 3; ++oz)
            inv_gen(X[9 * a + oz], X[9 * a + 3 + oz], X[9 * a + 6 + oz], uy, uyy, uym, uyp, s[9 * a + oz],
                    s[9 * a + 3 + oz], s[9 * a + 6 + oz]);
    const R uzz = uz * uz, uzm = uzz - uz, uzp = uzz + uz;
#pragma unroll
    for (int ab = 0; ab < 9; ++ab)
        inv_gen(s[3 * ab], s[3 * ab + 1], s[3 * ab + 2], uz, uzz, uzm, uzp, X[3 * ab], X[3 * ab + 1], X[3 * ab + 2]);
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = X[CELL3(q)];
}

// BGK with the closed-form equilibrium (fast variant).
template <class R>
LBM_HD void srt_fast(const Consts<R>& K, R* f, R rho, R ux, R uy, R uz, R omega)
{
    const R u2 = R(1.5) * (ux * ux + uy * uy + uz * uz);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const R uc = R(CX[q]) * ux + R(CY[q]) * uy + R(CZ[q]) * uz;
        const R fe = K.w[WCLASS[q]] * rho * (R(1) + R(3) * uc + R(4.5) * uc * uc - u2);
        f[q] = f[q] + omega * (fe - f[q]);
    }
}

}  // namespace lbm
