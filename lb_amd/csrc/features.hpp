// features.hpp -- inputs of the learned sub-grid closure, computed per cell
// from the non-equilibrium stress the collision already holds.
//
// Header-only, host and device.  The closure network maps 12 local
// quantities to a coefficient C(x) that replaces the Smagorinsky constant in
// the reference's rate formula (OmegaLES.h:103-131):
//   tau(x) = tau/2 + sqrt(tau^2 + 18 C(x) |Pi_neq| / rho) / 2,  omega = 1/tau(x).
// With that formula the coefficient is C = nu_t / |S|_F (lattice units,
// filter width one cell), which is what the training targets measure.
//
// The strain rate is estimated from Pi_neq = -(2/3) rho tau_s S with tau_s
// the Smagorinsky rate of the same cell (the baseline closure's own estimate),
// so the features need nothing beyond the cell's moments.
#pragma once

#include <cmath>

#include "collide.hpp"

namespace lbm {

constexpr int NFEAT = 12;

// x[0]  log10 |S|_F                   x[7]  tr S / |S|
// x[1..6] S_xx, S_yy, S_zz, S_xy, S_xz, S_yz over |S|
// x[8]  sqrt(6) tr(D^3) / |D|^3 (D = deviatoric part; shape, in [-1, 1])
// x[9]  100 (rho - 1)                 x[10] log10(|S| / nu0), nu0 = (tau0 - 1/2)/3
// x[11] |D| / |S|
// Strain components with |S| = 0 give x[0] = -30 (outside any envelope).
template <class R>
LBM_HD void strain_features(R sxx, R syy, R szz, R sxy, R sxz, R syz, R rho, R tau0, R* x)
{
    const R n2 = sxx * sxx + syy * syy + szz * szz + R(2) * (sxy * sxy + sxz * sxz + syz * syz);
    const R n = sqrt_r(n2);
    const R inv = n > R(0) ? R(1) / n : R(0);
    x[0] = n > R(0) ? log10_r(n) : R(-30);
    x[1] = sxx * inv;
    x[2] = syy * inv;
    x[3] = szz * inv;
    x[4] = sxy * inv;
    x[5] = sxz * inv;
    x[6] = syz * inv;
    const R tr = sxx + syy + szz;
    x[7] = tr * inv;
    const R t3 = tr * R(1.0 / 3.0);
    const R dxx = sxx - t3, dyy = syy - t3, dzz = szz - t3;
    const R d2 = dxx * dxx + dyy * dyy + dzz * dzz + R(2) * (sxy * sxy + sxz * sxz + syz * syz);
    const R dn = sqrt_r(d2);
    const R d3 = dxx * dxx * dxx + dyy * dyy * dyy + dzz * dzz * dzz + R(3) * dxx * (sxy * sxy + sxz * sxz) +
                 R(3) * dyy * (sxy * sxy + syz * syz) + R(3) * dzz * (sxz * sxz + syz * syz) + R(6) * sxy * sxz * syz;
    x[8] = dn > R(0) ? R(2.449489742783178) * d3 / (dn * dn * dn) : R(0);
    x[9] = R(100) * (rho - R(1));
    x[10] = n > R(0) ? x[0] - log10_r((tau0 - R(0.5)) * R(1.0 / 3.0)) : R(-30);
    x[11] = dn * inv;
}

// |Pi_neq| / rho from the central second moments (the same terms as
// cell_features and omega_les_fast).
template <class R>
LBM_HD R pin_over_rho(const Cen2<R>& c, R rho)
{
    LBM_FP_STRICT
    const R third = rho * R(1.0 / 3.0);
    const R pxx = c.kxx - third, pyy = c.kyy - third, pzz = c.kzz - third;
    const R n2 = pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (c.kxy * c.kxy + c.kxz * c.kxz + c.kyz * c.kyz);
    return sqrt_r(n2) / rho;
}

// Features of a lattice cell from its central second moments.  Also returns
// |Pi_neq| / rho (the argument the rate formula needs) and the Smagorinsky
// rate omega_s (the fallback).
template <class R>
LBM_HD void cell_features(const Cen2<R>& c, R rho, R tau0, R cles, R* x, R& pin_rho, R& omega_s)
{
    LBM_FP_STRICT
    const R third = rho * R(1.0 / 3.0);
    const R pxx = c.kxx - third, pyy = c.kyy - third, pzz = c.kzz - third;
    const R n2 = pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (c.kxy * c.kxy + c.kxz * c.kxz + c.kyz * c.kyz);
    pin_rho = sqrt_r(n2) / rho;
    const R tau_s = R(0.5) * tau0 + R(0.5) * sqrt_r(tau0 * tau0 + R(18) * cles * sqrt_r(n2) / rho);
    omega_s = R(1) / tau_s;
    const R k = R(-1.5) / (rho * tau_s);
    strain_features(k * pxx, k * pyy, k * pzz, k * c.kxy, k * c.kxz, k * c.kyz, rho, tau0, x);
}

// Rate from a closure coefficient C: the reference's formula with C(x).
template <class R>
LBM_HD R omega_from_coefficient(R C, R pin_rho, R tau0)
{
    return R(1) / (R(0.5) * tau0 + R(0.5) * sqrt_r(tau0 * tau0 + R(18) * C * pin_rho));
}

}  // namespace lbm
