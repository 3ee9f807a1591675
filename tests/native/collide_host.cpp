// Host build of the core's per-cell collision header (lb_amd/csrc/collide.hpp)
// for the CPU test suite: checks the product arithmetic against the oracle
// without a GPU.  Built with -ffp-contract=off, like the oracle.
#include <cstdint>

#include "collide.hpp"

using namespace lbm;

template <class R>
static Consts<R> consts(uint32_t compat)
{
    return make_consts<R>(compat);
}

template <class R>
static void run(int coll, int fast, uint32_t compat, const double* fin, double rho, double ux, double uy, double uz,
                double om, double* out)
{
    const Consts<R> K = consts<R>(compat);
    R f[Q];
    for (int q = 0; q < Q; ++q) f[q] = (R)fin[q];
    // coll: LBM_COLLISION_*; LBM_COLLISION_COMBINED stands for the Combined
    // operator's 2017 branch (a cell whose flag is not 1)
    if (!fast) {
        if (coll == LBM_COLLISION_CUM2015) cum_literal<CV_2015>(K, f, (R)rho, (R)ux, (R)uy, (R)uz, (R)om);
        else if (coll == LBM_COLLISION_CUM2017) cum_literal<CV_2017>(K, f, (R)rho, (R)ux, (R)uy, (R)uz, (R)om);
        else if (coll == LBM_COLLISION_COMBINED) cum_literal<CV_2017C>(K, f, (R)rho, (R)ux, (R)uy, (R)uz, (R)om);
        else srt_literal(K, f, (R)rho, (R)ux, (R)uy, (R)uz, (R)om);
    } else {
        const Raw2<R> r = raw_moments(f);
        const Cen2<R> c = central2(r, (R)ux, (R)uy, (R)uz);
        R in[Q];
        for (int q = 0; q < Q; ++q) in[q] = f[q];
        if (coll == LBM_COLLISION_CUM2015) cum2015_fast(K, c, r.m0, (R)rho, (R)ux, (R)uy, (R)uz, (R)om, f);
        else if (coll == LBM_COLLISION_CUM2017)
            cum2017_fast<CV_2017>(K, in, c, r.m0, (R)rho, (R)ux, (R)uy, (R)uz, (R)om, f);
        else if (coll == LBM_COLLISION_COMBINED)
            cum2017_fast<CV_2017C>(K, in, c, r.m0, (R)rho, (R)ux, (R)uy, (R)uz, (R)om, f);
        else srt_fast(K, f, (R)rho, (R)ux, (R)uy, (R)uz, (R)om);
    }
    for (int q = 0; q < Q; ++q) out[q] = (double)f[q];
}

extern "C" {

void host_collide(int coll, int fast, int use_double, uint32_t compat, const double* f, double rho, double ux,
                  double uy, double uz, double om, double* out)
{
    if (use_double) run<double>(coll, fast, compat, f, rho, ux, uy, uz, om, out);
    else run<float>(coll, fast, compat, f, rho, ux, uy, uz, om, out);
}

void host_moments(int use_double, const double* fin, double* out4)
{
    if (use_double) {
        double f[Q], r, x, y, z;
        for (int q = 0; q < Q; ++q) f[q] = fin[q];
        moments_literal(f, r, x, y, z);
        out4[0] = r; out4[1] = x; out4[2] = y; out4[3] = z;
    } else {
        float f[Q], r, x, y, z;
        for (int q = 0; q < Q; ++q) f[q] = (float)fin[q];
        moments_literal(f, r, x, y, z);
        out4[0] = r; out4[1] = x; out4[2] = y; out4[3] = z;
    }
}

double host_omega_les(int fast, int use_double, uint32_t compat, const double* fin, double rho, double ux, double uy,
                      double uz, double tau, double cles)
{
    if (use_double) {
        const Consts<double> K = consts<double>(compat);
        double f[Q];
        for (int q = 0; q < Q; ++q) f[q] = fin[q];
        if (!fast) return omega_les_literal(K, f, rho, ux, uy, uz, tau, cles);
        const Raw2<double> r = raw_moments(f);
        return omega_les_fast(central2(r, ux, uy, uz), rho, tau, cles);
    }
    const Consts<float> K = consts<float>(compat);
    float f[Q];
    for (int q = 0; q < Q; ++q) f[q] = (float)fin[q];
    if (!fast) return omega_les_literal(K, f, (float)rho, (float)ux, (float)uy, (float)uz, (float)tau, (float)cles);
    const Raw2<float> r = raw_moments(f);
    return omega_les_fast(central2(r, (float)ux, (float)uy, (float)uz), (float)rho, (float)tau, (float)cles);
}
}

#include "features.hpp"

extern "C" {

// Closure features of a strain tensor (xx, yy, zz, xy, xz, yz) and of a cell's
// central second moments, in double or float.
void host_strain_features(int use_double, const double* s6, double rho, double tau0, double* out12)
{
    if (use_double) {
        double x[NFEAT];
        strain_features(s6[0], s6[1], s6[2], s6[3], s6[4], s6[5], rho, tau0, x);
        for (int k = 0; k < NFEAT; ++k) out12[k] = x[k];
    } else {
        float x[NFEAT];
        strain_features((float)s6[0], (float)s6[1], (float)s6[2], (float)s6[3], (float)s6[4], (float)s6[5],
                        (float)rho, (float)tau0, x);
        for (int k = 0; k < NFEAT; ++k) out12[k] = x[k];
    }
}

void host_cell_features(const double* fin, double tau0, double cles, double* out12, double* pin_rho,
                        double* omega_s)
{
    double f[Q];
    for (int q = 0; q < Q; ++q) f[q] = fin[q];
    const Raw2<double> r = raw_moments(f);
    const double rho = r.m0, ux = r.jx / rho, uy = r.jy / rho, uz = r.jz / rho;
    const Cen2<double> c = central2(r, ux, uy, uz);
    cell_features(c, rho, tau0, cles, out12, *pin_rho, *omega_s);
}

}  // extern "C"
