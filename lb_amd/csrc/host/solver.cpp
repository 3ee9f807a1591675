// solver.cpp -- unit conversion and the SolverTurbulentLES time loop.
//
// NonDimensiolnaliseFactorsVelocity follows
// 3D/simulation/src/solvers/nonDimensionalisions/NonDimensiolnaliseFactorsVelocity.h:20-131
// in float arithmetic, including its tau = nu/3 + 1/2.  SolverTurbulentLES
// follows 3D/simulation/src/solvers/SolverTurbulentLES.h:106-324: the operator
// chain of one step (omega, collision, streaming, boundaries, moments) is a
// single lbm_step launch; the loop keeps the reference's cadences for the
// probe, the error norm with its NaN break, and VTK output.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>

#include "lbm_host.hpp"

namespace lbmh {

LBMData::~LBMData()
{
    if (ctx) lbm_destroy(ctx);
}

void Timer::start()
{
    t0 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    running = true;
}

void Timer::stop()
{
    if (!running) return;
    total += std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t0;
    running = false;
}

void NonDimensiolnaliseFactorsVelocity::nonDimensionalize(LBMConstants& C, bool verbose)
{
    if (C.U_lb > 0.1f)
        std::cout << "\n Lattice speed should not be higher than 0.1 due to stability close to speed of sound.\n";
    C.L_fyz = std::abs(C.BBmaxx - C.BBminx) * C.conversion_factor_fyz;
    C.Cl = C.L_fyz / C.dimX_int;
    const RealType U_fyz = C.u_guess_fyz;
    C.Cu = U_fyz / C.U_lb;
    C.Cu_inverse = 1 / C.Cu;
    C.Ct = C.Cl / C.Cu;
    C.Crho = C.rho_fyz;
    C.Cm = C.Crho * C.Cl * C.Cl * C.Cl;
    C.Cpressure = C.Cm / C.Cl / C.Ct / C.Ct;
    C.Re = U_fyz * C.L_fyz / C.ny_fyz;
    C.ny = C.ny_fyz * C.Ct / C.Cl / C.Cl;
    C.tau = C.ny / 3.f + 0.5f;  // as the reference writes it (nonDimensionalize :68)
    if (verbose) std::cout << "\n- Re is " << C.Re << "\n- Nu is " << C.ny << "\n- Tau is " << C.tau << "\n";
    C.omega = 1.f / C.tau;
}

void NonDimensiolnaliseFactorsVelocity::nonDimensionalizeInlet(LBMData& D, const LBMConstants& C)
{
    for (auto& b : D.meshBoundaryInletHost) {
        b.velocity.x = b.velocity.x * C.Cu_inverse;
        b.velocity.y = b.velocity.y * C.Cu_inverse;
        b.velocity.z = b.velocity.z * C.Cu_inverse;
    }
}

void NonDimensiolnaliseFactorsVelocity::nonDimensionalizeOutlet(LBMData& D, const LBMConstants& C)
{
    for (auto& b : D.meshBoundaryOutletHost) b.density = b.density / C.Cm * C.Cl * C.Cl * C.Cl;
}

void SolverTurbulentLES::convertToLattice(bool verbose) { NonDimensiolnaliseFactorsVelocity::nonDimensionalize(Constants, verbose); }

// initializeSimulationData (:276-324)
void SolverTurbulentLES::initializeSimulationData(bool verbose)
{
    if (Constants.plot_every_it == -1 && Constants.plot_every != -1)
        Constants.plot_every_it = (int)std::ceil(Constants.plot_every / Constants.Ct);
    if (Constants.err_every_it == -1 && Constants.err_every != -1)
        Constants.err_every_it = (int)std::ceil(Constants.err_every / Constants.Ct);
    if (Constants.iterations <= 0) Constants.iterations = (int)std::ceil(Constants.time / Constants.Ct);
    if (Constants.probe_every_it > 0) {
        const float ex = 2 * (Constants.additional_factor / Constants.resolution_factor);
        Constants.ProbeLocationLat[0] =
            (int)std::round(Constants.ProbeLocation.x / (Constants.BBmaxx - Constants.BBminx + ex) * Constants.dimX_int);
        Constants.ProbeLocationLat[1] =
            (int)std::round(Constants.ProbeLocation.y / (Constants.BBmaxy - Constants.BBminy + ex) * Constants.dimY_int);
        Constants.ProbeLocationLat[2] =
            (int)std::round(Constants.ProbeLocation.z / (Constants.BBmaxz - Constants.BBminz + ex) * Constants.dimZ_int);
    }
    if (verbose) {
        std::cout << "\nPlotting every " << Constants.plot_every_it << " iterations.\n";
        std::cout << "\nCalculation will run for " << Constants.iterations << " iterations.\n";
    }
}

// initializeSimulation (:106-146): sizes, inlet/outlet scaling, device
// lattice, equilibrium initialisation (InitializationEquilibriumConstVector).
int SolverTurbulentLES::initializeSimulation(bool verbose)
{
    Constants.Nvel = 27;
    initializeSimulationData(verbose);
    NonDimensiolnaliseFactorsVelocity::nonDimensionalizeInlet(Data, Constants);
    NonDimensiolnaliseFactorsVelocity::nonDimensionalizeOutlet(Data, Constants);
    std::vector<lbm_bc_node> bc;
    geometryMesherBoundary mesher(Constants, Data);
    mesher.arrayTransfer(false, bc);
    lbm_desc d;
    std::memset(&d, 0, sizeof(d));
    d.nx = Constants.dimX_int;
    d.ny = Constants.dimY_int;
    d.nz = Constants.dimZ_int;
    for (int a = 0; a < 3; ++a) d.periodic[a] = Constants.periodic[a];
    d.collision = Constants.collision;
    d.closure = Constants.closure;
    d.storage = Constants.storage;
    d.arith = Constants.arith;
    d.compat = Constants.compat;
    d.n_bc = (int32_t)bc.size();
    d.tau = Constants.tau;
    d.nu = Constants.ny;
    d.c_les = Constants.CLES;
    d.outlet_sponge = Constants.outlet_sponge;
    d.wall_mesh = Constants.wall_mesh;
    d.sponge_low = (float)Constants.omegaDumpingLow;
    d.sponge_high = (float)Constants.omegaDumpingHigh;
    d.rho0 = Constants.rho_fyz / Constants.Cm * Constants.Cl * Constants.Cl * Constants.Cl;
    d.flags = Data.meshFluidHost.data();
    d.bc = bc.empty() ? nullptr : bc.data();
    if (Data.ctx) {
        lbm_destroy(Data.ctx);
        Data.ctx = nullptr;
    }
    int st = lbm_create(&d, &Data.ctx);
    if (st != LBM_OK) return st;
    const size_t N = Data.meshFluidHost.size();
    std::vector<double> rho(N, 0.0), ux(N, 0.0), uy(N, 0.0), uz(N, 0.0);
    if (!Constants.InitFileName.empty()) {
        // InitializationEquilibriumVariables (:24-178): rho and u from a VTK
        // file in lattice units, inlet velocity on inlet cells
        const int rs = readInitFields(Constants.results_dir + Constants.InitFileName, Constants.dimX_int,
                                      Constants.dimY_int, Constants.dimZ_int, rho, ux, uy, uz);
        if (rs != 0) {
            std::cerr << "Error reading the initial fields (" << rs << ")." << std::endl;
            return LBM_E_ARG;
        }
    } else {
        // InitializationEquilibriumConstVector: rho = rho_fyz / Cm * Cl^3 and
        // u = u_0 * Cu^-1 on non-solid cells, inlet velocity on inlet cells (:74-109)
        const RealType rho0 = Constants.rho_fyz / Constants.Cm * Constants.Cl * Constants.Cl * Constants.Cl;
        for (size_t i = 0; i < N; ++i) {
            if (Data.meshFluidHost[i] == 0) continue;
            rho[i] = rho0;
            ux[i] = (RealType)(Constants.VelocityInit.x * Constants.Cu_inverse);
            uy[i] = (RealType)(Constants.VelocityInit.y * Constants.Cu_inverse);
            uz[i] = (RealType)(Constants.VelocityInit.z * Constants.Cu_inverse);
        }
    }
    for (const auto& b : Data.meshBoundaryInletHost) {
        const size_t i = Data.index(b.vertex.x, b.vertex.y, b.vertex.z);
        ux[i] = b.velocity.x;
        uy[i] = b.velocity.y;
        uz[i] = b.velocity.z;
    }
    st = lbm_init_equilibrium(Data.ctx, rho.data(), ux.data(), uy.data(), uz.data());
    if (verbose) std::cout << "Initialization of simulation done." << std::endl;
    return st;
}

int readInitFields(const std::string& path, int nx, int ny, int nz, std::vector<double>& rho,
                   std::vector<double>& ux, std::vector<double>& uy, std::vector<double>& uz)
{
    std::ifstream in(path);
    if (!in.is_open()) return 1;
    const size_t N = (size_t)nx * ny * nz;
    rho.assign(N, 0.0);
    ux.assign(N, 0.0);
    uy.assign(N, 0.0);
    uz.assign(N, 0.0);
    std::string line;
    while (std::getline(in, line))
        if (line.find("LOOKUP_TABLE default") != std::string::npos) break;
    for (size_t i = 0; i < N; ++i) {
        RealType v;
        if (!(in >> v)) return 2;
        rho[i] = v;
    }
    while (std::getline(in, line))
        if (line.find("VECTORS U double") != std::string::npos) break;
    for (size_t i = 0; i < N; ++i) {
        RealType a, b, c;
        if (!(in >> a >> b >> c)) return 3;
        ux[i] = a;
        uy[i] = b;
        uz[i] = c;
    }
    return 0;
}

// probeWrite (:150-177): the probe cell's rho, u are recorded on the device
// without waiting; probeFlush writes the records, oldest first, at the next
// point where the loop waits for the device anyway.
int SolverTurbulentLES::probeWrite()
{
    return lbm_probe_record(Data.ctx, Constants.ProbeLocationLat[0], Constants.ProbeLocationLat[1],
                            Constants.ProbeLocationLat[2]);
}

int SolverTurbulentLES::probeFlush()
{
    int64_t n = 0;
    int st = lbm_probe_fetch(Data.ctx, nullptr, 0, &n);
    if (st != LBM_OK || n == 0) return st;
    std::vector<double> v(4 * (size_t)n);
    st = lbm_probe_fetch(Data.ctx, v.data(), n, &n);
    if (st != LBM_OK) return st;
    for (int64_t i = 0; i < n; ++i)
        outputerVTK::probeWriteOut(Constants, (RealType)v[4 * i + 1], (RealType)v[4 * i + 2], (RealType)v[4 * i + 3],
                                   (RealType)v[4 * i]);
    return LBM_OK;
}

// ErrorQuadratic (ErrorQuadratic.h:33-90) on fluid cells with flag == 1,
// including the third-component typo (uy - u_prev.z): reduced on the device
// (lbm_error_quadratic keeps u_prev there), so only the two sums cross to the
// host at the loop's error cadence.
int SolverTurbulentLES::errorEvaluation()
{
    double e = 0.0;
    const int st = lbm_error_quadratic(Data.ctx, &e);
    if (st != LBM_OK && st != LBM_E_NAN) return st;
    Constants.err = (RealType)e;
    err_history.push_back(Constants.err);
    return st;
}

// runSimulation (:179-268).  Steps between two host events are queued as one
// lbm_step call; the event cadences are those of the reference loop.
int SolverTurbulentLES::runSimulation()
{
    if (!Data.ctx) return LBM_E_STATE;
    timer_loop.start();
    int k = 0;
    const int iters = Constants.iterations;
    auto is_event = [&](int kk) {
        const bool probe = Constants.probe_every_it > 0 && kk % Constants.probe_every_it == 0 &&
                           kk >= (iters - Constants.probe_iterations);
        const bool err = Constants.err_every_it > 0 && kk % Constants.err_every_it == 0 && kk != 0;
        const bool plot = write_output && Constants.plot_every_it > 0 && kk % Constants.plot_every_it == 0 && kk;
        return probe || err || plot;
    };
    // time-averaging window: momentAdd after the moment update of every step
    // k > iterations - iterationsMomentAvg (:216-221); a batch of steps never
    // straddles its start, so each batch is either counted or not
    const int avg_from = iters - Constants.iterationsMomentAvg;
    int status = LBM_OK;
    // VTK output at the plot cadence: with async_output the fields are
    // snapshotted behind the queued steps (lbm_snapshot_begin) and the file
    // is written after the next batch of steps has been queued, while the
    // device runs it; otherwise the loop waits for the device at each plot
    const bool async_out = Constants.async_output != 0;
    int pending_plot = -1;
    auto write_pending = [&]() -> int {
        if (pending_plot < 0) return LBM_OK;
        const size_t N = Data.meshFluidHost.size();
        Data.rho_out.resize(N);
        Data.ux_out.resize(N);
        Data.uy_out.resize(N);
        Data.uz_out.resize(N);
        timer_output.start();
        int st = lbm_snapshot_wait(Data.ctx, Data.rho_out.data(), Data.ux_out.data(), Data.uy_out.data(),
                                   Data.uz_out.data(), nullptr, nullptr);
        if (st == LBM_OK) outputerVTK::variablesVTK(Data, Constants, pending_plot, true, false);
        timer_output.stop();
        pending_plot = -1;
        return st;
    };
    while (k < iters) {
        // advance to the end of step k, or further while no host event is due
        int kk = k;
        while (kk + 1 < iters && !is_event(kk) && kk != avg_from) ++kk;
        timer_collision.start();
        status = k > avg_from ? lbm_step_moment_add(Data.ctx, kk - k + 1) : lbm_step(Data.ctx, kk - k + 1);
        timer_collision.stop();
        // the previous plot's file, while the device runs the steps just queued
        if (status == LBM_OK) status = write_pending();
        // wait for the device only where the host reads fields (error norm,
        // synchronous output) and at the end; probe records need no wait
        const bool err_due = Constants.err_every_it > 0 && kk % Constants.err_every_it == 0 && kk != 0;
        const bool plot_due = write_output && Constants.plot_every_it > 0 && kk % Constants.plot_every_it == 0 && kk;
        timer_collision.start();
        if (status == LBM_OK && (err_due || (plot_due && !async_out) || kk + 1 == iters))
            status = lbm_sync(Data.ctx);
        timer_collision.stop();
        if (status != LBM_OK) break;
        k = kk;
        if (Constants.probe_every_it > 0 && k % Constants.probe_every_it == 0 &&
            k >= (iters - Constants.probe_iterations)) {
            timer_output.start();
            status = probeWrite();
            timer_output.stop();
            if (status != LBM_OK) break;
        }
        if (err_due || plot_due) {
            timer_output.start();
            status = probeFlush();
            timer_output.stop();
            if (status != LBM_OK) break;
        }
        if (Constants.err_every_it > 0 && k % Constants.err_every_it == 0 && k != 0) {
            timer_err.start();
            status = errorEvaluation();
            timer_err.stop();
            std::printf("\n err=%e k=%d\n", (double)Constants.err, k);
            if (std::isnan(Constants.err)) {
                std::cout << "\n Error is NaN, breaking out.\n";
                status = LBM_E_NAN;
                break;
            }
        }
        if (write_output && Constants.plot_every_it > 0 && k % Constants.plot_every_it == 0 && k) {
            if (async_out) {
                status = lbm_snapshot_begin(Data.ctx);
                if (status != LBM_OK) break;
                pending_plot = k / Constants.plot_every_it;
            } else {
                timer_output.start();
                outputerVTK::variablesVTK(Data, Constants, k / Constants.plot_every_it, true);
                timer_output.stop();
            }
        }
        k++;
    }
    {
        const int st = write_pending();   // a plot at the last step, or before a break
        if (status == LBM_OK) status = st;
    }
    if (status == LBM_OK || status == LBM_E_NAN) {
        const int st = probeFlush();
        if (status == LBM_OK) status = st;
    }
    // momentAvg and its output (:255-264), when the window counted any state
    int64_t counted = 0;
    if (status == LBM_OK && Constants.iterationsMomentAvg > 0) {
        timer_timeAvg.start();
        const size_t N = Data.meshFluidHost.size();
        Data.rho_avg.resize(N);
        Data.ux_avg.resize(N);
        Data.uy_avg.resize(N);
        Data.uz_avg.resize(N);
        const int st = lbm_moment_avg(Data.ctx, Data.rho_avg.data(), Data.ux_avg.data(), Data.uy_avg.data(),
                                      Data.uz_avg.data(), &counted);
        timer_timeAvg.stop();
        Data.time_avg_count = counted;
        if (st != LBM_OK && st != LBM_E_STATE) status = st;
        if (st == LBM_OK && write_output) {
            timer_output.start();
            outputerVTK::variablesTimeAvgVTK(Data, Constants, true);
            timer_output.stop();
        }
    }
    timer_loop.stop();
    return status;
}

}  // namespace lbmh
