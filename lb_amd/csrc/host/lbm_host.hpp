// lbm_host.hpp -- C++ host side of the collide-stream core.
//
// Keeps the reference's setup API (class and member names, call order and
// arithmetic of 3D/simulation/src/{data,geometry,postprocesors,solvers} and
// 3D/preprocess/src/geometry) on top of the C ABI in include/lbm_core.h:
//
//   geometryHandlerOFF / geometryMesherOFF   OFF -> 0/1 grid (preprocess)
//   geometryMesherBoundary                   0/1 grid -> solver flags + BC lists
//   outputerMesh                             flag-grid text file in/out
//   outputerVTK                              legacy-VTK writers, probe CSV
//   NonDimensiolnaliseFactorsVelocity        physical -> lattice units
//   SolverTurbulentLES                       time loop over lbm_step
//
// The reference's device arrays (df, df_post, rho, u, omega) live inside one
// lbm_ctx; LBMData keeps the host-side arrays the setup path builds.
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/lbm_core.h"

namespace lbmh {

using RealType = float;  // LBMTraits::RealType (traits/LBMTraits.h:72)

struct d3 {
    double x, y, z;
};

// TNL::Containers::StaticVector<3, float> in the reference (LBMTraits.h:29)
struct Vector {
    float x = 0.f, y = 0.f, z = 0.f;
    bool operator==(const Vector& o) const { return x == o.x && y == o.y && z == o.z; }
};

struct Vertex {
    int x, y, z;
};

// Boundary PODs (LBMTraits.h:31-61).
struct boundaryConditionWall {
    Vertex vertex;
    bool regular;
};
struct boundaryConditionInlet {
    Vertex vertex;
    Vector normal;
    Vector velocity;
    bool regular;
    int kind;  // LBM_BC_INLET_MOVING_WALL or LBM_BC_INLET_EQUILIBRIUM
};
struct boundaryConditionOutlet {
    Vertex vertex;
    Vector normal;
    float density;
    bool regular;
};
struct boundaryConditionPeriodic {
    Vertex vertex;
    Vector normal;
    int periodicIndex;
    bool regular;
};

// LBMConstants (data/LBMConstants.h:4-113): same field names.
struct LBMConstants {
    // mesh
    RealType BBminx = 0, BBmaxx = 0, BBminy = 0, BBmaxy = 0, BBminz = 0, BBmaxz = 0;
    int dimX_int = 0, dimY_int = 0, dimZ_int = 0;
    RealType resolution_factor = 1, additional_factor = 1;
    d3 point_outside{-1, -1, -1};
    std::string file_name, mesh_name;
    std::string InitFileName;   // InitializationEquilibriumVariables: fields file under results_dir
    int inlet_num = 0, outlet_num = 0, symmetry_num = 0, periodic_num = 0, wall_num = 0;
    // physical
    RealType ny_fyz = 0, u_guess_fyz = 0, L_fyz = 0, rho_fyz = 1, Re = 0;
    RealType Fx_fyz = 0, Fy_fyz = 0, Fz_fyz = 0, conversion_factor_fyz = 1;
    // lattice
    RealType tau = 0, ny = 0, omega = 0, U_lb = 0.05f;
    RealType Cl = 1, Ct = 1, Cm = 1, Cu_inverse = 1, Crho = 1, Cpressure = 1, Cu = 1;
    Vector VelocityInit;
    RealType CLES = 0.173f;  // the reference's const member (:70); settable here
    RealType omegaDumpingLow = -1.f, omegaDumpingHigh = -1.f;  // LBMConstants.h:74-76
    int outlet_sponge = 0;   // 1: the ...Omega outlets' omega damping (outletOmega)
    int wall_mesh = 0;       // 1: BounceBackWallHalfMesh (flag-2 cells) instead of BounceBackWallHalf
    int async_output = 1;    // 1: VTK output from snapshots copied while the next steps run
    // simulation
    RealType err = 0;
    RealType time = 0, plot_every = -1.f, err_every = -1.f;
    int plot_every_it = -1, err_every_it = -1, iterations = 0;
    int iterationsMomentAvg = 0;
    int probe_every_it = -1, probe_iterations = -1;
    Vector ProbeLocation;
    int ProbeLocationLat[3] = {0, 0, 0};
    int Nvel = 27;
    const RealType cs = 1.f / std::sqrt(3.f);  // float overload, as the reference (:103)
    const RealType cs2 = 1.f / 3.f;
    const RealType cs4 = 1.f / 9.f;
    // core policies (the reference selects them as template arguments, main.cu:37-52)
    int collision = LBM_COLLISION_CUM2015;
    int closure = LBM_CLOSURE_SMAGORINSKY;
    int storage = LBM_STORAGE_F32;
    int arith = LBM_ARITH_FAST;
    uint32_t compat = LBM_COMPAT_REFERENCE;
    int periodic[3] = {0, 0, 0};  // index-wrap axes
    std::string results_dir = "results/";
};

// LBMData (data/LBMData.h): host arrays of the setup path and the device lattice.
struct LBMData {
    std::vector<int32_t> meshFluidHost;  // [z][y][x]
    std::vector<boundaryConditionWall> meshBoundaryWallHost;
    std::vector<boundaryConditionInlet> meshBoundaryInletHost;
    std::vector<boundaryConditionOutlet> meshBoundaryOutletHost;
    std::vector<boundaryConditionPeriodic> meshBoundaryPeriodicHost;
    int nx = 0, ny = 0, nz = 0;
    lbm_ctx* ctx = nullptr;  // device lattice (df/df_post/rho/u/omega of the reference)
    std::vector<double> rho_out, ux_out, uy_out, uz_out, omega_out;
    // rhoTimeAvg / uTimeAvg after momentAvg (MomentTimeAvg.h:47-77), lattice units
    std::vector<double> rho_avg, ux_avg, uy_avg, uz_avg;
    int64_t time_avg_count = 0;

    void setSizes(int x, int y, int z)
    {
        nx = x;
        ny = y;
        nz = z;
        meshFluidHost.assign((size_t)x * y * z, 0);
    }
    size_t index(int i, int j, int k) const { return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k); }
    int32_t& mesh(int i, int j, int k) { return meshFluidHost[index(i, j, k)]; }
    int32_t mesh(int i, int j, int k) const { return meshFluidHost[index(i, j, k)]; }
    ~LBMData();
};

// geometryObjectCuboid (geometry/geometryObjectCuboid.h:9-52)
class geometryObjectCuboid {
public:
    geometryObjectCuboid(d3 X1, d3 X2, d3 X3, int id);
    bool isInside(d3 X_ask) const;
    int id;

private:
    double x_min, y_min, z_min, x_max, y_max, z_max;
};

// geometryHandlerOFF (preprocess/src/geometry/geometryHandlerOFF.h): OFF
// polyhedron and the segment-parity inside test, with a bounding-volume
// hierarchy over the triangles in place of CGAL's AABB tree.
class geometryHandlerOFF {
public:
    geometryHandlerOFF();
    ~geometryHandlerOFF();
    // error: 0 ok, 1 empty polyhedron, 2 parse failure (the reference's codes, :79-97)
    void polyhedronFromFile(LBMConstants& Constants, bool verbose);
    void polyhedronBbox(bool verbose);
    void writeToConstants(LBMConstants& Constants) const;
    bool polyhedronInside(d3 pnt_ask, d3 pnt_ref) const;
    // polyhedronInside for every cell centre of a grid, on the GPU (lbm_voxelize)
    // with this handler's hierarchy; grid[cell] = 1 inside.  Returns an lbm_status.
    int insideGridDevice(const int32_t dims[3], double additional, double resolution, const double bbmin[3],
                         d3 pnt_ref, int32_t* grid) const;
    int error = 0;
    d3 d3_min_BB{0, 0, 0}, d3_max_BB{0, 0, 0};
    size_t triangles() const;

    struct Impl;

private:
    std::unique_ptr<Impl> impl;
};

// geometryMesherOFF (preprocess/src/geometry/geometryMesherOFF.h:22-150)
class geometryMesherOFF {
public:
    geometryMesherOFF(LBMConstants& C, LBMData& D) : Constants(C), Data(D) {}
    void initialization(bool verbose);
    // threads: CPU threads (0 = all); < 0 runs the inside tests on the GPU
    void meshingPolyhedron(const geometryHandlerOFF& handler, d3 pnt_outside, bool verbose, int threads = 0);
    int device_status = 0;   // lbm_status of the last GPU run

private:
    LBMConstants& Constants;
    LBMData& Data;
};

// geometryMesherBoundary (geometry/geometryMesherBoundary.h:24-668)
class geometryMesherBoundary {
public:
    geometryMesherBoundary(LBMConstants& C, LBMData& D) : Constants(C), Data(D) {}
    // returns false (and leaves the grid untouched) where the reference throws
    bool meshingBoundaryConditionInletUniform(const geometryObjectCuboid& cuboid, const Vector& normal_out,
                                              const Vector& velocity_in, bool verbose,
                                              int kind = LBM_BC_INLET_EQUILIBRIUM);
    // meshingBoundaryConditionInletParaboloidCircle / ...Rectangle (:111-274):
    // the cuboid's cells become inlet nodes (the uniform rule) and every inlet
    // entry whose centre lies in the cuboid gets the parabolic velocity
    // against its normal.  The rectangle version sets the extent along the
    // normal to 1, in the caller's variables too, as the reference does.
    void meshingBoundaryConditionInletParaboloidCircle(const geometryObjectCuboid& cuboid, const d3& inletCenter,
                                                       RealType inletRadius, const Vector& normal_out,
                                                       RealType mean_velocity_in, bool verbose,
                                                       int kind = LBM_BC_INLET_EQUILIBRIUM);
    void meshingBoundaryConditionInletParaboloidRectangle(const geometryObjectCuboid& cuboid, const d3& inletCenter,
                                                          RealType& inletDimX, RealType& inletDimY,
                                                          RealType& inletDimZ, const Vector& normal_out,
                                                          RealType mean_velocity_in, bool verbose,
                                                          int kind = LBM_BC_INLET_EQUILIBRIUM);
    void meshingBoundaryConditionOutlet(const geometryObjectCuboid& cuboid, const Vector& normal_out, RealType density,
                                        bool verbose);
    void meshingBoundaryConditionPeriodic(const geometryObjectCuboid& cuboid, const Vector& normal_out,
                                          int periodicIndex, bool verbose);
    void meshingBoundaryWall(bool verbose);
    void compileBoundaryArrayInlets(bool verbose);
    void compileBoundaryArrayOutlets(bool verbose);
    void compileBoundaryArrayWall(bool verbose);
    void compileBoundaryArrayPeriodic(bool verbose);
    // Build the core's flag grid + BC list (the reference uploads them here, :602-617).
    void arrayTransfer(bool verbose, std::vector<lbm_bc_node>& bc_out) const;

    std::vector<boundaryConditionInlet> boundary_vector_inlet;
    std::vector<boundaryConditionOutlet> boundary_vector_outlet;
    std::vector<boundaryConditionPeriodic> boundary_vector_periodic;
    std::vector<boundaryConditionWall> boundary_vector_wall;

private:
    d3 cellCentre(int i, int j, int k) const;
    // the uniform rule's cell selection without its velocity check
    int meshInletCells(const geometryObjectCuboid& cuboid, const Vector& normal_out, const Vector& velocity_in,
                       int kind);
    LBMConstants& Constants;
    LBMData& Data;
};

// The field reader of InitializationEquilibriumVariables
// (initializations/InitializationEquilibriumVariables.h:57-116): after the first
// line containing "LOOKUP_TABLE default", nx*ny*nz rho values, then after the
// first line containing "VECTORS U double", nx*ny*nz velocity triples, x
// fastest, each parsed into RealType.  Returns 0, 1 (cannot open), 2 (rho
// short), 3 (u short).
int readInitFields(const std::string& path, int nx, int ny, int nz, std::vector<double>& rho,
                   std::vector<double>& ux, std::vector<double>& uy, std::vector<double>& uz);

// outputerMesh (postprocesors/outputerMesh.h:10-213)
struct outputerMesh {
    static bool MeshMatrixOut(const LBMData& Data, const LBMConstants& Constants, const std::string& path);
    static bool MeshMatrixIn(LBMData& Data, LBMConstants& Constants, const std::string& path, bool verbose);
};

// outputerVTK (postprocesors/outputerVTK.h:14-852)
struct outputerVTK {
    static bool MeshVTK(const LBMData& Data, const LBMConstants& Constants, const std::string& filename);
    // physical units (rho*Crho, p = rho cs^2 * Cpressure, u*Cu)
    // fetch = false: the fields are already in Data.rho_out .. uz_out (a snapshot)
    static bool variablesVTK(LBMData& Data, const LBMConstants& Constants, int step, bool verbose, bool fetch = true);
    static bool variablesLatticeVTK(LBMData& Data, const LBMConstants& Constants, int step, bool verbose);
    static bool omegaVTK(LBMData& Data, const LBMConstants& Constants, bool verbose);
    // variablesTimeAvgVTK (outputerVTK.h:281-345): the time-averaged rho, p, u
    static bool variablesTimeAvgVTK(LBMData& Data, const LBMConstants& Constants, bool verbose);
    // variablesTimeAvgLatticeVTK (:449-576), distributionFunctionVTK (:660-740), tauVTK (:742-822)
    static bool variablesTimeAvgLatticeVTK(LBMData& Data, const LBMConstants& Constants, bool verbose);
    static bool distributionFunctionVTK(LBMData& Data, const LBMConstants& Constants, int velocity, bool verbose);
    static bool tauVTK(LBMData& Data, const LBMConstants& Constants, bool verbose);
    static bool probeWriteOut(const LBMConstants& Constants, RealType velX, RealType velY, RealType velZ,
                              RealType rho);
};

// NonDimensiolnaliseFactorsVelocity (solvers/nonDimensionalisions/...:20-131)
struct NonDimensiolnaliseFactorsVelocity {
    static void nonDimensionalize(LBMConstants& Constants, bool verbose);
    static void nonDimensionalizeInlet(LBMData& Data, const LBMConstants& Constants);
    static void nonDimensionalizeOutlet(LBMData& Data, const LBMConstants& Constants);
};

// Wall-clock phase timer (TNL::Timer in the reference).
struct Timer {
    double total = 0.0, t0 = 0.0;
    bool running = false;
    void start();
    void stop();
    double getRealTime() const { return total; }
};

// SolverTurbulentLES (solvers/SolverTurbulentLES.h:82-365): the whole per-step
// operator chain runs inside lbm_step; the loop keeps the reference's
// cadences (error norm + NaN break, VTK output, probe).
class SolverTurbulentLES {
public:
    SolverTurbulentLES(LBMConstants& C, LBMData& D) : Constants(C), Data(D) {}
    void convertToLattice(bool verbose);
    // returns an lbm_status; creates the device lattice and initialises it
    int initializeSimulation(bool verbose);
    int runSimulation();
    void initializeSimulationData(bool verbose);
    int probeWrite();
    int probeFlush();
    // ErrorQuadratic::errorEvaluation (errorEvaluations/ErrorQuadratic.h:33-90)
    int errorEvaluation();
    bool write_output = true;  // VTK at plot_every_it
    Timer timer_loop, timer_collision, timer_streaming, timer_bounceback, timer_momentsUpdate, timer_err,
        timer_output, timer_dumping, timer_timeAvg;
    std::vector<double> err_history;

private:
    LBMConstants& Constants;
    LBMData& Data;
};

}  // namespace lbmh
