// capi.cpp -- C entry points of liblbm_host.so (include/lbm_host.h).
#include <algorithm>
#include <cstring>
#include <map>
#include <string>

#include "../../../include/lbm_host.h"
#include "lbm_host.hpp"

using namespace lbmh;

struct lbmh_mesher {
    LBMConstants C;
    LBMData D;
    geometryMesherBoundary M{C, D};
    bool compiled = false;
};

struct lbmh_case {
    LBMConstants C;
    LBMData D;
    geometryMesherBoundary M{C, D};
    SolverTurbulentLES S{C, D};
};

static geometryObjectCuboid cuboid(const double* p1, const double* p2, const double* p3, int id)
{
    return geometryObjectCuboid({p1[0], p1[1], p1[2]}, {p2[0], p2[1], p2[2]}, {p3[0], p3[1], p3[2]}, id);
}

static void set_grid(LBMConstants& C, LBMData& D, const int32_t* dims, const int32_t* grid, float res, float add,
                     const float* bbmin, const float* bbmax)
{
    C.dimX_int = dims[0];
    C.dimY_int = dims[1];
    C.dimZ_int = dims[2];
    C.resolution_factor = res;
    C.additional_factor = add;
    C.BBminx = bbmin[0];
    C.BBminy = bbmin[1];
    C.BBminz = bbmin[2];
    if (bbmax) {
        C.BBmaxx = bbmax[0];
        C.BBmaxy = bbmax[1];
        C.BBmaxz = bbmax[2];
    }
    D.setSizes(dims[0], dims[1], dims[2]);
    std::memcpy(D.meshFluidHost.data(), grid, sizeof(int32_t) * D.meshFluidHost.size());
}

extern "C" {

int lbmh_voxelize_off(const char* path, float resolution, float additional, const double outside[3], int threads,
                      int32_t dims[3], float bbmin[3], float bbmax[3], int32_t* grid)
{
    if (!path || !outside || !dims) return 2;
    LBMConstants C;
    LBMData D;
    C.file_name = path;
    C.resolution_factor = resolution;
    C.additional_factor = additional;
    geometryHandlerOFF H;
    H.polyhedronFromFile(C, false);
    if (H.error != 0) return H.error;
    H.polyhedronBbox(false);
    H.writeToConstants(C);
    geometryMesherOFF M(C, D);
    M.initialization(false);
    dims[0] = C.dimX_int;
    dims[1] = C.dimY_int;
    dims[2] = C.dimZ_int;
    if (bbmin) { bbmin[0] = C.BBminx; bbmin[1] = C.BBminy; bbmin[2] = C.BBminz; }
    if (bbmax) { bbmax[0] = C.BBmaxx; bbmax[1] = C.BBmaxy; bbmax[2] = C.BBmaxz; }
    if (!grid) return 0;
    M.meshingPolyhedron(H, {outside[0], outside[1], outside[2]}, false, threads);
    if (threads < 0 && M.device_status != LBM_OK) return 10 + M.device_status;
    std::memcpy(grid, D.meshFluidHost.data(), sizeof(int32_t) * D.meshFluidHost.size());
    return 0;
}

int lbmh_mesh_write(const char* path, const int32_t dims[3], float resolution, float additional,
                    const float bbmin[3], const float bbmax[3], const int32_t* grid)
{
    if (!path || !dims || !grid || !bbmin || !bbmax) return 1;
    LBMConstants C;
    LBMData D;
    set_grid(C, D, dims, grid, resolution, additional, bbmin, bbmax);
    return outputerMesh::MeshMatrixOut(D, C, path) ? 0 : 1;
}

int lbmh_mesh_read(const char* path, int32_t dims[3], float* resolution, float* additional, float bbmin[3],
                   float bbmax[3], int32_t* grid)
{
    if (!path || !dims) return 1;
    LBMConstants C;
    LBMData D;
    if (!outputerMesh::MeshMatrixIn(D, C, path, false)) return 1;
    dims[0] = C.dimX_int;
    dims[1] = C.dimY_int;
    dims[2] = C.dimZ_int;
    if (resolution) *resolution = C.resolution_factor;
    if (additional) *additional = C.additional_factor;
    if (bbmin) { bbmin[0] = C.BBminx; bbmin[1] = C.BBminy; bbmin[2] = C.BBminz; }
    if (bbmax) { bbmax[0] = C.BBmaxx; bbmax[1] = C.BBmaxy; bbmax[2] = C.BBmaxz; }
    if (grid) std::memcpy(grid, D.meshFluidHost.data(), sizeof(int32_t) * D.meshFluidHost.size());
    return 0;
}

int lbmh_mesher_create(const int32_t dims[3], const int32_t* grid01, float resolution, float additional,
                       const float bbmin[3], lbmh_mesher** out)
{
    if (!dims || !grid01 || !bbmin || !out) return 1;
    lbmh_mesher* m = new lbmh_mesher();
    set_grid(m->C, m->D, dims, grid01, resolution, additional, bbmin, nullptr);
    *out = m;
    return 0;
}

int lbmh_mesher_wall(lbmh_mesher* m)
{
    if (!m) return 1;
    m->M.meshingBoundaryWall(false);
    return 0;
}

int lbmh_mesher_inlet(lbmh_mesher* m, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                      const float normal[3], const float velocity[3], int32_t kind)
{
    if (!m) return 1;
    return m->M.meshingBoundaryConditionInletUniform(cuboid(p1, p2, p3, id), {normal[0], normal[1], normal[2]},
                                                     {velocity[0], velocity[1], velocity[2]}, false, kind)
               ? 0
               : 1;
}

int lbmh_mesher_inlet_paraboloid_circle(lbmh_mesher* m, const double p1[3], const double p2[3], const double p3[3],
                                        int32_t id, const float normal[3], const double center[3], float radius,
                                        float mean_velocity, int32_t kind)
{
    if (!m || !center) return 1;
    m->M.meshingBoundaryConditionInletParaboloidCircle(cuboid(p1, p2, p3, id), {center[0], center[1], center[2]}, radius,
                                                       {normal[0], normal[1], normal[2]}, mean_velocity, false, kind);
    return 0;
}

int lbmh_mesher_inlet_paraboloid_rectangle(lbmh_mesher* m, const double p1[3], const double p2[3],
                                           const double p3[3], int32_t id, const float normal[3],
                                           const double center[3], float dims[3], float mean_velocity, int32_t kind)
{
    if (!m || !center || !dims) return 1;
    m->M.meshingBoundaryConditionInletParaboloidRectangle(cuboid(p1, p2, p3, id), {center[0], center[1], center[2]},
                                                          dims[0], dims[1], dims[2], {normal[0], normal[1], normal[2]},
                                                          mean_velocity, false, kind);
    return 0;
}

int lbmh_mesher_outlet(lbmh_mesher* m, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                       const float normal[3], float density)
{
    if (!m) return 1;
    m->M.meshingBoundaryConditionOutlet(cuboid(p1, p2, p3, id), {normal[0], normal[1], normal[2]}, density, false);
    return 0;
}

int lbmh_mesher_periodic(lbmh_mesher* m, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                         const float normal[3], int32_t periodic_index)
{
    if (!m) return 1;
    m->M.meshingBoundaryConditionPeriodic(cuboid(p1, p2, p3, id), {normal[0], normal[1], normal[2]}, periodic_index,
                                          false);
    return 0;
}

int lbmh_mesher_compile(lbmh_mesher* m, int32_t* flags, lbm_bc_node* bc, int32_t* n_bc)
{
    if (!m || !n_bc) return 1;
    if (!m->compiled) {
        m->M.compileBoundaryArrayWall(false);
        m->M.compileBoundaryArrayInlets(false);
        m->M.compileBoundaryArrayPeriodic(false);
        m->M.compileBoundaryArrayOutlets(false);
        m->compiled = true;
    }
    std::vector<lbm_bc_node> list;
    m->M.arrayTransfer(false, list);
    if (flags) std::memcpy(flags, m->D.meshFluidHost.data(), sizeof(int32_t) * m->D.meshFluidHost.size());
    if (bc) {
        if (*n_bc < (int32_t)list.size()) return 1;
        std::memcpy(bc, list.data(), sizeof(lbm_bc_node) * list.size());
    }
    *n_bc = (int32_t)list.size();
    return 0;
}

void lbmh_mesher_destroy(lbmh_mesher* m) { delete m; }

// ---- case ----

int lbmh_case_create(lbmh_case** out)
{
    if (!out) return 1;
    *out = new lbmh_case();
    return 0;
}

void lbmh_case_destroy(lbmh_case* c) { delete c; }

namespace {
struct Field {
    float* f = nullptr;
    int* i = nullptr;
    uint32_t* u = nullptr;
};

Field field(LBMConstants& C, const std::string& n)
{
    static const char* dummy = nullptr;
    (void)dummy;
    Field r;
#define F(name) if (n == #name) { r.f = &C.name; return r; }
#define I(name) if (n == #name) { r.i = &C.name; return r; }
    F(BBminx) F(BBmaxx) F(BBminy) F(BBmaxy) F(BBminz) F(BBmaxz) F(resolution_factor) F(additional_factor)
    F(ny_fyz) F(u_guess_fyz) F(L_fyz) F(rho_fyz) F(Re) F(conversion_factor_fyz) F(tau) F(ny) F(omega) F(U_lb)
    F(Cl) F(Ct) F(Cm) F(Cu_inverse) F(Crho) F(Cpressure) F(Cu) F(CLES) F(err) F(time) F(plot_every) F(err_every)
    F(omegaDumpingLow) F(omegaDumpingHigh) I(outlet_sponge) I(wall_mesh) I(async_output)
    I(dimX_int) I(dimY_int) I(dimZ_int) I(inlet_num) I(outlet_num) I(periodic_num) I(wall_num) I(plot_every_it)
    I(err_every_it) I(iterations) I(iterationsMomentAvg) I(probe_every_it) I(probe_iterations) I(Nvel)
    I(collision) I(closure) I(storage) I(arith)
#undef F
#undef I
    if (n == "compat") { r.u = &C.compat; return r; }
    if (n == "periodic_x") { r.i = &C.periodic[0]; return r; }
    if (n == "periodic_y") { r.i = &C.periodic[1]; return r; }
    if (n == "periodic_z") { r.i = &C.periodic[2]; return r; }
    if (n == "VelocityInit_x") { r.f = &C.VelocityInit.x; return r; }
    if (n == "VelocityInit_y") { r.f = &C.VelocityInit.y; return r; }
    if (n == "VelocityInit_z") { r.f = &C.VelocityInit.z; return r; }
    if (n == "ProbeLocation_x") { r.f = &C.ProbeLocation.x; return r; }
    if (n == "ProbeLocation_y") { r.f = &C.ProbeLocation.y; return r; }
    if (n == "ProbeLocation_z") { r.f = &C.ProbeLocation.z; return r; }
    return r;
}
}  // namespace

int lbmh_case_set(lbmh_case* c, const char* name, double value)
{
    if (!c || !name) return 1;
    Field f = field(c->C, name);
    if (f.f) *f.f = (float)value;
    else if (f.i) *f.i = (int)value;
    else if (f.u) *f.u = (uint32_t)value;
    else return 1;
    return 0;
}

int lbmh_case_get(lbmh_case* c, const char* name, double* value)
{
    if (!c || !name || !value) return 1;
    const std::string n = name;
    if (n == "cs") { *value = c->C.cs; return 0; }
    if (n == "cs2") { *value = c->C.cs2; return 0; }
    Field f = field(c->C, name);
    if (f.f) *value = *f.f;
    else if (f.i) *value = *f.i;
    else if (f.u) *value = *f.u;
    else return 1;
    return 0;
}

int lbmh_read_vtk_fields(const char* path, const int32_t dims[3], double* rho, double* ux, double* uy, double* uz)
{
    if (!path || !dims || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return 4;
    std::vector<double> r, x, y, z;
    const int st = readInitFields(path, dims[0], dims[1], dims[2], r, x, y, z);
    if (st != 0) return st;
    const std::vector<double>* src[4] = {&r, &x, &y, &z};
    double* dst[4] = {rho, ux, uy, uz};
    for (int k = 0; k < 4; ++k)
        if (dst[k]) std::copy(src[k]->begin(), src[k]->end(), dst[k]);
    return 0;
}

int lbmh_case_set_init_file(lbmh_case* c, const char* name)
{
    if (!c || !name) return 1;
    c->C.InitFileName = name;
    return 0;
}

int lbmh_case_set_results_dir(lbmh_case* c, const char* dir)
{
    if (!c || !dir) return 1;
    c->C.results_dir = dir;
    if (!c->C.results_dir.empty() && c->C.results_dir.back() != '/') c->C.results_dir += '/';
    return 0;
}

int lbmh_case_load_mesh(lbmh_case* c, const char* path)
{
    if (!c || !path) return 1;
    return outputerMesh::MeshMatrixIn(c->D, c->C, path, false) ? 0 : 1;
}

int lbmh_case_set_grid(lbmh_case* c, const int32_t dims[3], const int32_t* grid01, float resolution,
                       float additional, const float bbmin[3], const float bbmax[3])
{
    if (!c || !dims || !grid01 || !bbmin || !bbmax) return 1;
    set_grid(c->C, c->D, dims, grid01, resolution, additional, bbmin, bbmax);
    return 0;
}

int lbmh_case_mesh_wall(lbmh_case* c)
{
    if (!c) return 1;
    c->M.meshingBoundaryWall(false);
    return 0;
}

int lbmh_case_inlet(lbmh_case* c, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                    const float normal[3], const float velocity[3], int32_t kind)
{
    if (!c) return 1;
    return c->M.meshingBoundaryConditionInletUniform(cuboid(p1, p2, p3, id), {normal[0], normal[1], normal[2]},
                                                     {velocity[0], velocity[1], velocity[2]}, false, kind)
               ? 0
               : 1;
}

int lbmh_case_inlet_paraboloid_circle(lbmh_case* c, const double p1[3], const double p2[3], const double p3[3],
                                      int32_t id, const float normal[3], const double center[3], float radius,
                                      float mean_velocity, int32_t kind)
{
    if (!c || !center) return 1;
    c->M.meshingBoundaryConditionInletParaboloidCircle(cuboid(p1, p2, p3, id), {center[0], center[1], center[2]}, radius,
                                                       {normal[0], normal[1], normal[2]}, mean_velocity, false, kind);
    return 0;
}

int lbmh_case_inlet_paraboloid_rectangle(lbmh_case* c, const double p1[3], const double p2[3], const double p3[3],
                                         int32_t id, const float normal[3], const double center[3], float dims[3],
                                         float mean_velocity, int32_t kind)
{
    if (!c || !center || !dims) return 1;
    c->M.meshingBoundaryConditionInletParaboloidRectangle(cuboid(p1, p2, p3, id), {center[0], center[1], center[2]},
                                                          dims[0], dims[1], dims[2], {normal[0], normal[1], normal[2]},
                                                          mean_velocity, false, kind);
    return 0;
}

int lbmh_case_write_vtk(lbmh_case* c, const char* which, int32_t index)
{
    if (!c || !which) return 1;
    const std::string w(which);
    bool ok;
    if (w == "variables") ok = outputerVTK::variablesVTK(c->D, c->C, index, false);
    else if (w == "variablesLattice") ok = outputerVTK::variablesLatticeVTK(c->D, c->C, index, false);
    else if (w == "omega") ok = outputerVTK::omegaVTK(c->D, c->C, false);
    else if (w == "tau") ok = outputerVTK::tauVTK(c->D, c->C, false);
    else if (w == "distributionFunction") {
        if (index < 0 || index >= LBM_Q) return 1;
        ok = outputerVTK::distributionFunctionVTK(c->D, c->C, index, false);
    } else if (w == "variablesTimeAvg" || w == "variablesTimeAvgLattice") {
        if (c->D.rho_avg.size() != c->D.meshFluidHost.size()) return 1;
        ok = w == "variablesTimeAvg" ? outputerVTK::variablesTimeAvgVTK(c->D, c->C, false)
                                     : outputerVTK::variablesTimeAvgLatticeVTK(c->D, c->C, false);
    } else {
        return 1;
    }
    return ok ? 0 : 2;
}

int lbmh_case_outlet(lbmh_case* c, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                     const float normal[3], float density)
{
    if (!c) return 1;
    c->M.meshingBoundaryConditionOutlet(cuboid(p1, p2, p3, id), {normal[0], normal[1], normal[2]}, density, false);
    return 0;
}

int lbmh_case_periodic(lbmh_case* c, const double p1[3], const double p2[3], const double p3[3], int32_t id,
                       const float normal[3], int32_t periodic_index)
{
    if (!c) return 1;
    c->M.meshingBoundaryConditionPeriodic(cuboid(p1, p2, p3, id), {normal[0], normal[1], normal[2]}, periodic_index,
                                          false);
    return 0;
}

int lbmh_case_compile(lbmh_case* c)
{
    if (!c) return 1;
    c->M.compileBoundaryArrayWall(false);
    c->M.compileBoundaryArrayInlets(false);
    c->M.compileBoundaryArrayPeriodic(false);
    c->M.compileBoundaryArrayOutlets(false);
    return 0;
}

int lbmh_case_convert_to_lattice(lbmh_case* c)
{
    if (!c) return 1;
    c->S.convertToLattice(false);
    return 0;
}

int lbmh_case_initialize(lbmh_case* c) { return c ? c->S.initializeSimulation(false) : LBM_E_ARG; }

int lbmh_case_run(lbmh_case* c)
{
    if (!c) return LBM_E_ARG;
    c->S.write_output = c->C.plot_every_it > 0;
    return c->S.runSimulation();
}

int lbmh_case_flags(lbmh_case* c, int32_t* flags)
{
    if (!c || !flags) return 1;
    std::memcpy(flags, c->D.meshFluidHost.data(), sizeof(int32_t) * c->D.meshFluidHost.size());
    return 0;
}

lbm_ctx* lbmh_case_lattice(lbmh_case* c) { return c ? c->D.ctx : nullptr; }

int lbmh_case_time_avg(lbmh_case* c, double* rho, double* ux, double* uy, double* uz, int64_t* count)
{
    if (!c) return 1;
    if (count) *count = c->D.time_avg_count;
    const size_t N = c->D.rho_avg.size();
    if (N == 0) return 1;
    const std::vector<double>* src[4] = {&c->D.rho_avg, &c->D.ux_avg, &c->D.uy_avg, &c->D.uz_avg};
    double* dst[4] = {rho, ux, uy, uz};
    for (int k = 0; k < 4; ++k)
        if (dst[k]) std::copy(src[k]->begin(), src[k]->end(), dst[k]);
    return 0;
}

int lbmh_case_errors(lbmh_case* c, double* out, int32_t* n)
{
    if (!c || !n) return 1;
    const int32_t k = (int32_t)c->S.err_history.size();
    if (out) {
        if (*n < k) return 1;
        for (int32_t i = 0; i < k; ++i) out[i] = c->S.err_history[i];
    }
    *n = k;
    return 0;
}

}  // extern "C"
