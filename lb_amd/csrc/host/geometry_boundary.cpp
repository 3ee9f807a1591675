// geometry_boundary.cpp -- flag grid and boundary lists from the 0/1 grid.
//
// Follows 3D/simulation/src/geometry/geometryMesherBoundary.h and
// geometryObjectCuboid.h: fluid cells count their solid 26-neighbours
// (flag = 1 + count); cells with more than two solid neighbours form the
// wall list; boundary cuboids overwrite the flags of the cells they contain
// with their (negative) id.
#include <algorithm>
#include <iostream>
#include <stdexcept>
#include <cstdio>
#include <cstring>
#include <cmath>

#include "lbm_host.hpp"

namespace lbmh {

geometryObjectCuboid::geometryObjectCuboid(d3 X1, d3 X2, d3 X3, int id_) : id(id_)
{
    x_min = std::min(X1.x, std::min(X2.x, X3.x));
    x_max = std::max(X1.x, std::max(X2.x, X3.x));
    y_min = std::min(X1.y, std::min(X2.y, X3.y));
    y_max = std::max(X1.y, std::max(X2.y, X3.y));
    z_min = std::min(X1.z, std::min(X2.z, X3.z));
    z_max = std::max(X1.z, std::max(X2.z, X3.z));
}

bool geometryObjectCuboid::isInside(d3 X) const
{
    return X.x >= x_min && X.x <= x_max && X.y >= y_min && X.y <= y_max && X.z >= z_min && X.z <= z_max;
}

// Cell centre in physical coordinates, double (geometryMesherBoundary.h:291-293).
d3 geometryMesherBoundary::cellCentre(int i, int j, int k) const
{
    return {((double)i + 0.5 - Constants.additional_factor) / Constants.resolution_factor + Constants.BBminx,
            ((double)j + 0.5 - Constants.additional_factor) / Constants.resolution_factor + Constants.BBminy,
            ((double)k + 0.5 - Constants.additional_factor) / Constants.resolution_factor + Constants.BBminz};
}

// meshingBoundaryConditionInletUniform (:47-108).  The query point is held in
// RealType there (:62-78), so it is rounded to float before the cuboid test.
bool geometryMesherBoundary::meshingBoundaryConditionInletUniform(const geometryObjectCuboid& cuboid,
                                                                  const Vector& normal_out, const Vector& velocity_in,
                                                                  bool verbose, int kind)
{
    if (velocity_in.x == velocity_in.y && velocity_in.x == velocity_in.z) return false;  // reference throws (:52-54)
    if (std::abs(normal_out.x + normal_out.y + normal_out.z) != 1)
        std::printf("Your normal is not tangent to one of the axes.\n");
    const int num = meshInletCells(cuboid, normal_out, velocity_in, kind);
    if (verbose) std::cout << "\nMeshed cuboid id " << cuboid.id << " as Inlet.\nBoundary vertexes number: " << num << "\n";
    return true;
}

// The parabolic inlets start from the uniform rule with velocity (0, 0, 0)
// (:124, :200), which the uniform rule's own check rejects (:52-54: all
// components equal) -- as written they throw.  Here the cells are meshed
// without that check, which is what the two functions go on to assume.
void geometryMesherBoundary::meshingBoundaryConditionInletParaboloidCircle(
    const geometryObjectCuboid& cuboid, const d3& inletCenter, RealType inletRadius, const Vector& normal_out,
    RealType mean_velocity_in, bool verbose, int kind)
{
    if (std::abs(normal_out.x + normal_out.y + normal_out.z) != 1)
        std::printf("Your normal is not tangent to one of the axes.\n");
    meshInletCells(cuboid, normal_out, Vector{}, kind);
    int num = 0;
    for (boundaryConditionInlet& BC : boundary_vector_inlet) {
        const d3 c = cellCentre(BC.vertex.x, BC.vertex.y, BC.vertex.z);
        const RealType lookx = (RealType)c.x, looky = (RealType)c.y, lookz = (RealType)c.z;
        if (!cuboid.isInside({lookx, looky, lookz})) continue;
        // (:137-141): double differences and root, held in RealType
        const RealType distance = (RealType)std::sqrt(std::pow(lookx - inletCenter.x, 2) +
                                                      std::pow(looky - inletCenter.y, 2) +
                                                      std::pow(lookz - inletCenter.z, 2));
        if (distance <= inletRadius) {
            // (:147): 3 u_mean (1 - (r/R)^2), the power in double
            RealType v = (RealType)(3 * mean_velocity_in * (1.0 - std::pow(distance / inletRadius, 2)));
            if (v < 0) v = 0;
            BC.velocity.x = -v * normal_out.x;
            BC.velocity.y = -v * normal_out.y;
            BC.velocity.z = -v * normal_out.z;
            ++num;
        }
    }
    if (verbose) std::cout << "\nMeshed cuboid id " << cuboid.id << " as Inlet.\nBoundary vertexes number: " << num << "\n";
}

void geometryMesherBoundary::meshingBoundaryConditionInletParaboloidRectangle(
    const geometryObjectCuboid& cuboid, const d3& inletCenter, RealType& inletDimX, RealType& inletDimY,
    RealType& inletDimZ, const Vector& normal_out, RealType mean_velocity_in, bool verbose, int kind)
{
    if (std::abs(normal_out.x + normal_out.y + normal_out.z) != 1)
        std::printf("Your normal is not tangent to one of the axes.\n");
    int num = 0;
    meshInletCells(cuboid, normal_out, Vector{}, kind);
    for (boundaryConditionInlet& BC : boundary_vector_inlet) {
        const d3 c = cellCentre(BC.vertex.x, BC.vertex.y, BC.vertex.z);
        const RealType lookx = (RealType)c.x, looky = (RealType)c.y, lookz = (RealType)c.z;
        if (!cuboid.isInside({lookx, looky, lookz})) continue;
        // distances from the corner with the smallest coordinates (:214-216)
        const RealType distance_x = (RealType)(lookx - (inletCenter.x - inletDimX / 2));
        const RealType distance_y = (RealType)(looky - (inletCenter.y - inletDimY / 2));
        const RealType distance_z = (RealType)(lookz - (inletCenter.z - inletDimZ / 2));
        // a product of 1-D parabolas across the inlet (:225-252); the extent
        // along the normal becomes 1
        RealType Part_x, Part_y, Part_z;
        if (normal_out.x == 0) {
            Part_x = distance_x * (inletDimX - distance_x);
        } else {
            Part_x = 1.f;
            inletDimX = 1.f;
        }
        if (normal_out.y == 0) {
            Part_y = distance_y * (inletDimY - distance_y);
        } else {
            Part_y = 1.f;
            inletDimY = 1.f;
        }
        if (normal_out.z == 0) {
            Part_z = distance_z * (inletDimZ - distance_z);
        } else {
            Part_z = 1.f;
            inletDimZ = 1.f;
        }
        const RealType inletConstant = inletDimX * inletDimY * inletDimZ * inletDimX * inletDimY * inletDimZ / 16;
        RealType v = 2 * mean_velocity_in * Part_x * Part_y * Part_z / inletConstant;
        if (v < 0) v = 0;
        BC.velocity.x = -v * normal_out.x;
        BC.velocity.y = -v * normal_out.y;
        BC.velocity.z = -v * normal_out.z;
        ++num;
    }
    if (verbose) std::cout << "\nMeshed cuboid id " << cuboid.id << " as Inlet.\nBoundary vertexes number: " << num << "\n";
}

// Cells of the cuboid with a fluid flag become inlet nodes (:58-100).
int geometryMesherBoundary::meshInletCells(const geometryObjectCuboid& cuboid, const Vector& normal_out,
                                           const Vector& velocity_in, int kind)
{
    int num = 0;
    for (int k = 0; k < Constants.dimZ_int; k++)
        for (int j = 0; j < Constants.dimY_int; j++)
            for (int i = 0; i < Constants.dimX_int; i++) {
                const d3 c = cellCentre(i, j, k);
                const d3 p = {(double)(RealType)c.x, (double)(RealType)c.y, (double)(RealType)c.z};
                if (cuboid.isInside(p) && Data.mesh(i, j, k) > 0) {
                    boundaryConditionInlet b;
                    b.normal = normal_out;
                    b.velocity = velocity_in;
                    b.vertex = {i, j, k};
                    b.regular = !(Data.mesh(i, j, k) > 10);
                    b.kind = kind;
                    Data.mesh(i, j, k) = cuboid.id;
                    boundary_vector_inlet.push_back(b);
                    ++num;
                }
            }
    return num;
}

// meshingBoundaryConditionOutlet (:276-325)
void geometryMesherBoundary::meshingBoundaryConditionOutlet(const geometryObjectCuboid& cuboid,
                                                            const Vector& normal_out, RealType density, bool verbose)
{
    int num = 0;
    for (int k = 0; k < Constants.dimZ_int; k++)
        for (int j = 0; j < Constants.dimY_int; j++)
            for (int i = 0; i < Constants.dimX_int; i++) {
                if (cuboid.isInside(cellCentre(i, j, k)) && Data.mesh(i, j, k) != 0) {
                    boundaryConditionOutlet b;
                    b.normal = normal_out;
                    b.density = density;
                    b.vertex = {i, j, k};
                    b.regular = Data.mesh(i, j, k) == 10;
                    Data.mesh(i, j, k) = cuboid.id;
                    boundary_vector_outlet.push_back(b);
                    ++num;
                }
            }
    if (verbose) std::cout << "\nMeshed cuboid id " << cuboid.id << " as outlet.\nBoundary vertexes number: " << num << "\n";
}

// meshingBoundaryConditionPeriodic (:379-437)
void geometryMesherBoundary::meshingBoundaryConditionPeriodic(const geometryObjectCuboid& cuboid,
                                                              const Vector& normal_out, int periodicIndex,
                                                              bool verbose)
{
    int num = 0;
    for (int k = 0; k < Constants.dimZ_int; k++)
        for (int j = 0; j < Constants.dimY_int; j++)
            for (int i = 0; i < Constants.dimX_int; i++) {
                if (cuboid.isInside(cellCentre(i, j, k)) && Data.mesh(i, j, k) != 0) {
                    boundaryConditionPeriodic b;
                    b.vertex = {i, j, k};
                    b.normal = normal_out;
                    b.periodicIndex = periodicIndex;
                    b.regular = Data.mesh(i, j, k) == 10;
                    Data.mesh(i, j, k) = cuboid.id;
                    boundary_vector_periodic.push_back(b);
                    ++num;
                }
            }
    if (verbose) std::cout << "\nMeshed cuboid id " << cuboid.id << " as Periodic.\nBoundary vertexes number: " << num << "\n";
}

// meshingBoundaryWall (:439-484): solid counts are taken from the 0/1 grid
// before any cell of this pass is incremented (incremented cells stay > 0, so
// the in-place loop of the reference sees the same solids).
void geometryMesherBoundary::meshingBoundaryWall(bool verbose)
{
    const int nx = Constants.dimX_int, ny = Constants.dimY_int, nz = Constants.dimZ_int;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                if (Data.mesh(i, j, k) <= 0) continue;
                for (int l : {-1, 0, 1})
                    for (int m : {-1, 0, 1})
                        for (int n : {-1, 0, 1}) {
                            const int a = i + l, b = j + m, c = k + n;
                            if (a >= 0 && b >= 0 && c >= 0 && a < nx && b < ny && c < nz && Data.mesh(a, b, c) == 0)
                                Data.mesh(i, j, k) += 1;
                        }
                if (Data.mesh(i, j, k) > 2) boundary_vector_wall.push_back({{i, j, k}, !(Data.mesh(i, j, k) > 10)});
            }
    if (verbose) std::cout << "\nMeshed Wall.\nWall vertexes number: " << boundary_vector_wall.size() << "\n";
}

void geometryMesherBoundary::compileBoundaryArrayInlets(bool verbose)
{
    Constants.inlet_num = (int)boundary_vector_inlet.size();
    Data.meshBoundaryInletHost = boundary_vector_inlet;
    if (verbose) std::cout << "\nCreated boundary Array Inlets.\nBoundary vertexes number: " << Constants.inlet_num << "\n";
}

void geometryMesherBoundary::compileBoundaryArrayOutlets(bool verbose)
{
    Constants.outlet_num = (int)boundary_vector_outlet.size();
    Data.meshBoundaryOutletHost = boundary_vector_outlet;
    if (verbose) std::cout << "\nCreated boundary Array Outlet.\nBoundary vertexes number: " << Constants.outlet_num << "\n";
}

// compileBoundaryArrayWall (:521-561): entries whose flag became negative are erased.
void geometryMesherBoundary::compileBoundaryArrayWall(bool verbose)
{
    const size_t before = boundary_vector_wall.size();
    boundary_vector_wall.erase(std::remove_if(boundary_vector_wall.begin(), boundary_vector_wall.end(),
                                              [&](const boundaryConditionWall& b) {
                                                  return Data.mesh(b.vertex.x, b.vertex.y, b.vertex.z) < 0;
                                              }),
                               boundary_vector_wall.end());
    Constants.wall_num = (int)boundary_vector_wall.size();
    Data.meshBoundaryWallHost = boundary_vector_wall;
    if (verbose)
        std::cout << "\nCreated boundary Array Wall.\nErased " << before - boundary_vector_wall.size()
                  << " elements.\nWall vertexes number: " << Constants.wall_num << "\n";
}

void geometryMesherBoundary::compileBoundaryArrayPeriodic(bool verbose)
{
    Constants.periodic_num = (int)boundary_vector_periodic.size();
    Data.meshBoundaryPeriodicHost = boundary_vector_periodic;
    if (verbose) std::cout << "\nCreated boundary Array Periodic.\nPeriodic vertexes number: " << Constants.periodic_num << "\n";
}

void geometryMesherBoundary::arrayTransfer(bool verbose, std::vector<lbm_bc_node>& bc) const
{
    bc.clear();
    auto node = [](int x, int y, int z, int kind) {
        lbm_bc_node n;
        std::memset(&n, 0, sizeof(n));
        n.x = x;
        n.y = y;
        n.z = z;
        n.kind = kind;
        n.regular = 1;
        return n;
    };
    for (const auto& w : Data.meshBoundaryWallHost) {
        lbm_bc_node n = node(w.vertex.x, w.vertex.y, w.vertex.z, LBM_BC_WALL);
        n.regular = w.regular;
        bc.push_back(n);
    }
    for (const auto& b : Data.meshBoundaryInletHost) {
        lbm_bc_node n = node(b.vertex.x, b.vertex.y, b.vertex.z, b.kind);
        n.nx = b.normal.x; n.ny = b.normal.y; n.nz = b.normal.z;
        n.ux = b.velocity.x; n.uy = b.velocity.y; n.uz = b.velocity.z;
        n.regular = b.regular;
        bc.push_back(n);
    }
    for (const auto& b : Data.meshBoundaryOutletHost) {
        lbm_bc_node n = node(b.vertex.x, b.vertex.y, b.vertex.z, LBM_BC_OUTLET_INTERPOLATED);
        n.nx = b.normal.x; n.ny = b.normal.y; n.nz = b.normal.z;
        n.rho = b.density;
        n.regular = b.regular;
        bc.push_back(n);
    }
    for (const auto& b : Data.meshBoundaryPeriodicHost) {
        lbm_bc_node n = node(b.vertex.x, b.vertex.y, b.vertex.z, LBM_BC_PERIODIC_GHOST);
        n.nx = b.normal.x; n.ny = b.normal.y; n.nz = b.normal.z;
        n.per_idx = b.periodicIndex;
        n.regular = b.regular;
        bc.push_back(n);
    }
    if (verbose) std::cout << "\n Transferred " << bc.size() << " boundary nodes.\n";
}

}  // namespace lbmh
