// outputers.cpp -- flag-grid text files and legacy-VTK output.
//
// outputerMesh follows 3D/simulation/src/postprocesors/outputerMesh.h:19-211
// (header written with the default 6-significant-digit stream precision,
// values x fastest, one line per row, a blank line per z slice); the files
// are byte-compatible with the reference's.  outputerVTK follows
// outputerVTK.h (STRUCTURED_POINTS, x fastest); the pressure field is
// rho*cs^2, since the reference's p array is never computed in 3-D.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "lbm_host.hpp"

namespace lbmh {

bool outputerMesh::MeshMatrixOut(const LBMData& Data, const LBMConstants& Constants, const std::string& path)
{
    std::ofstream out(path);
    if (!out.is_open()) {
        std::cerr << "Failed to open the file for writing." << std::endl;
        return false;
    }
    out << "Dimensions: " << Constants.dimX_int << " " << Constants.dimY_int << " " << Constants.dimZ_int << " "
        << std::endl;
    out << "Resolution Factor: " << Constants.resolution_factor << std::endl;
    out << "Additional Factor: " << Constants.additional_factor << std::endl;
    out << "BBmaxx: " << Constants.BBmaxx << std::endl;
    out << "BBmaxy: " << Constants.BBmaxy << std::endl;
    out << "BBmaxz: " << Constants.BBmaxz << std::endl;
    out << "BBminx: " << Constants.BBminx << std::endl;
    out << "BBminy: " << Constants.BBminy << std::endl;
    out << "BBminz: " << Constants.BBminz << std::endl;
    std::string row;
    for (int z = 0; z < Constants.dimZ_int; ++z) {
        for (int y = 0; y < Constants.dimY_int; ++y) {
            row.clear();
            for (int x = 0; x < Constants.dimX_int; ++x) {
                row += std::to_string(Data.mesh(x, y, z));
                row += ' ';
            }
            out << row << '\n';
        }
        out << '\n';
    }
    return (bool)out;
}

bool outputerMesh::MeshMatrixIn(LBMData& Data, LBMConstants& Constants, const std::string& path, bool verbose)
{
    std::ifstream in(path);
    if (!in.is_open()) {
        std::cerr << "Failed to open the file for reading." << std::endl;
        return false;
    }
    std::string line;
    int got = 0;
    while (got < 9 && std::getline(in, line)) {
        float v;
        int a, b, c;
        if (std::sscanf(line.c_str(), "Dimensions: %d %d %d", &a, &b, &c) == 3) {
            Constants.dimX_int = a;
            Constants.dimY_int = b;
            Constants.dimZ_int = c;
            ++got;
        } else if (std::sscanf(line.c_str(), "Resolution Factor: %f", &v) == 1) { Constants.resolution_factor = v; ++got; }
        else if (std::sscanf(line.c_str(), "Additional Factor: %f", &v) == 1) { Constants.additional_factor = v; ++got; }
        else if (std::sscanf(line.c_str(), "BBmaxx: %f", &v) == 1) { Constants.BBmaxx = v; ++got; }
        else if (std::sscanf(line.c_str(), "BBmaxy: %f", &v) == 1) { Constants.BBmaxy = v; ++got; }
        else if (std::sscanf(line.c_str(), "BBmaxz: %f", &v) == 1) { Constants.BBmaxz = v; ++got; }
        else if (std::sscanf(line.c_str(), "BBminx: %f", &v) == 1) { Constants.BBminx = v; ++got; }
        else if (std::sscanf(line.c_str(), "BBminy: %f", &v) == 1) { Constants.BBminy = v; ++got; }
        else if (std::sscanf(line.c_str(), "BBminz: %f", &v) == 1) { Constants.BBminz = v; ++got; }
    }
    if (got < 9 || Constants.dimX_int <= 0 || Constants.dimY_int <= 0 || Constants.dimZ_int <= 0) return false;
    Data.setSizes(Constants.dimX_int, Constants.dimY_int, Constants.dimZ_int);
    for (int k = 0; k < Constants.dimZ_int; k++)
        for (int j = 0; j < Constants.dimY_int; j++)
            for (int i = 0; i < Constants.dimX_int; i++) {
                int v;
                if (!(in >> v)) return false;  // truncated file
                Data.mesh(i, j, k) = v;
            }
    if (verbose) {
        std::cout << "Resolution factor: " << Constants.resolution_factor << std::endl;
        std::cout << "Dimension X: " << Constants.dimX_int << std::endl;
        std::cout << "Dimension Y: " << Constants.dimY_int << std::endl;
        std::cout << "Dimension Z: " << Constants.dimZ_int << std::endl;
    }
    return true;
}

static void vtk_header(std::ostream& out, const LBMConstants& C, const std::string& title)
{
    out << "# vtk DataFile Version 2.0\n";
    out << "LBE " << title << "\n";
    out << "ASCII\n";
    out << "DATASET STRUCTURED_POINTS\n";
    out << "DIMENSIONS " << C.dimX_int << " " << C.dimY_int << " " << C.dimZ_int << "\n";
    out << "ASPECT_RATIO " << 1.f / C.resolution_factor << " " << 1.f / C.resolution_factor << " "
        << 1.f / C.resolution_factor << "\n";
    out << "ORIGIN " << C.BBminx - (C.additional_factor - 0.5) / C.resolution_factor << " "
        << C.BBminy - (C.additional_factor - 0.5) / C.resolution_factor << " "
        << C.BBminz - (C.additional_factor - 0.5) / C.resolution_factor << "\n";
    out << "POINT_DATA " << (long)C.dimX_int * C.dimY_int * C.dimZ_int << "\n";
}

bool outputerVTK::MeshVTK(const LBMData& Data, const LBMConstants& Constants, const std::string& filename)
{
    std::ofstream out(Constants.results_dir + filename + ".vtk");
    if (!out.is_open()) return false;
    vtk_header(out, Constants, filename);
    out << "SCALARS object double 1\nLOOKUP_TABLE default\n";
    for (int32_t v : Data.meshFluidHost) out << v << "\n";
    return (bool)out;
}

static bool fetch_fields(LBMData& Data)
{
    const size_t N = Data.meshFluidHost.size();
    Data.rho_out.resize(N);
    Data.ux_out.resize(N);
    Data.uy_out.resize(N);
    Data.uz_out.resize(N);
    return Data.ctx && lbm_get_macroscopic(Data.ctx, Data.rho_out.data(), Data.ux_out.data(), Data.uy_out.data(),
                                           Data.uz_out.data()) == LBM_OK;
}

static bool write_vars(LBMData& Data, const LBMConstants& C, const std::string& path, double crho, double cp,
                       double cu, const char* pname, bool fetch = true)
{
    if (fetch && !fetch_fields(Data)) return false;
    std::ofstream out(path);
    if (!out.is_open()) return false;
    vtk_header(out, C, "variables");
    out << "SCALARS rho double 1\nLOOKUP_TABLE default\n";
    for (double r : Data.rho_out) out << r * crho << "\n";
    out << "SCALARS " << pname << " double 1\nLOOKUP_TABLE default\n";
    for (double r : Data.rho_out) out << r * C.cs2 * cp << "\n";
    out << "VECTORS U double\n";
    for (size_t i = 0; i < Data.rho_out.size(); ++i)
        out << Data.ux_out[i] * cu << " " << Data.uy_out[i] * cu << " " << Data.uz_out[i] * cu << "\n";
    return (bool)out;
}

bool outputerVTK::variablesVTK(LBMData& Data, const LBMConstants& C, int step, bool verbose, bool fetch)
{
    if (verbose) std::cout << "\nDevice -> Host output variables\n";
    return write_vars(Data, C, C.results_dir + "variables." + std::to_string(step) + ".vtk", C.Crho, C.Cpressure, C.Cu,
                      "p", fetch);
}

bool outputerVTK::variablesLatticeVTK(LBMData& Data, const LBMConstants& C, int step, bool verbose)
{
    if (verbose) std::cout << "\nDevice -> Host output lattice units\n";
    return write_vars(Data, C, C.results_dir + "variablesLattice" + std::to_string(step) + ".vtk", 1.0, 1.0, 1.0,
                      "pressure");
}

bool outputerVTK::variablesTimeAvgVTK(LBMData& Data, const LBMConstants& C, bool verbose)
{
    if (verbose) std::cout << "\nWriting time averaged moments\n";
    const size_t N = Data.meshFluidHost.size();
    if (Data.rho_avg.size() != N) return false;
    std::ofstream out(C.results_dir + "variablesTimeAvg.vtk");
    if (!out.is_open()) return false;
    vtk_header(out, C, "variables");
    out << "SCALARS rho double 1\nLOOKUP_TABLE default\n";
    for (double r : Data.rho_avg) out << r * C.Crho << "\n";
    out << "SCALARS pressure double 1\nLOOKUP_TABLE default\n";
    for (double r : Data.rho_avg) out << r * C.cs2 * C.Cpressure << "\n";
    out << "VECTORS U double\n";
    for (size_t i = 0; i < N; ++i)
        out << Data.ux_avg[i] * C.Cu << " " << Data.uy_avg[i] * C.Cu << " " << Data.uz_avg[i] * C.Cu << "\n";
    return (bool)out;
}

// variablesTimeAvgLatticeVTK (outputerVTK.h:449-576): the time averages in
// lattice units, pressure rho cs^2.
bool outputerVTK::variablesTimeAvgLatticeVTK(LBMData& Data, const LBMConstants& C, bool verbose)
{
    if (verbose) std::cout << "\nWriting time averaged moments\n\nDevice -> Host output Lattice units\n";
    const size_t N = Data.meshFluidHost.size();
    if (Data.rho_avg.size() != N) return false;
    std::ofstream out(C.results_dir + "variablesTimeAvg.vtk");
    if (!out.is_open()) return false;
    vtk_header(out, C, "variables");
    out << "SCALARS rho double 1\nLOOKUP_TABLE default\n";
    for (double r : Data.rho_avg) out << r << "\n";
    out << "SCALARS pressure double 1\nLOOKUP_TABLE default\n";
    for (double r : Data.rho_avg) out << r * C.cs2 << "\n";
    out << "VECTORS U double\n";
    for (size_t i = 0; i < N; ++i) out << Data.ux_avg[i] << " " << Data.uy_avg[i] << " " << Data.uz_avg[i] << "\n";
    return (bool)out;
}

// distributionFunctionVTK (outputerVTK.h:660-740): one population plane of
// the current state (the scalar's name keeps the reference's spelling).
bool outputerVTK::distributionFunctionVTK(LBMData& Data, const LBMConstants& C, int velocity, bool verbose)
{
    if (verbose) std::cout << "\nDevice -> Host output distribution function\n";
    std::vector<double> f(Data.meshFluidHost.size());
    if (!Data.ctx || lbm_get_population(Data.ctx, velocity, f.data()) != LBM_OK) return false;
    std::ofstream out(C.results_dir + "distributionFunction" + std::to_string(velocity) + ".vtk");
    if (!out.is_open()) return false;
    vtk_header(out, C, "variables");
    out << "SCALARS DFunciton_" << velocity << " double 1\nLOOKUP_TABLE default\n";
    for (double v : f) out << v << "\n";
    return (bool)out;
}

// tauVTK (outputerVTK.h:742-822): 1 / omega of the next step's rate, in
// RealType (inf on solid cells, whose omega is 0).
bool outputerVTK::tauVTK(LBMData& Data, const LBMConstants& C, bool verbose)
{
    if (verbose) std::cout << "\nDevice -> Host output Tau.\n";
    Data.omega_out.resize(Data.meshFluidHost.size());
    if (!Data.ctx || lbm_get_omega(Data.ctx, Data.omega_out.data()) != LBM_OK) return false;
    std::ofstream out(C.results_dir + "tau.vtk");
    if (!out.is_open()) return false;
    vtk_header(out, C, "variables");
    out << "SCALARS Tau double 1\nLOOKUP_TABLE default\n";
    for (double w : Data.omega_out) out << 1.f / (RealType)w << "\n";
    return (bool)out;
}

bool outputerVTK::omegaVTK(LBMData& Data, const LBMConstants& C, bool verbose)
{
    if (verbose) std::cout << "\nDevice -> Host output omega\n";
    Data.omega_out.resize(Data.meshFluidHost.size());
    if (!Data.ctx || lbm_get_omega(Data.ctx, Data.omega_out.data()) != LBM_OK) return false;
    std::ofstream out(C.results_dir + "omega.vtk");
    if (!out.is_open()) return false;
    vtk_header(out, C, "variables");
    out << "SCALARS Omega double 1\nLOOKUP_TABLE default\n";
    for (double w : Data.omega_out) out << w << "\n";
    return (bool)out;
}

// probeWriteOut (outputerVTK.h:824-850): appended CSV, header on creation.
bool outputerVTK::probeWriteOut(const LBMConstants& C, RealType velX, RealType velY, RealType velZ, RealType rho)
{
    const std::string filename = C.results_dir + "probeData.csv";
    std::ifstream infile(filename);
    const bool exists = infile.good();
    infile.close();
    std::ofstream file(filename, std::ios_base::app);
    if (!file.is_open()) return false;
    if (!exists) {
        file << "Sampling frequency is: " << C.Ct * C.probe_every_it << "\n";
        file << "ux,uy,uz,rho,pressure\n";
    }
    file << static_cast<float>(velX * C.Cu) << "," << static_cast<float>(velY * C.Cu) << ","
         << static_cast<float>(velZ * C.Cu) << "," << static_cast<float>(rho * C.Crho) << ","
         << static_cast<float>(rho * C.cs2 * C.Cpressure) << "\n";
    return true;
}

}  // namespace lbmh
