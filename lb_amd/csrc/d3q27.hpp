// d3q27.hpp -- D3Q27 velocity set in the reference's slot order.
//
// The slot order is the population layout of the reference
// (3D/simulation/src/solvers/models/D3Q27/D3Q27.h:81-111) and therefore of
// every array this core exchanges with a caller, so it is kept.  Everything
// here is constexpr: loops over q unroll to immediates in the kernels.
#pragma once

#if defined(__HIPCC__)
#define LBM_HD __host__ __device__ __forceinline__
#else
#define LBM_HD inline
#endif

namespace lbm {

constexpr int Q = 27;

constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1};
constexpr int CY[Q] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1};
constexpr int CZ[Q] = {0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1};

// True opposite direction (c[OPP[q]] == -c[q]).
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};

// Number of non-zero components: weight class 0 (rest), 1 (axis), 2 (edge), 3 (corner).
constexpr int WCLASS[Q] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3};

// Slot of direction q in a 3x3x3 array indexed [cx+1][cy+1][cz+1].
constexpr int CELL3(int q) { return 9 * (CX[q] + 1) + 3 * (CY[q] + 1) + (CZ[q] + 1); }

// Reference bounce-back partner: equals OPP except slot 20, which the
// reference table maps to 21 (D3Q27.h:120-124).  Selected at run time.
LBM_HD int crev_of(int q, bool crev20_quirk) { return (crev20_quirk && q == 20) ? 21 : OPP[q]; }

}  // namespace lbm
