// voxel.hip -- the OFF voxelizer's inside test on the GPU.
//
// Same predicate as the host voxelizer (lb_amd/csrc/host/geometry_off.cpp,
// after geometryHandlerOFF::polyhedronInside, 3D/preprocess/src/geometry/
// geometryHandlerOFF.h:135-182, and geometryMesherOFF::meshingPolyhedron,
// geometryMesherOFF.h:90-103): one thread per cell casts the segment from the
// cell's query point to a point outside, walks the host's bounding-volume
// hierarchy in the host's order, and counts distinct crossing points.  All
// arithmetic is FP64 with contraction off, so every sign, crossing point and
// distance equals the host's.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/lbm_core.h"

#define LBM_FP_STRICT _Pragma("clang fp contract(off)")

namespace {

constexpr int VX_STACK = 64;   // hierarchy depth bound (median splits of 20 k triangles: ~13)
constexpr int VX_HITS = 64;    // crossing points held per segment

struct VoxArgs {
    const double* V;          // [nv][3]
    const int32_t* T;         // [nt][3]
    const lbm_bvh_node* nodes;
    const int32_t* order;
    int nx, ny, nz;
    double add, res, bx, by, bz;
    double qx, qy, qz;        // the outside point
    int32_t* grid;
    int32_t* overflow;        // set when a bound is exceeded
};

__device__ __forceinline__ bool seg_box(const double* p, const double* d, const double* lo, const double* hi)
{
    LBM_FP_STRICT
    double t0 = 0.0, t1 = 1.0;
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0.0) {
            if (p[a] < lo[a] || p[a] > hi[a]) return false;
            continue;
        }
        double ta = (lo[a] - p[a]) / d[a], tb = (hi[a] - p[a]) / d[a];
        if (ta > tb) {
            const double t = ta;
            ta = tb;
            tb = t;
        }
        t0 = (t0 < ta) ? ta : t0;   // std::max(t0, ta)
        t1 = (tb < t1) ? tb : t1;   // std::min(t1, tb)
        if (t0 > t1) return false;
    }
    return true;
}

// signed volume det(b - a, c - a, d - a)
__device__ __forceinline__ double vol(const double* a, const double* b, const double* c, const double* d)
{
    LBM_FP_STRICT
    const double b0 = b[0] - a[0], b1 = b[1] - a[1], b2 = b[2] - a[2];
    const double c0 = c[0] - a[0], c1 = c[1] - a[1], c2 = c[2] - a[2];
    const double d0 = d[0] - a[0], d1 = d[1] - a[1], d2 = d[2] - a[2];
    return b0 * (c1 * d2 - c2 * d1) - b1 * (c0 * d2 - c2 * d0) + b2 * (c0 * d1 - c1 * d0);
}

__device__ __forceinline__ int sgn(double v) { return (v > 0) - (v < 0); }

__device__ __forceinline__ int edge_side(const VoxArgs& a, const double* p, const double* q, int iu, int iv)
{
    if (iu < iv) return sgn(vol(p, q, a.V + 3 * iu, a.V + 3 * iv));
    return -sgn(vol(p, q, a.V + 3 * iv, a.V + 3 * iu));
}

__device__ __forceinline__ bool seg_tri(const VoxArgs& a, const double* p, const double* q, int t, double* x)
{
    LBM_FP_STRICT
    const int i0 = a.T[3 * t], i1 = a.T[3 * t + 1], i2 = a.T[3 * t + 2];
    const double *A = a.V + 3 * i0, *B = a.V + 3 * i1, *C = a.V + 3 * i2;
    const double vp = vol(A, B, C, p), vq = vol(A, B, C, q);
    const int sp = sgn(vp), sq = sgn(vq);
    if (sp == sq) return false;
    const int s1 = edge_side(a, p, q, i0, i1), s2 = edge_side(a, p, q, i1, i2), s3 = edge_side(a, p, q, i2, i0);
    const bool pos = s1 >= 0 && s2 >= 0 && s3 >= 0, neg = s1 <= 0 && s2 <= 0 && s3 <= 0;
    if (!pos && !neg) return false;
    const double t_ = vp / (vp - vq);
    for (int k = 0; k < 3; ++k) x[k] = p[k] + t_ * (q[k] - p[k]);
    return true;
}

__global__ __launch_bounds__(256) void k_voxelize(VoxArgs a)
{
    LBM_FP_STRICT
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    const size_t cell = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (cell >= n) return;
    const int i = (int)(cell % (size_t)a.nx);
    const int j = (int)((cell / (size_t)a.nx) % (size_t)a.ny);
    const int k = (int)(cell / ((size_t)a.nx * a.ny));
    // query point in double, held in a float (RealType), read back as double
    const double p[3] = {(double)(float)(((double)i + 0.5 - a.add) / a.res + a.bx),
                         (double)(float)(((double)j + 0.5 - a.add) / a.res + a.by),
                         (double)(float)(((double)k + 0.5 - a.add) / a.res + a.bz)};
    const double q[3] = {a.qx, a.qy, a.qz};
    if (p[0] == q[0] && p[1] == q[1] && p[2] == q[2]) {
        a.grid[cell] = 0;
        return;
    }
    const double d[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
    double pts[VX_HITS][3];
    int np = 0;
    bool over = false;
    int stack[VX_STACK];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const lbm_bvh_node nd = a.nodes[stack[--sp]];
        if (!seg_box(p, d, nd.lo, nd.hi)) continue;
        if (nd.left < 0) {
            for (int t = nd.first; t < nd.first + nd.count; ++t) {
                double x[3];
                if (seg_tri(a, p, q, a.order[t], x)) {
                    if (np < VX_HITS) {
                        pts[np][0] = x[0];
                        pts[np][1] = x[1];
                        pts[np][2] = x[2];
                        ++np;
                    } else {
                        over = true;
                    }
                }
            }
        } else if (sp + 2 <= VX_STACK) {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        } else {
            over = true;
        }
    }
    int nd_ = 0;
    for (int u = 0; u < np; ++u) {
        bool distinct = true;
        for (int v = 0; v < u; ++v) {
            const double dx = pts[u][0] - pts[v][0], dy = pts[u][1] - pts[v][1], dz = pts[u][2] - pts[v][2];
            distinct = (dx * dx + dy * dy + dz * dz) > 1e-10;
            if (!distinct) break;
        }
        if (distinct) ++nd_;
    }
    a.grid[cell] = nd_ % 2;
    if (over) atomicOr(a.overflow, 1);
}

}  // namespace

extern "C" lbm_status lbm_voxelize(const double* vertices, int32_t n_vertices, const int32_t* triangles,
                                   int32_t n_triangles, const lbm_bvh_node* nodes, int32_t n_nodes,
                                   const int32_t* order, const int32_t dims[3], double additional, double resolution,
                                   const double bbmin[3], const double outside[3], int32_t* grid)
{
    if (!dims || !bbmin || !outside || !grid || n_vertices < 0 || n_triangles < 0 || n_nodes < 0)
        return LBM_E_ARG;
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return LBM_E_ARG;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    if (n_triangles == 0 || n_nodes == 0) {   // an empty polyhedron: nothing is inside
        for (size_t c = 0; c < n; ++c) grid[c] = 0;
        return LBM_OK;
    }
    if (!vertices || !triangles || !nodes || !order) return LBM_E_ARG;
    // the traversal indexes these arrays with values the host built: check them
    for (int32_t t = 0; t < 3 * n_triangles; ++t)
        if (triangles[t] < 0 || triangles[t] >= n_vertices) return LBM_E_ARG;
    // a tree whose children follow their parent (as a depth-first build lays
    // it out): every walk ends, and visits each node at most once
    std::vector<uint8_t> has_parent((size_t)n_nodes, 0);
    for (int32_t k = 0; k < n_nodes; ++k) {
        const lbm_bvh_node& b = nodes[k];
        if (b.left >= 0) {
            if (b.left <= k || b.left >= n_nodes || b.right <= k || b.right >= n_nodes || b.left == b.right)
                return LBM_E_ARG;
            if (has_parent[b.left] || has_parent[b.right]) return LBM_E_ARG;
            has_parent[b.left] = has_parent[b.right] = 1;
        } else if (b.first < 0 || b.count < 0 || (int64_t)b.first + b.count > n_triangles) {
            return LBM_E_ARG;
        }
    }
    for (int32_t t = 0; t < n_triangles; ++t)
        if (order[t] < 0 || order[t] >= n_triangles) return LBM_E_ARG;

    void *dV = nullptr, *dT = nullptr, *dN = nullptr, *dO = nullptr, *dG = nullptr, *dF = nullptr;
    lbm_status st = LBM_OK;
    auto alloc = [&](void** p, size_t bytes) {
        if (st == LBM_OK && hipMalloc(p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            st = LBM_E_OOM;
        }
    };
    alloc(&dV, sizeof(double) * 3 * (size_t)n_vertices);
    alloc(&dT, sizeof(int32_t) * 3 * (size_t)n_triangles);
    alloc(&dN, sizeof(lbm_bvh_node) * (size_t)n_nodes);
    alloc(&dO, sizeof(int32_t) * (size_t)n_triangles);
    alloc(&dG, sizeof(int32_t) * n);
    alloc(&dF, sizeof(int32_t));
    int32_t over = 0;
    if (st == LBM_OK) {
        bool ok = hipMemcpy(dV, vertices, sizeof(double) * 3 * (size_t)n_vertices, hipMemcpyHostToDevice) == hipSuccess;
        ok = ok && hipMemcpy(dT, triangles, sizeof(int32_t) * 3 * (size_t)n_triangles, hipMemcpyHostToDevice) ==
                       hipSuccess;
        ok = ok && hipMemcpy(dN, nodes, sizeof(lbm_bvh_node) * (size_t)n_nodes, hipMemcpyHostToDevice) == hipSuccess;
        ok = ok && hipMemcpy(dO, order, sizeof(int32_t) * (size_t)n_triangles, hipMemcpyHostToDevice) == hipSuccess;
        ok = ok && hipMemcpy(dF, &over, sizeof(int32_t), hipMemcpyHostToDevice) == hipSuccess;
        if (ok) {
            VoxArgs a;
            a.V = (const double*)dV;
            a.T = (const int32_t*)dT;
            a.nodes = (const lbm_bvh_node*)dN;
            a.order = (const int32_t*)dO;
            a.nx = dims[0];
            a.ny = dims[1];
            a.nz = dims[2];
            a.add = additional;
            a.res = resolution;
            a.bx = bbmin[0];
            a.by = bbmin[1];
            a.bz = bbmin[2];
            a.qx = outside[0];
            a.qy = outside[1];
            a.qz = outside[2];
            a.grid = (int32_t*)dG;
            a.overflow = (int32_t*)dF;
            k_voxelize<<<(unsigned)((n + 255) / 256), 256>>>(a);
            ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
            ok = ok && hipMemcpy(&over, dF, sizeof(int32_t), hipMemcpyDeviceToHost) == hipSuccess;
            ok = ok && hipMemcpy(grid, dG, sizeof(int32_t) * n, hipMemcpyDeviceToHost) == hipSuccess;
        }
        if (!ok) st = LBM_E_HIP;
        else if (over) st = LBM_E_UNSUPPORTED;
    }
    void* bufs[] = {dV, dT, dN, dO, dG, dF};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    return st;
}
