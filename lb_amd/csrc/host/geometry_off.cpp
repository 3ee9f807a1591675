// geometry_off.cpp -- OFF polyhedron, inside test and voxelization.
//
// Follows the reference preprocess: geometryHandlerOFF
// (3D/preprocess/src/geometry/geometryHandlerOFF.h:63-182) and
// geometryMesherOFF (3D/preprocess/src/geometry/geometryMesherOFF.h:45-106).
// A point is inside when the segment from it to a point known to be outside
// meets the surface in an odd number of distinct points (points closer than
// sqrt(1e-10) are merged, as the reference does to count an edge or vertex
// hit once).  CGAL's AABB tree is replaced by a binary bounding-volume
// hierarchy; the segment/triangle test uses orientation signs evaluated with a
// canonical vertex order, so that the two triangles sharing an edge always
// agree on which side of it the segment passes.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "lbm_host.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace lbmh {

struct geometryHandlerOFF::Impl {
    std::vector<std::array<double, 3>> V;
    std::vector<std::array<int, 3>> T;
    struct Node {
        double lo[3], hi[3];
        int left, right;   // children (internal) or -1
        int first, count;  // triangle range (leaf)
    };
    std::vector<Node> nodes;
    std::vector<int> order;  // triangle ids in leaf order

    void box_of(int t, double* lo, double* hi) const
    {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min({V[T[t][0]][a], V[T[t][1]][a], V[T[t][2]][a]});
            hi[a] = std::max({V[T[t][0]][a], V[T[t][1]][a], V[T[t][2]][a]});
        }
    }

    int build(int first, int count)
    {
        Node n;
        for (int a = 0; a < 3; ++a) {
            n.lo[a] = 1e300;
            n.hi[a] = -1e300;
        }
        for (int k = first; k < first + count; ++k) {
            double lo[3], hi[3];
            box_of(order[k], lo, hi);
            for (int a = 0; a < 3; ++a) {
                n.lo[a] = std::min(n.lo[a], lo[a]);
                n.hi[a] = std::max(n.hi[a], hi[a]);
            }
        }
        const int id = (int)nodes.size();
        nodes.push_back(n);
        if (count <= 4) {
            nodes[id].left = nodes[id].right = -1;
            nodes[id].first = first;
            nodes[id].count = count;
            return id;
        }
        // split at the median centroid along the widest axis
        int axis = 0;
        double ext = -1;
        for (int a = 0; a < 3; ++a)
            if (n.hi[a] - n.lo[a] > ext) {
                ext = n.hi[a] - n.lo[a];
                axis = a;
            }
        auto cen = [&](int t) { return V[T[t][0]][axis] + V[T[t][1]][axis] + V[T[t][2]][axis]; };
        const int mid = first + count / 2;
        std::nth_element(order.begin() + first, order.begin() + mid, order.begin() + first + count,
                         [&](int a, int b) { return cen(a) < cen(b); });
        const int l = build(first, mid - first);
        const int r = build(mid, first + count - mid);
        nodes[id].left = l;
        nodes[id].right = r;
        nodes[id].first = first;
        nodes[id].count = 0;
        return id;
    }

    void build_all()
    {
        order.resize(T.size());
        for (size_t i = 0; i < T.size(); ++i) order[i] = (int)i;
        nodes.clear();
        nodes.reserve(2 * T.size() + 1);
        if (!T.empty()) build(0, (int)T.size());
    }

    // segment p + t d, t in [0, 1] against a box (slab test)
    static bool seg_box(const double* p, const double* d, const double* lo, const double* hi)
    {
        double t0 = 0.0, t1 = 1.0;
        for (int a = 0; a < 3; ++a) {
            if (d[a] == 0.0) {
                if (p[a] < lo[a] || p[a] > hi[a]) return false;
                continue;
            }
            double ta = (lo[a] - p[a]) / d[a], tb = (hi[a] - p[a]) / d[a];
            if (ta > tb) std::swap(ta, tb);
            t0 = std::max(t0, ta);
            t1 = std::min(t1, tb);
            if (t0 > t1) return false;
        }
        return true;
    }

    // signed volume of (a, b, c, d) = det(b - a, c - a, d - a)
    static double vol(const double* a, const double* b, const double* c, const double* d)
    {
        const double b0 = b[0] - a[0], b1 = b[1] - a[1], b2 = b[2] - a[2];
        const double c0 = c[0] - a[0], c1 = c[1] - a[1], c2 = c[2] - a[2];
        const double d0 = d[0] - a[0], d1 = d[1] - a[1], d2 = d[2] - a[2];
        return b0 * (c1 * d2 - c2 * d1) - b1 * (c0 * d2 - c2 * d0) + b2 * (c0 * d1 - c1 * d0);
    }

    static int sgn(double v) { return (v > 0) - (v < 0); }

    // side of the segment line (p, q) with respect to edge (u, v), evaluated
    // with the lower vertex id first so both triangles of an edge agree
    int edge_side(const double* p, const double* q, int iu, int iv) const
    {
        if (iu < iv) return sgn(vol(p, q, V[iu].data(), V[iv].data()));
        return -sgn(vol(p, q, V[iv].data(), V[iu].data()));
    }

    bool seg_tri(const double* p, const double* q, int t, double* x) const
    {
        const int i0 = T[t][0], i1 = T[t][1], i2 = T[t][2];
        const double *a = V[i0].data(), *b = V[i1].data(), *c = V[i2].data();
        const double vp = vol(a, b, c, p), vq = vol(a, b, c, q);
        const int sp = sgn(vp), sq = sgn(vq);
        if (sp == sq) return false;  // same side, or the segment lies in the plane
        const int s1 = edge_side(p, q, i0, i1), s2 = edge_side(p, q, i1, i2), s3 = edge_side(p, q, i2, i0);
        const bool pos = s1 >= 0 && s2 >= 0 && s3 >= 0, neg = s1 <= 0 && s2 <= 0 && s3 <= 0;
        if (!pos && !neg) return false;
        const double t_ = vp / (vp - vq);
        for (int k = 0; k < 3; ++k) x[k] = p[k] + t_ * (q[k] - p[k]);
        return true;
    }

    int count_crossings(const double* p, const double* q) const
    {
        if (nodes.empty()) return 0;
        const double d[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
        std::vector<std::array<double, 3>> pts;
        int stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const Node& n = nodes[stack[--sp]];
            if (!seg_box(p, d, n.lo, n.hi)) continue;
            if (n.left < 0) {
                for (int k = n.first; k < n.first + n.count; ++k) {
                    double x[3];
                    if (seg_tri(p, q, order[k], x)) pts.push_back({x[0], x[1], x[2]});
                }
            } else if (sp + 2 <= 128) {
                stack[sp++] = n.left;
                stack[sp++] = n.right;
            }
        }
        // distinct points (geometryHandlerOFF.h:166-179)
        int n_dist = 0;
        for (size_t i = 0; i < pts.size(); ++i) {
            bool distinct = true;
            for (size_t j = 0; j < i; ++j) {
                const double dx = pts[i][0] - pts[j][0], dy = pts[i][1] - pts[j][1], dz = pts[i][2] - pts[j][2];
                distinct = (dx * dx + dy * dy + dz * dz) > 1e-10;
                if (!distinct) break;
            }
            if (distinct) ++n_dist;
        }
        return n_dist;
    }
};

geometryHandlerOFF::geometryHandlerOFF() : impl(new Impl) {}
geometryHandlerOFF::~geometryHandlerOFF() = default;
size_t geometryHandlerOFF::triangles() const { return impl->T.size(); }

// Strips comments and returns the next non-empty line.
static bool next_line(std::istream& in, std::string& line)
{
    while (std::getline(in, line)) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line.erase(h);
        if (line.find_first_not_of(" \t\r") != std::string::npos) return true;
    }
    return false;
}

void geometryHandlerOFF::polyhedronFromFile(LBMConstants& Constants, bool verbose)
{
    impl->V.clear();
    impl->T.clear();
    std::ifstream in(Constants.file_name);
    if (verbose) std::cout << " Reading file " << Constants.file_name << " " << std::endl;
    std::string line;
    if (!in.is_open() || !next_line(in, line)) {
        error = in.is_open() ? 1 : 2;
        return;
    }
    std::istringstream hs(line);
    std::string key;
    hs >> key;
    if (key != "OFF" && key != "NOFF" && key != "COFF" && key != "CNOFF" && key != "NCOFF") {
        error = 2;
        return;
    }
    long nv = -1, nf = -1, ne = 0;
    if (!(hs >> nv >> nf)) {
        if (!next_line(in, line)) { error = 2; return; }
        std::istringstream cs(line);
        if (!(cs >> nv >> nf)) { error = 2; return; }
        cs >> ne;
    }
    if (nv <= 0 || nf <= 0) {
        error = 1;  // empty polyhedron (geometryHandlerOFF.h:93-98)
        return;
    }
    impl->V.resize(nv);
    for (long i = 0; i < nv; ++i) {
        if (!next_line(in, line)) { error = 2; return; }
        std::istringstream vs(line);
        if (!(vs >> impl->V[i][0] >> impl->V[i][1] >> impl->V[i][2])) { error = 2; return; }
    }
    for (long f = 0; f < nf; ++f) {
        if (!next_line(in, line)) { error = 2; return; }
        std::istringstream fs(line);
        int k;
        if (!(fs >> k) || k < 3) { error = 2; return; }
        std::vector<int> idx(k);
        for (int j = 0; j < k; ++j)
            if (!(fs >> idx[j]) || idx[j] < 0 || idx[j] >= nv) { error = 2; return; }
        for (int j = 1; j + 1 < k; ++j) impl->T.push_back({idx[0], idx[j], idx[j + 1]});
    }
    impl->build_all();
    if (verbose) {
        std::cout << " facets: " << nf << std::endl;
        std::cout << " vertices: " << nv << std::endl;
    }
    error = 0;
}

void geometryHandlerOFF::polyhedronBbox(bool verbose)
{
    if (impl->V.empty()) return;
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = impl->V[0][a];
    for (const auto& v : impl->V)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], v[a]);
            hi[a] = std::max(hi[a], v[a]);
        }
    d3_min_BB = {lo[0], lo[1], lo[2]};
    d3_max_BB = {hi[0], hi[1], hi[2]};
    if (verbose)
        std::cout << "BBox is [" << lo[0] << ", " << lo[1] << ", " << lo[2] << "] to [" << hi[0] << ", " << hi[1]
                  << ", " << hi[2] << "].\n";
}

// The double bounding box is truncated into the float constants
// (geometryHandlerOFF.h:124-133, LBMConstants BBmin/BBmax are RealType).
void geometryHandlerOFF::writeToConstants(LBMConstants& Constants) const
{
    Constants.BBminx = (RealType)d3_min_BB.x;
    Constants.BBmaxx = (RealType)d3_max_BB.x;
    Constants.BBminy = (RealType)d3_min_BB.y;
    Constants.BBmaxy = (RealType)d3_max_BB.y;
    Constants.BBminz = (RealType)d3_min_BB.z;
    Constants.BBmaxz = (RealType)d3_max_BB.z;
}

bool geometryHandlerOFF::polyhedronInside(d3 pnt_ask, d3 pnt_ref) const
{
    if (pnt_ask.x == pnt_ref.x && pnt_ask.y == pnt_ref.y && pnt_ask.z == pnt_ref.z) return false;
    const double p[3] = {pnt_ask.x, pnt_ask.y, pnt_ask.z}, q[3] = {pnt_ref.x, pnt_ref.y, pnt_ref.z};
    return impl->count_crossings(p, q) % 2 == 1;
}

int geometryHandlerOFF::insideGridDevice(const int32_t dims[3], double additional, double resolution,
                                         const double bbmin[3], d3 pnt_ref, int32_t* grid) const
{
    const Impl& I = *impl;
    std::vector<double> V(3 * I.V.size());
    for (size_t k = 0; k < I.V.size(); ++k)
        for (int a = 0; a < 3; ++a) V[3 * k + a] = I.V[k][a];
    std::vector<int32_t> T(3 * I.T.size());
    for (size_t k = 0; k < I.T.size(); ++k)
        for (int a = 0; a < 3; ++a) T[3 * k + a] = I.T[k][a];
    std::vector<lbm_bvh_node> N(I.nodes.size());
    for (size_t k = 0; k < I.nodes.size(); ++k) {
        for (int a = 0; a < 3; ++a) {
            N[k].lo[a] = I.nodes[k].lo[a];
            N[k].hi[a] = I.nodes[k].hi[a];
        }
        N[k].left = I.nodes[k].left;
        N[k].right = I.nodes[k].right;
        N[k].first = I.nodes[k].first;
        N[k].count = I.nodes[k].count;
    }
    const std::vector<int32_t> O(I.order.begin(), I.order.end());
    const double q[3] = {pnt_ref.x, pnt_ref.y, pnt_ref.z};
    return lbm_voxelize(V.data(), (int32_t)I.V.size(), T.data(), (int32_t)I.T.size(), N.data(), (int32_t)N.size(),
                        O.data(), dims, additional, resolution, bbmin, q, grid);
}

// Grid size in float arithmetic (geometryMesherOFF.h:45-79).
void geometryMesherOFF::initialization(bool verbose)
{
    Constants.dimX_int = (int)std::ceil((Constants.BBmaxx - Constants.BBminx) * Constants.resolution_factor +
                                        2 * Constants.additional_factor);
    Constants.dimY_int = (int)std::ceil((Constants.BBmaxy - Constants.BBminy) * Constants.resolution_factor +
                                        2 * Constants.additional_factor);
    Constants.dimZ_int = (int)std::ceil((Constants.BBmaxz - Constants.BBminz) * Constants.resolution_factor +
                                        2 * Constants.additional_factor);
    Data.setSizes(Constants.dimX_int, Constants.dimY_int, Constants.dimZ_int);
    if (verbose) {
        std::cout << "Resolution factor: " << Constants.resolution_factor << std::endl;
        std::cout << "Additional factor: " << Constants.additional_factor << std::endl;
        std::cout << "Dimension X: " << Constants.dimX_int << std::endl;
        std::cout << "Dimension Y: " << Constants.dimY_int << std::endl;
        std::cout << "Dimension Z: " << Constants.dimZ_int << std::endl;
    }
}

// Query point of each cell computed in double and held in a float
// (geometryMesherOFF.h:90-103); fluid = 1.
void geometryMesherOFF::meshingPolyhedron(const geometryHandlerOFF& handler, d3 pnt_outside, bool verbose,
                                          int threads)
{
    const int nx = Constants.dimX_int, ny = Constants.dimY_int, nz = Constants.dimZ_int;
    const double add = Constants.additional_factor, res = Constants.resolution_factor;
    const double bx = Constants.BBminx, by = Constants.BBminy, bz = Constants.BBminz;
    if (threads < 0) {
        // the same tests on the GPU; the grid is i fastest, like Data.mesh
        const int32_t dims[3] = {nx, ny, nz};
        const double bb[3] = {bx, by, bz};
        device_status = handler.insideGridDevice(dims, add, res, bb, pnt_outside, Data.meshFluidHost.data());
        if (verbose) std::cout << "Number of lattice points is " << (long)nx * ny * nz << "\n";
        return;
    }
#ifdef _OPENMP
    const int nth = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for collapse(2) schedule(dynamic, 4) num_threads(nth)
#endif
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const RealType lookx = (RealType)(((double)i + 0.5 - add) / res + bx);
                const RealType looky = (RealType)(((double)j + 0.5 - add) / res + by);
                const RealType lookz = (RealType)(((double)k + 0.5 - add) / res + bz);
                if (handler.polyhedronInside({lookx, looky, lookz}, pnt_outside)) Data.mesh(i, j, k) = 1;
            }
    (void)threads;
    if (verbose) std::cout << "Number of lattice points is " << (long)nx * ny * nz << "\n";
}

}  // namespace lbmh
