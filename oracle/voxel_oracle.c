/*
 * voxel_oracle.c -- brute-force OFF voxelizer (TEST INFRASTRUCTURE ONLY).
 * See voxel_oracle.h.
 */
#include "voxel_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void vox_free(void* p) { free(p); }

/* Reads the next non-empty, non-comment token line; returns NULL at EOF. */
static char* next_line(FILE* f, char* buf, int n)
{
    while (fgets(buf, n, f)) {
        char* s = buf;
        while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n') ++s;
        if (*s == '\0' || *s == '#') continue;
        return s;
    }
    return NULL;
}

int vox_read_off(const char* path, double** verts, int* nv, int** tris, int* nt)
{
    *verts = NULL;
    *tris = NULL;
    *nv = *nt = 0;
    FILE* f = fopen(path, "r");
    if (!f) return 2;
    char buf[4096];
    char* s = next_line(f, buf, sizeof buf);
    if (!s) { fclose(f); return 1; }
    int has_normals = 0;
    if (strncmp(s, "NOFF", 4) == 0) has_normals = 1;
    else if (strncmp(s, "COFF", 4) == 0 || strncmp(s, "OFF", 3) == 0) has_normals = 0;
    else { fclose(f); return 2; }
    /* header counts may follow the keyword on the same line */
    char* rest = s + (has_normals || strncmp(s, "COFF", 4) == 0 ? 4 : 3);
    int v = -1, fc = -1, e = 0;
    if (sscanf(rest, "%d %d %d", &v, &fc, &e) < 2) {
        s = next_line(f, buf, sizeof buf);
        if (!s || sscanf(s, "%d %d %d", &v, &fc, &e) < 2) { fclose(f); return 2; }
    }
    if (v <= 0 || fc <= 0) { fclose(f); return 1; }
    double* V = (double*)malloc(sizeof(double) * 3 * v);
    int* T = (int*)malloc(sizeof(int) * 3 * (size_t)fc * 4);
    int ntri = 0;
    for (int i = 0; i < v; ++i) {
        s = next_line(f, buf, sizeof buf);
        if (!s || sscanf(s, "%lf %lf %lf", &V[3 * i], &V[3 * i + 1], &V[3 * i + 2]) != 3) {
            free(V); free(T); fclose(f); return 2;
        }
    }
    for (int i = 0; i < fc; ++i) {
        s = next_line(f, buf, sizeof buf);
        int k, idx[64], off = 0, used = 0;
        if (!s || sscanf(s, "%d%n", &k, &used) != 1 || k < 3 || k > 64) { free(V); free(T); fclose(f); return 2; }
        off += used;
        for (int j = 0; j < k; ++j) {
            if (sscanf(s + off, "%d%n", &idx[j], &used) != 1 || idx[j] < 0 || idx[j] >= v) {
                free(V); free(T); fclose(f); return 2;
            }
            off += used;
        }
        /* fan triangulation of polygonal faces */
        for (int j = 1; j + 1 < k; ++j) {
            if (ntri >= fc * 4) break;
            T[3 * ntri] = idx[0];
            T[3 * ntri + 1] = idx[j];
            T[3 * ntri + 2] = idx[j + 1];
            ++ntri;
        }
    }
    fclose(f);
    *verts = V;
    *tris = T;
    *nv = v;
    *nt = ntri;
    return 0;
}

void vox_bbox(const double* V, int nv, double mn[3], double mx[3])
{
    for (int a = 0; a < 3; ++a) { mn[a] = V[a]; mx[a] = V[a]; }
    for (int i = 1; i < nv; ++i)
        for (int a = 0; a < 3; ++a) {
            if (V[3 * i + a] < mn[a]) mn[a] = V[3 * i + a];
            if (V[3 * i + a] > mx[a]) mx[a] = V[3 * i + a];
        }
}

void vox_dims(const float mn[3], const float mx[3], float res, float add, int dims[3])
{
    for (int a = 0; a < 3; ++a) {
        float v = (mx[a] - mn[a]) * res + 2 * add;
        dims[a] = (int)ceilf(v);
    }
}

/* sign of det(b-a, c-a, d-a) */
static double orient(const double* a, const double* b, const double* c, const double* d)
{
    double bx = b[0] - a[0], by = b[1] - a[1], bz = b[2] - a[2];
    double cx = c[0] - a[0], cy = c[1] - a[1], cz = c[2] - a[2];
    double dx = d[0] - a[0], dy = d[1] - a[1], dz = d[2] - a[2];
    return bx * (cy * dz - cz * dy) - by * (cx * dz - cz * dx) + bz * (cx * dy - cy * dx);
}

static int sgn(double v) { return (v > 0) - (v < 0); }

/* Segment pq against triangle abc: returns 1 and the point when they meet in
 * a single point (edge and vertex contacts included), 0 otherwise; a segment
 * lying in the triangle's plane is not a point intersection. */
static int seg_tri(const double* p, const double* q, const double* a, const double* b, const double* c,
                   double* out)
{
    const int op = sgn(orient(a, b, c, p));
    const int oq = sgn(orient(a, b, c, q));
    if (op == oq) return 0; /* same side, or coplanar */
    const int s1 = sgn(orient(p, q, a, b));
    const int s2 = sgn(orient(p, q, b, c));
    const int s3 = sgn(orient(p, q, c, a));
    const int pos = (s1 >= 0) && (s2 >= 0) && (s3 >= 0);
    const int neg = (s1 <= 0) && (s2 <= 0) && (s3 <= 0);
    if (!pos && !neg) return 0;
    /* intersection of line pq with the plane of abc */
    double n[3], ab[3], ac[3];
    for (int k = 0; k < 3; ++k) { ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; }
    n[0] = ab[1] * ac[2] - ab[2] * ac[1];
    n[1] = ab[2] * ac[0] - ab[0] * ac[2];
    n[2] = ab[0] * ac[1] - ab[1] * ac[0];
    double num = 0, den = 0;
    for (int k = 0; k < 3; ++k) { num += n[k] * (a[k] - p[k]); den += n[k] * (q[k] - p[k]); }
    if (den == 0) return 0;
    const double t = num / den;
    for (int k = 0; k < 3; ++k) out[k] = p[k] + t * (q[k] - p[k]);
    return 1;
}

int vox_inside(const double* V, const int* T, int nt, const double query[3], const double outside[3])
{
    if (query[0] == outside[0] && query[1] == outside[1] && query[2] == outside[2]) return 0;
    double* pts = (double*)malloc(sizeof(double) * 3 * (nt > 0 ? nt : 1));
    int np = 0;
    for (int t = 0; t < nt; ++t) {
        double x[3];
        if (seg_tri(query, outside, &V[3 * T[3 * t]], &V[3 * T[3 * t + 1]], &V[3 * T[3 * t + 2]], x)) {
            pts[3 * np] = x[0]; pts[3 * np + 1] = x[1]; pts[3 * np + 2] = x[2];
            ++np;
        }
    }
    int n_dist = 0;
    for (int i = 0; i < np; ++i) {
        int distinct = 1;
        for (int j = 0; j < i; ++j) {
            double dx = pts[3 * i] - pts[3 * j], dy = pts[3 * i + 1] - pts[3 * j + 1], dz = pts[3 * i + 2] - pts[3 * j + 2];
            distinct = (dx * dx + dy * dy + dz * dz) > 1e-10;
            if (!distinct) break;
        }
        if (distinct) ++n_dist;
    }
    free(pts);
    return n_dist % 2 == 1;
}

void vox_voxelize(const double* V, const int* T, int nt, const float mn[3], float res, float add, const int dims[3],
                  const double outside[3], int* grid)
{
    for (int i = 0; i < dims[0]; i++)
        for (int j = 0; j < dims[1]; j++)
            for (int k = 0; k < dims[2]; k++) {
                /* query point computed in double, held in float (geometryMesherOFF.h:90-103) */
                float lx = (float)(((double)i + 0.5 - (double)add) / (double)res + (double)mn[0]);
                float ly = (float)(((double)j + 0.5 - (double)add) / (double)res + (double)mn[1]);
                float lz = (float)(((double)k + 0.5 - (double)add) / (double)res + (double)mn[2]);
                double qp[3] = {lx, ly, lz};
                grid[(size_t)i + (size_t)dims[0] * ((size_t)j + (size_t)dims[1] * (size_t)k)] =
                    vox_inside(V, T, nt, qp, outside) ? 1 : 0;
            }
}
