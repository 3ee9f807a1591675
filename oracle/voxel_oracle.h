/*
 * voxel_oracle.h -- brute-force OFF voxelizer (TEST INFRASTRUCTURE ONLY).
 *
 * Restates the reference preprocess path: geometryHandlerOFF::polyhedronInside
 * (3D/preprocess/src/geometry/geometryHandlerOFF.h:135-182) and
 * geometryMesherOFF::{initialization, meshingPolyhedron}
 * (3D/preprocess/src/geometry/geometryMesherOFF.h:45-106), with every
 * triangle tested against every query segment instead of CGAL's AABB tree.
 * The product voxelizer (lb_amd/csrc/host) is checked against it and against
 * the reference's committed voxelization.
 */
#ifndef VOXEL_ORACLE_H
#define VOXEL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Parse an OFF / NOFF / COFF file.  Returns 0 on success, 1 on an empty
 * polyhedron, 2 on a parse failure (the reference's private error codes,
 * geometryHandlerOFF.h:79-97).  *verts (3*nv doubles) and *tris (3*nt ints)
 * are malloc'ed; free with vox_free. */
int vox_read_off(const char* path, double** verts, int* nv, int** tris, int* nt);
void vox_free(void* p);

/* Axis-aligned bounding box of the vertices. */
void vox_bbox(const double* verts, int nv, double bbmin[3], double bbmax[3]);

/* Grid size from geometryMesherOFF::initialization: float arithmetic,
 * ceil((BBmax - BBmin) * res + 2 * add). */
void vox_dims(const float bbmin[3], const float bbmax[3], float res, float add, int dims[3]);

/* Inside test of one query point: odd number of distinct intersection
 * points of the segment (query, outside) with the triangles; points closer
 * than sqrt(1e-10) are merged. */
int vox_inside(const double* verts, const int* tris, int nt, const double query[3], const double outside[3]);

/* 0/1 grid [z][y][x] of dims[0]*dims[1]*dims[2] ints (fluid = 1). */
void vox_voxelize(const double* verts, const int* tris, int nt, const float bbmin[3], float res, float add,
                  const int dims[3], const double outside[3], int* grid);

#ifdef __cplusplus
}
#endif
#endif
