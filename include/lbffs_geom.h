/*
 * lbffs_geom.h -- geometry preprocessing on the GPU, as a C ABI.
 *
 * The reference builds the solid flags and the wall distances on the host before its
 * time loop (D3Q19_opencl/ns.cpp:188-226, the same in D3Q27_opencl/ns.cpp): setSolid
 * from the walls STL, setNeiSolid, setSDF per patch when LES is on, then
 * setSDFnearBwalls.  These entry points replace setSolid and setSDF with HIP kernels;
 * setNeiSolid and setSDFnearBwalls are host functions of lbffs_case.h.
 *
 * Triangles come as three (or two) arrays of 3*n floats, one row per axis (x row, then
 * y row, then z row), like STLpatch::STLv0/STLv1/STLv2, STLc and STLnormal
 * (src/stl.hpp:116-120), in lattice units (vertices divided by the cell size L,
 * stl.hpp:158-172).  Host arrays in and out; each call runs on the current HIP device,
 * allocates its device buffers and frees them before it returns.  Returns 0 or a negative
 * LBFFS_E_* code (lbffs.h); lbffs_last_error() describes the last failure.
 */
#ifndef LBFFS_GEOM_H
#define LBFFS_GEOM_H

#include <stdint.h>

#include "lbffs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/walls.hpp:159-184 setSolid (with isInner, isCross and det, walls.hpp:18-157, and
 * the bounding boxes of stl.hpp:186-262 and stl.hpp:525-550).  n triangles of every
 * patch, concatenated.  Cells whose +x ray crosses the surface an odd number of times
 * get solid = 1, other cells keep their value; with invert != 0 every cell then becomes
 * (solid == 1 ? 0 : 1).  solid: nx*ny*nz bytes, in and out.  Bit-identical to the
 * reference host (FP32, no contraction). */
int lbffs_geom_solid(int n, const float *v0, const float *v1, const float *v2, int nx, int ny, int nz, int invert,
                     uint8_t *solid);

/* src/walls.hpp:218-297 setSDF (the per-patch overload called by ns.cpp:205-213).  Every
 * cell within int(dr) cells (a box, clamped to the lattice) of a cell with neiSolid == 1
 * gets the inverse-distance-weighted (weight r^-p) average of the signed distances to the
 * planes of the patch's triangles whose centroid lies within dr, or sdf_ini when there is
 * none; other cells keep their value.  centroid, normal: 3*n floats; sdf: nx*ny*nz floats,
 * in and out.  FP32 sums in triangle order; r^-p is rounded once from FP64. */
int lbffs_geom_sdf(int n, const float *centroid, const float *normal, float sdf_ini, float dr, float p, int nx, int ny,
                   int nz, const uint8_t *neiSolid, float *sdf);

/* q-distances of the solid links for the interpolated bounce-back (LBFFS_FIELD_QLINK of
 * lbffs.h), from the surface in FP64.  No counterpart runs in the reference, whose setQf
 * (walls.hpp:553-584) derives q from the signed distances and is never uploaded.  For every
 * cell with neiSolid == 1 and solid == 0 and every direction c_d whose upwind cell x - c_d
 * (periodic wrap, as setNeiSolid) is solid, q is the smallest t in [0, 1] at which the
 * segment x - t c_d meets one of the n triangles (Moller-Trumbore in FP64, vertices as
 * given: no periodic images); qlink[d*N + ic] = (float)q.  Links that meet no triangle keep
 * their value and are counted in *misses (may be NULL).  qlink: q*nx*ny*nz floats, in and
 * out, q-major. */
int lbffs_geom_qlinks(int q, int n, const float *v0, const float *v1, const float *v2, int nx, int ny, int nz,
                      const uint8_t *solid, const uint8_t *neiSolid, float *qlink, int *misses);

/* The same q-distances cast straight into a context's LBFFS_FIELD_QLINK (the lattice,
 * solid and neiSolid flags are the context's, as uploaded), so that no q*N host array
 * exists: at 1024 x 384 x 320 cells the D3Q27 array is 13.6 GB.  Links that meet no
 * triangle, and every link that is not a solid link, are set to init.  The array is
 * allocated on the first call and counts as uploaded: lbffs_step accepts ibb = 1 after it.
 * Replaces the setQf buffer fill and upload the reference leaves out (D3Q27_opencl/
 * ns.cpp:408, walls.hpp:553-584).  Waits for the context's queued work first. */
int lbffs_geom_qlinks_ctx(lbffs_ctx *ctx, int n, const float *v0, const float *v1, const float *v2, float init,
                          int *misses);

#ifdef __cplusplus
}
#endif
#endif
