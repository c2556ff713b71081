/*
 * lbffs_case.h -- host-side case setup of the LBFFS driver, as a C ABI.
 *
 * These functions build, on the CPU, the per-cell arrays that the reference host
 * prepares before its time loop and that the step kernel consumes (lbffs.h fields):
 * unit conversion, face BC codes and wall velocities, the domain-face flag, the
 * initial equilibrium populations, the solid-neighbour flag and the outlet-patch
 * omega sponge.  Results are bit-identical to the reference's host code for the same
 * inputs (x86-64, no FMA contraction).
 */
#ifndef LBFFS_CASE_H
#define LBFFS_CASE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Faces in the order of boundaryConditions.txt: nxMin, nxMax, nyMin, nyMax, nzMin, nzMax. */
enum { LBFFS_NXMIN = 0, LBFFS_NXMAX = 1, LBFFS_NYMIN = 2, LBFFS_NYMAX = 3, LBFFS_NZMIN = 4, LBFFS_NZMAX = 5 };

/* The input.txt quantities the step depends on (input.hpp:115-219), physical units. */
typedef struct {
    int nx, ny, nz;
    float Lx, uMax, rho0, U0, nu, dpdx;
    float uIni, vIni, wIni;
    int face_code[6][3];  /* <face>Boundary1..3 as BC codes (input.hpp:221-256) */
    float face_vel[6][3]; /* <face>U0,V0,W0 in m/s */
} lbffs_case_in;

/* Derived lattice quantities (ns.cpp:104-128, input.hpp:282-287). */
typedef struct {
    float L;        /* cell size Lx/(nx-1) */
    float c;        /* velocity scale uMax/U0 */
    float Re;       /* uMax*Lx/nu */
    float nu;       /* lattice viscosity */
    float dpdx;     /* lattice body force */
    float uIni, vIni, wIni;
    float omega0;   /* 1/(3 nu + 1/2) */
    float deltaT;   /* L/c */
    float face_vel[6][3]; /* face velocities in lattice units (divided by c) */
} lbffs_units;

/* ns.cpp:104-128 and the /c of input.hpp:282-287. */
int lbffs_case_units(const lbffs_case_in *in, lbffs_units *out);

/* setBoundaryConditions.hpp:1-186 and the face flag loop ns.cpp:164-177.
 * Arrays of nx*ny*nz entries; they are fully overwritten. */
int lbffs_case_faces(const lbffs_case_in *in, const lbffs_units *un, int32_t *boundary, int32_t *b1, int32_t *b2,
                     int32_t *b3, float *u0, float *v0, float *w0);

/* initialization.hpp:1-27: f = fTmp = (rho + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2) w_q, q-major. */
int lbffs_case_init_feq(int q, int nx, int ny, int nz, float rho0, float u, float v, float w, float *f);

/* walls.hpp:187-216 setNeiSolid (upwind neighbours with periodic wrap, faces excluded). */
int lbffs_case_neisolid(int q, int nx, int ny, int nz, const uint8_t *solid, uint8_t *neiSolid);

/* ns.cpp:343-375: cosine omega sponge of width spz cells around one outlet-patch cell
 * (i,j,k) with face codes (p1,p2,p3); omega is updated in place. */
int lbffs_case_outlet_sponge(int nx, int ny, int nz, int i, int j, int k, int p1, int p2, int p3, int spz,
                             float omega0, float *omega);

/* stl.hpp:317-453 STLpatch::findPoints: the lattice face cells an inlet or outlet patch
 * covers, for the patch loops of ns.cpp:278-377 (which then set the cells' face codes and
 * u0/v0/w0, and for outlets call lbffs_case_outlet_sponge per cell).  Per triangle, in
 * order: its plane is the axis of the largest |normal| component -- |.| is the reference's
 * unqualified abs() on a float, which resolves to int abs(int) with the headers stl.hpp
 * sees (a component is truncated toward zero first) -- and if its centroid truncates to 0
 * or n-1 of that axis (or to -1, n-2 with the +1 of stl.hpp), every cell of that lattice
 * face whose in-plane coordinates lie in the triangle (pointInElement, stl.hpp:95-111, FP32,
 * bounds inclusive) is listed; a cell in several triangles is listed once per triangle.
 * v0, v1, v2, normal, centroid: rows x, y, z of n floats each (3*n), lattice units.  Writes
 * at most cap indices ic = i + nx*(j + ny*k) to points (may be NULL with cap = 0) and
 * the full count to *count. */
int lbffs_case_find_points(int n, const float *v0, const float *v1, const float *v2, const float *normal,
                           const float *centroid, int nx, int ny, int nz, int32_t *points, long cap, long *count);

/* walls.hpp:372-469 setSDFnearBwalls: sdf = 0.5 + d in the int(dr)+1 layers d = 0, 1, ...
 * next to every domain face whose face cell has boundary code 1 (BounceBack) in the
 * face's normal-direction code array (b1 for x faces, b2 for y, b3 for z); a layer that
 * would reach the opposite face cell is skipped.  sdf: nx*ny*nz floats, updated in place. */
int lbffs_case_sdf_near_bwalls(float dr, const int32_t *b1, const int32_t *b2, const int32_t *b3, int nx, int ny, int nz,
                               float *sdf);

#ifdef __cplusplus
}
#endif
#endif
