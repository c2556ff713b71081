/*
 * geom_oracle.c -- geometry preprocessing of the CPU oracle.  TEST INFRASTRUCTURE ONLY:
 * the checker of lbffs_geom_solid / lbffs_geom_sdf; only tests/ may load it.
 *
 * Restates, in plain C with the reference's order of FP32 operations (the library is
 * built with -ffp-contract=off, as the reference host runs without FMA on x86-64):
 *   det, isCross, isInner, setSolid         src/walls.hpp:18-184
 *   bounding boxes of the surface/triangles src/stl.hpp:186-262, 525-550
 *   setSDF (per-patch overload)             src/walls.hpp:218-297
 * Triangles: v0, v1, v2 (and centroids, normals) as 3*n floats, rows x, y, z.
 * pow(x, 2.f) is written x * x: clang -O3, the reference's compiler
 * (D3Q19_opencl/CMakeLists.txt:4,10), folds pow(x, 2) to one multiplication.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"

/* walls.hpp:18-22 */
static float det3(const float *a, const float *b, const float *c)
{
    return (a[0] * b[1] * c[2]) + (a[1] * b[2] * c[0]) + (a[2] * b[0] * c[1]) - (a[0] * b[2] * c[1]) -
           (a[1] * b[0] * c[2]) - (a[2] * b[1] * c[0]);
}

/* walls.hpp:24-62 */
int orc_geom_is_cross(const float *v0, const float *v1, const float *v2, int i, int j, int k)
{
    const float a[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const float b[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    const float c[3] = {-1.f, 0.f, 0.f};
    const float d[3] = {(float)i - v0[0], (float)j - v0[1], (float)k - v0[2]};
    const float x = det3(d, b, c) / det3(a, b, c);
    const float y = det3(a, d, c) / det3(a, b, c);
    const float z = det3(a, b, d) / det3(a, b, c);
    if (x < 0.f || x > 1.f) return 0;
    if (y < 0.f || y > 1.f) return 0;
    if (x + y < 0.f || x + y > 1.f) return 0;
    if (z < 0.f) return 0;
    return 1;
}

static void vertex(const float *v, int n, int t, float *out)
{
    out[0] = v[t];
    out[1] = v[n + t];
    out[2] = v[2 * n + t];
}

/* walls.hpp:159-184 with isInner (walls.hpp:64-157); the surface box starts at the
 * origin (stl.hpp:536-550), the triangle boxes are fmin/fmax of the vertices */
int orc_geom_solid(int n, const float *v0, const float *v1, const float *v2, int nx, int ny, int nz, int invert,
                   unsigned char *solid)
{
    float bx0 = 0.f, bx1 = 0.f, by0 = 0.f, by1 = 0.f, bz0 = 0.f, bz1 = 0.f;
    float *box = (float *)malloc(sizeof(float) * 4 * (n > 0 ? n : 1));
    if (!box) return -1;
    for (int t = 0; t < n; t++) {
        float a[3], b[3], c[3];
        vertex(v0, n, t, a);
        vertex(v1, n, t, b);
        vertex(v2, n, t, c);
        bx0 = fminf(bx0, fminf(a[0], fminf(b[0], c[0])));
        bx1 = fmaxf(bx1, fmaxf(a[0], fmaxf(b[0], c[0])));
        box[4 * t + 0] = fminf(a[1], fminf(b[1], c[1]));
        box[4 * t + 1] = fmaxf(a[1], fmaxf(b[1], c[1]));
        box[4 * t + 2] = fminf(a[2], fminf(b[2], c[2]));
        box[4 * t + 3] = fmaxf(a[2], fmaxf(b[2], c[2]));
        by0 = fminf(by0, box[4 * t + 0]);
        by1 = fmaxf(by1, box[4 * t + 1]);
        bz0 = fminf(bz0, box[4 * t + 2]);
        bz1 = fmaxf(bz1, box[4 * t + 3]);
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                const float fi = (float)i, fj = (float)j, fk = (float)k;
                if (fi < bx0 || fi > bx1 || fj < by0 || fj > by1 || fk < bz0 || fk > bz1) continue;
                int ncross = 0;
                for (int t = 0; t < n; t++) {
                    if (fj >= box[4 * t + 0] && fj <= box[4 * t + 1] && fk >= box[4 * t + 2] && fk <= box[4 * t + 3]) {
                        float a[3], b[3], c[3];
                        vertex(v0, n, t, a);
                        vertex(v1, n, t, b);
                        vertex(v2, n, t, c);
                        ncross += orc_geom_is_cross(a, b, c, i, j, k);
                    }
                }
                if (ncross % 2 != 0) solid[orc_index(i, j, k, nx, ny)] = 1;
            }
    if (invert) {
        const long N = (long)nx * ny * nz;
        for (long ic = 0; ic < N; ic++) solid[ic] = (solid[ic] == 1) ? 0 : 1;
    }
    free(box);
    return 0;
}

/* walls.hpp:218-297: the near cells (box of int(dr) around each neiSolid == 1 cell,
 * clamped), then per near cell the weighted average over the patch's triangles */
int orc_geom_sdf(int n, const float *c, const float *nrm, float sdfIni, float dr, float p, int nx, int ny, int nz,
                 const unsigned char *neiSolid, float *sdf)
{
    const long N = (long)nx * ny * nz;
    const int drn = (int)dr;
    unsigned char *near = (unsigned char *)calloc((size_t)N, 1);
    if (!near) return -1;
    for (long ic = 0; ic < N; ic++) {
        if (neiSolid[ic] != 1) continue;
        const int i = (int)(ic % ((long)nx * ny) % nx), j = (int)(ic % ((long)nx * ny) / nx),
                  k = (int)(ic / ((long)nx * ny));
        for (int ii = -drn; ii <= drn; ii++) {
            const int in = i + ii;
            if (in < 0 || in > nx - 1) continue;
            for (int jj = -drn; jj <= drn; jj++) {
                const int jn = j + jj;
                if (jn < 0 || jn > ny - 1) continue;
                for (int kk = -drn; kk <= drn; kk++) {
                    const int kn = k + kk;
                    if (kn < 0 || kn > nz - 1) continue;
                    near[orc_index(in, jn, kn, nx, ny)] = 1;
                }
            }
        }
    }
#pragma omp parallel for schedule(dynamic, 4096)
    for (long ic = 0; ic < N; ic++) {
        if (!near[ic]) continue;
        const int i = (int)(ic % ((long)nx * ny) % nx), j = (int)(ic % ((long)nx * ny) / nx),
                  k = (int)(ic / ((long)nx * ny));
        float acc = 0.f, sumD = 0.f;
        for (int t = 0; t < n; t++) {
            const float dx = (float)i - c[t], dy = (float)j - c[n + t], dz = (float)k - c[2 * n + t];
            const float r = sqrtf(dx * dx + dy * dy + dz * dz);
            if (r < dr) {
                const float sd = dx * nrm[t] + dy * nrm[n + t] + dz * nrm[2 * n + t];
                sumD += powf(r, -p);
                acc += sd * powf(r, -p);
            }
        }
        sdf[ic] = sumD != 0.f ? acc / sumD : sdfIni;
    }
    free(near);
    return 0;
}

/* q-distances of the solid links (the engine's lbffs_geom_qlinks; the reference has no
 * counterpart in use): for each cell with neiSolid == 1 and solid == 0 and each direction
 * whose upwind cell (periodic wrap, walls.hpp:187-216) is solid, the smallest t in [0, 1]
 * at which x - t c meets a triangle, by the Moller-Trumbore test in FP64 over every
 * triangle.  Links without a hit keep their value and are counted. */
static double link_hit(const double *v0, const double *e1, const double *e2, const double *x, const double *c)
{
    const double dir[3] = {-c[0], -c[1], -c[2]};
    const double p[3] = {dir[1] * e2[2] - dir[2] * e2[1], dir[2] * e2[0] - dir[0] * e2[2], dir[0] * e2[1] - dir[1] * e2[0]};
    const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (fabs(det) < 1e-14) return 2.0;
    const double inv = 1.0 / det;
    const double sv[3] = {x[0] - v0[0], x[1] - v0[1], x[2] - v0[2]};
    const double u = (sv[0] * p[0] + sv[1] * p[1] + sv[2] * p[2]) * inv;
    if (u < -1e-12 || u > 1.0 + 1e-12) return 2.0;
    const double qv[3] = {sv[1] * e1[2] - sv[2] * e1[1], sv[2] * e1[0] - sv[0] * e1[2], sv[0] * e1[1] - sv[1] * e1[0]};
    const double v = (dir[0] * qv[0] + dir[1] * qv[1] + dir[2] * qv[2]) * inv;
    if (v < -1e-12 || u + v > 1.0 + 1e-12) return 2.0;
    const double t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
    return (t >= 0.0 && t <= 1.0) ? t : 2.0;
}

int orc_geom_qlinks(int Q, int n, const float *v0, const float *v1, const float *v2, int nx, int ny, int nz,
                    const unsigned char *solid, const unsigned char *neiSolid, float *qlink, int *misses)
{
    const int *cx = Q == 27 ? ORC27_CX : ORC_CX, *cy = Q == 27 ? ORC27_CY : ORC_CY, *cz = Q == 27 ? ORC27_CZ : ORC_CZ;
    const long N = (long)nx * ny * nz;
    double *T = (double *)malloc(sizeof(double) * 9 * (n > 0 ? n : 1));
    if (!T) return -1;
    for (int t = 0; t < n; t++)
        for (int a = 0; a < 3; a++) {
            const double p0 = v0[(long)a * n + t], p1 = v1[(long)a * n + t], p2 = v2[(long)a * n + t];
            T[9 * t + a] = p0;
            T[9 * t + 3 + a] = p1 - p0;
            T[9 * t + 6 + a] = p2 - p0;
        }
    int miss = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : miss)
    for (long ic = 0; ic < N; ic++) {
        if (neiSolid[ic] != 1 || solid[ic] != 0) continue;
        const int i = (int)(ic % ((long)nx * ny) % nx), j = (int)(ic % ((long)nx * ny) / nx),
                  k = (int)(ic / ((long)nx * ny));
        const double x[3] = {(double)i, (double)j, (double)k};
        for (int d = 1; d < Q; d++) {
            int si = i - cx[d], sj = j - cy[d], sk = k - cz[d];
            si = si < 0 ? nx - 1 : (si >= nx ? 0 : si);
            sj = sj < 0 ? ny - 1 : (sj >= ny ? 0 : sj);
            sk = sk < 0 ? nz - 1 : (sk >= nz ? 0 : sk);
            if (solid[orc_index(si, sj, sk, nx, ny)] != 1) continue;
            const double c[3] = {(double)cx[d], (double)cy[d], (double)cz[d]};
            double best = 2.0;
            for (int t = 0; t < n; t++) {
                const double h = link_hit(&T[9 * t], &T[9 * t + 3], &T[9 * t + 6], x, c);
                best = h < best ? h : best;
            }
            if (best <= 1.0)
                qlink[(long)d * N + ic] = (float)best;
            else
                miss++;
        }
    }
    free(T);
    if (misses) *misses = miss;
    return 0;
}
