// geom.hpp -- geometry preprocessing on the device: the solid cells enclosed by a triangle
// surface (setSolid with isInner/isCross/det, walls.hpp:18-184) and the signed distance
// near the surface as an inverse-distance-weighted average of the triangles' plane
// distances (setSDF, patch form, walls.hpp:218-297).
//
// Solid: a cell is inside when the ray from its centre along +x crosses the surface an
// odd number of times.  Each crossing test solves v0 + x a + y b = P + z e_x by Cramer's
// rule in FP32, in the reference's order of operations and without contraction, so the
// flags are bit-identical to the reference host, including its handling of rays that
// graze an edge (both triangles count) or lie in a triangle's plane (0/0 = NaN passes
// every range test and counts as a crossing).  Work: one work-item per cell, a work-group
// per 256 cells of one x row; the triangles are binned by the z layers their box covers,
// so a row reads only the triangles of its layer (uniform across the work-group: scalar
// loads, no divergence in the loop).
//
// SDF: cells within int(dr) cells (box, clamped) of a neiSolid cell get
//   sdf = sum_t w_t sd_t / sum_t w_t,  w_t = r_t^-p,  over triangles with centroid distance r_t < dr,
// sd_t the distance to the triangle's plane, summed in triangle order in FP32; sdfIni where
// no triangle is that near.  r^-p is evaluated in FP64 and rounded once (correctly rounded
// in practice); the reference calls the host's powf, so the two agree to the last bit
// except where that powf rounds the other way.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d3q27_step.hpp"

namespace lbffs {

// one triangle as the crossing test reads it (vertices, then the y/z extent of its box)
struct GeomTri {
    float v0[3], v1[3], v2[3];
    float ymin, ymax, zmin, zmax;
    float pad;
};

// walls.hpp:18-22, summed left to right as written
__device__ inline float geom_det(const float *a, const float *b, const float *c)
{
#pragma clang fp contract(off)
    return (a[0] * b[1] * c[2]) + (a[1] * b[2] * c[0]) + (a[2] * b[0] * c[1]) - (a[0] * b[2] * c[1])
           - (a[1] * b[0] * c[2]) - (a[2] * b[1] * c[0]);
}

// walls.hpp:24-62: does the +x ray from (fi, fj, fk) cross the triangle?
__device__ inline bool geom_is_cross(const GeomTri &t, float fi, float fj, float fk)
{
#pragma clang fp contract(off)
    const float a[3] = {t.v1[0] - t.v0[0], t.v1[1] - t.v0[1], t.v1[2] - t.v0[2]};
    const float b[3] = {t.v2[0] - t.v0[0], t.v2[1] - t.v0[1], t.v2[2] - t.v0[2]};
    const float c[3] = {-1.f, 0.f, 0.f};
    const float d[3] = {fi - t.v0[0], fj - t.v0[1], fk - t.v0[2]};
    const float den = geom_det(a, b, c);
    const float x = geom_det(d, b, c) / den;
    const float y = geom_det(a, d, c) / den;
    const float z = geom_det(a, b, d) / den;
    const float xy = x + y;
    // every test false for NaN: such a triangle counts, as in the reference
    return !(x < 0.f || x > 1.f || y < 0.f || y > 1.f || xy < 0.f || xy > 1.f || z < 0.f);
}

struct GeomBox {
    float xmin, xmax, ymin, ymax, zmin, zmax;
};

constexpr int kGeomRow = 256;

// walls.hpp:64-157 isInner for every cell; solid = 1 where inside, other cells unchanged.
// Work-group b covers cells [c * 256, c * 256 + 256) of row (j, k), b = c + chunks * (j + ny k).
__global__ void __launch_bounds__(kGeomRow)
    geom_solid_kernel(const GeomTri *__restrict__ tris, const int *__restrict__ bin_start,
                      const int *__restrict__ bin_tri, int nx, int ny, int chunks, GeomBox box, uint8_t *solid)
{
    const uint32_t row = blockIdx.x / (uint32_t)chunks;
    const int i = (int)((blockIdx.x - row * (uint32_t)chunks) * kGeomRow + threadIdx.x);
    const int j = (int)(row % (uint32_t)ny), k = (int)(row / (uint32_t)ny);
    if (i >= nx) return;
    const float fi = float(i), fj = float(j), fk = float(k);
    if (fi < box.xmin || fi > box.xmax || fj < box.ymin || fj > box.ymax || fk < box.zmin || fk > box.zmax) return;
    int ncross = 0;
    const int b1 = bin_start[k + 1];
    for (int b = bin_start[k]; b < b1; b++) {
        const GeomTri &t = tris[bin_tri[b]];
        if (fj >= t.ymin && fj <= t.ymax && fk >= t.zmin && fk <= t.zmax) ncross += geom_is_cross(t, fi, fj, fk) ? 1 : 0;
    }
    if (ncross & 1) solid[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)] = 1;
}

// walls.hpp:176-183: invertFluidSolid maps 1 -> 0 and every other value -> 1
__global__ void geom_invert_kernel(uint8_t *solid, size_t n)
{
    for (size_t ic = blockIdx.x * (size_t)blockDim.x + threadIdx.x; ic < n; ic += (size_t)gridDim.x * blockDim.x)
        solid[ic] = solid[ic] == 1 ? 0 : 1;
}

// walls.hpp:223-253: the cells within the (2 drn + 1)^3 box (clamped to the lattice)
// around some neiSolid cell, as three separable passes of a 1-D dilation along x, y, z.
// Pass along `axis` (0 x, 1 y, 2 z): dst = 1 where a source cell within drn along the axis
// is set (the first pass reads neiSolid == 1, the others != 0).
__global__ void geom_dilate_kernel(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, int nx, int ny, int nz,
                                   int axis, int drn, int first)
{
    const size_t n = (size_t)nx * ny * nz;
    const int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)nx : (size_t)nx * ny);
    for (size_t ic = blockIdx.x * (size_t)blockDim.x + threadIdx.x; ic < n; ic += (size_t)gridDim.x * blockDim.x) {
        const int pos = (int)(ic / stride % (size_t)len);
        const int lo = max(pos - drn, 0), hi = min(pos + drn, len - 1);
        const uint8_t *line = src + ic - (size_t)pos * stride;
        uint8_t hit = 0;
        for (int q = lo; q <= hi && !hit; q++) {
            const uint8_t v = line[(size_t)q * stride];
            hit = first ? (v == 1) : (v != 0);
        }
        dst[ic] = hit;
    }
}

// compact list of the marked cells (any order: every cell's value is independent)
__global__ void geom_near_list_kernel(const uint8_t *__restrict__ near_mark, size_t n, uint32_t *list, uint32_t *count)
{
    for (size_t ic = blockIdx.x * (size_t)blockDim.x + threadIdx.x; ic < n; ic += (size_t)gridDim.x * blockDim.x)
        if (near_mark[ic]) list[atomicAdd(count, 1u)] = (uint32_t)ic;
}

// walls.hpp:255-296 for the listed cells.  tri holds the centroids (rows cx, cy, cz) then
// the normals (rows nx, ny, nz), n entries each.
__global__ void __launch_bounds__(256)
    geom_sdf_kernel(const uint32_t *__restrict__ list, uint32_t count, const float *__restrict__ tri, int n, int nx,
                    int ny, float sdf_ini, float dr, float p, float *sdf)
{
#pragma clang fp contract(off)
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const uint32_t ic = list[idx];
    const float fi = float(ic % (uint32_t)nx), fj = float(ic / (uint32_t)nx % (uint32_t)ny);
    const float fk = float(ic / ((uint32_t)nx * (uint32_t)ny));
    const double mp = -(double)p;
    float s = 0.f, sumD = 0.f;
    for (int t = 0; t < n; t++) {
        const float dx = fi - tri[t], dy = fj - tri[n + t], dz = fk - tri[2 * n + t];
        // pow(x, 2.f) of the reference is x * x (exact square, rounded once)
        const float r = sqrtf(dx * dx + dy * dy + dz * dz);
        if (r < dr) {
            const float sd = dx * tri[3 * n + t] + dy * tri[4 * n + t] + dz * tri[5 * n + t];
            const float w = (float)pow((double)r, mp);
            sumD += w;
            s += sd * w;
        }
    }
    sdf[ic] = sumD != 0.f ? s / sumD : sdf_ini;
}

// ---- q-distances of the solid links from the surface, in FP64 ----
//
// For a fluid cell x with neiSolid == 1 and a direction c_d whose upwind cell x - c_d is
// solid, q is the smallest t in [0, 1] at which the segment x - t c_d meets a triangle
// (Moller-Trumbore in FP64, contraction off: the result is a function of the triangle and
// the link alone, so the order in which candidates are visited does not matter).  The
// triangles come binned on the host into cubic buckets of kQBucket cells: a triangle is
// listed in every bucket its box, grown by one cell, overlaps, so every triangle that can
// meet a link of x is listed in the bucket of x.

constexpr int kQBucket = 4;

struct GeomTri64 {
    double v0[3], e1[3], e2[3];
};

// smallest t in [0, 1] where x - t c meets the triangle, or 2 (no hit)
__device__ inline double geom_link_hit(const GeomTri64 &T, const double *x, const double *c)
{
#pragma clang fp contract(off)
    const double dir[3] = {-c[0], -c[1], -c[2]};
    const double p[3] = {dir[1] * T.e2[2] - dir[2] * T.e2[1], dir[2] * T.e2[0] - dir[0] * T.e2[2],
                         dir[0] * T.e2[1] - dir[1] * T.e2[0]};
    const double det = T.e1[0] * p[0] + T.e1[1] * p[1] + T.e1[2] * p[2];
    if (fabs(det) < 1e-14) return 2.0;  // link parallel to the triangle's plane
    const double inv = 1.0 / det;
    const double sv[3] = {x[0] - T.v0[0], x[1] - T.v0[1], x[2] - T.v0[2]};
    const double u = (sv[0] * p[0] + sv[1] * p[1] + sv[2] * p[2]) * inv;
    if (u < -1e-12 || u > 1.0 + 1e-12) return 2.0;
    const double qv[3] = {sv[1] * T.e1[2] - sv[2] * T.e1[1], sv[2] * T.e1[0] - sv[0] * T.e1[2],
                          sv[0] * T.e1[1] - sv[1] * T.e1[0]};
    const double v = (dir[0] * qv[0] + dir[1] * qv[1] + dir[2] * qv[2]) * inv;
    if (v < -1e-12 || u + v > 1.0 + 1e-12) return 2.0;
    const double t = (T.e2[0] * qv[0] + T.e2[1] * qv[1] + T.e2[2] * qv[2]) * inv;
    return (t >= 0.0 && t <= 1.0) ? t : 2.0;
}

// one work-item per listed neiSolid cell; Q directions; q written for the solid links
// found (no hit keeps the caller's value); misses counted
template <int Q, typename Out>
__global__ void __launch_bounds__(256)
    geom_qlinks_kernel(const uint32_t *__restrict__ cells, uint32_t ncells, const GeomTri64 *__restrict__ tris,
                       const int *__restrict__ bstart, const int *__restrict__ btri, int bx, int by, int bz, int nx,
                       int ny, int nz, const uint8_t *__restrict__ solid, Out *qlink, uint32_t *misses)
{
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ncells) return;
    const uint32_t ic = cells[idx];
    const int i = (int)(ic % (uint32_t)nx), j = (int)(ic / (uint32_t)nx % (uint32_t)ny);
    const int k = (int)(ic / ((uint32_t)nx * (uint32_t)ny));
    const double x[3] = {(double)i, (double)j, (double)k};
    const int bi = i / kQBucket, bj = j / kQBucket, bk = k / kQBucket;
    const int b = bi + bx * (bj + by * bk);
    const int t0 = bstart[b], t1 = bstart[b + 1];
    const size_t N = (size_t)nx * ny * nz;
    for (int d = 1; d < Q; d++) {
        const int cx = Q == 27 ? L27::cx[d] : L19::cx[d];
        const int cy = Q == 27 ? L27::cy[d] : L19::cy[d];
        const int cz = Q == 27 ? L27::cz[d] : L19::cz[d];
        // upwind cell with periodic wrap, as setNeiSolid (walls.hpp:187-216)
        int si = i - cx, sj = j - cy, sk = k - cz;
        si = si < 0 ? nx - 1 : (si >= nx ? 0 : si);
        sj = sj < 0 ? ny - 1 : (sj >= ny ? 0 : sj);
        sk = sk < 0 ? nz - 1 : (sk >= nz ? 0 : sk);
        if (solid[(size_t)si + (size_t)nx * ((size_t)sj + (size_t)ny * sk)] != 1) continue;
        const double c[3] = {(double)cx, (double)cy, (double)cz};
        double best = 2.0;
        for (int t = t0; t < t1; t++) best = fmin(best, geom_link_hit(tris[btri[t]], x, c));
        if (best <= 1.0)
            qlink[(size_t)d * N + ic] = (Out)(float)best;  // q rounded to FP32 in either output type
        else
            atomicAdd(misses, 1u);
    }
}

}  // namespace lbffs
