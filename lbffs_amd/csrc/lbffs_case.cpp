// lbffs_case.cpp -- host case setup of the LBFFS driver (include/lbffs_case.h).
//
// Builds the per-cell arrays the step kernel consumes, with the same float
// arithmetic as the reference host code so that codes, flags and initial
// populations are bit-identical.  Compiled with g++ -ffp-contract=off (no FMA
// contraction) so the host rounding matches a plain x86-64 build of ns.cpp.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/lbffs_case.h"

namespace {

// D3Q19 / D3Q27 direction tables (D3Q19.hpp:9-27, D3Q27.hpp)
const int CX27[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1};
const int CY27[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
const int CZ27[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, -1, 1};

float weight(int Q, int q)
{
    // the reference fills std::vector<float> from double quotients (D3Q19.hpp:9-12)
    if (Q == 19) return (float)(q == 0 ? 1.0 / 3.0 : (q <= 6 ? 1.0 / 18.0 : 1.0 / 36.0));
    return (float)(q == 0 ? 8.0 / 27.0 : (q <= 6 ? 2.0 / 27.0 : (q <= 18 ? 1.0 / 54.0 : 1.0 / 216.0)));
}

inline long idx(int i, int j, int k, int nx, int ny) { return (long)nx * ny * k + (long)nx * j + i; }

}  // namespace

extern "C" {

int lbffs_case_units(const lbffs_case_in *in, lbffs_units *o)
{
    if (!in || !o || in->nx < 2) return -1;
    const float L = in->Lx / float(in->nx - 1);
    const float c = in->uMax / in->U0;
    o->L = L;
    o->c = c;
    o->Re = in->uMax * in->Lx / in->nu;
    o->nu = in->nu / (L * c);
    o->dpdx = in->dpdx * (L / (c * c));
    o->uIni = in->uIni / c;
    o->vIni = in->vIni / c;
    o->wIni = in->wIni / c;
    o->deltaT = L / c;
    o->omega0 = (float)(1.0 / (3.0 * (double)o->nu + 0.5));
    for (int f = 0; f < 6; f++)
        for (int d = 0; d < 3; d++) o->face_vel[f][d] = in->face_vel[f][d] / c;
    return 0;
}

int lbffs_case_faces(const lbffs_case_in *in, const lbffs_units *un, int32_t *boundary, int32_t *b1, int32_t *b2,
                     int32_t *b3, float *u0, float *v0, float *w0)
{
    if (!in || !un || !boundary || !b1 || !b2 || !b3 || !u0 || !v0 || !w0) return -1;
    const int nx = in->nx, ny = in->ny, nz = in->nz;
    const long N = (long)nx * ny * nz;
    for (long ic = 0; ic < N; ic++) {
        boundary[ic] = b1[ic] = b2[ic] = b3[ic] = 0;
        u0[ic] = v0[ic] = w0[ic] = 0.f;
    }
    // wall velocity of a face in lattice units; uvw index 0/1/2
    auto V = [&](int face, int d) { return un->face_vel[face][d]; };
    // setBoundaryConditions.hpp:56-61 style edge rule: zero unless both faces move
    auto edge = [&](int fa, int fb, int d, bool gate) -> float {
        return (gate && V(fa, d) * V(fb, d) == 0.f) ? 0.f : V(fa, d);
    };
    auto edge3 = [&](int fa, int fb, int fc, int d) -> float {
        return (nz - 1 != 0 && V(fa, d) * V(fb, d) * V(fc, d) == 0.f) ? 0.f : V(fa, d);
    };
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const long ic = idx(i, j, k, nx, ny);
                auto set = [&](float a, float b, float c) {
                    u0[ic] = a;
                    v0[ic] = b;
                    w0[ic] = c;
                };
                if (i == 0) {
                    b1[ic] = in->face_code[LBFFS_NXMIN][0];
                    set(V(0, 0), V(0, 1), V(0, 2));
                } else if (i == nx - 1) {
                    b1[ic] = in->face_code[LBFFS_NXMAX][0];
                    set(V(1, 0), V(1, 1), V(1, 2));
                }
                if (j == 0) {
                    b2[ic] = in->face_code[LBFFS_NYMIN][1];
                    set(V(2, 0), V(2, 1), V(2, 2));
                } else if (j == ny - 1) {
                    b2[ic] = in->face_code[LBFFS_NYMAX][1];
                    set(V(3, 0), V(3, 1), V(3, 2));
                }
                if (k == 0) {
                    b3[ic] = in->face_code[LBFFS_NZMIN][2];
                    set(V(4, 0), V(4, 1), V(4, 2));
                } else if (k == nz - 1) {
                    b3[ic] = in->face_code[LBFFS_NZMAX][2];
                    set(V(5, 0), V(5, 1), V(5, 2));
                }
                // edges and corners: the reference tests each with an independent if, in
                // this order (setBoundaryConditions.hpp:56-183); x-y edges are not gated by nz
                const int xs[2] = {0, 1}, ys[2] = {2, 3}, zs[2] = {4, 5};
                const int ivals[2] = {0, nx - 1}, jvals[2] = {0, ny - 1}, kvals[2] = {0, nz - 1};
                // x-y: order (0,0), (0,ny-1), (nx-1,0), (nx-1,ny-1)
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                        if (i == ivals[a] && j == jvals[b])
                            set(edge(xs[a], ys[b], 0, true), edge(xs[a], ys[b], 1, true), edge(xs[a], ys[b], 2, true));
                // x-z: order (0,0), (0,nz-1), (nx-1,0), (nx-1,nz-1), gated by nz-1 != 0
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                        if (i == ivals[a] && k == kvals[b])
                            set(edge(xs[a], zs[b], 0, nz - 1 != 0), edge(xs[a], zs[b], 1, nz - 1 != 0),
                                edge(xs[a], zs[b], 2, nz - 1 != 0));
                // y-z
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                        if (j == jvals[a] && k == kvals[b])
                            set(edge(ys[a], zs[b], 0, nz - 1 != 0), edge(ys[a], zs[b], 1, nz - 1 != 0),
                                edge(ys[a], zs[b], 2, nz - 1 != 0));
                // corners (setBoundaryConditions.hpp:136-183), reference order
                const int corners[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                           {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
                for (int cidx = 0; cidx < 8; cidx++) {
                    const int a = corners[cidx][0], b = corners[cidx][1], cc = corners[cidx][2];
                    if (i == ivals[a] && j == jvals[b] && k == kvals[cc])
                        set(edge3(xs[a], ys[b], zs[cc], 0), edge3(xs[a], ys[b], zs[cc], 1),
                            edge3(xs[a], ys[b], zs[cc], 2));
                }
            }
    // ns.cpp:164-177: every cell on a lattice face; z faces only for 3-D lattices
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++)
                if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || (nz - 1 != 0 && (k == 0 || k == nz - 1)))
                    boundary[idx(i, j, k, nx, ny)] = 1;
    return 0;
}

int lbffs_case_init_feq(int Q, int nx, int ny, int nz, float rho, float u, float v, float w, float *f)
{
    if ((Q != 19 && Q != 27) || !f) return -1;
    const long N = (long)nx * ny * nz;
    for (int q = 0; q < Q; q++) {
        const float cx = (float)CX27[q], cy = (float)CY27[q], cz = (float)CZ27[q];
        const float uSqr = u * u + v * v + w * w;
        const float uDotC = u * cx + v * cy + w * cz;
        const float feq = (rho + 3.0f * uDotC + 4.5f * uDotC * uDotC - 1.5f * uSqr) * weight(Q, q);
        for (long ic = 0; ic < N; ic++) f[(long)q * N + ic] = feq;
    }
    return 0;
}

int lbffs_case_neisolid(int Q, int nx, int ny, int nz, const uint8_t *solid, uint8_t *nei)
{
    if ((Q != 19 && Q != 27) || !solid || !nei) return -1;
    const long N = (long)nx * ny * nz;
    for (long ic = 0; ic < N; ic++) nei[ic] = 0;
    for (long ic = 0; ic < N; ic++) {
        const int i = (int)(ic % ((long)nx * ny) % nx);
        const int j = (int)(ic % ((long)nx * ny) / nx);
        const int k = (int)(ic / ((long)nx * ny));
        const bool in_i = (i != 0 && i != nx - 1), in_j = (j != 0 && j != ny - 1), in_k = (k != 0 && k != nz - 1);
        if (!((nx == 1 && in_j && in_k) || (ny == 1 && in_i && in_k) || (nz == 1 && in_i && in_j) ||
              (in_i && in_j && in_k)))
            continue;
        for (int q = 0; q < Q; q++) {
            // upwind neighbour with periodic wrap (D3Q19.hpp:360-478)
            int si = i - CX27[q], sj = j - CY27[q], sk = k - CZ27[q];
            si = (si < 0) ? nx - 1 : (si >= nx ? 0 : si);
            sj = (sj < 0) ? ny - 1 : (sj >= ny ? 0 : sj);
            sk = (sk < 0) ? nz - 1 : (sk >= nz ? 0 : sk);
            if (solid[ic] == 0 && solid[idx(si, sj, sk, nx, ny)] == 1) nei[ic] = 1;
        }
    }
    return 0;
}

int lbffs_case_outlet_sponge(int nx, int ny, int nz, int i, int j, int k, int p1, int p2, int p3, int spz,
                             float omega0, float *omega)
{
    if (!omega || spz <= 0) return -1;
    // the patch cell must lie in the lattice: only the swept coordinate is clamped below
    if (nx < 1 || ny < 1 || nz < 1 || i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) return -1;
    const float tau0 = 0.55f;
    if (p1 != 0) {
        for (int s = (int)std::fmax(i - spz, 0); s < std::fmin(i + spz, nx); s++) {
            float fx = (i - spz < 0) ? float(s) / spz : float((nx - 1) - s) / spz;
            fx = 1.f - cos(0.5f * M_PI * fx);
            const float tau = fx * (1.f / omega0) + (1.f - fx) * tau0;
            omega[idx(s, j, k, nx, ny)] = 1.f / tau;
        }
    }
    if (p2 != 0) {
        for (int s = (int)std::fmax(j - spz, 0); s < std::fmin(j + spz, ny); s++) {
            float fx = (j - spz < 0) ? float(s) / spz : float((ny - 1) - s) / spz;
            fx = 1.f - cos(0.5f * M_PI * fx);
            const float tau = fx * (1.f / omega0) + (1.f - fx) * tau0;
            omega[idx(i, s, k, nx, ny)] = 1.f / tau;
        }
    }
    if (p3 != 0) {
        for (int s = (int)std::fmax(k - spz, 0); s < std::fmin(k + spz, nz); s++) {
            float fx = (k - spz < 0) ? float(s) / spz : float((nz - 1) - s) / spz;
            fx = 1.f - cos(0.5f * M_PI * fx);
            const float tau = fx * (1.f / omega0) + (1.f - fx) * tau0;
            omega[idx(i, j, s, nx, ny)] = 1.f / tau;
        }
    }
    return 0;
}

int lbffs_case_sdf_near_bwalls(float dr, const int32_t *b1, const int32_t *b2, const int32_t *b3, int nx, int ny, int nz,
                               float *sdf)
{
    if (!b1 || !b2 || !b3 || !sdf || nx < 1 || ny < 1 || nz < 1) return -1;
    const int n = int(dr) + 1;  // layers 0 .. int(dr) from each bounce-back face
    // x faces: the code on the face cell decides; the last layer before the opposite
    // face is left alone (i < nx-1, nx-1-i > 0)
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
            if (b1[idx(0, j, k, nx, ny)] == 1)
                for (int i = 0; i < n; i++)
                    if (i < nx - 1) sdf[idx(i, j, k, nx, ny)] = 0.5f + i;
            if (b1[idx(nx - 1, j, k, nx, ny)] == 1)
                for (int i = 0; i < n; i++)
                    if (nx - 1 - i > 0) sdf[idx(nx - 1 - i, j, k, nx, ny)] = 0.5f + i;
        }
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) {
            if (b2[idx(i, 0, k, nx, ny)] == 1)
                for (int j = 0; j < n; j++)
                    if (j < ny - 1) sdf[idx(i, j, k, nx, ny)] = 0.5f + j;
            if (b2[idx(i, ny - 1, k, nx, ny)] == 1)
                for (int j = 0; j < n; j++)
                    if (ny - 1 - j > 0) sdf[idx(i, ny - 1 - j, k, nx, ny)] = 0.5f + j;
        }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            if (b3[idx(i, j, 0, nx, ny)] == 1)
                for (int k = 0; k < n; k++)
                    if (k < nz - 1) sdf[idx(i, j, k, nx, ny)] = 0.5f + k;
            if (b3[idx(i, j, nz - 1, nx, ny)] == 1)
                for (int k = 0; k < n; k++)
                    if (nz - 1 - k > 0) sdf[idx(i, j, nz - 1 - k, nx, ny)] = 0.5f + k;
        }
    return 0;
}

}  // extern "C"

namespace {

// stl.hpp:95-111 pointInElement, in FP32 as written
bool point_in_element(float x, float y, float x1, float y1, float x2, float y2)
{
    const float s = (x * y2 - x2 * y) / (x1 * y2 - x2 * y1);
    const float t = (x * y1 - x1 * y) / (x2 * y1 - x1 * y2);
    return (s + t >= 0.f && s + t <= 1.f) && (s >= 0.f && s <= 1.f) && (t >= 0.f && t <= 1.f);
}

// the reference's abs() of a float component: int abs(int) of the truncated value
float int_abs(float v) { return (float)std::abs((int)v); }

}  // namespace

extern "C" int lbffs_case_find_points(int n, const float *v0, const float *v1, const float *v2, const float *normal,
                                      const float *centroid, int nx, int ny, int nz, int32_t *points, long cap,
                                      long *count)
{
    if (n < 0 || nx < 1 || ny < 1 || nz < 1 || !count || cap < 0 || (cap > 0 && !points)) return -1;
    if (n > 0 && (!v0 || !v1 || !v2 || !normal || !centroid)) return -1;
    long m = 0;
    auto push = [&](long ic) {
        if (m < cap) points[m] = (int32_t)ic;
        m++;
    };
    const int dims[3] = {nx, ny, nz};
    for (int t = 0; t < n; t++) {
        const float a0 = int_abs(normal[t]), a1 = int_abs(normal[n + t]), a2 = int_abs(normal[2 * n + t]);
        const float mx = std::fmax(a0, std::fmax(a1, a2));
        const int ax = (a0 == mx) ? 0 : ((a1 == mx) ? 1 : 2);
        // the two in-plane axes, in the reference's loop order
        const int u = ax == 0 ? 1 : 0, w = ax == 2 ? 1 : 2;
        const int ci = (int)centroid[ax * n + t];
        int face;
        if (ci == 0 || ci + 1 == 0)
            face = 0;
        else if (ci == dims[ax] - 1 || ci + 1 == dims[ax] - 1)
            face = dims[ax] - 1;
        else
            continue;
        const float x1 = v1[u * n + t] - v0[u * n + t], y1 = v1[w * n + t] - v0[w * n + t];
        const float x2 = v2[u * n + t] - v0[u * n + t], y2 = v2[w * n + t] - v0[w * n + t];
        for (int a = 0; a < dims[u]; a++)
            for (int b = 0; b < dims[w]; b++) {
                const float x = (float)a - v0[u * n + t], y = (float)b - v0[w * n + t];
                if (!point_in_element(x, y, x1, y1, x2, y2)) continue;
                int idx[3];
                idx[ax] = face;
                idx[u] = a;
                idx[w] = b;
                push((long)idx[0] + (long)nx * ((long)idx[1] + (long)ny * idx[2]));
            }
    }
    *count = m;
    return 0;
}
