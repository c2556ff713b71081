// ibm.hpp -- immersed-boundary moving walls: marker forcing, spreading, marker motion
// and the forcing of the populations (k_Gwall, k_Gibm, k_Force of
// D3Q19_opencl/kernel.cl:3648-3807 and D3Q27_opencl/kernel.cl:2828-2986, with the helpers
// Urot/Xrot/icBox of K19:1112-1279).
//
// Markers are points in lattice coordinates (xyz interleaved, as movingSTLcList).  After
// each step:
//   gwall:  G_m = U_wall(x_m) - sum_box u(x_b) delta(x_b - x_m)   (one work-item per marker)
//   spread: G(x_b) += G_m delta(x_b - x_m) over the 8-cell box, with hardware float atomics,
//           and each touched cell enters a compact list once per step (stamp per cell);
//           then the marker moves by the translation and turns by rotOmega about the axis
//   force:  f_q += rho w_q 3 (G . c_q) on the listed fluid cells only -- the reference
//           visits every cell, but G is zero away from the boxes, so the result is the same.
// The spread sums every contribution.  The reference writes the value atom_add_float
// returns (the value before the add) back into the cell (K19:3760-3762), which undoes the
// add unless another work-item interleaves: that result depends on scheduling and is not
// reproduced.  Urot/Xrot keep the reference's axis dot product OX_x*e0_z (K19:1213, 1256).
#pragma once

#include "d3q27_step.hpp"

namespace lbffs {

template <typename Real>
struct IbmMotion {
    Real trans[3];   // uMovingTrans, vMovingTrans, wMovingTrans (lattice units per step)
    Real centre[3];  // rotX, rotY, rotZ
    Real axis[3];    // rotAxisX, rotAxisY, rotAxisZ
    Real omega;      // rotOmega (radians per step)
};

// the frame of the rotation at a point (K19:1211-1233): the point's offset from the axis
// foot OOp, the radial unit vector e1, the tangential e2 = e0 x e1, and the radius
template <typename Real>
struct RotFrame {
    Real OOp[3], e1[3], e2[3], r;
};

template <typename Real>
__device__ inline RotFrame<Real> rot_frame(const IbmMotion<Real> &m, Real x, Real y, Real z)
{
    RotFrame<Real> F;
    const Real OX[3] = {x - m.centre[0], y - m.centre[1], z - m.centre[2]};
    const Real *e0 = m.axis;
    const Real OXdotE0 = OX[0] * e0[0] + OX[1] * e0[1] + OX[0] * e0[2];  // sic (K19:1213)
    Real OpX[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        F.OOp[a] = OXdotE0 * e0[a];
        OpX[a] = OX[a] - F.OOp[a];
    }
    const Real mag = sqrt(OpX[0] * OpX[0] + OpX[1] * OpX[1] + OpX[2] * OpX[2]);
#pragma unroll
    for (int a = 0; a < 3; a++) F.e1[a] = OpX[a] / mag;
    F.e2[0] = e0[1] * F.e1[2] - e0[2] * F.e1[1];
    F.e2[1] = e0[2] * F.e1[0] - e0[0] * F.e1[2];
    F.e2[2] = e0[0] * F.e1[1] - e0[1] * F.e1[0];
    F.r = sqrt(OpX[0] * OpX[0] + OpX[1] * OpX[1] + OpX[2] * OpX[2]);
    return F;
}

// trilinear weight of box corner b (0..7: +x, +y, +z bits as icBox, K19:1112-1155) of the
// cell (i, j, k) below the marker; -1 index when the corner lies outside (no wrap)
template <typename Real>
__device__ inline long box_corner(int b, int i, int j, int k, int nx, int ny, int nz, Real x, Real y, Real z, Real &d)
{
    const int di = (b == 1 || b == 4 || b == 5 || b == 7) ? 1 : 0;
    const int dj = (b == 2 || b == 4 || b == 6 || b == 7) ? 1 : 0;
    const int dk = (b == 3 || b == 5 || b == 6 || b == 7) ? 1 : 0;
    const int bi = i + di, bj = j + dj, bk = k + dk;
    if (bi >= nx || bj >= ny || bk >= nz) return -1;
    d = (Real(1) - fabs(Real(bi) - x)) * (Real(1) - fabs(Real(bj) - y)) * (Real(1) - fabs(Real(bk) - z));
    return (long)bi + (long)nx * ((long)bj + (long)ny * bk);
}

// k_Gwall (K19:3648-3716).  Thread 0 also clears the touched-cell count of this step.
template <typename Real>
__global__ void ibm_gwall_kernel(int n, const Real *xyz, int nx, int ny, int nz, const Real *u, const Real *v,
                                 const Real *w, IbmMotion<Real> m, Real *Gm, int *count)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) *count = 0;
    if (p >= n) return;
    const Real x = xyz[3 * p], y = xyz[3 * p + 1], z = xyz[3 * p + 2];
    const RotFrame<Real> F = rot_frame(m, x, y, z);
    Real U[3];
#pragma unroll
    for (int a = 0; a < 3; a++) U[a] = m.trans[a] + F.r * m.omega * F.e2[a];
    Real ui = 0, vi = 0, wi = 0;
    const int i = (int)x, j = (int)y, k = (int)z;  // truncation (K19:3671-3673)
    if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) {
#pragma unroll
        for (int b = 0; b < 8; b++) {
            Real d;
            const long ib = box_corner(b, i, j, k, nx, ny, nz, x, y, z, d);
            if (ib < 0) continue;
            ui += u[ib] * d;
            vi += v[ib] * d;
            wi += w[ib] * d;
        }
    }
    Gm[p] = U[0] - ui;
    Gm[n + p] = U[1] - vi;
    Gm[2 * n + p] = U[2] - wi;
}

// k_Gibm (K19:3718-3779) with a summing spread; cells touched this step are listed once
template <typename Real>
__global__ void ibm_spread_move_kernel(int n, Real *xyz, int nx, int ny, int nz, const Real *Gm, IbmMotion<Real> m,
                                       Real *Gx, Real *Gy, Real *Gz, uint32_t *mark, uint32_t stamp, int *list,
                                       int *count, int cap)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const Real x = xyz[3 * p], y = xyz[3 * p + 1], z = xyz[3 * p + 2];
    const int i = (int)x, j = (int)y, k = (int)z;
    if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) {
        const Real gx = Gm[p], gy = Gm[n + p], gz = Gm[2 * n + p];
#pragma unroll
        for (int b = 0; b < 8; b++) {
            Real d;
            const long ib = box_corner(b, i, j, k, nx, ny, nz, x, y, z, d);
            if (ib < 0) continue;
            atomicAdd(&Gx[ib], gx * d);
            atomicAdd(&Gy[ib], gy * d);
            atomicAdd(&Gz[ib], gz * d);
            if (atomicExch(&mark[ib], stamp) != stamp) {
                const int slot = atomicAdd(count, 1);
                if (slot < cap) list[slot] = (int)ib;
            }
        }
    }
    // translation, then the turn by rotOmega (Xrot, K19:1238-1279)
    const Real xt = x + m.trans[0], yt = y + m.trans[1], zt = z + m.trans[2];
    const RotFrame<Real> F = rot_frame(m, xt, yt, zt);
    const Real c = cos(m.omega), s = sin(m.omega);
    xyz[3 * p] = m.centre[0] + F.OOp[0] + F.r * (c * F.e1[0] + s * F.e2[0]);
    xyz[3 * p + 1] = m.centre[1] + F.OOp[1] + F.r * (c * F.e1[1] + s * F.e2[1]);
    xyz[3 * p + 2] = m.centre[2] + F.OOp[2] + F.r * (c * F.e1[2] + s * F.e2[2]);
}

// k_Force (K19:3781-3807) on the listed cells; zero_after clears G behind it so that the
// next step spreads into zeros
template <typename S, typename Real, int Q>
__global__ void ibm_force_kernel(const int *list, const int *count, int cap, size_t fs, S *f, const uint16_t *flag,
                                 const Real *rho, Real *Gx, Real *Gy, Real *Gz, int zero_after)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int cnt = min(*count, cap);
    if (idx >= cnt) return;
    const uint32_t ic = (uint32_t)list[idx];
    const Real gx = Gx[ic], gy = Gy[ic], gz = Gz[ic];
    if (zero_after) {
        Gx[ic] = Real(0);
        Gy[ic] = Real(0);
        Gz[ic] = Real(0);
    }
    if (flag[ic] & F_SOLID) return;
    const Real r3 = rho[ic] * Real(3);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int cx = Q == 27 ? L27::cx[q] : L19::cx[q];
        const int cy = Q == 27 ? L27::cy[q] : L19::cy[q];
        const int cz = Q == 27 ? L27::cz[q] : L19::cz[q];
        const Real wqq = wq<Real, Q>(q);
        S *p = f + pop_at<Q, S>(q, ic, fs);
        const Real fq = Codec<S, Real, Q>::load(p, q, wqq);
        Codec<S, Real, Q>::store(p, fq + r3 * wqq * (gx * Real(cx) + gy * Real(cy) + gz * Real(cz)), q, wqq);
    }
}

// G of the cells listed by the last spread back to zero (start of a call after a call
// whose last step kept them for output)
template <typename Real>
__global__ void ibm_zero_listed_kernel(const int *list, const int *count, int cap, Real *Gx, Real *Gy, Real *Gz)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= min(*count, cap)) return;
    const int ic = list[idx];
    Gx[ic] = Real(0);
    Gy[ic] = Real(0);
    Gz[ic] = Real(0);
}

template <typename Real>
__global__ void fill_kernel(Real *p, size_t n, Real v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace lbffs
