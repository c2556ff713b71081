// d3q27_step.hpp -- the fused D3Q27 time-step kernel with the parametrized cumulant
// collision.
//
// Same work decomposition and data layout as the D3Q19 kernel (one work-item per cell,
// q-major populations, interior pulls from SGPR bases).  Per fluid cell it does what
// k_streamingCollision of D3Q27_opencl/kernel.cl:2711-2826 does:
//   pull streaming (faces: wrap, or reflect/mirror through wall faces),
//   fixed-velocity / fixed-density / wall-function / outlet face conditions,
//   interpolated bounce-back on solid links (Filippova-Hanel, Mei-Luo-Shyy for q <= 1/2)
//   with momentum-exchange force, moments, WALE or Smagorinsky sub-grid model, sponge,
//   and the cumulant collision of Geier et al. (2017) with the velocity shifted by half
//   the body force.
// The collision runs on the central moments of the 27 populations obtained by three
// passes of 3-point transforms (one per axis, nine lines each) instead of 27x27 sums; the
// back-transform is the reference's three passes (x, then y, then z).
#pragma once

#include "d3q19_step.hpp"

namespace lbffs {

// D3Q27 directions (K27:2742-2747)
struct L27 {
    static constexpr int cx[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1};
    static constexpr int cy[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
    static constexpr int cz[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, -1, 1};
    // opposite direction: pairs (2m-1, 2m) as in D3Q19
    __host__ __device__ static constexpr int opp(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }
    // K27:186-528 mirrorQx/y/z: the component along the axis flipped, or the whole
    // direction reflected when it has none
    static constexpr int mirx[27] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 16, 15, 18, 17, 21, 22, 19, 20, 26, 25, 24, 23};
    static constexpr int miry[27] = {0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 12, 11, 14, 13, 18, 17, 16, 15, 23, 24, 26, 25, 19, 20, 22, 21};
    static constexpr int mirz[27] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 13, 14, 11, 12, 17, 18, 15, 16, 25, 26, 24, 23, 22, 21, 19, 20};
    // slot s = 9 (cx+1) + 3 (cy+1) + (cz+1) of each direction, and the direction of each slot
    static constexpr int slot[27] = {13, 22, 4, 16, 10, 14, 12, 25, 1, 19, 7, 23, 3, 21, 5, 17, 9, 15, 11, 26, 0, 8, 18, 20, 6, 24, 2};
    static constexpr int q_of_slot[27] = {20, 8, 26, 12, 2, 14, 24, 10, 21, 16, 4, 18, 6, 0, 5, 17, 3, 15, 22, 9, 23, 13, 1, 11, 25, 7, 19};
};

namespace detail {
constexpr bool l27_tables_consistent()
{
    for (int q = 0; q < 27; q++) {
        const int x = L27::cx[q], y = L27::cy[q], z = L27::cz[q];
        const int o = L27::opp(q);
        if (L27::cx[o] != -x || L27::cy[o] != -y || L27::cz[o] != -z) return false;
        const int mx = L27::mirx[q], my = L27::miry[q], mz = L27::mirz[q];
        if (x != 0 && (L27::cx[mx] != -x || L27::cy[mx] != y || L27::cz[mx] != z)) return false;
        if (y != 0 && (L27::cx[my] != x || L27::cy[my] != -y || L27::cz[my] != z)) return false;
        if (z != 0 && (L27::cx[mz] != x || L27::cy[mz] != y || L27::cz[mz] != -z)) return false;
        if (x == 0 && mx != o) return false;
        if (y == 0 && my != o) return false;
        if (z == 0 && mz != o) return false;
        const int s = 9 * (x + 1) + 3 * (y + 1) + (z + 1);
        if (L27::slot[q] != s || L27::q_of_slot[s] != q) return false;
    }
    return true;
}
static_assert(l27_tables_consistent(), "D3Q27 direction tables");
}  // namespace detail

// K27:2740 weights 8/27, 2/27, 1/54, 1/216
template <typename Real>
__host__ __device__ constexpr Real w27(int q)
{
    return q == 0 ? Real(8) / Real(27)
                  : (q <= 6 ? Real(2) / Real(27) : (q <= 18 ? Real(1) / Real(54) : Real(1) / Real(216)));
}

// lattice weight of direction q for either lattice
template <typename Real, int Q>
__host__ __device__ constexpr Real wq(int q)
{
    return Q == 27 ? w27<Real>(q) : w19<Real>(q);
}

template <int DIR>
__host__ __device__ constexpr int c27(int q)
{
    return DIR == 0 ? L27::cx[q] : (DIR == 1 ? L27::cy[q] : L27::cz[q]);
}

// K27:530-549 upwindID_B: pull source of a face cell, -1 through a wall face.  Formed from
// the interior source ic - c_q in 32-bit arithmetic as the D3Q19 face_upwind: a crossed
// periodic face moves it by a whole row, plane or lattice (unsigned wrap-around is exact,
// the result lying in the lattice).
__device__ inline long face_upwind27(int q, long ic, int i, int j, int k, int nx, int ny, int nz, int b1, int b2, int b3)
{
    const int cx = L27::cx[q], cy = L27::cy[q], cz = L27::cz[q];
    const bool crx = cx == 1 ? i == 0 : (cx == -1 ? i == nx - 1 : false);
    const bool cry = cy == 1 ? j == 0 : (cy == -1 ? j == ny - 1 : false);
    const bool crz = cz == 1 ? k == 0 : (cz == -1 ? k == nz - 1 : false);
    if ((crx && is_wall_code(b1)) || (cry && is_wall_code(b2)) || (crz && is_wall_code(b3))) return -1;
    const uint32_t nxny = (uint32_t)nx * (uint32_t)ny;
    uint32_t s = (uint32_t)ic - (uint32_t)cx - (uint32_t)cy * (uint32_t)nx - (uint32_t)cz * nxny;
    if (crx) s += (uint32_t)cx * (uint32_t)nx;
    if (cry) s += (uint32_t)cy * nxny;
    if (crz) s += (uint32_t)cz * nxny * (uint32_t)nz;
    return (long)s;
}

// Zou-He wall density numerator of a face: tangential populations plus twice those
// leaving through it
template <int DIR, typename Real>
__device__ inline Real zouhe27(const Real *ft, bool lo)
{
    Real t = Real(0), om = Real(0), op = Real(0);
#pragma unroll
    for (int q = 0; q < 27; q++) {
        const int c = c27<DIR>(q);
        if (c == 0) t += ft[q];
        else if (c < 0) om += ft[q];
        else op += ft[q];
    }
    return t + Real(2) * (lo ? om : op);
}

// Ladd moving-wall correction ft_q += 6 w_q rho_w (c_q . u_w) of the nine populations
// entering through a face.  The nzMax face of the reference divides its four
// corner-diagonal terms by 9 instead of 36 (K27:1376-1379, 1815-1818): zmax_quirk.
template <int DIR, typename Real>
__device__ inline void ladd27(Real *ft, bool lo, Real rhow, Real uw, Real vw, Real ww, bool zmax_quirk)
{
#pragma unroll
    for (int q = 1; q < 27; q++) {
        const int c = c27<DIR>(q);
        if (c == 0) continue;
        const Real cu = Real(L27::cx[q]) * uw + Real(L27::cy[q]) * vw + Real(L27::cz[q]) * ww;
        Real den = q <= 6 ? Real(9) / Real(4) : (q <= 18 ? Real(9) : Real(36));
        if (q > 18) den = zmax_quirk ? Real(9) : den;
        const Real d = rhow * cu / den;
        const bool enters = (c > 0) ? lo : !lo;
        ft[q] = enters ? ft[q] + d : ft[q];
    }
}

template <int DIR, typename S, typename Real>
__device__ inline void fixed_velocity27(const StepArgs<Real, 27> &a, Real *ft, long ic, int pos, long st, Real u0,
                                        Real v0, Real w0)
{
    const bool lo = (pos == 0);
    const Real un = DIR == 0 ? u0 : (DIR == 1 ? v0 : w0);
    Real rhow = zouhe27<DIR>(ft, lo) / (lo ? Real(1) - un : Real(1) + un);
    if (lo && un < Real(0)) rhow = a.rho_o[ic + st];
    // the x face takes the cell's own stored density (K27:1272-1276), y and z the inner cell
    if (!lo && un > Real(0)) rhow = (DIR == 0) ? a.rho_o[ic] : a.rho_o[ic - st];
    ladd27<DIR>(ft, lo, rhow, u0, v0, w0, DIR == 2 && !lo);
}

template <int DIR, typename Real>
__device__ inline void fixed_density27(const StepArgs<Real, 27> &a, Real *ft, long ic, int pos, int n, long st)
{
    const Real uc = a.u_o[ic], vc = a.v_o[ic], wc = a.w_o[ic];
    const long inner = (pos == 0) ? ic + st : ic + (long)(n - 2 - pos) * st;
    const Real ue = uc + Real(0.5) * (uc - a.u_o[inner]);
    const Real ve = vc + Real(0.5) * (vc - a.v_o[inner]);
    const Real we = wc + Real(0.5) * (wc - a.w_o[inner]);
    const Real uSqr = ue * ue + ve * ve + we * we;
    const bool atlo = (pos == 0), athi = (pos == n - 1);
#pragma unroll
    for (int q = 1; q < 27; q++) {
        const int c = c27<DIR>(q);
        if (c == 0) continue;
        const bool hit = (c > 0) ? atlo : athi;
        const Real uDotC = ue * Real(L27::cx[q]) + ve * Real(L27::cy[q]) + we * Real(L27::cz[q]);
        const Real nv = -ft[q] + Real(2) * w27<Real>(q) * Real(1) * (Real(1) + Real(4.5) * uDotC * uDotC - Real(1.5) * uSqr);
        ft[q] = hit ? nv : ft[q];
    }
}

template <int DIR, typename Real>
__device__ inline void wall_function27(const StepArgs<Real, 27> &a, Real *ft, long ic, int pos, int nz, long st,
                                       Real nu, Real Fx, Real Fy, Real Fz)
{
    const bool lo = (pos == 0);
    const Real rhow = zouhe27<DIR>(ft, lo);
    // neighbour sampled for the wall law: the next cell inward; the z faces of a 2-D
    // lattice sample the cell itself
    long nb = lo ? ic + st : ic - st;
    if (DIR == 2) nb = (nz - 1 != 0) ? ic + st : ic;
    const Real *t1 = DIR == 0 ? a.v_o : a.u_o;
    const Real *t2 = DIR == 2 ? a.v_o : a.w_o;
    Real p = t1[nb], s = t2[nb];
    Real magUp = sqrt(p * p + s * s);
    Real ut1 = 0, ut2 = 0;
    if (magUp != Real(0)) {
        const Real tauw = calc_tauw<Real>(rhow, magUp, nu);
        p = t1[ic];
        s = t2[ic];
        magUp = sqrt(p * p + s * s);
        if ((DIR == 2 && lo) || magUp != Real(0)) {  // the k == 0 branch has no zero guard
            const Real e1 = p / magUp, e2 = s / magUp;
            const Real Fw1 = DIR == 0 ? Fy : Fx;
            const Real Fw2 = DIR == 2 ? Fy : Fz;
            ut1 = Real(-3) * (tauw * e1 - Fw1) / rhow;
            ut2 = Real(-3) * (tauw * e2 - Fw2) / rhow;
        }
    }
    const Real uw = DIR == 0 ? Real(0) : ut1;
    const Real vw = DIR == 0 ? ut1 : (DIR == 1 ? Real(0) : ut2);
    const Real ww = DIR == 2 ? Real(0) : ut2;
    ladd27<DIR>(ft, lo, rhow, uw, vw, ww, DIR == 2 && !lo);
}

// Boundary handling of a domain-face cell (K27:2764-2777): streaming with walls, then
// the face conditions in the reference's order.
template <typename S, typename Real>
__device__ __forceinline__ void face_cell27(const StepArgs<Real, 27> &a, Real *ft, uint32_t fl, long ic, int i, int j,
                                            int k, Real omega)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const S *fs = (const S *)a.fsrc;
    const int b1 = flag_code(fl, F_B1_SHIFT), b2 = flag_code(fl, F_B2_SHIFT), b3 = flag_code(fl, F_B3_SHIFT);
    Real Fx = 0, Fy = 0, Fz = 0;  // wall momentum of this step's domain-face links
    ft[0] = Codec<S, Real, 27>::load(fs + pop_at<27, S>(0, (uint32_t)ic, a.fstride), 0, w27<Real>(0));
#pragma unroll
    for (int q = 1; q < 27; q++) {
        const long src = face_upwind27(q, ic, i, j, k, nx, ny, nz, b1, b2, b3);
        if (src >= 0) {
            ft[q] = Codec<S, Real, 27>::load(fs + pop_at<27, S>(q, (uint32_t)src, a.fstride), q, w27<Real>(q));
        } else {
            // reflected population; a Symmetry face code mirrors instead, the z test
            // overriding the y test overriding the x test (K27:1133-1145)
            int qb = L27::opp(q);
            qb = (b1 == BC_SYMMETRY) ? L27::mirx[q] : qb;
            qb = (b2 == BC_SYMMETRY) ? L27::miry[q] : qb;
            qb = (b3 == BC_SYMMETRY) ? L27::mirz[q] : qb;
            ft[q] = Codec<S, Real, 27>::load(fs + pop_at<27, S>(qb, (uint32_t)ic, a.fstride), qb, w27<Real>(qb));
            Fx += Real(-2) * ft[q] * Real(L27::cx[q]);
            Fy += Real(-2) * ft[q] * Real(L27::cy[q]);
            Fz += Real(-2) * ft[q] * Real(L27::cz[q]);
        }
    }
    const int corner = corner_flag(b3, i, j, k, nx, ny, nz);
    const long sy = nx, sz = (long)nx * ny;

    // fixedVelocityBC, K27:1243-1383
    if (corner == 0 && (b1 == BC_FIXEDVELOCITY || b2 == BC_FIXEDVELOCITY || b3 == BC_FIXEDVELOCITY)) {
        const Real u0 = a.u0[ic], v0 = a.v0[ic], w0 = a.w0[ic];
        if (b1 == BC_FIXEDVELOCITY && (i == 0 || i == nx - 1)) fixed_velocity27<0, S>(a, ft, ic, i, 1, u0, v0, w0);
        if (b2 == BC_FIXEDVELOCITY && (j == 0 || j == ny - 1)) fixed_velocity27<1, S>(a, ft, ic, j, sy, u0, v0, w0);
        if (b3 == BC_FIXEDVELOCITY && (k == 0 || k == nz - 1)) fixed_velocity27<2, S>(a, ft, ic, k, sz, u0, v0, w0);
    }
    // fixedDensityBC with rho_w = 1, K27:1385-1485
    if (corner == 0 && (b1 == BC_FIXEDDENSITY || b2 == BC_FIXEDDENSITY || b3 == BC_FIXEDDENSITY)) {
        if (b1 == BC_FIXEDDENSITY) fixed_density27<0>(a, ft, ic, i, nx, 1);
        if (b2 == BC_FIXEDDENSITY) fixed_density27<1>(a, ft, ic, j, ny, sy);
        if (b3 == BC_FIXEDDENSITY) fixed_density27<2>(a, ft, ic, k, nz, sz);
    }
    // wallFunctionBC, K27:1524-1821 (stationary wall, tangential slip velocity)
    if (corner == 0 && (b1 == BC_WALLFUNCTION || b2 == BC_WALLFUNCTION || b3 == BC_WALLFUNCTION)) {
        const Real nu = (Real(1) / omega - Real(0.5)) / Real(3);
        if (b1 == BC_WALLFUNCTION && (i == 0 || i == nx - 1)) wall_function27<0>(a, ft, ic, i, nz, 1, nu, Fx, Fy, Fz);
        if (b2 == BC_WALLFUNCTION && (j == 0 || j == ny - 1)) wall_function27<1>(a, ft, ic, j, nz, sy, nu, Fx, Fy, Fz);
        if (b3 == BC_WALLFUNCTION && (k == 0 || k == nz - 1)) wall_function27<2>(a, ft, ic, k, nz, sz, nu, Fx, Fy, Fz);
    }
    // outflowBC, K27:1900-1991: convective outlet with the cell's last velocity
    if (b1 == BC_OUTLET || b2 == BC_OUTLET || b3 == BC_OUTLET) {
        const Real is3 = Real(0.57735026918962576);
        const Real uo = a.u_o[ic], vo = a.v_o[ic], wo = a.w_o[ic];
#pragma unroll
        for (int q = 1; q < 27; q++) {
            const int cx = L27::cx[q], cy = L27::cy[q], cz = L27::cz[q];
#define OUTF(cond, inner, un)                                                                          \
    if (cond) {                                                                                        \
        const Real fi = Codec<S, Real, 27>::load(fs + pop_at<27, S>(q, (uint32_t)(inner), a.fstride), q, w27<Real>(q)); \
        const Real fc = Codec<S, Real, 27>::load(fs + pop_at<27, S>(q, (uint32_t)ic, a.fstride), q, w27<Real>(q)); \
        ft[q] = (is3 - (un)) * fi + (Real(1) - (is3 - (un))) * fc;                                      \
    }
            OUTF(i == 0 && b1 == BC_OUTLET && cx == 1, ic + 1, uo)
            OUTF(i == nx - 1 && b1 == BC_OUTLET && cx == -1, ic - 1, uo)
            OUTF(j == 0 && b2 == BC_OUTLET && cy == 1, ic + sy, vo)
            OUTF(j == ny - 1 && b2 == BC_OUTLET && cy == -1, ic - sy, vo)
            OUTF(k == 0 && b3 == BC_OUTLET && cz == 1, ic + sz, wo)
            OUTF(k == nz - 1 && b3 == BC_OUTLET && cz == -1, ic - sz, wo)
#undef OUTF
        }
    }
}

// First pass of the moment transform: the populations in slot order (9 z lines of 3),
// each line reduced to its raw z moments (sum f, sum z f, sum z^2 f).  Density, z momentum
// and the raw second moments follow from the nine lines with a few additions, so the
// collision's central moments and the cell's moments share this pass.
template <typename Real>
__device__ inline Mom19<Real> zpass_raw27(Real *K, const Real *ft)
{
#pragma unroll
    for (int s = 0; s < 27; s++) K[s] = ft[L27::q_of_slot[s]];
    // x and y momentum as sums of differences of mirrored populations: at low Mach
    // number the momentum is a small difference of order-one sums, and differencing
    // first keeps its rounding error at the level of the reference's signed sum in
    // direction order (opposite directions adjacent)
    Real mx = Real(0), my = Real(0);
#pragma unroll
    for (int s = 0; s < 9; s++) mx += K[18 + s] - K[s];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) my += K[9 * a + 6 + c] - K[9 * a + c];
#pragma unroll
    for (int p = 0; p < 9; p++) {
        const Real a = K[3 * p], b = K[3 * p + 1], c = K[3 * p + 2];
        const Real sum = c + a;
        K[3 * p] = sum + b;      // sum f
        K[3 * p + 1] = c - a;    // sum z f
        K[3 * p + 2] = sum;      // sum z^2 f
    }
    // line p = 3 (x+1) + (y+1)
#define L0(x, y) K[3 * (3 * ((x) + 1) + (y) + 1)]
#define L1(x, y) K[3 * (3 * ((x) + 1) + (y) + 1) + 1]
#define L2(x, y) K[3 * (3 * ((x) + 1) + (y) + 1) + 2]
    Mom19<Real> m;
    const Real xm = L0(-1, -1) + L0(-1, 0) + L0(-1, 1), x0 = L0(0, -1) + L0(0, 0) + L0(0, 1);
    const Real xp = L0(1, -1) + L0(1, 0) + L0(1, 1);
    const Real ym = L0(-1, -1) + L0(0, -1) + L0(1, -1), yp = L0(-1, 1) + L0(0, 1) + L0(1, 1);
    m.rho = xm + x0 + xp;
    m.mx = mx;
    m.my = my;
    m.mz = (L1(-1, -1) + L1(-1, 0) + L1(-1, 1)) + (L1(0, -1) + L1(0, 0) + L1(0, 1)) + (L1(1, -1) + L1(1, 0) + L1(1, 1));
    m.sxx = xm + xp;
    m.syy = ym + yp;
    m.szz = (L2(-1, -1) + L2(-1, 0) + L2(-1, 1)) + (L2(0, -1) + L2(0, 0) + L2(0, 1)) + (L2(1, -1) + L2(1, 0) + L2(1, 1));
    m.sxy = (L0(1, 1) + L0(-1, -1)) - (L0(1, -1) + L0(-1, 1));
    m.sxz = (L1(1, -1) + L1(1, 0) + L1(1, 1)) - (L1(-1, -1) + L1(-1, 0) + L1(-1, 1));
    m.syz = (L1(-1, 1) + L1(0, 1) + L1(1, 1)) - (L1(-1, -1) + L1(0, -1) + L1(1, -1));
#undef L0
#undef L1
#undef L2
    return m;
}

// 3-point central-moment transform along one axis: (f(-1), f(0), f(+1)) ->
// (k0, k1, k2) = sum_c (c - u)^n f(c), n = 0, 1, 2
template <typename Real>
__device__ inline void fwd3(Real &a, Real &b, Real &c, Real u)
{
    const Real s = c + a, d = c - a;
    const Real k0 = s + b;
    const Real k1 = d - u * k0;
    const Real k2 = s - u * (d + k1);
    a = k0;
    b = k1;
    c = k2;
}

// inverse of fwd3 (the reference's formulas, K27:2563-2662, regrouped):
// (k0, k1, k2) -> (f(-1), f(0), f(+1)) with hu = u/2, huu = u^2/2:
//   P = huu k0 + u k1 + k2/2,  D = hu k0 + k1/2,  f(-1) = P - D,  f(+1) = P + D,  f(0) = k0 - 2P
template <typename Real>
__device__ inline void bwd3(Real &a, Real &b, Real &c, Real u, Real hu, Real huu)
{
    const Real k0 = a, k1 = b, k2 = c;
    const Real P = huu * k0 + (u * k1 + Real(0.5) * k2);
    const Real D = hu * k0 + Real(0.5) * k1;
    a = P - D;
    c = P + D;
    b = k0 - Real(2) * P;
}

// K27:2352-2691 collisionCumulantGeier: cumulants of the central moments about the
// force-shifted velocity, relaxation with omega1 = omegaEff, omega2 = omegaB, omega3..5
// from the parametrization of Geier et al. (2017) with the limiter clim, the A/B
// fourth-order corrections, and back.  The back-transform products of Ks122, Ks212 and
// Ks221 use Ks022, Ks202 and Ks220 where the forward ones use K002, K200 and K020, as the
// reference does (K27:2557-2559).  ft (by direction) is replaced by the post-collision
// populations.
template <typename Real>
__device__ inline void cumulant27(Real *ft, Real *K, Real rho, Real u, Real v, Real w, Real omegaEff, Real omegaB,
                                  Real clim)
{
    // central moments.  K holds the raw z moments of the nine z lines (zpass_raw27):
    // shift them to central about w (k1 = k1r - w k0, k2 = k2r - w (k1r + k1), as fwd3),
    // then the y lines and the x lines; afterwards K[9a+3b+c] = K_abc
#pragma unroll
    for (int p = 0; p < 9; p++) {
        const Real k0 = K[3 * p], k1r = K[3 * p + 1], k2r = K[3 * p + 2];
        const Real k1 = k1r - w * k0;
        K[3 * p + 1] = k1;
        K[3 * p + 2] = k2r - w * (k1r + k1);
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) fwd3(K[9 * a + c], K[9 * a + 3 + c], K[9 * a + 6 + c], v);
#pragma unroll
    for (int p = 0; p < 9; p++) fwd3(K[p], K[9 + p], K[18 + p], u);
#define KM(a, b, c) K[9 * (a) + 3 * (b) + (c)]
    const Real K000 = rho;
    const Real K100 = KM(1, 0, 0), K010 = KM(0, 1, 0), K001 = KM(0, 0, 1);
    const Real K200 = KM(2, 0, 0), K020 = KM(0, 2, 0), K002 = KM(0, 0, 2);
    const Real K110 = KM(1, 1, 0), K101 = KM(1, 0, 1), K011 = KM(0, 1, 1);
    const Real K210 = KM(2, 1, 0), K201 = KM(2, 0, 1), K021 = KM(0, 2, 1);
    const Real K120 = KM(1, 2, 0), K102 = KM(1, 0, 2), K012 = KM(0, 1, 2), K111 = KM(1, 1, 1);
    const Real K220 = KM(2, 2, 0), K202 = KM(2, 0, 2), K022 = KM(0, 2, 2);
    const Real K211 = KM(2, 1, 1), K121 = KM(1, 2, 1), K112 = KM(1, 1, 2);
    const Real K221 = KM(2, 2, 1), K212 = KM(2, 1, 2), K122 = KM(1, 2, 2), K222 = KM(2, 2, 2);

    const Real invRho = rcp_nr(rho);
    // cumulants (orders 2 and 3 equal the central moments)
    const Real C200 = K200, C020 = K020, C002 = K002, C110 = K110, C101 = K101, C011 = K011;
    const Real C210 = K210, C201 = K201, C021 = K021, C120 = K120, C102 = K102, C012 = K012, C111 = K111;
    const Real C211 = K211 - (K200 * K011 + Real(2) * K110 * K101) * invRho;
    const Real C121 = K121 - (K020 * K101 + Real(2) * K110 * K011) * invRho;
    const Real C112 = K112 - (K002 * K110 + Real(2) * K101 * K011) * invRho;
    const Real C220 = K220 - (K200 * K020 + Real(2) * K110 * K110) * invRho;
    const Real C202 = K202 - (K002 * K200 + Real(2) * K101 * K101) * invRho;
    const Real C022 = K022 - (K020 * K002 + Real(2) * K011 * K011) * invRho;
    const Real C122 = K122 - (K002 * K120 + K020 * K102 + Real(4) * K011 * K111 + Real(2) * (K101 * K021 + K110 * K012)) * invRho;
    const Real C212 = K212 - (K200 * K012 + K002 * K210 + Real(4) * K101 * K111 + Real(2) * (K110 * K102 + K011 * K201)) * invRho;
    const Real C221 = K221 - (K020 * K201 + K200 * K021 + Real(4) * K110 * K111 + Real(2) * (K011 * K210 + K101 * K120)) * invRho;
    const Real C222 = K222
                      - (Real(4) * K111 * K111 + K200 * K022 + K020 * K202 + K002 * K220
                         + Real(4) * (K011 * K211 + K101 * K121 + K110 * K112)
                         + Real(2) * (K120 * K102 + K210 * K012 + K201 * K021)) * invRho
                      + (Real(16) * K110 * K101 * K011 + Real(4) * (K101 * K101 * K020 + K011 * K011 * K200 + K110 * K110 * K002)
                         + Real(2) * K200 * K020 * K002) * invRho * invRho;

    // relaxation rates (Geier et al. 2017, eqs. for omega3..omega5, A, B)
    const Real o1 = omegaEff, o2 = omegaB;
    const Real o3 = Real(8) * (o1 - Real(2)) * (o2 * (Real(3) * o1 - Real(1)) - Real(5) * o1)
                    * rcp(Real(8) * (Real(5) - Real(2) * o1) * o1 + o2 * (Real(8) + o1 * (Real(9) * o1 - Real(26))));
    const Real o4 = Real(8) * (o1 - Real(2)) * (o1 + o2 * (Real(3) * o1 - Real(7)))
                    * rcp(o2 * (Real(56) - Real(42) * o1 + Real(9) * o1 * o1) - Real(8) * o1);
    const Real o5 = Real(24) * (o1 - Real(2))
                    * (Real(4) * o1 * o1 + o1 * o2 * (Real(18) - Real(13) * o1) + o2 * o2 * (Real(2) + o1 * (Real(6) * o1 - Real(11))))
                    * rcp(Real(16) * o1 * o1 * (o1 - Real(6)) - Real(2) * o1 * o2 * (Real(216) + Real(5) * o1 * (Real(9) * o1 - Real(46)))
                          + o2 * o2 * (o1 * (Real(3) * o1 - Real(10)) * (Real(15) * o1 - Real(28)) - Real(48)));
    const Real climRho = clim * rho;
    auto lim = [&](Real om, Real c) -> Real {
        const Real ac = fabs(c);
        return ac > Real(0) ? om + (Real(1) - om) * ac * rcp(ac + climRho) : Real(1);
    };
    const Real o3a = lim(o3, C120 + C102), o3b = lim(o3, C210 + C012), o3c = lim(o3, C201 + C021);
    const Real o4a = lim(o4, C120 - C102), o4b = lim(o4, C210 - C012), o4c = lim(o4, C201 - C021);
    const Real o5l = lim(o5, C111);
    const Real denAB = rcp((o1 - o2) * (o2 * (Real(2) + Real(3) * o1) - Real(8) * o1));
    Real A = (Real(4) * o1 * o1 + Real(2) * o1 * o2 * (o1 - Real(6)) + o2 * o2 * (o1 * (Real(10) - Real(3) * o1) - Real(4))) * denAB;
    Real B = (Real(4) * o1 * o2 * (Real(9) * o1 - Real(16)) - Real(4) * o1 * o1 - Real(2) * o2 * o2 * (Real(2) + Real(9) * o1 * (o1 - Real(2))))
             * denAB * (Real(1) / Real(3));
    A = fmax(A, Real(0));
    B = fmax(B, Real(0));

    // second order: shear rates relax with omega1, the trace with omega2
    const Real Cs110 = (Real(1) - o1) * C110, Cs101 = (Real(1) - o1) * C101, Cs011 = (Real(1) - o1) * C011;
    const Real Dxu = Real(-0.5) * (o1 * (Real(2) * C200 - C020 - C002) + o2 * (C200 + C020 + C002 - K000)) * invRho;
    const Real Dyv = Dxu + Real(1.5) * o1 * (C200 - C020) * invRho;
    const Real Dzw = Dxu + Real(1.5) * o1 * (C200 - C002) * invRho;
    const Real Dxv_Dyu = Real(-3) * o1 * C110 * invRho, Dxw_Dzu = Real(-3) * o1 * C101 * invRho;
    const Real Dyw_Dzv = Real(-3) * o1 * C011 * invRho;
    Real X1 = (Real(1) - o1) * (C200 - C020) - Real(3) * rho * (Real(1) - Real(0.5) * o1) * (u * u * Dxu - v * v * Dyv);
    Real X2 = (Real(1) - o1) * (C200 - C002) - Real(3) * rho * (Real(1) - Real(0.5) * o1) * (u * u * Dxu - w * w * Dzw);
    Real X3 = K000 * o2 + (Real(1) - o2) * (C200 + C020 + C002)
              - Real(3) * rho * (Real(1) - Real(0.5) * o2) * (u * u * Dxu + v * v * Dyv + w * w * Dzw);
    const Real third = Real(1) / Real(3);
    const Real Cs200 = (X1 + X2 + X3) * third, Cs020 = Cs200 - X1, Cs002 = Cs200 - X2;
    // third order
    X1 = (Real(1) - o3a) * (C120 + C102);
    X2 = (Real(1) - o4a) * (C120 - C102);
    const Real Cs120 = Real(0.5) * (X1 + X2), Cs102 = Real(0.5) * (X1 - X2);
    X1 = (Real(1) - o3b) * (C210 + C012);
    X2 = (Real(1) - o4b) * (C210 - C012);
    const Real Cs210 = Real(0.5) * (X1 + X2), Cs012 = Real(0.5) * (X1 - X2);
    X1 = (Real(1) - o3c) * (C201 + C021);
    X2 = (Real(1) - o4c) * (C201 - C021);
    const Real Cs201 = Real(0.5) * (X1 + X2), Cs021 = Real(0.5) * (X1 - X2);
    const Real Cs111 = (Real(1) - o5l) * C111;
    // fourth order with omega6..8 = 1 (only the A/B terms remain); fifth, sixth order -> 0
    const Real vis = rcp(o1) - Real(0.5);
    const Real vA = Real(2) * vis * A * rho * third, vB = -vis * B * rho * third;
    X1 = vA * (Dxu - Real(2) * Dyv + Dzw);
    X2 = vA * (Dxu + Dyv - Real(2) * Dzw);
    X3 = Real(-2) * vA * (Dxu + Dyv + Dzw);
    const Real Cs220 = (X1 + X2 + X3) * third, Cs202 = (X3 - X1) * third, Cs022 = (X3 - X2) * third;
    const Real Cs211 = vB * Dyw_Dzv;
    const Real Cs121 = vB * Dxw_Dzu;
    const Real Cs112 = vB * Dxv_Dyu;

    // back to central moments
    const Real Ks200 = Cs200, Ks020 = Cs020, Ks002 = Cs002, Ks110 = Cs110, Ks101 = Cs101, Ks011 = Cs011;
    const Real Ks111 = Cs111, Ks210 = Cs210, Ks201 = Cs201, Ks021 = Cs021, Ks120 = Cs120, Ks102 = Cs102;
    const Real Ks012 = Cs012;
    const Real Ks211 = Cs211 + (Ks200 * Ks011 + Real(2) * Ks110 * Ks101) * invRho;
    const Real Ks121 = Cs121 + (Ks020 * Ks101 + Real(2) * Ks011 * Ks110) * invRho;
    const Real Ks112 = Cs112 + (Ks002 * Ks110 + Real(2) * Ks101 * Ks011) * invRho;
    const Real Ks220 = Cs220 + (Ks200 * Ks020 + Real(2) * Ks110 * Ks110) * invRho;
    const Real Ks202 = Cs202 + (Ks002 * Ks200 + Real(2) * Ks101 * Ks101) * invRho;
    const Real Ks022 = Cs022 + (Ks020 * Ks002 + Real(2) * Ks011 * Ks011) * invRho;
    const Real Ks122 = (Ks022 * Ks120 + Ks020 * Ks102 + Real(4) * Ks011 * Ks111 + Real(2) * (Ks101 * Ks021 + Ks110 * Ks012)) * invRho;
    const Real Ks212 = (Ks202 * Ks012 + Ks002 * Ks210 + Real(4) * Ks101 * Ks111 + Real(2) * (Ks110 * Ks102 + Ks011 * Ks201)) * invRho;
    const Real Ks221 = (Ks220 * Ks201 + Ks200 * Ks021 + Real(4) * Ks110 * Ks111 + Real(2) * (Ks011 * Ks210 + Ks101 * Ks120)) * invRho;
    const Real Ks222 = (Real(4) * Ks111 * Ks111 + Ks200 * Ks022 + Ks020 * Ks202 + Ks002 * Ks220
                        + Real(4) * (Ks011 * Ks211 + Ks101 * Ks121 + Ks110 * Ks112)
                        + Real(2) * (Ks120 * Ks102 + Ks210 * Ks012 + Ks201 * Ks021)) * invRho
                       - (Real(16) * Ks110 * Ks101 * Ks011
                          + Real(4) * (Ks101 * Ks101 * Ks020 + Ks011 * Ks011 * Ks200 + Ks110 * Ks110 * Ks002)
                          + Real(2) * Ks200 * Ks020 * Ks002) * invRho * invRho;
    KM(0, 0, 0) = K000;
    KM(1, 0, 0) = -K100;
    KM(0, 1, 0) = -K010;
    KM(0, 0, 1) = -K001;
    KM(2, 0, 0) = Ks200; KM(0, 2, 0) = Ks020; KM(0, 0, 2) = Ks002;
    KM(1, 1, 0) = Ks110; KM(1, 0, 1) = Ks101; KM(0, 1, 1) = Ks011;
    KM(2, 1, 0) = Ks210; KM(2, 0, 1) = Ks201; KM(0, 2, 1) = Ks021;
    KM(1, 2, 0) = Ks120; KM(1, 0, 2) = Ks102; KM(0, 1, 2) = Ks012; KM(1, 1, 1) = Ks111;
    KM(2, 2, 0) = Ks220; KM(2, 0, 2) = Ks202; KM(0, 2, 2) = Ks022;
    KM(2, 1, 1) = Ks211; KM(1, 2, 1) = Ks121; KM(1, 1, 2) = Ks112;
    KM(2, 2, 1) = Ks221; KM(2, 1, 2) = Ks212; KM(1, 2, 2) = Ks122; KM(2, 2, 2) = Ks222;
#undef KM
    // populations: x lines, y lines, z lines (K27:2563-2662)
    const Real hu = Real(0.5) * u, huu = hu * u, hv = Real(0.5) * v, hvv = hv * v, hw = Real(0.5) * w, hww = hw * w;
#pragma unroll
    for (int p = 0; p < 9; p++) bwd3(K[p], K[9 + p], K[18 + p], u, hu, huu);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) bwd3(K[9 * a + c], K[9 * a + 3 + c], K[9 * a + 6 + c], v, hv, hvv);
#pragma unroll
    for (int p = 0; p < 9; p++) bwd3(K[3 * p], K[3 * p + 1], K[3 * p + 2], w, hw, hww);
#pragma unroll
    for (int q = 0; q < 27; q++) ft[q] = K[L27::slot[q]];
}

// The recursive-regularized collision on the D3Q27 lattice, selectable at the D3Q27 call
// site (K27:2824): the structure of the reference's D3Q19 collisionRecursiveRegularized
// (K19:3156-3309) over the full D3Q27 Hermite set.  The second-order non-equilibrium
// moments come from the cell's shared raw moments; orders three to six are rebuilt
// recursively from them (each coefficient the sum, over the unordered index pairs of its
// multi-index, of N_pair times the velocities of the other indices -- the rule K19:3215-3224
// applies to x^2 y and x^2 y^2) and relaxed at omegaEff towards u^a v^b w^c; the second
// order splits its trace (omegaB) from the deviator (omegaEff) as K19:3226-3231.  Hermite
// coefficients become raw moments per axis (x^2 = H_xx + 1/3), and three passes of the
// 1-D inverse transform (f(-1), f(0), f(+1)) = ((m2 - m1)/2, m0 - m2, (m2 + m1)/2) return
// the populations.  u is the force-shifted velocity minus the half-step shift (K19:3641);
// the body force enters as S_q = rho w_q 3 c_qx dpdx (K19:3278-3287).  K is scratch.
template <typename Real>
__device__ inline void collide_rr27(Real *ft, Real *K, const Mom19<Real> &m, Real inv_rho, Real u, Real v, Real w,
                                    Real omegaEff, Real omegaB, Real dpdx)
{
    const Real cs2 = Real(1) / Real(3), cs4 = Real(1) / Real(9), cs6 = Real(1) / Real(27);
    const Real rho = m.rho;
    const Real R200 = (m.sxx - cs2 * rho) * inv_rho, R020 = (m.syy - cs2 * rho) * inv_rho;
    const Real R002 = (m.szz - cs2 * rho) * inv_rho;
    const Real R110 = m.sxy * inv_rho, R101 = m.sxz * inv_rho, R011 = m.syz * inv_rho;
    const Real uu = u * u, vv = v * v, ww = w * w, uv = u * v, uw = u * w, vw = v * w;
    const Real N200 = R200 - uu, N020 = R020 - vv, N002 = R002 - ww;
    const Real N110 = R110 - uv, N101 = R101 - uw, N011 = R011 - vw;
    const Real omegaM = (omegaB - omegaEff) * cs2, omegaP = omegaM + omegaEff, om1 = Real(1) - omegaEff;
    const Real C200 = R200 - omegaP * N200 - omegaM * N020 - omegaM * N002;
    const Real C020 = R020 - omegaM * N200 - omegaP * N020 - omegaM * N002;
    const Real C002 = R002 - omegaM * N200 - omegaM * N020 - omegaP * N002;
    const Real C110 = R110 - omegaEff * N110, C101 = R101 - omegaEff * N101, C011 = R011 - omegaEff * N011;
    // third order
    const Real C210 = uu * v + om1 * (v * N200 + Real(2) * u * N110);
    const Real C201 = uu * w + om1 * (w * N200 + Real(2) * u * N101);
    const Real C021 = vv * w + om1 * (w * N020 + Real(2) * v * N011);
    const Real C120 = vv * u + om1 * (u * N020 + Real(2) * v * N110);
    const Real C102 = ww * u + om1 * (u * N002 + Real(2) * w * N101);
    const Real C012 = ww * v + om1 * (v * N002 + Real(2) * w * N011);
    const Real C111 = uv * w + om1 * (w * N110 + v * N101 + u * N011);
    // fourth order
    const Real C220 = uu * vv + om1 * (vv * N200 + uu * N020 + Real(4) * uv * N110);
    const Real C202 = uu * ww + om1 * (ww * N200 + uu * N002 + Real(4) * uw * N101);
    const Real C022 = vv * ww + om1 * (ww * N020 + vv * N002 + Real(4) * vw * N011);
    const Real C211 = uu * vw + om1 * (vw * N200 + Real(2) * (uw * N110 + uv * N101) + uu * N011);
    const Real C121 = vv * uw + om1 * (uw * N020 + Real(2) * (vw * N110 + uv * N011) + vv * N101);
    const Real C112 = ww * uv + om1 * (uv * N002 + Real(2) * (vw * N101 + uw * N011) + ww * N110);
    // fifth and sixth order
    const Real uvw = uv * w;
    const Real C221 = uu * vv * w + om1 * (vv * w * N200 + uu * w * N020 + Real(4) * uvw * N110
                                           + Real(2) * (u * vv * N101 + uu * v * N011));
    const Real C212 = uu * v * ww + om1 * (v * ww * N200 + uu * v * N002 + Real(4) * uvw * N101
                                           + Real(2) * (u * ww * N110 + uu * w * N011));
    const Real C122 = u * vv * ww + om1 * (u * ww * N020 + u * vv * N002 + Real(4) * uvw * N011
                                           + Real(2) * (v * ww * N110 + vv * w * N101));
    const Real C222 = uu * vv * ww + om1 * (vv * ww * N200 + uu * ww * N020 + uu * vv * N002
                                            + Real(4) * (uvw * w * N110 + uvw * v * N101 + uvw * u * N011));
    // raw moments per unit density in the transform's layout K[9 a + 3 b + c] = M_abc
    K[0] = Real(1);
    K[9] = u;
    K[3] = v;
    K[1] = w;
    K[18] = C200 + cs2;
    K[6] = C020 + cs2;
    K[2] = C002 + cs2;
    K[12] = C110;
    K[10] = C101;
    K[4] = C011;
    K[21] = C210 + cs2 * v;
    K[19] = C201 + cs2 * w;
    K[7] = C021 + cs2 * w;
    K[15] = C120 + cs2 * u;
    K[11] = C102 + cs2 * u;
    K[5] = C012 + cs2 * v;
    K[13] = C111;
    K[24] = C220 + cs2 * (C200 + C020) + cs4;
    K[20] = C202 + cs2 * (C200 + C002) + cs4;
    K[8] = C022 + cs2 * (C020 + C002) + cs4;
    K[22] = C211 + cs2 * C011;
    K[16] = C121 + cs2 * C101;
    K[14] = C112 + cs2 * C110;
    K[25] = C221 + cs2 * (C201 + C021) + cs4 * w;
    K[23] = C212 + cs2 * (C210 + C012) + cs4 * v;
    K[17] = C122 + cs2 * (C120 + C102) + cs4 * u;
    K[26] = C222 + cs2 * (C220 + C202 + C022) + cs4 * (C200 + C020 + C002) + cs6;
    // back to populations: x, y, z passes of the 1-D inverse transform (no shift)
    auto inv3 = [](Real &a, Real &b, Real &c) {
        const Real m0 = a, h1 = Real(0.5) * b, h2 = Real(0.5) * c;
        a = h2 - h1;
        b = m0 - c;
        c = h2 + h1;
    };
#pragma unroll
    for (int p = 0; p < 9; p++) inv3(K[p], K[9 + p], K[18 + p]);
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int c = 0; c < 3; c++) inv3(K[9 * x + c], K[9 * x + 3 + c], K[9 * x + 6 + c]);
#pragma unroll
    for (int p = 0; p < 9; p++) inv3(K[3 * p], K[3 * p + 1], K[3 * p + 2]);
    const Real F = rho * Real(3) * dpdx;
#pragma unroll
    for (int q = 0; q < 27; q++) ft[q] = rho * K[L27::slot[q]] + F * w27<Real>(q) * Real(L27::cx[q]);
}

// At least seven waves per SIMD (72 VGPRs) for the 32-bit arithmetic: the cumulant update
// is VALU bound and fits without spilling (eight waves, 64 VGPRs, would spill); the
// compiler's own choice went as low as five.  FP32 populations with WALE: six (76-78 VGPRs),
// as the population blocks' pull spills 8 B at seven and the spill-free six ran the bluff
// body's step in 5.82 against 6.03 ms.  The FP64 form keeps the compiler's choice.
template <typename S, typename Real, int SGS>
constexpr int kStep27MinWaves = sizeof(Real) == 8 ? 1 : ((sizeof(S) == 4 && SGS == 0) ? 6 : 7);
// the batch kernel's extra indirection spills at seven: six
template <typename Real>
constexpr int kBatch27MinWaves = sizeof(Real) == 8 ? 1 : 6;

// PASS as step19_kernel: 1 the cells off the lattice faces, 2 the listed face cells, 3 both
// in one grid (the face-list work-groups first)
template <typename S, typename Real, int SGS, int PASS = 0, int COLL = COLL_CUMULANT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kStep27MinWaves<S, Real, SGS>, 8)))
step27_kernel(const StepArgs<Real, 27> a)
{
    const uint32_t nfb = (PASS == 3) ? face_blocks(a) : 0u;
    const bool face_wg = (PASS == 2) || (PASS == 3 && blockIdx.x < nfb);
    uint32_t icu;
    if (face_wg) {
        const uint32_t lin = blockIdx.x * blockDim.x + threadIdx.x;
        if (lin >= a.nface) return;
        icu = a.fcells[lin];
        if (face_pass_skips(cell_wave_class(a.wclass, icu))) return;
    } else {
        const uint32_t g = blockIdx.x - nfb;
        const uint32_t blk = ordered_block(g, gridDim.x - nfb, a.tune);
        const uint32_t lin = blk * blockDim.x + threadIdx.x;
        if (lin >= a.N) return;
        icu = (a.tune & TUNE_YTILE) ? ytile_cell(a, lin) : lin;
    }
#define CELL_RETURN return
#define CELL_WALE_INTERIOR true
#define CELL_SKIP_FACE (PASS == 1 || (PASS == 3 && !face_wg))
// (the 16-bit split interior pass would spill with the per-wave fast path)
#define CELL_PURE_OK (!face_wg && !(PASS == 1 && sizeof(S) == 2))
#define CELL_ALIGNED (!face_wg && a.wave_runs)
#define CELL_COLL COLL
#include "d3q27_cell.inc"
#undef CELL_COLL
#undef CELL_ALIGNED
#undef CELL_PURE_OK
#undef CELL_SKIP_FACE
#undef CELL_WALE_INTERIOR
#undef CELL_RETURN
}

}  // namespace lbffs
