// d3q19_step.hpp -- the fused D3Q19 time-step kernel.
//
// One work-item per cell, cells flattened as ic = i + nx*(j + ny*k) so that a
// wavefront's 64 lanes touch 64 consecutive cells of one direction (in a population block
// or a direction array, pop_at): every store is a coalesced run of 256 B (FP32) or 128 B
// (16-bit), every pull one or two.
// The kernel does, for each fluid cell, what k_streamingCollision does
// (D3Q19_opencl/kernel.cl:3547-3645):
//   pull streaming (interior: fixed offsets; faces: wrap or wall per face code),
//   fixed-velocity / fixed-density / wall-function / outlet face conditions,
//   simple bounce-back on solid links with momentum-exchange force,
//   moments, sub-grid model (WALE or Smagorinsky), sponge, recursive-regularized
//   collision with the dpdx body force.
// Deliberate differences from the reference's data flow (results are unchanged):
//   * neighbour reads of rho/u/v/w see the previous step (Jacobi) instead of a
//     scheduling-dependent mix, via two macro buffers that alternate each step;
//   * 19 per-cell side arrays are replaced by one 16-bit flag word; side arrays are
//     read only by the cells whose flag asks for them;
//   * rho/u/v/w/tauSGS/Fw/G are written for every cell only when the caller asks
//     (last step of a call or store_every), otherwise only where later steps read them.
#pragma once

#include "lbffs_common.hpp"

namespace lbffs {

// D3Q19 directions (kernel.cl:3558-3560) as compile-time tables, so that unrolled
// loops fold the lattice constants into the instruction stream
struct L19 {
    static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    __host__ __device__ static constexpr int opp(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }
};

template <typename Real>
__host__ __device__ constexpr Real w19(int q)
{
    return q == 0 ? Real(1) / Real(3) : (q <= 6 ? Real(1) / Real(18) : Real(1) / Real(36));
}

// Arguments of one step launch, shared by the D3Q19 and D3Q27 kernels.
template <typename Real, int Q>
struct StepArgs {
    uint32_t N;
    uint32_t fstride;  // 16-bit values: elements between two direction arrays (pop_at)
    int wave_runs;     // 1: a wave of the cell order holds one aligned run of 64 cells
    int nx, ny, nz;
    FastDiv dnx, dny;
    int off[Q];  // interior pull offset: source = ic - off[q]
    const void *fsrc;
    void *fdst;
    const uint16_t *flag;
    const Real *u0, *v0, *w0, *sdf, *omega;
    const Real *rho_o, *u_o, *v_o, *w_o;  // macros of the previous step
    Real *rho_n, *u_n, *v_n, *w_n;        // macros written by this step
    Real *tau, *Fwx, *Fwy, *Fwz, *Gx, *Gy, *Gz;
    Real dpdx, LES, omegaB, spz, Cs, omega0;
    Real clim;  // cumulant limiter (D3Q27)
    const Real *qlink;  // q-distances of the solid links (Q*N, q-major), read when ibb
    int ibb;            // interpolated bounce-back with qlink instead of the reference's walls
    int sround;         // 16-bit forms: stochastic rounding of the stored deviations
    int damping;
    int mac_all;      // write u/v/w for every fluid cell
    int rho_all;      // write rho for every fluid cell (implies mac_all)
    int aux_all;      // write tauSGS / Fw / G for every fluid cell
    int sponge_mask;  // faces whose centre cell is Outlet or FixedDensity (tauSpongeZone)
    Real inv_spx, inv_spy, inv_spz;  // 1 / (n * spzWidth) per axis
    int tune;         // TUNE_* launch options chosen by the autotuner
    // y-tiled cell order (TUNE_YTILE): tiles of ty rows, plane by plane inside a tile
    uint32_t tile_cells, tile_plane, nxny;  // nx*ty*nz, nx*ty, nx*ny
    FastDiv dtile, dplane;
    const uint32_t *fcells;  // the lattice's non-solid face cells (the face pass runs these)
    uint32_t nface;
    const uint32_t *wclass;  // per-wave summary (wave_class), or null: every wave reads its flags first
    int coll;                 // D3Q27: LBFFS_COLL_RR or LBFFS_COLL_CUMULANT (uniform per launch)
    uint32_t step;            // step number (stochastic rounding), see sr_step
    const uint32_t *step_ctr; // batches: device counter of the steps run in the current call
};

template <typename Real>
using Step19Args = StepArgs<Real, 19>;

// launch options (uniform per launch)
enum : int {
    TUNE_NT_STORE = 1,  // accepted, no effect: the populations are always stored non-temporally
    TUNE_NO_XCD = 2,    // keep the hardware work-group order (no XCD slab remap)
    TUNE_YTILE = 4,     // visit the cells in y-tiles (see ytile_cell)
    TUNE_SPLIT = 8,     // two passes per step: the cells off the faces, then the face cells
    TUNE_NO_PURE = 16,  // every wave waits for its flags before it pulls (no per-wave summary)
    TUNE_FACE_FIRST = 32,  // one launch: the face-list work-groups first, then the cells off the faces
    TUNE_XCD_CHUNK = 64,   // deal runs of kXcdChunk consecutive blocks to each XCD in turn
};

// the face pass skips the face cells that the interior pass updates in their own waves:
// those of XMIX and WRAP waves.  (Handing the x-face cells of XMIX waves to the face pass
// instead is slower on every lattice measured: consecutive x-face cells of the list lie a
// row apart, so each lane's pulls touch cache lines of their own, where in their own wave
// they share the lines of their neighbours' pulls.)
__device__ inline bool face_pass_skips(uint32_t cls)
{
    return cls == WAVE_WRAP || cls == WAVE_XMIX;
}

// Cell of position lin in the y-tiled order: tiles of ty full rows (ny a multiple of ty),
// each walked plane by plane (k), row by row inside a plane.  The velocity gradient of
// WALE reads the planes k - 1 and k + 1 of a cell's velocities: over a tile of 32 rows of
// 1024 cells those three planes of three arrays take 1.2 MB and stay in one XCD's L2
// between their uses, where whole planes (1.5 MB each) do not.
template <typename A>
__device__ inline uint32_t ytile_cell(const A &a, uint32_t lin)
{
    const uint32_t tile = fdiv(lin, a.dtile);
    const uint32_t rem = lin - tile * a.tile_cells;
    const uint32_t k = fdiv(rem, a.dplane);
    return (rem - k * a.tile_plane) + tile * a.tile_plane + k * a.nxny;
}

// XCD-aware block order: the hardware deals work-groups round-robin over the 8 XCDs;
// renumber them so that each XCD walks one contiguous slab of cells and the partial
// cache lines at block edges are shared inside one L2.
__device__ inline uint32_t xcd_block(uint32_t pid, uint32_t G)
{
    const uint32_t nx = 8;
    uint32_t x = pid % nx, local = pid / nx;
    uint32_t base = G / nx, r = G % nx;
    return x * base + (x < r ? x : r) + local;
}

// XCD-chunked block order: the k-th work-group of XCD x (hardware block 8k + x) takes block
// (k / C) 8C + x C + k mod C, so each XCD runs C consecutive blocks while the eight XCDs
// still sweep one contiguous front of the arrays.  A pull shifted by one cell in x reads
// one cache line of its neighbour block; in hardware order that neighbour runs on another
// XCD and the line is fetched twice (once per L2), in chunks only every C-th block edge
// crosses XCDs.  The blocks past the last whole group of 8C keep their own number.
constexpr uint32_t kXcdChunk = 4;
__device__ inline uint32_t xcd_chunk_block(uint32_t g, uint32_t G)
{
    constexpr uint32_t span = 8 * kXcdChunk;
    if (g >= G - G % span) return g;
    const uint32_t k = g / 8, x = g % 8;
    return (k / kXcdChunk) * span + x * kXcdChunk + k % kXcdChunk;
}

// block g of G in the order the launch option asks for; g % 8 must be the XCD that the
// hardware gives block g (true when the blocks before the first are a multiple of 8)
__device__ inline uint32_t ordered_block(uint32_t g, uint32_t G, uint32_t tune)
{
    if (tune & TUNE_XCD_CHUNK) return xcd_chunk_block(g, G);
    return (tune & TUNE_NO_XCD) ? g : xcd_block(g, G);
}

// pull source for a face cell (kernel.cl:328-575 upwindID_B): wrap across non-wall faces,
// -1 for a link through a wall face.
// Formed from the interior source ic - c_q in 32-bit arithmetic: a crossed periodic face moves
// the source by a whole row (x), plane (y) or lattice (z), c_x nx, c_y nx ny or c_z nx ny nz
// cells; unsigned wrap-around is exact because the result lies in the lattice.
__device__ inline long face_upwind(int q, long ic, int i, int j, int k, int nx, int ny, int nz, int b1, int b2, int b3)
{
    const int cx = L19::cx[q], cy = L19::cy[q], cz = L19::cz[q];
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

// kernel.cl:105-326 reflectOrMirrorQ as a table lookup: bounce-back partner, with
// Symmetry faces flipping only the normal component of diagonal directions.
__device__ inline int mirror_q(int q, int b1, int b2, int b3)
{
    if (q <= 6) return L19::opp(q);
    int cx = L19::cx[q], cy = L19::cy[q], cz = L19::cz[q];
    // flip the component normal to the first Symmetry face the reference tests
    if (q <= 10) {  // xy diagonals: test b1 then b2
        if (b1 == BC_SYMMETRY) cx = -cx;
        else if (b2 == BC_SYMMETRY) cy = -cy;
        else { cx = -cx; cy = -cy; }
    } else if (q <= 14) {  // xz diagonals: b1 then b3
        if (b1 == BC_SYMMETRY) cx = -cx;
        else if (b3 == BC_SYMMETRY) cz = -cz;
        else { cx = -cx; cz = -cz; }
    } else {  // yz diagonals: b2 then b3
        if (b2 == BC_SYMMETRY) cy = -cy;
        else if (b3 == BC_SYMMETRY) cz = -cz;
        else { cy = -cy; cz = -cz; }
    }
#pragma unroll
    for (int p = 7; p < 19; p++)
        if (L19::cx[p] == cx && L19::cy[p] == cy && L19::cz[p] == cz) return p;
    return L19::opp(q);
}

// kernel.cl:1346-1395 cornerFlag
__device__ inline int corner_flag(int b3, int i, int j, int k, int nx, int ny, int nz)
{
    bool xi = (i == 0 || i == nx - 1), yj = (j == 0 || j == ny - 1), zk = (k == 0 || k == nz - 1);
    if (nz - 1 == 0 || b3 == 0) return (xi && yj) ? 1 : 0;
    return ((xi && yj) || (xi && zk) || (yj && zk)) ? 1 : 0;
}

// kernel.cl:1621-1652 calcTauw (Newton on the wall law; the loop's error test uses
// the initial y+ as the reference does, K19:1646)
template <typename Real>
__device__ inline Real calc_tauw(Real rhow, Real magUp, Real nu)
{
    const Real y = Real(1.5), kappa = Real(0.42), E = Real(9.1);
    Real ut = sqrt(nu * magUp);
    int iter = 0;
    Real err;
    const Real yPlus = y * ut / nu;
    do {
        Real kUu = fmin(kappa * magUp / ut, Real(50));
        Real fkUu = exp(kUu) - Real(1) - kUu * (Real(1) + Real(0.5) * kUu);
        Real func = -ut * y / nu + magUp / ut + (fkUu - kUu * kUu * kUu / Real(6)) / E;
        Real dfunc = y / nu + magUp / (ut * ut) + (kUu * fkUu / ut) / E;
        ut = ut + func / dfunc;
        err = fabs(yPlus - y * ut / nu);
    } while (err > Real(0.005) && ++iter < 100);
    ut = fmax(ut, Real(0));
    return rhow * ut * ut;
}

// sum of the populations that leave through a face plus twice those arriving (Zou-He
// wall density numerator); dir 0/1/2 = x/y/z, lo = min face
template <typename Real>
__device__ inline Real zouhe_sum(const Real *ft, int dir, bool lo)
{
    Real s = Real(0), t = Real(0);
    // tangential populations in the reference's summation order
    if (dir == 0) {
        s = ft[0] + ft[3] + ft[4] + ft[5] + ft[6] + ft[15] + ft[16] + ft[17] + ft[18];
        t = lo ? (ft[2] + ft[10] + ft[8] + ft[14] + ft[12]) : (ft[1] + ft[7] + ft[9] + ft[11] + ft[13]);
    } else if (dir == 1) {
        s = ft[0] + ft[1] + ft[2] + ft[5] + ft[6] + ft[11] + ft[12] + ft[14] + ft[13];
        t = lo ? (ft[4] + ft[9] + ft[8] + ft[18] + ft[16]) : (ft[3] + ft[7] + ft[10] + ft[15] + ft[17]);
    } else {
        s = ft[0] + ft[1] + ft[2] + ft[3] + ft[4] + ft[7] + ft[8] + ft[9] + ft[10];
        t = lo ? (ft[6] + ft[12] + ft[13] + ft[16] + ft[17]) : (ft[5] + ft[14] + ft[11] + ft[18] + ft[15]);
    }
    return s + Real(2) * t;
}

// Ladd moving-wall correction of the five populations entering through a face with
// wall velocity (un, ut1, ut2) = (normal, first tangential, second tangential),
// kernel.cl:1397-1512.  Written with value selects (never "if (lo) ft[a] else ft[b]")
// so that ft[] stays in registers.
template <int A0, int A1, int A2, int A3, int A4, int B0, int B1, int B2, int B3, int B4, typename Real>
__device__ inline void ladd_face(Real *ft, bool lo, Real rhow, Real un, Real ut1, Real ut2)
{
    const Real d0 = rhow * un / Real(3);
    const Real d1 = rhow * (un + ut1) / Real(6);
    const Real d2 = rhow * (un - ut1) / Real(6);
    const Real d3 = rhow * (un + ut2) / Real(6);
    const Real d4 = rhow * (un - ut2) / Real(6);
    ft[A0] = lo ? ft[A0] + d0 : ft[A0];
    ft[A1] = lo ? ft[A1] + d1 : ft[A1];
    ft[A2] = lo ? ft[A2] + d2 : ft[A2];
    ft[A3] = lo ? ft[A3] + d3 : ft[A3];
    ft[A4] = lo ? ft[A4] + d4 : ft[A4];
    ft[B0] = lo ? ft[B0] : ft[B0] - d0;
    ft[B1] = lo ? ft[B1] : ft[B1] - d1;
    ft[B2] = lo ? ft[B2] : ft[B2] - d2;
    ft[B3] = lo ? ft[B3] : ft[B3] - d3;
    ft[B4] = lo ? ft[B4] : ft[B4] - d4;
}

template <typename Real>
__device__ inline void ladd(Real *ft, int dir, bool lo, Real rhow, Real un, Real ut1, Real ut2)
{
    if (dir == 0)  // tangentials v, w
        ladd_face<1, 7, 9, 11, 13, 2, 8, 10, 12, 14>(ft, lo, rhow, un, ut1, ut2);
    else if (dir == 1)  // tangentials u, w
        ladd_face<3, 7, 10, 15, 17, 4, 8, 9, 16, 18>(ft, lo, rhow, un, ut1, ut2);
    else  // tangentials u, v
        ladd_face<5, 11, 14, 15, 18, 6, 12, 13, 16, 17>(ft, lo, rhow, un, ut1, ut2);
}

template <typename Real>
struct Cell {
    long ic;
    int i, j, k;
};

// macro value of a neighbour: previous step's buffer; the cell itself: its new value
template <typename Real>
__device__ inline Real mac_at(const Real *old, long idx, long ic, Real self)
{
    return idx == ic ? self : old[idx];
}

// one component of grad(u) as kernel.cl:2275-2304 forms it
template <typename Real>
__device__ inline Real grad_comp(const Real *old, Real self, long ic, int i, int j, int k, int nx, int ny, int nz,
                                 int dir)
{
    const long sx = 1, sy = nx, sz = (long)nx * ny;
    const long base = ic - i * sx - j * sy - k * sz;
#define AT(ii, jj, kk) mac_at<Real>(old, base + (ii) * sx + (jj) * sy + (long)(kk) * sz, ic, self)
    if (dir == 0)
        return (i > 0 && i < nx - 1) ? Real(0.5) * (AT(i + 1, j, k) - AT(i - 1, j, k))
             : (i == 0) ? AT(1, j, k) - AT(0, j, k)
                        : AT(nx - 1, j, k) - AT(nx - 2, j, k);
    if (dir == 1)
        return (j > 0 && j < ny - 1) ? Real(0.5) * (AT(i, j + 1, k) - AT(i, j - 1, k))
             : (j == 0) ? AT(i, 1, k) - AT(i, 0, k)
                        : AT(i, ny - 1, k) - AT(i, ny - 2, k);
    if (k > 0 && k < nz - 1) return Real(0.5) * (AT(i, j, k + 1) - AT(i, j, k - 1));
    if (nz - 1 == 0) return Real(0);
    if (k == 0) {
        Real a = AT(i, j, 1);
        return a - a;
    }
    return AT(i, j, nz - 1) - AT(i, j, nz - 2);
#undef AT
}

// kernel.cl:2267-2365 waleTauSGS (Cw = 0.325) from the velocity gradient
template <typename Real>
__device__ inline Real wale_from_grad(Real dudx, Real dudy, Real dudz, Real dvdx, Real dvdy, Real dvdz, Real dwdx,
                                      Real dwdy, Real dwdz)
{
    const Real Cw = Real(0.325);
    Real S12 = Real(0.5) * (dudy + dvdx), S13 = Real(0.5) * (dudz + dwdx), S23 = Real(0.5) * (dvdz + dwdy);
    Real SS = dudx * dudx + Real(2) * S12 * S12 + Real(2) * S13 * S13 + dvdy * dvdy + Real(2) * S23 * S23
              + dwdz * dwdz;
    // g = grad(u)^2 (g_ab = sum_c d_c u_a ... as the reference composes it)
    Real g11 = dudx * dudx + dvdx * dudy + dwdx * dudz;
    Real g12 = dudx * dvdx + dvdx * dvdy + dwdx * dvdz;
    Real g13 = dudx * dwdx + dvdx * dwdy + dwdx * dwdz;
    Real g21 = dudy * dudx + dvdy * dudy + dwdy * dudz;
    Real g22 = dudy * dvdx + dvdy * dvdy + dwdy * dvdz;
    Real g23 = dudy * dwdx + dvdy * dwdy + dwdy * dwdz;
    Real g31 = dudz * dudx + dvdz * dudy + dwdz * dudz;
    Real g32 = dudz * dvdx + dvdz * dvdy + dwdz * dvdz;
    Real g33 = dudz * dwdx + dvdz * dwdy + dwdz * dwdz;
    Real tr3 = (g11 + g22 + g33) / Real(3);
    Real d11 = g11 - tr3, d22 = g22 - tr3, d33 = g33 - tr3;
    Real d12 = Real(0.5) * (g12 + g21), d13 = Real(0.5) * (g13 + g31), d23 = Real(0.5) * (g23 + g32);
    Real SdSd = d11 * d11 + d22 * d22 + d33 * d33 + Real(2) * (d12 * d12 + d13 * d13 + d23 * d23);
    if (!(SS > Real(1e-24) && SdSd > Real(1e-24))) return Real(0);
    Real sd15 = SdSd * sqrt(SdSd);          // SdSd^1.5
    Real ss25 = SS * SS * sqrt(SS);         // SS^2.5
    Real sd125 = SdSd * sqrt(sqrt(SdSd));        // SdSd^1.25
    return Real(9) * (Cw * Cw * sd15 / (ss25 + sd125));
}

// central difference 0.5 (p[o + d] - p[o - d]) at byte offset o (32-bit offsets from the
// array base; d = one element for x, so that it folds into the load's immediate offset)
template <typename Real>
__device__ inline Real cdiff(const Real *p, uint32_t o, uint32_t d)
{
    const char *b = (const char *)p;
    return Real(0.5) * (*(const Real *)(b + (size_t)(o + d)) - *(const Real *)(b + (size_t)(o - d)));
}

// Interior: take the central differences of interior cells directly (every step kernel;
// in the batch kernel it costs a wave per SIMD and still gains 2 %)
template <bool Interior, typename Real, int Q>
__device__ inline Real wale_tau(const StepArgs<Real, Q> &a, long ic, int i, int j, int k, Real u, Real v, Real w)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    if (Interior && Q == 27 && i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1) {
        // all six neighbours exist: central differences only, as grad_comp takes them there
        const uint32_t o = (uint32_t)ic * (uint32_t)sizeof(Real);
        const uint32_t dx = (uint32_t)sizeof(Real), dy = (uint32_t)nx * dx, dz = (uint32_t)ny * dy;
        const Real *U = a.u_o, *V = a.v_o, *W = a.w_o;
        return wale_from_grad<Real>(cdiff(U, o, dx), cdiff(U, o, dy), cdiff(U, o, dz), cdiff(V, o, dx), cdiff(V, o, dy),
                                    cdiff(V, o, dz), cdiff(W, o, dx), cdiff(W, o, dy), cdiff(W, o, dz));
    }
    if (Interior && Q == 19 && nx >= 3 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1) {
        // (D3Q19, whose waves of x-face and interior cells run together) the y and z
        // neighbours exist: central differences there, as grad_comp takes them; along x
        // central differences inside and, on an x face, grad_comp's one-sided difference with
        // the cell's own new velocity -- as a per-lane choice, so that such a wave runs one
        // path.  The x neighbours read on a face (the previous row's last or the next row's
        // first cell) exist, j being interior, and are not used there.
        const uint32_t o = (uint32_t)ic * (uint32_t)sizeof(Real);
        const uint32_t dx = (uint32_t)sizeof(Real), dy = (uint32_t)nx * dx, dz = (uint32_t)ny * dy;
        const Real *U = a.u_o, *V = a.v_o, *W = a.w_o;
        const bool xlo = (i == 0), xhi = (i == nx - 1);
        auto ddx = [&](const Real *P, Real self) {
            const char *b = (const char *)P;
            const Real p = *(const Real *)(b + (size_t)(o + dx)), m = *(const Real *)(b + (size_t)(o - dx));
            return xlo ? p - self : (xhi ? self - m : Real(0.5) * (p - m));
        };
        return wale_from_grad<Real>(ddx(U, u), cdiff(U, o, dy), cdiff(U, o, dz), ddx(V, v), cdiff(V, o, dy),
                                    cdiff(V, o, dz), ddx(W, w), cdiff(W, o, dy), cdiff(W, o, dz));
    }
    Real dudx = grad_comp<Real>(a.u_o, u, ic, i, j, k, nx, ny, nz, 0);
    Real dudy = grad_comp<Real>(a.u_o, u, ic, i, j, k, nx, ny, nz, 1);
    Real dudz = grad_comp<Real>(a.u_o, u, ic, i, j, k, nx, ny, nz, 2);
    Real dvdx = grad_comp<Real>(a.v_o, v, ic, i, j, k, nx, ny, nz, 0);
    Real dvdy = grad_comp<Real>(a.v_o, v, ic, i, j, k, nx, ny, nz, 1);
    Real dvdz = grad_comp<Real>(a.v_o, v, ic, i, j, k, nx, ny, nz, 2);
    Real dwdx = grad_comp<Real>(a.w_o, w, ic, i, j, k, nx, ny, nz, 0);
    Real dwdy = grad_comp<Real>(a.w_o, w, ic, i, j, k, nx, ny, nz, 1);
    Real dwdz = grad_comp<Real>(a.w_o, w, ic, i, j, k, nx, ny, nz, 2);
    return wale_from_grad<Real>(dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz);
}

// reciprocal: one v_rcp_f32 in FP32 (<= 1 ulp), an exact quotient in FP64
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double rcp(double x) { return 1.0 / x; }
// reciprocal refined by one Newton step (e = 1 - x r, r += r e): unbiased to within
// rounding, as a division would be.  Used for 1/rho, whose error the collision turns into
// momentum (the output momentum is rho times the velocity it was given), so that a
// one-sided error of the hardware reciprocal cannot accumulate over steps.
__device__ inline float rcp_nr(float x)
{
    const float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_fmaf(r, e, r);
}
__device__ inline double rcp_nr(double x) { return 1.0 / x; }
// square root: one v_sqrt_f32 in FP32 (<= 1 ulp), the correctly rounded one in FP64
__device__ inline float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline double fsqrt(double x) { return sqrt(x); }

// Density, momentum and the second moments sum_q c_a c_b f_q of the streamed
// populations; shared by the moments update, the Smagorinsky model and the collision.
template <typename Real>
struct Mom19 {
    Real rho, mx, my, mz;
    Real sxx, syy, szz, sxy, sxz, syz;
};

template <typename Real>
__device__ inline Mom19<Real> moments19(const Real *ft)
{
    Mom19<Real> m;
    const Real a12 = ft[1] + ft[2], a34 = ft[3] + ft[4], a56 = ft[5] + ft[6];
    const Real p78 = ft[7] + ft[8], p910 = ft[9] + ft[10], p1112 = ft[11] + ft[12], p1314 = ft[13] + ft[14];
    const Real p1516 = ft[15] + ft[16], p1718 = ft[17] + ft[18];
    m.rho = ft[0] + a12 + a34 + a56 + p78 + p910 + p1112 + p1314 + p1516 + p1718;
    const Real d78 = ft[7] - ft[8], d910 = ft[9] - ft[10], d1112 = ft[11] - ft[12], d1314 = ft[13] - ft[14];
    const Real d1516 = ft[15] - ft[16], d1718 = ft[17] - ft[18];
    m.mx = (ft[1] - ft[2]) + d78 + d910 + d1112 + d1314;
    m.my = (ft[3] - ft[4]) + d78 - d910 + d1516 + d1718;
    m.mz = (ft[5] - ft[6]) + d1112 - d1314 + d1516 - d1718;
    m.sxx = a12 + p78 + p910 + p1112 + p1314;
    m.syy = a34 + p78 + p910 + p1516 + p1718;
    m.szz = a56 + p1112 + p1314 + p1516 + p1718;
    m.sxy = p78 - p910;
    m.sxz = p1112 - p1314;
    m.syz = p1516 - p1718;
    return m;
}

// The same moments for populations held as 16-bit deviations from the weights: density
// and second moments are formed from the deviations f_q - w_q (exact differences, the
// populations being within a factor of two of their weights) plus the exact weight moments
// (sum w_q = 1, sum w_q c_a c_b = delta_ab / 3).  Summed as w_q + deviation, the density
// would carry the FP32 weights' excess sum_q fl(w_q) - 1 (about 2 ulp of 1) below every
// deviation a 16-bit value can hold, and its rounding to FP32 would go the same way in
// every cell and every step: a closed lattice would gain mass at about 1e-7 per cell per
// step, whatever the 16-bit rounding.  The momentum sums pairwise differences, as above.
template <typename Real>
__device__ inline Mom19<Real> moments19_dev(const Real *ft)
{
    Real d[19];
#pragma unroll
    for (int q = 0; q < 19; q++) d[q] = ft[q] - w19<Real>(q);
    Mom19<Real> m;
    const Real a12 = d[1] + d[2], a34 = d[3] + d[4], a56 = d[5] + d[6];
    const Real p78 = d[7] + d[8], p910 = d[9] + d[10], p1112 = d[11] + d[12], p1314 = d[13] + d[14];
    const Real p1516 = d[15] + d[16], p1718 = d[17] + d[18];
    m.rho = Real(1) + (d[0] + a12 + a34 + a56 + p78 + p910 + p1112 + p1314 + p1516 + p1718);
    const Real d78 = d[7] - d[8], d910 = d[9] - d[10], d1112 = d[11] - d[12], d1314 = d[13] - d[14];
    const Real d1516 = d[15] - d[16], d1718 = d[17] - d[18];
    m.mx = (d[1] - d[2]) + d78 + d910 + d1112 + d1314;
    m.my = (d[3] - d[4]) + d78 - d910 + d1516 + d1718;
    m.mz = (d[5] - d[6]) + d1112 - d1314 + d1516 - d1718;
    const Real third = Real(1) / Real(3);
    m.sxx = third + (a12 + p78 + p910 + p1112 + p1314);
    m.syy = third + (a34 + p78 + p910 + p1516 + p1718);
    m.szz = third + (a56 + p1112 + p1314 + p1516 + p1718);
    m.sxy = p78 - p910;
    m.sxz = p1112 - p1314;
    m.syz = p1516 - p1718;
    return m;
}

// van Driest wall damping of tauSGS (damping 2), the branch kernel.cl:2178-2197 keeps in
// comments beside the Fluent factor: y+ solves the log law y+ ln(E y+) = kappa U y / nu by
// Newton's method from y+ = 1 (stop when a step moves y+ by 1e-3 or less, or after 10
// steps; from y+ = 1 every iterate stays above 1/E, where 1 + ln(E y+) > 0), and the factor
// is min(1, (kappa / Cs)(1 + 1e-10 - exp(-y+ / A+)) y)^2, A+ = 26, E = 9.8,
// nu = (tau - 1/2) / 3.
template <typename Real>
__device__ inline Real van_driest(Real tau, Real U, Real y, Real Cs)
{
    const Real kappa = Real(0.41), Aplus = Real(26), E = Real(9.8);
    const Real nu = (tau - Real(0.5)) / Real(3);
    Real yp = Real(1), ypn;
    int it = 0;
    do {
        ypn = yp;
        yp = ((kappa * U * y / nu) + yp) / (Real(1) + log(E * yp));
        it++;
    } while (fabs(yp - ypn) > Real(0.001) && it < 10);
    const Real d = fmin(Real(1), (kappa / Cs) * ((Real(1) + Real(1e-10)) - exp(-yp / Aplus)) * y);
    return d * d;
}

// kernel.cl:2133-2201 smagorinskyTauSGS.  The reference sums c_a c_b (f_q - feq_q) over q
// with the incompressible feq (rho + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2) w_q, whose second
// moment is rho/3 delta_ab + u_a u_b; Pi_neq is formed from the shared moments instead.
template <typename Real>
__device__ inline Real smagorinsky_tau(const Mom19<Real> &m, Real u, Real v, Real w, Real tau, Real inv_rho,
                                       Real LES, Real sdf, Real Cs, int damping)
{
    const Real r3 = m.rho * (Real(1) / Real(3));
    const Real Pxx = m.sxx - r3 - u * u, Pyy = m.syy - r3 - v * v, Pzz = m.szz - r3 - w * w;
    const Real Pxy = m.sxy - u * v, Pxz = m.sxz - u * w, Pyz = m.syz - v * w;
    const Real PP = fsqrt(Pxx * Pxx + Pyy * Pyy + Pzz * Pzz + Real(2) * (Pxy * Pxy + Pxz * Pxz + Pyz * Pyz));
    Real t = LES * Real(0.5) *
             (-tau + fsqrt(tau * tau + Real(18) * Real(1.4142135623730951) * Cs * Cs * PP * inv_rho));
    if (damping == 1 && sdf != Real(10000) && sdf > Real(0)) {
        const Real d = fmin(Real(1), Real(0.41) * sdf / Cs);
        t *= d * d;
    } else if (damping == 2 && sdf != Real(10000) && sdf > Real(0)) {
        t *= van_driest<Real>(tau, sqrt(u * u + v * v + w * w), sdf, Cs);
    }
    return t;
}

// kernel.cl:2444-2503 tauSpongeZone
template <typename Real, int Q>
__device__ inline Real sponge_tau(const StepArgs<Real, Q> &a, Real tau, Real tauSGS, int i, int j, int k)
{
    const Real spz = a.spz;
    const int m = a.sponge_mask;
    if (m == 0) return tau;
    const Real one_m = Real(1) - tauSGS;
    if ((m & 1) && Real(i) < Real(a.nx) * spz) {
        const Real fx = Real(i) * a.inv_spx;
        tau = (Real(1) - fx) * one_m + fx * tau;
    }
    if ((m & 2) && Real(i) > Real(a.nx) * (Real(1) - spz)) {
        const Real fx = Real(a.nx - i) * a.inv_spx;
        tau = (Real(1) - fx) * one_m + fx * tau;
    }
    if ((m & 4) && Real(j) < Real(a.ny) * spz) {
        const Real fx = Real(j) * a.inv_spy;
        tau = (Real(1) - fx) * one_m + fx * tau;
    }
    if ((m & 8) && Real(j) > Real(a.ny) * (Real(1) - spz)) {
        const Real fx = Real(a.ny - j) * a.inv_spy;
        tau = (Real(1) - fx) * one_m + fx * tau;
    }
    if ((m & 16) && Real(k) < Real(a.nz) * spz) {
        const Real fx = Real(k) * a.inv_spz;
        tau = (Real(1) - fx) * one_m + fx * tau;
    }
    if ((m & 32) && Real(k) > Real(a.nz) * (Real(1) - spz)) {
        const Real fx = Real(a.nz - k) * a.inv_spz;
        tau = (Real(1) - fx) * one_m + fx * tau;
    }
    return tau;
}

// kernel.cl:3156-3309 recursive-regularized collision (bulk/shear split of the
// 2nd-order non-equilibrium moments, 3rd/4th order rebuilt recursively), with the
// body force S_q = rho w_q 3 c_qx dpdx.  u is the force-shifted velocity.
template <typename Real>
__device__ inline void collide_rr(Real *out, const Mom19<Real> &m, Real inv_rho, Real u, Real v, Real w,
                                  Real omegaEff, Real omegaB, Real dpdx)
{
    const Real cs2 = Real(1) / Real(3), cs4 = Real(1) / Real(9);
    const Real rho = m.rho;
    // second-order Hermite moments of ft per unit density
    const Real R200 = (m.sxx - cs2 * rho) * inv_rho;
    const Real R020 = (m.syy - cs2 * rho) * inv_rho;
    const Real R002 = (m.szz - cs2 * rho) * inv_rho;
    const Real R110 = m.sxy * inv_rho;
    const Real R101 = m.sxz * inv_rho;
    const Real R011 = m.syz * inv_rho;

    const Real E200 = u * u, E020 = v * v, E002 = w * w, E110 = u * v, E101 = u * w, E011 = v * w;
    const Real N200 = R200 - E200, N020 = R020 - E020, N002 = R002 - E002;
    const Real N110 = R110 - E110, N101 = R101 - E101, N011 = R011 - E011;
    const Real N210 = v * N200 + Real(2) * u * N110;
    const Real N201 = w * N200 + Real(2) * u * N101;
    const Real N021 = w * N020 + Real(2) * v * N011;
    const Real N120 = u * N020 + Real(2) * v * N110;
    const Real N102 = u * N002 + Real(2) * w * N101;
    const Real N012 = v * N002 + Real(2) * w * N011;
    const Real N220 = E020 * N200 + E200 * N020 + Real(4) * E110 * N110;
    const Real N202 = E002 * N200 + E200 * N002 + Real(4) * E101 * N101;
    const Real N022 = E002 * N020 + E020 * N002 + Real(4) * E011 * N011;

    const Real omegaM = (omegaB - omegaEff) * cs2;
    const Real omegaP = omegaM + omegaEff;
    const Real C200 = R200 - omegaP * N200 - omegaM * N020 - omegaM * N002;
    const Real C020 = R020 - omegaM * N200 - omegaP * N020 - omegaM * N002;
    const Real C002 = R002 - omegaM * N200 - omegaM * N020 - omegaP * N002;
    const Real C110 = R110 - omegaEff * N110;
    const Real C101 = R101 - omegaEff * N101;
    const Real C011 = R011 - omegaEff * N011;
    const Real om1 = Real(1) - omegaEff;
    const Real M210 = E200 * v + om1 * N210 + cs2 * v;
    const Real M201 = E200 * w + om1 * N201 + cs2 * w;
    const Real M021 = E020 * w + om1 * N021 + cs2 * w;
    const Real M120 = E020 * u + om1 * N120 + cs2 * u;
    const Real M102 = E002 * u + om1 * N102 + cs2 * u;
    const Real M012 = E002 * v + om1 * N012 + cs2 * v;
    const Real M220 = E200 * E020 + om1 * N220 + cs2 * (C200 + C020) + cs4;
    const Real M202 = E200 * E002 + om1 * N202 + cs2 * (C200 + C002) + cs4;
    const Real M022 = E020 * E002 + om1 * N022 + cs2 * (C020 + C002) + cs4;
    const Real M200 = C200 + cs2, M020 = C020 + cs2, M002 = C002 + cs2;
    const Real M110 = C110, M101 = C101, M011 = C011;

    const Real h = Real(0.5) * rho, qr = Real(0.25) * rho;
    const Real F1 = rho * w19<Real>(1) * Real(3) * dpdx;   // force on axis directions
    const Real F7 = rho * w19<Real>(7) * Real(3) * dpdx;   // force on diagonals
    const Real ax = h * (u + M200 - M120 - M102 - M220 - M202);
    const Real ay = h * (v + M020 - M210 - M012 - M220 - M022);
    const Real az = h * (w + M002 - M201 - M021 - M202 - M022);
    out[0] = rho * (Real(1) - M200 - M020 - M002 + M220 + M202 + M022);
    out[1] = ax + F1;
    out[2] = rho * (-u + M120 + M102) + ax - F1;
    out[3] = ay;
    out[4] = rho * (-v + M210 + M012) + ay;
    out[5] = az;
    out[6] = rho * (-w + M201 + M021) + az;
    const Real dxy = qr * (M110 + M210 + M120 + M220);
    const Real dxz = qr * (M101 + M201 + M102 + M202);
    const Real dyz = qr * (M011 + M021 + M012 + M022);
    out[7] = dxy + F7;
    out[8] = h * (-M210 - M120) + dxy - F7;
    out[9] = h * (-M110 - M210) + dxy + F7;
    out[10] = h * (-M110 - M120) + dxy - F7;
    out[11] = dxz + F7;
    out[12] = h * (-M201 - M102) + dxz - F7;
    out[13] = h * (-M101 - M201) + dxz + F7;
    out[14] = h * (-M101 - M102) + dxz - F7;
    out[15] = dyz;
    out[16] = h * (-M021 - M012) + dyz;
    out[17] = h * (-M011 - M021) + dyz;
    out[18] = h * (-M011 - M012) + dyz;
}

// Boundary handling of a domain-face cell: streaming with walls, then the face
// conditions in the reference's order (kernel.cl:3581-3592).
// xonly = false: every direction is streamed here.  xonly = true: a cell on an x face (and
// on no other face) whose directions were pulled from the interior offsets already; only
// the directions that cross its face (c_x = +1 at i = 0, c_x = -1 at i = nx - 1) are
// streamed again.  The face conditions follow in either case.
template <typename S, typename Real>
__device__ __forceinline__ void face_cell(const Step19Args<Real> &a, Real *ft, uint32_t fl, long ic, int i, int j, int k,
                                       Real omega, bool xonly)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const S *fs = (const S *)a.fsrc;
    const int b1 = flag_code(fl, F_B1_SHIFT), b2 = flag_code(fl, F_B2_SHIFT), b3 = flag_code(fl, F_B3_SHIFT);
    const bool sym = (b1 == BC_SYMMETRY || b2 == BC_SYMMETRY || b3 == BC_SYMMETRY);
    Real Fx = 0, Fy = 0, Fz = 0;  // wall momentum of this step's domain-face links
    const int xs = (i == 0) ? 1 : -1;
    if (!xonly) ft[0] = Codec<S, Real, 19>::load(fs + pop_at<19, S>(0, (uint32_t)ic, a.fstride), 0, w19<Real>(0));
#pragma unroll
    for (int q = 1; q < 19; q++) {
        if (xonly && L19::cx[q] != xs) continue;
        long src = face_upwind(q, ic, i, j, k, nx, ny, nz, b1, b2, b3);
        if (src >= 0) {
            ft[q] = Codec<S, Real, 19>::load(fs + pop_at<19, S>(q, (uint32_t)src, a.fstride), q, w19<Real>(q));
        } else {
            int qb = sym ? mirror_q(q, b1, b2, b3) : L19::opp(q);
            ft[q] = Codec<S, Real, 19>::load(fs + pop_at<19, S>(qb, (uint32_t)ic, a.fstride), qb, w19<Real>(qb));
            Fx += Real(-2) * ft[q] * Real(L19::cx[q]);
            Fy += Real(-2) * ft[q] * Real(L19::cy[q]);
            Fz += Real(-2) * ft[q] * Real(L19::cz[q]);
        }
    }
    const int corner = corner_flag(b3, i, j, k, nx, ny, nz);
    const long sy = nx, sz = (long)nx * ny;

    // fixedVelocityBC, kernel.cl:1397-1512
    if (corner == 0 && (b1 == BC_FIXEDVELOCITY || b2 == BC_FIXEDVELOCITY || b3 == BC_FIXEDVELOCITY)) {
        const Real u0 = a.u0[ic], v0 = a.v0[ic], w0 = a.w0[ic];
        if (b1 == BC_FIXEDVELOCITY && (i == 0 || i == nx - 1)) {
            const bool lo = (i == 0);
            Real rhow = zouhe_sum(ft, 0, lo) / (lo ? Real(1) - u0 : Real(1) + u0);
            if (lo && u0 < Real(0)) rhow = a.rho_o[ic + 1];
            if (!lo && u0 > Real(0)) rhow = a.rho_o[ic];  // own stored density (K19:1424-1428)
            ladd(ft, 0, lo, rhow, u0, v0, w0);
        }
        if (b2 == BC_FIXEDVELOCITY && (j == 0 || j == ny - 1)) {
            const bool lo = (j == 0);
            Real rhow = zouhe_sum(ft, 1, lo) / (lo ? Real(1) - v0 : Real(1) + v0);
            if (lo && v0 < Real(0)) rhow = a.rho_o[ic + sy];
            if (!lo && v0 > Real(0)) rhow = a.rho_o[ic - sy];
            ladd(ft, 1, lo, rhow, v0, u0, w0);
        }
        if (b3 == BC_FIXEDVELOCITY && (k == 0 || k == nz - 1)) {
            const bool lo = (k == 0);
            Real rhow = zouhe_sum(ft, 2, lo) / (lo ? Real(1) - w0 : Real(1) + w0);
            if (lo && w0 < Real(0)) rhow = a.rho_o[ic + sz];
            if (!lo && w0 > Real(0)) rhow = a.rho_o[ic - sz];
            ladd(ft, 2, lo, rhow, w0, u0, v0);
        }
    }

    // fixedDensityBC with rho_w = 1, kernel.cl:1514-1611
    if (corner == 0 && (b1 == BC_FIXEDDENSITY || b2 == BC_FIXEDDENSITY || b3 == BC_FIXEDDENSITY)) {
        const Real uc = a.u_o[ic], vc = a.v_o[ic], wc = a.w_o[ic];
#pragma unroll
        for (int dir = 0; dir < 3; dir++) {
            const int code = dir == 0 ? b1 : (dir == 1 ? b2 : b3);
            if (code != BC_FIXEDDENSITY) continue;
            const int pos = dir == 0 ? i : (dir == 1 ? j : k);
            const int n = dir == 0 ? nx : (dir == 1 ? ny : nz);
            const long st = dir == 0 ? 1 : (dir == 1 ? sy : sz);
            // inner neighbour: index 1 at the min face, n-2 otherwise (kernel.cl:1520, 1560, 1590)
            const long inner = (pos == 0) ? ic + st : ic + (long)(n - 2 - pos) * st;
            const Real ue = uc + Real(0.5) * (uc - a.u_o[inner]);
            const Real ve = vc + Real(0.5) * (vc - a.v_o[inner]);
            const Real we = wc + Real(0.5) * (wc - a.w_o[inner]);
            const Real uSqr = ue * ue + ve * ve + we * we;
            const bool atlo = (pos == 0), athi = (pos == n - 1);
#pragma unroll
            for (int q = 1; q < 19; q++) {
                const int cn = dir == 0 ? L19::cx[q] : (dir == 1 ? L19::cy[q] : L19::cz[q]);
                if ((atlo && cn == 1) || (athi && cn == -1)) {
                    const Real uDotC = ue * Real(L19::cx[q]) + ve * Real(L19::cy[q]) + we * Real(L19::cz[q]);
                    ft[q] = -ft[q] + Real(2) * w19<Real>(q) * Real(1)
                                         * (Real(1) + Real(4.5) * uDotC * uDotC - Real(1.5) * uSqr);
                }
            }
        }
    }

    // wallFunctionBC, kernel.cl:1654-1924 (stationary wall, tangential slip velocity)
    if (corner == 0 && (b1 == BC_WALLFUNCTION || b2 == BC_WALLFUNCTION || b3 == BC_WALLFUNCTION)) {
        const Real nu = (Real(1) / omega - Real(0.5)) / Real(3);
#pragma unroll
        for (int dir = 0; dir < 3; dir++) {
            const int code = dir == 0 ? b1 : (dir == 1 ? b2 : b3);
            const int pos = dir == 0 ? i : (dir == 1 ? j : k);
            const int n = dir == 0 ? nx : (dir == 1 ? ny : nz);
            if (code != BC_WALLFUNCTION || !(pos == 0 || pos == n - 1)) continue;
            const bool lo = (pos == 0);
            const long st = dir == 0 ? 1 : (dir == 1 ? sy : sz);
            const Real rhow = zouhe_sum(ft, dir, lo);
            // neighbour sampled for the wall law: the next cell inward; the z faces of a
            // 2-D lattice sample the cell itself (K19:1848)
            long nb = lo ? ic + st : ic - st;
            if (dir == 2) nb = (nz - 1 != 0) ? ic + st : ic;
            const Real *t1 = dir == 0 ? a.v_o : a.u_o;
            const Real *t2 = dir == 2 ? a.v_o : a.w_o;
            Real p = t1[nb], s = t2[nb];
            Real magUp = sqrt(p * p + s * s);
            Real ut1 = 0, ut2 = 0;
            if (magUp != Real(0)) {
                const Real tauw = calc_tauw<Real>(rhow, magUp, nu);
                p = t1[ic];
                s = t2[ic];
                magUp = sqrt(p * p + s * s);
                if ((dir == 2 && lo) || magUp != Real(0)) {
                    const Real e1 = p / magUp, e2 = s / magUp;
                    const Real Fw1 = dir == 0 ? Fy : Fx;
                    const Real Fw2 = dir == 2 ? Fy : Fz;
                    ut1 = Real(-3) * (tauw * e1 - Fw1) / rhow;
                    ut2 = Real(-3) * (tauw * e2 - Fw2) / rhow;
                }
            }
            ladd(ft, dir, lo, rhow, Real(0), ut1, ut2);
        }
    }

    // outflowBC, kernel.cl:2005-2097: convective outlet with the cell's last velocity
    if (b1 == BC_OUTLET || b2 == BC_OUTLET || b3 == BC_OUTLET) {
        const Real is3 = Real(0.57735026918962576);
        const Real uo = a.u_o[ic], vo = a.v_o[ic], wo = a.w_o[ic];
#pragma unroll
        for (int q = 1; q < 19; q++) {
            const int cx = L19::cx[q], cy = L19::cy[q], cz = L19::cz[q];
#define OUTF(cond, inner, un)                                                                          \
    if (cond) {                                                                                        \
        const Real fi = Codec<S, Real, 19>::load(fs + pop_at<19, S>(q, (uint32_t)(inner), a.fstride), q, w19<Real>(q)); \
        const Real fc = Codec<S, Real, 19>::load(fs + pop_at<19, S>(q, (uint32_t)ic, a.fstride), q, w19<Real>(q)); \
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

// IBB selects the interpolated bounce-back at compile time: its loop adds registers to the
// whole kernel, which the kernels without it should not carry.  AA: 0 two-lattice streaming, 1 and 2
// the two half-steps of in-place streaming (the cell body describes them).  PASS: 0 every
// cell; 1 every cell but the lattice-face cells; 2 the face cells of the list a.fcells.  A
// wave that holds face and interior cells runs both the face path and the interior pull;
// two passes keep them apart (TUNE_SPLIT runs 1 then 2; both read only the previous
// step's arrays).  It pays where such waves are many and the step is bound by instruction
// issue, and costs a second launch's drain where the step is bound by HBM.
// PASS 3 runs both passes in one grid: the first face_blocks(a) work-groups take the face
// cells of the list, the others the cells off the faces, so that no wave runs both paths
// and no second launch waits for the first to drain (the face work-groups are dispatched
// first and run beside the interior ones).
// (rounded up to a multiple of 8, so that the blocks after them keep their hardware XCD
// order; the work-groups past the list end at once)
template <typename A>
__device__ inline uint32_t face_blocks(const A &a)
{
    return merged_face_blocks(a.nface, blockDim.x);
}

// Waves per SIMD the step kernel is held to: eight (64 VGPRs), which the compiler reaches
// without spilling by interleaving each direction's pull address with its load; seven for
// the FP32 WALE kernel with interpolated bounce-back, which spills at eight; FP64 as the
// compiler chooses.
template <typename S, typename Real, int SGS, bool IBB>
constexpr int kStep19MinWaves = sizeof(Real) == 8 ? 1 : ((IBB && SGS == 0 && sizeof(S) == 4) ? 7 : 8);

template <typename S, typename Real, int SGS, bool IBB = false, int AA = 0, int PASS = 0>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kStep19MinWaves<S, Real, SGS, IBB>, 8)))
step19_kernel(const Step19Args<Real> a)
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
#define CELL_IBB IBB
#define CELL_WALE_INTERIOR true
#define CELL_AA AA
#define CELL_SKIP_FACE (PASS == 1 || (PASS == 3 && !face_wg))
#define CELL_PURE_OK (AA == 0 && !face_wg)
#define CELL_ALIGNED (!face_wg && a.wave_runs)
#include "d3q19_cell.inc"
#undef CELL_ALIGNED
#undef CELL_PURE_OK
#undef CELL_SKIP_FACE
#undef CELL_AA
#undef CELL_WALE_INTERIOR
#undef CELL_IBB
#undef CELL_RETURN
}

// In-place streaming: after the first half-step, out[p] of cell x sits in slot opp(p) of
// x + c_p and out[opp(p)] of x + c_p in slot p of x, for every link between two fluid cells
// that is not a wall link.  Exchanging the two restores the two-lattice layout (and
// exchanging again undoes it); cell x exchanges its odd directions p (opp(p) = p + 1), so
// every pair is exchanged once.  Links across a wall or to a solid cell already hold their
// value in the cell's own slot p.
template <typename S>
__global__ void aa_swap19_kernel(uint32_t N, int nx, int ny, int nz, FastDiv dnx, FastDiv dny,
                                 const uint16_t *__restrict__ flag, size_t FS, S *f)
{
    const uint32_t icu = blockIdx.x * blockDim.x + threadIdx.x;
    if (icu >= N) return;
    const uint32_t fl = flag[icu];
    if (fl & F_SOLID) return;
    const uint32_t t = fdiv(icu, dnx);
    const int i = (int)(icu - t * (uint32_t)nx);
    const uint32_t kk = fdiv(t, dny);
    const int j = (int)(t - kk * (uint32_t)ny);
    const int k = (int)kk;
    const int b1 = flag_code(fl, F_B1_SHIFT), b2 = flag_code(fl, F_B2_SHIFT), b3 = flag_code(fl, F_B3_SHIFT);
    const long sy = nx, sz = (nz == 1) ? 0 : (long)nx * ny;
#pragma unroll
    for (int p = 1; p < 19; p += 2) {
        const int po = p + 1;
        // the cell x + c_p: the pull source of direction opp(p)
        const long nb = (fl & F_FACE) ? face_upwind(po, (long)icu, i, j, k, nx, ny, nz, b1, b2, b3)
                                      : (long)icu + L19::cx[p] + L19::cy[p] * sy + L19::cz[p] * sz;
        if (nb < 0 || (flag[nb] & F_SOLID)) continue;
        S *x = f + pop_at<19, S>(p, icu, FS);
        S *y = f + pop_at<19, S>(po, (uint32_t)nb, FS);
        const S tx = *x;
        *x = *y;
        *y = tx;
    }
}

}  // namespace lbffs
