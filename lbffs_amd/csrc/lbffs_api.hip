// lbffs_api.hip -- the C ABI of the engine (include/lbffs.h): device context, field
// transfer, flag packing with input validation, the step launcher and FP64
// reductions.  Replaces the OpenCL host section of D3Q19_opencl/ns.cpp:382-603.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <memory>
#include <mutex>

#include "../../include/lbffs.h"
#include "../../include/lbffs_geom.h"
#include "d3q19_step.hpp"
#include "d3q27_step.hpp"
#include "ibm.hpp"
#include "launch.hpp"
#include "geom.hpp"

using namespace lbffs;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return fail(LBFFS_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

// validation failures reported by the pack kernel (bit set in a device word)
enum : int {
    PACK_BAD_CODE = 1,         // face code outside 0..7
    PACK_FACE_NOT_FLAGGED = 2, // lattice-face fluid cell with boundary==0 (its pulls would leave the lattice)
    PACK_WALLFN_ZMAX = 4,      // WallFunction on the nzMax face: the reference samples k+1 there
    PACK_2D_ZCODE = 8,         // nz==1 with a z-face condition that samples the layer k=1
    PACK_BAD_BOUNDARY = 16,    // boundary value other than 0/1
    PACK_NEISOLID_FACE = 32,   // D3Q27: neiSolid on a wall face (the reference reads solid[-1] there)
    PACK_AA_FACE = 64,         // in-place streaming: a face code other than Cyclic/BounceBack,
                               // or different codes on the two sides of an axis
    PACK_AA_NEISOLID = 128,    // in-place streaming: an interior fluid cell next to a solid one
                               // without neiSolid (its pulls from the solid cell have no place)
};

}  // namespace

struct lbffs_ctx {
    lbffs_desc d;
    long N = 0;
    size_t felems = 0;   // elements of one population array on the device (pop_elems)
    size_t fstride = 0;  // 16-bit values: elements between two direction arrays (pop_at)
    int real_size = 4;   // bytes of the host/compute real type
    int store_size = 4;  // bytes of one stored population
    void *f[2] = {nullptr, nullptr};  // one array (f[0]) with in-place streaming
    int cur = 0;
    int aa_odd = 0;  // in-place streaming: the array is in the layout of a first half-step
    uint32_t *fcells = nullptr;  // non-solid face cells (the face pass), rebuilt with the flags
    uint32_t nface = 0;
    uint32_t *wclass = nullptr;  // per-wave summary of the flags (wave_class), rebuilt with them
    uint64_t flags_gen = 0;      // bumped on every rebuild of the flags (a batch keeps its own copy)
    int device = 0;              // the HIP device the context was created on
    uint32_t nstep = 0;          // steps run (the step number of the stochastic rounding)
    int32_t *boundary = nullptr, *b1 = nullptr, *b2 = nullptr, *b3 = nullptr;
    uint8_t *solid = nullptr, *neiSolid = nullptr;
    void *sdf = nullptr, *u0 = nullptr, *v0 = nullptr, *w0 = nullptr, *omega = nullptr;
    void *Fwx = nullptr, *Fwy = nullptr, *Fwz = nullptr, *tau = nullptr;
    void *qlink = nullptr;  // q-distances of the solid links (Q*N reals), allocated on upload
    void *Gx = nullptr, *Gy = nullptr, *Gz = nullptr;
    void *rho[2] = {nullptr, nullptr}, *u[2] = {nullptr, nullptr}, *v[2] = {nullptr, nullptr},
         *w[2] = {nullptr, nullptr};
    int mcur = 0;
    uint16_t *flag = nullptr;
    int *d_err = nullptr;
    double *d_part = nullptr;  // reduction partials
    double *d_red = nullptr;
    bool flags_dirty = true;
    bool f_loaded = false;
    int sponge_mask = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = true;  // false for the lattices of a batch, which share its stream
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    int block = 256;  // work-group size of the step launch
    int tune = 0;     // TUNE_* launch options
    // immersed-boundary moving walls (active with more than one marker, ns.cpp:544)
    int n_markers = 0;
    void *markers = nullptr;   // 3n host reals, xyz interleaved
    void *Gm = nullptr;        // 3n host reals: G_x[n], G_y[n], G_z[n]
    uint32_t *mark = nullptr;  // per-cell stamp of the last spread that touched it
    int *list = nullptr;       // cells touched by the last spread (capacity 8n)
    int *count = nullptr;
    uint32_t stamp = 0;
    double motion[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool g_dirty = true;       // G may hold values outside the last list
};

namespace {

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

}  // namespace

// Scratch of lbffs_reduce_noalloc: the per-block partial sums of the FP64 reduction
// (kRedBlocks x LBFFS_REDUCE_N doubles) followed by the final sums (LBFFS_REDUCE_N), 64 KiB.
// Each GPU holds its own copy at its own address, kept until the process ends; host code
// never uses this name as an address -- it uses the pointer hipGetSymbolAddress gives for
// the device (scratch_ptr below).
static __device__ __align__(256) double scratch[kRedBlocks * LBFFS_REDUCE_N + LBFFS_REDUCE_N];

namespace {

constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU, set once
std::mutex g_scratch_mu;

// ---------------------------------------------------------------- kernels

template <typename Real>
__global__ void pack_flags_kernel(int Q, uint32_t N, int nx, int ny, int nz, FastDiv dnx, FastDiv dny,
                                  const int32_t *boundary, const int32_t *b1, const int32_t *b2, const int32_t *b3,
                                  const uint8_t *solid, const uint8_t *neiSolid, const Real *sdf, const Real *omega,
                                  Real omega0, const Real *w0, int aa, uint16_t *flag, int *err)
{
    uint32_t ic = blockIdx.x * blockDim.x + threadIdx.x;
    if (ic >= N) return;
    uint32_t t = fdiv(ic, dnx);
    int i = (int)(ic - t * (uint32_t)nx);
    uint32_t kk = fdiv(t, dny);
    int j = (int)(t - kk * (uint32_t)ny);
    int k = (int)kk;
    int e = 0;
    uint32_t fl = 0;
    const int c1 = b1[ic], c2 = b2[ic], c3 = b3[ic];
    const int bnd = boundary[ic];
    const bool sol = solid[ic] != 0;
    if (c1 < 0 || c1 > 7 || c2 < 0 || c2 > 7 || c3 < 0 || c3 > 7) e |= PACK_BAD_CODE;
    if (bnd != 0 && bnd != 1) e |= PACK_BAD_BOUNDARY;
    const bool geoface = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || (nz > 1 && (k == 0 || k == nz - 1)));
    if (sol) fl |= F_SOLID;
    if (bnd == 1) fl |= F_FACE;
    if (!sol) {
        if (geoface && bnd != 1) e |= PACK_FACE_NOT_FLAGGED;
        if (bnd == 1) {
            if (nz > 1 && k == nz - 1 && c3 == BC_WALLFUNCTION) e |= PACK_WALLFN_ZMAX;
            if (nz == 1 && (c3 == BC_OUTLET || c3 == BC_FIXEDDENSITY || (c3 == BC_FIXEDVELOCITY && w0[ic] < Real(0))))
                e |= PACK_2D_ZCODE;
        }
    }
    if (aa && !sol && bnd == 1) {
        // in-place streaming classifies a link from both of its cells: the faces of an axis
        // must agree, and only pulls that wrap or bounce back exist
        const bool okc = (c1 == BC_CYCLIC || c1 == BC_BOUNCEBACK) && (c2 == BC_CYCLIC || c2 == BC_BOUNCEBACK) &&
                         (c3 == BC_CYCLIC || c3 == BC_BOUNCEBACK);
        const long row = (long)ic - i, col = (long)i + (long)nx * ((long)ny * k);
        const long pl = (long)i + (long)nx * j;
        bool sym = true;
        if (i == 0 || i == nx - 1) sym = sym && b1[row + (nx - 1 - i)] == c1;
        if (j == 0 || j == ny - 1) sym = sym && b2[col + (long)nx * (ny - 1 - j)] == c2;
        if (nz > 1 && (k == 0 || k == nz - 1)) sym = sym && b3[pl + (long)nx * ny * (nz - 1 - k)] == c3;
        if (!okc || !sym) e |= PACK_AA_FACE;
    }
    if (aa && !sol && bnd != 1 && neiSolid[ic] != 1) {
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const int ii = i + dx, jj = j + dy, kz = (nz == 1) ? k : k + dz;
                    if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kz < 0 || kz >= nz) continue;
                    if (solid[(long)ii + (long)nx * ((long)jj + (long)ny * kz)]) e |= PACK_AA_NEISOLID;
                }
    }
    if (neiSolid[ic] == 1) fl |= F_NEISOLID;
    bool keep = (i <= 1 || i >= nx - 2 || j <= 1 || j >= ny - 2 || (nz > 1 && (k <= 1 || k >= nz - 2)));
    if (Q == 27) {
        // a NEISOLID cell tests solid[upID[q]] for all 26 links; on a wall face one of
        // them is -1 in the reference
        const bool on_wall = ((i == 0 || i == nx - 1) && is_wall_code(c1)) || ((j == 0 || j == ny - 1) && is_wall_code(c2)) ||
                             ((k == 0 || k == nz - 1) && is_wall_code(c3));
        if (!sol && bnd == 1 && neiSolid[ic] == 1 && on_wall) e |= PACK_NEISOLID_FACE;
    }
    // the interpolated bounce-back (D3Q27 always, D3Q19 with ibb = 1) reads the
    // previous-step moments of a NEISOLID cell and of its neighbours: keep them stored
    for (int dz = -1; dz <= 1 && !keep; dz++)
        for (int dy = -1; dy <= 1 && !keep; dy++)
            for (int dx = -1; dx <= 1 && !keep; dx++) {
                const int ii = i + dx, jj = j + dy, kz = k + dz;
                if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kz < 0 || kz >= nz) continue;
                if (neiSolid[(long)ii + (long)nx * ((long)jj + (long)ny * kz)] == 1) keep = true;
            }
    if (keep) fl |= F_KEEPMAC;
    if (sdf[ic] != Real(10000)) fl |= F_HASSDF;
    if (omega[ic] != omega0) fl |= F_OMEGAVAR;
    fl |= ((uint32_t)(c1 & 7) << F_B1_SHIFT) | ((uint32_t)(c2 & 7) << F_B2_SHIFT) | ((uint32_t)(c3 & 7) << F_B3_SHIFT);
    flag[ic] = (uint16_t)fl;
    if (e) atomicOr(err, e);
}

// host-real populations (q-major arrays of N) -> storage form in the device layout (pop_at)
template <typename S, typename Real, int Q>
__global__ void encode_kernel(uint32_t N, size_t fs, const Real *src, S *dst)
{
    uint32_t ic = blockIdx.x * blockDim.x + threadIdx.x;
    if (ic >= N) return;
#pragma unroll
    for (int q = 0; q < Q; q++)
        Codec<S, Real, Q>::store(dst + pop_at<Q, S>(q, ic, fs), src[(size_t)q * N + ic], q, wq<Real, Q>(q));
}

template <typename S, typename Real, int Q>
__global__ void decode_kernel(uint32_t N, size_t fs, const S *src, Real *dst)
{
    uint32_t ic = blockIdx.x * blockDim.x + threadIdx.x;
    if (ic >= N) return;
#pragma unroll
    for (int q = 0; q < Q; q++)
        dst[(size_t)q * N + ic] = Codec<S, Real, Q>::load(src + pop_at<Q, S>(q, ic, fs), q, wq<Real, Q>(q));
}

// density and velocity from the current populations, as the host computes them at an
// output step (cal_rhoUVW of D3Q19_opencl/D3Q19.hpp:647-666 and D3Q27_opencl/D3Q27.hpp:28-47):
// FP32 sums in direction order of f c (the products by 0 kept: they decide the sign of a
// zero), u /= rho, then u += (0.5 dpdx) / rho
template <typename S, typename Real, int Q>
__global__ void output_macros_kernel(uint32_t N, size_t fs, const S *f, Real dpdx, Real *rho, Real *u, Real *v,
                                     Real *w)
{
#pragma clang fp contract(off)
    const uint32_t ic = blockIdx.x * blockDim.x + threadIdx.x;
    if (ic >= N) return;
    Real r = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const Real fq = Codec<S, Real, Q>::load(f + pop_at<Q, S>(q, ic, fs), q, wq<Real, Q>(q));
        const Real cx = Real(Q == 27 ? L27::cx[q] : L19::cx[q]);
        const Real cy = Real(Q == 27 ? L27::cy[q] : L19::cy[q]);
        const Real cz = Real(Q == 27 ? L27::cz[q] : L19::cz[q]);
        r += fq;
        mx += fq * cx;
        my += fq * cy;
        mz += fq * cz;
    }
    mx /= r;
    my /= r;
    mz /= r;
    mx += Real(0.5) * dpdx / r;
    rho[ic] = r;
    u[ic] = mx;
    v[ic] = my;
    w[ic] = mz;
}

__device__ inline double wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    return x;
}

// pass 1: fixed grid, grid-stride over cells, per-block partial sums (deterministic order)
template <typename S, typename Real, int Q>
__global__ void __launch_bounds__(kRedThreads) reduce_kernel(uint32_t N, size_t fs, const uint16_t *flag, const S *f,
                                                             const Real *Fwx, const Real *Fwy, const Real *Fwz,
                                                             const Real *Gx, const Real *Gy, const Real *Gz,
                                                             double *part)
{
    double acc[LBFFS_REDUCE_N] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t ic = blockIdx.x * blockDim.x + threadIdx.x; ic < N; ic += gridDim.x * blockDim.x) {
        if (flag[ic] & F_SOLID) continue;
        acc[0] += (double)Fwx[ic];
        acc[1] += (double)Fwy[ic];
        acc[2] += (double)Fwz[ic];
        acc[3] += (double)Gx[ic];
        acc[4] += (double)Gy[ic];
        acc[5] += (double)Gz[ic];
        // summed in double: each decoded population is exact, and an FP32 sum near 1 would
        // round every cell's density alike (the weights' FP32 excess lies below its ulp)
        double rho = 0;
#pragma unroll
        for (int q = 0; q < Q; q++)
            rho += (double)Codec<S, Real, Q>::load(f + pop_at<Q, S>(q, ic, fs), q, wq<Real, Q>(q));
        acc[6] += rho - 1.0;
        acc[7] += 1.0;
    }
    __shared__ double sh[kRedThreads / 64][LBFFS_REDUCE_N];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < LBFFS_REDUCE_N; c++) {
        double s = wave_sum(acc[c]);
        if (lane == 0) sh[wid][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < LBFFS_REDUCE_N) {
        double s = 0;
        for (int wv = 0; wv < kRedThreads / 64; wv++) s += sh[wv][threadIdx.x];
        part[blockIdx.x * LBFFS_REDUCE_N + threadIdx.x] = s;
    }
}

// pass 2: one block sums the partials in block order
__global__ void reduce_final_kernel(const double *part, int nblocks, double *out)
{
    if (threadIdx.x < LBFFS_REDUCE_N) {
        double s = 0;
        for (int b = 0; b < nblocks; b++) s += part[b * LBFFS_REDUCE_N + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------- helpers

bool is_f64(const lbffs_ctx *c) { return c->d.storage == LBFFS_STORE_F64; }

size_t field_host_bytes(const lbffs_ctx *c, int field)
{
    const size_t N = (size_t)c->N;
    switch (field) {
    case LBFFS_FIELD_F:
    case LBFFS_FIELD_QLINK: return (size_t)c->d.q * N * c->real_size;
    case LBFFS_FIELD_BOUNDARY:
    case LBFFS_FIELD_BOUNDARY1:
    case LBFFS_FIELD_BOUNDARY2:
    case LBFFS_FIELD_BOUNDARY3: return N * sizeof(int32_t);
    case LBFFS_FIELD_SOLID:
    case LBFFS_FIELD_NEISOLID: return N;
    default: return N * c->real_size;
    }
}

void *field_ptr(lbffs_ctx *c, int field)
{
    switch (field) {
    case LBFFS_FIELD_BOUNDARY: return c->boundary;
    case LBFFS_FIELD_BOUNDARY1: return c->b1;
    case LBFFS_FIELD_BOUNDARY2: return c->b2;
    case LBFFS_FIELD_BOUNDARY3: return c->b3;
    case LBFFS_FIELD_SDF: return c->sdf;
    case LBFFS_FIELD_SOLID: return c->solid;
    case LBFFS_FIELD_NEISOLID: return c->neiSolid;
    case LBFFS_FIELD_U0: return c->u0;
    case LBFFS_FIELD_V0: return c->v0;
    case LBFFS_FIELD_W0: return c->w0;
    case LBFFS_FIELD_FWX: return c->Fwx;
    case LBFFS_FIELD_FWY: return c->Fwy;
    case LBFFS_FIELD_FWZ: return c->Fwz;
    case LBFFS_FIELD_TAUSGS: return c->tau;
    case LBFFS_FIELD_RHO: return c->rho[c->mcur];
    case LBFFS_FIELD_U: return c->u[c->mcur];
    case LBFFS_FIELD_V: return c->v[c->mcur];
    case LBFFS_FIELD_W: return c->w[c->mcur];
    case LBFFS_FIELD_QLINK: return c->qlink;
    case LBFFS_FIELD_OMEGA: return c->omega;
    case LBFFS_FIELD_GX: return c->Gx;
    case LBFFS_FIELD_GY: return c->Gy;
    case LBFFS_FIELD_GZ: return c->Gz;
    }
    return nullptr;
}

bool field_affects_flags(int field)
{
    switch (field) {
    case LBFFS_FIELD_BOUNDARY:
    case LBFFS_FIELD_BOUNDARY1:
    case LBFFS_FIELD_BOUNDARY2:
    case LBFFS_FIELD_BOUNDARY3:
    case LBFFS_FIELD_SDF:
    case LBFFS_FIELD_SOLID:
    case LBFFS_FIELD_NEISOLID:
    case LBFFS_FIELD_OMEGA:
    case LBFFS_FIELD_W0: return true;
    }
    return false;
}

inline unsigned grid_for(long n, int block) { return (unsigned)((n + block - 1) / block); }

// the device address of `scratch` on `device` (the current device), looked up on the first
// call for that device
int scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(LBFFS_E_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return 0;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(LBFFS_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = (double *)ptr;
    return 0;
}

int dalloc(void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(LBFFS_E_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    // hipMemset runs on the null stream and may return before it has finished; the
    // contexts' streams are non-blocking and do not wait for it, so a kernel that writes
    // the new buffer right away (the decode of a 16-bit download, for one) could be
    // overwritten by the zeros: wait for the fill here.
    e = hipMemset(*p, 0, bytes);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return fail(LBFFS_E_DEVICE, std::string("hipMemset: ") + hipGetErrorString(e));
    return 0;
}

// The class of the run of 64 cells from `first` (wave_class): PURE if all are fluid and off
// the lattice faces (PLAIN if besides every flag word is zero); on D3Q19 also XMIX if all are
// fluid and the face cells among them lie on an x face with j and k interior (so that every
// cell's interior-offset pulls stay in the lattice), or WRAP if all are fluid cells of one row
// with i interior, on a y face and/or a z face whose code is Cyclic for every cell (their
// pulls wrap like interior pulls shifted by whole rows or planes, and no face condition
// applies); GENERAL otherwise.
uint32_t wave_class_of(const lbffs_ctx *c, const uint16_t *fl, size_t first)
{
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
    bool face = false, plain = true;
    for (size_t l = 0; l < 64; l++) {
        if (fl[first + l] & F_SOLID) return WAVE_GENERAL;
        face = face || (fl[first + l] & F_FACE);
        plain = plain && fl[first + l] == 0;
    }
    if (!face) return plain ? WAVE_PLAIN : WAVE_PURE;
    if (c->d.q != 19) return WAVE_GENERAL;
    bool xmix = nz >= 3, wrap = true;
    const size_t i0 = first % (size_t)nx, r0 = first / (size_t)nx;
    for (size_t l = 0; l < 64; l++) {
        const size_t ic = first + l;
        const int i = (int)(ic % (size_t)nx), j = (int)((ic / (size_t)nx) % (size_t)ny);
        const int k = (int)(ic / ((size_t)nx * (size_t)ny));
        const uint32_t f = fl[ic];
        if (f & F_FACE) {
            const bool xonly = (i == 0 || i == nx - 1) && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2;
            xmix = xmix && xonly;
        }
        // WRAP: one row, i interior, every face it lies on periodic
        const bool same_row = ic / (size_t)nx == r0 && (l > 0 || i0 == (size_t)i);
        const bool jface = (j == 0 || j == ny - 1), kface = (k == 0 || k == nz - 1);
        wrap = wrap && same_row && i >= 1 && i <= nx - 2 && (jface || kface) &&
               (!jface || flag_code(f, F_B2_SHIFT) == BC_CYCLIC) && (!kface || flag_code(f, F_B3_SHIFT) == BC_CYCLIC);
    }
    if (xmix) return WAVE_XMIX;
    if (wrap) return WAVE_WRAP;
    return WAVE_GENERAL;
}

template <typename Real>
int pack_flags(lbffs_ctx *c)
{
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
    HIP_TRY(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
    Real om0 = (Real)c->d.omega0;
    hipLaunchKernelGGL((pack_flags_kernel<Real>), dim3(grid_for(c->N, 256)), dim3(256), 0, c->stream, c->d.q,
                       (uint32_t)c->N, nx, ny, nz, make_fastdiv(nx), make_fastdiv(ny), c->boundary, c->b1, c->b2,
                       c->b3, c->solid, c->neiSolid, (const Real *)c->sdf, (const Real *)c->omega, om0,
                       (const Real *)c->w0, c->d.streaming == LBFFS_STREAM_AA ? 1 : 0, c->flag, c->d_err);
    HIP_TRY(hipGetLastError());
    int err = 0;
    HIP_TRY(hipMemcpyAsync(&err, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    // face-centre codes that switch the sponge on (kernel.cl:2446-2451)
    int codes[6];
    const long cX0 = (long)nx * ny * (nz / 2) + (long)nx * (ny / 2);
    const long cXE = cX0 + nx - 1;
    const long cY0 = (long)nx * ny * (nz / 2) + nx / 2;
    const long cYE = cY0 + (long)nx * (ny - 1);
    const long cZ0 = (long)nx * (ny / 2) + nx / 2;
    const long cZE = cZ0 + (long)nx * ny * (nz - 1);
    HIP_TRY(hipMemcpyAsync(&codes[0], c->b1 + cX0, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&codes[1], c->b1 + cXE, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&codes[2], c->b2 + cY0, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&codes[3], c->b2 + cYE, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&codes[4], c->b3 + cZ0, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&codes[5], c->b3 + cZE, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (err) {
        std::string m = "invalid lattice description:";
        if (err & PACK_BAD_CODE) m += " face code outside 0..7;";
        if (err & PACK_BAD_BOUNDARY) m += " boundary value other than 0/1;";
        if (err & PACK_FACE_NOT_FLAGGED) m += " fluid cell on a lattice face with boundary==0;";
        if (err & PACK_WALLFN_ZMAX) m += " WallFunction on the nzMax face (the reference samples k+1 there);";
        if (err & PACK_2D_ZCODE) m += " nz==1 with an Outlet/FixedDensity/inflowing FixedVelocity z face;";
        if (err & PACK_NEISOLID_FACE) m += " neiSolid on a wall face (D3Q27 reads outside the lattice there);";
        if (err & PACK_AA_FACE)
            m += " in-place streaming needs Cyclic or BounceBack faces with the same code on both sides of an axis;";
        if (err & PACK_AA_NEISOLID) m += " in-place streaming needs neiSolid on interior fluid cells next to a solid cell;";
        return fail(LBFFS_E_INVALID, m);
    }
    c->sponge_mask = 0;
    for (int s = 0; s < 6; s++)
        if (codes[s] == BC_OUTLET || codes[s] == BC_FIXEDDENSITY) c->sponge_mask |= 1 << s;
    // the face pass's list: the non-solid lattice-face cells, in index order
    {
        std::vector<uint16_t> fl((size_t)c->N);
        HIP_TRY(hipMemcpy(fl.data(), c->flag, (size_t)c->N * sizeof(uint16_t), hipMemcpyDeviceToHost));
        std::vector<uint32_t> cells;
        for (long ic = 0; ic < c->N; ic++)
            if ((fl[ic] & F_FACE) && !(fl[ic] & F_SOLID)) cells.push_back((uint32_t)ic);
        if (c->fcells) (void)hipFree(c->fcells);
        c->fcells = nullptr;
        c->nface = 0;
        if (!cells.empty()) {
            if (int e = dalloc((void **)&c->fcells, cells.size() * sizeof(uint32_t))) return e;
            HIP_TRY(hipMemcpy(c->fcells, cells.data(), cells.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
            c->nface = (uint32_t)cells.size();
        }
        // per-wave summary (wave_class), 4 bits per run of 64 cells
        const size_t nw = (size_t)c->N / 64;
        std::vector<uint32_t> bits((nw + 7) / 8 + 1, 0u);
        for (size_t w = 0; w < nw; w++) {
            const uint32_t cls = wave_class_of(c, fl.data(), w * 64);
            bits[w >> 3] |= cls << (4 * (w & 7));
        }
        if (!c->wclass) {
            if (int e = dalloc((void **)&c->wclass, bits.size() * sizeof(uint32_t))) return e;
        }
        HIP_TRY(hipMemcpy(c->wclass, bits.data(), bits.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    c->flags_dirty = false;
    c->flags_gen++;
    return 0;
}

// rows per tile of the y-tiled cell order: 32, 16 or 8, the first that divides ny (0: the
// order is not available on this lattice)
uint32_t ytile_rows(const lbffs_ctx *c)
{
    for (uint32_t t : {32u, 16u, 8u})
        if ((uint32_t)c->d.ny % t == 0 && (uint32_t)c->d.ny > t) return t;
    return 0;
}

// whether the step kernels of this lattice read its per-wave flag summary: the summary
// names aligned runs of 64 cells, which a wave covers when its cells are consecutive (the
// plain order, or y-tiles whose rows are whole waves); off with TUNE_NO_PURE and in-place
// streaming
bool uses_wave_classes(const lbffs_ctx *c)
{
    const bool ytile = (c->tune & TUNE_YTILE) && ytile_rows(c) != 0;
    const bool runs = !ytile || c->d.nx % 64 == 0;
    return c->wclass && runs && !(c->tune & TUNE_NO_PURE) && c->d.streaming != LBFFS_STREAM_AA;
}

template <typename S, typename Real, int Q>
void set_args(lbffs_ctx *c, StepArgs<Real, Q> &a, const lbffs_step_params *p)
{
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
    a.N = (uint32_t)c->N;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.dnx = make_fastdiv(nx);
    a.dny = make_fastdiv(ny);
    for (int q = 0; q < Q; q++) {
        const int cx = Q == 27 ? L27::cx[q] : L19::cx[q];
        const int cy = Q == 27 ? L27::cy[q] : L19::cy[q];
        const int cz = Q == 27 ? L27::cz[q] : L19::cz[q];
        const int zoff = (nz == 1) ? 0 : cz;  // nz==1: z links stay in the layer (K19:907-989, K27:797-820)
        a.off[q] = cx + nx * cy + nx * ny * zoff;
    }
    a.fstride = (uint32_t)c->fstride;
    a.flag = c->flag;
    a.u0 = (const Real *)c->u0;
    a.v0 = (const Real *)c->v0;
    a.w0 = (const Real *)c->w0;
    a.sdf = (const Real *)c->sdf;
    a.omega = (const Real *)c->omega;
    a.tau = (Real *)c->tau;
    a.Fwx = (Real *)c->Fwx;
    a.Fwy = (Real *)c->Fwy;
    a.Fwz = (Real *)c->Fwz;
    a.Gx = (Real *)c->Gx;
    a.Gy = (Real *)c->Gy;
    a.Gz = (Real *)c->Gz;
    a.dpdx = (Real)p->dpdx;
    a.LES = (Real)p->LES;
    a.omegaB = (Real)p->omegaB;
    a.spz = (Real)p->spzWidth;
    a.Cs = (Real)p->Cs;
    a.clim = (Real)p->clim;
    a.ibb = p->ibb ? 1 : 0;
    a.sround = (c->store_size == 2 && c->d.rounding == LBFFS_ROUND_STOCHASTIC) ? 1 : 0;
    a.qlink = (const Real *)c->qlink;
    a.omega0 = (Real)c->d.omega0;
    a.damping = p->damping;
    a.sponge_mask = c->sponge_mask;
    const double spz = p->spzWidth;
    a.inv_spx = (Real)(spz > 0 ? 1.0 / ((double)nx * spz) : 0.0);
    a.inv_spy = (Real)(spz > 0 ? 1.0 / ((double)ny * spz) : 0.0);
    a.inv_spz = (Real)(spz > 0 ? 1.0 / ((double)nz * spz) : 0.0);
    a.tune = c->tune;
    a.fcells = c->fcells;
    a.nface = c->nface;
    a.step = c->nstep;
    a.step_ctr = nullptr;
    a.coll = c->d.collision;
    {
        const uint32_t ty = ytile_rows(c);
        const uint32_t t = ty ? ty : (uint32_t)ny;
        a.tile_plane = (uint32_t)nx * t;
        a.tile_cells = a.tile_plane * (uint32_t)nz;
        a.nxny = (uint32_t)nx * (uint32_t)ny;
        a.dtile = make_fastdiv(a.tile_cells);
        a.dplane = make_fastdiv(a.tile_plane);
        if (!ty) a.tune &= ~TUNE_YTILE;
    }
    // the plain order gives every wave one aligned run of 64 cells; y-tiles do where their
    // rows are whole waves
    a.wave_runs = (!(a.tune & TUNE_YTILE) || nx % 64 == 0) ? 1 : 0;
    a.wclass = uses_wave_classes(c) ? c->wclass : nullptr;
}

// In-place streaming: bring the array back to the two-lattice layout (after an odd number
// of half-steps) before anything reads it by cell; the flag array is the one the last step
// ran with
template <typename S>
int aa_restore_t(lbffs_ctx *c)
{
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
    hipLaunchKernelGGL((aa_swap19_kernel<S>), dim3(grid_for(c->N, 256)), dim3(256), 0, c->stream, (uint32_t)c->N, nx,
                       ny, nz, make_fastdiv(nx), make_fastdiv(ny), c->flag, c->fstride, (S *)c->f[0]);
    HIP_TRY(hipGetLastError());
    return 0;
}

int aa_restore(lbffs_ctx *c)
{
    if (c->d.streaming != LBFFS_STREAM_AA || !c->aa_odd) return 0;
    int e = 0;
    switch (c->d.storage) {
    case LBFFS_STORE_F32: e = aa_restore_t<float>(c); break;
    case LBFFS_STORE_F16DEV: e = aa_restore_t<__half>(c); break;
    case LBFFS_STORE_BF16DEV: e = aa_restore_t<bf16_t>(c); break;
    default: e = aa_restore_t<double>(c); break;
    }
    if (!e) c->aa_odd = 0;
    return e;
}

bool ibm_active(const lbffs_ctx *c);

int aa_step_ok(const lbffs_ctx *c, const lbffs_step_params *p)
{
    if (c->d.streaming != LBFFS_STREAM_AA) return 0;
    if (p->ibb) return fail(LBFFS_E_UNSUPPORTED, "in-place streaming runs without the interpolated bounce-back");
    if (ibm_active(c)) return fail(LBFFS_E_UNSUPPORTED, "in-place streaming runs without moving-wall markers");
    return 0;
}

template <typename Real>
IbmMotion<Real> ibm_motion(const lbffs_ctx *c)
{
    IbmMotion<Real> m;
    for (int a = 0; a < 3; a++) {
        m.trans[a] = (Real)c->motion[a];
        m.centre[a] = (Real)c->motion[3 + a];
        m.axis[a] = (Real)c->motion[6 + a];
    }
    m.omega = (Real)c->motion[9];
    return m;
}

bool ibm_active(const lbffs_ctx *c) { return c->n_markers > 1; }

// the moving-wall kernels after a step (ns.cpp:544-600): marker forcing from the new
// velocities, spread + marker motion, forcing of the new populations
template <typename S, typename Real, int Q>
int ibm_after_step(lbffs_ctx *c, bool last)
{
    const int n = c->n_markers, cap = 8 * n;
    const IbmMotion<Real> m = ibm_motion<Real>(c);
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
    c->stamp = (c->stamp == 0xffffffffu) ? 1u : c->stamp + 1u;
    hipLaunchKernelGGL((ibm_gwall_kernel<Real>), dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n,
                       (const Real *)c->markers, nx, ny, nz, (const Real *)c->u[c->mcur], (const Real *)c->v[c->mcur],
                       (const Real *)c->w[c->mcur], m, (Real *)c->Gm, c->count);
    hipLaunchKernelGGL((ibm_spread_move_kernel<Real>), dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n,
                       (Real *)c->markers, nx, ny, nz, (const Real *)c->Gm, m, (Real *)c->Gx, (Real *)c->Gy,
                       (Real *)c->Gz, c->mark, c->stamp, c->list, c->count, cap);
    hipLaunchKernelGGL((ibm_force_kernel<S, Real, Q>), dim3(grid_for(cap, 256)), dim3(256), 0, c->stream,
                       (const int *)c->list, (const int *)c->count, cap, c->fstride, (S *)c->f[c->cur], c->flag,
                       (const Real *)c->rho[c->mcur], (Real *)c->Gx, (Real *)c->Gy, (Real *)c->Gz, last ? 0 : 1);
    HIP_TRY(hipGetLastError());
    return 0;
}

// G = 0 where the next spread must start from zero
template <typename Real>
int ibm_clear_g(lbffs_ctx *c)
{
    if (c->g_dirty) {
        for (void *g : {c->Gx, c->Gy, c->Gz})
            hipLaunchKernelGGL((fill_kernel<Real>), dim3(1024), dim3(256), 0, c->stream, (Real *)g, (size_t)c->N,
                               Real(0));
        c->g_dirty = false;
    } else {
        const int cap = 8 * c->n_markers;
        hipLaunchKernelGGL((ibm_zero_listed_kernel<Real>), dim3(grid_for(cap, 256)), dim3(256), 0, c->stream,
                           (const int *)c->list, (const int *)c->count, cap, (Real *)c->Gx, (Real *)c->Gy,
                           (Real *)c->Gz);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename S, typename Real, int Q>
int run_steps_q(lbffs_ctx *c, int nsteps, const lbffs_step_params *p)
{
    StepArgs<Real, Q> a;
    set_args<S, Real, Q>(c, a, p);
    const int block = c->block;
    const unsigned grid = grid_for(c->N, block);
    const bool ibm = ibm_active(c);
    if (ibm) {
        if (int e = ibm_clear_g<Real>(c)) return e;
    }
    const bool aa = c->d.streaming == LBFFS_STREAM_AA;
    for (int s = 0; s < nsteps; s++) {
        const bool last = (s == nsteps - 1);
        a.fsrc = c->f[c->cur];
        a.fdst = aa ? c->f[c->cur] : c->f[1 - c->cur];
        a.rho_o = (const Real *)c->rho[c->mcur];
        a.u_o = (const Real *)c->u[c->mcur];
        a.v_o = (const Real *)c->v[c->mcur];
        a.w_o = (const Real *)c->w[c->mcur];
        a.rho_n = (Real *)c->rho[1 - c->mcur];
        a.u_n = (Real *)c->u[1 - c->mcur];
        a.v_n = (Real *)c->v[1 - c->mcur];
        a.w_n = (Real *)c->w[1 - c->mcur];
        a.aux_all = (p->store_every || last) ? 1 : 0;
        // WALE reads every neighbour's velocity (not its density); the moving-wall kernels
        // read the velocities of the marker boxes and the density of the cells they force
        a.mac_all = (a.aux_all || c->d.sgs == LBFFS_SGS_WALE || ibm) ? 1 : 0;
        a.rho_all = (a.aux_all || ibm) ? 1 : 0;
        a.step = c->nstep;
        const int phase = aa ? 1 + c->aa_odd : 0;
        switch (c->d.sgs) {
        case LBFFS_SGS_WALE: launch_step<S, Real, Q, 0>(a, grid, block, c->stream, phase); break;
        case LBFFS_SGS_SMAGORINSKY: launch_step<S, Real, Q, 1>(a, grid, block, c->stream, phase); break;
        default: launch_step<S, Real, Q, 2>(a, grid, block, c->stream, phase); break;
        }
        HIP_TRY(hipGetLastError());
        if (aa)
            c->aa_odd ^= 1;
        else
            c->cur ^= 1;
        c->mcur ^= 1;
        c->nstep++;
        if (ibm) {
            if (int e = ibm_after_step<S, Real, Q>(c, last)) return e;
        }
    }
    return 0;
}

template <typename S, typename Real>
int run_steps(lbffs_ctx *c, int nsteps, const lbffs_step_params *p)
{
    return c->d.q == 27 ? run_steps_q<S, Real, 27>(c, nsteps, p) : run_steps_q<S, Real, 19>(c, nsteps, p);
}

template <typename S, typename Real>
int upload_f(lbffs_ctx *c, const void *host)
{
    // host q-major arrays of N -> a device copy of them -> the storage form in the device
    // layout (encode_kernel)
    const size_t n = (size_t)c->d.q * c->N;
    void *tmp = nullptr;
    if (int e = dalloc(&tmp, n * sizeof(Real))) return e;
    hipError_t e1 = hipMemcpy(tmp, host, n * sizeof(Real), hipMemcpyHostToDevice);
    if (e1 == hipSuccess) {
        for (int b = 0; b < 2; b++) {
            if (!c->f[b]) continue;
            if (c->d.q == 27)
                hipLaunchKernelGGL((encode_kernel<S, Real, 27>), dim3(grid_for(c->N, 256)), dim3(256), 0, c->stream,
                                   (uint32_t)c->N, c->fstride, (const Real *)tmp, (S *)c->f[b]);
            else
                hipLaunchKernelGGL((encode_kernel<S, Real, 19>), dim3(grid_for(c->N, 256)), dim3(256), 0, c->stream,
                                   (uint32_t)c->N, c->fstride, (const Real *)tmp, (S *)c->f[b]);
        }
        e1 = hipGetLastError();
        if (e1 == hipSuccess) e1 = hipStreamSynchronize(c->stream);
    }
    hipFree(tmp);
    if (e1 != hipSuccess) return fail(LBFFS_E_DEVICE, std::string("upload f: ") + hipGetErrorString(e1));
    c->aa_odd = 0;
    return 0;
}

template <typename S, typename Real>
int download_f(lbffs_ctx *c, void *host)
{
    const size_t n = (size_t)c->d.q * c->N;
    HIP_TRY(hipStreamSynchronize(c->stream));
    void *tmp = nullptr;
    if (int e = dalloc(&tmp, n * sizeof(Real))) return e;
    if (c->d.q == 27)
        hipLaunchKernelGGL((decode_kernel<S, Real, 27>), dim3(grid_for(c->N, 256)), dim3(256), 0, c->stream,
                           (uint32_t)c->N, c->fstride, (const S *)c->f[c->cur], (Real *)tmp);
    else
        hipLaunchKernelGGL((decode_kernel<S, Real, 19>), dim3(grid_for(c->N, 256)), dim3(256), 0, c->stream,
                           (uint32_t)c->N, c->fstride, (const S *)c->f[c->cur], (Real *)tmp);
    hipError_t e1 = hipGetLastError();
    if (e1 == hipSuccess) e1 = hipStreamSynchronize(c->stream);
    if (e1 == hipSuccess) e1 = hipMemcpy(host, tmp, n * sizeof(Real), hipMemcpyDeviceToHost);
    hipFree(tmp);
    if (e1 != hipSuccess) return fail(LBFFS_E_DEVICE, std::string("download f: ") + hipGetErrorString(e1));
    return 0;
}

// part: kRedBlocks x LBFFS_REDUCE_N partial sums; red: the LBFFS_REDUCE_N final sums
template <typename S, typename Real>
int reduce_impl(lbffs_ctx *c, double *out, double *part, double *red)
{
    if (c->d.q == 27)
        hipLaunchKernelGGL((reduce_kernel<S, Real, 27>), dim3(kRedBlocks), dim3(kRedThreads), 0, c->stream,
                           (uint32_t)c->N, c->fstride, c->flag, (const S *)c->f[c->cur], (const Real *)c->Fwx,
                           (const Real *)c->Fwy, (const Real *)c->Fwz, (const Real *)c->Gx, (const Real *)c->Gy,
                           (const Real *)c->Gz, part);
    else
        hipLaunchKernelGGL((reduce_kernel<S, Real, 19>), dim3(kRedBlocks), dim3(kRedThreads), 0, c->stream,
                           (uint32_t)c->N, c->fstride, c->flag, (const S *)c->f[c->cur], (const Real *)c->Fwx,
                           (const Real *)c->Fwy, (const Real *)c->Fwz, (const Real *)c->Gx, (const Real *)c->Gy,
                           (const Real *)c->Gz, part);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(64), 0, c->stream, (const double *)part, kRedBlocks, red);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, red, sizeof(double) * LBFFS_REDUCE_N, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// dispatch on the storage form
template <template <typename, typename> class F, typename... A>
int dispatch(lbffs_ctx *c, A... args)
{
    switch (c->d.storage) {
    case LBFFS_STORE_F32: return F<float, float>::run(c, args...);
    case LBFFS_STORE_F16DEV: return F<__half, float>::run(c, args...);
    case LBFFS_STORE_BF16DEV: return F<bf16_t, float>::run(c, args...);
    case LBFFS_STORE_F64: return F<double, double>::run(c, args...);
    }
    return fail(LBFFS_E_UNSUPPORTED, "storage form");
}

template <typename S, typename Real>
struct StepOp {
    static int run(lbffs_ctx *c, int n, const lbffs_step_params *p) { return run_steps<S, Real>(c, n, p); }
};
template <typename S, typename Real>
struct UploadFOp {
    static int run(lbffs_ctx *c, const void *h) { return upload_f<S, Real>(c, h); }
};
template <typename S, typename Real>
struct DownloadFOp {
    static int run(lbffs_ctx *c, void *h) { return download_f<S, Real>(c, h); }
};
template <typename S, typename Real>
struct ReduceOp {
    static int run(lbffs_ctx *c, double *o, double *part, double *red) { return reduce_impl<S, Real>(c, o, part, red); }
};

template <typename S, typename Real>
int output_macros_impl(lbffs_ctx *c, double dpdx, void *const *host)
{
    const size_t n = (size_t)c->N;
    void *tmp = nullptr;
    if (int e = dalloc(&tmp, 4 * n * sizeof(Real))) return e;
    Real *t = (Real *)tmp;
    if (c->d.q == 27)
        hipLaunchKernelGGL((output_macros_kernel<S, Real, 27>), dim3(grid_for(c->N, 256)), dim3(256), 0, c->stream,
                           (uint32_t)c->N, c->fstride, (const S *)c->f[c->cur], (Real)dpdx, t, t + n, t + 2 * n,
                           t + 3 * n);
    else
        hipLaunchKernelGGL((output_macros_kernel<S, Real, 19>), dim3(grid_for(c->N, 256)), dim3(256), 0, c->stream,
                           (uint32_t)c->N, c->fstride, (const S *)c->f[c->cur], (Real)dpdx, t, t + n, t + 2 * n,
                           t + 3 * n);
    hipError_t e1 = hipGetLastError();
    if (e1 == hipSuccess) e1 = hipStreamSynchronize(c->stream);
    for (int a = 0; a < 4 && e1 == hipSuccess; a++) e1 = hipMemcpy(host[a], t + a * n, n * sizeof(Real), hipMemcpyDeviceToHost);
    hipFree(tmp);
    if (e1 != hipSuccess) return fail(LBFFS_E_DEVICE, std::string("output macros: ") + hipGetErrorString(e1));
    return 0;
}
template <typename S, typename Real>
struct OutputMacrosOp {
    static int run(lbffs_ctx *c, double dpdx, void *const *h) { return output_macros_impl<S, Real>(c, dpdx, h); }
};

void free_all(lbffs_ctx *c)
{
    void *ptrs[] = {c->f[0], c->f[1], c->boundary, c->b1, c->b2, c->b3, c->solid, c->neiSolid, c->sdf, c->u0, c->v0,
                    c->w0, c->omega, c->Fwx, c->Fwy, c->Fwz, c->tau, c->Gx, c->Gy, c->Gz, c->rho[0], c->rho[1],
                    c->u[0], c->u[1], c->v[0], c->v[1], c->w[0], c->w[1], c->flag, c->d_err, c->d_part, c->d_red,
                    c->markers, c->Gm, c->mark, c->list, c->count, c->qlink, c->fcells, c->wclass};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream && c->owns_stream) hipStreamDestroy(c->stream);
}

}  // namespace

struct lbffs_batch {
    std::vector<lbffs_ctx *> ctx;
    hipStream_t stream = nullptr;
    // launch arguments of every lattice: [parity + 2 * last step of a call]
    void *args_dev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<char> args_host[4];
    BatchBlock *table = nullptr;
    uint32_t nblocks = 0;
    // the face cells of every lattice (run by a launch of their own), rebuilt when flags change
    BatchBlock *ftable = nullptr;
    uint32_t nfblocks = 0;
    uint32_t *fcells = nullptr;
    bool ftable_ok = false;
    std::vector<uint64_t> gen;  // flags generation of each lattice the face table was built from
    std::vector<uint8_t> wc;    // whether each lattice used its wave summary then (its XMIX and
                                // WRAP waves' face cells are left out of the table)
    uint32_t *ctr = nullptr;    // device counter of the steps run in the current call
    int grid = 0;        // work-groups of the launch; 0 = one per table entry
    int graph_chunk = 0; // steps per captured hipGraph; 0 = plain launches
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int gexec_chunk[2] = {0, 0};
    int gexec_ibb[2] = {0, 0};  // kernel variant the graph was captured with
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
};

namespace {

// the batch's counter of steps within a call: reset at the start of a call, advanced after
// every step (a kernel, so that a captured graph advances it on every replay)
__global__ void step_ctr_kernel(uint32_t *ctr, int reset)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *ctr = reset ? 0u : *ctr + 1u;
}

void batch_drop_graphs(lbffs_batch *b)
{
    for (int p = 0; p < 2; p++) {
        if (b->gexec[p]) (void)hipGraphExecDestroy(b->gexec[p]);
        b->gexec[p] = nullptr;
        b->gexec_chunk[p] = 0;
    }
}

template <typename S, typename Real, int Q, int SGS, bool IBB>
void batch_launch(lbffs_batch *b, int parity, int last)
{
    const StepArgs<Real, Q> *args = (const StepArgs<Real, Q> *)b->args_dev[parity + 2 * last];
    launch_batch<S, Real, Q, SGS, IBB>(args, (const BatchBlock *)b->table, b->nblocks, (const BatchBlock *)b->ftable,
                                       b->nfblocks, (const uint32_t *)b->fcells, b->grid > 0 ? (unsigned)b->grid : 0u,
                                       b->ctx[0]->d.collision, b->stream);
}

// the face table: every non-solid face cell of every lattice, 256 to a block, except the face
// cells of XMIX and WRAP waves where the lattice uses its wave summary (the block-table pass
// runs those in their own waves; left in, they would idle lanes of the face pass's waves)
int batch_face_table(lbffs_batch *b)
{
    std::vector<uint32_t> cells;
    std::vector<BatchBlock> blocks;
    std::vector<uint16_t> fl;
    b->wc.assign(b->ctx.size(), 0);
    for (size_t d = 0; d < b->ctx.size(); d++) {
        lbffs_ctx *c = b->ctx[d];
        fl.resize((size_t)c->N);
        HIP_TRY(hipMemcpy(fl.data(), c->flag, (size_t)c->N * sizeof(uint16_t), hipMemcpyDeviceToHost));
        const bool wcs = uses_wave_classes(c);
        b->wc[d] = wcs ? 1 : 0;
        const size_t nw = (size_t)c->N / 64;  // whole runs; the tail run is GENERAL
        size_t cur_w = ~size_t(0);
        uint32_t cls = WAVE_GENERAL;
        const size_t start = cells.size();
        for (long ic = 0; ic < c->N; ic++) {
            if (!(fl[ic] & F_FACE) || (fl[ic] & F_SOLID)) continue;
            if (wcs) {
                const size_t w = (size_t)ic / 64;
                if (w != cur_w) {
                    cur_w = w;
                    cls = w < nw ? wave_class_of(c, fl.data(), w * 64) : WAVE_GENERAL;
                }
                if (cls == WAVE_XMIX || cls == WAVE_WRAP) continue;
            }
            cells.push_back((uint32_t)ic);
        }
        for (size_t f = start; f < cells.size(); f += kBatchBlock)
            blocks.push_back({(uint32_t)d, (uint32_t)f, (uint32_t)std::min<size_t>(kBatchBlock, cells.size() - f)});
    }
    if (cells.size() > 0xffffffffu) return fail(LBFFS_E_INVALID, "too many face cells in a batch");
    if (b->ftable) (void)hipFree(b->ftable);
    if (b->fcells) (void)hipFree(b->fcells);
    b->ftable = nullptr;
    b->fcells = nullptr;
    b->nfblocks = 0;
    if (!blocks.empty()) {
        if (int e = dalloc((void **)&b->ftable, blocks.size() * sizeof(BatchBlock))) return e;
        if (int e = dalloc((void **)&b->fcells, cells.size() * sizeof(uint32_t))) return e;
        HIP_TRY(hipMemcpy(b->ftable, blocks.data(), blocks.size() * sizeof(BatchBlock), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b->fcells, cells.data(), cells.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        b->nfblocks = (uint32_t)blocks.size();
    }
    batch_drop_graphs(b);  // the captured launches name the old table
    b->gen.resize(b->ctx.size());
    for (size_t d = 0; d < b->ctx.size(); d++) b->gen[d] = b->ctx[d]->flags_gen;
    b->ftable_ok = true;
    return 0;
}

template <typename S, typename Real, int Q, int SGS, bool IBB>
int batch_steps_sgs(lbffs_batch *b, int nsteps, int cur0)
{
    int s = 0;
    // all steps but the last, in captured chunks when asked
    if (b->graph_chunk > 0) {
        const int chunk = b->graph_chunk;
        while (nsteps - 1 - s >= chunk) {
            const int p = (cur0 + s) & 1;
            if (!b->gexec[p] || b->gexec_chunk[p] != chunk || b->gexec_ibb[p] != (int)IBB) {
                if (b->gexec[p]) (void)hipGraphExecDestroy(b->gexec[p]);
                b->gexec[p] = nullptr;
                hipGraph_t g = nullptr;
                HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                for (int k = 0; k < chunk; k++) {
                    batch_launch<S, Real, Q, SGS, IBB>(b, (p + k) & 1, 0);
                    hipLaunchKernelGGL(step_ctr_kernel, dim3(1), dim3(64), 0, b->stream, b->ctr, 0);
                }
                HIP_TRY(hipStreamEndCapture(b->stream, &g));
                hipError_t e = hipGraphInstantiate(&b->gexec[p], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) return fail(LBFFS_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
                b->gexec_chunk[p] = chunk;
                b->gexec_ibb[p] = (int)IBB;
            }
            HIP_TRY(hipGraphLaunch(b->gexec[p], b->stream));
            s += chunk;
        }
    }
    for (; s < nsteps; s++) {
        const int p = (cur0 + s) & 1;
        batch_launch<S, Real, Q, SGS, IBB>(b, p, s == nsteps - 1 ? 1 : 0);
        hipLaunchKernelGGL(step_ctr_kernel, dim3(1), dim3(64), 0, b->stream, b->ctr, 0);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

template <typename S, typename Real, int Q>
int batch_steps_q(lbffs_batch *b, int nsteps, const lbffs_step_params *params)
{
    const int cur0 = b->ctx[0]->cur;
    for (int v = 0; v < 4; v++) {
        const int p = v & 1, last = v >> 1;
        b->args_host[v].assign(b->ctx.size() * sizeof(StepArgs<Real, Q>), 0);
        StepArgs<Real, Q> *A = (StepArgs<Real, Q> *)b->args_host[v].data();
        for (size_t d = 0; d < b->ctx.size(); d++) {
            lbffs_ctx *c = b->ctx[d];
            StepArgs<Real, Q> &a = A[d];
            set_args<S, Real, Q>(c, a, &params[d]);
            a.fsrc = c->f[p];
            a.fdst = c->f[1 - p];
            a.rho_o = (const Real *)c->rho[p];
            a.u_o = (const Real *)c->u[p];
            a.v_o = (const Real *)c->v[p];
            a.w_o = (const Real *)c->w[p];
            a.rho_n = (Real *)c->rho[1 - p];
            a.u_n = (Real *)c->u[1 - p];
            a.v_n = (Real *)c->v[1 - p];
            a.w_n = (Real *)c->w[1 - p];
            a.aux_all = (params[d].store_every || last) ? 1 : 0;
            a.mac_all = (a.aux_all || c->d.sgs == LBFFS_SGS_WALE) ? 1 : 0;
            a.rho_all = a.aux_all;
            a.step = c->nstep;  // + the device counter of steps within this call
            a.step_ctr = b->ctr;
        }
        HIP_TRY(hipMemcpy(b->args_dev[v], b->args_host[v].data(), b->args_host[v].size(), hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(step_ctr_kernel, dim3(1), dim3(64), 0, b->stream, b->ctr, 1);
    // D3Q19's interpolated bounce-back is a kernel variant (its loop costs registers in
    // every cell); D3Q27 selects the q source per link at run time
    bool ibb = false;
    for (size_t d = 0; d < b->ctx.size(); d++) ibb = ibb || params[d].ibb;
    if constexpr (Q == 19) {
        if (ibb) {
            switch (b->ctx[0]->d.sgs) {
            case LBFFS_SGS_WALE: return batch_steps_sgs<S, Real, Q, 0, true>(b, nsteps, cur0);
            case LBFFS_SGS_SMAGORINSKY: return batch_steps_sgs<S, Real, Q, 1, true>(b, nsteps, cur0);
            default: return batch_steps_sgs<S, Real, Q, 2, true>(b, nsteps, cur0);
            }
        }
    }
    switch (b->ctx[0]->d.sgs) {
    case LBFFS_SGS_WALE: return batch_steps_sgs<S, Real, Q, 0, false>(b, nsteps, cur0);
    case LBFFS_SGS_SMAGORINSKY: return batch_steps_sgs<S, Real, Q, 1, false>(b, nsteps, cur0);
    default: return batch_steps_sgs<S, Real, Q, 2, false>(b, nsteps, cur0);
    }
}

template <typename S, typename Real>
struct BatchStepOp {
    static int run(lbffs_batch *b, int n, const lbffs_step_params *p)
    {
        return b->ctx[0]->d.q == 27 ? batch_steps_q<S, Real, 27>(b, n, p) : batch_steps_q<S, Real, 19>(b, n, p);
    }
};

template <template <typename, typename> class F, typename... A>
int batch_dispatch(lbffs_batch *b, A... args)
{
    switch (b->ctx[0]->d.storage) {
    case LBFFS_STORE_F32: return F<float, float>::run(b, args...);
    case LBFFS_STORE_F16DEV: return F<__half, float>::run(b, args...);
    case LBFFS_STORE_BF16DEV: return F<bf16_t, float>::run(b, args...);
    case LBFFS_STORE_F64: return F<double, double>::run(b, args...);
    }
    return fail(LBFFS_E_UNSUPPORTED, "storage form");
}

void batch_free(lbffs_batch *b)
{
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graphs(b);
    for (lbffs_ctx *c : b->ctx) {
        free_all(c);
        delete c;
    }
    for (void *p : {b->args_dev[0], b->args_dev[1], b->args_dev[2], b->args_dev[3], (void *)b->table,
                    (void *)b->ftable, (void *)b->fcells, (void *)b->ctr})
        if (p) (void)hipFree(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

// a device copy of everything a step writes (the current populations, moments, relaxation
// times, wall forces, marker forcing and markers) and the step counters, so that a tuner can
// run trial steps and leave the lattice as it found it
struct StateCopy {
    lbffs_ctx *c = nullptr;
    void *f = nullptr, *r = nullptr, *m = nullptr;
    int cur = 0, mcur = 0, aa_odd = 0;
    uint32_t nstep = 0;

    StateCopy() = default;
    StateCopy(const StateCopy &) = delete;
    StateCopy &operator=(const StateCopy &) = delete;
    ~StateCopy()
    {
        for (void *p : {f, r, m})
            if (p) (void)hipFree(p);
    }
    size_t fbytes() const { return c->felems * c->store_size; }
    size_t rbytes() const { return (size_t)c->N * c->real_size; }
    size_t mbytes() const { return (size_t)3 * (size_t)c->n_markers * c->real_size; }
    void fields(void *k[11], int mc) const
    {
        void *a[11] = {c->rho[mc], c->u[mc], c->v[mc], c->w[mc], c->tau, c->Fwx, c->Fwy, c->Fwz, c->Gx, c->Gy, c->Gz};
        for (int i = 0; i < 11; i++) k[i] = a[i];
    }
    // allocate and enqueue the copies on the lattice's stream
    int save(lbffs_ctx *ctx)
    {
        c = ctx;
        cur = c->cur;
        mcur = c->mcur;
        aa_odd = c->aa_odd;
        nstep = c->nstep;
        if (int e = dalloc(&f, fbytes())) return e;
        if (int e = dalloc(&r, 11 * rbytes())) return e;
        if (c->n_markers > 0)
            if (int e = dalloc(&m, mbytes())) return e;
        void *k[11];
        fields(k, mcur);
        hipError_t he = hipMemcpyAsync(f, c->f[cur], fbytes(), hipMemcpyDeviceToDevice, c->stream);
        for (int i = 0; i < 11 && he == hipSuccess; i++)
            he = hipMemcpyAsync((char *)r + i * rbytes(), k[i], rbytes(), hipMemcpyDeviceToDevice, c->stream);
        if (m && he == hipSuccess) he = hipMemcpyAsync(m, c->markers, mbytes(), hipMemcpyDeviceToDevice, c->stream);
        if (he != hipSuccess) return fail(LBFFS_E_DEVICE, std::string("state copy: ") + hipGetErrorString(he));
        return 0;
    }
    // put the step counters back and enqueue the copies back
    hipError_t restore()
    {
        c->cur = cur;
        c->mcur = mcur;
        c->aa_odd = aa_odd;
        c->nstep = nstep;
        c->g_dirty = true;  // the restored G are the saved ones, not those of the last list
        c->timed = false;
        void *k[11];
        fields(k, mcur);
        hipError_t he = hipMemcpyAsync(c->f[cur], f, fbytes(), hipMemcpyDeviceToDevice, c->stream);
        for (int i = 0; i < 11 && he == hipSuccess; i++)
            he = hipMemcpyAsync(k[i], (char *)r + i * rbytes(), rbytes(), hipMemcpyDeviceToDevice, c->stream);
        if (m && he == hipSuccess) he = hipMemcpyAsync(c->markers, m, mbytes(), hipMemcpyDeviceToDevice, c->stream);
        return he;
    }
};

}  // namespace

extern "C" {

int lbffs_create(const lbffs_desc *desc, lbffs_ctx **out)
{
    if (!desc || !out) return fail(LBFFS_E_INVALID, "null argument");
    *out = nullptr;
    if (desc->abi_version != LBFFS_ABI_VERSION) return fail(LBFFS_E_INVALID, "abi_version mismatch");
    if (desc->q != 19 && desc->q != 27) return fail(LBFFS_E_UNSUPPORTED, "q must be 19 or 27");
    if (desc->nx < 2 || desc->ny < 2 || desc->nz < 1) return fail(LBFFS_E_INVALID, "lattice needs nx,ny >= 2, nz >= 1");
    const long N = (long)desc->nx * desc->ny * desc->nz;
    if (N >= (1L << 31) / 1 || (long)desc->q * N < 0) return fail(LBFFS_E_INVALID, "cell count must be below 2^31");
    if (desc->storage < 0 || desc->storage > 3) return fail(LBFFS_E_INVALID, "storage");
    if (desc->sgs < 0 || desc->sgs > 2) return fail(LBFFS_E_INVALID, "sgs");
    if (desc->rounding != LBFFS_ROUND_STOCHASTIC && desc->rounding != LBFFS_ROUND_NEAREST)
        return fail(LBFFS_E_INVALID, "rounding");
    if (desc->streaming != LBFFS_STREAM_TWO_LATTICE && desc->streaming != LBFFS_STREAM_AA)
        return fail(LBFFS_E_INVALID, "streaming");
    if (desc->streaming == LBFFS_STREAM_AA && desc->q != 19)
        return fail(LBFFS_E_UNSUPPORTED, "in-place streaming runs D3Q19 lattices");
    if (desc->q == 19 && desc->collision != LBFFS_COLL_RR)
        return fail(LBFFS_E_UNSUPPORTED, "D3Q19 runs the recursive-regularized collision (K19:3637)");
    if (desc->q == 27 && desc->collision != LBFFS_COLL_CUMULANT && desc->collision != LBFFS_COLL_RR)
        return fail(LBFFS_E_UNSUPPORTED, "D3Q27 runs the cumulant (K27:2824) or the recursive-regularized collision");
    if (desc->block != 0 && (desc->block % 64 != 0 || desc->block > 256 || desc->block < 64))
        return fail(LBFFS_E_INVALID, "block must be 64, 128, 192 or 256");
    // the step kernels address a population array, and the interior velocity gradient a
    // macroscopic array (FP32, FP64 for FP64 storage), with 32-bit byte offsets from its base
    {
        const long elem = (desc->storage == LBFFS_STORE_F64) ? 8 : (desc->storage == LBFFS_STORE_F32 ? 4 : 2);
        const long real = (desc->storage == LBFFS_STORE_F64) ? 8 : 4;
        if (N * std::max(elem, real) > 0xFFFFFFFFL)
            return fail(LBFFS_E_INVALID, "one population or macroscopic array must stay below 4 GiB (" +
                                             std::to_string(N) + " cells of " + std::to_string(std::max(elem, real)) +
                                             " bytes)");
    }

    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    if (int e = scratch_setup(device)) return e;
    lbffs_ctx *c = new lbffs_ctx();
    c->d = *desc;
    c->device = device;
    c->N = N;
    c->block = desc->block > 0 ? desc->block : 256;
    c->real_size = (desc->storage == LBFFS_STORE_F64) ? 8 : 4;
    c->store_size = (desc->storage == LBFFS_STORE_F64) ? 8 : (desc->storage == LBFFS_STORE_F32 ? 4 : 2);
    const size_t n = (size_t)N, rs = c->real_size;
    int e = 0;
    hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete c;
        return fail(LBFFS_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    c->felems = pop_elems(desc->q, c->store_size, n);
    c->fstride = pop_stride(c->store_size, n);
    if (!e) e = dalloc(&c->f[0], c->felems * c->store_size);
    if (!e && desc->streaming != LBFFS_STREAM_AA) e = dalloc(&c->f[1], c->felems * c->store_size);
    if (!e) e = dalloc((void **)&c->boundary, n * 4);
    if (!e) e = dalloc((void **)&c->b1, n * 4);
    if (!e) e = dalloc((void **)&c->b2, n * 4);
    if (!e) e = dalloc((void **)&c->b3, n * 4);
    if (!e) e = dalloc((void **)&c->solid, n);
    if (!e) e = dalloc((void **)&c->neiSolid, n);
    void **reals[] = {&c->sdf, &c->u0, &c->v0, &c->w0, &c->omega, &c->Fwx, &c->Fwy, &c->Fwz, &c->tau, &c->Gx,
                      &c->Gy, &c->Gz, &c->rho[0], &c->rho[1], &c->u[0], &c->u[1], &c->v[0], &c->v[1], &c->w[0],
                      &c->w[1]};
    for (void **p : reals)
        if (!e) e = dalloc(p, n * rs);
    if (!e) e = dalloc((void **)&c->flag, n * sizeof(uint16_t));
    if (!e) e = dalloc((void **)&c->d_err, sizeof(int) * 4);
    if (!e) e = dalloc((void **)&c->d_part, sizeof(double) * LBFFS_REDUCE_N * kRedBlocks);
    if (!e) e = dalloc((void **)&c->d_red, sizeof(double) * LBFFS_REDUCE_N);
    if (!e) {
        he = hipEventCreate(&c->ev0);
        if (he == hipSuccess) he = hipEventCreate(&c->ev1);
        if (he != hipSuccess) e = fail(LBFFS_E_DEVICE, std::string("hipEventCreate: ") + hipGetErrorString(he));
    }
    if (e) {
        free_all(c);
        delete c;
        return e;
    }
    // sdf defaults to the reference's "no wall" value 10000 (ns.cpp:201-202); omega to omega0
    if (rs == 4) {
        std::vector<float> v(n, 10000.f), o(n, desc->omega0);
        hipMemcpy(c->sdf, v.data(), n * 4, hipMemcpyHostToDevice);
        hipMemcpy(c->omega, o.data(), n * 4, hipMemcpyHostToDevice);
    } else {
        std::vector<double> v(n, 10000.0), o(n, (double)desc->omega0);
        hipMemcpy(c->sdf, v.data(), n * 8, hipMemcpyHostToDevice);
        hipMemcpy(c->omega, o.data(), n * 8, hipMemcpyHostToDevice);
    }
    he = hipDeviceSynchronize();
    if (he != hipSuccess) {
        free_all(c);
        delete c;
        return fail(LBFFS_E_DEVICE, std::string("create: ") + hipGetErrorString(he));
    }
    *out = c;
    return 0;
}

int lbffs_upload(lbffs_ctx *c, int field, const void *host, size_t bytes)
{
    if (!c || !host) return fail(LBFFS_E_INVALID, "null argument");
    if (field < 0 || field >= LBFFS_FIELD_COUNT) return fail(LBFFS_E_INVALID, "field id");
    if (bytes != field_host_bytes(c, field))
        return fail(LBFFS_E_INVALID, "upload size " + std::to_string(bytes) + " != " +
                                         std::to_string(field_host_bytes(c, field)));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (field == LBFFS_FIELD_F) {
        int e = dispatch<UploadFOp>(c, host);
        if (!e) c->f_loaded = true;
        return e;
    }
    if (field == LBFFS_FIELD_QLINK && !c->qlink) {
        if (int e = dalloc(&c->qlink, bytes)) return e;
    }
    HIP_TRY(hipMemcpy(field_ptr(c, field), host, bytes, hipMemcpyHostToDevice));
    if (field_affects_flags(field)) c->flags_dirty = true;
    if (field == LBFFS_FIELD_GX || field == LBFFS_FIELD_GY || field == LBFFS_FIELD_GZ) c->g_dirty = true;
    return 0;
}

// step scalars that select code paths
static int params_ok(const lbffs_step_params *p)
{
    if (p->damping < 0 || p->damping > 2) return fail(LBFFS_E_INVALID, "damping must be 0, 1 or 2");
    return 0;
}

int lbffs_step(lbffs_ctx *c, int nsteps, const lbffs_step_params *p)
{
    if (!c || !p) return fail(LBFFS_E_INVALID, "null argument");
    if (nsteps < 0) return fail(LBFFS_E_INVALID, "nsteps < 0");
    if (int e = params_ok(p)) return e;
    if (!c->f_loaded) return fail(LBFFS_E_STATE, "populations were never uploaded");
    if (p->ibb && !c->qlink) return fail(LBFFS_E_STATE, "ibb = 1 needs the q-distances (LBFFS_FIELD_QLINK)");
    if (int e = aa_step_ok(c, p)) return e;
    if (c->flags_dirty) {
        if (int e = aa_restore(c)) return e;  // with the flags the array was streamed with
        int e = is_f64(c) ? pack_flags<double>(c) : pack_flags<float>(c);
        if (e) return e;
    }
    if (nsteps == 0) return 0;
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    int e = dispatch<StepOp>(c, nsteps, p);
    if (e) return e;
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    c->timed = true;
    return 0;
}

int lbffs_autotune(lbffs_ctx *c, const lbffs_step_params *p, int trial_steps, int *best_block, int *best_tune,
                   float *best_ms)
{
    if (!c || !p) return fail(LBFFS_E_INVALID, "null argument");
    if (trial_steps < 1) return fail(LBFFS_E_INVALID, "trial_steps < 1");
    if (int e = params_ok(p)) return e;
    if (!c->f_loaded) return fail(LBFFS_E_STATE, "populations were never uploaded");
    if (p->ibb && !c->qlink) return fail(LBFFS_E_STATE, "ibb = 1 needs the q-distances (LBFFS_FIELD_QLINK)");
    if (int e = aa_step_ok(c, p)) return e;
    if (c->flags_dirty) {
        if (int e = aa_restore(c)) return e;
        int e = is_f64(c) ? pack_flags<double>(c) : pack_flags<float>(c);
        if (e) return e;
    }
    if (int e = aa_restore(c)) return e;  // the saved array then restarts at a first half-step
    HIP_TRY(hipStreamSynchronize(c->stream));
    // save everything a step writes, so that tuning leaves the lattice state unchanged
    StateCopy saved;
    if (int e = saved.save(c)) return e;
    const int block0 = c->block, tune0 = c->tune;
    int rc = 0;
    hipError_t he = hipSuccess;
    const int blocks[3] = {64, 128, 256};
    // launch options: the work-group order (XCD slabs, XCD chunks or the hardware's), the
    // face cells in the same waves as their neighbours, in a second launch or in the first
    // work-groups of the one launch, and the y-tiled cell order where the lattice allows it
    int tunes[12] = {0,
                     TUNE_NO_XCD,
                     TUNE_XCD_CHUNK,
                     TUNE_FACE_FIRST,
                     TUNE_FACE_FIRST | TUNE_NO_XCD,
                     TUNE_FACE_FIRST | TUNE_XCD_CHUNK,
                     TUNE_SPLIT,
                     TUNE_SPLIT | TUNE_NO_XCD,
                     TUNE_SPLIT | TUNE_XCD_CHUNK,
                     TUNE_YTILE,
                     TUNE_FACE_FIRST | TUNE_YTILE,
                     TUNE_SPLIT | TUNE_YTILE};
    int ntunes = ytile_rows(c) ? 12 : 9;
    if (c->d.streaming == LBFFS_STREAM_AA) {  // one pass per half-step
        tunes[3] = TUNE_YTILE;
        ntunes = ytile_rows(c) ? 4 : 3;
    }
    // device time per step of one shape over n steps, after one warming step
    auto time_shape = [&](int block, int tune, int n, float *ms) {
        c->block = block;
        c->tune = tune;
        rc = dispatch<StepOp>(c, 1, p);
        if (!rc) he = hipEventRecord(c->ev0, c->stream);
        if (!rc && he == hipSuccess) rc = dispatch<StepOp>(c, n, p);
        if (!rc && he == hipSuccess) he = hipEventRecord(c->ev1, c->stream);
        if (!rc && he == hipSuccess) he = hipEventSynchronize(c->ev1);
        *ms = 0.f;
        if (!rc && he == hipSuccess) he = hipEventElapsedTime(ms, c->ev0, c->ev1);
        *ms /= n;
        return !rc && he == hipSuccess;
    };
    // every shape once; then the three fastest again over twice the steps, in two interleaved
    // rounds, keeping each one's best: shapes within a few per cent of each other are told
    // apart by more than one short trial
    struct Shape {
        float ms;
        int block, tune;
    };
    std::vector<Shape> shapes;
    for (int bi = 0; bi < 3; bi++) {
        for (int ti = 0; ti < ntunes; ti++) {
            float ms;
            if (!time_shape(blocks[bi], tunes[ti], trial_steps, &ms)) break;
            shapes.push_back({ms, blocks[bi], tunes[ti]});
        }
        if (rc || he != hipSuccess) break;
    }
    std::sort(shapes.begin(), shapes.end(), [](const Shape &x, const Shape &y) { return x.ms < y.ms; });
    const size_t nfinal = std::min<size_t>(3, shapes.size());
    std::vector<float> best(nfinal, 1e30f);
    for (int round = 0; round < 2 && rc == 0 && he == hipSuccess; round++)
        for (size_t k = 0; k < nfinal; k++) {
            float ms;
            if (!time_shape(shapes[k].block, shapes[k].tune, 2 * trial_steps, &ms)) break;
            best[k] = std::min(best[k], ms);
        }
    float bms = 1e30f;
    int bb = block0, bt = tune0;
    for (size_t k = 0; k < nfinal; k++)
        if (best[k] < bms) {
            bms = best[k];
            bb = shapes[k].block;
            bt = shapes[k].tune;
        }
    // restore the state
    const hipError_t hr = saved.restore();
    if (he == hipSuccess) he = hr;
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (rc) {
        c->block = block0;
        c->tune = tune0;
        return rc;
    }
    if (he != hipSuccess) {
        c->block = block0;
        c->tune = tune0;
        return fail(LBFFS_E_DEVICE, std::string("autotune: ") + hipGetErrorString(he));
    }
    c->block = bb;
    c->tune = bt;
    c->timed = false;
    if (best_block) *best_block = bb;
    if (best_tune) *best_tune = bt;
    if (best_ms) *best_ms = bms;
    return 0;
}

int lbffs_set_launch(lbffs_ctx *c, int block, int tune)
{
    if (!c) return fail(LBFFS_E_INVALID, "null argument");
    if (block % 64 != 0 || block < 64 || block > 256) return fail(LBFFS_E_INVALID, "block must be 64, 128, 192 or 256");
    if (tune & ~(TUNE_NT_STORE | TUNE_NO_XCD | TUNE_YTILE | TUNE_SPLIT | TUNE_NO_PURE | TUNE_FACE_FIRST |
                 TUNE_XCD_CHUNK))
        return fail(LBFFS_E_INVALID, "unknown launch option");
    c->block = block;
    c->tune = tune;
    return 0;
}

int lbffs_ibm_set_markers(lbffs_ctx *c, const void *xyz, int n)
{
    if (!c || (n > 0 && !xyz)) return fail(LBFFS_E_INVALID, "null argument");
    if (n < 0 || n > (1 << 27)) return fail(LBFFS_E_INVALID, "marker count out of range");
    HIP_TRY(hipStreamSynchronize(c->stream));
    void *ptrs[] = {c->markers, c->Gm, c->mark, c->list, c->count};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    c->markers = c->Gm = nullptr;
    c->mark = nullptr;
    c->list = c->count = nullptr;
    c->n_markers = 0;
    if (n == 0) return 0;
    const size_t rs = c->real_size, m3 = (size_t)3 * n * rs;
    int e = dalloc(&c->markers, m3);
    if (!e) e = dalloc(&c->Gm, m3);
    if (!e) e = dalloc((void **)&c->mark, (size_t)c->N * sizeof(uint32_t));
    if (!e) e = dalloc((void **)&c->list, (size_t)8 * n * sizeof(int));
    if (!e) e = dalloc((void **)&c->count, sizeof(int));
    if (e) return e;
    HIP_TRY(hipMemcpy(c->markers, xyz, m3, hipMemcpyHostToDevice));
    c->n_markers = n;
    c->stamp = 0;
    c->g_dirty = true;
    return 0;
}

int lbffs_ibm_set_motion(lbffs_ctx *c, const lbffs_ibm_motion *m)
{
    if (!c || !m) return fail(LBFFS_E_INVALID, "null argument");
    for (int a = 0; a < 3; a++) {
        c->motion[a] = m->trans[a];
        c->motion[3 + a] = m->rot_center[a];
        c->motion[6 + a] = m->rot_axis[a];
    }
    c->motion[9] = m->rot_omega;
    return 0;
}

int lbffs_ibm_download(lbffs_ctx *c, int what, void *host, int n)
{
    if (!c || !host) return fail(LBFFS_E_INVALID, "null argument");
    if (n != c->n_markers || n == 0) return fail(LBFFS_E_INVALID, "marker count differs from the markers set");
    if (what != LBFFS_IBM_POSITIONS && what != LBFFS_IBM_FORCING) return fail(LBFFS_E_INVALID, "what");
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(host, what == LBFFS_IBM_POSITIONS ? c->markers : c->Gm, (size_t)3 * n * c->real_size,
                      hipMemcpyDeviceToHost));
    return 0;
}

int lbffs_download(lbffs_ctx *c, int field, void *host, size_t bytes)
{
    if (!c || !host) return fail(LBFFS_E_INVALID, "null argument");
    if (field < 0 || field >= LBFFS_FIELD_COUNT) return fail(LBFFS_E_INVALID, "field id");
    if (bytes != field_host_bytes(c, field))
        return fail(LBFFS_E_INVALID, "download size " + std::to_string(bytes) + " != " +
                                         std::to_string(field_host_bytes(c, field)));
    if (field == LBFFS_FIELD_QLINK && !c->qlink) return fail(LBFFS_E_STATE, "q-distances were never uploaded");
    if (field == LBFFS_FIELD_F) {
        if (int e = aa_restore(c)) return e;
        return dispatch<DownloadFOp>(c, host);
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(host, field_ptr(c, field), bytes, hipMemcpyDeviceToHost));
    return 0;
}

int lbffs_reduce(lbffs_ctx *c, double out[LBFFS_REDUCE_N])
{
    if (!c || !out) return fail(LBFFS_E_INVALID, "null argument");
    if (c->flags_dirty) {
        if (int e = aa_restore(c)) return e;
        int e = is_f64(c) ? pack_flags<double>(c) : pack_flags<float>(c);
        if (e) return e;
    }
    if (int e = aa_restore(c)) return e;  // per-cell density sums in the two-lattice layout
    return dispatch<ReduceOp>(c, out, c->d_part, c->d_red);
}

int lbffs_reduce_noalloc(lbffs_ctx *c, double out[LBFFS_REDUCE_N])
{
    if (!c || !out) return fail(LBFFS_E_INVALID, "null argument");
    // rebuilding the flags (after an upload that changes them) allocates: refused here
    if (c->flags_dirty)
        return fail(LBFFS_E_STATE, "the lattice's flags must be rebuilt first (lbffs_step, lbffs_reduce or "
                                   "lbffs_autotune does; lbffs_reduce_noalloc allocates nothing)");
    // the partials fill exactly kRedBlocks rows of the fixed scratch array, whatever the
    // lattice size (grid-stride kernel); a reduction asking for more rows is refused
    const long rows = kRedBlocks;
    if (rows * LBFFS_REDUCE_N + LBFFS_REDUCE_N > (long)(sizeof(scratch) / sizeof(double)))
        return fail(LBFFS_E_INVALID, "reduction needs more partial rows than the scratch array holds");
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != c->device) HIP_TRY(hipSetDevice(c->device));
    double *base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
    int e = base ? aa_restore(c) : fail(LBFFS_E_STATE, "no scratch address for this device");
    if (!e) e = dispatch<ReduceOp>(c, out, base, base + rows * LBFFS_REDUCE_N);
    if (cur != c->device) (void)hipSetDevice(cur);
    return e;
}

int lbffs_output_macros(lbffs_ctx *c, float dpdx, void *rho, void *u, void *v, void *w, size_t bytes)
{
    if (!c || !rho || !u || !v || !w) return fail(LBFFS_E_INVALID, "null argument");
    if (!c->f_loaded) return fail(LBFFS_E_STATE, "populations not uploaded");
    if (bytes != (size_t)c->N * (size_t)c->real_size)
        return fail(LBFFS_E_INVALID, "output_macros size " + std::to_string(bytes) + " != " +
                                         std::to_string((size_t)c->N * (size_t)c->real_size));
    if (int e = aa_restore(c)) return e;
    HIP_TRY(hipStreamSynchronize(c->stream));
    void *const h[4] = {rho, u, v, w};
    return dispatch<OutputMacrosOp>(c, (double)dpdx, h);
}

int lbffs_synchronize(lbffs_ctx *c)
{
    if (!c) return fail(LBFFS_E_INVALID, "null argument");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int lbffs_last_step_ms(lbffs_ctx *c, float *ms)
{
    if (!c || !ms) return fail(LBFFS_E_INVALID, "null argument");
    if (!c->timed) return fail(LBFFS_E_STATE, "no step has run");
    HIP_TRY(hipEventSynchronize(c->ev1));
    HIP_TRY(hipEventElapsedTime(ms, c->ev0, c->ev1));
    return 0;
}

int lbffs_destroy(lbffs_ctx *c)
{
    if (!c) return 0;
    hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return 0;
}

const char *lbffs_strerror(int code)
{
    switch (code) {
    case LBFFS_OK: return "ok";
    case LBFFS_E_INVALID: return "invalid argument";
    case LBFFS_E_NOMEM: return "device out of memory";
    case LBFFS_E_DEVICE: return "HIP runtime error";
    case LBFFS_E_UNSUPPORTED: return "unsupported configuration";
    case LBFFS_E_STATE: return "call out of order";
    }
    return "unknown error";
}

const char *lbffs_last_error(void) { return g_last_error.c_str(); }

int lbffs_batch_create(const lbffs_desc *descs, int n, lbffs_batch **out)
{
    if (!descs || !out) return fail(LBFFS_E_INVALID, "null argument");
    *out = nullptr;
    if (n < 1) return fail(LBFFS_E_INVALID, "a batch needs at least one lattice");
    for (int d = 1; d < n; d++)
        if (descs[d].q != descs[0].q || descs[d].storage != descs[0].storage || descs[d].sgs != descs[0].sgs ||
            descs[d].collision != descs[0].collision)
            return fail(LBFFS_E_INVALID, "the lattices of a batch share q, storage, sgs and collision");
    for (int d = 0; d < n; d++)
        if (descs[d].streaming != LBFFS_STREAM_TWO_LATTICE)
            return fail(LBFFS_E_UNSUPPORTED, "the lattices of a batch stream between two arrays");
    lbffs_batch *b = new lbffs_batch();
    hipError_t he = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreate(&b->ev0);
    if (he == hipSuccess) he = hipEventCreate(&b->ev1);
    if (he != hipSuccess) {
        batch_free(b);
        delete b;
        return fail(LBFFS_E_DEVICE, std::string("batch stream: ") + hipGetErrorString(he));
    }
    std::vector<BatchBlock> table;
    for (int d = 0; d < n; d++) {
        lbffs_ctx *c = nullptr;
        int e = lbffs_create(&descs[d], &c);
        if (e) {
            std::string m = g_last_error;
            batch_free(b);
            delete b;
            return fail(e, "lattice " + std::to_string(d) + ": " + m);
        }
        (void)hipStreamDestroy(c->stream);
        c->stream = b->stream;
        c->owns_stream = false;
        b->ctx.push_back(c);
        for (long first = 0; first < c->N; first += kBatchBlock) table.push_back({(uint32_t)d, (uint32_t)first, 0u});
    }
    b->nblocks = (uint32_t)table.size();
    const size_t amax = sizeof(StepArgs<double, 27>) * (size_t)n;
    int e = 0;
    for (int v = 0; v < 4 && !e; v++) e = dalloc(&b->args_dev[v], amax);
    if (!e) e = dalloc((void **)&b->ctr, sizeof(uint32_t));
    if (!e) e = dalloc((void **)&b->table, table.size() * sizeof(BatchBlock));
    if (!e && hipMemcpy(b->table, table.data(), table.size() * sizeof(BatchBlock), hipMemcpyHostToDevice) != hipSuccess)
        e = fail(LBFFS_E_DEVICE, "batch table upload");
    if (e) {
        batch_free(b);
        delete b;
        return e;
    }
    *out = b;
    return 0;
}

int lbffs_batch_context(lbffs_batch *b, int i, lbffs_ctx **ctx)
{
    if (!b || !ctx) return fail(LBFFS_E_INVALID, "null argument");
    if (i < 0 || i >= (int)b->ctx.size()) return fail(LBFFS_E_INVALID, "lattice index");
    *ctx = b->ctx[i];
    return 0;
}

int lbffs_batch_set_launch(lbffs_batch *b, int grid, int graph_chunk)
{
    if (!b) return fail(LBFFS_E_INVALID, "null argument");
    if (grid < 0 || graph_chunk < 0 || graph_chunk > 4096) return fail(LBFFS_E_INVALID, "grid/graph_chunk");
    b->grid = grid;
    b->graph_chunk = graph_chunk;
    batch_drop_graphs(b);
    return 0;
}

int lbffs_batch_step(lbffs_batch *b, int nsteps, const lbffs_step_params *params)
{
    if (!b || !params) return fail(LBFFS_E_INVALID, "null argument");
    if (nsteps < 0) return fail(LBFFS_E_INVALID, "nsteps < 0");
    HIP_TRY(hipStreamSynchronize(b->stream));  // the argument arrays are rewritten below
    const int cur0 = b->ctx[0]->cur;
    for (size_t d = 0; d < b->ctx.size(); d++) {
        lbffs_ctx *c = b->ctx[d];
        if (!c->f_loaded) return fail(LBFFS_E_STATE, "lattice " + std::to_string(d) + ": populations never uploaded");
        if (ibm_active(c)) return fail(LBFFS_E_UNSUPPORTED, "moving-wall markers in a batch");
        if (params[d].ibb && !c->qlink)
            return fail(LBFFS_E_STATE, "lattice " + std::to_string(d) + ": ibb = 1 needs the q-distances");
        if (int e = params_ok(&params[d])) return e;
        if (c->cur != cur0 || c->mcur != cur0)
            return fail(LBFFS_E_STATE, "lattice " + std::to_string(d) + " was stepped on its own: step parity differs");
        if (c->flags_dirty) {
            int e = is_f64(c) ? pack_flags<double>(c) : pack_flags<float>(c);
            if (e) return fail(e, "lattice " + std::to_string(d) + ": " + g_last_error);
        }
        // the flags may also have been rebuilt by a call on the lattice's own context
        // (lbffs_reduce, lbffs_step, lbffs_autotune): compare generations, not the dirty bit
        if (b->gen.size() != b->ctx.size() || b->gen[d] != c->flags_gen) b->ftable_ok = false;
        // a lattice whose kernels stopped (or started) reading its wave summary: the table
        // must (or may) hold its XMIX and WRAP face cells
        if (b->wc.size() != b->ctx.size() || b->wc[d] != (uint8_t)uses_wave_classes(c)) b->ftable_ok = false;
    }
    if (!b->ftable_ok) {
        if (int e = batch_face_table(b)) return e;
    }
    if (nsteps == 0) return 0;
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    int e = batch_dispatch<BatchStepOp>(b, nsteps, params);
    if (e) return e;
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    for (lbffs_ctx *c : b->ctx) {
        c->cur ^= (nsteps & 1);
        c->mcur ^= (nsteps & 1);
        c->nstep += (uint32_t)nsteps;
        c->timed = false;
    }
    b->timed = true;
    return 0;
}

int lbffs_batch_last_step_ms(lbffs_batch *b, float *ms)
{
    if (!b || !ms) return fail(LBFFS_E_INVALID, "null argument");
    if (!b->timed) return fail(LBFFS_E_STATE, "no step has run");
    HIP_TRY(hipEventSynchronize(b->ev1));
    HIP_TRY(hipEventElapsedTime(ms, b->ev0, b->ev1));
    return 0;
}

int lbffs_batch_autotune(lbffs_batch *b, const lbffs_step_params *params, int trial_steps, int *best_grid,
                         int *best_chunk, float *best_ms)
{
    if (!b || !params) return fail(LBFFS_E_INVALID, "null argument");
    if (trial_steps < 1) return fail(LBFFS_E_INVALID, "trial_steps < 1");
    // checks every lattice, packs changed flags and rebuilds the face table; steps nothing
    if (int e = lbffs_batch_step(b, 0, params)) return e;
    const size_t n = b->ctx.size();
    std::unique_ptr<StateCopy[]> saved(new StateCopy[n]);
    for (size_t d = 0; d < n; d++)
        if (int e = saved[d].save(b->ctx[d])) return e;
    const int grid0 = b->grid, chunk0 = b->graph_chunk;
    // one work-group per 256-cell block, or persistent grids of 8, 16 and 32 work-groups per
    // CU where the table has more blocks than that; plain launches or captured 16-step chunks
    const int grids[4] = {0, 2048, 4096, 8192};
    const int chunks[2] = {0, 16};
    float bms = 1e30f;
    int bg = grid0, bc = chunk0, rc = 0;
    for (int gi = 0; gi < 4 && !rc; gi++) {
        if (grids[gi] > 0 && (uint32_t)grids[gi] >= b->nblocks) continue;
        for (int ci = 0; ci < 2 && !rc; ci++) {
            b->grid = grids[gi];
            b->graph_chunk = chunks[ci];
            batch_drop_graphs(b);  // a captured graph keeps the grid it was captured with
            // two warm calls of odd length capture the graphs of both step parities
            const int warm = chunks[ci] + 1;
            rc = lbffs_batch_step(b, warm, params);
            if (!rc) rc = lbffs_batch_step(b, warm, params);
            if (!rc) rc = lbffs_batch_step(b, trial_steps, params);
            float ms = 0.f;
            if (!rc) rc = lbffs_batch_last_step_ms(b, &ms);
            if (!rc && ms / trial_steps < bms) {
                bms = ms / trial_steps;
                bg = grids[gi];
                bc = chunks[ci];
            }
        }
    }
    hipError_t he = hipStreamSynchronize(b->stream);
    for (size_t d = 0; d < n; d++) {
        const hipError_t hr = saved[d].restore();
        if (he == hipSuccess) he = hr;
    }
    if (he == hipSuccess) he = hipStreamSynchronize(b->stream);
    b->timed = false;
    batch_drop_graphs(b);
    if (rc || he != hipSuccess) {
        b->grid = grid0;
        b->graph_chunk = chunk0;
        return rc ? rc : fail(LBFFS_E_DEVICE, std::string("batch autotune: ") + hipGetErrorString(he));
    }
    b->grid = bg;
    b->graph_chunk = bc;
    if (best_grid) *best_grid = bg;
    if (best_chunk) *best_chunk = bc;
    if (best_ms) *best_ms = bms;
    return 0;
}

int lbffs_batch_destroy(lbffs_batch *b)
{
    if (!b) return 0;
    batch_free(b);
    delete b;
    return 0;
}

}  // extern "C"

// ---- geometry preprocessing (include/lbffs_geom.h) ----

namespace {

// device buffers of one geometry call, freed on every return path
struct GeomScratch {
    std::vector<void *> bufs;
    template <typename T>
    int alloc(T **p, size_t count)
    {
        void *q = nullptr;
        int e = dalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 1);
        if (e) return e;
        bufs.push_back(q);
        *p = (T *)q;
        return 0;
    }
    ~GeomScratch()
    {
        for (void *q : bufs) hipFree(q);
    }
};

int geom_dims_ok(int nx, int ny, int nz)
{
    if (nx < 1 || ny < 1 || nz < 1) return fail(LBFFS_E_INVALID, "lattice dimensions must be positive");
    if ((double)nx * ny * nz >= 4294967295.0) return fail(LBFFS_E_INVALID, "lattice too large (2^32 cells or more)");
    return 0;
}

unsigned geom_grid(size_t n)
{
    return (unsigned)std::min<size_t>((n + 255) / 256, 65536);
}

// The surface of a q-distance cast: triangles in FP64 and their buckets. Every
// kQBucket^3-cell bucket that a triangle's box grown by one cell overlaps lists it, so the
// links of a cell (at most one cell long per axis) meet only triangles of the cell's bucket.
struct QlinkSurface {
    std::vector<lbffs::GeomTri64> tris;
    std::vector<int> bstart, btri;
    int bx = 0, by = 0, bz = 0;
};

int qlink_surface(int n, const float *v0, const float *v1, const float *v2, int nx, int ny, int nz, QlinkSurface &s)
{
    const int B = lbffs::kQBucket;
    s.bx = (nx + B - 1) / B;
    s.by = (ny + B - 1) / B;
    s.bz = (nz + B - 1) / B;
    const size_t nb = (size_t)s.bx * s.by * s.bz;
    s.tris.resize((size_t)n);
    s.bstart.assign(nb + 1, 0);
    std::vector<std::array<int, 6>> range((size_t)n);
    for (int t = 0; t < n; t++) {
        double a[3], b[3], c[3];
        for (int ax = 0; ax < 3; ax++) {
            a[ax] = v0[(size_t)ax * n + t];
            b[ax] = v1[(size_t)ax * n + t];
            c[ax] = v2[(size_t)ax * n + t];
            s.tris[t].v0[ax] = a[ax];
            s.tris[t].e1[ax] = b[ax] - a[ax];
            s.tris[t].e2[ax] = c[ax] - a[ax];
        }
        const int dims[3] = {s.bx, s.by, s.bz};
        bool ok = true;
        for (int ax = 0; ax < 3; ax++) {
            const double lo = std::min(a[ax], std::min(b[ax], c[ax])) - 1.0 - 1e-9;
            const double hi = std::max(a[ax], std::max(b[ax], c[ax])) + 1.0 + 1e-9;
            if (!(lo <= hi)) ok = false;  // NaN vertex
            const double l = std::max(std::floor(lo / B), 0.0), h = std::min(std::floor(hi / B), (double)dims[ax] - 1);
            if (l > h) ok = false;
            range[t][2 * ax] = ok ? (int)l : 0;
            range[t][2 * ax + 1] = ok ? (int)h : -1;
        }
        if (!ok) range[t] = {0, -1, 0, -1, 0, -1};
        for (int kk = range[t][4]; kk <= range[t][5]; kk++)
            for (int jj = range[t][2]; jj <= range[t][3]; jj++)
                for (int ii = range[t][0]; ii <= range[t][1]; ii++)
                    s.bstart[(size_t)ii + s.bx * ((size_t)jj + s.by * kk) + 1]++;
    }
    for (size_t b = 0; b < nb; b++) s.bstart[b + 1] += s.bstart[b];
    if ((size_t)s.bstart[nb] > 0x7fffffffu) return fail(LBFFS_E_INVALID, "too many bucket entries");
    s.btri.resize((size_t)s.bstart[nb]);
    std::vector<int> fill(s.bstart.begin(), s.bstart.end() - 1);
    for (int t = 0; t < n; t++)
        for (int kk = range[t][4]; kk <= range[t][5]; kk++)
            for (int jj = range[t][2]; jj <= range[t][3]; jj++)
                for (int ii = range[t][0]; ii <= range[t][1]; ii++)
                    s.btri[(size_t)fill[(size_t)ii + s.bx * ((size_t)jj + s.by * kk)]++] = t;
    return 0;
}

// the fluid cells next to a solid one (neiSolid, walls.hpp:187-216): the cells with solid links
std::vector<uint32_t> qlink_cells(size_t N, const uint8_t *solid, const uint8_t *neiSolid)
{
    std::vector<uint32_t> cells;
    for (size_t ic = 0; ic < N; ic++)
        if (neiSolid[ic] == 1 && solid[ic] == 0) cells.push_back((uint32_t)ic);
    return cells;
}

// Casts every solid link of the listed cells against the surface; q of the links that meet
// a triangle is written to d_q (q-major, q*N values), the others keep what d_q holds and
// are counted in *misses. Returns after the kernel has finished.
template <typename T>
int qlink_cast(int q, const QlinkSurface &s, int nx, int ny, int nz, const std::vector<uint32_t> &cells,
               const uint8_t *d_solid, T *d_q, hipStream_t st, int *misses)
{
    GeomScratch g;
    lbffs::GeomTri64 *d_tris;
    int *d_bstart, *d_btri;
    uint32_t *d_cells, *d_miss;
    if (int e = g.alloc(&d_tris, s.tris.size())) return e;
    if (int e = g.alloc(&d_bstart, s.bstart.size())) return e;
    if (int e = g.alloc(&d_btri, s.btri.size())) return e;
    if (int e = g.alloc(&d_cells, cells.size())) return e;
    if (int e = g.alloc(&d_miss, 1)) return e;  // zeroed by the allocation
    if (!s.tris.empty())
        HIP_TRY(hipMemcpy(d_tris, s.tris.data(), s.tris.size() * sizeof(lbffs::GeomTri64), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_bstart, s.bstart.data(), s.bstart.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!s.btri.empty()) HIP_TRY(hipMemcpy(d_btri, s.btri.data(), s.btri.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!cells.empty())
        HIP_TRY(hipMemcpy(d_cells, cells.data(), cells.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (!cells.empty()) {
        const unsigned grid = (unsigned)((cells.size() + 255) / 256);
        if (q == 27)
            hipLaunchKernelGGL((lbffs::geom_qlinks_kernel<27, T>), dim3(grid), dim3(256), 0, st, d_cells,
                               (uint32_t)cells.size(), d_tris, d_bstart, d_btri, s.bx, s.by, s.bz, nx, ny, nz, d_solid,
                               d_q, d_miss);
        else
            hipLaunchKernelGGL((lbffs::geom_qlinks_kernel<19, T>), dim3(grid), dim3(256), 0, st, d_cells,
                               (uint32_t)cells.size(), d_tris, d_bstart, d_btri, s.bx, s.by, s.bz, nx, ny, nz, d_solid,
                               d_q, d_miss);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(st));
    uint32_t miss = 0;
    HIP_TRY(hipMemcpy(&miss, d_miss, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (misses) *misses = (int)miss;
    return 0;
}

}  // namespace

extern "C" {

int lbffs_geom_solid(int n, const float *v0, const float *v1, const float *v2, int nx, int ny, int nz, int invert,
                     uint8_t *solid)
{
    if (!solid || (n > 0 && (!v0 || !v1 || !v2))) return fail(LBFFS_E_INVALID, "null argument");
    if (n < 0) return fail(LBFFS_E_INVALID, "negative triangle count");
    if (int e = geom_dims_ok(nx, ny, nz)) return e;
    const size_t N = (size_t)nx * ny * nz;
    // bounding box of the surface, grown from the origin (stl.hpp:525-550), and of each
    // triangle (stl.hpp:224-262); min/max are exact, so the order of the reduction is free
    lbffs::GeomBox box = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    std::vector<lbffs::GeomTri> tris((size_t)n);
    for (int t = 0; t < n; t++) {
        lbffs::GeomTri &g = tris[t];
        for (int a = 0; a < 3; a++) {
            g.v0[a] = v0[(size_t)a * n + t];
            g.v1[a] = v1[(size_t)a * n + t];
            g.v2[a] = v2[(size_t)a * n + t];
        }
        g.ymin = fminf(g.v0[1], fminf(g.v1[1], g.v2[1]));
        g.ymax = fmaxf(g.v0[1], fmaxf(g.v1[1], g.v2[1]));
        g.zmin = fminf(g.v0[2], fminf(g.v1[2], g.v2[2]));
        g.zmax = fmaxf(g.v0[2], fmaxf(g.v1[2], g.v2[2]));
        g.pad = 0.f;
        box.xmin = fminf(box.xmin, fminf(g.v0[0], fminf(g.v1[0], g.v2[0])));
        box.xmax = fmaxf(box.xmax, fmaxf(g.v0[0], fmaxf(g.v1[0], g.v2[0])));
        box.ymin = fminf(box.ymin, g.ymin);
        box.ymax = fmaxf(box.ymax, g.ymax);
        box.zmin = fminf(box.zmin, g.zmin);
        box.zmax = fmaxf(box.zmax, g.zmax);
    }
    // triangles by the z layers their box may cover (a superset: the kernel applies the
    // exact test of walls.hpp:116-121); a box with a NaN bound covers no layer, as the
    // reference's comparisons fail for it
    std::vector<int> bin_start((size_t)nz + 1, 0), bin_tri;
    {
        std::vector<std::pair<int, int>> span((size_t)n, {1, 0});
        for (int t = 0; t < n; t++) {
            const float lo = tris[t].zmin, hi = tris[t].zmax;
            if (!(lo <= hi)) continue;
            const double klo = std::max(std::floor((double)lo) - 1.0, 0.0);
            const double khi = std::min(std::ceil((double)hi) + 1.0, (double)(nz - 1));
            if (klo > khi) continue;
            span[t] = {(int)klo, (int)khi};
            for (int k = (int)klo; k <= (int)khi; k++) bin_start[k + 1]++;
        }
        for (int k = 0; k < nz; k++) bin_start[k + 1] += bin_start[k];
        bin_tri.resize((size_t)bin_start[nz]);
        std::vector<int> fill(bin_start.begin(), bin_start.end() - 1);
        for (int t = 0; t < n; t++)
            for (int k = span[t].first; k <= span[t].second; k++) bin_tri[(size_t)fill[k]++] = t;
    }
    GeomScratch g;
    uint8_t *d_solid;
    lbffs::GeomTri *d_tris;
    int *d_start, *d_bins;
    if (int e = g.alloc(&d_solid, N)) return e;
    if (int e = g.alloc(&d_tris, tris.size())) return e;
    if (int e = g.alloc(&d_start, bin_start.size())) return e;
    if (int e = g.alloc(&d_bins, bin_tri.size())) return e;
    HIP_TRY(hipMemcpy(d_solid, solid, N, hipMemcpyHostToDevice));
    if (n) HIP_TRY(hipMemcpy(d_tris, tris.data(), tris.size() * sizeof(lbffs::GeomTri), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_start, bin_start.data(), bin_start.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!bin_tri.empty()) HIP_TRY(hipMemcpy(d_bins, bin_tri.data(), bin_tri.size() * sizeof(int), hipMemcpyHostToDevice));
    const int chunks = (nx + lbffs::kGeomRow - 1) / lbffs::kGeomRow;
    const size_t blocks = (size_t)chunks * ny * nz;
    if (blocks > 0x7fffffffu) return fail(LBFFS_E_INVALID, "lattice too large for the solid launch");
    hipLaunchKernelGGL(lbffs::geom_solid_kernel, dim3((unsigned)blocks), dim3(lbffs::kGeomRow), 0, nullptr, d_tris,
                       d_start, d_bins, nx, ny, chunks, box, d_solid);
    HIP_TRY(hipGetLastError());
    if (invert) {
        hipLaunchKernelGGL(lbffs::geom_invert_kernel, dim3(geom_grid(N)), dim3(256), 0, nullptr, d_solid, N);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpy(solid, d_solid, N, hipMemcpyDeviceToHost));
    return 0;
}

int lbffs_geom_qlinks(int q, int n, const float *v0, const float *v1, const float *v2, int nx, int ny, int nz,
                      const uint8_t *solid, const uint8_t *neiSolid, float *qlink, int *misses)
{
    if (!solid || !neiSolid || !qlink || (n > 0 && (!v0 || !v1 || !v2))) return fail(LBFFS_E_INVALID, "null argument");
    if (q != 19 && q != 27) return fail(LBFFS_E_INVALID, "q must be 19 or 27");
    if (n < 0) return fail(LBFFS_E_INVALID, "negative triangle count");
    if (int e = geom_dims_ok(nx, ny, nz)) return e;
    const size_t N = (size_t)nx * ny * nz;
    if ((double)q * N >= 4294967295.0) return fail(LBFFS_E_INVALID, "lattice too large for the q-distance array");
    QlinkSurface surf;
    if (int e = qlink_surface(n, v0, v1, v2, nx, ny, nz, surf)) return e;
    const std::vector<uint32_t> cells = qlink_cells(N, solid, neiSolid);
    GeomScratch g;
    uint8_t *d_solid;
    float *d_q;
    if (int e = g.alloc(&d_solid, N)) return e;
    if (int e = g.alloc(&d_q, (size_t)q * N)) return e;
    HIP_TRY(hipMemcpy(d_solid, solid, N, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_q, qlink, (size_t)q * N * sizeof(float), hipMemcpyHostToDevice));
    if (int e = qlink_cast(q, surf, nx, ny, nz, cells, d_solid, d_q, nullptr, misses)) return e;
    HIP_TRY(hipMemcpy(qlink, d_q, (size_t)q * N * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int lbffs_geom_qlinks_ctx(lbffs_ctx *c, int n, const float *v0, const float *v1, const float *v2, float init,
                          int *misses)
{
    if (!c || (n > 0 && (!v0 || !v1 || !v2))) return fail(LBFFS_E_INVALID, "null argument");
    if (n < 0) return fail(LBFFS_E_INVALID, "negative triangle count");
    const int q = c->d.q, nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
    const size_t N = c->N;
    QlinkSurface surf;
    if (int e = qlink_surface(n, v0, v1, v2, nx, ny, nz, surf)) return e;
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<uint8_t> solid(N), nei(N);
    HIP_TRY(hipMemcpy(solid.data(), c->solid, N, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(nei.data(), c->neiSolid, N, hipMemcpyDeviceToHost));
    const std::vector<uint32_t> cells = qlink_cells(N, solid.data(), nei.data());
    if (!c->qlink) {
        if (int e = dalloc(&c->qlink, (size_t)q * N * c->real_size)) return e;
    }
    if (c->real_size == 8) {
        hipLaunchKernelGGL((fill_kernel<double>), dim3(4096), dim3(256), 0, c->stream, (double *)c->qlink, (size_t)q * N,
                           (double)init);
        HIP_TRY(hipGetLastError());
        return qlink_cast(q, surf, nx, ny, nz, cells, c->solid, (double *)c->qlink, c->stream, misses);
    }
    hipLaunchKernelGGL((fill_kernel<float>), dim3(4096), dim3(256), 0, c->stream, (float *)c->qlink, (size_t)q * N, init);
    HIP_TRY(hipGetLastError());
    return qlink_cast(q, surf, nx, ny, nz, cells, c->solid, (float *)c->qlink, c->stream, misses);
}

int lbffs_geom_sdf(int n, const float *centroid, const float *normal, float sdf_ini, float dr, float p, int nx, int ny,
                   int nz, const uint8_t *neiSolid, float *sdf)
{
    if (!neiSolid || !sdf || (n > 0 && (!centroid || !normal))) return fail(LBFFS_E_INVALID, "null argument");
    if (n < 0) return fail(LBFFS_E_INVALID, "negative triangle count");
    if (int e = geom_dims_ok(nx, ny, nz)) return e;
    if (!(dr >= 0.f) || dr > 1e6f) return fail(LBFFS_E_INVALID, "dr must be in [0, 1e6]");
    const size_t N = (size_t)nx * ny * nz;
    const int drn = (int)dr;
    GeomScratch g;
    uint8_t *d_nei, *d_m1, *d_m2;
    uint32_t *d_list, *d_count;
    float *d_tri, *d_sdf;
    if (int e = g.alloc(&d_nei, N)) return e;
    if (int e = g.alloc(&d_m1, N)) return e;
    if (int e = g.alloc(&d_m2, N)) return e;
    if (int e = g.alloc(&d_list, N)) return e;
    if (int e = g.alloc(&d_count, 1)) return e;
    if (int e = g.alloc(&d_tri, (size_t)6 * n)) return e;
    if (int e = g.alloc(&d_sdf, N)) return e;
    HIP_TRY(hipMemcpy(d_nei, neiSolid, N, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_sdf, sdf, N * sizeof(float), hipMemcpyHostToDevice));
    if (n) {
        HIP_TRY(hipMemcpy(d_tri, centroid, (size_t)3 * n * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_tri + (size_t)3 * n, normal, (size_t)3 * n * sizeof(float), hipMemcpyHostToDevice));
    }
    const unsigned gr = geom_grid(N);
    hipLaunchKernelGGL(lbffs::geom_dilate_kernel, dim3(gr), dim3(256), 0, nullptr, d_nei, d_m1, nx, ny, nz, 0, drn, 1);
    hipLaunchKernelGGL(lbffs::geom_dilate_kernel, dim3(gr), dim3(256), 0, nullptr, d_m1, d_m2, nx, ny, nz, 1, drn, 0);
    hipLaunchKernelGGL(lbffs::geom_dilate_kernel, dim3(gr), dim3(256), 0, nullptr, d_m2, d_m1, nx, ny, nz, 2, drn, 0);
    hipLaunchKernelGGL(lbffs::geom_near_list_kernel, dim3(gr), dim3(256), 0, nullptr, d_m1, N, d_list, d_count);
    HIP_TRY(hipGetLastError());
    uint32_t count = 0;
    HIP_TRY(hipMemcpy(&count, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (count > N) return fail(LBFFS_E_DEVICE, "near-cell count exceeds the lattice");
    if (count) {
        hipLaunchKernelGGL(lbffs::geom_sdf_kernel, dim3((count + 255) / 256), dim3(256), 0, nullptr, d_list, count, d_tri,
                           n, nx, ny, sdf_ini, dr, p, d_sdf);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpy(sdf, d_sdf, N * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
