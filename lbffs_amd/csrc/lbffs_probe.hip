// lbffs_probe.hip -- HBM bandwidth probes: plain streaming kernels of known size.
// Their durations give the practical HBM ceilings next to the 8 TB/s peak, and the
// copy kernel's FETCH_SIZE / WRITE_SIZE counters calibrate the counter-to-bytes
// factor used to turn the step kernel's counters into HBM traffic per launch.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/lbffs.h"
#include "lbffs_common.hpp"

namespace {

using lbffs::pop_arrays;
using lbffs::pop_lb;

// mode 0: grid-stride float4 copy
__global__ void __launch_bounds__(256) lbffs_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                         size_t n4)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

// mode 1: four float4 per thread, all loads issued before the stores
__global__ void __launch_bounds__(256) lbffs_copy4_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                          size_t n4)
{
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const size_t i = base + (size_t)k * 256;
        if (i < n4) v[k] = src[i];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const size_t i = base + (size_t)k * 256;
        if (i < n4) dst[i] = v[k];
    }
}

// mode 2: read only (per-thread sum, one float written per thread)
__global__ void __launch_bounds__(256) lbffs_read_kernel(const float4 *__restrict__ src, float *__restrict__ sink,
                                                         size_t n4)
{
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const size_t i = base + (size_t)k * 256;
        if (i < n4) {
            float4 v = src[i];
            s += v.x + v.y + v.z + v.w;
        }
    }
    if (s == 12345.678f) sink[blockIdx.x] = s;  // keeps the loads alive, never true for zeros
}

// mode 3: write only
__global__ void __launch_bounds__(256) lbffs_fill_kernel(float4 *__restrict__ dst, size_t n4)
{
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const size_t i = base + (size_t)k * 256;
        if (i < n4) dst[i] = make_float4(1.f, 2.f, 3.f, 4.f);
    }
}

// ---- configurable streaming probes (lbffs_probe_stream) ----
typedef float pf4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ inline pf4 ld4(const pf4 *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ inline void st4(pf4 *p, pf4 v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// KIND 0 copy, 1 read, 2 write.  Tiles of U * blockDim.x float4; each work-group walks
// tiles blockIdx.x, blockIdx.x + gridDim.x, ... with all U loads of a tile issued before
// its stores.
template <int KIND, int U, bool NT>
__global__ void __launch_bounds__(1024) probe_stream_kernel(const pf4 *__restrict__ src, pf4 *__restrict__ dst,
                                                            size_t n4, size_t ntiles, float *__restrict__ sink)
{
    const size_t tile = (size_t)U * blockDim.x;
    pf4 acc = {0.f, 0.f, 0.f, 0.f};
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const size_t base = t * tile + threadIdx.x;
        pf4 v[U];
        if constexpr (KIND != 2) {
#pragma unroll
            for (int k = 0; k < U; k++) {
                const size_t i = base + (size_t)k * blockDim.x;
                v[k] = i < n4 ? ld4<NT>(src + i) : pf4{0.f, 0.f, 0.f, 0.f};
            }
        } else {
#pragma unroll
            for (int k = 0; k < U; k++) v[k] = pf4{1.f, 2.f, 3.f, (float)k};
        }
        if constexpr (KIND == 1) {
#pragma unroll
            for (int k = 0; k < U; k++) acc += v[k];
        } else {
#pragma unroll
            for (int k = 0; k < U; k++) {
                const size_t i = base + (size_t)k * blockDim.x;
                if (i < n4) st4<NT>(dst + i, v[k]);
            }
        }
    }
    if (KIND == 1 && acc.x + acc.y + acc.z + acc.w == 12345.678f) sink[blockIdx.x] = acc.x;
}

template <int KIND, int U>
void launch_stream_u(bool nt, unsigned grid, int block, const pf4 *a, pf4 *b, size_t n4, size_t ntiles, float *sink)
{
    if (nt)
        hipLaunchKernelGGL((probe_stream_kernel<KIND, U, true>), dim3(grid), dim3(block), 0, 0, a, b, n4, ntiles, sink);
    else
        hipLaunchKernelGGL((probe_stream_kernel<KIND, U, false>), dim3(grid), dim3(block), 0, 0, a, b, n4, ntiles, sink);
}

template <int KIND>
bool launch_stream_k(int unroll, bool nt, unsigned grid, int block, const pf4 *a, pf4 *b, size_t n4, size_t ntiles,
                     float *sink)
{
    switch (unroll) {
    case 1: launch_stream_u<KIND, 1>(nt, grid, block, a, b, n4, ntiles, sink); return true;
    case 2: launch_stream_u<KIND, 2>(nt, grid, block, a, b, n4, ntiles, sink); return true;
    case 4: launch_stream_u<KIND, 4>(nt, grid, block, a, b, n4, ntiles, sink); return true;
    case 8: launch_stream_u<KIND, 8>(nt, grid, block, a, b, n4, ntiles, sink); return true;
    case 16: launch_stream_u<KIND, 16>(nt, grid, block, a, b, n4, ntiles, sink); return true;
    }
    return false;
}

// ---- memory skeleton of the lattice step (lbffs_probe_lattice) ----
// Every cell pulls its Q populations from the previous array with the step's offsets and
// stores them, unchanged, to its own slots of the next array: the step's memory traffic
// without the collision.  Arrays carry a guard of `guard` elements on both sides so that
// every pull of every cell stays inside the allocation.
// VARIANT 0: pulls with the lattice offsets (x-shifted runs straddle cache lines);
// VARIANT 1: the x component of every offset dropped (every run aligned: the traffic of
//            an ideal pull);
// VARIANT 2: x-shifted directions loaded aligned and moved one lane with a DPP wave shift;
//            the lane that leaves the wave loads its value alone;
// VARIANT 3: as 2, with the leaving lane's value passed between the waves of a work-group
//            through LDS (only the work-group's first / last wave loads it).
template <int Q>
struct ProbeDirs;
template <>
struct ProbeDirs<19> {
    static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
};
template <>
struct ProbeDirs<27> {
    static constexpr int cx[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    static constexpr int cy[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
    static constexpr int cz[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
};

template <typename T>
__device__ inline T dpp_shift(T old, T v, bool right)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 2, "32- or 16-bit element");
    int o, x;
    if constexpr (sizeof(T) == 4) {
        o = __builtin_bit_cast(int, old);
        x = __builtin_bit_cast(int, v);
    } else {
        o = (int)__builtin_bit_cast(unsigned short, old);
        x = (int)__builtin_bit_cast(unsigned short, v);
    }
    // wave_shr:1 (lane l takes lane l - 1) / wave_shl:1 (lane l takes lane l + 1); the lane
    // with no source keeps `old`
    const int r = right ? __builtin_amdgcn_update_dpp(o, x, 0x138, 0xf, 0xf, false)
                        : __builtin_amdgcn_update_dpp(o, x, 0x130, 0xf, 0xf, false);
    if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, r);
    else return __builtin_bit_cast(T, (unsigned short)r);
}

template <typename T, int Q, int VARIANT>
__global__ void __launch_bounds__(256) probe_lattice_kernel(const T *__restrict__ src, T *__restrict__ dst, size_t FS,
                                                            int sy, int sz)
{
    __shared__ T edge[2][Q][4];
    const long ic = (long)blockIdx.x * 256 + threadIdx.x;
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
    T v[Q], ed[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int cx = ProbeDirs<Q>::cx[q];
        const long yz = (long)ProbeDirs<Q>::cy[q] * sy + (long)ProbeDirs<Q>::cz[q] * sz;
        const T *s = src + (size_t)q * FS;
        ed[q] = T(0);
        if (VARIANT == 0 || cx == 0) {
            v[q] = s[ic - yz - (VARIANT == 0 ? cx : 0)];
        } else if (VARIANT == 1) {
            v[q] = s[ic - yz];
        } else {
            v[q] = s[ic - yz];  // aligned run; shifted by one lane below
            if constexpr (VARIANT == 2) {
                if (cx == 1 && lane == 0) ed[q] = s[ic - yz - 1];
                if (cx == -1 && lane == 63) ed[q] = s[ic - yz + 1];
            } else {
                if (cx == 1 && lane == 63) edge[0][q][wave] = v[q];
                if (cx == -1 && lane == 0) edge[1][q][wave] = v[q];
                if (cx == 1 && lane == 0 && wave == 0) ed[q] = s[ic - yz - 1];
                if (cx == -1 && lane == 63 && wave == 3) ed[q] = s[ic - yz + 1];
            }
        }
    }
    if constexpr (VARIANT >= 2) {
        if constexpr (VARIANT == 3) __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int cx = ProbeDirs<Q>::cx[q];
            if (cx == 0) continue;
            T old = ed[q];
            if constexpr (VARIANT == 3) {
                if (cx == 1 && lane == 0 && wave > 0) old = edge[0][q][wave - 1];
                if (cx == -1 && lane == 63 && wave < 3) old = edge[1][q][wave + 1];
            }
            v[q] = dpp_shift<T>(old, v[q], cx == 1);
        }
    }
#pragma unroll
    for (int q = 0; q < Q; q++) __builtin_nontemporal_store(v[q], dst + (size_t)q * FS + ic);
}

// direction q of cell c in population blocks of 2^pop_lb cells (the engine's layout for the
// widths that pop_arrays leaves to blocks)
template <int Q, typename T>
__device__ inline size_t blk_at(int q, uint32_t c)
{
    constexpr uint32_t lb = pop_lb(sizeof(T));
    return ((((size_t)(c >> lb)) * (size_t)Q + (size_t)q) << lb) + (c & ((1u << lb) - 1u));
}

// the same skeleton on population blocks (blk_at): cell c of the lattice is
// cell c + guard of the arrays, guard a whole number of blocks >= the largest pull offset
template <typename T, int Q>
__global__ void __launch_bounds__(256) probe_blocks_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                           uint32_t guard, int sy, int sz)
{
    const uint32_t ic = blockIdx.x * 256u + threadIdx.x + guard;
    T v[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int off = ProbeDirs<Q>::cx[q] + ProbeDirs<Q>::cy[q] * sy + ProbeDirs<Q>::cz[q] * sz;
        v[q] = src[blk_at<Q, T>(q, (uint32_t)((int)ic - off))];
    }
#pragma unroll
    for (int q = 0; q < Q; q++) __builtin_nontemporal_store(v[q], dst + blk_at<Q, T>(q, ic));
}

// the same skeleton on population blocks of 2^lb cells, lb chosen at run time (the sweep
// behind the engine's block sizes)
template <typename T, int Q>
__global__ void __launch_bounds__(256) probe_blocks_rt_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                              uint32_t guard, int sy, int sz, uint32_t lb)
{
    const uint32_t ic = blockIdx.x * 256u + threadIdx.x + guard;
    const uint32_t m = (1u << lb) - 1u;
    T v[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const uint32_t c = (uint32_t)((int)ic - (ProbeDirs<Q>::cx[q] + ProbeDirs<Q>::cy[q] * sy + ProbeDirs<Q>::cz[q] * sz));
        v[q] = src[((((size_t)(c >> lb)) * Q + q) << lb) + (c & m)];
    }
#pragma unroll
    for (int q = 0; q < Q; q++)
        __builtin_nontemporal_store(v[q], dst + ((((size_t)(ic >> lb)) * Q + q) << lb) + (ic & m));
}

template <typename T, int Q>
void launch_lattice(int variant, unsigned grid, const T *a, T *b, size_t FS, int sy, int sz)
{
    switch (variant) {
    case 4: hipLaunchKernelGGL((probe_lattice_kernel<T, Q, 0>), dim3(grid), dim3(256), 0, 0, a, b, FS, sy, sz); break;
    case 1: hipLaunchKernelGGL((probe_lattice_kernel<T, Q, 1>), dim3(grid), dim3(256), 0, 0, a, b, FS, sy, sz); break;
    case 2: hipLaunchKernelGGL((probe_lattice_kernel<T, Q, 2>), dim3(grid), dim3(256), 0, 0, a, b, FS, sy, sz); break;
    default: hipLaunchKernelGGL((probe_lattice_kernel<T, Q, 3>), dim3(grid), dim3(256), 0, 0, a, b, FS, sy, sz); break;
    }
}

void launch(int mode, const float4 *a, float4 *b, size_t n4)
{
    const unsigned g4 = (unsigned)((n4 + 1023) / 1024);
    switch (mode) {
    case 0: hipLaunchKernelGGL(lbffs_copy_kernel, dim3(256u * 32u), dim3(256), 0, 0, a, b, n4); break;
    case 1: hipLaunchKernelGGL(lbffs_copy4_kernel, dim3(g4), dim3(256), 0, 0, a, b, n4); break;
    case 2: hipLaunchKernelGGL(lbffs_read_kernel, dim3(g4), dim3(256), 0, 0, a, (float *)b, n4); break;
    default: hipLaunchKernelGGL(lbffs_fill_kernel, dim3(g4), dim3(256), 0, 0, b, n4); break;
    }
}

}  // namespace

extern "C" int lbffs_probe_bandwidth(int mode, size_t bytes, int reps, float *ms_per_pass)
{
    if (mode < 0 || mode > 3 || bytes < 16 || bytes % 16 != 0 || bytes / 16 / 1024 >= (1ull << 31) || reps < 1 ||
        !ms_per_pass)
        return LBFFS_E_INVALID;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LBFFS_OK;
    const size_t n4 = bytes / 16;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        rc = LBFFS_E_NOMEM;
    } else if (hipMemset(a, 0, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
               hipEventCreate(&e1) != hipSuccess) {
        rc = LBFFS_E_DEVICE;
    } else {
        launch(mode, (const float4 *)a, (float4 *)b, n4);  // untimed pass
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < reps; r++) launch(mode, (const float4 *)a, (float4 *)b, n4);
        (void)hipEventRecord(e1, 0);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = LBFFS_E_DEVICE;
        } else {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *ms_per_pass = ms / reps;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

extern "C" int lbffs_probe_copy(size_t bytes, int reps, float *ms_per_copy)
{
    return lbffs_probe_bandwidth(0, bytes, reps, ms_per_copy);
}

extern "C" int lbffs_probe_stream(int kind, int unroll, int nt, int block, int grid, size_t bytes, int reps,
                                  float *ms_per_pass)
{
    if (kind < 0 || kind > 2 || block < 64 || block > 1024 || block % 64 != 0 || grid < 0 || bytes < 16 ||
        bytes % 16 != 0 || reps < 1 || !ms_per_pass)
        return LBFFS_E_INVALID;
    if (unroll != 1 && unroll != 2 && unroll != 4 && unroll != 8 && unroll != 16) return LBFFS_E_INVALID;
    const size_t n4 = bytes / 16;
    const size_t ntiles = (n4 + (size_t)unroll * block - 1) / ((size_t)unroll * block);
    if (ntiles >= (1ull << 31)) return LBFFS_E_INVALID;
    const unsigned g = grid > 0 ? (unsigned)grid : (unsigned)ntiles;
    void *a = nullptr, *b = nullptr, *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LBFFS_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
        hipMalloc(&sink, (size_t)g * sizeof(float)) != hipSuccess) {
        rc = LBFFS_E_NOMEM;
    } else if (hipMemset(a, 0, bytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
               hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = LBFFS_E_DEVICE;
    } else {
        auto go = [&]() {
            const pf4 *s4 = (const pf4 *)a;
            pf4 *d4 = (pf4 *)b;
            if (kind == 0) launch_stream_k<0>(unroll, nt != 0, g, block, s4, d4, n4, ntiles, (float *)sink);
            else if (kind == 1) launch_stream_k<1>(unroll, nt != 0, g, block, s4, d4, n4, ntiles, (float *)sink);
            else launch_stream_k<2>(unroll, nt != 0, g, block, s4, d4, n4, ntiles, (float *)sink);
        };
        go();  // untimed pass
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < reps; r++) go();
        (void)hipEventRecord(e1, 0);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = LBFFS_E_DEVICE;
        } else {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *ms_per_pass = ms / reps;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void *p : {a, b, sink})
        if (p) (void)hipFree(p);
    return rc;
}

extern "C" int lbffs_probe_lattice(int q, int store_bytes, int variant, int nx, int ny, int nz, int reps,
                                   float *ms_per_pass)
{
    const bool rt = variant >= 106 && variant <= 118;  // blocks of 2^(variant - 100) cells
    if ((q != 19 && q != 27) || (store_bytes != 4 && store_bytes != 2) || variant < 0 || (variant > 4 && !rt) || nx < 2 ||
        ny < 2 || nz < 2 || reps < 1 || !ms_per_pass)
        return LBFFS_E_INVALID;
    const size_t N = (size_t)nx * ny * nz;
    if (N % 256 != 0 || N >= (1ull << 31)) return LBFFS_E_INVALID;
    const int sy = nx, sz = nx * ny;
    // guard >= the largest pull offset (1 + nx + nx*ny), rounded to whole 256-byte lines
    const size_t guard = ((size_t)(1 + sy + sz) * store_bytes + 255) / 256 * 256 / store_bytes;
    // per-direction stride skewed as the engine's arrays (8 KiB past a multiple of 256 KiB)
    const size_t span = size_t(1) << 18, skew = 8192;
    const size_t raw = (N + 2 * guard) * store_bytes;
    const size_t FS = (((raw - skew + span - 1) / span) * span + skew) / store_bytes;
    // variant 0, the engine's layout: population blocks for 4-byte values, direction arrays
    // (variant 4) for 2-byte ones; blocks get a guard of whole blocks on either side
    const int v = (variant == 0 && pop_arrays((size_t)store_bytes)) ? 4 : variant;
    const uint32_t lb = rt ? (uint32_t)(variant - 100) : pop_lb((size_t)store_bytes);
    const size_t blk = (size_t)1 << lb;
    const size_t bguard = (1 + (size_t)sy + (size_t)sz + blk - 1) / blk * blk;
    if (N + 2 * bguard >= (1ull << 31)) return LBFFS_E_INVALID;
    const size_t bbytes = (N + 2 * bguard + blk - 1) / blk * blk * (size_t)q * store_bytes;
    const size_t bytes = std::max(((size_t)q * FS + 2 * guard) * store_bytes, bbytes);
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LBFFS_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        rc = LBFFS_E_NOMEM;
    } else if (hipMemset(a, 0, bytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
               hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = LBFFS_E_DEVICE;
    } else {
        const unsigned grid = (unsigned)(N / 256);
        auto go = [&]() {
            if (rt) {
                const uint32_t g = (uint32_t)bguard;
                if (store_bytes == 4 && q == 19)
                    hipLaunchKernelGGL((probe_blocks_rt_kernel<float, 19>), dim3(grid), dim3(256), 0, 0,
                                       (const float *)a, (float *)b, g, sy, sz, lb);
                else if (store_bytes == 4)
                    hipLaunchKernelGGL((probe_blocks_rt_kernel<float, 27>), dim3(grid), dim3(256), 0, 0,
                                       (const float *)a, (float *)b, g, sy, sz, lb);
                else if (q == 19)
                    hipLaunchKernelGGL((probe_blocks_rt_kernel<uint16_t, 19>), dim3(grid), dim3(256), 0, 0,
                                       (const uint16_t *)a, (uint16_t *)b, g, sy, sz, lb);
                else
                    hipLaunchKernelGGL((probe_blocks_rt_kernel<uint16_t, 27>), dim3(grid), dim3(256), 0, 0,
                                       (const uint16_t *)a, (uint16_t *)b, g, sy, sz, lb);
            } else if (v == 0) {
                const uint32_t g = (uint32_t)bguard;
                if (store_bytes == 4 && q == 19)
                    hipLaunchKernelGGL((probe_blocks_kernel<float, 19>), dim3(grid), dim3(256), 0, 0, (const float *)a,
                                       (float *)b, g, sy, sz);
                else if (store_bytes == 4)
                    hipLaunchKernelGGL((probe_blocks_kernel<float, 27>), dim3(grid), dim3(256), 0, 0, (const float *)a,
                                       (float *)b, g, sy, sz);
                else if (q == 19)
                    hipLaunchKernelGGL((probe_blocks_kernel<uint16_t, 19>), dim3(grid), dim3(256), 0, 0,
                                       (const uint16_t *)a, (uint16_t *)b, g, sy, sz);
                else
                    hipLaunchKernelGGL((probe_blocks_kernel<uint16_t, 27>), dim3(grid), dim3(256), 0, 0,
                                       (const uint16_t *)a, (uint16_t *)b, g, sy, sz);
            } else if (store_bytes == 4) {
                const float *s = (const float *)a + guard;
                float *d = (float *)b + guard;
                if (q == 19) launch_lattice<float, 19>(v, grid, s, d, FS, sy, sz);
                else launch_lattice<float, 27>(v, grid, s, d, FS, sy, sz);
            } else {
                const uint16_t *s = (const uint16_t *)a + guard;
                uint16_t *d = (uint16_t *)b + guard;
                if (q == 19) launch_lattice<uint16_t, 19>(v, grid, s, d, FS, sy, sz);
                else launch_lattice<uint16_t, 27>(v, grid, s, d, FS, sy, sz);
            }
        };
        go();
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < reps; r++) go();
        (void)hipEventRecord(e1, 0);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = LBFFS_E_DEVICE;
        } else {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *ms_per_pass = ms / reps;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void *p : {a, b})
        if (p) (void)hipFree(p);
    return rc;
}
