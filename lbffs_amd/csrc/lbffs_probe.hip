// lbffs_probe.hip -- HBM bandwidth probes: plain streaming kernels of known size.
// Their durations give the practical HBM ceilings next to the 8 TB/s peak, and the
// copy kernel's FETCH_SIZE / WRITE_SIZE counters calibrate the counter-to-bytes
// factor used to turn the step kernel's counters into HBM traffic per launch.
#include <hip/hip_runtime.h>

#include "../../include/lbffs.h"

namespace {

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
