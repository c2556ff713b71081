// lbffs_probe.hip -- HBM copy probe: a plain streaming copy of known size.  Its
// duration gives the practical HBM ceiling next to the 8 TB/s peak, and its
// FETCH_SIZE / WRITE_SIZE counters calibrate the counter-to-bytes factor used to turn
// the step kernel's counters into HBM traffic per launch.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/lbffs.h"

namespace {

__global__ void __launch_bounds__(256) lbffs_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                         size_t n4)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" int lbffs_probe_copy(size_t bytes, int reps, float *ms_per_copy)
{
    if (bytes < 16 || bytes % 16 != 0 || reps < 1 || !ms_per_copy) return LBFFS_E_INVALID;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LBFFS_OK;
    const size_t n4 = bytes / 16;
    const unsigned grid = 256u * 32u;  // 32 work-groups per CU on 256 CUs, grid-stride beyond
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        rc = LBFFS_E_NOMEM;
    } else if (hipMemset(a, 0, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
               hipEventCreate(&e1) != hipSuccess) {
        rc = LBFFS_E_DEVICE;
    } else {
        // one untimed pass, then reps timed passes
        hipLaunchKernelGGL(lbffs_copy_kernel, dim3(grid), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n4);
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(lbffs_copy_kernel, dim3(grid), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n4);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = LBFFS_E_DEVICE;
        } else {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *ms_per_copy = ms / reps;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
