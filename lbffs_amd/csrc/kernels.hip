// kernels.hip -- definitions and explicit instantiations of the step launchers of
// launch.hpp for one lattice (KERN_Q), one storage form (KERN_STORE: 0 FP32, 1 FP16
// deviations, 2 BF16 deviations, 3 FP64) and either the single-lattice step
// (KERN_BATCH = 0) or the batched step (KERN_BATCH = 1).  The build compiles this file
// once per combination.
#include "launch.hpp"

#if !defined(KERN_Q) || !defined(KERN_STORE) || !defined(KERN_BATCH)
#error "kernels.hip is compiled with -DKERN_Q=19|27 -DKERN_STORE=0..3 -DKERN_BATCH=0|1"
#endif

namespace lbffs {

template <typename S, typename Real, int SGS, int COLL>
void launch_step27(const StepArgs<Real, 27> &a, unsigned grid, unsigned fgrid, unsigned mfgrid, int block,
                   hipStream_t st, bool merged, bool split)
{
    if (merged) {
        hipLaunchKernelGGL((step27_kernel<S, Real, SGS, 3, COLL>), dim3(mfgrid + grid), dim3(block), 0, st, a);
    } else if (split) {
        hipLaunchKernelGGL((step27_kernel<S, Real, SGS, 1, COLL>), dim3(grid), dim3(block), 0, st, a);
        if (fgrid) hipLaunchKernelGGL((step27_kernel<S, Real, SGS, 2, COLL>), dim3(fgrid), dim3(block), 0, st, a);
    } else {
        hipLaunchKernelGGL((step27_kernel<S, Real, SGS, 0, COLL>), dim3(grid), dim3(block), 0, st, a);
    }
}

template <typename S, typename Real, int Q, int SGS>
void launch_step(const StepArgs<Real, Q> &a, unsigned grid, int block, hipStream_t st, int aa)
{
    const bool merged = (a.tune & TUNE_FACE_FIRST) && aa == 0;
    const bool split = (a.tune & TUNE_SPLIT) && aa == 0 && !merged;
    const unsigned fgrid = (a.nface + (unsigned)block - 1) / (unsigned)block;
    // the one-launch form leads with the face blocks, padded to a multiple of 8
    const unsigned mfgrid = merged_face_blocks(a.nface, (unsigned)block);
    if constexpr (Q == 27) {
        if (a.coll == COLL_RR)
            launch_step27<S, Real, SGS, COLL_RR>(a, grid, fgrid, mfgrid, block, st, merged, split);
        else
            launch_step27<S, Real, SGS, COLL_CUMULANT>(a, grid, fgrid, mfgrid, block, st, merged, split);
    } else {
        if (aa == 1)
            hipLaunchKernelGGL((step19_kernel<S, Real, SGS, false, 1>), dim3(grid), dim3(block), 0, st, a);
        else if (aa == 2)
            hipLaunchKernelGGL((step19_kernel<S, Real, SGS, false, 2>), dim3(grid), dim3(block), 0, st, a);
        else if (a.ibb && merged)
            hipLaunchKernelGGL((step19_kernel<S, Real, SGS, true, 0, 3>), dim3(mfgrid + grid), dim3(block), 0, st, a);
        else if (a.ibb && split) {
            hipLaunchKernelGGL((step19_kernel<S, Real, SGS, true, 0, 1>), dim3(grid), dim3(block), 0, st, a);
            if (fgrid)
                hipLaunchKernelGGL((step19_kernel<S, Real, SGS, true, 0, 2>), dim3(fgrid), dim3(block), 0, st, a);
        } else if (a.ibb) {
            hipLaunchKernelGGL((step19_kernel<S, Real, SGS, true, 0, 0>), dim3(grid), dim3(block), 0, st, a);
        } else if (merged) {
            hipLaunchKernelGGL((step19_kernel<S, Real, SGS, false, 0, 3>), dim3(mfgrid + grid), dim3(block), 0, st, a);
        } else if (split) {
            hipLaunchKernelGGL((step19_kernel<S, Real, SGS, false, 0, 1>), dim3(grid), dim3(block), 0, st, a);
            if (fgrid)
                hipLaunchKernelGGL((step19_kernel<S, Real, SGS, false, 0, 2>), dim3(fgrid), dim3(block), 0, st, a);
        } else {
            hipLaunchKernelGGL((step19_kernel<S, Real, SGS, false, 0, 0>), dim3(grid), dim3(block), 0, st, a);
        }
    }
}

template <typename S, typename Real, int Q, int SGS, bool IBB>
void launch_batch(const StepArgs<Real, Q> *args, const BatchBlock *table, uint32_t nblocks, const BatchBlock *ftable,
                  uint32_t nfblocks, const uint32_t *fcells, unsigned grid, int coll, hipStream_t st)
{
    const unsigned g = batch_grid(grid, nblocks);
    const unsigned fg = batch_grid(grid, nfblocks);
    if constexpr (Q == 27) {
        if (coll == COLL_RR) {
            hipLaunchKernelGGL((batch_step_kernel<S, Real, SGS, Q, IBB, 1, COLL_RR>), dim3(g), dim3(kBatchBlock), 0,
                               st, args, table, nblocks, (const uint32_t *)nullptr);
            if (nfblocks)
                hipLaunchKernelGGL((batch_step_kernel<S, Real, SGS, Q, IBB, 2, COLL_RR>), dim3(fg), dim3(kBatchBlock),
                                   0, st, args, ftable, nfblocks, fcells);
            return;
        }
    }
    hipLaunchKernelGGL((batch_step_kernel<S, Real, SGS, Q, IBB, 1>), dim3(g), dim3(kBatchBlock), 0, st, args, table,
                       nblocks, (const uint32_t *)nullptr);
    if (nfblocks)
        hipLaunchKernelGGL((batch_step_kernel<S, Real, SGS, Q, IBB, 2>), dim3(fg), dim3(kBatchBlock), 0, st, args,
                           ftable, nfblocks, fcells);
}

#if KERN_STORE == 0
using KS = float;
using KR = float;
#elif KERN_STORE == 1
using KS = __half;
using KR = float;
#elif KERN_STORE == 2
using KS = bf16_t;
using KR = float;
#else
using KS = double;
using KR = double;
#endif

#if KERN_BATCH == 0
template void launch_step<KS, KR, KERN_Q, 0>(const StepArgs<KR, KERN_Q> &, unsigned, int, hipStream_t, int);
template void launch_step<KS, KR, KERN_Q, 1>(const StepArgs<KR, KERN_Q> &, unsigned, int, hipStream_t, int);
template void launch_step<KS, KR, KERN_Q, 2>(const StepArgs<KR, KERN_Q> &, unsigned, int, hipStream_t, int);
#else
#define LBFFS_BATCH_INST(SGS, IBB)                                                                                  \
    template void launch_batch<KS, KR, KERN_Q, SGS, IBB>(const StepArgs<KR, KERN_Q> *, const BatchBlock *, uint32_t, \
                                                         const BatchBlock *, uint32_t, const uint32_t *, unsigned,   \
                                                         int, hipStream_t);
LBFFS_BATCH_INST(0, false)
LBFFS_BATCH_INST(1, false)
LBFFS_BATCH_INST(2, false)
#if KERN_Q == 19
LBFFS_BATCH_INST(0, true)
LBFFS_BATCH_INST(1, true)
LBFFS_BATCH_INST(2, true)
#endif
#undef LBFFS_BATCH_INST
#endif

}  // namespace lbffs
