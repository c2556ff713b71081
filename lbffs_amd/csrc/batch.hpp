// batch.hpp -- one launch advances many independent lattices of uneven size.
//
// The lattices of a batch share the lattice type, storage form and sub-grid model, so one
// kernel instantiation serves them all.  A packed table lists every 256-cell block of every
// lattice as (lattice, first cell); the kernel walks the table with a grid-stride loop
// (persistent work-groups when the grid is smaller than the table) and runs the same
// per-cell update as the single-lattice kernels, with that lattice's launch arguments read
// from a device array (uniform per work-group, so they come in through scalar loads).
#pragma once

#include "d3q27_step.hpp"

namespace lbffs {

struct BatchBlock {
    uint32_t lattice;  // index into the argument array
    uint32_t first;    // first cell of this 256-cell block in its lattice
};

constexpr int kBatchBlock = 256;

// LAST = 1 on the last step of a call: every cell stores its moments and diagnostics
template <typename S, typename Real, int SGS, int Q, int LAST>
__global__ void __launch_bounds__(kBatchBlock)
    batch_step_kernel(const StepArgs<Real, Q> *__restrict__ args, const BatchBlock *__restrict__ table,
                      uint32_t nblocks)
{
    for (uint32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const BatchBlock e = table[b];
        StepArgs<Real, Q> a = args[e.lattice];
        const uint32_t icu = e.first + threadIdx.x;
        if (icu >= a.N) continue;
        if (LAST) {
            a.aux_all = 1;
            a.mac_all = 1;
        }
        if constexpr (Q == 27) step27_cell<S, Real, SGS>(a, icu);
        else step19_cell<S, Real, SGS>(a, icu);
    }
}

}  // namespace lbffs
