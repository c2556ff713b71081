// launch.hpp -- host-side launchers of the step kernels.
//
// The step and batch kernels are instantiated for every lattice, storage form, sub-grid
// model and variant; kernels.hip defines these launchers and is compiled once per
// (lattice, storage form, single/batch) so that the instantiations build in parallel.
// The API translation unit sees only the declarations below.
#pragma once

#include "batch.hpp"

namespace lbffs {

// aa: 0 two-lattice streaming, 1 or 2 the first or second in-place half-step (D3Q19).
// With TUNE_SPLIT a two-lattice step runs in two passes: the cells off the lattice faces,
// then the face cells of the list (a.fcells); otherwise, and for in-place half-steps, one.
template <typename S, typename Real, int Q, int SGS>
void launch_step(const StepArgs<Real, Q> &a, unsigned grid, int block, hipStream_t st, int aa);

// One batched step: the block table (cells off the lattice faces), then the face table.
// grid = 0: one work-group per table entry; otherwise a grid-stride walk with that many.
// coll: the lattices' collision (COLL_RR or COLL_CUMULANT; D3Q19 runs RR whatever it says).
template <typename S, typename Real, int Q, int SGS, bool IBB>
void launch_batch(const StepArgs<Real, Q> *args, const BatchBlock *table, uint32_t nblocks, const BatchBlock *ftable,
                  uint32_t nfblocks, const uint32_t *fcells, unsigned grid, int coll, hipStream_t st);

}  // namespace lbffs
