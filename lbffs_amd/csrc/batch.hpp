// batch.hpp -- one launch advances many independent lattices of uneven size.
//
// The lattices of a batch share the lattice type, storage form and sub-grid model, so one
// kernel instantiation serves them all.  A packed table lists every 256-cell block of every
// lattice as (lattice, first cell); each XCD walks its own slab of the table (persistent
// work-groups when the grid is smaller than the table) and the kernel runs the same
// per-cell update as the single-lattice kernels, with that lattice's launch arguments read
// from a device array (uniform per work-group, so they come in through scalar loads).
#pragma once

#include "d3q27_step.hpp"

namespace lbffs {

struct BatchBlock {
    uint32_t lattice;  // index into the argument array
    uint32_t first;    // first cell of this 256-cell block in its lattice, or (face table)
                       // first entry of the block in the face-cell list
    uint32_t count;    // face table: cells of the block (the others: unused)
};

constexpr int kBatchBlock = 256;

// The launch arguments of every lattice are read from device memory where they are used
// (uniform per work-group, so through scalar loads); the host prepares a separate argument
// array for the last step of a call, whose cells store every moment and diagnostic.
// IBB: D3Q19's interpolated bounce-back compiled in (chosen when any lattice of the batch
// asks for it); the lattices without it keep simple bounce-back through their own ibb flag.
// PASS 1 runs the cells of the block table that are not on a lattice face; PASS 2 runs the
// face cells of the face table (cell indices from `cells`).  On small lattices nearly every
// wave holds a face cell (i = 0 or nx - 1), and a wave with both kinds runs the face path
// and the interior pull: split in two launches, each wave runs one of them.  Both passes
// read only the previous step's arrays, so their order does not matter.
// Table entries of this work-group.  The table is cut into 8 slabs of consecutive entries,
// one per XCD (the hardware deals work-group w to XCD w mod 8), and the work-groups of an
// XCD walk its slab in order.  Consecutive entries are mostly blocks of one lattice, whose
// pulls and velocity stencil reach the neighbour rows and planes: in one slab those reads
// find their cache lines in the L2 that fetched them; dealt round-robin, a plane's
// neighbours run on other XCDs and every XCD fetches the same lines from HBM.
// The grid is either one work-group per entry or a multiple of 8 below the entry count.
struct TableWalk {
    uint32_t first, end, stride;
};
__device__ inline TableWalk table_walk(uint32_t nblocks)
{
    const uint32_t G = gridDim.x;
    if (G >= nblocks) {
        const uint32_t e = xcd_block(blockIdx.x, G);
        return {e, e + 1, 1};
    }
    const uint32_t x = blockIdx.x & 7u, k = blockIdx.x >> 3;
    const uint32_t base = nblocks >> 3, r = nblocks & 7u;
    const uint32_t lo = x * base + (x < r ? x : r);
    return {lo + k, lo + base + (x < r ? 1u : 0u), G >> 3};
}

// work-groups of a batch launch over n table entries for a requested grid (0: one per entry)
__host__ inline unsigned batch_grid(unsigned grid, uint32_t n)
{
    if (grid == 0 || grid >= n) return n;
    const unsigned g = (grid + 7u) & ~7u;
    return g >= n ? n : g;
}

// Waves per SIMD the D3Q19 block-table pass is held to: eight with WALE (64 VGPRs, no spill;
// the ensemble's step 6.63 -> 6.22 ms), seven without a sub-grid model, where eight loads
// fewer populations in flight and ran slower (5.30 -> 5.92 ms).
template <int SGS, bool IBB, int PASS, typename Real>
constexpr int kBatch19MinWaves = (IBB || PASS != 1 || sizeof(Real) == 8) ? 1 : (SGS == 0 ? 8 : 7);

template <typename S, typename Real, int SGS, int Q, bool IBB, int PASS, int COLL = COLL_CUMULANT>
__global__ void __launch_bounds__(kBatchBlock)
__attribute__((amdgpu_waves_per_eu(Q == 27 ? kBatch27MinWaves<Real> : kBatch19MinWaves<SGS, IBB, PASS, Real>, 8)))
    batch_step_kernel(const StepArgs<Real, Q> *__restrict__ args, const BatchBlock *__restrict__ table,
                      uint32_t nblocks, const uint32_t *__restrict__ cells)
{
    const TableWalk tw = table_walk(nblocks);
    for (uint32_t b = tw.first; b < tw.end; b += tw.stride) {
        const BatchBlock e = table[b];
        const StepArgs<Real, Q> &a = args[e.lattice];
        uint32_t icu;
        if constexpr (PASS == 2) {
            if (threadIdx.x >= e.count) continue;
            icu = cells[e.first + threadIdx.x];
            if (face_pass_skips(cell_wave_class(a.wclass, icu))) continue;
        } else {
            icu = e.first + threadIdx.x;
            if (icu >= a.N) continue;
        }
        do {
#define CELL_RETURN continue
#define CELL_IBB (IBB && a.ibb != 0)
#define CELL_WALE_INTERIOR true
#define CELL_AA 0
#define CELL_SKIP_FACE (PASS == 1)
#define CELL_PURE_OK (PASS == 1)
#define CELL_ALIGNED (PASS == 1)
#define CELL_COLL COLL
            if constexpr (Q == 27) {
#include "d3q27_cell.inc"
            } else {
#include "d3q19_cell.inc"
            }
#undef CELL_COLL
#undef CELL_ALIGNED
#undef CELL_PURE_OK
#undef CELL_SKIP_FACE
#undef CELL_AA
#undef CELL_WALE_INTERIOR
#undef CELL_IBB
#undef CELL_RETURN
        } while (0);
    }
}

}  // namespace lbffs
