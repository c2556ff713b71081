/*
 * lbffs.h -- C ABI of the MI355X lattice-Boltzmann time-step engine.
 *
 * The engine replaces the per-step device work of LBFFS: the OpenCL program build,
 * the device buffers and the k_streamingCollision launch of the reference host
 * (D3Q19_opencl/ns.cpp:382-603, D3Q27_opencl/ns.cpp same structure).  The host keeps
 * its own arrays in the reference layout (populations q-major, f[q*N + ic],
 * ic = i + nx*(j + ny*k)) and hands them across this boundary as plain pointers.
 *
 * Conventions
 *  - every entry point returns int: 0 on success, a negative LBFFS_E* code on error
 *    (as cl_int does for the OpenCL calls it replaces); lbffs_strerror() names it and
 *    lbffs_last_error() gives the detail text of the last failure on this thread;
 *  - "host real" means float, or double when the context stores LBFFS_STORE_F64;
 *  - calls on one context are not thread-safe; lbffs_download/lbffs_reduce block
 *    until the device work before them has finished (like cl::copy after
 *    queue.finish(), ns.cpp:537 / write.hpp:3-15); lbffs_step returns as soon as the
 *    work is queued.
 */
#ifndef LBFFS_H
#define LBFFS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBFFS_ABI_VERSION 1

/* ---- error codes ---- */
enum {
    LBFFS_OK = 0,
    LBFFS_E_INVALID = -1,     /* bad argument or size */
    LBFFS_E_NOMEM = -2,       /* device allocation failed */
    LBFFS_E_DEVICE = -3,      /* HIP runtime error */
    LBFFS_E_UNSUPPORTED = -4, /* configuration this build does not implement */
    LBFFS_E_STATE = -5        /* call out of order (e.g. step before upload of f) */
};

/* ---- storage forms of the populations in HBM ---- */
enum {
    LBFFS_STORE_F32 = 0,     /* reference form: float f_q */
    LBFFS_STORE_F16DEV = 1,  /* IEEE half of (f_q - w_q) * 2^e_q, arithmetic in FP32 */
    LBFFS_STORE_BF16DEV = 2, /* bfloat16 of (f_q - w_q) * 2^e_q, arithmetic in FP32 */
    LBFFS_STORE_F64 = 3      /* double storage and double arithmetic (accuracy reference) */
};

/* ---- sub-grid-scale models ---- */
enum {
    LBFFS_SGS_WALE = 0,        /* waleTauSGS, the model the reference kernel runs (K19:3631) */
    LBFFS_SGS_SMAGORINSKY = 1, /* smagorinskyTauSGS with wall damping (K19:2133-2201; see damping) */
    LBFFS_SGS_NONE = 2         /* tauSGS = 0 */
};

/* ---- collision operators ---- */
enum {
    LBFFS_COLL_RR = 0,      /* recursive regularized, K19:3156-3309 (D3Q19) */
    LBFFS_COLL_CUMULANT = 1 /* parametrized cumulant, D3Q27_opencl/kernel.cl:2352-2691 (D3Q27) */
};

/* ---- fields exchanged with the host (names of the reference buffers, ns.cpp:405-436) ---- */
enum {
    LBFFS_FIELD_F = 0,         /* Q*N host reals, q-major */
    LBFFS_FIELD_BOUNDARY = 1,  /* N int32: 1 on domain faces (ns.cpp:164-177) */
    LBFFS_FIELD_BOUNDARY1 = 2, /* N int32: x-face BC code (input.hpp:221-256) */
    LBFFS_FIELD_BOUNDARY2 = 3, /* N int32: y-face BC code */
    LBFFS_FIELD_BOUNDARY3 = 4, /* N int32: z-face BC code */
    LBFFS_FIELD_SDF = 5,       /* N host reals */
    LBFFS_FIELD_SOLID = 6,     /* N uint8 */
    LBFFS_FIELD_NEISOLID = 7,  /* N uint8 */
    LBFFS_FIELD_U0 = 8,        /* N host reals, lattice units */
    LBFFS_FIELD_V0 = 9,
    LBFFS_FIELD_W0 = 10,
    LBFFS_FIELD_FWX = 11, /* N host reals: momentum-exchange force on solid links */
    LBFFS_FIELD_FWY = 12,
    LBFFS_FIELD_FWZ = 13,
    LBFFS_FIELD_TAUSGS = 14, /* N host reals */
    LBFFS_FIELD_RHO = 15,    /* N host reals: pre-collision moments of the last step */
    LBFFS_FIELD_U = 16,
    LBFFS_FIELD_V = 17,
    LBFFS_FIELD_W = 18,
    LBFFS_FIELD_OMEGA = 19, /* N host reals: 1/tau per cell */
    LBFFS_FIELD_GX = 20,    /* N host reals: IBM forcing density */
    LBFFS_FIELD_GY = 21,
    LBFFS_FIELD_GZ = 22,
    LBFFS_FIELD_QLINK = 23, /* Q*N host reals, q-major: q-distance of each solid link (fraction
                               of the link from the fluid cell to the wall), read by the
                               interpolated bounce-back when lbffs_step_params.ibb = 1 (the
                               layout of the reference's unused qf array, ns.cpp:215-217);
                               device memory is allocated on the first upload */
    LBFFS_FIELD_COUNT = 24
};

/* Fixed per-run description of a lattice.  Replaces the program build and buffer
 * creation of ns.cpp:385-436. */
typedef struct {
    int abi_version; /* LBFFS_ABI_VERSION */
    int q;           /* 19 (D3Q19_opencl) or 27 (D3Q27_opencl) */
    int nx, ny, nz;
    int storage;   /* LBFFS_STORE_* */
    int sgs;       /* LBFFS_SGS_* */
    int collision; /* LBFFS_COLL_* */
    float omega0;  /* 1/(3 nu + 1/2): cells whose omega equals it skip the omega read */
    int block;     /* work-group size; 0 = engine default */
    int rounding;  /* 16-bit storage forms: LBFFS_ROUND_STOCHASTIC (default) or LBFFS_ROUND_NEAREST */
    int streaming; /* LBFFS_STREAM_TWO_LATTICE (default) or LBFFS_STREAM_AA */
    int reserved[6];
} lbffs_desc;

/* Rounding of the 16-bit storage forms (no counterpart in the reference, which stores
 * FP32).  Round-to-nearest removes every perturbation below half a unit in the last place
 * at each step, which can hold a flow on a symmetric, non-physical branch (a cylinder wake
 * at Re 100 never sheds); stochastic rounding (gfx950 v_cvt_sr_*) is unbiased, with the
 * random bits hashed from the cell's moments: deterministic for a given state. */
enum { LBFFS_ROUND_STOCHASTIC = 0, LBFFS_ROUND_NEAREST = 1 };

/* Streaming.  TWO_LATTICE pulls from one population array into another and swaps them,
 * as ns.cpp:602 does.  AA streams in place on one array (half the population memory):
 * odd and even steps alternate two kernels that each read and write the same locations
 * per cell; results are bit-identical to TWO_LATTICE.  AA runs D3Q19 lattices whose faces
 * are Cyclic or BounceBack with the same code on both sides of an axis, without the
 * interpolated bounce-back (ibb) and moving walls (LBFFS_E_UNSUPPORTED / LBFFS_E_INVALID
 * otherwise).  Downloads, lbffs_output_macros and lbffs_reduce see the two-lattice layout
 * whatever the step count (after an odd step they first restore it in place). */
enum { LBFFS_STREAM_TWO_LATTICE = 0, LBFFS_STREAM_AA = 1 };

/* Scalars of one k_streamingCollision launch (D3Q19 ns.cpp:528-535, K19:3541-3545;
 * D3Q27 ns.cpp:555-577, K27:2725-2731). */
typedef struct {
    float dpdx;      /* body force along x, lattice units */
    float LES;       /* 1 or 0: multiplies the Smagorinsky tauSGS (WALE ignores it) */
    float omegaB;    /* bulk relaxation rate (RR collision; omega2 of the cumulant collision) */
    float spzWidth;  /* sponge width as a fraction of the domain (tauSpongeZone) */
    float Cs;        /* Smagorinsky constant */
    int damping;     /* Smagorinsky wall damping at the distance sdf: 0 off; 1 the Fluent-style
                        min(1, 0.41 y / Cs)^2 the reference applies (K19:2170-2177); 2 van
                        Driest, the branch K19:2178-2197 keeps in comments (y+ from the log
                        law by Newton's method, A+ = 26) */
    int store_every; /* 1: write rho/u/v/w/tauSGS/Fw/G for every cell on every step
                        (reference semantics throughout); 0: only on the last step of
                        the call and for the cells later steps read */
    float clim;      /* limiter of the D3Q27 cumulant collision (input.hpp default 0.01) */
    int ibb;         /* 0: the reference's solid links (D3Q19 half-way bounce-back, D3Q27
                        interpolated with q from the signed distances); 1: Filippova-Hanel /
                        Mei-Luo-Shyy interpolated bounce-back (the formula of K27:1846-1888) on
                        both lattices with q from LBFFS_FIELD_QLINK */
    int reserved[6];
} lbffs_step_params;

typedef struct lbffs_ctx lbffs_ctx;

/* Create a context on the current HIP device (the OpenCL context + buffers of
 * ns.cpp:385-436).  Every field is zero-initialised except f, which must be uploaded.
 * One population array (N cells in the storage form) and one macroscopic array (N FP32
 * values, FP64 for LBFFS_STORE_F64) must each stay below 4 GiB: about 1.07e9 cells for
 * FP32 and the 16-bit forms, 5.4e8 for FP64 (LBFFS_E_INVALID otherwise). */
int lbffs_create(const lbffs_desc *desc, lbffs_ctx **out);

/* Copy a host array into a device field (the cl::Buffer fills of ns.cpp:405-436).
 * bytes must equal the field's host size. */
int lbffs_upload(lbffs_ctx *ctx, int field, const void *host, size_t bytes);

/* Advance nsteps time steps (the loop body ns.cpp:514-537 + swap ns.cpp:602). */
int lbffs_step(lbffs_ctx *ctx, int nsteps, const lbffs_step_params *params);

/* Copy a device field to the host (cl::copy of write.hpp:3-15).  F is returned in
 * the host real type whatever the storage form. */
int lbffs_download(lbffs_ctx *ctx, int field, void *host, size_t bytes);

/* FP64 sums over fluid cells (the host loops of write.hpp:16-47, on the device):
 * out[0..2] = sum Fw{x,y,z}; out[3..5] = sum G{x,y,z}; out[6] = sum (rho - 1) with rho
 * summed from the current populations; out[7] = number of fluid cells. */
#define LBFFS_REDUCE_N 8
int lbffs_reduce(lbffs_ctx *ctx, double out[LBFFS_REDUCE_N]);

/* No counterpart in the reference: lbffs_reduce for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same sums, bit for bit, into the same out[8]; nothing is allocated or freed.  The
 * reduction's partial sums live in a fixed device array of the library (1024 x 8 + 8
 * doubles, 64 KiB, one per GPU, held until the process ends), whatever the lattice size.
 * Every call on one GPU shares that array: calls that overlap on one GPU race.  Returns
 * LBFFS_E_STATE while the lattice's flags are to be rebuilt (after an upload of a field
 * they derive from; lbffs_step, lbffs_reduce or lbffs_autotune rebuilds them, and that
 * rebuild allocates). */
int lbffs_reduce_noalloc(lbffs_ctx *ctx, double out[LBFFS_REDUCE_N]);

/* Density and velocity of every cell from the current populations, as the host computes
 * them for an output step (cal_rhoUVW, D3Q19_opencl/D3Q19.hpp:647-666 and
 * D3Q27_opencl/D3Q27.hpp:28-47, called at write.hpp:27): sums in direction order,
 * u = sum(f cx)/rho + 0.5 dpdx/rho (lattice units).  Four arrays of nx*ny*nz host reals,
 * `bytes` each; bit-identical to the reference host for FP32 storage. */
int lbffs_output_macros(lbffs_ctx *ctx, float dpdx, void *rho, void *u, void *v, void *w, size_t bytes);

/* Launch options of the step (no counterpart in the reference, whose work-group size
 * is left to the OpenCL runtime, ns.cpp:516). */
enum {
    LBFFS_TUNE_NT_STORE = 1, /* accepted, no effect: new populations are always stored non-temporally */
    LBFFS_TUNE_NO_XCD = 2,   /* keep the hardware work-group order across the 8 XCDs */
    LBFFS_TUNE_YTILE = 4,    /* visit the cells in tiles of 32, 16 or 8 rows (the first that
                                divides ny), plane by plane: the neighbour planes WALE reads
                                stay in L2; ignored where no such tile size exists */
    LBFFS_TUNE_SPLIT_FACES = 8, /* two launches per step: the cells off the lattice faces, then
                                   the face cells, so that no wave runs both paths (two-lattice
                                   streaming; ignored by in-place half-steps) */
    LBFFS_TUNE_NO_PURE = 16,    /* every wave reads its flags before it pulls; by default a wave
                                   of 64 interior fluid cells (a per-wave summary of the flags
                                   says so) issues its pulls together with the flag load */
    LBFFS_TUNE_FACE_FIRST = 32, /* one launch per step whose first work-groups run the face cells
                                   (from a list) and the others the cells off the faces, so that
                                   no wave runs both paths (two-lattice streaming; takes
                                   precedence over LBFFS_TUNE_SPLIT_FACES) */
    LBFFS_TUNE_XCD_CHUNK = 64   /* deal runs of 4 consecutive work-groups to each XCD in turn:
                                   the cache line a block shares with its x-neighbour block is
                                   fetched once, not once per XCD (takes precedence over
                                   LBFFS_TUNE_NO_XCD) */
};

/* Autotuner: time every candidate work-group size (64, 128, 256) and launch option on
 * this lattice for trial_steps steps each, time the three fastest again over 2 trial_steps
 * in two interleaved rounds, keep the best of those for later lbffs_step calls and report
 * it (best_ms from the second stage).  Populations and every field a step writes are
 * restored afterwards, so the simulation state is unchanged. */
int lbffs_autotune(lbffs_ctx *ctx, const lbffs_step_params *params, int trial_steps, int *best_block,
                   int *best_tune, float *best_ms_per_step);

/* Set the launch shape directly (work-group size multiple of 64 up to 256, LBFFS_TUNE_*). */
int lbffs_set_launch(lbffs_ctx *ctx, int block, int tune);

/* ---- immersed-boundary moving walls (k_Gwall / k_Gibm / k_Force, K19:3648-3807,
 * K27:2828-2986) ----
 * Markers are points in lattice coordinates (movingSTLcList, ns.cpp:248-254).  With more
 * than one marker every step is followed, as in ns.cpp:544-600, by: G_m = U_wall - the
 * trilinear interpolation of u over the marker's 8-cell box; G spread back over the box;
 * the marker moved by the translation and turned by rot_omega about the axis; and
 * f_q += rho w_q 3 (G . c_q) on the forced fluid cells.  The spread sums every marker's
 * contribution (the reference's atom_add_float result is written back over the sum,
 * K19:3760-3762).  GX/GY/GZ hold the spread G after a call, as the reference's GxIBM. */
typedef struct {
    float trans[3];      /* uMovingTrans, vMovingTrans, wMovingTrans per step (walls.hpp:606-611) */
    float rot_center[3]; /* rotX, rotY, rotZ in lattice units (walls.hpp:615-620) */
    float rot_axis[3];   /* rotAxisX, rotAxisY, rotAxisZ (walls.hpp:621-626) */
    float rot_omega;     /* rotOmega in radians per step (walls.hpp:613-614) */
    int reserved[6];
} lbffs_ibm_motion;

enum { LBFFS_IBM_POSITIONS = 0, LBFFS_IBM_FORCING = 1 };

/* Set n markers (3n host reals, xyz interleaved; n <= 1 switches the moving walls off, as
 * the reference host does when movingWalls.stl has one entry or none, ns.cpp:239-245, 544). */
int lbffs_ibm_set_markers(lbffs_ctx *ctx, const void *xyz, int n);
/* Motion of the markers (readMotions, walls.hpp:586-628, already in lattice units). */
int lbffs_ibm_set_motion(lbffs_ctx *ctx, const lbffs_ibm_motion *motion);
/* Copy out the marker positions (3n, xyz interleaved) or the marker forcing of the last
 * step (3n: G_x of all markers, then G_y, then G_z; GxMovingWall...). */
int lbffs_ibm_download(lbffs_ctx *ctx, int what, void *host, int n);

/* ---- batches: many independent lattices advanced by one launch per step ----
 * No counterpart in the reference, which runs one lattice per process (ns.cpp).  The
 * lattices share q, storage, sgs and collision; sizes, omega0 and step scalars are their
 * own.  A packed table of 256-cell blocks over all lattices drives one kernel per step;
 * the table is cut into 8 slabs of consecutive blocks, one per XCD, so that the blocks of a
 * lattice share one L2; a grid smaller than the table gives persistent work-groups that
 * walk their XCD's slab.  Each lattice is an ordinary context (lbffs_batch_context) for
 * uploads, downloads and reductions; the batch owns and destroys them, and they share the
 * batch's stream.  Moving-wall markers are not supported inside a batch. */
typedef struct lbffs_batch lbffs_batch;
int lbffs_batch_create(const lbffs_desc *descs, int n, lbffs_batch **out);
int lbffs_batch_context(lbffs_batch *batch, int i, lbffs_ctx **ctx);
/* grid: work-groups per launch (0 = one per 256-cell block; otherwise rounded up to a
 * multiple of 8, at most one per block); graph_chunk: steps captured
 * into one hipGraph and replayed (0 = plain launches; the last step of a call is always a
 * plain launch) */
int lbffs_batch_set_launch(lbffs_batch *batch, int grid, int graph_chunk);
/* nsteps steps of every lattice; params holds one entry per lattice */
int lbffs_batch_step(lbffs_batch *batch, int nsteps, const lbffs_step_params *params);
int lbffs_batch_last_step_ms(lbffs_batch *batch, float *ms);
/* Time the launch shapes (grid 0 and persistent grids of 2048, 4096 and 8192 work-groups
 * where the table has more blocks; graph_chunk 0 and 16) over trial_steps steps each, keep
 * the fastest as lbffs_batch_set_launch would, and put every lattice's populations, moments
 * and step counters back as they were.  Output pointers may be null.  Like the batch itself
 * it has no counterpart in the reference. */
int lbffs_batch_autotune(lbffs_batch *batch, const lbffs_step_params *params, int trial_steps, int *best_grid,
                         int *best_chunk, float *best_ms);
int lbffs_batch_destroy(lbffs_batch *batch);

/* Wait for queued device work. */
int lbffs_synchronize(lbffs_ctx *ctx);

/* Device time of the last lbffs_step call in milliseconds (HIP events around it). */
int lbffs_last_step_ms(lbffs_ctx *ctx, float *ms);

int lbffs_destroy(lbffs_ctx *ctx);

const char *lbffs_strerror(int code);
const char *lbffs_last_error(void);

/* Diagnostics, no counterpart in the reference: time plain streaming kernels over
 * `bytes` (multiple of 16) on the current device, averaged over `reps` passes.  They give
 * the practical HBM ceilings and the counter calibration for traffic measurements.
 * mode 0: grid-stride copy; 1: copy, four 16-byte loads per thread issued before the
 * stores; 2: read only; 3: write only.  lbffs_probe_copy is mode 0. */
int lbffs_probe_bandwidth(int mode, size_t bytes, int reps, float *ms_per_pass);
int lbffs_probe_copy(size_t bytes, int reps, float *ms_per_copy);

/* Diagnostics, no counterpart in the reference: a configurable streaming probe over
 * `bytes` (multiple of 16).  kind 0 copy, 1 read only, 2 write only; unroll = 1, 2, 4, 8
 * or 16 float4 per work-item, all loads issued before the stores; nt = 1: non-temporal
 * loads and stores; block = work-group size (multiple of 64, at most 1024); grid = work-
 * groups (0: one per tile of unroll*block float4, otherwise a grid-stride walk over tiles). */
int lbffs_probe_stream(int kind, int unroll, int nt, int block, int grid, size_t bytes, int reps, float *ms_per_pass);

/* Diagnostics, no counterpart in the reference: the memory skeleton of one lattice step --
 * every cell of an nx*ny*nz lattice (a multiple of 256 cells) pulls its q (19 or 27)
 * populations of store_bytes (4 or 2) with the step's offsets and stores them to its own
 * slots of a second array, with no collision.  variant 0: the engine's layout and the
 * lattice offsets (4-byte values in population blocks of 8192 cells, each holding its q
 * directions one after another; 2-byte values as variant 4).  The others lay the
 * populations out as q arrays of the whole lattice (direction-major): 4 the lattice
 * offsets (x-shifted runs
 * straddle cache lines); 1 the x component of every offset dropped (aligned runs: the ideal
 * pull); 2 x-shifted directions loaded aligned and moved one lane by a DPP wave shift, the
 * lane leaving the wave loading its value alone; 3 as 2, with that lane's value passed
 * between the four waves of a work-group through LDS.  100 + lb (lb = 6..18): population
 * blocks of 2^lb cells, whatever the width (the sweep behind the engine's block sizes). */
int lbffs_probe_lattice(int q, int store_bytes, int variant, int nx, int ny, int nz, int reps, float *ms_per_pass);

#ifdef __cplusplus
}
#endif
#endif
