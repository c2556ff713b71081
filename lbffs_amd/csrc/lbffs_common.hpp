// lbffs_common.hpp -- shared device definitions of the MI355X LBM engine:
// per-cell flag word, 32-bit division by run-time constants, and the storage codecs
// for populations held in HBM as float, double, or 16-bit deviations from the
// lattice weights.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace lbffs {

// ---- per-cell flag word (uint16), built once from the reference's side arrays ----
// bit 0      SOLID      solid cell: no work (k_streamingCollision skips it, kernel.cl:3549)
// bit 1      FACE       boundary[ic]==1: domain-face streaming + BCs (ns.cpp:164-177)
// bit 2      NEISOLID   fluid cell with a solid upwind neighbour (walls.hpp:187-216)
// bit 3      KEEPMAC    later steps read this cell's rho/u/v/w (faces and the layer inside)
// bit 4      HASSDF     sdf != 10000: Smagorinsky wall damping may apply
// bit 5      OMEGAVAR   omega differs from the run's omega0: read the omega field
// bits 6-8   x-face BC code, 9-11 y-face code, 12-14 z-face code (codes 0..7)
enum : uint32_t {
    F_SOLID = 1u << 0,
    F_FACE = 1u << 1,
    F_NEISOLID = 1u << 2,
    F_KEEPMAC = 1u << 3,
    F_HASSDF = 1u << 4,
    F_OMEGAVAR = 1u << 5,
    F_B1_SHIFT = 6,
    F_B2_SHIFT = 9,
    F_B3_SHIFT = 12,
};

__host__ __device__ inline int flag_code(uint32_t fl, int shift) { return (int)((fl >> shift) & 7u); }

// Per-wave summary of the flags: 4 bits per aligned run of 64 cells (bits 4 (w % 8) of word
// w / 8 for the run 64w .. 64w + 63), built with the flags:
//   WAVE_GENERAL  every cell reads its flags first and takes its own path;
//   WAVE_PURE     64 fluid cells off the lattice faces: fixed interior offsets;
//   WAVE_XMIX     64 fluid cells whose face cells lie on an x face only (their j and k
//                 interior): every cell pulls from the interior offsets (in bounds), then the
//                 face cells redo the directions that cross their face and apply its conditions;
//   WAVE_WRAP     64 fluid cells of one row, none on an x face, inside a y and/or z face plane
//                 whose code is Cyclic for all of them: interior offsets with a periodic wrap
//                 uniform over the wave;
//   WAVE_PLAIN    a PURE run whose 64 flag words are all zero (no solid link, no field to keep
//                 or read): the wave skips its flag load.
// Such waves issue their pulls together with the flag load instead of one memory round trip
// after it, and no lane runs the general face path.  Read through the scalar cache (the
// wave's first active cell names the run; the caller launches 64-cell aligned waves).
enum : uint32_t { WAVE_GENERAL = 0, WAVE_PURE = 1, WAVE_XMIX = 2, WAVE_WRAP = 3, WAVE_PLAIN = 4 };

__device__ inline uint32_t wave_class(const uint32_t *bits, uint32_t icu)
{
    if (!bits) return WAVE_GENERAL;
    const uint32_t w = __builtin_amdgcn_readfirstlane(icu) >> 6;
    typedef const __attribute__((address_space(4))) uint32_t const_u32;
    const uint32_t word = *((const_u32 *)bits + (w >> 3));
    return (word >> (4u * (w & 7u))) & 15u;
}

// the class of the run holding cell ic, per lane (a vector load)
__device__ inline uint32_t cell_wave_class(const uint32_t *bits, uint32_t ic)
{
    if (!bits) return WAVE_GENERAL;
    return (bits[ic >> 9] >> (4u * ((ic >> 6) & 7u))) & 15u;
}

// BC codes (input.hpp:221-256)
enum { BC_CYCLIC = 0, BC_BOUNCEBACK = 1, BC_OUTLET = 3, BC_SYMMETRY = 4, BC_FIXEDVELOCITY = 5,
       BC_FIXEDDENSITY = 6, BC_WALLFUNCTION = 7 };

// collision operators (the values of LBFFS_COLL_* in lbffs.h)
enum { COLL_RR = 0, COLL_CUMULANT = 1 };

// a face code that reflects populations instead of wrapping them (kernel.cl:335-337)
__host__ __device__ inline bool is_wall_code(int c) { return c == 1 || c == 4 || c == 5 || c == 6 || c == 7; }

// ---- unsigned 32-bit division by a run-time constant: q = umulhi(n, m) >> s ----
// Valid for n < 2^31 (checked by the host: cell counts are below 2^31).
struct FastDiv {
    uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv r{d, 0u, 0u};
    if (d <= 1) return r;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;  // ceil(log2 d)
    uint32_t p = 31 + l;
    r.m = (uint32_t)(((1ull << p) + d - 1) / d);
    r.s = p - 32;
    return r;
}

__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv &f)
{
    if (f.d <= 1) return n;
#ifdef __HIP_DEVICE_COMPILE__
    return __umulhi(n, f.m) >> f.s;
#else
    return (uint32_t)(((uint64_t)n * f.m) >> 32) >> f.s;
#endif
}

// p held as an opaque wave-uniform value in scalar registers.  Stores of one lane offset
// from several bases (new populations: direction q at its own base + offset) would otherwise
// be rewritten by the compiler into one 64-bit vector address advanced by a 64-bit vector
// add per store; through this barrier each store takes its array's scalar base and the
// shared 32-bit lane offset.  (An empty asm statement: no instruction is emitted.)
template <typename T>
__device__ inline T *scalar_base(T *p)
{
    uint64_t a = (uint64_t)p;
    asm("" : "+s"(a));
    // back as a global-memory pointer (a plain integer cast would make the stores flat ones)
    typedef __attribute__((address_space(1))) T gT;
    return (T *)(gT *)a;
}

// x held as an opaque wave-uniform value in a scalar register: the compiler cannot relate a
// value computed from it back to the lane's cell index (it would otherwise rewrite the
// pulls' uniform base plus lane offset into per-lane 64-bit arithmetic).  No instruction.
__device__ inline uint32_t scalar_u32(uint32_t x)
{
    asm("" : "+s"(x));
    return x;
}

// work-groups of the face-cell list at the head of a one-launch step (face cells first):
// rounded up to a multiple of 8, the number of XCDs the hardware deals blocks over in turn
__host__ __device__ inline uint32_t merged_face_blocks(uint32_t nface, uint32_t block)
{
    return ((nface + block - 1) / block + 7u) & ~7u;
}

// ---- population layout on the device ----
// 32- and 64-bit values: the cells are grouped in blocks of 2^lb consecutive cells, and a
// block holds its Q directions one after another, each as 2^lb consecutive values: direction
// q of cell c is element ((c >> lb) * Q + q) * 2^lb + (c mod 2^lb).  A wave of 64 aligned
// cells stores each direction as one contiguous run, and the Q streams a block reads and
// writes lie within Q * 2^lb values of each other rather than a whole lattice apart (8192
// cells for 32-bit values, 4096 for 64-bit: the fastest of the memory skeleton of the step).
// 16-bit values: Q arrays of the whole lattice, fs values apart (direction-major, fs from
// population_stride).  Blocks of 64 or 8192 such values made the skeleton faster but the
// 16-bit step slower, in every form of the block pull measured.

// whether values of this width live in direction arrays (otherwise in population blocks)
__host__ __device__ constexpr bool pop_arrays(size_t store_size)
{
    return store_size == 2;
}
__host__ __device__ constexpr uint32_t pop_lb(size_t store_size)
{
    return store_size == 8 ? 12u : 13u;
}
template <int Q, typename S>
__host__ __device__ inline size_t pop_at(int q, uint32_t c, size_t fs)
{
    if constexpr (pop_arrays(sizeof(S))) {
        return (size_t)q * fs + c;
    } else {
        constexpr uint32_t lb = pop_lb(sizeof(S));
        return ((((size_t)(c >> lb)) * (size_t)Q + (size_t)q) << lb) + (c & ((1u << lb) - 1u));
    }
}

// Elements between two direction arrays of 16-bit values: the first count >= n whose byte
// size is 8 KiB past a multiple of 256 KiB.  Arrays whose sizes are large powers of two
// apart fall into the same cache sets and HBM channels when a cell's 19 or 27 pulls go
// out together; the 8 KiB skew spreads the directions over 32 slots of a 256 KiB span.
// (Measured on 256^3 lattices, whose arrays are 2^25 bytes: the 16-bit-storage step
// ran 27 % slower than on 255 x 256 x 256.)
inline size_t population_stride(size_t n, size_t elem)
{
    const size_t span = size_t(1) << 18, skew = 8192;
    const size_t bytes = n * elem;
    const size_t padded = (bytes <= skew) ? skew : ((bytes - skew + span - 1) / span) * span + skew;
    return padded / elem;
}

// elements of the population array of N cells: whole blocks, or Q strided arrays
inline size_t pop_elems(int Q, size_t store_size, size_t N)
{
    if (pop_arrays(store_size)) return (size_t)Q * population_stride(N, store_size);
    const uint32_t lb = pop_lb(store_size);
    return (((N + ((size_t)1 << lb) - 1) >> lb) * (size_t)Q) << lb;
}
// the stride pop_at takes: the direction arrays' for 16-bit values, unused by blocks
inline size_t pop_stride(size_t store_size, size_t N)
{
    return pop_arrays(store_size) ? population_stride(N, store_size) : 0;
}

// ---- storage codecs ----
// 16-bit forms hold (f_q - w_q) * 2^e_q; the power-of-two scale keeps the stored value
// O(0.1..1) for every direction and is exact in both directions of the conversion.
template <int Q>
struct DevScale;

template <>
struct DevScale<19> {
    // w_q = 1/3, 1/18 (axis), 1/36 (diagonal)  ->  scale 4, 16, 32
    __host__ __device__ static constexpr float scale(int q) { return q == 0 ? 4.0f : (q <= 6 ? 16.0f : 32.0f); }
    __host__ __device__ static constexpr float inv(int q) { return q == 0 ? 0.25f : (q <= 6 ? 0.0625f : 0.03125f); }
};

template <>
struct DevScale<27> {
    // w = 8/27, 2/27, 1/54, 1/216 -> scale 4, 16, 64, 256
    __host__ __device__ static constexpr float scale(int q)
    {
        return q == 0 ? 4.0f : (q <= 6 ? 16.0f : (q <= 18 ? 64.0f : 256.0f));
    }
    __host__ __device__ static constexpr float inv(int q)
    {
        return q == 0 ? 0.25f : (q <= 6 ? 0.0625f : (q <= 18 ? 0.015625f : 0.00390625f));
    }
};

// (v - w) scale(q) as one multiply-add: the scale is a power of two, so v scale and
// w scale are exact and fma(v, s, -w s) rounds s (v - w) once -- the same float as the
// subtraction followed by the (exact) scaling, in one instruction
template <int Q>
__device__ inline float dev_scaled(float v, float w, int q)
{
    const float s = DevScale<Q>::scale(q);
    return __builtin_fmaf(v, s, -(w * s));
}

struct bf16_t {
    uint16_t bits;
};

__device__ inline float bf16_to_float(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

// round to nearest even; NaN stays NaN
__device__ inline uint16_t float_to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Stochastic rounding of the 16-bit forms.  Rounding each stored deviation to the nearest
// 16-bit value removes perturbations below half a unit in the last place every step, and a
// flow held symmetric by that rounding never leaves it (a cylinder wake at Re 100 does not
// shed); rounding up with the probability of the remainder keeps the stored value unbiased.
// The random bits come from a hash of the cell's moments of this step, its index and the
// step number -- a deterministic function of the state and the step count, the same in
// every launch shape, batch and graph replay.  The step number makes the bits of a cell
// whose state no longer changes (a steady flow) differ from step to step, so that its
// rounding errors average out over time instead of repeating.
__device__ inline uint32_t sr_hash(float a, float b, float c, float d, uint32_t ic, uint32_t step)
{
    const uint32_t ub = __float_as_uint(b), uc = __float_as_uint(c), ud = __float_as_uint(d);
    uint32_t h = __float_as_uint(a) ^ ((ub << 11) | (ub >> 21)) ^ ((uc << 19) | (uc >> 13)) ^ ((ud << 27) | (ud >> 5)) ^
                 (ic * 0x9E3779B1u);
    h *= 0x2C1B3C6Du;
    h ^= h >> 15;
    h ^= step * 0x85EBCA77u;
    h *= 0xC2B2AE3Du;
    h ^= h >> 13;
    return h;
}

// the step number a launch runs: `step` (its own, or the first step of a batched call) plus
// the batch's device counter of steps within the call (null outside a batch; read through
// the scalar cache)
__device__ inline uint32_t sr_step(uint32_t step, const uint32_t *ctr)
{
    if (!ctr) return step;
    typedef const __attribute__((address_space(4))) uint32_t const_u32;
    return step + *(const_u32 *)ctr;
}

// the bits of direction q: the cell's hash advanced by q golden-ratio steps, which spreads
// the 19 (27) rounding thresholds of a cell evenly over [0, 1) -- one add per direction
__device__ inline uint32_t sr_bits(uint32_t h, int q) { return h + (uint32_t)q * 0x9E3779B9u; }

// Store<S,Real>: load/store of population q with weight w as compute type Real;
// store_nt_sr(..., r) is the non-temporal store with stochastic rounding from the random
// bits r (the 32-bit forms store exactly and ignore r).
template <typename S, typename Real, int Q>
struct Codec;

template <int Q>
struct Codec<float, float, Q> {
    __device__ static inline float load(const float *p, int, float) { return *p; }
    __device__ static inline float load_nt(const float *p, int, float) { return __builtin_nontemporal_load(p); }
    __device__ static inline void store(float *p, float v, int, float) { *p = v; }
    __device__ static inline void store_nt(float *p, float v, int, float) { __builtin_nontemporal_store(v, p); }
    __device__ static inline void store_nt_sr(float *p, float v, int q, float w, uint32_t) { store_nt(p, v, q, w); }
    __device__ static inline void store_nt_sr2(float *p0, float *p1, float v0, float v1, int q0, int q1, float w0,
                                               float w1, uint32_t, uint32_t)
    {
        store_nt(p0, v0, q0, w0);
        store_nt(p1, v1, q1, w1);
    }
};

template <int Q>
struct Codec<double, double, Q> {
    __device__ static inline double load(const double *p, int, double) { return *p; }
    __device__ static inline double load_nt(const double *p, int, double) { return __builtin_nontemporal_load(p); }
    __device__ static inline void store(double *p, double v, int, double) { *p = v; }
    __device__ static inline void store_nt(double *p, double v, int, double) { __builtin_nontemporal_store(v, p); }
    __device__ static inline void store_nt_sr(double *p, double v, int q, double w, uint32_t) { store_nt(p, v, q, w); }
    __device__ static inline void store_nt_sr2(double *p0, double *p1, double v0, double v1, int q0, int q1, double w0,
                                               double w1, uint32_t, uint32_t)
    {
        store_nt(p0, v0, q0, w0);
        store_nt(p1, v1, q1, w1);
    }
};

template <int Q>
struct Codec<__half, float, Q> {
    __device__ static inline float load(const __half *p, int q, float w)
    {
        return w + __half2float(*p) * DevScale<Q>::inv(q);
    }
    __device__ static inline float load_nt(const __half *p, int q, float w)
    {
        const unsigned short b = __builtin_nontemporal_load((const unsigned short *)p);
        return w + __half2float(__ushort_as_half(b)) * DevScale<Q>::inv(q);
    }
    __device__ static inline void store(__half *p, float v, int q, float w)
    {
        *p = __float2half_rn(dev_scaled<Q>(v, w, q));
    }
    __device__ static inline void store_nt(__half *p, float v, int q, float w)
    {
        __half h = __float2half_rn(dev_scaled<Q>(v, w, q));
        __builtin_nontemporal_store(__half_as_ushort(h), (unsigned short *)p);
    }
    __device__ static inline void store_nt_sr(__half *p, float v, int q, float w, uint32_t r)
    {
        typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
        h2_t h = {0, 0};
        h = __builtin_amdgcn_cvt_sr_f16_f32(h, dev_scaled<Q>(v, w, q), r, false);
        const _Float16 e = h[0];
        unsigned short bits;
        __builtin_memcpy(&bits, &e, 2);
        __builtin_nontemporal_store(bits, (unsigned short *)p);
    }
    // two populations rounded into the two halves of one register (no register to clear
    // for the conversion's other half), stored from its low and high half
    __device__ static inline void store_nt_sr2(__half *p0, __half *p1, float v0, float v1, int q0, int q1, float w0,
                                               float w1, uint32_t r0, uint32_t r1)
    {
        typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
        // both halves are written below: start from the register of the random bits, which
        // the first conversion consumes, rather than from a cleared one
        h2_t h = __builtin_bit_cast(h2_t, r0);
        h = __builtin_amdgcn_cvt_sr_f16_f32(h, dev_scaled<Q>(v0, w0, q0), r0, false);
        h = __builtin_amdgcn_cvt_sr_f16_f32(h, dev_scaled<Q>(v1, w1, q1), r1, true);
        const _Float16 e0 = h[0], e1 = h[1];
        unsigned short b0, b1;
        __builtin_memcpy(&b0, &e0, 2);
        __builtin_memcpy(&b1, &e1, 2);
        __builtin_nontemporal_store(b0, (unsigned short *)p0);
        __builtin_nontemporal_store(b1, (unsigned short *)p1);
    }
};

template <int Q>
struct Codec<bf16_t, float, Q> {
    __device__ static inline float load(const bf16_t *p, int q, float w)
    {
        return w + bf16_to_float(p->bits) * DevScale<Q>::inv(q);
    }
    __device__ static inline float load_nt(const bf16_t *p, int q, float w)
    {
        return w + bf16_to_float(__builtin_nontemporal_load(&p->bits)) * DevScale<Q>::inv(q);
    }
    __device__ static inline void store(bf16_t *p, float v, int q, float w)
    {
        p->bits = float_to_bf16(dev_scaled<Q>(v, w, q));
    }
    __device__ static inline void store_nt(bf16_t *p, float v, int q, float w)
    {
        __builtin_nontemporal_store(float_to_bf16(dev_scaled<Q>(v, w, q)), (unsigned short *)p);
    }
    __device__ static inline void store_nt_sr(bf16_t *p, float v, int q, float w, uint32_t r)
    {
        typedef short b2_t __attribute__((ext_vector_type(2)));
        b2_t h = {0, 0};
        h = __builtin_amdgcn_cvt_sr_bf16_f32(h, dev_scaled<Q>(v, w, q), r, false);
        __builtin_nontemporal_store((unsigned short)h[0], (unsigned short *)p);
    }
    __device__ static inline void store_nt_sr2(bf16_t *p0, bf16_t *p1, float v0, float v1, int q0, int q1, float w0,
                                               float w1, uint32_t r0, uint32_t r1)
    {
        typedef short b2_t __attribute__((ext_vector_type(2)));
        b2_t h = __builtin_bit_cast(b2_t, r0);  // both halves are written below (as for FP16)
        h = __builtin_amdgcn_cvt_sr_bf16_f32(h, dev_scaled<Q>(v0, w0, q0), r0, false);
        h = __builtin_amdgcn_cvt_sr_bf16_f32(h, dev_scaled<Q>(v1, w1, q1), r1, true);
        __builtin_nontemporal_store((unsigned short)h[0], (unsigned short *)p0);
        __builtin_nontemporal_store((unsigned short)h[1], (unsigned short *)p1);
    }
};

}  // namespace lbffs
