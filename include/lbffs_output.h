/*
 * lbffs_output.h -- the output step of the LBFFS driver, as a C ABI (host code).
 *
 * At every output time the reference host (write.hpp, included at the top of the time
 * loop, D3Q19_opencl/ns.cpp:511; D3Q27_opencl/write.hpp is the same with 27 directions)
 * downloads f, tauSGS and omega (and the force fields), recomputes rho and u from f
 * (lbffs_output_macros of lbffs.h does that on the device), prints diagnostics, adds
 * u to the running means, writes ./out/resNNNNNNNN.vtk and overwrites ./out/data.dat,
 * the restart file read by restart.hpp.  These functions reproduce the file formats and
 * the diagnostics' arithmetic; the caller owns the arrays (nx*ny*nz entries, cell
 * index ic = i + nx*(j + ny*k), populations q-major).
 *
 * Returns 0, or a negative LBFFS_E_* code of lbffs.h (LBFFS_E_INVALID for bad arguments
 * and for files that cannot be opened, read or written; lbffs_last_error() does not
 * cover these host functions).
 */
#ifndef LBFFS_OUTPUT_H
#define LBFFS_OUTPUT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The fields of write.hpp:119-600, in file order, lattice units as the engine holds them;
 * the writer applies the reference's scalings (u*c, 3*(3 tauSGS L c), sdf*L, uMean*c/timeMean). */
typedef struct {
    int nx, ny, nz;
    float L, c;       /* cell size and velocity scale (ns.cpp:103-104) */
    float timeMean;   /* averaging time of uMean/vMean/wMean (write.hpp:68) */
    const float *u, *v, *w, *rho, *tauSGS;
    const int32_t *boundary1, *boundary2, *boundary3;
    const float *sdf;
    const uint8_t *solid, *neiSolid;
    const float *GxIBM, *GyIBM, *GzIBM, *omega;
    const float *uMean, *vMean, *wMean;
} lbffs_vtk_fields;

enum {
    LBFFS_VTK_ASCII = 0,       /* writeBinary false: values printed as std::ostream prints floats */
    LBFFS_VTK_BINARY = 1,      /* legacy VTK binary: big-endian float32 */
    LBFFS_VTK_BINARY_REF = 2   /* the reference's byte order (base.hpp:35-43), whose first byte
                                  repeats the float's low byte: files identical to the reference's */
};

/* write.hpp:76-600: one legacy-VTK STRUCTURED_POINTS file with the 18 scalar fields. */
int lbffs_out_write_vtk(const char *path, const lbffs_vtk_fields *fields, int mode);

/* write.hpp:602-641: int32 nt, nx, ny, nz, then q*nx*ny*nz float32 populations, q-major
 * (native byte order).  f holds FP32 values whatever the engine's storage form. */
int lbffs_out_write_restart(const char *path, int nt, int q, int nx, int ny, int nz, const float *f);

/* restart.hpp:1-28: reads the header into *nt, *nx, *ny, *nz; when f is not NULL also the
 * q*nx*ny*nz populations, which must fill exactly f_count floats. */
int lbffs_out_read_restart(const char *path, int q, int *nt, int *nx, int *ny, int *nz, float *f, size_t f_count);

/* Inputs of the output-step diagnostics (write.hpp:16-70, 93-99). */
typedef struct {
    int nx, ny, nz;
    float c, L, uMax, Dref;
    int forceCoeffs;
    const float *rho, *u;              /* lbffs_output_macros results */
    const uint8_t *solid;
    const float *Fwx, *Fwy, *Fwz;      /* read only when forceCoeffs != 0 */
    const float *GxIBM, *GyIBM, *GzIBM;
} lbffs_out_stats_in;

typedef struct {
    float rho_av;            /* mean density of the fluid cells (sum of rho-1 in double) */
    float rhoMax, rhoMin;    /* over every cell */
    float uAve;              /* mean of u*c over the i = 0 plane */
    int eleFluid;
    float sumFw[3], sumG[3]; /* FP32 sums over fluid cells, in cell order */
    double Cx, Cy, Cz;       /* sumFw c^2 L^2 / (0.5 rho_av uMax^2 Dref L) */
    double CxIBM, CyIBM, CzIBM; /* -sumG ... */
} lbffs_out_stats;

/* write.hpp:16-60 and 93-99 with the reference's FP32/FP64 arithmetic. */
int lbffs_out_stats_compute(const lbffs_out_stats_in *in, lbffs_out_stats *out);

/* The same coefficients from the engine's FP64 device sums (lbffs_reduce of lbffs.h:
 * sums[0..2] Fw, [3..5] G, [6] sum of rho-1, [7] fluid cells), in FP64 throughout.
 * out: rho_av, Cx, Cy, Cz, CxIBM, CyIBM, CzIBM. */
int lbffs_out_coeffs_fp64(const double sums[8], float c, float L, float uMax, float Dref, double out[7]);

/* write.hpp:68-74: timeMean += interval; uMean += u*interval (FP32), for n cells. */
int lbffs_out_accumulate_mean(long n, int interval, const float *u, const float *v, const float *w, float *uMean,
                              float *vMean, float *wMean, float *timeMean);

#ifdef __cplusplus
}
#endif
#endif
