/*
 * output_oracle.c -- the host's recomputation of the macroscopic fields at an output
 * step, restated for the CPU oracle.  TEST INFRASTRUCTURE ONLY: the checker of
 * lbffs_output_macros; only tests/ may load it.
 *
 * cal_rhoUVW of D3Q19_opencl/D3Q19.hpp:647-666 and D3Q27_opencl/D3Q27.hpp:28-47: FP32
 * sums over the directions in order of f, f cx, f cy, f cz; u, v, w divided by rho; then
 * u += 0.5 dpdx / rho.  Populations q-major, f[q*N + ic].
 */
#include "lbm_oracle.h"

void orc_cal_rhouvw(int Q, long N, const float *f, float dpdx, float *rho, float *u, float *v, float *w)
{
    const int *cx = Q == 27 ? ORC27_CX : ORC_CX, *cy = Q == 27 ? ORC27_CY : ORC_CY, *cz = Q == 27 ? ORC27_CZ : ORC_CZ;
    for (long ic = 0; ic < N; ic++) {
        float r = 0.0f, a = 0.0f, b = 0.0f, c = 0.0f;
        for (int q = 0; q < Q; q++) {
            const float fq = f[(long)q * N + ic];
            r += fq;
            a += fq * (float)cx[q];
            b += fq * (float)cy[q];
            c += fq * (float)cz[q];
        }
        a /= r;
        b /= r;
        c /= r;
        a += 0.5f * dpdx / r;
        rho[ic] = r;
        u[ic] = a;
        v[ic] = b;
        w[ic] = c;
    }
}
