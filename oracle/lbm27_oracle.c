/*
 * lbm27_oracle.c -- CPU restatement of the reference D3Q19/D3Q27 time step, D3Q27 part.
 *
 * TEST INFRASTRUCTURE ONLY (see lbm19_oracle.c).  Follows D3Q27_opencl/kernel.cl (K27).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"

/* K27:2742-2747 direction order */
const int ORC27_CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1};
const int ORC27_CY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
const int ORC27_CZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, -1, 1};

static int c2q27(int cx, int cy, int cz)
{
    for (int q = 0; q < 27; q++)
        if (ORC27_CX[q] == cx && ORC27_CY[q] == cy && ORC27_CZ[q] == cz) return q;
    orc_fatal("c2q27: no such direction");
    return -1;
}

/* K27:48-160 reflectQ */
int orc27_reflect(int q) { return c2q27(-ORC27_CX[q], -ORC27_CY[q], -ORC27_CZ[q]); }

/* K27:186-528 mirrorQx / mirrorQy / mirrorQz (the tables, not the unused mirrorQ*s
 * variants): a direction with a component along the axis has that component flipped;
 * one without (parallel to the mirror plane) is reflected whole, e.g. mirrorQx(3) = 4. */
int orc27_mirror(int q, int axis)
{
    int c[3] = {ORC27_CX[q], ORC27_CY[q], ORC27_CZ[q]};
    if (c[axis] != 0) {
        c[axis] = -c[axis];
    } else {
        c[0] = -c[0];
        c[1] = -c[1];
        c[2] = -c[2];
    }
    return c2q27(c[0], c[1], c[2]);
}

/* K27:530-549 upwindID_B */
long orc27_upwind_B(int q, int i, int j, int k, int nx, int ny, int nz, int b1, int b2, int b3)
{
    const int w1 = !(b1 == 1 || b1 == 4 || b1 == 5 || b1 == 6 || b1 == 7);
    const int w2 = !(b2 == 1 || b2 == 4 || b2 == 5 || b2 == 6 || b2 == 7);
    const int w3 = !(b3 == 1 || b3 == 4 || b3 == 5 || b3 == 6 || b3 == 7);
    int upi = i - ORC27_CX[q], upj = j - ORC27_CY[q], upk = k - ORC27_CZ[q];
    upi = (upi == -1) ? (w1 ? nx - 1 : -1) : ((upi == nx) ? (w1 ? 0 : -1) : upi);
    upj = (upj == -1) ? (w2 ? ny - 1 : -1) : ((upj == ny) ? (w2 ? 0 : -1) : upj);
    upk = (upk == -1) ? (w3 ? nz - 1 : -1) : ((upk == nz) ? (w3 ? 0 : -1) : upk);
    return (upi == -1 || upj == -1 || upk == -1) ? -1 : orc_index(upi, upj, upk, nx, ny);
}

#define REAL float
#define SFX(name) name##_f32
#define SQRT sqrtf
#define POW powf
#define EXP expf
#define FABS fabsf
#define FMIN fminf
#define FMAX fmaxf
#include "lbm27_body.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef POW
#undef EXP
#undef FABS
#undef FMIN
#undef FMAX

#define REAL double
#define SFX(name) name##_f64
#define SQRT sqrt
#define POW pow
#define EXP exp
#define FABS fabs
#define FMIN fmin
#define FMAX fmax
#include "lbm27_body.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef POW
#undef EXP
#undef FABS
#undef FMIN
#undef FMAX
