/*
 * ibm_oracle.c -- CPU restatement of the reference's immersed-boundary kernels
 * (k_Gwall, k_Gibm, k_Force and helpers).  TEST INFRASTRUCTURE ONLY (see lbm19_oracle.c).
 * Motion parameters m[10] = {uMovingTrans, vMovingTrans, wMovingTrans, rotX, rotY, rotZ,
 * rotAxisX, rotAxisY, rotAxisZ, rotOmega} in lattice units (walls.hpp:586-628).
 */
#include <math.h>
#include <stdlib.h>

#include "lbm_oracle.h"

/* K19:1112-1155 icBox: corner b of the unit box above cell ic (no wrap; -1 outside) */
long orc_ibm_icbox(long ic, int b, int nx, int ny, int nz)
{
    const int i = (int)(ic % ((long)nx * ny) % nx), j = (int)(ic % ((long)nx * ny) / nx), k = (int)(ic / ((long)nx * ny));
    const int di = (b == 1 || b == 4 || b == 5 || b == 7), dj = (b == 2 || b == 4 || b == 6 || b == 7),
              dk = (b == 3 || b == 5 || b == 6 || b == 7);
    if (i + di >= nx || j + dj >= ny || k + dk >= nz) return -1;
    return orc_index(i + di, j + dj, k + dk, nx, ny);
}

#define RDIV(a, b) ((REAL)(a) / (REAL)(b))

#define REAL float
#define SFX(name) name##_f32
#define SQRT sqrtf
#define FABS fabsf
#define COS cosf
#define SIN sinf
#include "ibm_body.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS
#undef COS
#undef SIN

#define REAL double
#define SFX(name) name##_f64
#define SQRT sqrt
#define FABS fabs
#define COS cos
#define SIN sin
#include "ibm_body.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS
#undef COS
#undef SIN
