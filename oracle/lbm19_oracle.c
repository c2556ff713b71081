/*
 * lbm19_oracle.c -- CPU restatement of the reference D3Q19 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product path
 * (lbffs_amd/csrc); only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it.  It follows D3Q19_opencl/kernel.cl (K19:line) function by
 * function; the integer helpers below are literal restatements of the reference's
 * branch chains so that the product kernel's table-driven formulation is checked
 * against an independent form.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"

const int ORC_CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const int ORC_CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
const int ORC_CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};

void orc_fatal(const char *msg)
{
    fprintf(stderr, "lbm oracle: %s\n", msg);
    abort();
}

/* K19:17-20 index1d */
long orc_index(int i, int j, int k, int nx, int ny) { return (long)nx * ny * k + (long)nx * j + i; }

/* K19:32-103 reflectQ: opposite direction */
int orc_reflect(int q)
{
    if (q == 0) return 0;
    return (q % 2 == 1) ? q + 1 : q - 1;
}

/* K19:105-326 reflectOrMirrorQ: Symmetry faces mirror the diagonal directions */
int orc_reflect_or_mirror(int q, int b1, int b2, int b3)
{
    switch (q) {
    case 0: return 0;
    case 1: return 2;
    case 2: return 1;
    case 3: return 4;
    case 4: return 3;
    case 5: return 6;
    case 6: return 5;
    case 7: return b1 == 4 ? 10 : b2 == 4 ? 9 : 8;
    case 8: return b1 == 4 ? 9 : b2 == 4 ? 10 : 7;
    case 9: return b1 == 4 ? 8 : b2 == 4 ? 7 : 10;
    case 10: return b1 == 4 ? 7 : b2 == 4 ? 8 : 9;
    case 11: return b1 == 4 ? 14 : b3 == 4 ? 13 : 12;
    case 12: return b1 == 4 ? 13 : b3 == 4 ? 14 : 11;
    case 13: return b1 == 4 ? 12 : b3 == 4 ? 11 : 14;
    case 14: return b1 == 4 ? 11 : b3 == 4 ? 12 : 13;
    case 15: return b2 == 4 ? 18 : b3 == 4 ? 17 : 16;
    case 16: return b2 == 4 ? 17 : b3 == 4 ? 18 : 15;
    case 17: return b2 == 4 ? 16 : b3 == 4 ? 15 : 18;
    case 18: return b2 == 4 ? 15 : b3 == 4 ? 16 : 17;
    }
    orc_fatal("reflect_or_mirror: q out of range");
    return -1;
}

/* K19:335-337 */
static int not_wall(int code) { return !(code == 1 || code == 4 || code == 5 || code == 6 || code == 7); }

/* K19:328-575 upwindID_B: pull source of direction q for a face cell; -1 marks a
 * wall link (face code BounceBack/Symmetry/FixedVelocity/FixedDensity/WallFunction).
 * Non-wall faces (Cyclic, Outlet) wrap periodically. */
long orc_upwind_B(int q, int i, int j, int k, int nx, int ny, int nz, int b1, int b2, int b3)
{
    const int w1 = not_wall(b1), w2 = not_wall(b2), w3 = not_wall(b3);
#define IX(a, b, c) orc_index((a), (b), (c), nx, ny)
    switch (q) {
    case 0: return IX(i, j, k);
    case 1: return i != 0 ? IX(i - 1, j, k) : (w1 ? IX(nx - 1, j, k) : -1);
    case 2: return i != nx - 1 ? IX(i + 1, j, k) : (w1 ? IX(0, j, k) : -1);
    case 3: return j != 0 ? IX(i, j - 1, k) : (w2 ? IX(i, ny - 1, k) : -1);
    case 4: return j != ny - 1 ? IX(i, j + 1, k) : (w2 ? IX(i, 0, k) : -1);
    case 5: return k != 0 ? IX(i, j, k - 1) : (w3 ? IX(i, j, nz - 1) : -1);
    case 6: return k != nz - 1 ? IX(i, j, k + 1) : (w3 ? IX(i, j, 0) : -1);
    case 7:
        return (i != 0 && j != 0) ? IX(i - 1, j - 1, k)
             : (i == 0 && j != 0) ? (w1 ? IX(nx - 1, j - 1, k) : -1)
             : (i != 0 && j == 0) ? (w2 ? IX(i - 1, ny - 1, k) : -1)
                                  : ((w1 && w2) ? IX(nx - 1, ny - 1, k) : -1);
    case 8:
        return (i != nx - 1 && j != ny - 1) ? IX(i + 1, j + 1, k)
             : (i == nx - 1 && j != ny - 1) ? (w1 ? IX(0, j + 1, k) : -1)
             : (i != nx - 1 && j == ny - 1) ? (w2 ? IX(i + 1, 0, k) : -1)
                                            : ((w1 && w2) ? IX(0, 0, k) : -1);
    case 9:
        return (i != 0 && j != ny - 1) ? IX(i - 1, j + 1, k)
             : (i == 0 && j != ny - 1) ? (w1 ? IX(nx - 1, j + 1, k) : -1)
             : (i != 0 && j == ny - 1) ? (w2 ? IX(i - 1, 0, k) : -1)
                                       : ((w1 && w2) ? IX(nx - 1, 0, k) : -1);
    case 10:
        return (i != nx - 1 && j != 0) ? IX(i + 1, j - 1, k)
             : (i == nx - 1 && j != 0) ? (w1 ? IX(0, j - 1, k) : -1)
             : (i != nx - 1 && j == 0) ? (w2 ? IX(i + 1, ny - 1, k) : -1)
                                       : ((w1 && w2) ? IX(0, ny - 1, k) : -1);
    case 11:
        return (i != 0 && k != 0) ? IX(i - 1, j, k - 1)
             : (i == 0 && k != 0) ? (w1 ? IX(nx - 1, j, k - 1) : -1)
             : (i != 0 && k == 0) ? (w3 ? IX(i - 1, j, nz - 1) : -1)
                                  : ((w1 && w3) ? IX(nx - 1, j, nz - 1) : -1);
    case 12:
        return (i != nx - 1 && k != nz - 1) ? IX(i + 1, j, k + 1)
             : (i == nx - 1 && k != nz - 1) ? (w1 ? IX(0, j, k + 1) : -1)
             : (i != nx - 1 && k == nz - 1) ? (w3 ? IX(i + 1, j, 0) : -1)
                                            : ((w1 && w3) ? IX(0, j, 0) : -1);
    case 13:
        return (i != 0 && k != nz - 1) ? IX(i - 1, j, k + 1)
             : (i == 0 && k != nz - 1) ? (w1 ? IX(nx - 1, j, k + 1) : -1)
             : (i != 0 && k == nz - 1) ? (w3 ? IX(i - 1, j, 0) : -1)
                                       : ((w1 && w3) ? IX(nx - 1, j, 0) : -1);
    case 14:
        return (i != nx - 1 && k != 0) ? IX(i + 1, j, k - 1)
             : (i == nx - 1 && k != 0) ? (w1 ? IX(0, j, k - 1) : -1)
             : (i != nx - 1 && k == 0) ? (w3 ? IX(i + 1, j, nz - 1) : -1)
                                       : ((w1 && w3) ? IX(0, j, nz - 1) : -1);
    case 15:
        return (j != 0 && k != 0) ? IX(i, j - 1, k - 1)
             : (j == 0 && k != 0) ? (w2 ? IX(i, ny - 1, k - 1) : -1)
             : (j != 0 && k == 0) ? (w3 ? IX(i, j - 1, nz - 1) : -1)
                                  : ((w2 && w3) ? IX(i, ny - 1, nz - 1) : -1);
    case 16:
        return (j != ny - 1 && k != nz - 1) ? IX(i, j + 1, k + 1)
             : (j == ny - 1 && k != nz - 1) ? (w2 ? IX(i, 0, k + 1) : -1)
             : (j != ny - 1 && k == nz - 1) ? (w3 ? IX(i, j + 1, 0) : -1)
                                            : ((w2 && w3) ? IX(i, 0, 0) : -1);
    case 17:
        return (j != 0 && k != nz - 1) ? IX(i, j - 1, k + 1)
             : (j == 0 && k != nz - 1) ? (w2 ? IX(i, ny - 1, k + 1) : -1)
             : (j != 0 && k == nz - 1) ? (w3 ? IX(i, j - 1, 0) : -1)
                                       : ((w2 && w3) ? IX(i, ny - 1, 0) : -1);
    case 18:
        return (j != ny - 1 && k != 0) ? IX(i, j + 1, k - 1)
             : (j == ny - 1 && k != 0) ? (w2 ? IX(i, 0, k - 1) : -1)
             : (j != ny - 1 && k == 0) ? (w3 ? IX(i, j + 1, nz - 1) : -1)
                                       : ((w2 && w3) ? IX(i, 0, nz - 1) : -1);
    }
#undef IX
    return 0;
}

/* K19:823-905 upwindID_internal (interior cells, no wrap) */
long orc_upwind_internal(int q, int i, int j, int k, int nx, int ny)
{
    return orc_index(i - ORC_CX[q], j - ORC_CY[q], k - ORC_CZ[q], nx, ny);
}

/* K19:907-989 upwindID_2D (nz==1: z neighbours are the single layer) */
long orc_upwind_2D(int q, int i, int j, int k, int nx, int ny, int nz)
{
    (void)k;
    int kk = (ORC_CZ[q] == 1) ? nz - 1 : (ORC_CZ[q] == -1) ? 0 : k;
    return orc_index(i - ORC_CX[q], j - ORC_CY[q], kk, nx, ny);
}

/* K19:1346-1395 cornerFlag: edge cells skip the velocity/density/wall-function BCs */
int orc_corner(int b3, int i, int j, int k, int nx, int ny, int nz)
{
    int xy = (i == 0 && j == 0) || (i == 0 && j == ny - 1) || (i == nx - 1 && j == 0) || (i == nx - 1 && j == ny - 1);
    if (nz - 1 == 0 || b3 == 0) return xy ? 1 : 0;
    int xz = (i == 0 && k == 0) || (i == 0 && k == nz - 1) || (i == nx - 1 && k == 0) || (i == nx - 1 && k == nz - 1);
    int yz = (j == 0 && k == 0) || (j == 0 && k == nz - 1) || (j == ny - 1 && k == 0) || (j == ny - 1 && k == nz - 1);
    return (xy || xz || yz) ? 1 : 0;
}

/* ---- FP32 instantiation ---- */
#define REAL float
#define SFX(name) name##_f32
#define SQRT sqrtf
#define POW powf
#define EXP expf
#define LOG logf
#define FABS fabsf
#define FMIN fminf
#define FMAX fmaxf
#include "lbm19_body.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef POW
#undef EXP
#undef LOG
#undef FABS
#undef FMIN
#undef FMAX

/* ---- FP64 instantiation ---- */
#define REAL double
#define SFX(name) name##_f64
#define SQRT sqrt
#define POW pow
#define EXP exp
#define LOG log
#define FABS fabs
#define FMIN fmin
#define FMAX fmax
#include "lbm19_body.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef POW
#undef EXP
#undef LOG
#undef FABS
#undef FMIN
#undef FMAX
