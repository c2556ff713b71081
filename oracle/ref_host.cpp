// ref_host.cpp -- harness that compiles two textual fragments of the reference host,
// src/setBoundaryConditions.hpp and D3Q19_opencl/initialization.hpp, exactly as
// ns.cpp:162 and ns.cpp:185 include them: inside a function that declares the locals
// they use.  TEST INFRASTRUCTURE ONLY: the resulting oracle/_ref/libref_host.so pins the
// product's host case setup (lbffs_case.cpp) bit for bit.
// the standard headers ns.cpp:1-12 includes before D3Q19.hpp
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "D3Q19.hpp"  // D3Q19_opencl/D3Q19.hpp (+ src/base.hpp): setWt, setCx.., index1d, idf

extern "C" {

// codes[6][3], vel[6][3]: faces nxMin,nxMax,nyMin,nyMax,nzMin,nzMax; vel already in lattice units
void ref_set_boundary_conditions(int nx, int ny, int nz, const int *codes, const float *vel, int32_t *out_b1,
                                 int32_t *out_b2, int32_t *out_b3, float *out_u0, float *out_v0, float *out_w0)
{
    const size_t N = (size_t)nx * ny * nz;
    std::vector<int> boundary1(N, 0), boundary2(N, 0), boundary3(N, 0);
    std::vector<float> u0(N, 0.f), v0(N, 0.f), w0(N, 0.f);
#define C(f, d) codes[(f)*3 + (d)]
#define V(f, d) vel[(f)*3 + (d)]
    int nxMinBoundary1 = C(0, 0), nxMaxBoundary1 = C(1, 0);
    int nyMinBoundary2 = C(2, 1), nyMaxBoundary2 = C(3, 1);
    int nzMinBoundary3 = C(4, 2), nzMaxBoundary3 = C(5, 2);
    float nxMinU0 = V(0, 0), nxMinV0 = V(0, 1), nxMinW0 = V(0, 2);
    float nxMaxU0 = V(1, 0), nxMaxV0 = V(1, 1), nxMaxW0 = V(1, 2);
    float nyMinU0 = V(2, 0), nyMinV0 = V(2, 1), nyMinW0 = V(2, 2);
    float nyMaxU0 = V(3, 0), nyMaxV0 = V(3, 1), nyMaxW0 = V(3, 2);
    float nzMinU0 = V(4, 0), nzMinV0 = V(4, 1), nzMinW0 = V(4, 2);
    float nzMaxU0 = V(5, 0), nzMaxV0 = V(5, 1), nzMaxW0 = V(5, 2);
#undef C
#undef V
#include "setBoundaryConditions.hpp"
    for (size_t ic = 0; ic < N; ic++) {
        out_b1[ic] = boundary1[ic];
        out_b2[ic] = boundary2[ic];
        out_b3[ic] = boundary3[ic];
        out_u0[ic] = u0[ic];
        out_v0[ic] = v0[ic];
        out_w0[ic] = w0[ic];
    }
}

void ref_initialization(int nx, int ny, int nz, float rho0, float uIni, float vIni, float wIni, float *out_f,
                        float *out_fTmp)
{
    const std::vector<float> wt = setWt();
    const std::vector<float> cx = setCx();
    const std::vector<float> cy = setCy();
    const std::vector<float> cz = setCz();
    const size_t N = (size_t)nx * ny * nz;
    std::vector<float> f(19 * N), fTmp(19 * N);
#include "initialization.hpp"
    std::memcpy(out_f, f.data(), 19 * N * sizeof(float));
    std::memcpy(out_fTmp, fTmp.data(), 19 * N * sizeof(float));
}

}  // extern "C"
