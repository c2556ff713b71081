// lbffs_output.cpp -- the output step of the LBFFS driver (include/lbffs_output.h): the
// legacy-VTK file and the restart file of write.hpp / restart.hpp, and the diagnostics
// printed at an output step.  Host code, built without FP contraction so that the FP32
// arithmetic matches the reference host.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/lbffs.h"
#include "../../include/lbffs_output.h"

namespace {

// one scalar field of the VTK file: its name and its value at cell ic, as written
struct VtkField {
    const char *name;
    bool is_int;  // ASCII: printed as an integer (the boundary codes)
    float (*value)(const lbffs_vtk_fields &, long ic);
    int (*ivalue)(const lbffs_vtk_fields &, long ic);
};

// write.hpp:127-599, in file order, with the reference's scalings
const VtkField kFields[] = {
    {"u", false, [](const lbffs_vtk_fields &F, long ic) { return F.u[ic] * F.c; }, nullptr},
    {"v", false, [](const lbffs_vtk_fields &F, long ic) { return F.v[ic] * F.c; }, nullptr},
    {"w", false, [](const lbffs_vtk_fields &F, long ic) { return F.w[ic] * F.c; }, nullptr},
    {"rho", false, [](const lbffs_vtk_fields &F, long ic) { return F.rho[ic]; }, nullptr},
    // nuSGS = 3 tauSGS L c, written times 3 (write.hpp:236-241)
    {"nuSGS", false,
     [](const lbffs_vtk_fields &F, long ic) {
         const float nuSGS = 3.f * F.tauSGS[ic] * F.L * F.c;
         return 3.f * nuSGS;
     },
     nullptr},
    {"boundary1", true, [](const lbffs_vtk_fields &F, long ic) { return (float)F.boundary1[ic]; },
     [](const lbffs_vtk_fields &F, long ic) { return (int)F.boundary1[ic]; }},
    {"boundary2", true, [](const lbffs_vtk_fields &F, long ic) { return (float)F.boundary2[ic]; },
     [](const lbffs_vtk_fields &F, long ic) { return (int)F.boundary2[ic]; }},
    {"boundary3", true, [](const lbffs_vtk_fields &F, long ic) { return (float)F.boundary3[ic]; },
     [](const lbffs_vtk_fields &F, long ic) { return (int)F.boundary3[ic]; }},
    {"sdf", false, [](const lbffs_vtk_fields &F, long ic) { return F.sdf[ic] * F.L; }, nullptr},
    {"solid", false, [](const lbffs_vtk_fields &F, long ic) { return (float)F.solid[ic]; }, nullptr},
    {"neiSolid", false, [](const lbffs_vtk_fields &F, long ic) { return (float)F.neiSolid[ic]; }, nullptr},
    {"GxIBM", false, [](const lbffs_vtk_fields &F, long ic) { return F.GxIBM[ic]; }, nullptr},
    {"GyIBM", false, [](const lbffs_vtk_fields &F, long ic) { return F.GyIBM[ic]; }, nullptr},
    {"GzIBM", false, [](const lbffs_vtk_fields &F, long ic) { return F.GzIBM[ic]; }, nullptr},
    {"omega", false, [](const lbffs_vtk_fields &F, long ic) { return F.omega[ic]; }, nullptr},
    {"uMean", false, [](const lbffs_vtk_fields &F, long ic) { return F.uMean[ic] * F.c / F.timeMean; }, nullptr},
    {"vMean", false, [](const lbffs_vtk_fields &F, long ic) { return F.vMean[ic] * F.c / F.timeMean; }, nullptr},
    {"wMean", false, [](const lbffs_vtk_fields &F, long ic) { return F.wMean[ic] * F.c / F.timeMean; }, nullptr},
};

bool fields_complete(const lbffs_vtk_fields &F)
{
    const void *p[] = {F.u, F.v, F.w, F.rho, F.tauSGS, F.boundary1, F.boundary2, F.boundary3, F.sdf,
                       F.solid, F.neiSolid, F.GxIBM, F.GyIBM, F.GzIBM, F.omega, F.uMean, F.vMean, F.wMean};
    for (const void *q : p)
        if (!q) return false;
    return true;
}

// the four bytes of a float in the chosen binary order
inline void float_bytes(float x, int mode, char *out)
{
    unsigned char b[4];
    std::memcpy(b, &x, 4);  // little-endian host
    out[0] = (char)b[3];
    out[1] = (char)b[2];
    out[2] = (char)b[1];
    out[3] = (char)(mode == LBFFS_VTK_BINARY_REF ? b[3] : b[0]);  // base.hpp:40 repeats byte 3
}

}  // namespace

extern "C" {

int lbffs_out_write_vtk(const char *path, const lbffs_vtk_fields *F, int mode)
{
    if (!path || !F || !fields_complete(*F) || F->nx < 1 || F->ny < 1 || F->nz < 1) return LBFFS_E_INVALID;
    if (mode != LBFFS_VTK_ASCII && mode != LBFFS_VTK_BINARY && mode != LBFFS_VTK_BINARY_REF) return LBFFS_E_INVALID;
    const bool binary = mode != LBFFS_VTK_ASCII;
    std::ofstream out(path, binary ? (std::ios::out | std::ios::binary) : std::ios::out);
    if (!out) return LBFFS_E_INVALID;
    const long N = (long)F->nx * F->ny * F->nz;
    // write.hpp:109-125
    out << "# vtk DataFile Version 3.0\n"
        << "vtk output\n"
        << (binary ? "BINARY\n" : "ASCII\n")
        << "DATASET STRUCTURED_POINTS\n"
        << "DIMENSIONS " << F->nx << " " << F->ny << " " << F->nz << "\n"
        << "ORIGIN " << 0.0 << " " << 0.0 << " " << 0.0 << "\n"
        << "SPACING " << F->L << " " << F->L << " " << F->L << "\n\n"
        << "POINT_DATA " << F->nx * F->ny * F->nz << "\n";
    std::vector<char> buf;
    for (const VtkField &fd : kFields) {
        out << "SCALARS " << fd.name << " float\n"
            << "LOOKUP_TABLE default\n";
        if (binary) {
            buf.resize((size_t)N * 4);
            for (long ic = 0; ic < N; ic++) float_bytes(fd.value(*F, ic), mode, &buf[(size_t)ic * 4]);
            out.write(buf.data(), (std::streamsize)buf.size());
        } else if (fd.is_int) {
            for (long ic = 0; ic < N; ic++) out << fd.ivalue(*F, ic) << "\n";
        } else {
            for (long ic = 0; ic < N; ic++) out << fd.value(*F, ic) << "\n";
        }
        out << "\n";
    }
    out.close();
    return out ? 0 : LBFFS_E_INVALID;
}

int lbffs_out_write_restart(const char *path, int nt, int q, int nx, int ny, int nz, const float *f)
{
    if (!path || !f || (q != 19 && q != 27) || nx < 1 || ny < 1 || nz < 1) return LBFFS_E_INVALID;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return LBFFS_E_INVALID;
    const int32_t head[4] = {nt, nx, ny, nz};
    const size_t n = (size_t)q * nx * ny * nz;
    bool ok = std::fwrite(head, sizeof(int32_t), 4, fp) == 4 && std::fwrite(f, sizeof(float), n, fp) == n;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? 0 : LBFFS_E_INVALID;
}

int lbffs_out_read_restart(const char *path, int q, int *nt, int *nx, int *ny, int *nz, float *f, size_t f_count)
{
    if (!path || !nt || !nx || !ny || !nz || (q != 19 && q != 27)) return LBFFS_E_INVALID;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return LBFFS_E_INVALID;
    int32_t head[4];
    bool ok = std::fread(head, sizeof(int32_t), 4, fp) == 4;
    if (ok) {
        *nt = head[0];
        *nx = head[1];
        *ny = head[2];
        *nz = head[3];
        ok = head[1] > 0 && head[2] > 0 && head[3] > 0;
    }
    if (ok && f) {
        const size_t n = (size_t)q * (size_t)head[1] * (size_t)head[2] * (size_t)head[3];
        ok = n == f_count && std::fread(f, sizeof(float), n, fp) == n;
    }
    std::fclose(fp);
    return ok ? 0 : LBFFS_E_INVALID;
}

int lbffs_out_stats_compute(const lbffs_out_stats_in *in, lbffs_out_stats *out)
{
    if (!in || !out || !in->rho || !in->u || !in->solid || in->nx < 1 || in->ny < 1 || in->nz < 1)
        return LBFFS_E_INVALID;
    if (in->forceCoeffs && (!in->Fwx || !in->Fwy || !in->Fwz || !in->GxIBM || !in->GyIBM || !in->GzIBM))
        return LBFFS_E_INVALID;
    const long N = (long)in->nx * in->ny * in->nz;
    // write.hpp:16-47: rho - 1 summed in double, the forces in float, fluid cells in order
    double rho_av_tmp = 0.f;
    int eleFluid = 0;
    float sF[3] = {0.f, 0.f, 0.f}, sG[3] = {0.f, 0.f, 0.f};
    for (long ic = 0; ic < N; ic++) {
        if (in->solid[ic] != 0) continue;
        rho_av_tmp += in->rho[ic] - 1.0;
        if (in->forceCoeffs) {
            sF[0] += in->Fwx[ic];
            sF[1] += in->Fwy[ic];
            sF[2] += in->Fwz[ic];
            sG[0] += in->GxIBM[ic];
            sG[1] += in->GyIBM[ic];
            sG[2] += in->GzIBM[ic];
        }
        eleFluid++;
    }
    const float rho_av = rho_av_tmp / float(eleFluid) + 1.f;
    float rmax = in->rho[0], rmin = in->rho[0];
    for (long ic = 1; ic < N; ic++) {
        if (rmax < in->rho[ic]) rmax = in->rho[ic];  // std::max_element: first maximum
        if (in->rho[ic] < rmin) rmin = in->rho[ic];
    }
    // write.hpp:52-61: u*c over the i = 0 plane, divided by ny*nz
    float uAve = 0.f;
    for (int j = 0; j < in->ny; j++)
        for (int k = 0; k < in->nz; k++) uAve += in->u[(long)in->nx * (j + (long)in->ny * k)] * in->c;
    uAve /= in->ny * in->nz;
    out->rho_av = rho_av;
    out->rhoMax = rmax;
    out->rhoMin = rmin;
    out->uAve = uAve;
    out->eleFluid = eleFluid;
    for (int a = 0; a < 3; a++) {
        out->sumFw[a] = sF[a];
        out->sumG[a] = sG[a];
    }
    // write.hpp:93-99: float numerator, double denominator
    const float c = in->c, L = in->L;
    const double den = 0.5 * rho_av * in->uMax * in->uMax * in->Dref * L;
    out->Cx = sF[0] * c * c * L * L / den;
    out->Cy = sF[1] * c * c * L * L / den;
    out->Cz = sF[2] * c * c * L * L / den;
    out->CxIBM = -sG[0] * c * c * L * L / den;
    out->CyIBM = -sG[1] * c * c * L * L / den;
    out->CzIBM = -sG[2] * c * c * L * L / den;
    return 0;
}

int lbffs_out_coeffs_fp64(const double sums[8], float c, float L, float uMax, float Dref, double out[7])
{
    if (!sums || !out || !(sums[7] > 0)) return LBFFS_E_INVALID;
    const double rho_av = sums[6] / sums[7] + 1.0;
    const double scale = (double)c * c * L * L / (0.5 * rho_av * uMax * uMax * Dref * L);
    out[0] = rho_av;
    for (int a = 0; a < 3; a++) {
        out[1 + a] = sums[a] * scale;
        out[4 + a] = -sums[3 + a] * scale;
    }
    return 0;
}

int lbffs_out_accumulate_mean(long n, int interval, const float *u, const float *v, const float *w, float *uMean,
                              float *vMean, float *wMean, float *timeMean)
{
    if (n < 0 || !u || !v || !w || !uMean || !vMean || !wMean || !timeMean) return LBFFS_E_INVALID;
    *timeMean += interval;
    for (long ic = 0; ic < n; ic++) {
        uMean[ic] += u[ic] * interval;
        vMean[ic] += v[ic] * interval;
        wMean[ic] += w[ic] * interval;
    }
    return 0;
}

}  // extern "C"
