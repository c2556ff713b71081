// Exercises every host entry point of lbffs_case.cpp and lbffs_output.cpp (valid inputs,
// edge sizes and the rejected ones) so that a build with AddressSanitizer and
// UndefinedBehaviorSanitizer checks them; test_host_sanitizers.py builds and runs it.
// Exits 0 when every call returned what it should; the sanitizers abort on a finding.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lbffs_case.h"
#include "lbffs_output.h"

static int failures = 0;
#define EXPECT(cond)                                                                 \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #cond); \
            failures++;                                                              \
        }                                                                            \
    } while (0)

static lbffs_case_in make_case(int nx, int ny, int nz, int code_set)
{
    lbffs_case_in in;
    std::memset(&in, 0, sizeof(in));
    in.nx = nx;
    in.ny = ny;
    in.nz = nz;
    in.Lx = 1.0f;
    in.uMax = 1.0f;
    in.rho0 = 1.0f;
    in.U0 = 0.1f;
    in.nu = 1e-3f;
    in.dpdx = 1e-6f;
    in.uIni = 0.01f;
    // every BC code on some face: BounceBack, FixedVelocity, Outlet, FixedDensity,
    // Symmetry, WallFunction and Cyclic pairs
    static const int sets[2][6] = {{1, 3, 1, 5, 0, 0}, {5, 6, 4, 7, 1, 1}};
    for (int f = 0; f < 6; f++) {
        for (int d = 0; d < 3; d++) in.face_code[f][d] = sets[code_set][f];
        in.face_vel[f][0] = 0.05f * (float)(f + 1);
        in.face_vel[f][1] = -0.01f * (float)f;
    }
    return in;
}

static void case_calls(int nx, int ny, int nz, int code_set)
{
    const size_t N = (size_t)nx * ny * nz;
    lbffs_case_in in = make_case(nx, ny, nz, code_set);
    lbffs_units un;
    EXPECT(lbffs_case_units(&in, &un) == 0);
    std::vector<int32_t> bnd(N), b1(N), b2(N), b3(N);
    std::vector<float> u0(N), v0(N), w0(N);
    EXPECT(lbffs_case_faces(&in, &un, bnd.data(), b1.data(), b2.data(), b3.data(), u0.data(), v0.data(), w0.data()) == 0);
    for (int q : {19, 27}) {
        std::vector<float> f((size_t)q * N);
        EXPECT(lbffs_case_init_feq(q, nx, ny, nz, 1.0f, 0.02f, -0.01f, 0.005f, f.data()) == 0);
        std::vector<uint8_t> solid(N), nei(N);
        for (size_t ic = 0; ic < N; ic++) solid[ic] = (uint8_t)((ic * 2654435761u >> 7) % 5 == 0);
        EXPECT(lbffs_case_neisolid(q, nx, ny, nz, solid.data(), nei.data()) == 0);
    }
    std::vector<float> omega(N, 1.6f);
    // a sponge around every corner cell of the x faces, and a few rejected centres
    for (int i : {0, nx - 1})
        for (int j : {0, ny - 1})
            for (int k : {0, nz - 1})
                EXPECT(lbffs_case_outlet_sponge(nx, ny, nz, i, j, k, 3, 1, 1, 4, 1.6f, omega.data()) == 0);
    EXPECT(lbffs_case_outlet_sponge(nx, ny, nz, nx, 0, 0, 3, 1, 1, 4, 1.6f, omega.data()) == -1);
    EXPECT(lbffs_case_outlet_sponge(nx, ny, nz, 0, -1, 0, 3, 1, 1, 4, 1.6f, omega.data()) == -1);
    EXPECT(lbffs_case_outlet_sponge(nx, ny, nz, 0, 0, nz, 3, 1, 1, 4, 1.6f, omega.data()) == -1);
    std::vector<float> sdf(N, 10000.0f);
    EXPECT(lbffs_case_sdf_near_bwalls(2.5f, b1.data(), b2.data(), b3.data(), nx, ny, nz, sdf.data()) == 0);
    EXPECT(lbffs_case_sdf_near_bwalls(50.0f, b1.data(), b2.data(), b3.data(), nx, ny, nz, sdf.data()) == 0);

    // findPoints: triangles on every face, one off the lattice, one degenerate
    const int n = 8;
    std::vector<float> v0t(3 * n), v1t(3 * n), v2t(3 * n), nrm(3 * n), cen(3 * n);
    const float X = (float)(nx - 1), Y = (float)(ny - 1), Z = (float)(nz - 1);
    const float tri[n][10] = {
        // plane axis, plane coordinate, normal sign, in-plane triangle (a0,b0,a1,b1,a2,b2), unused
        {0, 0, -1, -1, -1, Y + 1, -1, -1, Z + 1, 0},  {0, X, 1, 0, 0, Y, 0, 0, Z, 0},
        {1, 0, -1, 0, 0, X, 0, 0, Z, 0},               {1, Y, 1, -2, -2, X + 2, -2, -2, Z + 2, 0},
        {2, 0, -1, 0, 0, X, 0, 0, Y, 0},               {2, Z, 1, 1, 1, X - 1, 1, 1, Y - 1, 0},
        {0, X + 7, 1, 0, 0, Y, 0, 0, Z, 0},            {2, 0, -1, 2, 2, 2, 2, 2, 2, 0}};
    for (int t = 0; t < n; t++) {
        const int ax = (int)tri[t][0], a = (ax + 1) % 3, b = (ax + 2) % 3;
        float p[3][3];
        for (int v = 0; v < 3; v++) {
            p[v][ax] = tri[t][1];
            p[v][a] = tri[t][3 + 2 * v];
            p[v][b] = tri[t][4 + 2 * v];
        }
        for (int d = 0; d < 3; d++) {
            v0t[d * n + t] = p[0][d];
            v1t[d * n + t] = p[1][d];
            v2t[d * n + t] = p[2][d];
            nrm[d * n + t] = d == ax ? tri[t][2] : 0.0f;
            cen[d * n + t] = (p[0][d] + p[1][d] + p[2][d]) / 3.0f;
        }
    }
    long count = -1;
    EXPECT(lbffs_case_find_points(n, v0t.data(), v1t.data(), v2t.data(), nrm.data(), cen.data(), nx, ny, nz, nullptr, 0,
                                  &count) == 0);
    EXPECT(count > 0);
    std::vector<int32_t> pts((size_t)(count > 0 ? count : 1));
    long count2 = -1;
    EXPECT(lbffs_case_find_points(n, v0t.data(), v1t.data(), v2t.data(), nrm.data(), cen.data(), nx, ny, nz, pts.data(),
                                  count, &count2) == 0);
    EXPECT(count2 == count);
    for (long p = 0; p < count2; p++) EXPECT(pts[p] >= 0 && (size_t)pts[p] < N);
    // a buffer shorter than the count: written up to its capacity only
    if (count > 2) {
        std::vector<int32_t> few(2);
        EXPECT(lbffs_case_find_points(n, v0t.data(), v1t.data(), v2t.data(), nrm.data(), cen.data(), nx, ny, nz,
                                      few.data(), 2, &count2) == 0);
        EXPECT(count2 == count);
    }
}

static void output_calls(const std::string &dir, int nx, int ny, int nz)
{
    const size_t N = (size_t)nx * ny * nz;
    std::vector<float> u(N), v(N), w(N), rho(N), tau(N), sdf(N, 10000.0f), g(N), om(N, 1.7f);
    std::vector<float> um(N), vm(N), wm(N);
    std::vector<int32_t> b(N, 1);
    std::vector<uint8_t> solid(N), nei(N);
    for (size_t ic = 0; ic < N; ic++) {
        u[ic] = 0.01f * std::sin(0.1f * (float)ic);
        v[ic] = -0.5f * u[ic];
        w[ic] = 0.25f * u[ic];
        rho[ic] = 1.0f + 1e-3f * u[ic];
        tau[ic] = 1e-4f;
        g[ic] = 1e-6f * (float)(ic % 7);
        solid[ic] = (uint8_t)(ic % 11 == 0);
    }
    float timeMean = 0.0f;
    for (int step = 0; step < 4; step++)
        EXPECT(lbffs_out_accumulate_mean((long)N, 10, u.data(), v.data(), w.data(), um.data(), vm.data(), wm.data(),
                                         &timeMean) == 0);
    lbffs_vtk_fields F;
    std::memset(&F, 0, sizeof(F));
    F.nx = nx;
    F.ny = ny;
    F.nz = nz;
    F.L = 0.01f;
    F.c = 10.0f;
    F.timeMean = timeMean;
    F.u = u.data();
    F.v = v.data();
    F.w = w.data();
    F.rho = rho.data();
    F.tauSGS = tau.data();
    F.boundary1 = F.boundary2 = F.boundary3 = b.data();
    F.sdf = sdf.data();
    F.solid = solid.data();
    F.neiSolid = nei.data();
    F.GxIBM = F.GyIBM = F.GzIBM = g.data();
    F.omega = om.data();
    F.uMean = um.data();
    F.vMean = vm.data();
    F.wMean = wm.data();
    for (int mode : {LBFFS_VTK_ASCII, LBFFS_VTK_BINARY, LBFFS_VTK_BINARY_REF})
        EXPECT(lbffs_out_write_vtk((dir + "/out" + std::to_string(mode) + ".vtk").c_str(), &F, mode) == 0);
    EXPECT(lbffs_out_write_vtk((dir + "/no/such/dir/x.vtk").c_str(), &F, LBFFS_VTK_ASCII) != 0);

    for (int q : {19, 27}) {
        std::vector<float> f((size_t)q * N), back((size_t)q * N);
        for (size_t i = 0; i < f.size(); i++) f[i] = 1e-3f * (float)(i % 97);
        const std::string path = dir + "/restart" + std::to_string(q) + ".bin";
        EXPECT(lbffs_out_write_restart(path.c_str(), 123, q, nx, ny, nz, f.data()) == 0);
        int nt = 0, rx = 0, ry = 0, rz = 0;
        EXPECT(lbffs_out_read_restart(path.c_str(), q, &nt, &rx, &ry, &rz, back.data(), back.size()) == 0);
        EXPECT(nt == 123 && rx == nx && ry == ny && rz == nz && back == f);
        // a buffer too small for the file, the other lattice, a missing file: refused
        EXPECT(lbffs_out_read_restart(path.c_str(), q, &nt, &rx, &ry, &rz, back.data(), back.size() - 1) != 0);
        EXPECT(lbffs_out_read_restart(path.c_str(), q == 19 ? 27 : 19, &nt, &rx, &ry, &rz, back.data(), back.size()) != 0);
        EXPECT(lbffs_out_read_restart((dir + "/missing.bin").c_str(), q, &nt, &rx, &ry, &rz, back.data(),
                                      back.size()) != 0);
    }
    // a truncated restart file
    {
        const std::string path = dir + "/restart19.bin", cut = dir + "/cut.bin";
        FILE *in = std::fopen(path.c_str(), "rb");
        FILE *out = std::fopen(cut.c_str(), "wb");
        EXPECT(in && out);
        if (in && out) {
            std::vector<char> buf(64);
            const size_t got = std::fread(buf.data(), 1, buf.size(), in);
            std::fwrite(buf.data(), 1, got, out);
        }
        if (in) std::fclose(in);
        if (out) std::fclose(out);
        std::vector<float> back((size_t)19 * N);
        int nt = 0, rx = 0, ry = 0, rz = 0;
        EXPECT(lbffs_out_read_restart(cut.c_str(), 19, &nt, &rx, &ry, &rz, back.data(), back.size()) != 0);
    }

    lbffs_out_stats_in si;
    std::memset(&si, 0, sizeof(si));
    si.nx = nx;
    si.ny = ny;
    si.nz = nz;
    si.c = 10.0f;
    si.L = 0.01f;
    si.uMax = 1.0f;
    si.Dref = 0.1f;
    si.rho = rho.data();
    si.u = u.data();
    si.solid = solid.data();
    si.GxIBM = si.GyIBM = si.GzIBM = g.data();
    lbffs_out_stats st;
    EXPECT(lbffs_out_stats_compute(&si, &st) == 0);
    si.forceCoeffs = 1;
    si.Fwx = si.Fwy = si.Fwz = g.data();
    EXPECT(lbffs_out_stats_compute(&si, &st) == 0);
    const double sums[8] = {1e-3, -2e-4, 3e-5, 1.0, 0.5, 0.25, 0.125, 1.0};
    double co[7];
    EXPECT(lbffs_out_coeffs_fp64(sums, 10.0f, 0.01f, 1.0f, 0.1f, co) == 0);
    for (double x : co) EXPECT(std::isfinite(x));
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    // 2-D and 3-D lattices, the smallest ones each face rule allows, uneven sizes
    const int sizes[][3] = {{3, 3, 1}, {4, 3, 3}, {9, 7, 5}, {16, 12, 10}};
    for (const auto &s : sizes)
        for (int set = 0; set < 2; set++) case_calls(s[0], s[1], s[2], set);
    for (const auto &s : sizes) output_calls(dir, s[0], s[1], s[2]);
    std::printf("host sanitize driver: %d failed expectations\n", failures);
    return failures == 0 ? 0 : 1;
}
