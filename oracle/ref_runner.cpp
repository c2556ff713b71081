// ref_runner.cpp -- runs the reference's own k_streamingCollision (D3Q19_opencl/kernel.cl and
// D3Q27_opencl/kernel.cl, built by `make -C oracle ref` into oracle/_ref/kernel19.co and
// kernel27.co) on the GPU with the reference host loop: launch over nx*ny*nz work-items,
// wait, swap f/fTmp (D3Q19 ns.cpp:509-603, D3Q27 ns.cpp:540-640, without output and moving
// walls).
//
// TEST INFRASTRUCTURE ONLY: the checker that pins the CPU oracle and the HIP product to
// the reference kernel on the same inputs.  The reference kernels have no bounds guard
// (K19:3548, K27:2737) and a few branches that index outside the lattice; the runner
// refuses inputs that would reach them and launches exactly N work-items.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {
std::string g_err;

int err(const std::string &m)
{
    g_err = m;
    return -1;
}

#define H(x)                                                                          \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) return err(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Buf {
    void *p = nullptr;
    ~Buf()
    {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

extern "C" {

const char *refrun_error(void) { return g_err.c_str(); }

namespace {
// Arrays are host arrays in the reference layout.  In/out arrays are overwritten with
// the state after nsteps: f (Q*N, current populations after the swaps), Fw*, tauSGS,
// rho/u/v/w, G*.  Q = 27 appends the limiter clim to the kernel's arguments (K27:2731).
int run_steps(int Q, const char *co_path, int nx, int ny, int nz, int nsteps, int block, float *f,
              const int32_t *boundary, const int32_t *b1, const int32_t *b2, const int32_t *b3, const float *sdf,
              const uint8_t *solid, const uint8_t *neiSolid, const float *u0, const float *v0, const float *w0,
              float *Fwx, float *Fwy, float *Fwz, float *tauSGS, float *rho, float *u, float *v, float *w,
              const float *omega, float *Gx, float *Gy, float *Gz, float dpdx, float rho_av, float LES,
              int isReadMovingWalls, float omegaB, float spzWidth, float clim)
{
    const long N = (long)nx * ny * nz;
    if (nx < 2 || ny < 2 || nz < 1 || N <= 0 || (long)Q * N >= (1L << 31)) return err("lattice size");
    // the kernels are built without a work-group size attribute: at most 256 work-items each
    if (block <= 0 || block > 256 || N % block != 0) return err("work-group size above 256, or not dividing N");
    // inputs that would send the reference kernel outside its buffers
    for (long ic = 0; ic < N; ic++) {
        const int i = (int)(ic % ((long)nx * ny) % nx), j = (int)(ic % ((long)nx * ny) / nx),
                  k = (int)(ic / ((long)nx * ny));
        if (solid[ic]) continue;
        const bool face = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || (nz > 1 && (k == 0 || k == nz - 1)));
        if (face && boundary[ic] != 1) return err("fluid face cell without boundary flag");
        if (boundary[ic] == 1 && neiSolid[ic]) return err("neiSolid on a face cell");
        if (boundary[ic] == 1 && nz > 1 && k == nz - 1 && b3[ic] == 7) return err("WallFunction on nzMax");
        if (boundary[ic] == 1 && nz == 1 && (b3[ic] == 3 || b3[ic] == 6 || (b3[ic] == 5 && w0[ic] < 0.f)))
            return err("2-D lattice with a z-face condition sampling k=1");
    }
    hipModule_t mod;
    H(hipModuleLoad(&mod, co_path));
    hipFunction_t fn;
    hipError_t ge = hipModuleGetFunction(&fn, mod, "k_streamingCollision");
    if (ge != hipSuccess) {
        (void)hipModuleUnload(mod);
        return err(std::string("hipModuleGetFunction: ") + hipGetErrorString(ge));
    }
    const size_t nf = (size_t)N * sizeof(float), ni = (size_t)N * sizeof(int32_t);
    Buf F, FT, B, B1, B2, B3, SDF, SOL, NEI, U0, V0, W0, FX, FY, FZ, TAU, RHO, U, V, W, OM, GX, GY, GZ;
    struct {
        Buf *b;
        const void *src;
        size_t bytes;
    } up[] = {{&F, f, Q * nf},      {&FT, f, Q * nf},   {&B, boundary, ni}, {&B1, b1, ni},     {&B2, b2, ni},
              {&B3, b3, ni},        {&SDF, sdf, nf},     {&SOL, solid, (size_t)N}, {&NEI, neiSolid, (size_t)N},
              {&U0, u0, nf},        {&V0, v0, nf},       {&W0, w0, nf},      {&FX, Fwx, nf},    {&FY, Fwy, nf},
              {&FZ, Fwz, nf},       {&TAU, tauSGS, nf},  {&RHO, rho, nf},    {&U, u, nf},       {&V, v, nf},
              {&W, w, nf},          {&OM, omega, nf},    {&GX, Gx, nf},      {&GY, Gy, nf},     {&GZ, Gz, nf}};
    int rc = 0;
    for (auto &x : up) {
        if (hipMalloc(&x.b->p, x.bytes) != hipSuccess) {
            rc = err("hipMalloc");
            break;
        }
        if (hipMemcpy(x.b->p, x.src, x.bytes, hipMemcpyHostToDevice) != hipSuccess) {
            rc = err("hipMemcpy H2D");
            break;
        }
    }
    if (rc == 0) {
        unsigned elements = (unsigned)N;
        void *fsrc = F.p, *fdst = FT.p;
        for (int s = 0; s < nsteps && rc == 0; s++) {
            void *params[] = {&fsrc, &fdst, &B.p, &B1.p, &B2.p, &B3.p, &SDF.p, &SOL.p, &NEI.p, &U0.p, &V0.p, &W0.p,
                              &FX.p, &FY.p, &FZ.p, &TAU.p, &RHO.p, &U.p, &V.p, &W.p, &OM.p, &GX.p, &GY.p, &GZ.p,
                              &elements, &dpdx, &rho_av, &nx, &ny, &nz, &LES, &isReadMovingWalls, &omegaB,
                              &spzWidth, &clim};
            const int nargs = (int)(sizeof(params) / sizeof(params[0])) - (Q == 27 ? 0 : 1);
            std::vector<void *> args(params, params + nargs);
            hipError_t e = hipModuleLaunchKernel(fn, (unsigned)(N / block), 1, 1, (unsigned)block, 1, 1, 0, nullptr,
                                                 args.data(), nullptr);
            if (e == hipSuccess) e = hipDeviceSynchronize();  // queue.finish(), ns.cpp:537
            if (e != hipSuccess) rc = err(std::string("launch: ") + hipGetErrorString(e));
            std::swap(fsrc, fdst);  // ns.cpp:602
        }
        struct {
            void *dst;
            const void *src;
            size_t bytes;
        } down[] = {{f, fsrc, Q * nf}, {Fwx, FX.p, nf}, {Fwy, FY.p, nf}, {Fwz, FZ.p, nf}, {tauSGS, TAU.p, nf},
                    {rho, RHO.p, nf},   {u, U.p, nf},    {v, V.p, nf},    {w, W.p, nf},    {Gx, GX.p, nf},
                    {Gy, GY.p, nf},     {Gz, GZ.p, nf}};
        for (auto &x : down)
            if (rc == 0 && hipMemcpy(x.dst, x.src, x.bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = err("D2H");
    }
    (void)hipModuleUnload(mod);
    return rc;
}
}  // namespace

int refrun_step19(const char *co_path, int nx, int ny, int nz, int nsteps, int block, float *f,
                  const int32_t *boundary, const int32_t *b1, const int32_t *b2, const int32_t *b3, const float *sdf,
                  const uint8_t *solid, const uint8_t *neiSolid, const float *u0, const float *v0, const float *w0,
                  float *Fwx, float *Fwy, float *Fwz, float *tauSGS, float *rho, float *u, float *v, float *w,
                  const float *omega, float *Gx, float *Gy, float *Gz, float dpdx, float rho_av, float LES,
                  int isReadMovingWalls, float omegaB, float spzWidth)
{
    return run_steps(19, co_path, nx, ny, nz, nsteps, block, f, boundary, b1, b2, b3, sdf, solid, neiSolid, u0, v0,
                     w0, Fwx, Fwy, Fwz, tauSGS, rho, u, v, w, omega, Gx, Gy, Gz, dpdx, rho_av, LES,
                     isReadMovingWalls, omegaB, spzWidth, 0.f);
}

int refrun_step27(const char *co_path, int nx, int ny, int nz, int nsteps, int block, float *f,
                  const int32_t *boundary, const int32_t *b1, const int32_t *b2, const int32_t *b3, const float *sdf,
                  const uint8_t *solid, const uint8_t *neiSolid, const float *u0, const float *v0, const float *w0,
                  float *Fwx, float *Fwy, float *Fwz, float *tauSGS, float *rho, float *u, float *v, float *w,
                  const float *omega, float *Gx, float *Gy, float *Gz, float dpdx, float rho_av, float LES,
                  int isReadMovingWalls, float omegaB, float spzWidth, float clim)
{
    return run_steps(27, co_path, nx, ny, nz, nsteps, block, f, boundary, b1, b2, b3, sdf, solid, neiSolid, u0, v0,
                     w0, Fwx, Fwy, Fwz, tauSGS, rho, u, v, w, omega, Gx, Gy, Gz, dpdx, rho_av, LES,
                     isReadMovingWalls, omegaB, spzWidth, clim);
}

// One pass of the reference's moving-wall kernels (K19:3648-3807 / K27:2828-2986) on given
// inputs: k_Gwall (marker forcing -> Gm, 3n as GxMovingWall, GyMovingWall, GzMovingWall),
// k_Gibm (its spread goes to a scratch lattice that is not returned -- it writes the value
// before the atomic add back over the sum, K19:3760-3762 -- and its marker motion updates
// xyz), and k_Force on the caller's G (Gin, 3N) adding to f (Q*N).  k_Force declares its
// solid argument float* (K19:3784): it is given the flags as floats.
int refrun_ibm(const char *co_path, int Q, int nx, int ny, int nz, int n, float *xyz, const float *u, const float *v,
               const float *w, const float *rho, const float *motion, float *Gm, const float *Gin, float *f,
               const float *solid_f)
{
    const long N = (long)nx * ny * nz;
    if (n < 1 || N <= 0 || (long)Q * N >= (1L << 31)) return err("sizes");
    auto divisor = [](long m) {
        for (int b : {256, 128, 64, 32, 16, 8, 4, 2, 1})
            if (m % b == 0) return b;
        return 1;
    };
    hipModule_t mod;
    H(hipModuleLoad(&mod, co_path));
    hipFunction_t fg, fi, ff;
    if (hipModuleGetFunction(&fg, mod, "k_Gwall") != hipSuccess || hipModuleGetFunction(&fi, mod, "k_Gibm") != hipSuccess ||
        hipModuleGetFunction(&ff, mod, "k_Force") != hipSuccess) {
        (void)hipModuleUnload(mod);
        return err("k_Gwall/k_Gibm/k_Force not found");
    }
    const size_t nf = (size_t)N * sizeof(float), m3 = (size_t)3 * n * sizeof(float);
    Buf RHO, U, V, W, XYZ, GMX, GMY, GMZ, SGX, SGY, SGZ, GX, GY, GZ, F, SOL;
    struct {
        Buf *b;
        const void *src;
        size_t bytes;
    } up[] = {{&RHO, rho, nf}, {&U, u, nf}, {&V, v, nf}, {&W, w, nf}, {&XYZ, xyz, m3},
              {&GMX, nullptr, m3 / 3}, {&GMY, nullptr, m3 / 3}, {&GMZ, nullptr, m3 / 3},
              {&SGX, nullptr, nf}, {&SGY, nullptr, nf}, {&SGZ, nullptr, nf},
              {&GX, Gin, nf}, {&GY, Gin + N, nf}, {&GZ, Gin + 2 * N, nf}, {&F, f, (size_t)Q * nf}, {&SOL, solid_f, nf}};
    int rc = 0;
    for (auto &x : up) {
        if (hipMalloc(&x.b->p, x.bytes) != hipSuccess) {
            rc = err("hipMalloc");
            break;
        }
        hipError_t e = x.src ? hipMemcpy(x.b->p, x.src, x.bytes, hipMemcpyHostToDevice) : hipMemset(x.b->p, 0, x.bytes);
        if (e != hipSuccess) {
            rc = err("upload");
            break;
        }
    }
    if (rc == 0) {
        int nM = n;
        float ut = motion[0], vt = motion[1], wt = motion[2], rx = motion[3], ry = motion[4], rz = motion[5],
              ax = motion[6], ay = motion[7], az = motion[8], om = motion[9];
        unsigned elements = (unsigned)N;
        void *pg[] = {&RHO.p, &U.p, &V.p, &W.p, &XYZ.p, &GMX.p, &GMY.p, &GMZ.p, &nM, &nx, &ny, &nz,
                      &ut, &vt, &wt, &rx, &ry, &rz, &ax, &ay, &az, &om};
        void *pi[] = {&RHO.p, &SGX.p, &SGY.p, &SGZ.p, &XYZ.p, &GMX.p, &GMY.p, &GMZ.p, &nM, &nx, &ny, &nz,
                      &ut, &vt, &wt, &rx, &ry, &rz, &ax, &ay, &az, &om};
        void *pf[] = {&F.p, &SOL.p, &RHO.p, &GX.p, &GY.p, &GZ.p, &elements};
        const int bm = divisor(n), bn = divisor(N);
        hipError_t e = hipModuleLaunchKernel(fg, (unsigned)(n / bm), 1, 1, bm, 1, 1, 0, nullptr, pg, nullptr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipModuleLaunchKernel(fi, (unsigned)(n / bm), 1, 1, bm, 1, 1, 0, nullptr, pi, nullptr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipModuleLaunchKernel(ff, (unsigned)(N / bn), 1, 1, bn, 1, 1, 0, nullptr, pf, nullptr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = err(std::string("launch: ") + hipGetErrorString(e));
        if (rc == 0 && (hipMemcpy(xyz, XYZ.p, m3, hipMemcpyDeviceToHost) != hipSuccess ||
                        hipMemcpy(Gm, GMX.p, m3 / 3, hipMemcpyDeviceToHost) != hipSuccess ||
                        hipMemcpy(Gm + n, GMY.p, m3 / 3, hipMemcpyDeviceToHost) != hipSuccess ||
                        hipMemcpy(Gm + 2 * n, GMZ.p, m3 / 3, hipMemcpyDeviceToHost) != hipSuccess ||
                        hipMemcpy(f, F.p, (size_t)Q * nf, hipMemcpyDeviceToHost) != hipSuccess))
            rc = err("D2H");
    }
    (void)hipModuleUnload(mod);
    return rc;
}

}  // extern "C"
