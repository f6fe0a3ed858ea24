// A C++ caller of libidefix_amd.so built with g++ against include/idefix_amd.h only, doing
// what a Fluid<Phys> binding does (context from the problem description, device arrays
// owned by the caller, TimeIntegrator::Cycle through idfx_cycle, read back).  The GPU test
// tests/test_gpu_consumer.py runs it on the 2-D Orszag-Tang configuration and compares the
// result with the oracle.
//
// usage: idfx_consumer NX NY NCYCLES in.bin out.bin
//   in.bin : Vc [6][NY+4][NX+4] then Vs [2][NY+5][NX+5] (doubles, LayoutRight)
//   out.bin: t, dt, then Vc and Vs after NCYCLES cycles
// Configuration: 2-D MHD (two components), PLM, HLLD, UCT-HLLD, RK2, CFL 0.6,
// CFL_max_var 1.1, first_dt 1e-4, gamma 5/3, periodic [0,1]^2.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "idefix_amd.h"

#define CHECK(x)                                                          \
  do {                                                                    \
    int s_ = (x);                                                         \
    if (s_ != IDFX_OK) {                                                  \
      std::fprintf(stderr, "%s failed: %s\n", #x, idfx_strerror(s_));     \
      return 1;                                                           \
    }                                                                     \
  } while (0)
#define HCHECK(x)                                                         \
  do {                                                                    \
    if ((x) != hipSuccess) {                                              \
      std::fprintf(stderr, "%s failed\n", #x);                            \
      return 1;                                                           \
    }                                                                     \
  } while (0)

int main(int argc, char **argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s NX NY NCYCLES in.bin out.bin\n", argv[0]);
    return 2;
  }
  const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]);
  const long ncycles = std::atol(argv[3]);
  idfx_desc d{};
  d.dim = 2;
  d.comps = 2;
  d.mhd = 1;
  d.nx[0] = nx;
  d.nx[1] = ny;
  d.nx[2] = 1;
  d.recon = IDFX_RECON_PLM;
  d.solver = IDFX_SOLVER_HLLD;
  d.emf = IDFX_EMF_UCT_HLLD;
  d.precision = IDFX_PREC_FP64;
  for (int q = 0; q < 3; q++) d.lbound[q] = d.rbound[q] = IDFX_BC_PERIODIC;
  d.gamma = 5.0 / 3.0;
  // cell widths of the uniform grid (gridHost.cpp:79-86), ghosts included
  const int ng = 2;
  std::vector<double> dx1(nx + 2 * ng, 1.0 / nx), dx2(ny + 2 * ng, 1.0 / ny), dx3(1, 1.0);
  idfx_ctx *ctx = nullptr;
  CHECK(idfx_ctx_create(&d, dx1.data(), dx2.data(), dx3.data(), &ctx));
  int np_tot[3], nghost[3], nvar, vs[3];
  CHECK(idfx_ctx_layout(ctx, np_tot, nghost, &nvar, vs));
  const size_t n3 = (size_t)np_tot[0] * np_tot[1] * np_tot[2];
  const size_t nvc = n3 * nvar, nvs = (size_t)2 * vs[0] * vs[1] * vs[2];
  std::vector<double> hVc(nvc), hVs(nvs);
  FILE *f = std::fopen(argv[4], "rb");
  if (!f || std::fread(hVc.data(), sizeof(double), nvc, f) != nvc ||
      std::fread(hVs.data(), sizeof(double), nvs, f) != nvs) {
    std::fprintf(stderr, "cannot read %s\n", argv[4]);
    return 1;
  }
  std::fclose(f);
  double *Vc, *Vs, *Uc, *InvDt;
  HCHECK(hipMalloc(&Vc, nvc * sizeof(double)));
  HCHECK(hipMalloc(&Vs, nvs * sizeof(double)));
  HCHECK(hipMalloc(&Uc, nvc * sizeof(double)));
  HCHECK(hipMalloc(&InvDt, n3 * sizeof(double)));
  HCHECK(hipMemcpy(Vc, hVc.data(), nvc * sizeof(double), hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(Vs, hVs.data(), nvs * sizeof(double), hipMemcpyHostToDevice));
  hipStream_t stream;
  HCHECK(hipStreamCreate(&stream));
  double t = 0.0, dt = 1e-4;
  for (long n = 0; n < ncycles; n++)
    CHECK(idfx_cycle(ctx, 2, 0.6, 1.1, n, &t, &dt, Vc, Vs, Uc, InvDt, stream));
  HCHECK(hipStreamSynchronize(stream));
  HCHECK(hipMemcpy(hVc.data(), Vc, nvc * sizeof(double), hipMemcpyDeviceToHost));
  HCHECK(hipMemcpy(hVs.data(), Vs, nvs * sizeof(double), hipMemcpyDeviceToHost));
  double divb = -1.0;
  CHECK(idfx_check_divb(ctx, Vs, &divb, stream));
  f = std::fopen(argv[5], "wb");
  if (!f) return 1;
  std::fwrite(&t, sizeof(double), 1, f);
  std::fwrite(&dt, sizeof(double), 1, f);
  std::fwrite(hVc.data(), sizeof(double), nvc, f);
  std::fwrite(hVs.data(), sizeof(double), nvs, f);
  std::fclose(f);
  std::printf("consumer: %ld cycles, t = %.17g, dt = %.17g, divB = %.3e\n", ncycles, t, dt, divb);
  CHECK(idfx_ctx_destroy(ctx));
  HCHECK(hipFree(Vc));
  HCHECK(hipFree(Vs));
  HCHECK(hipFree(Uc));
  HCHECK(hipFree(InvDt));
  HCHECK(hipStreamDestroy(stream));
  return 0;
}
