// idfx_compute_timestep_noalloc: DataBlock::ComputeTimestep (dataBlock.cpp:333-364) for
// callers that cannot allocate device or pinned memory.  Its reduction scratch is the
// file-scope device array `scratch` below (kNoAllocSlots per-block partial maxima and one
// result word, about 32 KiB, held by each device for the life of the process).  Host code
// never passes the array's name to a kernel: it takes the array's device address from
// hipGetSymbolAddress, once per device, and uses that address everywhere.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../include/idefix_amd.h"
#include "launch.hpp"

namespace idfx {
const GridDesc &ctx_grid(const idfx_ctx *c);
constexpr int kNoAllocSlots = 4096;
constexpr long long kNoAllocCellsPerBlock = 65536;  // at most 4096 * 65536 = 2^28 cells
constexpr int kNoAllocMaxDevices = 64;
}  // namespace idfx

static __device__ __align__(256) unsigned long long scratch[idfx::kNoAllocSlots + 1];

namespace idfx {
namespace {
std::mutex g_scratch_mu;
unsigned long long *g_scratch_ptr[kNoAllocMaxDevices] = {};

// device address of `scratch` on the current device, looked up on the first call there
int scratch_address(unsigned long long **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return IDFX_E_LAUNCH;
  if (dev < 0 || dev >= kNoAllocMaxDevices) return IDFX_E_ARGS;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return IDFX_E_LAUNCH;
    g_scratch_ptr[dev] = static_cast<unsigned long long *>(p);
  }
  *out = g_scratch_ptr[dev];
  return IDFX_OK;
}

__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long v) {
  __shared__ unsigned long long red[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long m = red[0];
#pragma unroll
  for (int q = 1; q < 4; q++) m = red[q] > m ? red[q] : m;
  return m;
}

// Block b reduces interior cells [b per, (b + 1) per) (i fastest) to the largest InvDt bit
// pattern (InvDt >= 0, so the bit order is the value order) and writes its slot, 0 for an
// empty range: no slot keeps a value from an earlier call.
__global__ __launch_bounds__(256) void k_invdt_partial(GridDesc g, const double *InvDt,
                                                       unsigned long long *part, long long per) {
  const int ni = g.end[0] - g.beg[0], nj = g.end[1] - g.beg[1], nk = g.end[2] - g.beg[2];
  const long long n = (long long)ni * nj * nk;
  const long long t0 = (long long)blockIdx.x * per;
  const long long t1 = t0 + per < n ? t0 + per : n;
  unsigned long long m = 0;
  for (long long t = t0 + threadIdx.x; t < t1; t += 256) {
    const int i = g.beg[0] + (int)(t % ni);
    const int j = g.beg[1] + (int)((t / ni) % nj);
    const int k = g.beg[2] + (int)(t / ((long long)ni * nj));
    const unsigned long long b = __double_as_longlong(InvDt[idx3(g, k, j, i)]);
    m = b > m ? b : m;
  }
  m = block_max_u64(m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ __launch_bounds__(256) void k_invdt_final(const unsigned long long *part, int nb,
                                                     unsigned long long *out) {
  unsigned long long m = 0;
  for (int q = threadIdx.x; q < nb; q += 256) m = part[q] > m ? part[q] : m;
  m = block_max_u64(m);
  if (threadIdx.x == 0) *out = m;
}
}  // namespace
}  // namespace idfx

extern "C" int idfx_compute_timestep_noalloc(idfx_ctx *c, const double *InvDt, double *dtmin,
                                             void *stream) {
  using namespace idfx;
  if (!c || !InvDt || !dtmin) return IDFX_E_ARGS;
  const GridDesc &g = ctx_grid(c);
  const long long n = (long long)(g.end[0] - g.beg[0]) * (g.end[1] - g.beg[1]) *
                      (g.end[2] - g.beg[2]);
  const long long nb = (n + kNoAllocCellsPerBlock - 1) / kNoAllocCellsPerBlock;
  if (nb < 1 || nb > kNoAllocSlots) {
    std::fprintf(stderr,
                 "idfx_compute_timestep_noalloc: %lld interior cells exceed its static "
                 "scratch (at most %lld)\n",
                 n, (long long)kNoAllocSlots * kNoAllocCellsPerBlock);
    return IDFX_E_ARGS;
  }
  unsigned long long *sp = nullptr;
  const int st = scratch_address(&sp);
  if (st != IDFX_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_invdt_partial, dim3((unsigned)nb), dim3(256), 0, s, g, InvDt, sp,
                     kNoAllocCellsPerBlock);
  hipLaunchKernelGGL(k_invdt_final, dim3(1), dim3(256), 0, s, sp, (int)nb, sp + kNoAllocSlots);
  if (hipGetLastError() != hipSuccess) return IDFX_E_LAUNCH;
  unsigned long long b = 0;  // a pageable host word: no pinned buffer is allocated
  if (hipMemcpyAsync(&b, sp + kNoAllocSlots, sizeof(b), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return IDFX_E_LAUNCH;
  double m;
  std::memcpy(&m, &b, sizeof(m));
  *dtmin = 1.0 / m;  // min(1/x) == 1/max(x) under correct rounding
  return IDFX_OK;
}
