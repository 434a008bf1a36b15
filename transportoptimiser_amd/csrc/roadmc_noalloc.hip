// Expected PV without device allocation: roadmc_exp_pv_noalloc.
//
// roadmc_exp_pv allocates a per-path value buffer on every call.  This
// variant keeps that buffer, and the two statistics reduced from it, in a
// fixed file-scope device array, so that callers that cannot allocate device
// memory (code under stream capture, tight serving loops) can run the same
// evaluation.  The path kernel is exp_pv_kernel (rov_kernels.hip), the one
// roadmc_exp_pv launches; mean and SD are reduced on the device in FP64 in a
// fixed order, so results agree with roadmc_exp_pv (a host reduction) to
// FP64 rounding.
//
// Every call on a device shares that device's array, which keeps the last
// call's contents: calls that overlap on one device (two contexts on one
// GPU, two host threads) race.  Nothing has to start at zero: the path
// kernel writes every value the reduction reads, and the reduction writes
// both statistics.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "ctx.h"
#include "rng.h"
#include "rov_kernels.h"

namespace {

// Paths one call may take; the array below holds their values (8 MiB) and the
// two statistics.
constexpr int64_t kNoAllocMaxPaths = int64_t(1) << 20;
static_assert(kNoAllocMaxPaths == ROADMC_NOALLOC_MAX_PATHS, "header limit");

static __device__ __align__(256) double scratch[kNoAllocMaxPaths + 2];

// The array's device address, per device, found once with that device
// current (the array's name is not a device address in host code).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_address(roadmc_ctx* ctx, double** out) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED, "roadmc_exp_pv_noalloc: device index >= 64");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return rmc_fail(ctx, ROADMC_ERR_HIP,
                      std::string("roadmc_exp_pv_noalloc: hipGetSymbolAddress: ") +
                          hipGetErrorString(e));
    g_scratch[ctx->device] = static_cast<double*>(p);
  }
  *out = g_scratch[ctx->device];
  return ROADMC_OK;
}

constexpr int kStatThreads = 1024;

// Mean, then SD with 1/N, of v[0 .. n): one workgroup, each thread summing a
// fixed stride of the values, then a fixed-shape tree in LDS.
__global__ __launch_bounds__(kStatThreads) void exp_pv_stats_kernel(const double* __restrict__ v,
                                                                    int n, double* out) {
  __shared__ double red[kStatThreads];
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int i = tid; i < n; i += kStatThreads) s += v[i];
  red[tid] = s;
  __syncthreads();
  for (int w = kStatThreads / 2; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const double mean = red[0] / (double)n;
  __syncthreads();
  double q = 0.0;
  for (int i = tid; i < n; i += kStatThreads) {
    const double d = v[i] - mean;
    q += d * d;
  }
  red[tid] = q;
  __syncthreads();
  for (int w = kStatThreads / 2; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) {
    out[0] = mean;
    out[1] = sqrt(red[0] / (double)n);
  }
}

}  // namespace

extern "C" int roadmc_exp_pv_noalloc(roadmc_ctx* ctx, const roadmc_uncertainty* unc,
                                     int32_t n_steps, double time_step, double rrr,
                                     double growth, int64_t n_paths, uint64_t seed, double* mean,
                                     double* sd) {
  if (!ctx) return ROADMC_ERR_INVALID;
  if (!unc || !mean || !sd || n_steps < 1 || n_paths < 1)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "roadmc_exp_pv_noalloc: bad arguments");
  if (n_paths > kNoAllocMaxPaths)
    return rmc_fail(ctx, ROADMC_ERR_INVALID,
                    "roadmc_exp_pv_noalloc: n_paths exceeds ROADMC_NOALLOC_MAX_PATHS (2^20); "
                    "use roadmc_exp_pv");
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  double* buf = nullptr;
  int rc = scratch_address(ctx, &buf);
  if (rc) return rc;
  roadmc::ExpPvArgs a{};
  a.N = (int)n_paths;
  a.T = n_steps;
  a.x0 = unc->x0;
  a.mean = unc->mean;
  a.sd = unc->sd;
  a.reversion = unc->reversion;
  a.jump = unc->jump_size;
  a.jump_prob = unc->jump_prob;
  a.dt = time_step;
  a.rrr = rrr;
  a.growth = growth;
  a.tag = rmc_tag(1, 0, RMC_KIND_ROV);  // the stream of roadmc_exp_pv
  a.k0 = (uint32_t)(seed & 0xffffffffu);
  a.k1 = (uint32_t)(seed >> 32);
  a.out = buf;
  RMC_HIP(ctx, roadmc::launch_exp_pv(a, ctx->stream));
  hipLaunchKernelGGL(exp_pv_stats_kernel, dim3(1), dim3(kStatThreads), 0, ctx->stream, buf,
                     a.N, buf + kNoAllocMaxPaths);
  RMC_HIP(ctx, hipGetLastError());
  double stats[2];
  RMC_HIP(ctx, hipMemcpyAsync(stats, buf + kNoAllocMaxPaths, sizeof(stats),
                              hipMemcpyDeviceToHost, ctx->stream));
  RMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *mean = stats[0];
  *sd = stats[1];
  return ROADMC_OK;
}
