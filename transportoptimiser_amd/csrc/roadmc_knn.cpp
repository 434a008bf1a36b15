// C-ABI host code for the grid surrogate (knn_kernels.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.h"
#include "knn_kernels.h"
#include "roadmc.h"

static_assert(ROADMC_KNN_MAX_K == roadmc::kKnnMaxK,
              "the header's neighbour limit must be the one the kernels enforce");

// The reference's rules, capped at the kernels' limit: its ROV rule grows
// without bound (25 + ceil(0.02 (n - 500)) passes 128 above 5 650 samples).
extern "C" int roadmc_knn_k(int32_t n, int32_t rule, int32_t* k_out) {
  if (!k_out || n < 0 || (rule != ROADMC_KNN_RULE_MTE && rule != ROADMC_KNN_RULE_ROV))
    return ROADMC_ERR_INVALID;
  int k;
  if (rule == ROADMC_KNN_RULE_MTE) {  // SimulateGPU.cu:5545-5554
    if (n < 50)
      k = 5;
    else if (n < 100)
      k = 7;
    else if (n < 500)
      k = (int)std::ceil(0.02 * (n - 100) + 7);
    else
      k = (int)std::ceil(std::log((double)n) + 8);
  } else {  // :5711-5724; the inner quotients are integer divisions there
    if (n < 50)
      k = 5;
    else if (n < 100)
      k = 5 + 5 * ((n - 50) / 50);
    else if (n < 500)
      k = 10 + 15 * ((n - 100) / 400);
    else
      k = (int)(25 + std::ceil((n - 500) * 0.02));
  }
  *k_out = std::min({k, n, roadmc::kKnnMaxK});
  return ROADMC_OK;
}

namespace {

bool grid_size(int d, int R, long long* G) {
  long long g = 1;
  for (int i = 0; i < d; ++i) {
    g *= R;
    if (g > (1LL << 24)) return false;
  }
  *G = g;
  return true;
}

}  // namespace

extern "C" int roadmc_knn_surrogate_fit(roadmc_ctx* ctx, int32_t n, int32_t d, const double* x,
                                        const double* y, int32_t R, int32_t k, int32_t flags,
                                        double* surrogate) {
  if (!ctx) return ROADMC_ERR_INVALID;
  long long G = 0;
  if (d < 1 || d > ROADMC_KNN_MAX_DIMS) return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_dims out of range");
  if (R < 2 || !grid_size(d, R, &G))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "dim_res must be >= 2 with dim_res^n_dims <= 2^24");
  if (n < 0 || n > (1 << 24)) return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_samples out of range");
  if (k < 1 || k > roadmc::kKnnMaxK)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "k must be in [1, ROADMC_KNN_MAX_K]");
  if (flags & ~(ROADMC_KNN_NORMALIZE | ROADMC_KNN_NONPOSITIVE))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "unknown flags");
  if (!surrogate || (n > 0 && (!x || !y)))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "NULL array");
  for (size_t i = 0; i < (size_t)n * d; ++i)
    if (!std::isfinite(x[i])) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite sample point");
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(y[i])) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite sample value");
  const bool norm = flags & ROADMC_KNN_NORMALIZE;
  // Ranges, grid coordinates (createQueryPoints, SimulateGPU.cu:1176-1190),
  // and the coordinates of the search and the fit.
  std::vector<double> mn(d, 0.0), mx(d, 0.0), span(d, 0.0), grid((size_t)d * R), qtab((size_t)d * R);
  for (int dd = 0; dd < d; ++dd) {
    if (n > 0) mn[dd] = mx[dd] = x[dd];
    for (int j = 1; j < n; ++j) {
      mn[dd] = std::fmin(mn[dd], x[(size_t)j * d + dd]);
      mx[dd] = std::fmax(mx[dd], x[(size_t)j * d + dd]);
    }
    span[dd] = mx[dd] - mn[dd];
    for (int i = 0; i < R; ++i) {
      grid[(size_t)dd * R + i] = (double)i * span[dd] / (double)(R - 1) + mn[dd];
      qtab[(size_t)dd * R + i] =
          norm ? (span[dd] > 0.0 ? (double)i / (double)(R - 1) : 0.0) : grid[(size_t)dd * R + i];
    }
  }
  std::vector<double> vals((size_t)G, 0.0);
  if (n >= 3) {  // fewer than 3 samples: every value is 0 (multiLocLinReg, :2943-2947)
    std::vector<double> u((size_t)n * d);
    for (int j = 0; j < n; ++j)
      for (int dd = 0; dd < d; ++dd) {
        const double v = x[(size_t)j * d + dd];
        u[(size_t)j * d + dd] = norm ? (span[dd] > 0.0 ? (v - mn[dd]) / span[dd] : 0.0) : v;
      }
    RMC_HIP(ctx, hipSetDevice(ctx->device));
    auto& B = ctx->knn_bufs;
    int rc;
    if ((rc = rmc_buf_ensure(ctx, B.a, u.size() * sizeof(double))) ||
        (rc = rmc_buf_ensure(ctx, B.b, (size_t)n * sizeof(double))) ||
        (rc = rmc_buf_ensure(ctx, B.c, qtab.size() * sizeof(double))) ||
        (rc = rmc_buf_ensure(ctx, B.d, (size_t)G * sizeof(double))))
      return rc;
    hipStream_t s = ctx->stream;
    RMC_HIP(ctx, hipMemcpyAsync(B.a.ptr, u.data(), u.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
    RMC_HIP(ctx, hipMemcpyAsync(B.b.ptr, y, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
    RMC_HIP(ctx, hipMemcpyAsync(B.c.ptr, qtab.data(), qtab.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
    roadmc::KnnFitArgs a{};
    a.n = n;
    a.d = d;
    a.R = R;
    a.k = std::min(k, n);
    a.cap = std::min(a.k, (n + roadmc::kKnnThreads - 1) / roadmc::kKnnThreads);
    a.clamp = (flags & ROADMC_KNN_NONPOSITIVE) ? 1 : 0;
    a.G = G;
    a.u = static_cast<const double*>(B.a.ptr);
    a.y = static_cast<const double*>(B.b.ptr);
    a.qtab = static_cast<const double*>(B.c.ptr);
    a.val = static_cast<double*>(B.d.ptr);
    RMC_HIP(ctx, roadmc::launch_knn_fit(a, s));
    RMC_HIP(ctx, hipMemcpyAsync(vals.data(), a.val, (size_t)G * sizeof(double),
                                hipMemcpyDeviceToHost, s));
    RMC_HIP(ctx, hipStreamSynchronize(s));
    for (double v : vals)
      if (!std::isfinite(v)) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite fitted value");
  }
  std::memcpy(surrogate, grid.data(), grid.size() * sizeof(double));
  std::memcpy(surrogate + grid.size(), vals.data(), vals.size() * sizeof(double));
  return ROADMC_OK;
}

extern "C" int roadmc_knn_surrogate_eval(roadmc_ctx* ctx, int32_t d, int32_t R,
                                         const double* surrogate, int32_t m,
                                         const double* pred, double* out) {
  if (!ctx) return ROADMC_ERR_INVALID;
  long long G = 0;
  if (d < 1 || d > ROADMC_KNN_MAX_DIMS) return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_dims out of range");
  if (R < 2 || !grid_size(d, R, &G))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "dim_res must be >= 2 with dim_res^n_dims <= 2^24");
  if (m < 0) return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_points < 0");
  if (m == 0) return ROADMC_OK;
  if (!surrogate || !pred || !out) return rmc_fail(ctx, ROADMC_ERR_INVALID, "NULL array");
  for (size_t i = 0; i < (size_t)m * d; ++i)
    if (!std::isfinite(pred[i])) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite predictor");
  const size_t ns = (size_t)d * R + (size_t)G;
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  auto& B = ctx->knn_bufs;
  int rc;
  if ((rc = rmc_buf_ensure(ctx, B.a, ns * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, B.b, (size_t)m * d * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, B.d, (size_t)m * sizeof(double))))
    return rc;
  hipStream_t s = ctx->stream;
  RMC_HIP(ctx, hipMemcpyAsync(B.a.ptr, surrogate, ns * sizeof(double), hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(B.b.ptr, pred, (size_t)m * d * sizeof(double),
                              hipMemcpyHostToDevice, s));
  roadmc::KnnEvalArgs a{};
  a.d = d;
  a.R = R;
  a.m = m;
  a.surrogate = static_cast<const double*>(B.a.ptr);
  a.pred = static_cast<const double*>(B.b.ptr);
  a.out = static_cast<double*>(B.d.ptr);
  RMC_HIP(ctx, roadmc::launch_knn_eval(a, s));
  std::vector<double> res(m);
  RMC_HIP(ctx, hipMemcpyAsync(res.data(), a.out, (size_t)m * sizeof(double),
                              hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipStreamSynchronize(s));
  std::memcpy(out, res.data(), (size_t)m * sizeof(double));
  return ROADMC_OK;
}
