// Context lifecycle, device memory and the host-side helpers restated from
// the reference (dense2Sparse, NormalCDFInverse, the MTE penalty).
#include "ctx.h"

#include <cmath>
#include <cstring>

int rmc_fail(roadmc_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int rmc_malloc(roadmc_ctx* ctx, void** p, size_t bytes) {
  *p = nullptr;
  if (ctx->fail_allocs > 0) {
    --ctx->fail_allocs;
    return rmc_fail(ctx, ROADMC_ERR_OOM, "device allocation failed (injected)");
  }
  if (bytes == 0) bytes = 16;
  const hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return rmc_fail(ctx, ROADMC_ERR_OOM,
                    std::string("hipMalloc(") + std::to_string(bytes) +
                        "): " + hipGetErrorString(e));
  }
  return ROADMC_OK;
}

void rmc_buf_free(rmc_devbuf& b) {
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

int rmc_buf_ensure(roadmc_ctx* ctx, rmc_devbuf& b, size_t bytes) {
  if (bytes <= b.bytes && b.ptr) return ROADMC_OK;
  // The old buffer may still be read by queued work.
  RMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  rmc_buf_free(b);
  const size_t want = bytes + bytes / 4 + 256;
  int rc = rmc_malloc(ctx, &b.ptr, want);
  if (rc) return rc;
  b.bytes = want;
  return ROADMC_OK;
}

int rmc_buf_alloc_exact(roadmc_ctx* ctx, rmc_devbuf& b, size_t bytes) {
  rmc_buf_free(b);
  int rc = rmc_malloc(ctx, &b.ptr, bytes ? bytes : 16);
  if (rc) return rc;
  b.bytes = bytes;
  return ROADMC_OK;
}

extern "C" int roadmc_device_count(int* count) {
  if (!count) return ROADMC_ERR_INVALID;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return ROADMC_OK;
}

extern "C" int roadmc_ctx_create(int device, roadmc_ctx** out) {
  if (!out) return ROADMC_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return ROADMC_ERR_HIP;
  if (device < 0 || device >= n) return ROADMC_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return ROADMC_ERR_HIP;
  roadmc_ctx* ctx = new roadmc_ctx();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete ctx;
    return ROADMC_ERR_HIP;
  }
  ctx->n_cu = prop.multiProcessorCount;
  // Default priority: streams of other priorities are not used here.
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return ROADMC_ERR_HIP;
  }
  *out = ctx;
  return ROADMC_OK;
}

extern "C" void roadmc_ctx_destroy(roadmc_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  rmc_devbuf* bufs[] = {&ctx->mte_bufs.mpack, &ctx->mte_bufs.cols, &ctx->mte_bufs.partial,
                        &ctx->mte_bufs.out,   &ctx->mte_bufs.endpop, &ctx->mte_bufs.jobs,
                        &ctx->mte_bufs.work,  &ctx->mte_bufs.off,  &ctx->mte_bufs.counter,
                        &ctx->rov_bufs.a,     &ctx->rov_bufs.b,    &ctx->rov_bufs.c,
                        &ctx->rov_bufs.d,     &ctx->rov_bufs.e,    &ctx->rov_bufs.f,
                        &ctx->rov_bufs.g,     &ctx->rov_bufs.h,    &ctx->adj_bufs.ns,
                        &ctx->adj_bufs.np,    &ctx->adj_bufs.mu,   &ctx->adj_bufs.lam,
                        &ctx->adj_bufs.gradr, &ctx->adj_bufs.gradm, &ctx->adj_bufs.gpart,
                        &ctx->adj_bufs.rpart, &ctx->adj_bufs.ckpt_spare, &ctx->road_bufs.a,   &ctx->road_bufs.b,
                        &ctx->road_bufs.c,    &ctx->road_bufs.d,   &ctx->road_bufs.e,
                        &ctx->road_bufs.f,    &ctx->road_bufs.g,   &ctx->road_bufs.h,
                        &ctx->road_bufs.tail, &ctx->road_bufs.in,
                        &ctx->knn_bufs.a,     &ctx->knn_bufs.b,    &ctx->knn_bufs.c,
                        &ctx->knn_bufs.d,     &ctx->rovs_bufs.mpack, &ctx->rovs_bufs.cols,
                        &ctx->rovs_bufs.prices, &ctx->rovs_bufs.controls, &ctx->rovs_bufs.tp,
                        &ctx->rovs_bufs.aars, &ctx->rovs_bufs.profit, &ctx->rovs_bufs.V,
                        &ctx->rovs_bufs.optc, &ctx->rovs_bufs.adj, &ctx->rovs_bufs.flag,
                        &ctx->rovs_bufs.blk,  &ctx->rovs_bufs.xs,  &ctx->rovs_bufs.ys,
                        &ctx->rovs_bufs.u,    &ctx->rovs_bufs.small, &ctx->rovs_bufs.grid,
                        &ctx->rovs_bufs.args, &ctx->rovs_bufs.ckpt, &ctx->rovs_bufs.grm,
                        &ctx->rovs_bufs.rvec, &ctx->rovs_bufs.grids,
                        &ctx->rovs_grad_bufs.mtpack, &ctx->rovs_grad_bufs.policy,
                        &ctx->rovs_grad_bufs.coef, &ctx->rovs_grad_bufs.w,
                        &ctx->rovs_grad_bufs.gpart, &ctx->rovs_grad_bufs.grad,
                        &ctx->rovs_grad_bufs.jpart, &ctx->rovs_grad_bufs.rows,
                        &ctx->rovs_grad_bufs.blk};
  for (rmc_devbuf* b : bufs) rmc_buf_free(*b);
  for (auto& l : ctx->road_lanes) {
    for (rmc_devbuf* b : {&l.cross, &l.trans, &l.surv, &l.blockmax, &l.part}) rmc_buf_free(*b);
    if (l.done) (void)hipEventDestroy(l.done);
    if (l.stream) (void)hipStreamDestroy(l.stream);
  }
  if (ctx->road_ready) (void)hipEventDestroy(ctx->road_ready);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  if (ctx->rov_graph.exec) (void)hipGraphExecDestroy(ctx->rov_graph.exec);
  rmc_buf_free(ctx->rov_graph.param);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char* roadmc_last_error(const roadmc_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int roadmc_debug_fail_allocs(roadmc_ctx* ctx, int n) {
  if (!ctx || n < 0) return ROADMC_ERR_INVALID;
  ctx->fail_allocs = n;
  return ROADMC_OK;
}

extern "C" int roadmc_dense_to_csc(const float* dense, int32_t rows, int32_t cols, float tol,
                                   float* val, int32_t* row_idx, int32_t* per_col,
                                   int32_t* nnz_out) {
  // Column by column (the order the CSC product walks), keeping values
  // strictly above tol: SimulateGPU::dense2Sparse (SimulateGPU.cu:5973-5996).
  if (!dense || !val || !row_idx || !per_col || !nnz_out || rows < 0 || cols < 0)
    return ROADMC_ERR_INVALID;
  int32_t nnz = 0;
  for (int32_t c = 0; c < cols; ++c) {
    int32_t here = 0;
    const float* col = dense + (size_t)c * rows;
    for (int32_t r = 0; r < rows; ++r) {
      if (col[r] > tol) {
        val[nnz] = col[r];
        row_idx[nnz] = r;
        ++nnz;
        ++here;
      }
    }
    per_col[c] = here;
  }
  *nnz_out = nnz;
  return ROADMC_OK;
}

namespace {
// Abramowitz and Stegun 26.2.23 rational approximation of the upper-tail
// normal quantile in t = sqrt(-2 ln p); |error| < 4.5e-4.
double as26_2_23(double t) {
  static const double c[3] = {2.515517, 0.802853, 0.010328};
  static const double d[3] = {1.432788, 0.189269, 0.001308};
  const double num = (c[2] * t + c[1]) * t + c[0];
  const double den = ((d[2] * t + d[1]) * t + d[0]) * t + 1.0;
  return t - num / den;
}
}  // namespace

extern "C" int roadmc_normal_cdf_inverse(double p, double* out) {
  // Utility::NormalCDFInverse (Utility.cpp:165-184) throws
  // std::invalid_argument outside (0, 1); this returns ROADMC_ERR_INVALID.
  if (!out || !(p > 0.0 && p < 1.0)) return ROADMC_ERR_INVALID;
  *out = p < 0.5 ? -as26_2_23(std::sqrt(-2.0 * std::log(p)))
                 : as26_2_23(std::sqrt(-2.0 * std::log(1.0 - p)));
  return ROADMC_OK;
}

extern "C" int roadmc_mte_penalty(int32_t n_species, const double* threshold,
                                  const double* end_mean, const double* end_sd,
                                  double confidence_interval, const double* initial_population,
                                  const double* cost_per_animal, double* penalty) {
  // Simulator::simulateMTE (Simulator.cpp:197-212): for each species the
  // shortfall of the (1 - CI) quantile level z*SD + mean below the threshold,
  // priced per animal of the initial population.
  if (!threshold || !end_mean || !end_sd || !initial_population || !cost_per_animal ||
      !penalty || n_species < 0)
    return ROADMC_ERR_INVALID;
  double z = 0.0;
  if (roadmc_normal_cdf_inverse(1.0 - confidence_interval, &z) != ROADMC_OK)
    return ROADMC_ERR_INVALID;
  double total = 0.0;
  for (int s = 0; s < n_species; ++s) {
    const double level = z * end_sd[s] + end_mean[s];
    if (threshold[s] > level)
      total += (threshold[s] - level) * initial_population[s] * cost_per_animal[s];
  }
  *penalty = total;
  return ROADMC_OK;
}

int rmc_pinned_ensure(roadmc_ctx* ctx, size_t bytes) {
  if (ctx->pinned_bytes >= bytes) return ROADMC_OK;
  if (ctx->pinned) {
    (void)hipHostFree(ctx->pinned);
    ctx->pinned = nullptr;
    ctx->pinned_bytes = 0;
  }
  if (ctx->fail_allocs > 0) {
    --ctx->fail_allocs;
    return rmc_fail(ctx, ROADMC_ERR_OOM, "pinned host allocation failed (injected)");
  }
  if (hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault) != hipSuccess) {
    ctx->pinned = nullptr;
    return rmc_fail(ctx, ROADMC_ERR_OOM, "pinned host allocation failed");
  }
  ctx->pinned_bytes = bytes;
  return ROADMC_OK;
}
