// C-ABI host code for the surrogate MLP ensemble.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"
#include "mlp_kernels.h"
#include "roadmc.h"

struct roadmc_mlp {
  roadmc_ctx* ctx = nullptr;
  roadmc::MlpArgs a{};
  int np = 0;    // parameters per member
  long step = 0; // Adam steps taken
  rmc_devbuf params, m, v, x, y, partial, lpart, loss, grad, out;
};

namespace {

int mlp_buffers(roadmc_mlp* net, int n) {
  roadmc_ctx* ctx = net->ctx;
  const int nchunk = (n + 63) / 64;
  const size_t E = net->a.E;
  int rc;
  if ((rc = rmc_buf_ensure(ctx, net->x, (size_t)n * net->a.n_in * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, net->y, (size_t)n * net->a.n_out * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, net->partial, E * nchunk * net->np * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, net->lpart, E * nchunk * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, net->loss, E * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, net->grad, E * net->np * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, net->out, E * (size_t)n * net->a.n_out * sizeof(float))))
    return rc;
  return ROADMC_OK;
}

bool finite(const float* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

}  // namespace

extern "C" int roadmc_mlp_create(roadmc_ctx* ctx, int32_t n_in, int32_t n_hidden, int32_t n_out,
                                 int32_t n_members, roadmc_mlp** out) {
  if (!ctx) return ROADMC_ERR_INVALID;
  if (!out) return rmc_fail(ctx, ROADMC_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (n_in < 1 || n_in > roadmc::kMlpMaxIn || n_out < 1 || n_out > roadmc::kMlpMaxOut ||
      n_members < 1 || n_members > roadmc::kMlpMaxMembers)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "mlp: n_in <= 16, n_out <= 4, members <= 8");
  if (n_hidden != 64) return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED, "mlp: hidden width is 64");
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  roadmc_mlp* net = new roadmc_mlp();
  net->ctx = ctx;
  net->a.n_in = n_in;
  net->a.n_out = n_out;
  net->a.E = n_members;
  net->np = roadmc::mlp_param_count(n_in, n_out);
  const size_t bytes = (size_t)n_members * net->np * sizeof(float);
  int rc;
  if ((rc = rmc_buf_alloc_exact(ctx, net->params, bytes)) ||
      (rc = rmc_buf_alloc_exact(ctx, net->m, bytes)) ||
      (rc = rmc_buf_alloc_exact(ctx, net->v, bytes))) {
    roadmc_mlp_destroy(net);
    return rc;
  }
  net->a.params = static_cast<float*>(net->params.ptr);
  net->a.m = static_cast<float*>(net->m.ptr);
  net->a.v = static_cast<float*>(net->v.ptr);
  hipError_t e = hipMemsetAsync(net->params.ptr, 0, bytes, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(net->m.ptr, 0, bytes, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(net->v.ptr, 0, bytes, ctx->stream);
  if (e == hipSuccess) e = roadmc::mlp_set_lds_limit(n_in, n_out);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    roadmc_mlp_destroy(net);
    return rmc_fail(ctx, ROADMC_ERR_HIP, std::string("mlp create: ") + hipGetErrorString(e));
  }
  *out = net;
  return ROADMC_OK;
}

extern "C" void roadmc_mlp_destroy(roadmc_mlp* net) {
  if (!net) return;
  (void)hipStreamSynchronize(net->ctx->stream);
  rmc_devbuf* all[] = {&net->params, &net->m,     &net->v,    &net->x,    &net->y,
                       &net->partial, &net->lpart, &net->loss, &net->grad, &net->out};
  for (rmc_devbuf* b : all) rmc_buf_free(*b);
  delete net;
}

extern "C" int roadmc_mlp_n_params(const roadmc_mlp* net, int32_t* n) {
  if (!net || !n) return ROADMC_ERR_INVALID;
  *n = net->np;
  return ROADMC_OK;
}

extern "C" int roadmc_mlp_set_params(roadmc_mlp* net, const float* params) {
  if (!net) return ROADMC_ERR_INVALID;
  roadmc_ctx* ctx = net->ctx;
  const size_t n = (size_t)net->a.E * net->np;
  if (!params || !finite(params, n)) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "bad parameters");
  RMC_HIP(ctx, hipMemcpy(net->params.ptr, params, n * sizeof(float), hipMemcpyHostToDevice));
  RMC_HIP(ctx, hipMemset(net->m.ptr, 0, n * sizeof(float)));
  RMC_HIP(ctx, hipMemset(net->v.ptr, 0, n * sizeof(float)));
  net->step = 0;
  return ROADMC_OK;
}

extern "C" int roadmc_mlp_get_params(roadmc_mlp* net, float* params) {
  if (!net || !params) return ROADMC_ERR_INVALID;
  RMC_HIP(net->ctx, hipMemcpy(params, net->params.ptr, (size_t)net->a.E * net->np * sizeof(float),
                              hipMemcpyDeviceToHost));
  return ROADMC_OK;
}

// One gradient evaluation (no update): grad [E][n_params], loss [E].
extern "C" int roadmc_mlp_gradient(roadmc_mlp* net, const float* x, const float* y, int32_t n,
                                   float* grad, float* loss) {
  if (!net) return ROADMC_ERR_INVALID;
  roadmc_ctx* ctx = net->ctx;
  if (!x || !y || n < 1 || (!grad && !loss))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "mlp gradient: bad arguments");
  int rc = mlp_buffers(net, n);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  const int nchunk = (n + 63) / 64;
  RMC_HIP(ctx, hipMemcpyAsync(net->x.ptr, x, (size_t)n * net->a.n_in * sizeof(float),
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(net->y.ptr, y, (size_t)n * net->a.n_out * sizeof(float),
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, roadmc::launch_mlp_grad(net->a, static_cast<float*>(net->x.ptr),
                                       static_cast<float*>(net->y.ptr), n,
                                       static_cast<float*>(net->partial.ptr),
                                       static_cast<float*>(net->lpart.ptr), s));
  RMC_HIP(ctx, roadmc::launch_mlp_update(net->a, static_cast<float*>(net->partial.ptr), nchunk,
                                         static_cast<float*>(net->grad.ptr), 0.0f, 1.0f, 1.0f,
                                         static_cast<float*>(net->lpart.ptr),
                                         static_cast<float*>(net->loss.ptr), s));
  std::vector<float> g((size_t)net->a.E * net->np), l(net->a.E);
  RMC_HIP(ctx, hipMemcpyAsync(g.data(), net->grad.ptr, g.size() * sizeof(float),
                              hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipMemcpyAsync(l.data(), net->loss.ptr, l.size() * sizeof(float),
                              hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipStreamSynchronize(s));
  if (grad) std::memcpy(grad, g.data(), g.size() * sizeof(float));
  if (loss) std::memcpy(loss, l.data(), l.size() * sizeof(float));
  return ROADMC_OK;
}

// Full-batch Adam for `steps` steps on (x, y); loss [E] after the last step.
extern "C" int roadmc_mlp_train(roadmc_mlp* net, const float* x, const float* y, int32_t n,
                                int32_t steps, float lr, float* loss) {
  if (!net) return ROADMC_ERR_INVALID;
  roadmc_ctx* ctx = net->ctx;
  if (!x || !y || n < 1 || steps < 0 || !(lr > 0.0f))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "mlp train: bad arguments");
  if (!finite(x, (size_t)n * net->a.n_in) || !finite(y, (size_t)n * net->a.n_out))
    return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "mlp train: non-finite data");
  int rc = mlp_buffers(net, n);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  const int nchunk = (n + 63) / 64;
  RMC_HIP(ctx, hipMemcpyAsync(net->x.ptr, x, (size_t)n * net->a.n_in * sizeof(float),
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(net->y.ptr, y, (size_t)n * net->a.n_out * sizeof(float),
                              hipMemcpyHostToDevice, s));
  for (int it = 0; it < steps; ++it) {
    ++net->step;
    const float bc1 = 1.0f - std::pow(0.9f, (float)net->step);
    const float bc2 = 1.0f - std::pow(0.999f, (float)net->step);
    RMC_HIP(ctx, roadmc::launch_mlp_grad(net->a, static_cast<float*>(net->x.ptr),
                                         static_cast<float*>(net->y.ptr), n,
                                         static_cast<float*>(net->partial.ptr),
                                         static_cast<float*>(net->lpart.ptr), s));
    RMC_HIP(ctx, roadmc::launch_mlp_update(net->a, static_cast<float*>(net->partial.ptr), nchunk,
                                           nullptr, lr, bc1, bc2,
                                           static_cast<float*>(net->lpart.ptr),
                                           it + 1 == steps ? static_cast<float*>(net->loss.ptr)
                                                           : nullptr,
                                           s));
  }
  std::vector<float> l(net->a.E, 0.0f);
  if (steps > 0)
    RMC_HIP(ctx, hipMemcpyAsync(l.data(), net->loss.ptr, l.size() * sizeof(float),
                                hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipStreamSynchronize(s));
  if (loss) std::memcpy(loss, l.data(), l.size() * sizeof(float));
  return ROADMC_OK;
}

// Predictions of every member: out [E][n][n_out].
extern "C" int roadmc_mlp_predict(roadmc_mlp* net, const float* x, int32_t n, float* out) {
  if (!net) return ROADMC_ERR_INVALID;
  roadmc_ctx* ctx = net->ctx;
  if (!x || !out || n < 1) return rmc_fail(ctx, ROADMC_ERR_INVALID, "mlp predict: bad arguments");
  int rc = mlp_buffers(net, n);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  RMC_HIP(ctx, hipMemcpyAsync(net->x.ptr, x, (size_t)n * net->a.n_in * sizeof(float),
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, roadmc::launch_mlp_predict(net->a, static_cast<float*>(net->x.ptr), n,
                                          static_cast<float*>(net->out.ptr), s));
  RMC_HIP(ctx, hipMemcpyAsync(out, net->out.ptr, (size_t)net->a.E * n * net->a.n_out * sizeof(float),
                              hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipStreamSynchronize(s));
  return ROADMC_OK;
}
