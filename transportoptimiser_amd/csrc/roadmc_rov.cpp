// C-ABI host code for the ROV path and the expected-PV helper.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"
#include "rng.h"
#include "roadmc.h"
#include "rov_kernels.h"

namespace {

int basis_size(int n, int deg) {
  // C(n + deg, deg)
  long r = 1;
  for (int i = 1; i <= deg; ++i) r = r * (n + i) / i;
  return (int)r;
}

// Lower Cholesky factor of a correlation matrix; false if not positive
// definite.
bool cholesky(const double* C, int n, double* L) {
  for (int i = 0; i < n * n; ++i) L[i] = 0.0;
  for (int j = 0; j < n; ++j) {
    double d = C[j * n + j];
    for (int k = 0; k < j; ++k) d -= L[j * n + k] * L[j * n + k];
    if (!(d > 1e-14)) return false;
    L[j * n + j] = std::sqrt(d);
    for (int i = j + 1; i < n; ++i) {
      double s = C[i * n + j];
      for (int k = 0; k < j; ++k) s -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = s / L[j * n + j];
    }
  }
  return true;
}

int validate_rov(roadmc_ctx* ctx, const roadmc_rov_problem* p) {
  if (!p) return rmc_fail(ctx, ROADMC_ERR_INVALID, "rov problem is NULL");
  if (p->n_uncertainties < 1 || p->n_uncertainties > ROADMC_MAX_UNCERTAINTIES || !p->unc)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_uncertainties out of range");
  if (p->mode != ROADMC_ROV_SWITCHING && p->mode != ROADMC_ROV_AMERICAN)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "unknown ROV mode");
  if (p->value_mode != ROADMC_VALUE_REGRESSED && p->value_mode != ROADMC_VALUE_REALIZED)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "unknown value mode");
  if (p->mode == ROADMC_ROV_AMERICAN && p->value_mode != ROADMC_VALUE_REALIZED)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED,
                    "ROADMC_ROV_AMERICAN carries realized cash flows back: "
                    "value_mode must be ROADMC_VALUE_REALIZED");
  if (p->mode == ROADMC_ROV_SWITCHING) {
    if (p->n_controls < 1 || p->n_controls > ROADMC_MAX_CONTROLS || !p->flow)
      return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_controls/flow invalid");
    if (p->n_steps < 2) return rmc_fail(ctx, ROADMC_ERR_INVALID, "switching needs n_steps >= 2");
    if (p->traffic_index >= p->n_uncertainties)
      return rmc_fail(ctx, ROADMC_ERR_INVALID, "traffic_index out of range");
  } else if (p->n_steps < 1) {
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_steps must be >= 1");
  }
  if (p->n_paths < 1 || p->n_paths > 0x7fffffffLL)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_paths must be in [1, 2^31)");
  if (p->basis_degree < 1 || p->basis_degree > 3)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "basis_degree must be 1..3");
  if (basis_size(p->n_uncertainties, p->basis_degree) > ROADMC_MAX_BASIS)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "basis larger than ROADMC_MAX_BASIS");
  if (!(p->time_step > 0.0) || !std::isfinite(p->rrr) || !std::isfinite(p->hysteresis))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "time_step/rrr/hysteresis invalid");
  for (int u = 0; u < p->n_uncertainties; ++u) {
    const roadmc_uncertainty& q = p->unc[u];
    if (q.kind < ROADMC_SDE_MRJD || q.kind > ROADMC_SDE_GBM)
      return rmc_fail(ctx, ROADMC_ERR_INVALID, "unknown SDE kind");
    const double v[6] = {q.x0, q.mean, q.sd, q.reversion, q.jump_size, q.jump_prob};
    for (double x : v)
      if (!std::isfinite(x)) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite parameter");
  }
  return ROADMC_OK;
}

}  // namespace

extern "C" int roadmc_rov_basis_size(int32_t n_vars, int32_t degree, int32_t* size) {
  if (!size || n_vars < 1 || degree < 0 || degree > 3) return ROADMC_ERR_INVALID;
  *size = basis_size(n_vars, degree);
  return ROADMC_OK;
}

extern "C" int roadmc_rov_run(roadmc_ctx* ctx, const roadmc_rov_problem* pr,
                              roadmc_rov_result* res) {
  if (!ctx) return ROADMC_ERR_INVALID;
  if (!res) return rmc_fail(ctx, ROADMC_ERR_INVALID, "result is NULL");
  int rc = validate_rov(ctx, pr);
  if (rc) return rc;
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  const int N = (int)pr->n_paths, T = pr->n_steps, U = pr->n_uncertainties;
  const bool american = pr->mode == ROADMC_ROV_AMERICAN;
  const int C = american ? 2 : pr->n_controls;
  const int K = basis_size(U, pr->basis_degree);
  const int E = K * (K + 1) / 2 + K + 1;
  const int G = (american || pr->pooled) ? 1 : C;
  const int nblk = (N + roadmc::kRovStepPaths - 1) / roadmc::kRovStepPaths;
  const int nblk256 = (N + 255) / 256;
  const int nblk_gram = (N + roadmc::kRovGramPaths - 1) / roadmc::kRovGramPaths;
  const int t_term = american ? T : T - 1;
  const int n_fit = t_term - 1;  // steps 1 .. t_term - 1 are fitted
  const size_t GE = (size_t)G * E, GK = (size_t)G * K;

  double L[roadmc::kRovMaxU * roadmc::kRovMaxU] = {0};
  if (pr->correlation) {
    std::vector<double> Lu((size_t)U * U);
    if (!cholesky(pr->correlation, U, Lu.data()))
      return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "correlation matrix is not positive definite");
    for (int i = 0; i < U; ++i)
      for (int j = 0; j < U; ++j) L[i * roadmc::kRovMaxU + j] = Lu[(size_t)i * U + j];
  } else {
    for (int i = 0; i < U; ++i) L[i * roadmc::kRovMaxU + i] = 1.0;
  }

  auto& B = ctx->rov_bufs;
  const size_t n_states = (size_t)(T + 1) * U * N;
  if ((rc = rmc_buf_ensure(ctx, B.a, n_states * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, B.b, (size_t)T * N * sizeof(int))) ||
      (rc = rmc_buf_ensure(ctx, B.c, (size_t)(T + 1) * N * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, B.d, (size_t)T * N * sizeof(int))) ||
      (rc = rmc_buf_ensure(ctx, B.e, (size_t)(T + 1) * U * 18 * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, B.f, ((size_t)std::max(n_fit, 1) * nblk_gram * GE +
                                       (size_t)T * GK * K + 2 * (size_t)nblk * GK +
                                       (size_t)nblk256 * 3 * roadmc::kRovMaxC) *
                                          sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, B.g, (size_t)T * G * K * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, B.h, ((size_t)T * G + 1) * sizeof(int) + 128)))
    return rc;
  double* d_states = static_cast<double*>(B.a.ptr);
  int* d_ctrl = static_cast<int*>(B.b.ptr);
  double* d_V = static_cast<double*>(B.c.ptr);
  int* d_optc = static_cast<int*>(B.d.ptr);
  double* d_scale = static_cast<double*>(B.e.ptr);
  // B.h: status [T][G] ints, out [3] doubles (16-byte aligned), tickets.
  int* d_status = static_cast<int*>(B.h.ptr);
  const size_t out_off = ((size_t)T * G * sizeof(int) + 15) & ~(size_t)15;
  double* d_out = reinterpret_cast<double*>(static_cast<char*>(B.h.ptr) + out_off);
  int* d_ticket = reinterpret_cast<int*>(static_cast<char*>(B.h.ptr) + out_off + 32);
  hipStream_t s = ctx->stream;

  const uint32_t k0 = (uint32_t)(pr->seed & 0xffffffffu), k1 = (uint32_t)(pr->seed >> 32);
  const uint32_t tag = rmc_tag(0, 0, RMC_KIND_ROV);
  if (pr->states) {
    RMC_HIP(ctx, hipMemcpyAsync(d_states, pr->states, n_states * sizeof(double),
                                hipMemcpyHostToDevice, s));
  }
  if (pr->controls) {
    RMC_HIP(ctx, hipMemcpyAsync(d_ctrl, pr->controls, (size_t)T * N * sizeof(int),
                                hipMemcpyHostToDevice, s));
  }
  if (!pr->states || !pr->controls) {
    roadmc::RovForwardArgs fa{};
    fa.N = N;
    fa.T = T;
    fa.U = U;
    fa.C = C;
    fa.dt = pr->time_step;
    for (int u = 0; u < U; ++u) {
      const roadmc_uncertainty& q = pr->unc[u];
      const double v[7] = {(double)q.kind, q.x0, q.mean, q.sd, q.reversion, q.jump_size,
                           q.jump_prob};
      for (int i = 0; i < 7; ++i) fa.par[u * 7 + i] = v[i];
    }
    std::memcpy(fa.chol, L, sizeof(L));
    fa.tag = tag;
    fa.k0 = k0;
    fa.k1 = k1;
    // With caller states, the forward pass only draws the controls into a
    // scratch copy of the states.
    fa.states = pr->states ? static_cast<double*>(B.c.ptr) : d_states;
    fa.controls = pr->controls ? nullptr : d_ctrl;
    if (pr->states) {
      // B.c (V) is at least (T+1)*N doubles; the scratch needs (T+1)*U*N.
      if ((rc = rmc_buf_ensure(ctx, B.c, std::max((size_t)(T + 1) * N, n_states) *
                                             sizeof(double))))
        return rc;
      d_V = static_cast<double*>(B.c.ptr);
      fa.states = d_V;
    }
    RMC_HIP(ctx, roadmc::launch_rov_forward(fa, s));
  }
  if (pr->standardize) RMC_HIP(ctx, roadmc::launch_rov_moments(d_states, N, U, T + 1, d_scale, s));

  roadmc::RovStepArgs a{};
  a.N = N;
  a.T = T;
  a.U = U;
  a.C = C;
  a.G = G;
  a.K = K;
  a.E = E;
  a.deg = pr->basis_degree;
  a.mode = pr->mode;
  a.value_mode = pr->value_mode;
  a.pooled = pr->pooled;
  a.itm_only = pr->itm_only;
  a.traffic = american ? -1 : pr->traffic_index;
  a.disc = pr->discount > 0.0 ? pr->discount : 1.0 / (1.0 + pr->rrr * pr->time_step / 100.0);
  a.hysteresis = pr->hysteresis;
  a.unit_cost = pr->unit_cost;
  a.strike = pr->strike;
  for (int u = 0; u < U; ++u) a.weight[u] = pr->profit_weight ? pr->profit_weight[u] : 0.0;
  for (int c = 0; c < C; ++c) a.flow[c] = (!american && pr->flow) ? pr->flow[c] : 0.0;
  a.states = d_states;
  a.scale = pr->standardize ? d_scale : nullptr;
  a.controls = d_ctrl;
  a.V = d_V;
  a.optc = d_optc;
  a.nblk = nblk;
  a.partial = static_cast<double*>(B.f.ptr);
  a.ainv = a.partial + (size_t)std::max(n_fit, 1) * nblk_gram * GE;
  a.bpart = a.ainv + (size_t)T * GK * K;
  a.fpart = a.bpart + 2 * (size_t)nblk * GK;
  a.ticket = d_ticket;
  a.beta = static_cast<double*>(B.g.ptr);
  a.status = d_status;

  // Zero the coefficients so steps without a fit read zeros, and the arrival
  // ticket (each launch leaves it zero again; cleared here in case the buffer
  // was regrown).  Both before, not inside, the captured graph.
  RMC_HIP(ctx, hipMemsetAsync(a.beta, 0, (size_t)T * G * K * sizeof(double), s));
  RMC_HIP(ctx, hipMemsetAsync(d_status, 0, (size_t)T * G * sizeof(int), s));
  RMC_HIP(ctx, hipMemsetAsync(d_ticket, 0, sizeof(int), s));

  RMC_HIP(ctx, roadmc::rov_set_lds_limits());
  a.out = d_out;
  // Parameters go to the device block the captured kernels read.
  auto& RG = ctx->rov_graph;
  if ((rc = rmc_buf_ensure(ctx, RG.param, sizeof(roadmc::RovStepArgs)))) return rc;
  const roadmc::RovStepArgs* pa = static_cast<const roadmc::RovStepArgs*>(RG.param.ptr);
  RMC_HIP(ctx, hipMemcpyAsync(RG.param.ptr, &a, sizeof(a), hipMemcpyHostToDevice, s));
  // The graph depends only on the launch shapes and the parameter block's
  // address.
  const long long key[8] = {N, T, K, G * 1000LL + E, C * 100LL + U, american ? 1 : 0,
                            (long long)(uintptr_t)pa, a.nblk};
  if (!RG.exec || std::memcmp(key, RG.key, sizeof(key)) != 0) {
    if (RG.exec) {
      (void)hipGraphExecDestroy(RG.exec);
      RG.exec = nullptr;
    }
    // Backward induction as one graph: the regression matrices of every
    // fitted step (t_term - 1 .. 1) and their inverses, then one launch per
    // step (the coefficients of the step after it, its decisions, the
    // right-hand sides of this step), then step 1 and time 0.
    RMC_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = roadmc::launch_rov_gram(a, pa, 1, n_fit, s);
    for (int t = t_term - 1; t >= 1 && e == hipSuccess; --t)
      e = roadmc::launch_rov_step(a, pa, t, t + 1 == t_term ? 2 : 1, s);
    if (e == hipSuccess) e = roadmc::launch_rov_final(a, pa, t_term == 1, s);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (e != hipSuccess || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return rmc_fail(ctx, ROADMC_ERR_HIP,
                      std::string("rov capture: ") + hipGetErrorString(e != hipSuccess ? e : ec));
    }
    e = hipGraphInstantiate(&RG.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      RG.exec = nullptr;
      return rmc_fail(ctx, ROADMC_ERR_HIP, std::string("rov graph: ") + hipGetErrorString(e));
    }
    std::memcpy(RG.key, key, sizeof(key));
  }
  hipError_t e = hipGraphLaunch(RG.exec, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess)
    return rmc_fail(ctx, ROADMC_ERR_HIP, std::string("rov graph: ") + hipGetErrorString(e));

  double out[3];
  RMC_HIP(ctx, hipMemcpy(out, d_out, sizeof(out), hipMemcpyDeviceToHost));
  if (!std::isfinite(out[0]) || !std::isfinite(out[2]))
    return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite ROV value");
  // Optional bulk outputs (copied before any scalar output is written).
  std::vector<double> coeffs, cond;
  std::vector<int32_t> optc;
  if (res->coeffs) {
    coeffs.resize((size_t)T * C * K, 0.0);
    std::vector<double> bg((size_t)T * G * K);
    RMC_HIP(ctx, hipMemcpy(bg.data(), a.beta, bg.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int t = 0; t < T; ++t)
      for (int c = 0; c < C; ++c)
        for (int k = 0; k < K; ++k)
          coeffs[((size_t)t * C + c) * K + k] = bg[((size_t)t * G + (G == 1 ? 0 : c)) * K + k];
  }
  if (res->cond_exp) {
    cond.resize((size_t)(T + 1) * N);
    RMC_HIP(ctx, hipMemcpy(cond.data(), d_V, cond.size() * sizeof(double), hipMemcpyDeviceToHost));
    // Time 0 carries the first-period value on every path (the reference
    // assigns it to all paths, SimulateGPU.cu:2314-2321).
    for (int p = 0; p < N; ++p) cond[p] = out[0];
  }
  if (res->opt_control) {
    optc.resize((size_t)T * N);
    RMC_HIP(ctx, hipMemcpy(optc.data(), d_optc, optc.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (!american)
      for (int p = 0; p < N; ++p) optc[p] = (int)out[1];
  }
  if (res->states)
    RMC_HIP(ctx, hipMemcpy(res->states, d_states, n_states * sizeof(double),
                           hipMemcpyDeviceToHost));
  if (res->coeffs) std::memcpy(res->coeffs, coeffs.data(), coeffs.size() * sizeof(double));
  if (res->cond_exp) std::memcpy(res->cond_exp, cond.data(), cond.size() * sizeof(double));
  if (res->opt_control) std::memcpy(res->opt_control, optc.data(), optc.size() * sizeof(int));
  if (res->value) *res->value = out[0];
  if (res->first_control) *res->first_control = (int32_t)out[1];
  if (res->value_sd) *res->value_sd = out[2];
  return ROADMC_OK;
}

extern "C" int roadmc_exp_pv(roadmc_ctx* ctx, const roadmc_uncertainty* unc, int32_t n_steps,
                             double time_step, double rrr, double growth, int64_t n_paths,
                             uint64_t seed, double* mean, double* sd) {
  if (!ctx) return ROADMC_ERR_INVALID;
  if (!unc || !mean || !sd || n_steps < 1 || n_paths < 1 || n_paths > 0x7fffffffLL)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "roadmc_exp_pv: bad arguments");
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  const int N = (int)n_paths;
  void* d = nullptr;
  int rc = rmc_malloc(ctx, &d, (size_t)N * sizeof(double));
  if (rc) return rc;
  roadmc::ExpPvArgs a{};
  a.N = N;
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
  a.tag = rmc_tag(1, 0, RMC_KIND_ROV);
  a.k0 = (uint32_t)(seed & 0xffffffffu);
  a.k1 = (uint32_t)(seed >> 32);
  a.out = static_cast<double*>(d);
  std::vector<double> h(N);
  hipError_t e = roadmc::launch_exp_pv(a, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), d, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess)
    return rmc_fail(ctx, ROADMC_ERR_HIP, std::string("exp_pv: ") + hipGetErrorString(e));
  double s1 = 0.0;
  for (double v : h) s1 += v;
  const double m = s1 / N;
  double s2 = 0.0;
  for (double v : h) s2 += (v - m) * (v - m);
  *mean = m;
  *sd = std::sqrt(s2 / N);
  return ROADMC_OK;
}
