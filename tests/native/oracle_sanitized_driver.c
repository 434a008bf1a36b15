/*
 * Runs every function of the CPU oracle once on small seeded inputs, for a
 * build with AddressSanitizer and UndefinedBehaviorSanitizer (see
 * tests/test_oracle_sanitized.py).  Exit status 0 when every call succeeds
 * and every output is finite; a sanitizer report aborts the run.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint64_t g_state = 0x9e3779b97f4a7c15ull;

static double urand(void) {
  g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(g_state >> 11) * (1.0 / 9007199254740992.0);
}

static int finite_all(const double* a, size_t n, const char* what) {
  for (size_t i = 0; i < n; ++i)
    if (!isfinite(a[i])) {
      fprintf(stderr, "%s[%zu] is not finite\n", what, i);
      return 0;
    }
  return 1;
}

#define CHECK(cond, msg)                     \
  do {                                       \
    if (!(cond)) {                           \
      fprintf(stderr, "failed: %s\n", msg);  \
      return 1;                              \
    }                                        \
  } while (0)

/* Row-stochastic transition matrix, column-major with (i, j) = origin i ->
 * destination j. */
static void stochastic(int P, double* m) {
  for (int i = 0; i < P; ++i) {
    double s = 0.0;
    for (int j = 0; j < P; ++j) s += (m[(size_t)j * P + i] = 0.1 + urand());
    for (int j = 0; j < P; ++j) m[(size_t)j * P + i] /= s;
  }
}

static int run_mte(void) {
  enum { P = 7, N = 40, T = 5 };
  double trans[P * P], surv[P * P], init[P], cap[P], gr[P], end[N], patch[P];
  stochastic(P, trans);
  for (int i = 0; i < P * P; ++i) surv[i] = 0.7 + 0.3 * urand();
  for (int i = 0; i < P; ++i) {
    cap[i] = 500.0 + 4000.0 * urand();
    init[i] = 0.8 * cap[i];
    gr[i] = 0.02 + 0.08 * urand();
  }
  orc_species sp;
  memset(&sp, 0, sizeof sp);
  sp.n_patches = P;
  sp.trans = trans;
  sp.surv = surv;
  sp.init_pop = init;
  sp.capacity = cap;
  sp.growth = gr;
  sp.noise_sd = 0.1;
  const double params[8] = {0.05, 0.05, 0.02, 0.5, 0.3, 0.1, 0.05, 0.1};
  memcpy(sp.params, params, sizeof params);
  for (int law = 0; law < 3; ++law)
    for (int ep = 0; ep < 2; ++ep)
      for (int fp32 = 0; fp32 < 2; ++fp32)
        for (int drop = 0; drop < 2; ++drop) {
          memset(patch, 0, sizeof patch);
          CHECK(orc_mte_paths(&sp, 1, 3, T, 1.0, law, ep, drop ? 1e-3 : 0.0, 99, 5, 5 + N, fp32, 2,
                              end, patch) == 0,
                "orc_mte_paths");
          if (!finite_all(end, N, "end_totals") || !finite_all(patch, P, "patch_sum")) return 1;
        }
  float dense[P * P], val[P * P];
  int32_t rows[P * P], per_col[P];
  for (int i = 0; i < P * P; ++i) dense[i] = (float)(trans[i] * surv[i]);
  const int32_t nnz = orc_dense_to_csc(dense, P, P, 1e-5f, val, rows, per_col);
  CHECK(nnz > 0 && nnz <= P * P, "orc_dense_to_csc");
  CHECK(isfinite(orc_normal_cdf_inverse(0.975)) && isnan(orc_normal_cdf_inverse(1.5)),
        "orc_normal_cdf_inverse");
  CHECK(isfinite(orc_mte_penalty(0.5, 1000.0, 50.0, 0.95, 2000.0, 10.0)), "orc_mte_penalty");
  double obj, gm[P * P], grr[P];
  sp.surv = surv;
  CHECK(orc_mte_adjoint(&sp, 2, T, N, 7, &obj, gm, grr) == 0, "orc_mte_adjoint");
  if (!finite_all(&obj, 1, "objective") || !finite_all(gm, P * P, "grad_m") ||
      !finite_all(grr, P, "grad_r"))
    return 1;
  return 0;
}

static int run_rov(void) {
  enum { U = 2, T = 6, N = 300, C = 2, K = 35 };
  const double par[U * 7] = {1, 1.0, 1.2, 0.2, 0.5, 0.0, 0.0, 0, 1.0, 1.0, 0.15, 0.4, 0.2, 0.1};
  const double chol[U * U] = {1.0, 0.0, 0.3, 0.9539392014169456};
  double* states = malloc(sizeof(double) * (T + 1) * U * N);
  int32_t* controls = malloc(sizeof(int32_t) * T * N);
  double* V = malloc(sizeof(double) * (T + 1) * N);
  double* beta = malloc(sizeof(double) * T * C * K);
  CHECK(states && controls && V && beta, "malloc");
  CHECK(orc_rov_forward(U, par, chol, C, T, 1.0, N, 11, states, controls) == 0,
        "orc_rov_forward");
  if (!finite_all(states, (size_t)(T + 1) * U * N, "states")) return 1;
  for (int mode = 0; mode < 2; ++mode)
    for (int value_mode = 0; value_mode < 2; ++value_mode) {
      orc_rov_spec s;
      memset(&s, 0, sizeof s);
      s.n_unc = U;
      s.n_steps = T;
      s.n_controls = C;
      s.deg = 2;
      s.mode = mode;
      s.value_mode = value_mode;
      s.itm_only = 1;
      s.standardize = 1;
      s.traffic = mode == 0 ? 1 : -1;
      s.n_paths = N;
      s.dt = 1.0;
      s.rrr = 7.0;
      s.hysteresis = 1.0;
      s.unit_cost = 0.5;
      s.strike = 1.1;
      s.weight[0] = -1.0;
      s.weight[1] = 0.5;
      s.flow[0] = 0.0;
      s.flow[1] = 100.0;
      double out[3];
      CHECK(orc_rov_backward(&s, states, controls, out, V, beta) == 0, "orc_rov_backward");
      if (!finite_all(out, 3, "rov out")) return 1;
    }
  double A[9] = {4, 1, 2, 1, 3, 0, 2, 0, 5}, b[3] = {1, 2, 3}, x[3];
  CHECK(orc_solve(3, A, b, x) == 0 && finite_all(x, 3, "solve"), "orc_solve");
  free(states);
  free(controls);
  free(V);
  free(beta);
  return 0;
}

static int run_mlp(void) {
  enum { NI = 3, H = 8, NO = 2, NS = 10 };
  const int nw = NI * H + H + H * H + H + H * NO + NO;
  double* w = malloc(sizeof(double) * nw);
  double* g = malloc(sizeof(double) * nw);
  double X[NS * NI], Y[NS * NO];
  CHECK(w && g, "malloc");
  for (int i = 0; i < nw; ++i) w[i] = urand() - 0.5;
  for (int i = 0; i < NS * NI; ++i) X[i] = urand();
  for (int i = 0; i < NS * NO; ++i) Y[i] = urand();
  const double loss = orc_mlp_loss_grad(NI, H, NO, w, X, Y, NS, g);
  CHECK(isfinite(loss) && finite_all(g, nw, "mlp grad"), "orc_mlp_loss_grad");
  free(w);
  free(g);
  return 0;
}

static int run_road(void) {
  enum { P = 9, NP = 6, C = 3 };
  double x[P], y[P], area[P], pm[P], ps[P], rx[NP], ry[NP], flow[C] = {0.0, 50.0, 200.0};
  int32_t rt[NP - 1] = {0, 0, 1, 0, 2};
  for (int i = 0; i < P; ++i) {
    x[i] = 10.0 * urand();
    y[i] = 10.0 * urand();
    area[i] = 1.0 + urand();
    pm[i] = urand();
    ps[i] = 0.1;
  }
  for (int i = 0; i < NP; ++i) {
    rx[i] = 2.0 * i;
    ry[i] = 10.0 * urand();
  }
  int32_t cross[P * P];
  double tr[P * P], sv[C * P * P];
  for (int drop = 0; drop < 2; ++drop)
    CHECK(orc_road_crossings(P, x, y, 0.5, 0.1, 0.2, NP, rx, ry, rt, drop, cross) >= 0,
          "orc_road_crossings");
  for (int form = 0; form < 2; ++form) {
    orc_road_transitions(P, x, y, area, pm, ps, 0.3, 0.5, 0.1, 0.2, form, tr);
    if (!finite_all(tr, P * P, "transitions")) return 1;
  }
  orc_road_survival(P, cross, C, flow, 2.0, 1.0, 10.0, sv);
  return finite_all(sv, C * P * P, "survival") ? 0 : 1;
}

static int run_knn(void) {
  enum { N = 300, D = 2, R = 6 };
  double x[N * D], y[N], surr[D * R + R * R], pred[5 * D], out[5];
  for (int i = 0; i < N; ++i) {
    x[i * D] = urand();
    x[i * D + 1] = 2.0 * urand();
    y[i] = sin(3.0 * x[i * D]) - x[i * D + 1];
  }
  CHECK(orc_knn_k(N, 0) > 0 && orc_knn_k(N, 1) > 0, "orc_knn_k");
  for (int flags = 0; flags < 4; ++flags) {
    CHECK(orc_knn_fit(N, D, x, y, R, 25, flags, surr) == 0, "orc_knn_fit");
    if (!finite_all(surr, D * R + R * R, "surrogate")) return 1;
  }
  for (int i = 0; i < 5 * D; ++i) pred[i] = 3.0 * urand() - 0.5;
  orc_knn_eval(D, R, surr, 5, pred, out);
  return finite_all(out, 5, "knn eval") ? 0 : 1;
}

static int run_rovs(void) {
  enum { P = 5, S = 2, C = 3, T = 6, N = 60, R = 4, D = S + 1 };
  double init[S * P], cap[S * P], trans[S * P * P], surv[S * C * P * P], params[S * 8];
  for (int s = 0; s < S; ++s) {
    stochastic(P, trans + s * P * P);
    for (int i = 0; i < P; ++i) {
      cap[s * P + i] = 500.0 + 4000.0 * urand();
      init[s * P + i] = 0.8 * cap[s * P + i];
    }
    for (int c = 0; c < C; ++c)
      for (int i = 0; i < P * P; ++i) surv[(s * C + c) * P * P + i] = 1.0 - 0.1 * c * urand();
    const double prm[8] = {5.0, 5.0, 0.5, 0.97 * 0.8 * 2000.0 * P, 0.3, 0.1, 0.05, 0.2};
    memcpy(params + s * 8, prm, sizeof prm);
  }
  const double flow[C] = {0.0, 50.0, 100.0};
  const double unc[2 * 7] = {0, 1.5, 1.5, 0.2, 0.5, 0.1, 0.05, 0, 1.0, 1.0, 0.2, 0.5, 0.0, 0.0};
  const double ore_mean[1] = {0.9}, ore_sd[1] = {0.1}, fuel_cost[1] = {0.2};
  const int32_t fuel_idx[1] = {0};
  orc_rovs_spec sp;
  memset(&sp, 0, sizeof sp);
  sp.P = P;
  sp.S = S;
  sp.C = C;
  sp.T = T;
  sp.R = R;
  sp.N = N;
  sp.init_pop = init;
  sp.capacity = cap;
  sp.trans = trans;
  sp.surv = surv;
  sp.params = params;
  sp.flow = flow;
  sp.n_fuels = 1;
  sp.n_comm = 1;
  sp.unc = unc;
  sp.ore_mean = ore_mean;
  sp.ore_sd = ore_sd;
  sp.n_veh = 1;
  sp.fuel_cost = fuel_cost;
  sp.fuel_idx = fuel_idx;
  sp.unit_cost = 1.0;
  sp.unit_revenue = 2.0;
  sp.dt = 1.0;
  sp.rrr = 7.0;
  sp.scaling = 2.0;
  sp.hysteresis = 0.975;
  sp.seed = 5;
  sp.road_id = 3;
  int32_t* controls = malloc(sizeof(int32_t) * T * N);
  int32_t* optc = malloc(sizeof(int32_t) * T * N);
  double* tp = malloc(sizeof(double) * S * (T + 1) * N);
  double* aars = malloc(sizeof(double) * S * T * C * N);
  double* profit = malloc(sizeof(double) * T * N);
  double* V = malloc(sizeof(double) * T * N);
  const size_t gsz = (size_t)D * R + (size_t)R * R * R;
  double* grids = malloc(sizeof(double) * T * C * gsz);
  CHECK(controls && optc && tp && aars && profit && V && grids, "malloc");
  CHECK(orc_rovs_forward(&sp, controls, tp, aars, profit) == 0, "orc_rovs_forward");
  if (!finite_all(tp, (size_t)S * (T + 1) * N, "total_pops") ||
      !finite_all(aars, (size_t)S * T * C * N, "aars"))
    return 1;
  double out[3];
  CHECK(orc_rovs_backward(&sp, controls, tp, aars, profit, V, optc, out) == 0,
        "orc_rovs_backward");
  if (!finite_all(V, (size_t)T * N, "V") || !finite_all(out, 3, "out")) return 1;
  for (int method = 4; method <= 6; method += 2) {
    CHECK(orc_rovs_backward_method(&sp, method, controls, tp, aars, profit, V, optc, grids, out) ==
              0,
          "orc_rovs_backward_method");
    if (!finite_all(V, (size_t)T * N, "V") || !finite_all(out, 3, "out")) return 1;
  }
  {
    /* the smoothed-validity value and gradient under the drawn controls */
    double J = 0.0;
    double* gs = malloc(sizeof(double) * (size_t)S * C * P * P);
    CHECK(gs != NULL, "malloc");
    CHECK(orc_rovs_smooth_grad(&sp, controls, 0.05 * sp.params[3], &J, gs) == 0,
          "orc_rovs_smooth_grad");
    if (!finite_all(&J, 1, "J") || !finite_all(gs, (size_t)S * C * P * P, "grad_surv")) return 1;
    CHECK(orc_rovs_smooth_grad(&sp, controls, 0.0, &J, gs) == -2, "orc_rovs_smooth_grad tau");
    free(gs);
  }
  free(controls);
  free(optc);
  free(tp);
  free(aars);
  free(profit);
  free(V);
  free(grids);
  return 0;
}

int main(void) {
  uint32_t ctr[4] = {0, 0, 0, 0}, key[2] = {0, 0}, o[4];
  orc_philox4x32_10(ctr, key, o);
  double sc[3];
  orc_path_scalars(1, 5, 2, 3, sc);
  if (!isfinite(orc_patch_normal(1, 4, 2, 3, 4, 0)) || !isfinite(orc_patch_normal(1, 4, 2, 3, 4, 1)) ||
      !finite_all(sc, 3, "path scalars"))
    return 1;
  if (run_mte() || run_rov() || run_mlp() || run_road() || run_knn() || run_rovs()) return 1;
  printf("oracle sanitized run: ok\n");
  return 0;
}
