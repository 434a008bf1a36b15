/*
 * CPU oracle: species-state ROV, the reference's ALGO4.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates, in FP64 and one path at
 * a time:
 *   - forwardPathKernel (include/gpuCode/SimulateGPU.cu:789-956) with
 *     randControls (:720-730): each path's populations moved under its drawn
 *     control, the surviving fraction under every control, the growth-rate
 *     SDE (jump probability from its own field) and n = n'(1 + g (K - n') /
 *     K / 100), g = grMean + z local_variability; prices on the laws of the
 *     ROV oracle;
 *   - computePathStates' unit profit (:985-1002) with the ore composition
 *     drawn per path, year and commodity;
 *   - the ALGO4 backward loop (:4401-4998): allocateXYRegressionData
 *     (:1010-1068), the ranges of computeStateMinMax (:1073-1147), the scaled
 *     kNN search (:4582-4607) with k = 10 (ln n + 1) (:4522-4525), the
 *     local-linear grid fit (multiLocLinReg, as knn_oracle.c), and
 *     backwardInduction's decision (:2333-2784) with multilinear
 *     interpolation of the grids; firstPeriodInduction (:2210-2330);
 *   - ALGO6 (:5011-5444, optimalForwardPaths :1199-2206): after each step's
 *     fit, every path re-simulated from that step under the fitted policy,
 *     its value the realized discounted cash flows.
 * The library's decisions (documented with its API) are followed where the
 * reference is inconsistent.
 */
#pragma GCC optimize("fp-contract=off")

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double rv_u01(uint32_t x) { return ((double)(x >> 9) + 0.5) / 8388608.0; }
static const double kRvTwoPi = 6.283185307179586476925286766559;

static void rv_words(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t tag,
                     uint32_t out[4]) {
  uint32_t ctr[4], key[2];
  ctr[0] = c0;
  ctr[1] = c1;
  ctr[2] = c2;
  ctr[3] = tag;
  key[0] = (uint32_t)(seed & 0xffffffffu);
  key[1] = (uint32_t)(seed >> 32);
  orc_philox4x32_10(ctr, key, out);
}

static void rv_normal4(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t tag,
                       double z[4]) {
  uint32_t x[4];
  double r;
  rv_words(seed, c0, c1, c2, tag, x);
  r = sqrt(-2.0 * log(rv_u01(x[0])));
  z[0] = r * cos(kRvTwoPi * rv_u01(x[1]));
  z[1] = r * sin(kRvTwoPi * rv_u01(x[1]));
  r = sqrt(-2.0 * log(rv_u01(x[2])));
  z[2] = r * cos(kRvTwoPi * rv_u01(x[3]));
  z[3] = r * sin(kRvTwoPi * rv_u01(x[3]));
}

static uint32_t rv_tag(uint32_t road, uint32_t species, uint32_t kind) {
  return (road << 8) | ((species & 63u) << 2) | (kind & 3u);
}

/* Every control's product y_c = n M_c (M_c = trans o surv_c, y[c][j] =
 * sum_k n_k M_c(k, j)) and its total rs[c]. */
static void rv_products(const orc_rovs_spec* s, int sp, const double* n, double* y, double* rs) {
  const int P = s->P, C = s->C;
  const double* tr = s->trans + (size_t)sp * P * P;
  for (int c = 0; c < C; ++c) {
    const double* sv = s->surv + ((size_t)sp * s->C + c) * P * P;
    double r = 0.0;
    for (int j = 0; j < P; ++j) {
      double acc = 0.0;
      for (int k = 0; k < P; ++k) acc += n[k] * (tr[(size_t)j * P + k] * sv[(size_t)j * P + k]);
      y[(size_t)c * P + j] = acc;
      r += acc;
    }
    rs[c] = r;
  }
}

/* Year t of species sp on path p after migration to yc: the growth-rate SDE
 * step (jump probability from its own field), then
 * n = y (1 + g (K - y) / K / 100) with g = grMean + z local_variability.
 * Returns the new total. */
static double rv_grow(const orc_rovs_spec* s, int sp, int64_t p, int t, const double* yc,
                      double* n, double* grm) {
  const int P = s->P;
  const double* prm = s->params + sp * 8;
  const double* K = s->capacity + (size_t)sp * P;
  const uint32_t tag0 = rv_tag(s->road_id, (uint32_t)sp, 0);
  const uint32_t tag1 = rv_tag(s->road_id, (uint32_t)sp, 1);
  double sc[3];
  orc_path_scalars(s->seed, tag1, (uint32_t)p, (uint32_t)t, sc);
  const double jump = sc[2] < prm[6] ? 1.0 : 0.0;
  const double jsz = sc[1] * prm[5] * prm[5] - prm[5] * prm[5] / 2.0;
  const double g0 = *grm;
  *grm = g0 + prm[4] * (prm[1] - g0) * s->dt + g0 * (sc[0] * prm[2]) + (exp(jsz) - 1.0) * g0 * jump;
  double tot = 0.0;
  for (int j = 0; j < P; ++j) {
    const double xv = yc[j];
    const double z = orc_patch_normal(s->seed, tag0, (uint32_t)p, (uint32_t)t, (uint32_t)j, 0);
    const double g = *grm + z * prm[7];
    n[j] = xv * (1.0 + g * (K[j] - xv) / K[j] / 100.0);
    tot += n[j];
  }
  return tot;
}

int orc_rovs_forward(const orc_rovs_spec* s, int32_t* controls, double* total_pops, double* aars,
                     double* profit) {
  const int P = s->P, S = s->S, C = s->C, T = s->T, U = s->n_fuels + s->n_comm;
  const int64_t N = s->N;
  const uint32_t tag_rov = rv_tag(s->road_id, 0, 2), tag_ore = rv_tag(s->road_id, 0, 3);
  double* price = (double*)malloc(sizeof(double) * (size_t)(T + 1) * (U > 0 ? U : 1));
  double* n = (double*)malloc(sizeof(double) * (size_t)P);
  double* y = (double*)malloc(sizeof(double) * (size_t)P * C);
  if (!price || !n || !y) {
    free(price);
    free(n);
    free(y);
    return -1;
  }
  for (int64_t p = 0; p < N; ++p) {
    /* prices and controls (the ROV stream: blocks 0-1 diffusion normals, 2-3
       jump-size normals, 4-5 jump uniforms, 6 the control uniform) */
    double x[8];
    for (int u = 0; u < U; ++u) {
      x[u] = s->unc[u * 7 + 1];
      price[u] = x[u];
    }
    for (int t = 0; t < T; ++t) {
      double z[8], zj[8], uj[8];
      for (int b = 0; 4 * b < U; ++b) {
        double w[4];
        uint32_t r[4];
        rv_normal4(s->seed, (uint32_t)p, (uint32_t)t, (uint32_t)b, tag_rov, w);
        for (int i = 0; i < 4 && 4 * b + i < U; ++i) z[4 * b + i] = w[i];
        rv_normal4(s->seed, (uint32_t)p, (uint32_t)t, (uint32_t)(2 + b), tag_rov, w);
        for (int i = 0; i < 4 && 4 * b + i < U; ++i) zj[4 * b + i] = w[i];
        rv_words(s->seed, (uint32_t)p, (uint32_t)t, (uint32_t)(4 + b), tag_rov, r);
        for (int i = 0; i < 4 && 4 * b + i < U; ++i) uj[4 * b + i] = rv_u01(r[i]);
      }
      for (int u = 0; u < U; ++u) {
        const double* q = s->unc + u * 7;
        const int kind = (int)q[0];
        const double xv = x[u];
        if (kind == 1) {
          x[u] = xv + q[4] * (q[2] - xv) * s->dt + q[3] * sqrt(s->dt) * z[u];
        } else if (kind == 2) {
          x[u] = xv * exp((q[2] - 0.5 * q[3] * q[3]) * s->dt + q[3] * sqrt(s->dt) * z[u]);
        } else {
          const double jump = uj[u] < q[6] ? 1.0 : 0.0;
          const double jsz = zj[u] * q[5] * q[5] - q[5] * q[5] / 2.0;
          x[u] = xv + q[4] * (q[2] - xv) * s->dt + xv * q[3] * z[u] + (exp(jsz) - 1.0) * xv * jump;
        }
        price[(size_t)(t + 1) * U + u] = x[u];
      }
      uint32_t r[4];
      rv_words(s->seed, (uint32_t)p, (uint32_t)t, 6u, tag_rov, r);
      int c = (int)(rv_u01(r[0]) * C);
      if (c >= C) c = C - 1;
      controls[(size_t)t * N + p] = c;
    }
    /* unit profit */
    for (int t = 0; t < T; ++t) {
      double fuel = 0.0, ore = 0.0;
      for (int v = 0; v < s->n_veh; ++v) fuel += s->fuel_cost[v] * price[(size_t)t * U + s->fuel_idx[v]];
      for (int b = 0; 4 * b < s->n_comm; ++b) {
        double w[4];
        rv_normal4(s->seed, (uint32_t)p, (uint32_t)t, (uint32_t)b, tag_ore, w);
        for (int i = 0; i < 4 && 4 * b + i < s->n_comm; ++i) {
          const int c = 4 * b + i;
          ore += (s->ore_mean[c] + s->ore_sd[c] * w[i]) * price[(size_t)t * U + s->n_fuels + c];
        }
      }
      profit[(size_t)t * N + p] = s->unit_cost + fuel - s->unit_revenue * ore;
    }
    /* species */
    for (int sp = 0; sp < S; ++sp) {
      double tot = 0.0, grm = s->params[sp * 8 + 0], rs[8];
      for (int j = 0; j < P; ++j) {
        n[j] = s->init_pop[(size_t)sp * P + j];
        tot += n[j];
      }
      total_pops[((size_t)sp * (T + 1)) * N + p] = tot;
      for (int t = 0; t < T; ++t) {
        const int ctl = controls[(size_t)t * N + p];
        rv_products(s, sp, n, y, rs);
        for (int c = 0; c < C; ++c) aars[(((size_t)sp * T + t) * C + c) * N + p] = rs[c] / tot;
        tot = rv_grow(s, sp, p, t, y + (size_t)ctl * P, n, &grm);
        total_pops[((size_t)sp * (T + 1) + t + 1) * N + p] = tot;
      }
    }
  }
  free(price);
  free(n);
  free(y);
  return 0;
}

static double rv_adj(const orc_rovs_spec* s, const double* tp, const double* aars, int sp, int t,
                     int c, int64_t p) {
  const int64_t N = s->N;
  return tp[((size_t)sp * (s->T + 1) + t) * N + p] *
         aars[(((size_t)sp * s->T + t) * s->C + c) * N + p];
}

/* The decision of one path at one step: each valid control's payoff
 * (control 0 always valid; control c > 0 needs every species' adjusted
 * population adj[sp][c] at its threshold) is Q_c x unit profit plus, with a
 * grid, the discounted continuation interpolated at the highest-flow state x
 * and clamped to <= 0 (backwardInduction, SimulateGPU.cu:2453-2771), without
 * one the terminal payoff min(Q_c x unit profit, 0) (:2375-2452); a control
 * replaces the running best only below hysteresis x best (1 at the terminal
 * step).  Returns the control; *best is its payoff. */
static int rv_decide(const orc_rovs_spec* s, const double* grid, const double* x,
                     const double* adj, double* best_out) {
  const int S = s->S, C = s->C, D = S + 1, R = s->R;
  size_t G = 1;
  for (int d = 0; d < D; ++d) G *= (size_t)R;
  const size_t gsz = (size_t)D * R + G;
  const double disc = 1.0 / (1.0 + s->rrr * s->dt / 100.0);
  const double hyst = grid ? s->hysteresis : 1.0;
  double best = 0.0;
  int bc = 0;
  for (int c = 0; c < C; ++c) {
    int valid = 1;
    for (int sp = 0; sp < S && c > 0; ++sp)
      if (adj[sp * C + c] < s->params[sp * 8 + 3]) valid = 0;
    double pay = 0.0;
    if (valid) {
      if (grid) {
        double cont;
        orc_knn_eval(D, R, grid + (size_t)c * gsz, 1, x, &cont);
        pay = s->flow[c] * x[S] + fmin(cont * disc, 0.0);
      } else {
        pay = fmin(s->flow[c] * x[S], 0.0);
      }
    }
    if (c == 0) {
      best = pay;
    } else if (valid && pay < hyst * best) {
      best = pay;
      bc = c;
    }
  }
  *best_out = best;
  return bc;
}

/* The samples of step t (allocateXYRegressionData, :1010-1068), their
 * ranges (computeStateMinMax, :1073-1147), the scaled kNN coordinates
 * (:4582-4607) and one local-linear grid per control into grid [C][gsz]. */
static int rv_fit_step(const orc_rovs_spec* s, int t, const int32_t* controls, const double* tp,
                       const double* aars, const double* profit, const double* V, double* xs,
                       double* us, double* ys, double* qtab, double* grid) {
  const int S = s->S, C = s->C, D = s->S + 1, R = s->R;
  const int64_t N = s->N;
  size_t G = 1;
  for (int d = 0; d < D; ++d) G *= (size_t)R;
  const size_t gsz = (size_t)D * R + G;
  int64_t cnt[8];
  double lo[8], hi[8];
  for (int c = 0; c < C; ++c) cnt[c] = 0;
  for (int64_t p = 0; p < N; ++p) {
    double x[8];
    int valid = 1;
    for (int sp = 0; sp < S; ++sp) {
      x[sp] = rv_adj(s, tp, aars, sp, t, C - 1, p);
      if (x[sp] < s->params[sp * 8 + 3]) valid = 0;
    }
    x[S] = profit[(size_t)t * N + p];
    const int c = controls[(size_t)t * N + p];
    if (!(valid || c == 0)) continue;
    for (int d = 0; d < D; ++d) xs[((size_t)c * N + cnt[c]) * D + d] = x[d];
    ys[(size_t)c * N + cnt[c]] = V[(size_t)(t + 1) * N + p];
    ++cnt[c];
  }
  for (int d = 0; d < D; ++d) {
    lo[d] = INFINITY;
    hi[d] = -INFINITY;
    for (int c = 0; c < C; ++c) {
      if (c > 0 && cnt[c] <= D + 1) continue;
      for (int64_t j = 0; j < cnt[c]; ++j) {
        lo[d] = fmin(lo[d], xs[((size_t)c * N + j) * D + d]);
        hi[d] = fmax(hi[d], xs[((size_t)c * N + j) * D + d]);
      }
    }
    if (!(lo[d] <= hi[d])) lo[d] = hi[d] = 0.0;
  }
#define RV_SCALE(dd, v)                                                                    \
  (!((hi[dd] - lo[dd]) > 0.0) ? 0.0                                                        \
   : (dd) == D - 1             ? ((v) - lo[dd]) / (s->scaling * (hi[dd] - lo[dd]))         \
                               : ((v) - lo[dd]) / (hi[dd] - lo[dd]))
  for (int c = 0; c < C; ++c)
    for (int64_t j = 0; j < cnt[c]; ++j)
      for (int d = 0; d < D; ++d)
        us[((size_t)c * N + j) * D + d] = RV_SCALE(d, xs[((size_t)c * N + j) * D + d]);
  for (int d = 0; d < D; ++d)
    for (int i = 0; i < R; ++i) {
      const double q = (double)i * (hi[d] - lo[d]) / (double)(R - 1) + lo[d];
      qtab[d * R + i] = RV_SCALE(d, q);
      for (int c = 0; c < C; ++c) grid[(size_t)c * gsz + (size_t)d * R + i] = q;
    }
#undef RV_SCALE
  for (int c = 0; c < C; ++c) {
    const int64_t nc = cnt[c];
    int k = nc > 3 ? (int)(10.0 * (log((double)nc) + 1.0)) : 0;
    if (k > 128) k = 128;
    if (k > nc) k = (int)nc;
    double* val = grid + (size_t)c * gsz + (size_t)D * R;
    if (nc < 3 || k < 1) {
      for (size_t g = 0; g < G; ++g) val[g] = 0.0;
    } else if (orc_knn_fit_scaled((int32_t)nc, D, us + (size_t)c * N * D, ys + (size_t)c * N, R, k,
                                  qtab, val) != 0) {
      return -2;
    }
  }
  return 0;
}

/* The recomputation of path p from step t (optimalForwardPaths,
 * SimulateGPU.cu:1199-2206, the ALGO6 loop :5011-5444): the populations at t
 * re-simulated under the drawn controls, then per year u = t .. T-1 the
 * decision on the recomputed state with the grid of step u (the terminal rule
 * at T-1), the chosen control's cash flow Q_c x unit profit (the terminal
 * payoff at T-1) discounted to t, and the migration and growth under that
 * control.  Returns the realized value; *c_t is the decision at t and x_t the
 * highest-flow state at t. */
static double rv_recompute_path(const orc_rovs_spec* s, int t, int64_t p,
                                const int32_t* controls, const double* profit,
                                const double* grids, double* n /*[S][P]*/, double* y /*[C][P]*/,
                                int* c_t, double* x_t) {
  const int S = s->S, C = s->C, T = s->T, P = s->P, D = S + 1, R = s->R;
  const int64_t N = s->N;
  size_t G = 1;
  for (int d = 0; d < D; ++d) G *= (size_t)R;
  const size_t gsz = (size_t)D * R + G;
  const double disc = 1.0 / (1.0 + s->rrr * s->dt / 100.0);
  double grm[8], rs[8], adj[64], x[8];
  for (int sp = 0; sp < S; ++sp) {
    grm[sp] = s->params[sp * 8 + 0];
    for (int j = 0; j < P; ++j) n[(size_t)sp * P + j] = s->init_pop[(size_t)sp * P + j];
    for (int u = 0; u < t; ++u) {
      rv_products(s, sp, n + (size_t)sp * P, y, rs);
      rv_grow(s, sp, p, u, y + (size_t)controls[(size_t)u * N + p] * P, n + (size_t)sp * P,
              &grm[sp]);
    }
  }
  double value = 0.0, f = 1.0;
  for (int u = t; u < T; ++u) {
    /* states: every control's total after one migration */
    for (int sp = 0; sp < S; ++sp) {
      rv_products(s, sp, n + (size_t)sp * P, y, rs);
      for (int c = 0; c < C; ++c) adj[sp * C + c] = rs[c];
      x[sp] = rs[C - 1];
    }
    x[S] = profit[(size_t)u * N + p];
    double best;
    const int last = u == T - 1;
    const int bc = rv_decide(s, last ? NULL : grids + (size_t)u * C * gsz, x, adj, &best);
    value += f * (last ? best : s->flow[bc] * x[S]);
    f *= disc;
    if (u == t) {
      *c_t = bc;
      for (int d = 0; d < S; ++d) x_t[d] = x[d];
    }
    if (last) break;
    for (int sp = 0; sp < S; ++sp) {
      rv_products(s, sp, n + (size_t)sp * P, y, rs);
      rv_grow(s, sp, p, u, y + (size_t)bc * P, n + (size_t)sp * P, &grm[sp]);
    }
  }
  return value;
}

static int rv_backward(const orc_rovs_spec* s, int method, const int32_t* controls,
                       const double* tp, const double* aars, const double* profit, double* V,
                       int32_t* optc, double* grids_out, double out[3]) {
  const int S = s->S, C = s->C, T = s->T, D = s->S + 1, R = s->R, P = s->P;
  const int64_t N = s->N;
  size_t G = 1;
  for (int d = 0; d < D; ++d) G *= (size_t)R;
  const size_t gsz = (size_t)D * R + G;
  const double disc = 1.0 / (1.0 + s->rrr * s->dt / 100.0);
  const int recompute = method == 6;
  double* xs = (double*)malloc(sizeof(double) * (size_t)C * N * D);
  double* us = (double*)malloc(sizeof(double) * (size_t)C * N * D);
  double* ys = (double*)malloc(sizeof(double) * (size_t)C * N);
  double* grids = (double*)calloc((size_t)T * C * gsz, sizeof(double));
  double* qtab = (double*)malloc(sizeof(double) * (size_t)D * R);
  double* n = (double*)malloc(sizeof(double) * (size_t)S * P);
  double* y = (double*)malloc(sizeof(double) * (size_t)C * P);
  if (!xs || !us || !ys || !grids || !qtab || !n || !y) {
    free(xs);
    free(us);
    free(ys);
    free(grids);
    free(qtab);
    free(n);
    free(y);
    return -1;
  }
  /* terminal step */
  for (int64_t p = 0; p < N; ++p) {
    const int t = T - 1;
    double x[8], adj[64], best;
    for (int sp = 0; sp < S; ++sp)
      for (int c = 0; c < C; ++c) adj[sp * C + c] = rv_adj(s, tp, aars, sp, t, c, p);
    x[S] = profit[(size_t)t * N + p];
    optc[(size_t)t * N + p] = rv_decide(s, NULL, x, adj, &best);
    V[(size_t)t * N + p] = best;
  }
  for (int t = T - 2; t >= 1; --t) {
    double* grid = grids + (size_t)t * C * gsz;
    if (rv_fit_step(s, t, controls, tp, aars, profit, V, xs, us, ys, qtab, grid) != 0) {
      free(xs);
      free(us);
      free(ys);
      free(grids);
      free(qtab);
      free(n);
      free(y);
      return -2;
    }
    for (int64_t p = 0; p < N; ++p) {
      if (recompute) {
        int bc;
        double xt[8];
        V[(size_t)t * N + p] = rv_recompute_path(s, t, p, controls, profit, grids, n, y, &bc, xt);
        optc[(size_t)t * N + p] = bc;
        continue;
      }
      double x[8], adj[64], best;
      for (int sp = 0; sp < S; ++sp) {
        x[sp] = rv_adj(s, tp, aars, sp, t, C - 1, p);
        for (int c = 0; c < C; ++c) adj[sp * C + c] = rv_adj(s, tp, aars, sp, t, c, p);
      }
      x[S] = profit[(size_t)t * N + p];
      optc[(size_t)t * N + p] = rv_decide(s, grid, x, adj, &best);
      V[(size_t)t * N + p] = best;
    }
  }
  if (grids_out) memcpy(grids_out, grids, sizeof(double) * (size_t)T * C * gsz);
  /* time 0 */
  {
    double sums[8], ns[8], up = 0.0, best = 0.0;
    int bc = -1;
    for (int c = 0; c < C; ++c) sums[c] = ns[c] = 0.0;
    for (int64_t p = 0; p < N; ++p) {
      const int c = controls[p];
      sums[c] += V[(size_t)N + p];
      ns[c] += 1.0;
      up += profit[p];
    }
    up /= (double)N;
    for (int c = 0; c < C; ++c) {
      if (ns[c] == 0.0) continue;
      int valid = 1;
      for (int sp = 0; sp < S && c > 0; ++sp)
        if (rv_adj(s, tp, aars, sp, 0, c, 0) < s->params[sp * 8 + 3]) valid = 0;
      if (!valid) continue;
      const double pay = fmin(sums[c] / ns[c] * disc + s->flow[c] * up, 0.0);
      if (bc < 0 || pay < best) {
        best = pay;
        bc = c;
      }
    }
    if (bc < 0) bc = 0;
    double m = 0.0, q = 0.0;
    const double base = s->flow[bc] * up;
    for (int64_t p = 0; p < N; ++p)
      if (controls[p] == bc) m += V[(size_t)N + p] * disc + base;
    const double n = ns[bc] > 0.0 ? ns[bc] : 1.0;
    m /= n;
    for (int64_t p = 0; p < N; ++p)
      if (controls[p] == bc) {
        const double d = V[(size_t)N + p] * disc + base - m;
        q += d * d;
      }
    for (int64_t p = 0; p < N; ++p) {
      V[p] = best;
      optc[p] = bc;
    }
    out[0] = best;
    out[1] = sqrt(q / n);
    out[2] = (double)bc;
  }
  free(xs);
  free(us);
  free(ys);
  free(grids);
  free(qtab);
  free(n);
  free(y);
  return 0;
}

int orc_rovs_backward(const orc_rovs_spec* s, const int32_t* controls, const double* tp,
                      const double* aars, const double* profit, double* V, int32_t* optc,
                      double out[3]) {
  return rv_backward(s, 4, controls, tp, aars, profit, V, optc, NULL, out);
}

int orc_rovs_backward_method(const orc_rovs_spec* s, int32_t method, const int32_t* controls,
                             const double* tp, const double* aars, const double* profit,
                             double* V, int32_t* optc, double* grids, double out[3]) {
  if (method != 4 && method != 6) return -3;
  return rv_backward(s, method, controls, tp, aars, profit, V, optc, grids, out);
}

/* Smoothed-validity value of the species-state valuation under a fixed
 * policy and its gradient with respect to the survival matrices.  Not in the
 * reference: its validity test (every adjusted population at its threshold,
 * SimulateGPU.cu:1046-1054, 2757-2771) is an indicator, so the value's
 * pathwise derivative with respect to mortality vanishes.  Here that
 * indicator becomes prod_s sigmoid((A_s - theta_s) / tau) and the decisions
 * are the given policy, so that
 *   J = (1/N) sum_p sum_{t<T} d^t UP_t [Q_0 + (Q_u - Q_0) prod_s sigma_s,t],
 *   u = policy[t][p], A_s,t = sum_j (n_s,t M_s,u)_j (the adjusted population
 *   under u), d = 1 / (1 + rrr dt / 100),
 * with every species' populations moved under the policy's controls and grown
 * as in orc_rovs_forward (same streams); UP is orc_rovs_forward's unit profit.
 * The gradient is reverse mode over each path: with w_t = lambda_t+1 o G'(x_t)
 * + c_t (G'(x) = 1 + g (K - 2x) / K / 100, c_t = dJ_p/dA_s,t), lambda_t =
 * M_u w_t and dJ/dM_u(k, j) += n_t(k) w_t(j); dJ/dsurv = trans o dJ/dM.
 * grad_surv [S][C][P*P] column-major like surv (may be NULL). */
int orc_rovs_smooth_grad(const orc_rovs_spec* s, const int32_t* policy, double tau, double* J,
                         double* grad_surv) {
  const int P = s->P, S = s->S, C = s->C, T = s->T;
  const int64_t N = s->N;
  if (!(tau > 0.0) || N < 1) return -2;
  const double disc = 1.0 / (1.0 + s->rrr * s->dt / 100.0);
  int32_t* ctl = (int32_t*)malloc(sizeof(int32_t) * (size_t)T * N);
  double* tp = (double*)malloc(sizeof(double) * (size_t)S * (T + 1) * N);
  double* aa = (double*)malloc(sizeof(double) * (size_t)S * T * C * N);
  double* up = (double*)malloc(sizeof(double) * (size_t)T * N);
  double* n = (double*)malloc(sizeof(double) * (size_t)S * (T + 1) * P);  /* n_s,t */
  double* x = (double*)malloc(sizeof(double) * (size_t)S * T * P);        /* moved, pre-growth */
  double* gm = (double*)malloc(sizeof(double) * (size_t)S * T);          /* growth-rate means */
  double* A = (double*)malloc(sizeof(double) * (size_t)S * T);
  double* y = (double*)malloc(sizeof(double) * (size_t)C * P);
  double* lam = (double*)malloc(sizeof(double) * (size_t)P * 2);
  int rc = 0;
  if (!ctl || !tp || !aa || !up || !n || !x || !gm || !A || !y || !lam) {
    rc = -1;
    goto done;
  }
  if ((rc = orc_rovs_forward(s, ctl, tp, aa, up)) != 0) goto done;
  if (grad_surv) memset(grad_surv, 0, sizeof(double) * (size_t)S * C * P * P);
  double jsum = 0.0;
  for (int64_t p = 0; p < N; ++p) {
    /* forward of every species under the policy */
    for (int sp = 0; sp < S; ++sp) {
      double* n0 = n + (size_t)sp * (T + 1) * P;
      double grm = s->params[sp * 8 + 0], rs[8];
      for (int j = 0; j < P; ++j) n0[j] = s->init_pop[(size_t)sp * P + j];
      for (int t = 0; t < T; ++t) {
        const int u = policy[(size_t)t * N + p];
        rv_products(s, sp, n0 + (size_t)t * P, y, rs);
        A[sp * T + t] = rs[u];
        memcpy(x + ((size_t)sp * T + t) * P, y + (size_t)u * P, sizeof(double) * P);
        rv_grow(s, sp, p, t, y + (size_t)u * P, n0 + (size_t)(t + 1) * P, &grm);
        gm[sp * T + t] = grm;
      }
    }
    /* the path's value and the coefficients dJ_p / dA_s,t */
    double f = 1.0, jp = 0.0;
    for (int t = 0; t < T; ++t) {
      const int u = policy[(size_t)t * N + p];
      const double upt = up[(size_t)t * N + p];
      double v = 1.0;
      for (int sp = 0; sp < S; ++sp)
        v *= 1.0 / (1.0 + exp(-(A[sp * T + t] - s->params[sp * 8 + 3]) / tau));
      jp += f * upt * (s->flow[0] + (s->flow[u] - s->flow[0]) * v);
      for (int sp = 0; sp < S; ++sp) {
        const double sg = 1.0 / (1.0 + exp(-(A[sp * T + t] - s->params[sp * 8 + 3]) / tau));
        A[sp * T + t] = f * upt * (s->flow[u] - s->flow[0]) * v * (1.0 - sg) / tau;  /* c_t */
      }
      f *= disc;
    }
    jsum += jp;
    if (!grad_surv) continue;
    /* reverse sweep per species */
    for (int sp = 0; sp < S; ++sp) {
      const double* tr = s->trans + (size_t)sp * P * P;
      const double* K = s->capacity + (size_t)sp * P;
      const double lv = s->params[sp * 8 + 7];
      const uint32_t tag0 = rv_tag(s->road_id, (uint32_t)sp, 0);
      double* l = lam;       /* lambda_t+1 */
      double* w = lam + P;   /* w_t */
      for (int j = 0; j < P; ++j) l[j] = 0.0;
      for (int t = T - 1; t >= 0; --t) {
        const int u = policy[(size_t)t * N + p];
        const double* xt = x + ((size_t)sp * T + t) * P;
        const double* nt = n + ((size_t)sp * (T + 1) + t) * P;
        const double* sv = s->surv + ((size_t)sp * C + u) * P * P;
        double* gs = grad_surv + ((size_t)sp * C + u) * P * P;
        for (int j = 0; j < P; ++j) {
          const double z = orc_patch_normal(s->seed, tag0, (uint32_t)p, (uint32_t)t, (uint32_t)j, 0);
          const double g = gm[sp * T + t] + z * lv;
          w[j] = l[j] * (1.0 + g * (K[j] - 2.0 * xt[j]) / K[j] / 100.0) + A[sp * T + t];
        }
        for (int k = 0; k < P; ++k) {
          double acc = 0.0;
          for (int j = 0; j < P; ++j) {
            const size_t cm = (size_t)j * P + k;
            acc += tr[cm] * sv[cm] * w[j];
            gs[cm] += tr[cm] * nt[k] * w[j] / (double)N;
          }
          l[k] = acc;
        }
      }
    }
  }
  *J = jsum / (double)N;
done:
  free(ctl);
  free(tp);
  free(aa);
  free(up);
  free(n);
  free(x);
  free(gm);
  free(A);
  free(y);
  free(lam);
  return rc;
}
