// C-ABI host code for the species-state ROV (rovs_kernels.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.h"
#include "adj_kernels.h"
#include "knn_kernels.h"
#include "mte_common.h"
#include "mte_wide.h"
#include "mte_kernels.h"
#include "rng.h"
#include "roadmc.h"
#include "rov_kernels.h"
#include "rovs_kernels.h"

namespace {

int round_up(int x, int m) { return (x + m - 1) / m * m; }

int validate(roadmc_ctx* ctx, const roadmc_rovs_problem* p, const roadmc_rovs_result* r) {
  auto bad = [&](const char* m) { return rmc_fail(ctx, ROADMC_ERR_INVALID, m); };
  if (!p || !r) return bad("problem or result is NULL");
  if (!r->value) return bad("result.value is NULL");
  // Beyond one workgroup's tile (512 patches) the ALGO4 forward pass keeps
  // the populations in HBM (rovs_forward_wide), in FP32.
  const bool algo6 = p->method == ROADMC_ROVS_ALGO6;
  if (p->n_patches < 1 || p->n_patches > ROADMC_MAX_PATCHES_WIDE)
    return bad("n_patches must be in [1, ROADMC_MAX_PATCHES_WIDE]");
  if (p->n_patches > roadmc::kRovsMaxPatches && (algo6 || p->precision != ROADMC_PREC_FP32))
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED,
                    "more than 512 patches: ALGO4 with FP32 forward products only");
  if (p->n_species < 1 || p->n_species + 1 > ROADMC_KNN_MAX_DIMS || !p->species)
    return bad("n_species must be in [1, 7]");
  if (p->n_controls < 1 || p->n_controls > ROADMC_MAX_CONTROLS || !p->flow)
    return bad("n_controls must be in [1, 8] with flow rates");
  if (p->n_fuels < 0 || p->n_commodities < 0 || p->n_fuels + p->n_commodities > ROADMC_MAX_UNCERTAINTIES)
    return bad("n_fuels + n_commodities must be <= 8");
  if (p->n_fuels + p->n_commodities > 0 && !p->uncertainties) return bad("uncertainties is NULL");
  if (p->n_commodities > 0 && (!p->ore_mean || !p->ore_sd)) return bad("ore composition is NULL");
  if (p->n_vehicles < 0 || p->n_vehicles > roadmc::kRovsMaxFuel ||
      (p->n_vehicles > 0 && (!p->fuel_cost || !p->fuel_idx)))
    return bad("n_vehicles must be in [0, 8] with fuel costs and indices");
  for (int v = 0; v < p->n_vehicles; ++v)
    if (p->fuel_idx[v] < 0 || p->fuel_idx[v] >= std::max(p->n_fuels, 1) ||
        (p->n_fuels == 0))
      return bad("fuel_idx out of range");
  if (p->n_years < 3) return bad("n_years must be >= 3");
  if (p->n_paths < 1 || p->n_paths > (1 << 24)) return bad("n_paths must be in [1, 2^24]");
  if (p->dim_res < 2) return bad("dim_res must be >= 2");
  long long G = 1;
  for (int d = 0; d <= p->n_species; ++d) {
    G *= p->dim_res;
    if (G > (1LL << 24)) return bad("dim_res^(n_species+1) must be <= 2^24");
  }
  if (p->method != 0 && p->method != ROADMC_ROVS_ALGO4 && p->method != ROADMC_ROVS_ALGO6)
    return bad("method must be ROADMC_ROVS_ALGO4 or ROADMC_ROVS_ALGO6");
  if (p->precision == ROADMC_PREC_FP64)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED,
                    "the species-state forward runs FP32 or split-bf16 products");
  if (p->precision != ROADMC_PREC_FP32 && p->precision != ROADMC_PREC_BF16X3)
    return bad("unknown precision");
  // The ALGO4 value moves by about 0.7% between FP32 and split-bf16 forward
  // products at 100 000 paths, more than its FP32 sampling error explains;
  // until that is bounded ALGO4 runs FP32 products only.
  if (p->precision == ROADMC_PREC_BF16X3 && p->method != ROADMC_ROVS_ALGO6)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED,
                    "ALGO4 runs FP32 forward products only; split-bf16 is ALGO6's option");
  if (p->checkpoint_years < 0) return bad("checkpoint_years must be >= 0");
  if (!(p->time_step > 0) || !(p->profit_scaling > 0) || !std::isfinite(p->rrr) ||
      !std::isfinite(p->hysteresis) || !std::isfinite(p->unit_cost) ||
      !std::isfinite(p->unit_revenue))
    return bad("time_step, profit_scaling > 0 and finite scalars required");
  for (int s = 0; s < p->n_species; ++s) {
    const roadmc_rovs_species& sp = p->species[s];
    if (!sp.init_pop || !sp.capacity || !sp.trans || !sp.surv) return bad("species array is NULL");
    for (int i = 0; i < p->n_patches; ++i)
      if (!(sp.capacity[i] > 0) || !std::isfinite(sp.init_pop[i]))
        return bad("capacities must be positive, initial populations finite");
  }
  return ROADMC_OK;
}

// One species' forward pass beyond the LDS tile: the populations of every
// path in HBM ([N][Pp], Pp = P rounded up to 128), and per year the growth-
// rate step, every control's product with the drawn control's growth
// (rovs_wide_year_kernel) and the surviving fractions and totals
// (rovs_wide_tally_kernel), into the tiled pass's total_pops [T+1][N] and
// aars [T][C][N].
int rovs_forward_wide(roadmc_ctx* ctx, const roadmc_rovs_problem* pr, int sp, const int* controls,
                      float* total_pops, float* aars, uint32_t k0, uint32_t k1) {
  const roadmc_rovs_species& sc = pr->species[sp];
  const int P = pr->n_patches, C = pr->n_controls, T = pr->n_years, N = (int)pr->n_paths;
  const int Pp = round_up(P, roadmc::kWideBlock), nb = Pp / roadmc::kWideBlock;
  auto& W = ctx->wide_bufs;
  const size_t plane = (size_t)N * Pp, mt_elems = (size_t)C * Pp * Pp;
  int rc;
  if ((rc = rmc_buf_ensure(ctx, W.mt, mt_elems * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, W.cols, 4 * (size_t)Pp * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, W.s0, plane * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, W.s1, plane * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, W.grm, (size_t)N * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, W.rs, (size_t)nb * C * N * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, W.tot, (size_t)nb * N * sizeof(float))))
    return rc;
  hipStream_t s = ctx->stream;
  // The previous user of the pinned staging is done before it is refilled.
  RMC_HIP(ctx, hipStreamSynchronize(s));
  const size_t mt_bytes = mt_elems * sizeof(float), cols_bytes = 4 * (size_t)Pp * sizeof(float);
  if ((rc = rmc_pinned_ensure(ctx, mt_bytes + cols_bytes + (size_t)N * sizeof(float)))) return rc;
  float* mt_h = static_cast<float*>(ctx->pinned);
  float* cols_h = reinterpret_cast<float*>(static_cast<char*>(ctx->pinned) + mt_bytes);
  float* tot0_h = reinterpret_cast<float*>(static_cast<char*>(ctx->pinned) + mt_bytes + cols_bytes);
  // mt[c][j][k] = trans o surv_c (origin k, destination j), as the tiled
  // pass packs it (the FP64 product rounded to binary32).
  for (int c = 0; c < C; ++c) {
    const double* sv = sc.surv + (size_t)c * P * P;
    float* m = mt_h + (size_t)c * Pp * Pp;
    for (int j = 0; j < Pp; ++j)
      for (int k = 0; k < Pp; ++k)
        m[(size_t)j * Pp + k] =
            (j < P && k < P) ? (float)(sc.trans[(size_t)j * P + k] * sv[(size_t)j * P + k]) : 0.f;
  }
  double tot = 0.0;
  for (int j = 0; j < Pp; ++j) {
    const bool in = j < P;
    cols_h[j] = in ? (float)sc.capacity[j] : 1.f;
    cols_h[Pp + j] = in ? (float)(1.0 / sc.capacity[j]) : 1.f;
    cols_h[2 * Pp + j] = 0.f;
    cols_h[3 * Pp + j] = in ? (float)sc.init_pop[j] : 0.f;
    if (in) tot += (double)(float)sc.init_pop[j];
  }
  std::fill(tot0_h, tot0_h + N, (float)tot);
  RMC_HIP(ctx, hipMemcpyAsync(W.mt.ptr, mt_h, mt_bytes, hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(W.cols.ptr, cols_h, cols_bytes, hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(total_pops, tot0_h, (size_t)N * sizeof(float),
                              hipMemcpyHostToDevice, s));
  const double* prm = sc.params;
  roadmc::MteWideArgs g{};
  g.cols = static_cast<const float*>(W.cols.ptr);
  g.grm = static_cast<float*>(W.grm.ptr);
  g.P = P;
  g.Pp = Pp;
  g.N = N;
  g.law = roadmc::kLawRefGpu;  // the init pass then sets the growth-rate state
  g.dt = (float)pr->time_step;
  g.mu = (float)prm[1];
  g.sd = (float)prm[2];
  g.rev = (float)prm[4];
  g.jmp = (float)prm[5];
  g.jprob = (float)prm[6];
  g.tag1 = rmc_tag(pr->road_id, (uint32_t)sp, RMC_KIND_PATH);
  g.key0 = k0;
  g.key1 = k1;
  float* cur = static_cast<float*>(W.s0.ptr);
  float* nxt = static_cast<float*>(W.s1.ptr);
  RMC_HIP(ctx, roadmc::launch_mte_wide_init(g, (float)prm[0], cur, s));
  roadmc::RovsWideArgs a{};
  a.mt = static_cast<const float*>(W.mt.ptr);
  a.cols = static_cast<const float*>(W.cols.ptr);
  a.grm = static_cast<const float*>(W.grm.ptr);
  a.rs_part = static_cast<float*>(W.rs.ptr);
  a.tot_part = static_cast<float*>(W.tot.ptr);
  a.P = P;
  a.Pp = Pp;
  a.N = N;
  a.C = C;
  a.lv = (float)prm[7];
  a.tag0 = rmc_tag(pr->road_id, (uint32_t)sp, RMC_KIND_PATCH);
  a.key0 = k0;
  a.key1 = k1;
  for (int t = 0; t < T; ++t) {
    g.year = t;
    RMC_HIP(ctx, roadmc::launch_mte_wide_grm(g, s));
    a.year = t;
    a.s_in = cur;
    a.s_out = nxt;
    a.controls = controls + (size_t)t * N;
    RMC_HIP(ctx, roadmc::launch_rovs_wide_year(a, s));
    RMC_HIP(ctx, roadmc::launch_rovs_wide_tally(a.rs_part, a.tot_part, nb, C, N,
                                                total_pops + (size_t)t * N,
                                                total_pops + (size_t)(t + 1) * N,
                                                aars + (size_t)t * C * N, s));
    std::swap(cur, nxt);
  }
  // The pinned staging stays in use until the uploads above are done.
  RMC_HIP(ctx, hipStreamSynchronize(s));
  return ROADMC_OK;
}

}  // namespace

extern "C" int roadmc_rovs_run(roadmc_ctx* ctx, const roadmc_rovs_problem* pr,
                               roadmc_rovs_result* res) {
  if (!ctx) return ROADMC_ERR_INVALID;
  int rc = validate(ctx, pr, res);
  if (rc) return rc;
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  const int P = pr->n_patches, S = pr->n_species, C = pr->n_controls, T = pr->n_years;
  const int N = (int)pr->n_paths, U = pr->n_fuels + pr->n_commodities, D = S + 1, R = pr->dim_res;
  // Split-bf16 products contract over Pp (an even count of 16-k chunks).
  const bool split = pr->precision == ROADMC_PREC_BF16X3;
  const bool wide = P > roadmc::kRovsMaxPatches;
  const int Pp = round_up(P, 32), Pk = split ? Pp : round_up(P, 16);
  size_t G = 1;
  for (int d = 0; d < D; ++d) G *= (size_t)R;
  const size_t gsz = (size_t)D * R + G;
  const int nblk = (N + roadmc::kRovsScanBlock - 1) / roadmc::kRovsScanBlock;
  const uint32_t k0 = (uint32_t)(pr->seed & 0xffffffffu), k1 = (uint32_t)(pr->seed >> 32);
  const bool algo6 = pr->method == ROADMC_ROVS_ALGO6;
  // The recomputation keeps every species' tile of 32 paths in LDS, starts
  // from a checkpoint of the forward populations every ck years (every year
  // while the checkpoints fit in 16 GiB) and reads every step's grids.
  constexpr size_t kLdsMax = 160 * 1024, kCkptBudget = (size_t)16 << 30;
  if (algo6 && roadmc::rovs_recompute_lds(Pp, C, S, split) > kLdsMax)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED,
                    "ALGO6: the species' population tiles exceed the LDS of a workgroup");
  const size_t ck_all = (size_t)S * T * N * Pp * sizeof(float);
  const int ck = !algo6                     ? 1
                 : pr->checkpoint_years > 0 ? std::min(pr->checkpoint_years, T)
                                            : (int)std::max<size_t>(1, (ck_all + kCkptBudget - 1) /
                                                                           kCkptBudget);
  const int nck = (T + ck - 1) / ck;

  // Device workspace.
  auto& B = ctx->rovs_bufs;
  const size_t Uc = std::max(U, 1);
  struct Need {
    rmc_devbuf* b;
    size_t bytes;
  } need[] = {
      {&B.mpack, wide ? 8 : (size_t)S * C * Pp * Pk * sizeof(float)},
      {&B.cols, wide ? 8 : (size_t)S * 3 * Pp * sizeof(float)},
      {&B.prices, (size_t)(T + 1) * Uc * N * sizeof(double)},
      {&B.controls, (size_t)T * N * sizeof(int)},
      {&B.tp, (size_t)S * (T + 1) * N * sizeof(float)},
      {&B.aars, (size_t)S * T * C * N * sizeof(float)},
      {&B.profit, (size_t)T * N * sizeof(double)},
      {&B.V, (size_t)T * N * sizeof(double)},
      {&B.optc, (size_t)T * N * sizeof(int)},
      {&B.adj, res->adj_pops ? (size_t)T * S * N * sizeof(double) : 8},
      {&B.flag, (size_t)N * sizeof(int)},
      {&B.blk, (size_t)2 * nblk * C * sizeof(int) + 64},
      {&B.xs, (size_t)C * N * D * sizeof(double)},
      {&B.ys, (size_t)C * N * sizeof(double)},
      {&B.u, (size_t)C * N * D * sizeof(double)},
      {&B.small, (size_t)(64 * D * 2 + D * 2 + D * R + 8) * sizeof(double)},
      {&B.grid, (size_t)C * gsz * sizeof(double)},
      {&B.args, sizeof(roadmc::RovsStepArgs)},
      {&B.ckpt, algo6 ? (size_t)S * nck * N * Pp * sizeof(float) : 8},
      {&B.grm, algo6 ? (size_t)S * T * N * sizeof(float) : 8},
      {&B.rvec, algo6 ? (size_t)S * C * Pp * sizeof(float) : 8},
      {&B.grids, algo6 ? (size_t)T * C * gsz * sizeof(double) : 8},
  };
  for (auto& n : need)
    if ((rc = rmc_buf_ensure(ctx, *n.b, n.bytes))) return rc;
  hipStream_t s = ctx->stream;

  // Packed movement matrices trans o surv_c (FP32, MFMA B order) and columns
  // (the wide forward pass packs its own, species by species).
  std::vector<float> mp(wide ? 0 : (size_t)S * C * Pp * Pk, 0.f), cols(wide ? 0 : (size_t)S * 3 * Pp);
  std::vector<float> rvec(algo6 ? (size_t)S * C * Pp : 0, 0.f);
  std::vector<float> init_total(S);
  for (int sp = 0; sp < (wide ? 0 : S); ++sp) {
    const roadmc_rovs_species& sc = pr->species[sp];
    for (int c = 0; c < C; ++c) {
      float* m = mp.data() + ((size_t)sp * C + c) * Pp * Pk;
      const double* sv = sc.surv + (size_t)c * P * P;
      for (int k = 0; k < P; ++k) {
        double rs = 0.0;
        for (int j = 0; j < P; ++j) {
          const float v = (float)(sc.trans[(size_t)j * P + k] * sv[(size_t)j * P + k]);
          m[roadmc::mte_pack_index(false, Pk, k, j)] = v;
          rs += v;
        }
        // Row sum of origin k, at k's column of the population tile.
        if (algo6) rvec[((size_t)sp * C + c) * Pp + roadmc::rovs_tile_col(k, split)] = (float)rs;
      }
    }
    float* cl = cols.data() + (size_t)sp * 3 * Pp;
    double tot = 0.0;
    for (int j = 0; j < Pp; ++j) {
      const bool in = j < P;
      cl[j] = in ? (float)sc.capacity[j] : 1.f;
      cl[Pp + j] = in ? (float)(1.0 / sc.capacity[j]) : 1.f;
      cl[2 * Pp + j] = in ? (float)sc.init_pop[j] : 0.f;
      if (in) tot += (double)(float)sc.init_pop[j];
    }
    init_total[sp] = (float)tot;
  }
  if (!wide) {
    RMC_HIP(ctx, hipMemcpyAsync(B.mpack.ptr, mp.data(), mp.size() * sizeof(float),
                                hipMemcpyHostToDevice, s));
    if (split)
      RMC_HIP(ctx, roadmc::launch_mte_split_b3(static_cast<float*>(B.mpack.ptr), mp.size(), s));
    RMC_HIP(ctx, hipMemcpyAsync(B.cols.ptr, cols.data(), cols.size() * sizeof(float),
                                hipMemcpyHostToDevice, s));
  }
  if (algo6)
    RMC_HIP(ctx, hipMemcpyAsync(B.rvec.ptr, rvec.data(), rvec.size() * sizeof(float),
                                hipMemcpyHostToDevice, s));

  // Prices and drawn controls (the ROV forward pass, independent prices).
  roadmc::RovForwardArgs fa{};
  fa.N = N;
  fa.T = T;
  fa.U = U;
  fa.C = C;
  fa.dt = pr->time_step;
  for (int u = 0; u < U; ++u) {
    const roadmc_uncertainty& q = pr->uncertainties[u];
    double* pp = fa.par + u * 7;
    pp[0] = q.kind;
    pp[1] = q.x0;
    pp[2] = q.mean;
    pp[3] = q.sd;
    pp[4] = q.reversion;
    pp[5] = q.jump_size;
    pp[6] = q.jump_prob;
    fa.chol[u * roadmc::kRovMaxU + u] = 1.0;
  }
  fa.tag = rmc_tag(pr->road_id, 0, RMC_KIND_ROV);
  fa.k0 = k0;
  fa.k1 = k1;
  fa.states = static_cast<double*>(B.prices.ptr);
  fa.controls = static_cast<int*>(B.controls.ptr);
  RMC_HIP(ctx, roadmc::launch_rov_forward(fa, s));

  // Species populations under the drawn controls.
  for (int sp = 0; sp < S && wide; ++sp)
    if ((rc = rovs_forward_wide(ctx, pr, sp, static_cast<const int*>(B.controls.ptr),
                                static_cast<float*>(B.tp.ptr) + (size_t)sp * (T + 1) * N,
                                static_cast<float*>(B.aars.ptr) + (size_t)sp * T * C * N, k0, k1)))
      return rc;
  for (int sp = 0; sp < (wide ? 0 : S); ++sp) {
    roadmc::RovsForwardArgs a{};
    a.N = N;
    a.T = T;
    a.P = P;
    a.Pp = Pp;
    a.Pk = Pk;
    a.C = C;
    a.mpack = static_cast<const float*>(B.mpack.ptr) + (size_t)sp * C * Pp * Pk;
    a.cols = static_cast<const float*>(B.cols.ptr) + (size_t)sp * 3 * Pp;
    for (int i = 0; i < 8; ++i) a.params[i] = pr->species[sp].params[i];
    a.dt = pr->time_step;
    a.init_total = init_total[sp];
    a.tag0 = rmc_tag(pr->road_id, (uint32_t)sp, RMC_KIND_PATCH);
    a.tag1 = rmc_tag(pr->road_id, (uint32_t)sp, RMC_KIND_PATH);
    a.key0 = k0;
    a.key1 = k1;
    a.controls = static_cast<const int*>(B.controls.ptr);
    a.total_pops = static_cast<float*>(B.tp.ptr) + (size_t)sp * (T + 1) * N;
    a.aars = static_cast<float*>(B.aars.ptr) + (size_t)sp * T * C * N;
    a.ck = 1;
    a.split = split;
    if (algo6) {
      a.ckpt = static_cast<float*>(B.ckpt.ptr) + (size_t)sp * nck * N * Pp;
      a.ck = ck;
      a.grm_out = static_cast<float*>(B.grm.ptr) + (size_t)sp * T * N;
    }
    RMC_HIP(ctx, roadmc::launch_rovs_forward(a, s));
  }

  // Unit profit per path and year.
  roadmc::RovsProfitArgs pa{};
  pa.N = N;
  pa.T = T;
  pa.U = U;
  pa.n_vehicles = pr->n_vehicles;
  pa.n_fuels = pr->n_fuels;
  pa.n_comm = pr->n_commodities;
  pa.unit_cost = pr->unit_cost;
  pa.unit_revenue = pr->unit_revenue;
  for (int v = 0; v < pr->n_vehicles; ++v) {
    pa.fuel_cost[v] = pr->fuel_cost[v];
    pa.fuel_idx[v] = pr->fuel_idx[v];
  }
  for (int c = 0; c < pr->n_commodities; ++c) {
    pa.ore_mean[c] = pr->ore_mean[c];
    pa.ore_sd[c] = pr->ore_sd[c];
  }
  pa.tag = rmc_tag(pr->road_id, 0, RMC_KIND_ORE);
  pa.key0 = k0;
  pa.key1 = k1;
  pa.prices = static_cast<const double*>(B.prices.ptr);
  pa.profit = static_cast<double*>(B.profit.ptr);
  RMC_HIP(ctx, roadmc::launch_rovs_profit(pa, s));

  // Backward induction.
  roadmc::RovsStepArgs h{};
  h.N = N;
  h.T = T;
  h.S = S;
  h.C = C;
  h.D = D;
  h.R = R;
  h.disc = 1.0 / (1.0 + pr->rrr * pr->time_step / 100.0);
  h.hysteresis = pr->hysteresis;
  h.scaling = pr->profit_scaling;
  for (int sp = 0; sp < S; ++sp) h.thr[sp] = pr->species[sp].params[3];
  for (int c = 0; c < C; ++c) h.flow[c] = pr->flow[c];
  h.total_pops = static_cast<const float*>(B.tp.ptr);
  h.aars = static_cast<const float*>(B.aars.ptr);
  h.profit = static_cast<const double*>(B.profit.ptr);
  h.controls = static_cast<const int*>(B.controls.ptr);
  h.V = static_cast<double*>(B.V.ptr);
  h.optc = static_cast<int*>(B.optc.ptr);
  h.adj_pops = res->adj_pops ? static_cast<double*>(B.adj.ptr) : nullptr;
  h.flag = static_cast<int*>(B.flag.ptr);
  h.blk_count = static_cast<int*>(B.blk.ptr);
  h.blk_off = h.blk_count + (size_t)nblk * C;
  h.count = h.blk_off + (size_t)nblk * C;
  h.xs = static_cast<double*>(B.xs.ptr);
  h.ys = static_cast<double*>(B.ys.ptr);
  h.u = static_cast<double*>(B.u.ptr);
  h.mm = static_cast<double*>(B.small.ptr);
  h.range = h.mm + 64 * D * 2;
  h.qtab = h.range + D * 2;
  double* d_out = h.qtab + (size_t)D * R;
  h.grid = static_cast<double*>(B.grid.ptr);
  auto* d_args = static_cast<roadmc::RovsStepArgs*>(B.args.ptr);
  RMC_HIP(ctx, hipMemcpyAsync(d_args, &h, sizeof(h), hipMemcpyHostToDevice, s));

  std::vector<double> grids_host;
  if (res->grids) grids_host.assign((size_t)T * C * gsz, 0.0);
  roadmc::RovsRecomputeArgs ra{};
  if (algo6) {
    ra.N = N;
    ra.T = T;
    ra.P = P;
    ra.Pp = Pp;
    ra.Pk = Pk;
    ra.C = C;
    ra.S = S;
    ra.D = D;
    ra.R = R;
    ra.ck = ck;
    ra.split = split;
    ra.mpack = static_cast<const float*>(B.mpack.ptr);
    ra.cols = static_cast<const float*>(B.cols.ptr);
    ra.rvec = static_cast<const float*>(B.rvec.ptr);
    ra.ckpt = static_cast<const float*>(B.ckpt.ptr);
    ra.ckpt_species = (size_t)nck * N * Pp;
    ra.grm = static_cast<const float*>(B.grm.ptr);
    ra.controls = static_cast<const int*>(B.controls.ptr);
    ra.profit = static_cast<const double*>(B.profit.ptr);
    ra.grids = static_cast<const double*>(B.grids.ptr);
    for (int sp = 0; sp < S; ++sp) {
      ra.lv[sp] = (float)pr->species[sp].params[7];
      ra.thr[sp] = pr->species[sp].params[3];
      ra.tag0[sp] = rmc_tag(pr->road_id, (uint32_t)sp, RMC_KIND_PATCH);
    }
    for (int c = 0; c < C; ++c) ra.flow[c] = pr->flow[c];
    ra.disc = h.disc;
    ra.hysteresis = h.hysteresis;
    ra.key0 = k0;
    ra.key1 = k1;
    ra.V = h.V;
    ra.optc = h.optc;
    ra.adj_pops = h.adj_pops;
  }
  RMC_HIP(ctx, roadmc::launch_rovs_terminal(d_args, N, s));
  // Each step's control counts come back through the pinned staging (resized
  // while the stream is idle).  The range kernels read the counts on the
  // device, so they are queued before the host waits for them.
  RMC_HIP(ctx, hipStreamSynchronize(s));
  if ((rc = rmc_pinned_ensure(ctx, roadmc::kRovsMaxC * sizeof(int)))) return rc;
  int* const counts = static_cast<int*>(ctx->pinned);
  for (int t = T - 2; t >= 1; --t) {
    RMC_HIP(ctx, roadmc::launch_rovs_collect(d_args, N, t, s));
    RMC_HIP(ctx, hipMemcpyAsync(counts, h.count, C * sizeof(int), hipMemcpyDeviceToHost, s));
    RMC_HIP(ctx, roadmc::launch_rovs_ranges(d_args, N, D, s));
    RMC_HIP(ctx, roadmc::launch_rovs_normalise(d_args, N, D, R, s));
    RMC_HIP(ctx, hipStreamSynchronize(s));
    std::vector<roadmc::KnnFitArgs> fits;
    for (int c = 0; c < C; ++c) {
      const int n = counts[c];
      // k = 10 (ln n + 1) for more than 3 samples (SimulateGPU.cu:4522-4525).
      int k = n > 3 ? (int)(10.0 * (std::log((double)n) + 1.0)) : 0;
      k = std::min({k, roadmc::kKnnMaxK, n});
      double* vals = h.grid + (size_t)c * gsz + (size_t)D * R;
      if (n < 3 || k < 1) {
        RMC_HIP(ctx, hipMemsetAsync(vals, 0, G * sizeof(double), s));
        continue;
      }
      roadmc::KnnFitArgs ka{};
      ka.n = n;
      ka.d = D;
      ka.R = R;
      ka.k = k;
      ka.clamp = 0;
      ka.cap = std::min(k, (n + roadmc::kKnnThreads - 1) / roadmc::kKnnThreads);
      ka.G = (long long)G;
      ka.u = h.u + (size_t)c * N * D;
      ka.y = h.ys + (size_t)c * N;
      ka.qtab = h.qtab;
      ka.val = vals;
      fits.push_back(ka);
    }
    // The step's fits in one launch per fit kernel.
    RMC_HIP(ctx, roadmc::launch_knn_fit_batch(fits.data(), (int)fits.size(), s));
    if (algo6) {
      double* gt = static_cast<double*>(B.grids.ptr) + (size_t)t * C * gsz;
      RMC_HIP(ctx, hipMemcpyAsync(gt, h.grid, (size_t)C * gsz * sizeof(double),
                                  hipMemcpyDeviceToDevice, s));
      ra.t = t;
      ra.y0 = t / ck * ck;
      RMC_HIP(ctx, roadmc::launch_rovs_recompute(ra, s));
    } else {
      RMC_HIP(ctx, roadmc::launch_rovs_decide(d_args, N, D, t, s));
    }
    if (res->grids)
      RMC_HIP(ctx, hipMemcpyAsync(grids_host.data() + (size_t)t * C * gsz, h.grid,
                                  (size_t)C * gsz * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  RMC_HIP(ctx, roadmc::launch_rovs_first(d_args, d_out, s));
  double out3[3];
  RMC_HIP(ctx, hipMemcpyAsync(out3, d_out, sizeof(out3), hipMemcpyDeviceToHost, s));

  // Optional outputs, staged, then written.
  std::vector<double> ce, adj, up, tp, aa;
  std::vector<int32_t> oc, ctl;
  std::vector<float> tpf, aaf;
  if (res->cond_exp) {
    ce.resize((size_t)T * N);
    RMC_HIP(ctx, hipMemcpyAsync(ce.data(), h.V, ce.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  if (res->opt_control) {
    oc.resize((size_t)T * N);
    RMC_HIP(ctx, hipMemcpyAsync(oc.data(), h.optc, oc.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  }
  if (res->adj_pops) {
    adj.resize((size_t)T * S * N);
    RMC_HIP(ctx, hipMemcpyAsync(adj.data(), h.adj_pops, adj.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  if (res->unit_profit) {
    up.resize((size_t)T * N);
    RMC_HIP(ctx, hipMemcpyAsync(up.data(), h.profit, up.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  if (res->controls) {
    ctl.resize((size_t)T * N);
    RMC_HIP(ctx, hipMemcpyAsync(ctl.data(), h.controls, ctl.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  }
  if (res->total_pops) {
    tpf.resize((size_t)S * (T + 1) * N);
    RMC_HIP(ctx, hipMemcpyAsync(tpf.data(), h.total_pops, tpf.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  if (res->aars) {
    aaf.resize((size_t)S * T * C * N);
    RMC_HIP(ctx, hipMemcpyAsync(aaf.data(), h.aars, aaf.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  RMC_HIP(ctx, hipStreamSynchronize(s));
  if (!std::isfinite(out3[0]) || !std::isfinite(out3[1]))
    return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite value");
  *res->value = out3[0];
  if (res->value_sd) *res->value_sd = out3[1];
  if (res->first_control) *res->first_control = (int32_t)out3[2];
  if (res->cond_exp) std::memcpy(res->cond_exp, ce.data(), ce.size() * sizeof(double));
  if (res->opt_control) std::memcpy(res->opt_control, oc.data(), oc.size() * sizeof(int32_t));
  if (res->adj_pops) std::memcpy(res->adj_pops, adj.data(), adj.size() * sizeof(double));
  if (res->unit_profit) std::memcpy(res->unit_profit, up.data(), up.size() * sizeof(double));
  if (res->controls) std::memcpy(res->controls, ctl.data(), ctl.size() * sizeof(int32_t));
  if (res->total_pops)
    for (size_t i = 0; i < tpf.size(); ++i) res->total_pops[i] = tpf[i];
  if (res->aars)
    for (size_t i = 0; i < aaf.size(); ++i) res->aars[i] = aaf[i];
  if (res->grids) std::memcpy(res->grids, grids_host.data(), grids_host.size() * sizeof(double));
  return ROADMC_OK;
}

// ---------------------------------------------------------------------------
// Smoothed-validity value under a fixed policy and its gradient with respect
// to the survival matrices (no counterpart in the reference).

extern "C" int roadmc_rovs_smooth_grad(roadmc_ctx* ctx, const roadmc_rovs_problem* pr,
                                       const int32_t* policy, double tau, double* value,
                                       double* grad_surv) {
  if (!ctx) return ROADMC_ERR_INVALID;
  double v0 = 0.0;
  roadmc_rovs_result probe{};
  probe.value = &v0;
  int rc = validate(ctx, pr, &probe);
  if (rc) return rc;
  auto bad = [&](const char* m) { return rmc_fail(ctx, ROADMC_ERR_INVALID, m); };
  if (!policy || !value) return bad("policy or value is NULL");
  if (!(tau > 0.0) || !std::isfinite(tau)) return bad("tau must be positive and finite");
  if (pr->n_patches > roadmc::kRovsMaxPatches)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED, "the smoothed gradient takes P <= 512");
  if (pr->precision != ROADMC_PREC_FP32)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED, "the smoothed gradient runs FP32 products");
  const int P = pr->n_patches, S = pr->n_species, C = pr->n_controls, T = pr->n_years;
  const int N = (int)pr->n_paths, U = pr->n_fuels + pr->n_commodities;
  // The contraction counts its (year, path) rows in 32-bit integers.
  if ((long long)T * N > (1LL << 30)) return bad("n_years x n_paths must be <= 2^30");
  for (size_t i = 0; i < (size_t)T * N; ++i)
    if (policy[i] < 0 || policy[i] >= C) return bad("policy controls must be in [0, n_controls)");
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  // Pp a multiple of 64: the gradient contraction works in 64 x 64 blocks.
  const int Pp = round_up(P, 64), Pk = round_up(P, 16);
  const uint32_t k0 = (uint32_t)(pr->seed & 0xffffffffu), k1 = (uint32_t)(pr->seed >> 32);
  const int nblk = (N + 255) / 256, nsplit = 128;
  auto& B = ctx->rovs_bufs;
  auto& G = ctx->rovs_grad_bufs;
  const size_t Uc = std::max(U, 1);
  const bool want_grad = grad_surv != nullptr;
  struct Need {
    rmc_devbuf* b;
    size_t bytes;
  } need[] = {
      {&B.mpack, (size_t)S * C * Pp * Pk * sizeof(float)},
      {&G.mtpack, (size_t)S * C * Pp * Pk * sizeof(float)},
      {&B.rvec, (size_t)S * C * Pp * sizeof(float)},
      {&B.cols, (size_t)S * 3 * Pp * sizeof(float)},
      {&B.prices, (size_t)(T + 1) * Uc * N * sizeof(double)},
      {&B.controls, (size_t)T * N * sizeof(int)},
      {&G.policy, (size_t)T * N * sizeof(int)},
      {&B.tp, (size_t)S * (T + 1) * N * sizeof(float)},
      {&B.aars, (size_t)S * T * C * N * sizeof(float)},
      {&B.profit, (size_t)T * N * sizeof(double)},
      {&B.ckpt, want_grad ? (size_t)S * T * N * Pp * sizeof(float) : 8},
      {&B.grm, want_grad ? (size_t)S * T * N * sizeof(float) : 8},
      {&G.coef, (size_t)S * T * N * sizeof(float)},
      {&G.jpart, (size_t)nblk * sizeof(double)},
      {&G.w, want_grad ? (size_t)T * N * Pp * sizeof(float) : 8},
      {&G.rows, want_grad ? (size_t)T * N * sizeof(int) : 8},
      {&G.blk, want_grad ? ((size_t)2 * ((T * N + 1023) / 1024) * C + 8) * sizeof(int) : 8},
      {&G.gpart, want_grad ? (size_t)nsplit * Pp * Pp * sizeof(double) : 8},
      {&G.grad, want_grad ? (size_t)Pp * Pp * sizeof(double) : 8},
  };
  for (auto& n : need)
    if ((rc = rmc_buf_ensure(ctx, *n.b, n.bytes))) return rc;
  hipStream_t s = ctx->stream;

  // trans o surv_c and its transpose in MFMA B order; the per-patch columns.
  // The row sums of each origin (for the forward pass's surviving fractions),
  // at the origin's column of the population tile.
  std::vector<float> mp((size_t)S * C * Pp * Pk, 0.f), mt((size_t)S * C * Pp * Pk, 0.f);
  std::vector<float> cols((size_t)S * 3 * Pp), init_total(S), rvec((size_t)S * C * Pp, 0.f);
  for (int sp = 0; sp < S; ++sp) {
    const roadmc_rovs_species& sc = pr->species[sp];
    for (int c = 0; c < C; ++c) {
      const size_t off = ((size_t)sp * C + c) * Pp * Pk;
      const double* sv = sc.surv + (size_t)c * P * P;
      for (int k = 0; k < P; ++k) {
        double rs = 0.0;
        for (int j = 0; j < P; ++j) {
          const float v = (float)(sc.trans[(size_t)j * P + k] * sv[(size_t)j * P + k]);
          mp[off + roadmc::mte_pack_index(false, Pk, k, j)] = v;
          mt[off + roadmc::mte_pack_index(false, Pk, j, k)] = v;
          rs += v;
        }
        rvec[((size_t)sp * C + c) * Pp + roadmc::rovs_tile_col(k, false)] = (float)rs;
      }
    }
    float* cl = cols.data() + (size_t)sp * 3 * Pp;
    double tot = 0.0;
    for (int j = 0; j < Pp; ++j) {
      const bool in = j < P;
      cl[j] = in ? (float)sc.capacity[j] : 1.f;
      cl[Pp + j] = in ? (float)(1.0 / sc.capacity[j]) : 1.f;
      cl[2 * Pp + j] = in ? (float)sc.init_pop[j] : 0.f;
      if (in) tot += (double)(float)sc.init_pop[j];
    }
    init_total[sp] = (float)tot;
  }
  RMC_HIP(ctx, hipMemcpyAsync(B.mpack.ptr, mp.data(), mp.size() * sizeof(float),
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(G.mtpack.ptr, mt.data(), mt.size() * sizeof(float),
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(B.rvec.ptr, rvec.data(), rvec.size() * sizeof(float),
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(B.cols.ptr, cols.data(), cols.size() * sizeof(float),
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(G.policy.ptr, policy, (size_t)T * N * sizeof(int),
                              hipMemcpyHostToDevice, s));

  // Prices (the valuation's streams; its drawn controls are not used) and
  // unit profit, as roadmc_rovs_run computes them.
  roadmc::RovForwardArgs fa{};
  fa.N = N;
  fa.T = T;
  fa.U = U;
  fa.C = C;
  fa.dt = pr->time_step;
  for (int u = 0; u < U; ++u) {
    const roadmc_uncertainty& q = pr->uncertainties[u];
    double* pp = fa.par + u * 7;
    pp[0] = q.kind;
    pp[1] = q.x0;
    pp[2] = q.mean;
    pp[3] = q.sd;
    pp[4] = q.reversion;
    pp[5] = q.jump_size;
    pp[6] = q.jump_prob;
    fa.chol[u * roadmc::kRovMaxU + u] = 1.0;
  }
  fa.tag = rmc_tag(pr->road_id, 0, RMC_KIND_ROV);
  fa.k0 = k0;
  fa.k1 = k1;
  fa.states = static_cast<double*>(B.prices.ptr);
  fa.controls = static_cast<int*>(B.controls.ptr);
  RMC_HIP(ctx, roadmc::launch_rov_forward(fa, s));
  roadmc::RovsProfitArgs pa{};
  pa.N = N;
  pa.T = T;
  pa.U = U;
  pa.n_vehicles = pr->n_vehicles;
  pa.n_fuels = pr->n_fuels;
  pa.n_comm = pr->n_commodities;
  pa.unit_cost = pr->unit_cost;
  pa.unit_revenue = pr->unit_revenue;
  for (int v = 0; v < pr->n_vehicles; ++v) {
    pa.fuel_cost[v] = pr->fuel_cost[v];
    pa.fuel_idx[v] = pr->fuel_idx[v];
  }
  for (int c = 0; c < pr->n_commodities; ++c) {
    pa.ore_mean[c] = pr->ore_mean[c];
    pa.ore_sd[c] = pr->ore_sd[c];
  }
  pa.tag = rmc_tag(pr->road_id, 0, RMC_KIND_ORE);
  pa.key0 = k0;
  pa.key1 = k1;
  pa.prices = static_cast<const double*>(B.prices.ptr);
  pa.profit = static_cast<double*>(B.profit.ptr);
  RMC_HIP(ctx, roadmc::launch_rovs_profit(pa, s));

  // Every species' populations under the policy, each year's tile kept.
  std::vector<uint32_t> tag0(S);
  for (int sp = 0; sp < S; ++sp) {
    roadmc::RovsForwardArgs a{};
    a.N = N;
    a.T = T;
    a.P = P;
    a.Pp = Pp;
    a.Pk = Pk;
    a.C = C;
    a.mpack = static_cast<const float*>(B.mpack.ptr) + (size_t)sp * C * Pp * Pk;
    a.cols = static_cast<const float*>(B.cols.ptr) + (size_t)sp * 3 * Pp;
    for (int i = 0; i < 8; ++i) a.params[i] = pr->species[sp].params[i];
    a.dt = pr->time_step;
    a.init_total = init_total[sp];
    tag0[sp] = rmc_tag(pr->road_id, (uint32_t)sp, RMC_KIND_PATCH);
    a.tag0 = tag0[sp];
    a.tag1 = rmc_tag(pr->road_id, (uint32_t)sp, RMC_KIND_PATH);
    a.key0 = k0;
    a.key1 = k1;
    a.controls = static_cast<const int*>(G.policy.ptr);
    a.total_pops = static_cast<float*>(B.tp.ptr) + (size_t)sp * (T + 1) * N;
    a.aars = static_cast<float*>(B.aars.ptr) + (size_t)sp * T * C * N;
    a.ck = 1;
    a.split = 0;
    // Controls fixed by the policy: the dots' form (surviving fractions from
    // the row sums, products only for the controls a tile takes).
    a.rvec = static_cast<const float*>(B.rvec.ptr) + (size_t)sp * C * Pp;
    if (want_grad) {
      a.ckpt = static_cast<float*>(B.ckpt.ptr) + (size_t)sp * T * N * Pp;
      a.grm_out = static_cast<float*>(B.grm.ptr) + (size_t)sp * T * N;
    }
    RMC_HIP(ctx, roadmc::launch_rovs_forward(a, s));
  }

  // The value and the coefficients dJ_p / dA.
  roadmc::RovsSmoothArgs va{};
  va.N = N;
  va.T = T;
  va.S = S;
  va.C = C;
  va.disc = 1.0 / (1.0 + pr->rrr * pr->time_step / 100.0);
  va.tau = tau;
  for (int sp = 0; sp < S; ++sp) va.thr[sp] = pr->species[sp].params[3];
  for (int c = 0; c < C; ++c) va.flow[c] = pr->flow[c];
  va.total_pops = static_cast<const float*>(B.tp.ptr);
  va.aars = static_cast<const float*>(B.aars.ptr);
  va.profit = static_cast<const double*>(B.profit.ptr);
  va.policy = static_cast<const int*>(G.policy.ptr);
  va.coef = static_cast<float*>(G.coef.ptr);
  va.jpart = static_cast<double*>(G.jpart.ptr);
  RMC_HIP(ctx, roadmc::launch_rovs_smooth_value(va, s));
  std::vector<double> jpart(nblk);
  RMC_HIP(ctx, hipMemcpyAsync(jpart.data(), G.jpart.ptr, nblk * sizeof(double),
                              hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipStreamSynchronize(s));
  double jsum = 0.0;
  for (double v : jpart) jsum += v;
  if (!std::isfinite(jsum)) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite value");
  *value = jsum / (double)N;
  if (!want_grad) return ROADMC_OK;

  // The (year, path) rows of each control, in order: the contraction of
  // control c runs over its own rows only.
  const int nrows = T * N, nblk_rows = (nrows + 1023) / 1024;
  int* blk_count = static_cast<int*>(G.blk.ptr);
  int* blk_off = blk_count + (size_t)nblk_rows * C;
  int* d_total = blk_off + (size_t)nblk_rows * C;
  RMC_HIP(ctx, roadmc::launch_ctl_rows(static_cast<const int*>(G.policy.ptr), nrows, C, blk_count,
                                       blk_off, d_total, nullptr, nullptr, s, false));
  std::vector<int> total(C), first(C);
  RMC_HIP(ctx, hipMemcpyAsync(total.data(), d_total, C * sizeof(int), hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipStreamSynchronize(s));
  for (int c = 0, run = 0; c < C; ++c) {
    first[c] = run;
    run += total[c];
  }
  if (first[C - 1] + total[C - 1] != nrows)
    return rmc_fail(ctx, ROADMC_ERR_HIP, "control row lists do not cover the rows");
  RMC_HIP(ctx, roadmc::launch_ctl_rows(static_cast<const int*>(G.policy.ptr), nrows, C, blk_count,
                                       blk_off, d_total, first.data(),
                                       static_cast<int*>(G.rows.ptr), s, true));

  // Per species: the reverse sweep, then per control dJ/dM_c = sum over the
  // years and paths under c of n_t^T w_t (rows in the tile's stored order).
  std::vector<double> gm((size_t)Pp * Pp);
  std::vector<int> pos(P);
  for (int k = 0; k < P; ++k) pos[k] = roadmc::rovs_tile_col(k, false);
  for (int sp = 0; sp < S; ++sp) {
    const roadmc_rovs_species& sc = pr->species[sp];
    roadmc::RovsAdjArgs a{};
    a.N = N;
    a.T = T;
    a.P = P;
    a.Pp = Pp;
    a.Pk = Pk;
    a.C = C;
    a.mpack = static_cast<const float*>(B.mpack.ptr) + (size_t)sp * C * Pp * Pk;
    a.mtpack = static_cast<const float*>(G.mtpack.ptr) + (size_t)sp * C * Pp * Pk;
    a.cols = static_cast<const float*>(B.cols.ptr) + (size_t)sp * 3 * Pp;
    a.ckpt = static_cast<const float*>(B.ckpt.ptr) + (size_t)sp * T * N * Pp;
    a.grm = static_cast<const float*>(B.grm.ptr) + (size_t)sp * T * N;
    a.policy = static_cast<const int*>(G.policy.ptr);
    a.coef = static_cast<const float*>(G.coef.ptr) + (size_t)sp * T * N;
    a.lv = (float)sc.params[7];
    a.tag0 = tag0[sp];
    a.key0 = k0;
    a.key1 = k1;
    a.w_out = static_cast<float*>(G.w.ptr);
    RMC_HIP(ctx, roadmc::launch_rovs_adj(a, s));
    for (int c = 0; c < C; ++c) {
      if (total[c] == 0) {  // no path ever under c: no dependence on its survival
        std::fill(gm.begin(), gm.end(), 0.0);
      } else {
        RMC_HIP(ctx, hipMemsetAsync(G.grad.ptr, 0, (size_t)Pp * Pp * sizeof(double), s));
        RMC_HIP(ctx, roadmc::launch_adj_gradm(false, a.ckpt, G.w.ptr, total[c], Pp, nsplit,
                                              static_cast<double*>(G.gpart.ptr),
                                              static_cast<double*>(G.grad.ptr), s,
                                              static_cast<const int*>(G.rows.ptr) + first[c]));
        RMC_HIP(ctx, hipMemcpyAsync(gm.data(), G.grad.ptr, gm.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
        RMC_HIP(ctx, hipStreamSynchronize(s));
      }
      // dJ/dsurv = trans o dJ/dM / N, column-major like surv.
      double* out = grad_surv + ((size_t)sp * C + c) * P * P;
      for (int j = 0; j < P; ++j)
        for (int k = 0; k < P; ++k)
          out[(size_t)j * P + k] =
              sc.trans[(size_t)j * P + k] * gm[(size_t)pos[k] * Pp + j] / (double)N;
    }
  }
  for (size_t i = 0; i < (size_t)S * C * P * P; ++i)
    if (!std::isfinite(grad_surv[i])) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite gradient");
  return ROADMC_OK;
}
