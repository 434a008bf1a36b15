// Host-visible declarations of the species-state ROV kernels (rovs_kernels.hip).
#ifndef ROADMC_ROVS_KERNELS_H
#define ROADMC_ROVS_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace roadmc {

constexpr int kRovsRows = 32;          // paths per forward workgroup
constexpr int kRovsNT = 2;            // 32-column tiles per forward wave
constexpr int kRovsMaxThreads = 512;  // 8 waves: 16 column tiles, P <= 512
constexpr int kRovsMaxPatches = 512;
constexpr int kRovsMaxC = 8;
constexpr int kRovsMaxS = 7;       // S + 1 state dimensions <= 8
constexpr int kRovsMaxFuel = 8;
constexpr int kRovsMaxComm = 8;
constexpr int kRovsScanBlock = 1024;

// One species' forward pass: populations under the randomly drawn control of
// each path and year, the surviving fraction under every control.
struct RovsForwardArgs {
  int N, T, P, Pp, Pk, C;
  const float* mpack;   // [C][Pp * Pk] trans o surv_c in MFMA B order
  const float* cols;    // [3][Pp]: K, 1/K, initial population (K = 1 on padding)
  double params[8];     // growth current, mean, noise SD, threshold, reversion,
                        // jump size, jump probability, local variability
  double dt;
  float init_total;
  uint32_t tag0, tag1, key0, key1;
  const int* controls;  // [T][N]
  float* total_pops;    // [T+1][N]
  float* aars;          // [T][C][N]
  // For the recomputation (nullptr otherwise): the population tile at the
  // start of every ck-th year, [T / ck rounded up][N][Pp] in the tile's
  // column order, and the growth-rate mean of each year's growth, [T][N].
  float* ckpt;
  int ck;
  float* grm_out;
  // Split-bf16 products (mpack in that order, Pk == Pp); the checkpoints then
  // hold each entry's hi | lo << 16 bits.
  int split;
  // Optional (FP32 only): the row sums of each control's matrix, [C][Pp] in
  // the tile's column order.  When given, the surviving fractions are dots of
  // the year-start tile with them, and a tile forms only the products of the
  // controls its paths take that year (for controls fixed by a policy).
  const float* rvec;
};

size_t rovs_forward_lds(int Pp, int C, bool split, bool dots = false);
hipError_t launch_rovs_forward(const RovsForwardArgs& a, hipStream_t s);

// Unit profit per path and year: unit_cost + sum_v fuel_cost_v price_{fuel_idx_v}
// - unit_revenue sum_c composition_c price_{n_fuels + c}.
struct RovsProfitArgs {
  int N, T, U, n_vehicles, n_fuels, n_comm;
  double unit_cost, unit_revenue;
  double fuel_cost[kRovsMaxFuel];
  int fuel_idx[kRovsMaxFuel];
  double ore_mean[kRovsMaxComm], ore_sd[kRovsMaxComm];
  uint32_t tag, key0, key1;
  const double* prices;  // [(T+1)][U][N]
  double* profit;        // [T][N]
  double* composition;   // [T][n_comm][N] or nullptr
};
hipError_t launch_rovs_profit(const RovsProfitArgs& a, hipStream_t s);

// Per-step arguments of the backward induction (device copy, read by pointer).
struct RovsStepArgs {
  int N, T, S, C, D, R;
  double disc, hysteresis, scaling;
  double thr[kRovsMaxS];
  double flow[kRovsMaxC];
  const float* total_pops;  // [S][T+1][N]
  const float* aars;        // [S][T][C][N]
  const double* profit;     // [T][N]
  const int* controls;      // [T][N]
  double* V;                // [T][N]
  int* optc;                // [T][N]
  double* adj_pops;         // [T][S][N] (highest-flow adjusted populations)
  // per-step scratch
  int* flag;                // [N] control of an included sample, or -1
  int* blk_count;           // [nblk][C]
  int* blk_off;             // [nblk][C]
  int* count;               // [C]
  int* pos;                 // [N] position of the sample within its control
  double* xs;               // [C][N][D] included samples (state order), compacted
  double* ys;               // [C][N]
  double* mm;               // [nblk][D][2] block min / max
  double* range;            // [D][2] global min / max
  double* u;                // [C][N][D] normalised samples
  double* qtab;             // [D][R] normalised query coordinates
  double* grid;             // [C][D*R + R^D] grid coordinates then fitted values
};

hipError_t launch_rovs_terminal(const RovsStepArgs* a, int N, hipStream_t s);
hipError_t launch_rovs_collect(const RovsStepArgs* a, int N, int t, hipStream_t s);
hipError_t launch_rovs_ranges(const RovsStepArgs* a, int N, int D, hipStream_t s);
hipError_t launch_rovs_normalise(const RovsStepArgs* a, int N, int D, int R, hipStream_t s);
hipError_t launch_rovs_decide(const RovsStepArgs* a, int N, int D, int t, hipStream_t s);

// One backward step of the recomputation (ALGO6): every path re-simulated
// from the checkpoint year y0 <= t under its drawn controls up to t, then
// from t to T - 1 under the fitted policy (the decision of each year on the
// recomputed state with that year's grids, the terminal rule at T - 1);
// writes V[t] (the realized value), optc[t] and adj_pops[t].
struct RovsRecomputeArgs {
  int N, T, P, Pp, Pk, C, S, D, R;
  int t, y0, ck;
  const float* mpack;    // [S][C][Pp * Pk]
  const float* cols;     // [S][3][Pp]
  const float* rvec;     // [S][C][Pp] row sums of trans o surv_c, tile column order
  const float* ckpt;     // [S][nck][N][Pp]
  size_t ckpt_species;   // floats per species in ckpt
  const float* grm;      // [S][T][N]
  const int* controls;   // [T][N]
  const double* profit;  // [T][N]
  const double* grids;   // [T][C][D R + R^D]
  float lv[kRovsMaxS];
  double thr[kRovsMaxS];
  double flow[kRovsMaxC];
  double disc, hysteresis;
  uint32_t tag0[kRovsMaxS];
  uint32_t key0, key1;
  double* V;
  int* optc;
  double* adj_pops;      // [T][S][N] or nullptr
  int split;             // split-bf16 products, as the forward pass ran them
};
size_t rovs_recompute_lds(int Pp, int C, int S, bool split);

// Smoothed-validity value under a fixed policy (no counterpart in the
// reference): per path, the discounted cash flows d^t UP_t (Q_0 + (Q_u - Q_0)
// prod_s sigmoid((A_s,t - thr_s) / tau)), A_s,t = total_pops x aars of the
// policy's control u, summed per block in a fixed order into jpart
// [ceil(N / 256)]; coef [S][T][N] = dJ_p / dA_s,t.
struct RovsSmoothArgs {
  int N, T, S, C;
  double disc, tau;
  double thr[kRovsMaxS];
  double flow[kRovsMaxC];
  const float* total_pops;  // [S][T+1][N] under the policy
  const float* aars;        // [S][T][C][N]
  const double* profit;     // [T][N]
  const int* policy;        // [T][N]
  float* coef;              // [S][T][N]
  double* jpart;            // [ceil(N / 256)]
};
hipError_t launch_rovs_smooth_value(const RovsSmoothArgs& a, hipStream_t s);

// Reverse sweep of one species' populations under the policy, years T-1 .. 0,
// one 32-path tile per workgroup: x_t = n_t M_u (every control's product, the
// path's selected), w_t = lambda_t+1 o G'(x_t) + coef_t on the patches (zero
// on padding), lambda_t = w_t M_u^T; w_t goes to w_out for dJ/dM_c = sum over
// (t, paths with u = c) of n_t^T w_t.  FP32 products on the forward pass's
// packed operands.
struct RovsAdjArgs {
  int N, T, P, Pp, Pk, C;
  const float* mpack;     // [C][Pp * Pk] trans o surv_c, MFMA B order
  const float* mtpack;    // [C][Pp * Pk] its transpose, same order
  const float* cols;      // [3][Pp]: K, 1/K, initial population
  const float* ckpt;      // [T][N][Pp] populations entering each year, tile order
  const float* grm;       // [T][N] growth-rate means of each year's growth
  const int* policy;      // [T][N]
  const float* coef;      // [T][N] this species' dJ_p / dA_t
  float lv;               // local variability
  uint32_t tag0, key0, key1;
  float* w_out;           // [T][N][Pp]
};
size_t rovs_adj_lds(int Pp);
hipError_t launch_rovs_adj(const RovsAdjArgs& a, hipStream_t s);

// Lists of the n rows of ctl [n] by control, stable.  First call (scatter
// false): blk_count and blk_off [ceil(n / 1024)][C] and the totals [C].  Second
// call (scatter true): with host_off [C] the start of each control's list in
// rows [n], each row's index at its place.
hipError_t launch_ctl_rows(const int* ctl, int n, int C, int* blk_count, int* blk_off,
                           int* total, const int* host_off, int* rows, hipStream_t s,
                           bool scatter);
// Position of patch k within a row of the population tile.
int rovs_tile_col(int k, bool split);
hipError_t launch_rovs_recompute(const RovsRecomputeArgs& a, hipStream_t s);
// out (device): {value, SD, control}.
hipError_t launch_rovs_first(const RovsStepArgs* a, double* out, hipStream_t s);

}  // namespace roadmc

#endif
