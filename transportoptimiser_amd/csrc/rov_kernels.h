// Host-visible declarations of the ROV kernels (rov_kernels.hip).
#ifndef ROADMC_ROV_KERNELS_H
#define ROADMC_ROV_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace roadmc {

constexpr int kRovMaxU = 8;   // uncertainties
constexpr int kRovMaxC = 8;   // controls
constexpr int kRovMaxK = 35;  // basis functions

struct RovForwardArgs {
  int N, T, U, C;
  double dt;
  double par[kRovMaxU * 7];          // kind, x0, mean, sd, reversion, J, p
  double chol[kRovMaxU * kRovMaxU];  // lower Cholesky factor of the correlation
  uint32_t tag, k0, k1;
  double* states;  // [(T+1)][U][N]
  int* controls;   // [T][N] or nullptr
};

constexpr int kRovStepPaths = 512;   // paths per block of the step kernel (two per thread)
constexpr int kRovGramPaths = 1024;  // paths per block of the Gram kernel (four per thread)

struct RovStepArgs {
  int N, T, U, C, G, K, E, deg, mode, value_mode, pooled, itm_only, traffic;
  double disc, hysteresis, unit_cost, strike;
  double weight[kRovMaxU];
  double flow[kRovMaxC];
  const double* states;  // [(T+1)][U][N]
  const double* scale;   // [(T+1)][U][2] mean, 1/sd, or nullptr
  const int* controls;   // [T][N]
  double* V;             // [(T+1)][N]
  int* optc;             // [T][N]
  int nblk;              // blocks of the step kernel, ceil(N / kRovStepPaths)
  double* partial;       // [step - 1][ceil(N / kRovGramPaths)][G][E] block sums of the matrices A
  double* ainv;          // [T][G][K][K] inverses of A
  double* bpart;         // [2][nblk][G][K] block sums of the right-hand sides, by step parity
  double* fpart;         // [ceil(N / 256)][3][kRovMaxC] time-0 block sums
  int* ticket;           // [1] arrival counter of the time-0 kernel, zero between launches
  double* beta;          // [T][G][K]
  int* status;           // [T][G]
  double* out;           // [3] value, control, SD of time 0
};

struct ExpPvArgs {
  int N, T;
  double x0, mean, sd, reversion, jump, jump_prob, dt, rrr, growth;
  uint32_t tag, k0, k1;
  double* out;  // [N]
};

hipError_t rov_set_lds_limits();
hipError_t launch_rov_forward(const RovForwardArgs& a, hipStream_t s);
// scale holds T1 U (mean, 1/SD) pairs followed by 16 T1 U doubles of scratch.
hipError_t launch_rov_moments(const double* states, int N, int U, int T1, double* scale,
                              hipStream_t s);
// The kernels below read their parameters from the device copy `pa` of `a`
// (a supplies the launch shapes), so a captured graph of them stays valid
// for any problem of the same shape.
// The matrices A and their inverses of steps t_lo .. t_lo + n_steps - 1, all
// at once (two launches).
hipError_t launch_rov_gram(const RovStepArgs& a, const RovStepArgs* pa, int t_lo, int n_steps,
                           hipStream_t s);
// Step t of the backward induction, one launch: with decide_next == 1 the
// coefficients of step t + 1 and its decisions, with decide_next == 2 the
// terminal decisions of step t + 1; then step t's partial right-hand sides.
hipError_t launch_rov_step(const RovStepArgs& a, const RovStepArgs* pa, int t, int decide_next,
                           hipStream_t s);
// The coefficients and decisions of step 1 (terminal when `terminal`), then
// time 0 (firstPeriodInduction) in the last block to arrive.
hipError_t launch_rov_final(const RovStepArgs& a, const RovStepArgs* pa, bool terminal,
                            hipStream_t s);
hipError_t launch_exp_pv(const ExpPvArgs& a, hipStream_t s);

}  // namespace roadmc

#endif
