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
  double* partial;       // [G][E][nblk] block partial sums of one step
  double* sums;          // [G][E] their fixed-order totals
  int nblk;              // assembly blocks
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

size_t rov_assemble_lds(int K);
hipError_t rov_set_lds_limits();
hipError_t launch_rov_forward(const RovForwardArgs& a, hipStream_t s);
hipError_t launch_rov_moments(const double* states, int N, int U, int T1, double* scale,
                              hipStream_t s);
// The step kernels read their parameters from the device copy `pa` of `a`
// (a supplies the launch shapes), so a captured graph of them stays valid
// for any problem of the same shape.
// Fit of step t: normal equations, their reduction and solve.  With
// decide_next >= 0 the assembly first takes the decision of step t + 1
// (decide_next == 2: the terminal rule) for every path.
hipError_t launch_rov_fit(const RovStepArgs& a, const RovStepArgs* pa, int t, int decide_next,
                          hipStream_t s);
hipError_t launch_rov_decide(const RovStepArgs& a, const RovStepArgs* pa, int t, bool terminal,
                             hipStream_t s);
hipError_t launch_rov_first(const RovStepArgs* pa, hipStream_t s);
hipError_t launch_exp_pv(const ExpPvArgs& a, hipStream_t s);

}  // namespace roadmc

#endif
