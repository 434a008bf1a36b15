// MTE simulation beyond one workgroup's LDS tile (more than 1024 patches):
// host-visible declarations of mte_wide.hip.
#ifndef ROADMC_MTE_WIDE_H
#define ROADMC_MTE_WIDE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace roadmc {

// Paths and patches per output block of a year's product; Pp (the padded
// patch count of the HBM state and of the matrix) is a multiple of it.
constexpr int kWideBlock = 128;

// One year of one species: S_out = G(S_in M) with the state path-major,
// [N][Pp], and M destination-major, mt[j Pp + k] = M[k][j] (origin k).
struct MteWideArgs {
  const float* mt;    // [Pp][Pp]
  const float* cols;  // [4][Pp]: capacity, 1/capacity, growth rate, initial population
  const float* s_in;  // [N][Pp]
  float* s_out;       // [N][Pp]
  float* grm;         // [N] per-path growth-rate state (REF_GPU law)
  int32_t P, Pp, N, year, law, keep_pre;
  // LOGNORMAL noise SD and half variance, both scaled by log2(e); the
  // reference laws' growth parameters (SimulateGPU.cu:686-709).
  float sig, half_var, dt, mu, sd, rev, jmp, jprob, lv;
  uint32_t tag0, tag1, key0, key1;
};

// State of year 0 (the initial populations in every path) and the growth-rate
// state g0.
hipError_t launch_mte_wide_init(const MteWideArgs& a, float g0, float* s0, hipStream_t s);
// The REF_GPU growth-rate step of year a.year for every path (before the
// year's product).
hipError_t launch_mte_wide_grm(const MteWideArgs& a, hipStream_t s);
// Product and growth of year a.year.
hipError_t launch_mte_wide_year(const MteWideArgs& a, hipStream_t s);
// Tallies of the final state s (layout of the tiled kernels' partials, one
// tile per kWideBlock paths): [n_tiles][3 + Pp] count of totals >= thr,
// sum and sum of squares of the totals, per-patch sums; the totals into
// end_pops when it is not null.
hipError_t launch_mte_wide_tally(const float* s, int N, int Pp, double thr, double* partial,
                                 double* end_pops, hipStream_t st);
// mt (destination-major, Pp x Pp, float) from trans o surv (column-major
// P x P, FP64, surv may be null), entries whose binary32 value is not above
// tol dropped when tol > 0.
hipError_t launch_mte_wide_pack(const double* trans, const double* surv, int P, int Pp,
                                double tol, float* mt, hipStream_t s);

// One year of one species of the species-state valuation's forward pass
// beyond its LDS tile: for every control c the product S_in M_c (the block
// grid's z index), the per-block row sums of each product (for the surviving
// fractions), and for each path the growth of its drawn control's product
// into S_out (SimulateGPU.cu:839-916).
struct RovsWideArgs {
  const float* mt;        // [C][Pp][Pp] destination-major trans o surv_c
  const float* cols;      // [4][Pp]: capacity, 1/capacity, (unused), initial population
  const float* s_in;      // [N][Pp]
  float* s_out;           // [N][Pp]
  const float* grm;       // [N] growth-rate means of the year (after its SDE step)
  const int* controls;    // [N] drawn control of each path this year
  float* rs_part;         // [Pp / kWideBlock][C][N] row sums of each block's products
  float* tot_part;        // [Pp / kWideBlock][N] row sums of the grown block
  int32_t P, Pp, N, C, year;
  float lv;               // local variability of the growth rate
  uint32_t tag0, key0, key1;
};

hipError_t launch_rovs_wide_year(const RovsWideArgs& a, hipStream_t s);
// Surviving fractions aars[c][n] = (sum over blocks of rs_part) / tot_prev[n]
// and the new totals tot_next[n] (sums over blocks of tot_part), in block order.
hipError_t launch_rovs_wide_tally(const float* rs_part, const float* tot_part, int nblk, int C,
                                  int N, const float* tot_prev, float* tot_next, float* aars,
                                  hipStream_t s);

}  // namespace roadmc

#endif
