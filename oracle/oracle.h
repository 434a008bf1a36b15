/*
 * CPU oracle for the Monte Carlo road-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker (or
 * the timed CPU baseline).  The product library libroadmc never links it.
 *
 * It restates, in plain C with FP64 (and an FP32 mode), the reference
 * functions on the path:
 *   - Simulator::simulateMTEPath + naturalBirthDeath
 *     (include/main/Simulator.cpp:686-775): the REF_CPU growth law;
 *   - mteKernel (include/gpuCode/SimulateGPU.cu:595-717): the REF_GPU growth
 *     law and the pre-growth endpoint;
 *   - SimulateGPU::dense2Sparse (SimulateGPU.cu:5973-5996): CSC with a drop
 *     tolerance;
 *   - Utility::NormalCDFInverse (include/utilities/Utility.cpp:165-193) and the
 *     penalty of Simulator::simulateMTE (Simulator.cpp:184-212);
 *   - the price SDEs of forwardPathKernel / expPVPath / singlePathValue
 *     (SimulateGPU.cu:256-279, 930-952; include/valuation/Uncertainty.cpp:83-112);
 *   - the regression normal equations and the linear solve of multiQuadReg /
 *     solveLinearSystem (SimulateGPU.cu:6092-6207, 6297-6358), with correct
 *     partial pivoting;
 *   - SpeciesRoadPatches::roadCrossings, computeTransitionProbabilities and
 *     computeSurvivalProbabilities (include/road/SpeciesRoadPatches.cpp:
 *     270-451) with Utility::lineSegmentIntersect (Utility.cpp:3-91);
 *   - the kNN local-linear grid surrogate of buildSurrogateMTECUDA /
 *     buildSurrogateROVCUDA (SimulateGPU.cu:5533-5917) and its multilinear
 *     evaluation (Utility.cpp:253-363);
 *   - the species-state ROV of simulateROVCUDA's ALGO4 (SimulateGPU.cu:
 *     789-1148, 2210-2784, 4401-4998) and ALGO6's forward recomputation
 *     (optimalForwardPaths, :1199-2206, the loop :5011-5444);
 *   - the adjoint and MLP references used by the gradient tests.
 *
 * Random numbers follow the Philox4x32-10 stream contract documented in
 * transportoptimiser_amd/csrc/rng.h; the Philox here is an independent
 * implementation, pinned by the Random123 known-answer vectors.
 *
 * Pinning status of each function is listed in oracle/README.md.
 */
#ifndef ROADMC_ORACLE_H
#define ROADMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10: out = philox(ctr, key). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);

/* Patch-noise normal z(path, year, patch) of the stream contract, FP64
 * transform (fp32 != 0: binary32 transform). */
double orc_patch_normal(uint64_t seed, uint32_t tag, uint32_t path,
                        uint32_t year, uint32_t patch, int fp32);

/* Path-scalar draws of kind 1: z_brownian, z_jumpsize, u_jump. */
void orc_path_scalars(uint64_t seed, uint32_t tag, uint32_t path,
                      uint32_t year, double out[3]);

/* One species of one road.  Matrices are column-major (Eigen default) with
 * entry (i, j) = origin i -> destination j. */
typedef struct {
  int32_t n_patches;
  const double* trans;   /* [P*P] */
  const double* surv;    /* [P*P] or NULL */
  const double* init_pop;
  const double* capacity;
  const double* growth;  /* [P] or NULL */
  double params[8];      /* SimulateGPU.cu:3690-3714 order */
  double noise_sd;
} orc_species;

enum { ORC_LAW_LOGNORMAL = 0, ORC_LAW_REF_CPU = 1, ORC_LAW_REF_GPU = 2 };
enum { ORC_END_POST = 0, ORC_END_PRE = 1 };

/* Simulates paths [path_begin, path_end) of one species.  Writes each path's
 * end total to end_totals[path - path_begin] and, when patch_sum is non-NULL,
 * adds the end vector of every path to patch_sum[P].  Returns 0 on success. */
int orc_mte_paths(const orc_species* sp, uint32_t species_index,
                  uint32_t road_id, int32_t n_years, double time_step,
                  int32_t law, int32_t endpoint, double drop_tol,
                  uint64_t seed, int64_t path_begin, int64_t path_end,
                  int fp32, int n_threads, double* end_totals,
                  double* patch_sum);

/* SimulateGPU::dense2Sparse on a column-major float matrix: keeps entries
 * strictly greater than tol.  Returns nnz. */
int32_t orc_dense_to_csc(const float* dense, int32_t rows, int32_t cols,
                         float tol, float* val, int32_t* row_idx,
                         int32_t* per_col);

/* Utility::NormalCDFInverse (Abramowitz and Stegun 26.2.23).  Returns NAN
 * outside (0, 1). */
double orc_normal_cdf_inverse(double p);

/* Penalty of Simulator::simulateMTE for one species. */
double orc_mte_penalty(double threshold, double end_mean, double end_sd,
                       double confidence_interval, double initial_population,
                       double cost_per_animal);

/* ---- ROV ------------------------------------------------------------- */

/* Price paths of U uncertainties (par[u*7..] = kind, x0, mean, sd, reversion,
 * jump size, jump probability; chol = lower Cholesky factor, U x U) and the
 * random controls, on the ROV Philox stream.  states [(T+1)][U][N],
 * controls [T][N] (may be NULL). */
int orc_rov_forward(int32_t n_unc, const double* par, const double* chol,
                    int32_t n_controls, int32_t n_steps, double dt,
                    int64_t n_paths, uint64_t seed, double* states,
                    int32_t* controls);
/* Expected PV of an uncontrolled MRJD price (SimulateGPU::expPV,
 * Uncertainty::singlePathValue); values [n_paths] (may be NULL: then sd is
 * NaN). */
int orc_exp_pv(const double* par, int32_t n_steps, double dt, double rrr, double growth,
               int64_t n_paths, uint64_t seed, double* values, double* mean, double* sd);

typedef struct {
  int32_t n_unc, n_steps, n_controls, deg, mode, value_mode, pooled, itm_only,
      standardize, traffic;
  int64_t n_paths;
  double dt, rrr, discount, hysteresis, unit_cost, strike;
  double weight[8];
  double flow[8];
} orc_rov_spec;

/* Backward induction (see rov_oracle.c).  out = {value, control, SD}. */
int orc_rov_backward(const orc_rov_spec* s, const double* states,
                     const int32_t* controls, double* out, double* V,
                     double* beta);

/* A x = b by Gaussian elimination with partial pivoting (A, b overwritten). */
int orc_solve(int n, double* A, double* b, double* x);

/* ---- Adjoint ---------------------------------------------------------- */

/* J = mean total end population (LOGNORMAL law, post-growth), its gradient
 * with respect to the crossing mortality (column-major [P*P], S = 1 - m) and
 * the growth rates [P]. */
int orc_mte_adjoint(const orc_species* sp, uint32_t road_id, int32_t n_years,
                    int64_t n_paths, uint64_t seed, double* objective,
                    double* grad_m, double* grad_r);

/* ---- Surrogate MLP ------------------------------------------------------ */

/* Mean squared error over n samples and its gradient (grad may be NULL). */
double orc_mlp_loss_grad(int n_in, int H, int n_out, const double* w,
                         const double* X, const double* Y, int n, double* grad);

/* Road -> matrices (road_oracle.c); column-major P x P, row = origin.
 * orc_road_crossings returns the number of in-range pairs, or -1 when out of
 * memory. */
int64_t orc_road_crossings(int32_t P, const double* x, const double* y, double lambda_mean,
                           double lambda_sd, double lambda_mult, int32_t n_points,
                           const double* road_x, const double* road_y, const int32_t* road_type,
                           int32_t drop_last, int32_t* out);
void orc_road_transitions(int32_t P, const double* x, const double* y, const double* area,
                          const double* pref_mean, const double* pref_sd, double pref_mult,
                          double lambda_mean, double lambda_sd, double lambda_mult,
                          int32_t form, double* out);
void orc_road_survival(int32_t P, const int32_t* crossings, int32_t n_controls,
                       const double* flow, double vehicle_width, double animal_length,
                       double animal_speed, double* out);

/* kNN local-linear grid surrogate (knn_oracle.c).  orc_knn_fit returns 0,
 * or a negative value for bad arguments or when out of memory. */
int32_t orc_knn_k(int32_t n, int32_t rule);
int orc_knn_fit(int32_t n, int32_t d, const double* x, const double* y, int32_t dim_res,
                int32_t k, int32_t flags, double* surrogate);
void orc_knn_eval(int32_t d, int32_t dim_res, const double* surrogate, int32_t m,
                  const double* pred, double* out);
/* Fit at the query coordinates qtab [d][dim_res] from samples already in the
 * search coordinates u [n][d]; values into val [dim_res^d]. */
int orc_knn_fit_scaled(int32_t n, int32_t d, const double* u, const double* y, int32_t dim_res,
                       int32_t k, const double* qtab, double* val);

/* Species-state ROV, the reference's ALGO4 (rovs_oracle.c). */
typedef struct {
  int32_t P, S, C, T, R;
  int64_t N;
  const double* init_pop;   /* [S][P] */
  const double* capacity;   /* [S][P] */
  const double* trans;      /* [S][P*P] column-major, origin = row */
  const double* surv;       /* [S][C][P*P] */
  const double* params;     /* [S][8] */
  const double* flow;       /* [C] */
  int32_t n_fuels, n_comm;
  const double* unc;        /* [n_fuels + n_comm][7]: kind, x0, mean, sd, reversion, J, p */
  const double* ore_mean;   /* [n_comm] */
  const double* ore_sd;     /* [n_comm] */
  int32_t n_veh;
  const double* fuel_cost;  /* [n_veh] */
  const int32_t* fuel_idx;  /* [n_veh] */
  double unit_cost, unit_revenue, dt, rrr, scaling, hysteresis;
  uint64_t seed;
  uint32_t road_id;
} orc_rovs_spec;

/* Forward pass of paths [0, N): controls [T][N], total_pops [S][T+1][N],
 * aars [S][T][C][N], profit [T][N]. */
int orc_rovs_forward(const orc_rovs_spec* s, int32_t* controls, double* total_pops, double* aars,
                     double* profit);
/* Backward induction on given forward arrays: V [T][N], optc [T][N],
 * out = {value, SD, control}. */
int orc_rovs_backward(const orc_rovs_spec* s, const int32_t* controls, const double* total_pops,
                      const double* aars, const double* profit, double* V, int32_t* optc,
                      double out[3]);
/* The same with the method named: 4 as above; 6 values every path at each
 * step by recomputing it forward from that step under the fitted policy
 * (realized cash flows, V [t] the realized value).  grids [T][C][gsz] (steps
 * 1 .. T-2 filled) may be NULL. */
int orc_rovs_backward_method(const orc_rovs_spec* s, int32_t method, const int32_t* controls,
                             const double* total_pops, const double* aars, const double* profit,
                             double* V, int32_t* optc, double* grids, double out[3]);
/* Smoothed-validity value J of the species-state valuation under a fixed
 * policy [T][N] (the validity indicator replaced by prod_s sigmoid((A_s -
 * threshold_s) / tau), realized cash flows d^t UP_t (Q_0 + (Q_u - Q_0) v)) and
 * dJ/dsurv [S][C][P*P] (column-major, may be NULL) by reverse mode over each
 * path.  Returns 0, -1 out of memory, -2 bad arguments. */
int orc_rovs_smooth_grad(const orc_rovs_spec* s, const int32_t* policy, double tau, double* J,
                         double* grad_surv);

#ifdef __cplusplus
}
#endif

#endif
