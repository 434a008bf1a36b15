/*
 * libroadmc: MI355X (gfx950) Monte Carlo road evaluation, C ABI.
 *
 * The entry points replace the reference's GPU dispatch functions
 * (include/gpuCode/SimulateGPU.h) over plain data: the reference takes Eigen
 * objects and a Simulator object graph; this ABI takes the same arrays as
 * pointers and sizes.  Inputs and outputs are host memory owned by the caller;
 * the context owns device memory and reuses it across calls.  Every function
 * returns ROADMC_OK (0) or an error code and, on error, writes no output;
 * roadmc_last_error() returns the message.
 *
 * Random streams are Philox4x32-10 keyed by the problem's seed; the layout is
 * documented in transportoptimiser_amd/csrc/rng.h.  Results are deterministic
 * for a given problem.
 */
#ifndef ROADMC_H
#define ROADMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ROADMC_OK 0
#define ROADMC_ERR_INVALID 1     /* bad argument or size */
#define ROADMC_ERR_HIP 2         /* HIP runtime error */
#define ROADMC_ERR_OOM 3         /* device allocation failed */
#define ROADMC_ERR_NUMERIC 4     /* singular system, non-finite input */
#define ROADMC_ERR_UNSUPPORTED 5 /* valid request this build does not run */

/* Growth laws (the reference has two different ones, see the header of
 * mte_kernels.hip). */
#define ROADMC_LAW_LOGNORMAL 0 /* n = n'(1 + r(1 - n'/K)) exp(s z - s^2/2) */
#define ROADMC_LAW_REF_CPU 1   /* Simulator.cpp:712-728 */
#define ROADMC_LAW_REF_GPU 2   /* SimulateGPU.cu:691-709 */

/* Which population a path returns. */
#define ROADMC_END_POST_GROWTH 0 /* Simulator.cpp:772-774 */
#define ROADMC_END_PRE_GROWTH 1  /* SimulateGPU.cu:712-715 */

#define ROADMC_PREC_FP32 0
#define ROADMC_PREC_FP64 1
/* Tiled MTE kernels only: FP32 state and accumulation, products on bf16
 * matrix cores as hi.m_hi + hi.m_lo + lo.m_hi of the two-term bf16 splits
 * x = hi + lo (about 17 significant bits of state and movement matrix). */
#define ROADMC_PREC_BF16X3 2

/* Largest patch count the tiled kernels take (the population tile of one
 * workgroup must fit in LDS): the batched entry points, FP64 and BF16X3,
 * the adjoint and the reference-design baseline. */
#define ROADMC_MAX_PATCHES 1024
/* Largest patch count of roadmc_mte_run in FP32: above ROADMC_MAX_PATCHES
 * the state of every path is kept in HBM and each year is one blocked
 * product (the matrix must fit the device: P^2 floats). */
#define ROADMC_MAX_PATCHES_WIDE 16384

typedef struct roadmc_ctx roadmc_ctx;

struct roadmc_road_species; /* declared with roadmc_build_species_road below */

/* One species on one road: the arrays SpeciesRoadPatches holds
 * (include/road/SpeciesRoadPatches.h:294-306) and the packed parameters of
 * SimulateGPU.cu:3690-3714. */
typedef struct {
  int32_t n_patches;        /* P, 1..ROADMC_MAX_PATCHES (roadmc_mte_run in FP32:
                               ..ROADMC_MAX_PATCHES_WIDE) */
  const double* init_pop;   /* [P] initial population per patch */
  const double* capacity;   /* [P] carrying capacity, > 0 */
  const double* growth;     /* [P] intrinsic growth rate (LOGNORMAL law);
                               NULL reads as 0 */
  const double* trans;      /* [P*P] column-major, (i, j) = P(origin i ->
                               destination j), rows sum to 1 */
  const double* surv;       /* [P*P] column-major survival for the control in
                               force (survProbs[last] in the reference); NULL
                               means no road mortality */
  double params[8];         /* 0 growth current, 1 growth mean, 2 growth noise
                               SD, 3 threshold, 4 reversion, 5 jump size,
                               6 jump probability, 7 local variability */
  double noise_sd;          /* LOGNORMAL year noise sigma */
  double threshold_ratio;   /* MTE: P(total_T >= ratio * total_0) */
  const struct roadmc_road_species* road; /* with trans == NULL (roadmc_mte_run,
                               roadmc_mte_run_batch, roadmc_batch_prepare): trans
                               and the survival of road_control are built on the
                               device from this species' patches and road (as
                               roadmc_build_species_road builds them) and never
                               leave it; NULL otherwise */
  int32_t road_control;     /* that flow control; negative: the last (the
                               reference's survProbs[last]) */
} roadmc_species;

/* One road.  The scalars the reference reads from the optimiser object
 * (SimulateGPU.cu:3667-3675). */
typedef struct {
  int32_t n_species;
  const roadmc_species* species; /* [n_species] */
  int32_t n_years;               /* T >= 1 */
  double time_step;              /* dt */
  int64_t n_paths;               /* N >= 1, < 2^31 */
  int32_t growth_law;            /* ROADMC_LAW_* */
  int32_t endpoint;              /* ROADMC_END_* */
  int32_t precision;             /* ROADMC_PREC_* */
  double drop_tol;               /* movement entries <= drop_tol are removed
                                    (dense2Sparse uses 1e-5); 0 keeps all */
  uint64_t seed;
  uint32_t road_id;              /* < 2^24, selects an independent stream */
} roadmc_mte_problem;

/* Outputs; each pointer may be NULL. */
typedef struct {
  double* mte_prob;   /* [S] fraction of paths with total_T >= ratio*total_0 */
  double* end_mean;   /* [S] mean end total (Simulator.cpp:184-186) */
  double* end_sd;     /* [S] population SD of end totals (1/N) */
  double* patch_mean; /* [sum_s P_s] mean end population per patch */
  double* end_pops;   /* [N*S] column-major per-path end totals (the
                         reference's endPops matrix) */
} roadmc_mte_result;

/* ---- context ---------------------------------------------------------- */

/* SimulateGPU::deviceCount (SimulateGPU.h:44). */
int roadmc_device_count(int* count);
int roadmc_ctx_create(int device, roadmc_ctx** out);
void roadmc_ctx_destroy(roadmc_ctx* ctx);
const char* roadmc_last_error(const roadmc_ctx* ctx);
/* Test hook: the next n device (or pinned host) allocations of this context
 * fail with ROADMC_ERR_OOM. */
int roadmc_debug_fail_allocs(roadmc_ctx* ctx, int n);

/* ---- MTE ---------------------------------------------------------------- */

/* SimulateGPU::simulateMTECUDA (SimulateGPU.h:152-156, SimulateGPU.cu:
 * 3650-3864): all species of one road on the fused tiled kernels. */
int roadmc_mte_run(roadmc_ctx* ctx, const roadmc_mte_problem* problem,
                   roadmc_mte_result* result);

/* The same evaluation with the reference's GPU design: per-call device
 * allocation, random streams pre-generated into HBM, one thread per path over
 * a CSC matrix (SimulateGPU.cu:595-717, 3720-3843).  Kept as the baseline of
 * the batched path; results agree with roadmc_mte_run to FP32 rounding. */
int roadmc_mte_run_plain(roadmc_ctx* ctx, const roadmc_mte_problem* problem,
                         roadmc_mte_result* result);

/* A whole GA generation: every species of every road in one persistent
 * launch (RoadGA.cpp:1687-1748 evaluates such roads one call at a time). */
int roadmc_mte_run_batch(roadmc_ctx* ctx, const roadmc_mte_problem* roads,
                         int32_t n_roads, roadmc_mte_result* results);

/* Batched evaluation split into upload / launch / download, so a caller can
 * relaunch a generation whose inputs already sit in HBM.  launch() replays a
 * captured HIP graph on the context's stream and returns without waiting. */
typedef struct roadmc_batch roadmc_batch;
int roadmc_batch_prepare(roadmc_ctx* ctx, const roadmc_mte_problem* roads,
                         int32_t n_roads, roadmc_batch** out);
int roadmc_batch_launch(roadmc_batch* batch);
int roadmc_batch_sync(roadmc_batch* batch);
int roadmc_batch_fetch(roadmc_batch* batch, roadmc_mte_result* results);
/* Average duration in ms of the persistent kernel over the last n launches,
 * measured with HIP events on the stream the kernel runs on. */
int roadmc_batch_kernel_ms(roadmc_batch* batch, double* ms);
void roadmc_batch_destroy(roadmc_batch* batch);

/* Tile settings of the MTE launches, per precision (ROADMC_PREC_FP32 or
 * ROADMC_PREC_BF16X3; FP64 tiles follow LDS) and patch-count bucket
 * (P <= 128, 256, 512, 1024): rows of paths per workgroup (a power of two in
 * [32, 256], rows x P <= 32768); split 1 lets remainder tiles (the tiles left
 * after an even share per SIMD) be computed as four k-quarters, 0 not;
 * per_cu 1 or 2 workgroups per CU of the persistent launch (0: two when LDS
 * allows; a batch takes the smallest of its buckets' values).
 * Rows and per_cu never change results (each path's products and draws are
 * the same in any tile); a split changes the k-summation order of the split
 * tiles' paths, so their end totals move within FP32 rounding.
 * roadmc_autotune_ex times every valid setting of each bucket on a synthetic
 * road of the bucket's largest P (n_paths paths, n_years years), stores the
 * fastest in the context, and reports up to max_trials trials (n_trials: all
 * of them).  roadmc_autotune does so for FP32 and BF16X3 without a report;
 * roadmc_get_tiles reports the FP32 rows. */
typedef struct {
  int32_t precision;     /* ROADMC_PREC_FP32 or ROADMC_PREC_BF16X3 */
  int32_t bucket_max_p;  /* 128, 256, 512 or 1024 */
  int32_t rows, split, per_cu;
  int32_t chosen;        /* 1: the bucket's fastest, now in the context */
  double ms;             /* average persistent-kernel time of two launches */
} roadmc_tile_trial;
int roadmc_autotune_ex(roadmc_ctx* ctx, int32_t precision, int32_t n_paths, int32_t n_years,
                       roadmc_tile_trial* trials, int32_t max_trials, int32_t* n_trials);
int roadmc_autotune(roadmc_ctx* ctx, int32_t n_paths, int32_t n_years);
int roadmc_set_tile(roadmc_ctx* ctx, int32_t precision, int32_t bucket_max_p, int32_t rows,
                    int32_t split, int32_t per_cu);
int roadmc_get_tile(const roadmc_ctx* ctx, int32_t precision, int32_t bucket_max_p,
                    int32_t* rows, int32_t* split, int32_t* per_cu);
int roadmc_get_tiles(const roadmc_ctx* ctx, int32_t* bucket_max_p,
                     int32_t* rows, int32_t* n_buckets);

/* ---- helpers restated from the reference -------------------------------- */

/* SimulateGPU::dense2Sparse (SimulateGPU.cu:5973-5996): column-major dense
 * to CSC, keeping entries > tol.  nnz_out receives the count; val and row_idx
 * need rows*cols capacity. */
int roadmc_dense_to_csc(const float* dense, int32_t rows, int32_t cols,
                        float tol, float* val, int32_t* row_idx,
                        int32_t* per_col, int32_t* nnz_out);

/* Utility::NormalCDFInverse (Utility.cpp:165-193). */
int roadmc_normal_cdf_inverse(double p, double* out);

/* Penalty of Simulator::simulateMTE (Simulator.cpp:197-212) summed over
 * species. */
int roadmc_mte_penalty(int32_t n_species, const double* threshold,
                       const double* end_mean, const double* end_sd,
                       double confidence_interval,
                       const double* initial_population,
                       const double* cost_per_animal, double* penalty);

/* Writes the patch-noise normals z(path, year, patch) for paths [0, n_paths),
 * years [0, n_years), patches [0, n_patches) to out[(path*n_years + year)*
 * n_patches + patch] (the reference's rgr layout, SimulateGPU.cu:703-705),
 * computed on the device. */
int roadmc_patch_normals(roadmc_ctx* ctx, uint64_t seed, uint32_t road_id,
                         uint32_t species, int32_t n_paths, int32_t n_years,
                         int32_t n_patches, float* out);

/* ---- Road -> movement and survival matrices -------------------------------- */

/* One species' patches and one road, as SpeciesRoadPatches reads them
 * (SpeciesRoadPatches.cpp:269-451).  Transition weights use
 * lambda = lambda_mean + lambda_sd * lambda_mult; the distance cut-off is
 * max_dist = ln(0.05) / -(lambda_mean + lambda_sd * x), with x = lambda_mult
 * for the crossing counts (roadCrossings, :328-332) and x = lambda for the
 * transitions (computeTransitionProbabilities, :389-395), as there. */
#define ROADMC_SEG_ROAD 0   /* RoadSegments::ROAD: counted */
#define ROADMC_SEG_BRIDGE 1 /* not counted */
#define ROADMC_SEG_TUNNEL 2 /* not counted */

#define ROADMC_TRANS_REFERENCE 0 /* off-diagonal weight / exp(-lambda d), i.e. x exp(+lambda d)
                                    (SpeciesRoadPatches.cpp:413-416) */
#define ROADMC_TRANS_DECAY 1     /* off-diagonal weight x exp(-lambda d) */

typedef struct roadmc_road_species {
  int32_t n_patches;         /* P */
  const double* x;           /* [P] patch centroids */
  const double* y;           /* [P] */
  const double* area;        /* [P] */
  const double* pref_mean;   /* [P] habitat-preference mean of the patch's type */
  const double* pref_sd;     /* [P] habitat-preference SD of the patch's type */
  double pref_mult;          /* habitat-preference multiplier (scenario) */
  double lambda_mean, lambda_sd, lambda_mult;
  int32_t n_road_points;     /* road vertices; segment s joins s and s + 1 */
  const double* road_x;      /* [n_road_points] */
  const double* road_y;      /* [n_road_points] */
  const int32_t* road_type;  /* [n_road_points - 1] ROADMC_SEG_*, or NULL: all road */
  int32_t n_controls;        /* C flow rates */
  const double* flow;        /* [C] vehicles per hour */
  double vehicle_width;      /* sum of vehicle width x proportion */
  double animal_length;      /* species length mean */
  double animal_speed;       /* species speed mean */
  int32_t transition_form;   /* ROADMC_TRANS_* */
  int32_t drop_last_link;    /* 1: as roadCrossings, which passes all but the last
                                in-range pair to the crossing count
                                (SpeciesRoadPatches.cpp:369-374); 0: count all */
} roadmc_road_species;

/* SpeciesRoadPatches::roadCrossings + computeTransitionProbabilities +
 * computeSurvivalProbabilities (SpeciesRoadPatches.cpp:296-451) with the
 * segment test of Utility::lineSegmentIntersect (Utility.cpp:3-92; GPU form
 * SimulateGPU.cu:454-493) in FP64 on the GPU.  Outputs are column-major P x P
 * (Eigen's order), origin = row; any may be NULL:
 *   crossings [P*P], trans [P*P], surv [C][P*P]. */
int roadmc_build_species_road(roadmc_ctx* ctx, const roadmc_road_species* rs,
                              int32_t* crossings, double* trans, double* surv);

/* ---- Adjoint ---------------------------------------------------------------- */

/* Gradient of J = d_objective * mean total end population (one species,
 * LOGNORMAL law, post-growth endpoint, common random numbers) with respect to
 * the crossing mortality m (surv = 1 - m) and the growth rates r.  No
 * counterpart in the reference, which has no derivatives; it extends
 * simulateMTECUDA (SimulateGPU.cu:3650-3864).  The forward pass stores the
 * state every checkpoint_every years; the backward pass recomputes each
 * segment from its checkpoint. */
typedef struct roadmc_adjoint roadmc_adjoint;
int roadmc_mte_adjoint_fwd(roadmc_ctx* ctx, const roadmc_mte_problem* problem,
                           int32_t checkpoint_every, roadmc_adjoint** out,
                           double* objective);
/* grad_mortality: [P*P] column-major like trans; grad_growth: [P]. */
int roadmc_mte_adjoint_bwd(roadmc_adjoint* adj, double d_objective,
                           double* grad_mortality, double* grad_growth);
void roadmc_adjoint_destroy(roadmc_adjoint* adj);

/* ---- ROV: price paths and regression Monte Carlo --------------------------- */

/* Price laws of one exogenous uncertainty. */
#define ROADMC_SDE_MRJD 0 /* x += k(m - x)dt + x s z + (e^J' - 1) x 1[u < p],
                             J' = J^2 z' - J^2/2 (SimulateGPU.cu:930-952) */
#define ROADMC_SDE_OU 1   /* x += k(m - x)dt + s sqrt(dt) z */
#define ROADMC_SDE_GBM 2  /* x *= exp((m - s^2/2)dt + s sqrt(dt) z) */

#define ROADMC_ROV_SWITCHING 0 /* flow control per step (backwardInduction,
                                  SimulateGPU.cu:2333-2784) */
#define ROADMC_ROV_AMERICAN 1  /* optimal stopping of a put on uncertainty 0
                                  (Longstaff and Schwartz 2001); realized cash
                                  flows only: ROADMC_VALUE_REGRESSED returns
                                  ROADMC_ERR_UNSUPPORTED */

#define ROADMC_VALUE_REGRESSED 0 /* carry the regressed value back
                                    (SimulateGPU.cu:2757-2771) */
#define ROADMC_VALUE_REALIZED 1  /* carry realized cash flows back */

#define ROADMC_MAX_UNCERTAINTIES 8
#define ROADMC_MAX_CONTROLS 8
#define ROADMC_MAX_BASIS 35 /* cubic in 4 dimensions */

/* The six packed parameters of SimulateGPU.cu:4020-4070 plus the law. */
typedef struct {
  int32_t kind;          /* ROADMC_SDE_* */
  double x0;             /* current value */
  double mean;           /* long-run mean (MRJD/OU) or drift (GBM) */
  double sd;             /* volatility */
  double reversion;      /* mean-reversion strength */
  double jump_size;      /* MRJD jump size J */
  double jump_prob;      /* MRJD jump probability per step */
} roadmc_uncertainty;

typedef struct {
  int32_t n_uncertainties;            /* U, 1..ROADMC_MAX_UNCERTAINTIES */
  const roadmc_uncertainty* unc;      /* [U] */
  const double* correlation;          /* [U*U] row-major, NULL = independent */
  int32_t n_controls;                 /* C, 1..ROADMC_MAX_CONTROLS (SWITCHING) */
  const double* flow;                 /* [C] flow rate per control, Q[c] */
  double unit_cost;                   /* unit profit = unit_cost + sum_u w_u x_u */
  const double* profit_weight;        /* [U] (fuel costs, -revenue * ore) */
  int32_t traffic_index;              /* uncertainty scaling the flow, or -1 */
  int32_t n_steps;                    /* T decision steps */
  double time_step;                   /* dt */
  double rrr;                         /* per-step discount 1/(1 + rrr dt / 100) */
  double discount;                    /* if > 0, the per-step discount factor */
  int64_t n_paths;                    /* N */
  int32_t basis_degree;               /* 1..3: all monomials up to this degree */
  int32_t standardize;                /* scale features per step to zero mean,
                                         unit variance before the fit */
  int32_t itm_only;                   /* fit only in-the-money paths
                                         (keepValidPoints, SimulateGPU.cu:5999) */
  int32_t pooled;                     /* 1: one fit on all paths; 0: one fit per
                                         randomly drawn control (randControls) */
  double hysteresis;                  /* switch only if payoff < h * best
                                         (0.975 in SimulateGPU.cu:2763) */
  int32_t mode;                       /* ROADMC_ROV_* */
  int32_t value_mode;                 /* ROADMC_VALUE_* */
  double strike;                      /* AMERICAN: payoff max(strike - x_0, 0) */
  uint64_t seed;
  const double* states;               /* optional pre-generated paths
                                         [(T+1)][U][N]; NULL simulates them */
  const int32_t* controls;            /* optional random controls [T][N] */
} roadmc_rov_problem;

typedef struct {
  double* value;         /* [1] time-0 value (SWITCHING: expected cost, <= 0
                            is profit; AMERICAN: option value) */
  double* value_sd;      /* [1] SD of the time-0 value over paths */
  int32_t* first_control;/* [1] optimal control at time 0 */
  double* coeffs;        /* [T][C][K] fitted coefficients (raw basis when
                            standardize is 0); K from roadmc_rov_basis_size */
  double* cond_exp;      /* [(T+1)][N] value per path and step */
  int32_t* opt_control;  /* [T][N] chosen control per path and step */
  double* states;        /* [(T+1)][U][N] simulated paths */
} roadmc_rov_result;

/* Number of basis functions of a degree-d polynomial in n variables. */
int roadmc_rov_basis_size(int32_t n_vars, int32_t degree, int32_t* size);

/* SimulateGPU::simulateROVCUDA (SimulateGPU.h:187-191): forward paths, then
 * backward induction with per-step FP64 regression, captured in one HIP
 * graph. */
int roadmc_rov_run(roadmc_ctx* ctx, const roadmc_rov_problem* problem,
                   roadmc_rov_result* result);

/* SimulateGPU::expPV (SimulateGPU.h:53, SimulateGPU.cu:3163-3259): expected
 * present value sum_t (1 + g)^t x_t / (1 + r)^t of one uncontrolled MRJD
 * price with Brownian SD s*dt (the reference's parameterisation), and its SD
 * over paths (1/N; the reference omits the 1/N, Uncertainty.cpp:75). */
int roadmc_exp_pv(roadmc_ctx* ctx, const roadmc_uncertainty* unc,
                  int32_t n_steps, double time_step, double rrr, double growth,
                  int64_t n_paths, uint64_t seed, double* mean, double* sd);

/* roadmc_exp_pv without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under stream
 * capture, serving loops).  The per-path values and the statistics live in a
 * fixed device array of the library (8 MiB per device, held until the process
 * ends), so n_paths <= ROADMC_NOALLOC_MAX_PATHS; larger requests return
 * ROADMC_ERR_INVALID.  Calls on one device share that array: calls that
 * overlap on one device race.  Same streams and results as roadmc_exp_pv to
 * FP64 rounding (mean and SD are reduced on the device). */
#define ROADMC_NOALLOC_MAX_PATHS 1048576
int roadmc_exp_pv_noalloc(roadmc_ctx* ctx, const roadmc_uncertainty* unc,
                          int32_t n_steps, double time_step, double rrr, double growth,
                          int64_t n_paths, uint64_t seed, double* mean, double* sd);

/* ---- ROV with species states (the reference's ALGO4) --------------------- */

/* One species of the species-state valuation.  Every species has the same
 * patches (SimulateGPU.cu:3965-3975). */
typedef struct {
  const double* init_pop;  /* [P] */
  const double* capacity;  /* [P] */
  const double* trans;     /* [P*P] column-major, origin = row */
  const double* surv;      /* [C][P*P] survival under each control, column-major */
  double params[8];        /* the packing of SimulateGPU.cu:3990-4016: growth current,
                              mean, noise SD, threshold, reversion, jump size, jump
                              probability, local variability */
} roadmc_rovs_species;

/* The valuation methods (Optimiser::ROVType, Optimiser.h:74-82).  ALGO6
 * (SimulateGPU.cu:5011-5444, optimalForwardPaths :1199-2206): after each
 * step's grids are fitted, every path is recomputed from that step to the
 * horizon, re-simulated to the step under its drawn controls and then under
 * the fitted policy (each year's decision on the recomputed state with that
 * year's grids, the terminal rule in the last year), and its value at the step
 * is the realized cash flows, Q_c x unit profit of each year's chosen control
 * discounted by 1 / (1 + rrr dt / 100) per year; those values are the next
 * step's regression responses.  The reference's kernel accumulates that value
 * but leaves its store commented out (:2191) and migrates with the stored
 * rather than the recomputed decisions; here the value is kept and the policy
 * followed.  S 32 x (P rounded up to 32) + 4 float tiles must fit the 160 KiB
 * of LDS of one workgroup (S <= 2 at P = 512, S <= 4 at P = 256). */
#define ROADMC_ROVS_ALGO4 4
#define ROADMC_ROVS_ALGO6 6

/* SimulateGPU::simulateROVCUDA with ROVMethod ALGO4 (SimulateGPU.cu:3870-4998):
 *  - forward (forwardPathKernel, :789-956): per path and year a control drawn
 *    uniformly (randControls, :720-730); each species' patch populations moved
 *    by trans o surv of that control, the surviving fraction recorded under
 *    every control (the AARs); growth with the growth-rate SDE and
 *    n = n'(1 + g (K - n') / K / 100), g = grMean + z local_variability; fuel
 *    and commodity prices on the MRJD law; ore composition N(mean, sd) per
 *    path, year and commodity;
 *  - unit profit = unit_cost + sum_v fuel_cost[v] price[fuel_idx[v]]
 *    - unit_revenue sum_c composition_c price[n_fuels + c];
 *  - backward (:4401-4998, backwardInduction :2333-2784): state = (adjusted
 *    population of each species under the highest flow, unit profit); per
 *    step and control a kNN local-linear grid (k = 10 (ln n + 1), capped at
 *    ROADMC_KNN_MAX_K) over the paths that drew that control and are in the
 *    money (every adjusted population at its threshold) or drew control 0;
 *    each path takes the valid control minimising Q_c unit profit plus the
 *    discounted, interpolated continuation (<= 0), switching from the running
 *    best only below hysteresis x best;
 *  - time 0 (firstPeriodInduction, :2210-2330).
 * P <= 512 (ALGO4 with FP32 products: P <= ROADMC_MAX_PATCHES_WIDE, the forward
 * populations then kept in HBM), n_species + 1 <= ROADMC_KNN_MAX_DIMS, n_controls <=
 * ROADMC_MAX_CONTROLS, dim_res^(n_species+1) <= 2^24. */
typedef struct {
  int32_t n_patches;                      /* P */
  int32_t n_species;                      /* S */
  const roadmc_rovs_species* species;     /* [S] */
  int32_t n_controls;                     /* C */
  const double* flow;                     /* [C] Q_c; control C - 1 is the highest flow */
  int32_t n_fuels;                        /* fuel prices: uncertainties 0 .. n_fuels - 1 */
  int32_t n_commodities;                  /* ore prices: the next n_commodities */
  const roadmc_uncertainty* uncertainties;/* [n_fuels + n_commodities] */
  const double* ore_mean;                 /* [n_commodities] composition mean */
  const double* ore_sd;                   /* [n_commodities] composition SD */
  int32_t n_vehicles;
  const double* fuel_cost;                /* [n_vehicles] fuel use per unit traffic */
  const int32_t* fuel_idx;                /* [n_vehicles] fuel of each vehicle class */
  double unit_cost, unit_revenue;
  int32_t n_years;                        /* T */
  double time_step;                       /* dt */
  double rrr;                             /* discount 1 / (1 + rrr dt / 100) per step */
  int64_t n_paths;                        /* N */
  int32_t dim_res;                        /* grid points per state dimension */
  double profit_scaling;                  /* unit-profit scale of the kNN search
                                             (2 in SimulateGPU.cu:3900) */
  double hysteresis;                      /* 0.975 in SimulateGPU.cu:2763 */
  uint64_t seed;
  uint32_t road_id;
  int32_t method;                         /* ROADMC_ROVS_ALGO4 (0 means the same) or
                                             ROADMC_ROVS_ALGO6 */
  int32_t checkpoint_years;               /* ALGO6: forward populations kept every this
                                             many years, the recomputation replaying from
                                             the last one (0: every year while they fit
                                             in 16 GiB) */
  int32_t precision;                      /* products of the forward pass and the
                                             recomputation: ROADMC_PREC_FP32, or
                                             ROADMC_PREC_BF16X3 with ALGO6 only
                                             (ALGO4 returns ROADMC_ERR_UNSUPPORTED) */
} roadmc_rovs_problem;

typedef struct {
  double* value;          /* [1] time-0 expected cost (<= 0 is profit) */
  double* value_sd;       /* [1] SD over the paths of the chosen control */
  int32_t* first_control; /* [1] */
  double* cond_exp;       /* [T][N] value per step and path (time 0 included), or NULL */
  int32_t* opt_control;   /* [T][N], or NULL */
  double* adj_pops;       /* [T][S][N] highest-flow adjusted populations, or NULL */
  double* unit_profit;    /* [T][N], or NULL */
  int32_t* controls;      /* [T][N] drawn controls, or NULL */
  double* total_pops;     /* [S][T+1][N] forward total populations, or NULL */
  double* aars;           /* [S][T][C][N] surviving fraction per control, or NULL */
  double* grids;          /* [T][C][(S+1) dim_res + dim_res^(S+1)] fitted grids
                             (steps 1 .. T-2), or NULL */
} roadmc_rovs_result;

int roadmc_rovs_run(roadmc_ctx* ctx, const roadmc_rovs_problem* problem,
                    roadmc_rovs_result* result);

/* Smoothed-validity value of the species-state valuation under a fixed
 * policy, and its gradient with respect to every species' survival matrices.
 * No counterpart in the reference: there a control is valid only when every
 * adjusted population is at its threshold (SimulateGPU.cu:1046-1054,
 * 2757-2771), an indicator, so the value's pathwise derivative with respect to
 * mortality is zero almost everywhere.  Here the indicator is replaced by
 * prod_s sigmoid((A_s - threshold_s) / tau) and the decisions by the given
 * policy [T][N] (for instance roadmc_rovs_run's opt_control):
 *   value = (1/N) sum_paths sum_{t<T} d^t UP_t (Q_0 + (Q_u - Q_0) prod_s sigma_s,t)
 * with u = policy[t][path], A_s,t the adjusted population of species s under u
 * (total population x surviving fraction), d = 1 / (1 + rrr dt / 100), UP the
 * unit profit of roadmc_rovs_run's streams, and every species' populations
 * moved by trans o surv_u and grown as roadmc_rovs_run's forward pass does.
 * grad_surv [S][C][P*P] (column-major like surv, or NULL for the value only)
 * receives dvalue/dsurv; the crossing mortality m = 1 - surv has the opposite
 * gradient.  FP32 products, P <= 512 (ROADMC_ERR_UNSUPPORTED otherwise), tau >
 * 0, n_years x n_paths <= 2^30, every policy entry in [0, n_controls);
 * problem->method and dim_res are not used. */
int roadmc_rovs_smooth_grad(roadmc_ctx* ctx, const roadmc_rovs_problem* problem,
                            const int32_t* policy, double tau, double* value,
                            double* grad_surv);

/* ---- On-disk formats ------------------------------------------------------- */

/* The surrogate training data file of a scenario ("surrogate_data"):
 * RoadGA::writeOutSurrogateData (RoadGA.cpp:2977-3070) and
 * RoadGA::getExistingSurrogateData (RoadGA.cpp:4553-4611).  Four header lines
 * (three banner lines, one column line), then one sample per line. */
#define ROADMC_SURR_MTE 0        /* per species: initial AAR, mean end population, its SD */
#define ROADMC_SURR_CONTROLLED 1 /* initial AAR per species, initial unit profit, mean
                                    value, value SD */

/* Appends n_rows samples (rows: [n_rows][3 n_species] for MTE, [n_rows][n_species + 3]
 * for CONTROLLED), creating the file with the reference's header when it does not
 * exist.  CONTROLLED samples are kept only when |value| and |SD| <= |max_benefit|
 * (and max_benefit is finite), as the reference filters them.  Numbers are written
 * as the reference's stream writes them: width 20, 6 significant digits. */
int roadmc_surrogate_data_write(const char* path, int32_t kind, int32_t n_species,
                                int32_t n_rows, const double* rows, double max_benefit,
                                int32_t* n_written);
/* Reads up to max_rows samples into rows (same layout); *n_rows receives the
 * number of samples in the file. */
int roadmc_surrogate_data_read(const char* path, int32_t kind, int32_t n_species,
                               int32_t max_rows, double* rows, int32_t* n_rows);

/* Input files of the reference's data directory, read by key ("Key : value"
 * lines; RoadGA::initialiseFromTextInput, RoadGA.cpp:3464-4550, reads the same
 * fields by position), so that Monte Carlo inputs can be built from them. */
#define ROADMC_NAME_LEN 64
#define ROADMC_MAX_HABITAT_TYPES 8

typedef struct { /* Commodities/<name> */
  char name[ROADMC_NAME_LEN];
  double current, mean, sd, reversion, jump_prob, poisson_jump;
  double ore_mean, ore_sd; /* contribution to the transported ore: mean, SD */
} roadmc_commodity_file;

typedef struct { /* Vehicles/<name> */
  char name[ROADMC_NAME_LEN];
  double width, length, proportion, load;
  int32_t fuel; /* index into the fuels of the main input file */
  double a, agr, av, avsq, travel_cost;
} roadmc_vehicle_file;

typedef struct { /* Species/<name>; fields absent from the file are NaN */
  char name[ROADMC_NAME_LEN];
  int32_t n_habitat_types;
  char habitat_name[ROADMC_MAX_HABITAT_TYPES][16];
  double max_pop[ROADMC_MAX_HABITAT_TYPES], cost_per_m2[ROADMC_MAX_HABITAT_TYPES];
  double pref_mean[ROADMC_MAX_HABITAT_TYPES], pref_sd[ROADMC_MAX_HABITAT_TYPES];
  double growth_current, growth_mean, growth_sd, growth_reversion, growth_jump_prob,
      growth_poisson_jump;
} roadmc_species_file;

int roadmc_read_commodity(const char* path, roadmc_commodity_file* out);
int roadmc_read_vehicle(const char* path, roadmc_vehicle_file* out);
int roadmc_read_species(const char* path, roadmc_species_file* out);

/* The Monte Carlo subset of the main input file (INPUT_FILE, parsed by
 * position in RoadGA::initialiseFromTextInput, RoadGA.cpp:3464-3890): the
 * economic horizon, path count and grid resolutions, the method codes, the
 * variable-parameter multiplier lists of the scenario grid, the traffic
 * programs and the commodity, vehicle and species file lists.  Absent fields
 * read as NaN, -1 or empty lists. */
#define ROADMC_MAX_LIST 16     /* values of one multiplier list */
#define ROADMC_MAX_PROGRAMS 8  /* traffic programs */
#define ROADMC_MAX_FILES 16    /* commodity, vehicle or species files */
#define ROADMC_PATH_LEN 256

typedef struct {
  int32_t n;
  double v[ROADMC_MAX_LIST];
} roadmc_list;

typedef struct {
  double confidence_interval;  /* "Confidence Interval": quantile of the MTE penalty */
  int32_t use_gpu;             /* "Use GPU": 1 TRUE, 0 FALSE */
  int32_t surrogate_dim_res;   /* "Surrogate Dimension Resolution" */
  int32_t interpolation;       /* "Interpolation Routine": 0 cubic spline, 1 multiple
                                  local linear regression (the kNN grid) */
  double traffic_growth;       /* "Traffic Growth Rate (p.a.)" */
  double rrr;                  /* "Required Rate of Return (per cen per annum)", per cent */
  int32_t n_years;             /* "Number of Years" */
  double time_step;            /* "Step Size (Years)" */
  int32_t n_paths;             /* "Number of Simulation Paths" */
  int32_t rov_dim_res;         /* "ROV Dimension Resolution" */
  int32_t optimisation_method; /* "Optimisation Method" as written.  The file's legend
                                  counts from 1; the reference casts the value to
                                  Optimiser::Type, which counts from 0 (RoadGA.cpp:
                                  3797-3799, Optimiser.h:67-72) */
  int32_t rov_method;          /* "ROV Method" as written, with the same mismatch: 4
                                  ("State & Control" in the legend) selects ALGO5, an
                                  empty case, in the reference (RoadGA.cpp:3803-3805,
                                  Optimiser.h:74-82, SimulateGPU.cu:5005-5009) */
  /* "VARIABLE PARAMETERS" (the scenario grid, VariableParameters) */
  roadmc_list animal_bridges, hab_pref_sd, movement_sd, ranging_sd, growth_mean,
      growth_sd, price_mean, price_sd, ore_sd, pop_level, comparison_cost;
  /* "Program k : flows : switching costs [: bridge]" ("Programs" gives the
     count where present).  The bridge flag is read from the fourth field; the
     reference compares the flow field instead (RoadGA.cpp:3918), so its flag
     is always false */
  int32_t n_programs;
  int32_t n_flows[ROADMC_MAX_PROGRAMS];
  double flows[ROADMC_MAX_PROGRAMS][ROADMC_MAX_CONTROLS];
  int32_t n_switching[ROADMC_MAX_PROGRAMS];
  double switching[ROADMC_MAX_PROGRAMS][1 << ROADMC_MAX_CONTROLS];
  int32_t bridge[ROADMC_MAX_PROGRAMS];
  /* "Number of Fuels" and "Number of Transported Commodities" (-1 when absent;
     older files give only "Number of Commodities", read into n_transported);
     the "Commodity k" files list the fuels first.  "Vehicle k", "Species k":
     paths as written, relative to the program's root directory */
  int32_t n_fuels, n_transported;
  int32_t n_commodities, n_vehicles, n_species;
  char commodity[ROADMC_MAX_FILES][ROADMC_PATH_LEN];
  char vehicle[ROADMC_MAX_FILES][ROADMC_PATH_LEN];
  char species[ROADMC_MAX_FILES][ROADMC_PATH_LEN];
  /* "EXISTING ROAD DATA" (the CONTROLLED method's comparison road) */
  int32_t road_exists;         /* 1 TRUE, 0 FALSE, -1 absent */
  double travel_time, road_length, accident_cost;
  roadmc_list fuel_per_journey;
} roadmc_run_settings;

int roadmc_read_run_settings(const char* path, roadmc_run_settings* out);

/* ---- Learned surrogate --------------------------------------------------- */

/* An ensemble of n_members MLPs n_in -> 64 -> 64 -> n_out (tanh), trained in
 * FP32 by full-batch Adam on the mean squared error with this library's
 * kernels.  It stands in for the reference's kNN/local-linear surrogate fit
 * (SimulateGPU::buildSurrogateMTECUDA/ROVCUDA, SimulateGPU.cu:5533-5968,
 * interpolated by Utility::interpolateSurrogate, Utility.cpp:253-363);
 * callers fall back to full Monte Carlo where the members disagree.
 * Parameters of a member: W1[n_in][64], b1[64], W2[64][64], b2[64],
 * W3[64][n_out], b3[n_out]. */
typedef struct roadmc_mlp roadmc_mlp;
int roadmc_mlp_create(roadmc_ctx* ctx, int32_t n_in, int32_t n_hidden,
                      int32_t n_out, int32_t n_members, roadmc_mlp** out);
void roadmc_mlp_destroy(roadmc_mlp* net);
int roadmc_mlp_n_params(const roadmc_mlp* net, int32_t* n);
/* params [n_members][n_params]; setting them resets the Adam state. */
int roadmc_mlp_set_params(roadmc_mlp* net, const float* params);
int roadmc_mlp_get_params(roadmc_mlp* net, float* params);
/* Gradient of the mean squared error over (x [n][n_in], y [n][n_out]):
 * grad [n_members][n_params], loss [n_members]. */
int roadmc_mlp_gradient(roadmc_mlp* net, const float* x, const float* y,
                        int32_t n, float* grad, float* loss);
int roadmc_mlp_train(roadmc_mlp* net, const float* x, const float* y,
                     int32_t n, int32_t steps, float lr, float* loss);
/* out [n_members][n][n_out]. */
int roadmc_mlp_predict(roadmc_mlp* net, const float* x, int32_t n, float* out);

/* ---- Grid surrogate: kNN local-linear regression -------------------------- */

/* The reference's surrogate fit, SimulateGPU::buildSurrogateMTECUDA and
 * buildSurrogateROVCUDA (SimulateGPU.cu:5533-5917) with createQueryPoints
 * (:1152-1193), multiLocLinReg (:2934-3053) and rovCorrection (:3055-3073),
 * evaluated by multilinear interpolation (Utility::interpolateSurrogate,
 * Utility.cpp:253-363; interpolateMulti, SimulateGPU.cu:3077-3155).
 *
 * The surrogate vector holds the grid coordinates of each dimension, then
 * the fitted values: surrogate[dd*dim_res + i] = x_min[dd] + i (x_max[dd] -
 * x_min[dd]) / (dim_res - 1), and value[idx] at surrogate[n_dims*dim_res +
 * idx] with idx = sum_dd i_dd dim_res^(n_dims-1-dd) (dimension 0 most
 * significant), as the reference lays it out.  Each grid value is the
 * intercept of a weighted linear fit over the k nearest samples (ties to the
 * lower sample index), weights exp(-(d/h)^2/2)/sqrt(2 pi) with h the k-th
 * neighbour's distance; solved by FP64 Cholesky.  When the fit is singular
 * (neighbours that do not span the dimensions) the value is the weighted
 * mean; with fewer than 3 samples every value is 0, as the reference. */
#define ROADMC_KNN_MAX_DIMS 8
#define ROADMC_KNN_MAX_K 128
#define ROADMC_KNN_NORMALIZE 1   /* neighbours and fit in [0,1]-scaled coordinates
                                    (buildSurrogateROVCUDA, :5807-5822) */
#define ROADMC_KNN_NONPOSITIVE 2 /* clamp fitted values to <= 0 (rovCorrection) */
#define ROADMC_KNN_RULE_MTE 0    /* neighbour count of buildSurrogateMTECUDA (:5545-5554) */
#define ROADMC_KNN_RULE_ROV 1    /* neighbour count of buildSurrogateROVCUDA (:5711-5724) */

/* The reference's neighbour count for n_samples samples, at most n_samples and
 * at most ROADMC_KNN_MAX_K.  The reference has no cap: its ROV rule
 * (25 + ceil(0.02 (n - 500))) passes 128 above 5 650 samples, where this
 * returns 128. */
int roadmc_knn_k(int32_t n_samples, int32_t rule, int32_t* k);
/* x: [n_samples][n_dims] row-major, y: [n_samples]; surrogate:
 * [n_dims*dim_res + dim_res^n_dims].  dim_res >= 2, dim_res^n_dims <= 2^24,
 * 1 <= k <= ROADMC_KNN_MAX_K. */
int roadmc_knn_surrogate_fit(roadmc_ctx* ctx, int32_t n_samples, int32_t n_dims,
                             const double* x, const double* y, int32_t dim_res, int32_t k,
                             int32_t flags, double* surrogate);
/* predictors: [n_points][n_dims]; out: [n_points].  Multilinear
 * interpolation in the grid cell of each point (cells at the ends of each
 * dimension extend linearly beyond the grid). */
int roadmc_knn_surrogate_eval(roadmc_ctx* ctx, int32_t n_dims, int32_t dim_res,
                              const double* surrogate, int32_t n_points,
                              const double* predictors, double* out);

#ifdef __cplusplus
}
#endif

#endif /* ROADMC_H */
