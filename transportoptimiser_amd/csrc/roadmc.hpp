// C++ host interface over the C ABI, with the call shape of the reference's
// SimulateGPU class (include/gpuCode/SimulateGPU.h:44-191) so that a
// Simulator-style caller can switch to it: the same argument order and
// meaning, caller-sized outputs, and errors thrown as `const char*` as
// gpuAssert does (SimulateGPU.cu:14-26).  Eigen and the Simulator object graph
// are replaced by plain containers; the scalars the reference reads through
// sim->getRoad()->getOptimiser() (SimulateGPU.cu:3667-3675) are passed in
// SimulatorInputs.
//
// Header-only; link with -lroadmc.
#ifndef ROADMC_HPP
#define ROADMC_HPP

#include <algorithm>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "roadmc.h"

namespace roadmc {

// Column-major dense matrix (Eigen::MatrixXd's storage order).
struct Matrix {
  int64_t rows = 0, cols = 0;
  std::vector<double> data;
  Matrix() = default;
  Matrix(int64_t r, int64_t c) : rows(r), cols(c), data((size_t)(r * c), 0.0) {}
  double& operator()(int64_t r, int64_t c) { return data[(size_t)(c * rows + r)]; }
  double operator()(int64_t r, int64_t c) const { return data[(size_t)(c * rows + r)]; }
};

// What SimulateGPU reads from a SpeciesRoadPatches object
// (SpeciesRoadPatches.h:294-306) and its Species (SimulateGPU.cu:3690-3714).
struct SpeciesRoadPatches {
  Matrix transProbs;                 // P x P, origin rows
  std::vector<Matrix> survProbs;     // one P x P matrix per control
  std::vector<double> growthRates;   // per patch (LOGNORMAL law), may be empty
  double speciesParams[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double noiseSD = 0.1;
  double thresholdRatio = 0.5;
  // Patch populations and capacities: the species-state simulateROVCUDA
  // reads them from the road's habitat patches (SimulateGPU.cu:3965-3989);
  // simulateMTECUDA takes them as arguments, as the reference does.
  std::vector<double> initPops;
  std::vector<double> capacities;
};

// Column-major int matrix (Eigen::MatrixXi's storage order).
struct MatrixI {
  int64_t rows = 0, cols = 0;
  std::vector<int32_t> data;
  MatrixI() = default;
  MatrixI(int64_t r, int64_t c) : rows(r), cols(c), data((size_t)(r * c), 0) {}
  int32_t& operator()(int64_t r, int64_t c) { return data[(size_t)(c * rows + r)]; }
  int32_t operator()(int64_t r, int64_t c) const { return data[(size_t)(c * rows + r)]; }
};

// What SimulateGPU::expPV reads from an Uncertainty (SimulateGPU.cu:3163-3259,
// multipliers already applied) and the two results it stores back on it.
struct Uncertainty {
  roadmc_uncertainty params{};
  double expPV = 0.0;
  double expPVSD = 0.0;
};

// Economic inputs of the valuation; the reference reads these through
// sim->getRoad()->getOptimiser() (SimulateGPU.cu:3903-3936).
struct RovInputs {
  std::vector<roadmc_uncertainty> uncertainties;
  std::vector<double> correlation;   // U x U row-major, empty = independent
  std::vector<double> flows;         // per control
  double unitCost = 0.0;
  std::vector<double> profitWeights; // per uncertainty
  int trafficIndex = -1;
  int nYears = 50;
  double timeStep = 1.0;
  double rrr = 7.0;
  int64_t noPaths = 10000;
  int basisDegree = 3;
  bool standardize = true;
  bool itmOnly = true;
  bool pooled = false;
  double hysteresis = 1.0;
  int valueMode = ROADMC_VALUE_REGRESSED;
  uint64_t seed = 0;
};

// Inputs of the species-state valuation (ROVMethod ALGO4 / ALGO6) that the
// reference reads through sim->getRoad()->getOptimiser(), its traffic
// program and its commodities (SimulateGPU.cu:3880-4070).
struct RovSpeciesInputs {
  std::vector<double> flows;                      // per control; the last is the highest flow
  int nFuels = 0;                                 // uncertainties 0 .. nFuels - 1 are fuels
  std::vector<roadmc_uncertainty> uncertainties;  // fuels, then one per commodity
  std::vector<double> oreMean, oreSD;             // per commodity
  std::vector<double> fuelCost;                   // per vehicle class
  std::vector<int32_t> fuelIdx;                   // per vehicle class: its fuel
  double unitCost = 0.0;
  double unitRevenue = 0.0;
  int nYears = 50;
  double timeStep = 1.0;
  double rrr = 7.0;
  int64_t noPaths = 10000;
  int dimRes = 20;
  double profitScaling = 2.0;  // SimulateGPU.cu:3900
  double hysteresis = 0.975;   // SimulateGPU.cu:2763
  uint64_t seed = 0;
  uint32_t roadId = 0;
  int method = ROADMC_ROVS_ALGO4;
  int checkpointYears = 0;
  int precision = ROADMC_PREC_FP32;
};

// The road attributes simulateROVCUDA stores (SimulateGPU.cu:5461-5490).
struct RovStats {
  double value = 0.0;
  double valueSD = 0.0;
  int firstControl = 0;
};

struct SimulatorInputs {
  int nYears = 50;
  double timeStep = 1.0;
  int64_t noPaths = 10000;
  int growthLaw = ROADMC_LAW_LOGNORMAL;
  int endpoint = ROADMC_END_POST_GROWTH;
  int precision = ROADMC_PREC_FP32;
  double dropTol = 0.0;
  uint64_t seed = 0;
  uint32_t roadId = 0;
};

class SimulateGPU {
 public:
  // SimulateGPU::deviceCount (SimulateGPU.h:44).
  static int deviceCount() {
    int n = 0;
    roadmc_device_count(&n);
    return n;
  }

  // SimulateGPU::simulateMTECUDA (SimulateGPU.h:152-156): endPops must be
  // noPaths x srp.size(); column s receives species s's end totals.
  static void simulateMTECUDA(const SimulatorInputs& sim, std::vector<SpeciesRoadPatches>& srp,
                              std::vector<std::vector<double>>& initPops,
                              std::vector<std::vector<double>>& capacities, Matrix& endPops,
                              int device = 0) {
    if (srp.size() != initPops.size() || srp.size() != capacities.size())
      fail("simulateMTECUDA: species arrays differ in length");
    if (endPops.rows != sim.noPaths || endPops.cols != (int64_t)srp.size())
      fail("simulateMTECUDA: endPops must be noPaths x species");
    std::vector<roadmc_species> sp(srp.size());
    for (size_t s = 0; s < srp.size(); ++s) {
      const SpeciesRoadPatches& r = srp[s];
      roadmc_species& o = sp[s];
      o.n_patches = (int32_t)initPops[s].size();
      if ((int64_t)capacities[s].size() != o.n_patches || r.transProbs.rows != o.n_patches ||
          r.transProbs.cols != o.n_patches)
        fail("simulateMTECUDA: patch arrays differ in size");
      o.init_pop = initPops[s].data();
      o.capacity = capacities[s].data();
      o.growth = r.growthRates.empty() ? nullptr : r.growthRates.data();
      o.trans = r.transProbs.data.data();
      // The last control is the one the reference simulates (survProbs[last]).
      o.surv = r.survProbs.empty() ? nullptr : r.survProbs.back().data.data();
      for (int i = 0; i < 8; ++i) o.params[i] = r.speciesParams[i];
      o.noise_sd = r.noiseSD;
      o.threshold_ratio = r.thresholdRatio;
    }
    roadmc_mte_problem pr{};
    pr.n_species = (int32_t)sp.size();
    pr.species = sp.data();
    pr.n_years = sim.nYears;
    pr.time_step = sim.timeStep;
    pr.n_paths = sim.noPaths;
    pr.growth_law = sim.growthLaw;
    pr.endpoint = sim.endpoint;
    pr.precision = sim.precision;
    pr.drop_tol = sim.dropTol;
    pr.seed = sim.seed;
    pr.road_id = sim.roadId;
    roadmc_mte_result res{};
    res.end_pops = endPops.data.data();
    roadmc_ctx* ctx = context(device);
    if (roadmc_mte_run(ctx, &pr, &res) != ROADMC_OK) fail(roadmc_last_error(ctx));
  }

  // SimulateGPU::simulateROVCUDA (SimulateGPU.h:187-191) for the price-state
  // model: condExp is (nYears+1) x noPaths and optCont nYears x noPaths
  // (step-major, like the reference's T x N matrices); regressions receives
  // nYears * controls * K coefficients.  Outputs are resized here.
  static void simulateROVCUDA(const RovInputs& in, Matrix& condExp, MatrixI& optCont,
                              std::vector<double>& regressions, RovStats& stats,
                              int device = 0) {
    const int U = (int)in.uncertainties.size();
    const int C = (int)in.flows.size();
    if ((int)in.profitWeights.size() != U) fail("simulateROVCUDA: one profit weight per uncertainty");
    if (!in.correlation.empty() && (int)in.correlation.size() != U * U)
      fail("simulateROVCUDA: correlation must be U x U");
    int32_t K = 0;
    if (roadmc_rov_basis_size(U, in.basisDegree, &K) != ROADMC_OK) fail("simulateROVCUDA: basis");
    roadmc_rov_problem pr{};
    pr.n_uncertainties = U;
    pr.unc = in.uncertainties.data();
    pr.correlation = in.correlation.empty() ? nullptr : in.correlation.data();
    pr.n_controls = C;
    pr.flow = in.flows.data();
    pr.unit_cost = in.unitCost;
    pr.profit_weight = in.profitWeights.data();
    pr.traffic_index = in.trafficIndex;
    pr.n_steps = in.nYears;
    pr.time_step = in.timeStep;
    pr.rrr = in.rrr;
    pr.n_paths = in.noPaths;
    pr.basis_degree = in.basisDegree;
    pr.standardize = in.standardize ? 1 : 0;
    pr.itm_only = in.itmOnly ? 1 : 0;
    pr.pooled = in.pooled ? 1 : 0;
    pr.hysteresis = in.hysteresis;
    pr.mode = ROADMC_ROV_SWITCHING;
    pr.value_mode = in.valueMode;
    pr.seed = in.seed;
    Matrix ce(in.nYears + 1, in.noPaths);
    MatrixI oc(in.nYears, in.noPaths);
    std::vector<double> reg((size_t)in.nYears * C * K);
    // The library writes [step][path] row-major, which is the transpose of
    // a column-major (steps x paths) matrix; stage and reorder.
    std::vector<double> ce_rm(ce.data.size());
    std::vector<int32_t> oc_rm(oc.data.size());
    RovStats st;
    roadmc_rov_result res{};
    res.value = &st.value;
    res.value_sd = &st.valueSD;
    res.first_control = &st.firstControl;
    res.coeffs = reg.data();
    res.cond_exp = ce_rm.data();
    res.opt_control = oc_rm.data();
    roadmc_ctx* ctx = context(device);
    if (roadmc_rov_run(ctx, &pr, &res) != ROADMC_OK) fail(roadmc_last_error(ctx));
    for (int64_t t = 0; t <= in.nYears; ++t)
      for (int64_t i = 0; i < in.noPaths; ++i) ce(t, i) = ce_rm[(size_t)(t * in.noPaths + i)];
    for (int64_t t = 0; t < in.nYears; ++t)
      for (int64_t i = 0; i < in.noPaths; ++i) oc(t, i) = oc_rm[(size_t)(t * in.noPaths + i)];
    condExp = std::move(ce);
    optCont = std::move(oc);
    regressions = std::move(reg);
    stats = st;
  }

  // SimulateGPU::simulateROVCUDA (SimulateGPU.h:187-191, SimulateGPU.cu:
  // 3870-5525) for the reference's species states, ALGO4 or ALGO6 by
  // sim.method.  Outputs are resized here to the reference's shapes:
  // adjPops[t] is noPaths x species (the highest-flow adjusted populations of
  // year t), unitProfits, condExp and optCont are noPaths x nYears, and
  // regressions holds each year's and control's fitted grid, nYears x
  // controls x (dims dimRes + dimRes^dims) with dims = species + 1 (the
  // grid axes, then the grid values; the reference's vector reserves a
  // second dimRes^dims block per grid that it never fills).  plotResults is
  // accepted and ignored (the reference only plots with it).  stats receives
  // the value, its SD and the first control, which the reference stores on
  // the road (SimulateGPU.cu:5485-5488).
  static void simulateROVCUDA(const RovSpeciesInputs& sim, std::vector<SpeciesRoadPatches>& srp,
                              std::vector<Matrix>& adjPops, Matrix& unitProfits, Matrix& condExp,
                              MatrixI& optCont, std::vector<double>& regressions,
                              bool plotResults, RovStats& stats, int device = 0) {
    (void)plotResults;
    const int S = (int)srp.size();
    const int C = (int)sim.flows.size();
    if (S < 1) fail("simulateROVCUDA: no species");
    const int64_t P = (int64_t)srp[0].initPops.size();
    const int nComm = (int)sim.uncertainties.size() - sim.nFuels;
    if (nComm < 0 || (int)sim.oreMean.size() != nComm || (int)sim.oreSD.size() != nComm)
      fail("simulateROVCUDA: one ore mean and SD per commodity uncertainty");
    if (sim.fuelCost.size() != sim.fuelIdx.size())
      fail("simulateROVCUDA: fuel cost and fuel index per vehicle class");
    std::vector<roadmc_rovs_species> sp((size_t)S);
    std::vector<std::vector<double>> surv((size_t)S);
    for (int s = 0; s < S; ++s) {
      const SpeciesRoadPatches& r = srp[(size_t)s];
      if ((int64_t)r.initPops.size() != P || (int64_t)r.capacities.size() != P ||
          r.transProbs.rows != P || r.transProbs.cols != P || (int)r.survProbs.size() != C)
        fail("simulateROVCUDA: every species needs P patches and one survival matrix per control");
      surv[(size_t)s].reserve((size_t)(C * P * P));
      for (const Matrix& m : r.survProbs) {
        if (m.rows != P || m.cols != P) fail("simulateROVCUDA: survival matrices must be P x P");
        surv[(size_t)s].insert(surv[(size_t)s].end(), m.data.begin(), m.data.end());
      }
      roadmc_rovs_species& o = sp[(size_t)s];
      o.init_pop = r.initPops.data();
      o.capacity = r.capacities.data();
      o.trans = r.transProbs.data.data();
      o.surv = surv[(size_t)s].data();
      for (int i = 0; i < 8; ++i) o.params[i] = r.speciesParams[i];
    }
    roadmc_rovs_problem pr{};
    pr.n_patches = (int32_t)P;
    pr.n_species = S;
    pr.species = sp.data();
    pr.n_controls = C;
    pr.flow = sim.flows.data();
    pr.n_fuels = sim.nFuels;
    pr.n_commodities = nComm;
    pr.uncertainties = sim.uncertainties.data();
    pr.ore_mean = sim.oreMean.data();
    pr.ore_sd = sim.oreSD.data();
    pr.n_vehicles = (int32_t)sim.fuelCost.size();
    pr.fuel_cost = sim.fuelCost.data();
    pr.fuel_idx = sim.fuelIdx.data();
    pr.unit_cost = sim.unitCost;
    pr.unit_revenue = sim.unitRevenue;
    pr.n_years = sim.nYears;
    pr.time_step = sim.timeStep;
    pr.rrr = sim.rrr;
    pr.n_paths = sim.noPaths;
    pr.dim_res = sim.dimRes;
    pr.profit_scaling = sim.profitScaling;
    pr.hysteresis = sim.hysteresis;
    pr.seed = sim.seed;
    pr.road_id = sim.roadId;
    pr.method = sim.method;
    pr.checkpoint_years = sim.checkpointYears;
    pr.precision = sim.precision;
    const int64_t N = sim.noPaths, T = sim.nYears;
    const int D = S + 1;
    int64_t grid = (int64_t)D * sim.dimRes, cells = 1;
    for (int d = 0; d < D; ++d) cells *= sim.dimRes;
    grid += cells;
    // [T][N] row-major is the column-major noPaths x nYears matrix itself.
    Matrix ce(N, T), up(N, T);
    MatrixI oc(N, T);
    std::vector<double> adj((size_t)(T * S * N));
    std::vector<double> reg((size_t)(T * C * grid));
    RovStats st;
    roadmc_rovs_result res{};
    res.value = &st.value;
    res.value_sd = &st.valueSD;
    res.first_control = &st.firstControl;
    res.cond_exp = ce.data.data();
    res.opt_control = oc.data.data();
    res.unit_profit = up.data.data();
    res.adj_pops = adj.data();
    res.grids = reg.data();
    roadmc_ctx* ctx = context(device);
    if (roadmc_rovs_run(ctx, &pr, &res) != ROADMC_OK) fail(roadmc_last_error(ctx));
    // Year t's [S][N] block is the column-major noPaths x species matrix.
    std::vector<Matrix> ap((size_t)T);
    for (int64_t t = 0; t < T; ++t) {
      ap[(size_t)t] = Matrix(N, S);
      std::copy(adj.begin() + t * S * N, adj.begin() + (t + 1) * S * N, ap[(size_t)t].data.begin());
    }
    adjPops = std::move(ap);
    unitProfits = std::move(up);
    condExp = std::move(ce);
    optCont = std::move(oc);
    regressions = std::move(reg);
    stats = st;
  }

  // SimulateGPU::expPV (SimulateGPU.h:53): stores the expected present value
  // and its SD on the uncertainty.
  static void expPV(Uncertainty& u, int nYears, double timeStep, double rrr, double growth,
                    int64_t noPaths, uint64_t seed, int device = 0) {
    roadmc_ctx* ctx = context(device);
    double m = 0.0, sd = 0.0;
    if (roadmc_exp_pv(ctx, &u.params, nYears, timeStep, rrr, growth, noPaths, seed, &m, &sd) !=
        ROADMC_OK)
      fail(roadmc_last_error(ctx));
    u.expPV = m;
    u.expPVSD = sd;
  }

 private:
  [[noreturn]] static void fail(const std::string& msg) {
    thread_local std::string buf;
    buf = msg;
    throw buf.c_str();
  }

  // One context per host thread and device: contexts are not shared between
  // threads, so calls from a thread pool do not contend.
  static roadmc_ctx* context(int device) {
    struct Holder {
      std::map<int, roadmc_ctx*> m;
      ~Holder() {
        for (auto& kv : m) roadmc_ctx_destroy(kv.second);
      }
    };
    thread_local Holder h;
    auto it = h.m.find(device);
    if (it != h.m.end()) return it->second;
    roadmc_ctx* ctx = nullptr;
    if (roadmc_ctx_create(device, &ctx) != ROADMC_OK) fail("CUDA Error.");
    h.m[device] = ctx;
    return ctx;
  }
};

}  // namespace roadmc

#endif
