// C++ host interface over the C ABI, with the call shape of the reference's
// SimulateGPU class (include/gpuCode/SimulateGPU.h:197-344) so that a
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
  // SimulateGPU::deviceCount (SimulateGPU.h:197).
  static int deviceCount() {
    int n = 0;
    roadmc_device_count(&n);
    return n;
  }

  // SimulateGPU::simulateMTECUDA (SimulateGPU.h:305-309): endPops must be
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
