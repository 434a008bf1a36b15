// Drives the C++ shim (roadmc.hpp) the way Simulator::simulateMTE and
// Simulator::simulateROV drive SimulateGPU (Simulator.cpp:108-119): caller-
// sized outputs, errors caught as const char*.
//
//   shim_driver --no-gpu-expected   the built-in MTE road; the call must throw
//   shim_driver                     the built-in MTE road on the GPU
//   shim_driver --mte IN OUT        an MTE road read from IN; end totals to OUT
//   shim_driver --rovs IN OUT       a species-state valuation read from IN;
//                                   its outputs to OUT
//
// IN is whitespace-separated numbers in the order the readers below take
// them (matrices row-major, origin = row); OUT is one number per line.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "roadmc.hpp"

namespace {

struct Reader {
  std::FILE* f;
  double num() {
    double v;
    if (std::fscanf(f, "%lf", &v) != 1) throw std::runtime_error("short input");
    return v;
  }
  int64_t integer() { return (int64_t)std::llround(num()); }
  std::vector<double> vec(int64_t n) {
    std::vector<double> v((size_t)n);
    for (double& x : v) x = num();
    return v;
  }
  roadmc::Matrix square(int64_t P) {  // row-major in the file
    roadmc::Matrix m(P, P);
    for (int64_t i = 0; i < P; ++i)
      for (int64_t j = 0; j < P; ++j) m(i, j) = num();
    return m;
  }
};

int builtin_mte(bool no_gpu_expected) {
  const int P = 8;
  roadmc::SpeciesRoadPatches sp;
  sp.transProbs = roadmc::Matrix(P, P);
  for (int i = 0; i < P; ++i)
    for (int j = 0; j < P; ++j) sp.transProbs(i, j) = (i == j) ? 0.6 : 0.4 / (P - 1);
  sp.survProbs.push_back(roadmc::Matrix(P, P));
  for (double& v : sp.survProbs[0].data) v = 0.95;
  sp.growthRates.assign(P, 0.05);
  std::vector<roadmc::SpeciesRoadPatches> srp{sp};
  std::vector<std::vector<double>> init{std::vector<double>(P, 800.0)};
  std::vector<std::vector<double>> caps{std::vector<double>(P, 1000.0)};
  roadmc::SimulatorInputs sim;
  sim.nYears = 10;
  sim.noPaths = 64;
  sim.seed = 3;
  roadmc::Matrix endPops(sim.noPaths, 1);
  try {
    roadmc::SimulateGPU::simulateMTECUDA(sim, srp, init, caps, endPops);
  } catch (const char* err) {
    std::printf("caught: %s\n", err);
    return no_gpu_expected ? 0 : 1;
  }
  if (no_gpu_expected) {
    std::printf("expected an error without a GPU\n");
    return 1;
  }
  double mean = 0.0;
  for (int64_t i = 0; i < sim.noPaths; ++i) mean += endPops(i, 0);
  mean /= (double)sim.noPaths;
  std::printf("mean end population %.6f\n", mean);
  // Malformed call: endPops of the wrong shape must throw.
  roadmc::Matrix bad(3, 1);
  try {
    roadmc::SimulateGPU::simulateMTECUDA(sim, srp, init, caps, bad);
    return 1;
  } catch (const char* err) {
    std::printf("caught: %s\n", err);
  }
  return std::isfinite(mean) && mean > 0.0 ? 0 : 1;
}

// P nYears noPaths seed growthLaw; trans; surv; growth; init; cap; params[8];
// noiseSD thresholdRatio.
int file_mte(std::FILE* in, std::FILE* out) {
  Reader r{in};
  const int64_t P = r.integer();
  roadmc::SimulatorInputs sim;
  sim.nYears = (int)r.integer();
  sim.noPaths = r.integer();
  sim.seed = (uint64_t)r.integer();
  sim.growthLaw = (int)r.integer();
  roadmc::SpeciesRoadPatches sp;
  sp.transProbs = r.square(P);
  sp.survProbs.push_back(r.square(P));
  sp.growthRates = r.vec(P);
  std::vector<std::vector<double>> init{r.vec(P)};
  std::vector<std::vector<double>> caps{r.vec(P)};
  for (double& v : sp.speciesParams) v = r.num();
  sp.noiseSD = r.num();
  sp.thresholdRatio = r.num();
  std::vector<roadmc::SpeciesRoadPatches> srp{sp};
  roadmc::Matrix endPops(sim.noPaths, 1);
  roadmc::SimulateGPU::simulateMTECUDA(sim, srp, init, caps, endPops);
  for (int64_t i = 0; i < sim.noPaths; ++i) std::fprintf(out, "%.17g\n", endPops(i, 0));
  return 0;
}

// P S C; per species: trans, C surv matrices, init, cap, params[8]; flows[C];
// nFuels nCommodities; per uncertainty: kind x0 mean sd reversion jumpSize
// jumpProb; oreMean, oreSD; nVehicles, fuelCost, fuelIdx; unitCost
// unitRevenue nYears timeStep rrr noPaths dimRes profitScaling hysteresis
// seed roadId method.
int file_rovs(std::FILE* in, std::FILE* out) {
  Reader r{in};
  const int64_t P = r.integer();
  const int S = (int)r.integer(), C = (int)r.integer();
  std::vector<roadmc::SpeciesRoadPatches> srp((size_t)S);
  for (auto& sp : srp) {
    sp.transProbs = r.square(P);
    for (int c = 0; c < C; ++c) sp.survProbs.push_back(r.square(P));
    sp.initPops = r.vec(P);
    sp.capacities = r.vec(P);
    for (double& v : sp.speciesParams) v = r.num();
  }
  roadmc::RovSpeciesInputs sim;
  sim.flows = r.vec(C);
  sim.nFuels = (int)r.integer();
  const int nComm = (int)r.integer();
  for (int u = 0; u < sim.nFuels + nComm; ++u) {
    roadmc_uncertainty q{};
    q.kind = (int32_t)r.integer();
    q.x0 = r.num();
    q.mean = r.num();
    q.sd = r.num();
    q.reversion = r.num();
    q.jump_size = r.num();
    q.jump_prob = r.num();
    sim.uncertainties.push_back(q);
  }
  sim.oreMean = r.vec(nComm);
  sim.oreSD = r.vec(nComm);
  const int64_t nVeh = r.integer();
  sim.fuelCost = r.vec(nVeh);
  for (int64_t v = 0; v < nVeh; ++v) sim.fuelIdx.push_back((int32_t)r.integer());
  sim.unitCost = r.num();
  sim.unitRevenue = r.num();
  sim.nYears = (int)r.integer();
  sim.timeStep = r.num();
  sim.rrr = r.num();
  sim.noPaths = r.integer();
  sim.dimRes = (int)r.integer();
  sim.profitScaling = r.num();
  sim.hysteresis = r.num();
  sim.seed = (uint64_t)r.integer();
  sim.roadId = (uint32_t)r.integer();
  sim.method = (int)r.integer();
  std::vector<roadmc::Matrix> adjPops;
  roadmc::Matrix unitProfits, condExp;
  roadmc::MatrixI optCont;
  std::vector<double> regressions;
  roadmc::RovStats stats;
  roadmc::SimulateGPU::simulateROVCUDA(sim, srp, adjPops, unitProfits, condExp, optCont,
                                       regressions, false, stats);
  const int64_t N = sim.noPaths, T = sim.nYears;
  std::fprintf(out, "%.17g\n%.17g\n%d\n", stats.value, stats.valueSD, stats.firstControl);
  // Year-major, path-minor: the [T][N] order of the C ABI.
  for (int64_t t = 0; t < T; ++t)
    for (int64_t i = 0; i < N; ++i) std::fprintf(out, "%.17g\n", condExp(i, t));
  for (int64_t t = 0; t < T; ++t)
    for (int64_t i = 0; i < N; ++i) std::fprintf(out, "%d\n", optCont(i, t));
  for (int64_t t = 0; t < T; ++t)
    for (int64_t i = 0; i < N; ++i) std::fprintf(out, "%.17g\n", unitProfits(i, t));
  for (int64_t t = 0; t < T; ++t)
    for (int s = 0; s < S; ++s)
      for (int64_t i = 0; i < N; ++i) std::fprintf(out, "%.17g\n", adjPops[(size_t)t](i, s));
  std::fprintf(out, "%zu\n", regressions.size());
  for (double v : regressions) std::fprintf(out, "%.17g\n", v);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 4 && (std::strcmp(argv[1], "--mte") == 0 || std::strcmp(argv[1], "--rovs") == 0)) {
    std::FILE* in = std::fopen(argv[2], "r");
    std::FILE* out = std::fopen(argv[3], "w");
    if (!in || !out) {
      std::printf("cannot open %s or %s\n", argv[2], argv[3]);
      return 2;
    }
    int rc = 1;
    try {
      rc = std::strcmp(argv[1], "--mte") == 0 ? file_mte(in, out) : file_rovs(in, out);
    } catch (const char* err) {
      std::printf("caught: %s\n", err);
    } catch (const std::exception& e) {
      std::printf("input error: %s\n", e.what());
    }
    std::fclose(in);
    std::fclose(out);
    return rc;
  }
  return builtin_mte(argc > 1 && std::strcmp(argv[1], "--no-gpu-expected") == 0);
}
