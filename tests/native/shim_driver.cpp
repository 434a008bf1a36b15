// Drives the C++ shim (roadmc.hpp) the way Simulator::simulateMTE drives
// SimulateGPU::simulateMTECUDA (Simulator.cpp:108-119): caller-sized endPops,
// errors caught as const char*.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "roadmc.hpp"

int main(int argc, char** argv) {
  const bool no_gpu_expected = argc > 1 && std::strcmp(argv[1], "--no-gpu-expected") == 0;
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
