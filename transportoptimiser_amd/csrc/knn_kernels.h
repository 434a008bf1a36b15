// Host-visible declarations of the grid-surrogate kernels (knn_kernels.hip).
#ifndef ROADMC_KNN_KERNELS_H
#define ROADMC_KNN_KERNELS_H

#include <hip/hip_runtime.h>

namespace roadmc {

struct KnnFitArgs {
  int n, d, R, k, clamp;  // samples, dimensions, grid points per dimension, neighbours
  int cap;                // per-lane list capacity: min(k, ceil(n / 64))
  long long G;            // grid points, R^d
  const double* u;        // [n][d] sample coordinates of the search and the fit
  const double* y;        // [n]
  const double* qtab;     // [d][R] query coordinate of grid index i in dimension dd
  double* val;            // [G]
};

struct KnnEvalArgs {
  int d, R, m;
  const double* surrogate;  // [d*R + R^d]
  const double* pred;       // [m][d]
  double* out;              // [m]
};

constexpr int kKnnThreads = 64;  // one wave per workgroup
constexpr int kKnnMaxK = 128;

constexpr int kKnnSelectMin = 512;   // above this many samples: the radix-select kernel
constexpr int kKnnSelectCap = 2048;  // largest bucket of the k-th key sorted in LDS
constexpr int kKnnMaxBatch = 8;      // fits of one launch (blockIdx.y)

// Fits launched together, one per blockIdx.y; grid points beyond a fit's G
// exit at once.
struct KnnFitBatch {
  int nfits;
  KnnFitArgs f[kKnnMaxBatch];
};

// Dynamic LDS of the wave-per-point fit kernel: per-lane lists of cap
// entries, then the k merged neighbours.
inline size_t knn_fit_lds(int cap, int k) {
  return ((size_t)cap * kKnnThreads + (size_t)k) * (sizeof(double) + sizeof(int));
}
// The fit's per-neighbour terms and sums after the lists (D <= 3): nterms x k
// doubles and nterms more, starting knn_fit_term_offset doubles past nbd (the
// k neighbours' distances, then their k indices, rounded up to doubles).
__host__ __device__ inline size_t knn_fit_term_offset(int cap, int k) {
  (void)cap;
  return (size_t)k + ((size_t)k + 1) / 2;
}
inline size_t knn_fit_term_lds(int k, int nterms) {
  return ((size_t)nterms * k + nterms + 1) * sizeof(double);
}

hipError_t launch_knn_fit(const KnnFitArgs& a, hipStream_t s);
// Up to kKnnMaxBatch fits of the same dimension d in at most two launches
// (one per fit kernel), so that small fits fill the machine together.
hipError_t launch_knn_fit_batch(const KnnFitArgs* fits, int nfits, hipStream_t s);
hipError_t launch_knn_eval(const KnnEvalArgs& a, hipStream_t s);

}  // namespace roadmc

#endif
