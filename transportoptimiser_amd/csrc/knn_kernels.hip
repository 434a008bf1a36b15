// Grid surrogate for gfx950: kNN local-linear regression on a dim_res^d grid
// and its multilinear evaluation.
//
// Reference: SimulateGPU::buildSurrogateMTECUDA / buildSurrogateROVCUDA
// (include/gpuCode/SimulateGPU.cu:5533-5917) run a kNN search
// (knn_cuda_with_indexes, FP32, host round trip) and then multiLocLinReg
// (:2934-3053), one thread per grid point with per-thread device malloc and a
// Gaussian-elimination solve in FP32; interpolateMulti (:3077-3155) and
// Utility::interpolateSurrogate (Utility.cpp:253-363) evaluate the grid.
//
// Here one wave per grid point does both the search and the fit: lane L
// scans the samples j = L (mod 64) and keeps its best (squared distance,
// index) pairs sorted in LDS (laid out [rank][lane], conflict-free); k rounds
// of a wave-wide minimum over the list heads then select the same k
// neighbours, in the same order, as a stable sort of all samples; the
// (d+1) x (d+1) weighted normal equations are summed per lane, added across
// lanes in a fixed butterfly, and Cholesky-solved in FP64 registers
// (templated on d so that nothing spills).
// Compiled without multiply-add contraction, so that the neighbour ranking
// and the sums follow the same roundings as the CPU statement.
#include <hip/hip_runtime.h>

#include "knn_kernels.h"

namespace roadmc {

template <int D>
__global__ __launch_bounds__(kKnnThreads) void knn_fit_kernel(KnnFitArgs a) {
  extern __shared__ double knn_lds[];
  const int cap = a.cap, k = a.k, R = a.R, n = a.n;
  double* topd = knn_lds;                                              // [cap][64]
  int* topi = reinterpret_cast<int*>(knn_lds + (size_t)cap * kKnnThreads);  // [cap][64]
  const int lane = threadIdx.x;
  double q[D];
  {
    long long rem = blockIdx.x;  // one wave per grid point
#pragma unroll
    for (int dd = D - 1; dd >= 0; --dd) {
      const int i = (int)(rem % R);
      rem /= R;
      q[dd] = a.qtab[dd * R + i];
    }
  }
  // 1. Lane-local sorted lists of the best `cap` samples j = lane (mod 64),
  //    ranked by (squared distance, index): samples arrive in index order and
  //    an equal distance never displaces an earlier one.
  int cnt = 0;
  double worst = __builtin_inf();
  for (int j = lane; j < n; j += kKnnThreads) {
    double s = 0.0;
#pragma unroll
    for (int dd = 0; dd < D; ++dd) {
      const double e = a.u[(size_t)j * D + dd] - q[dd];
      s += e * e;
    }
    if (s < worst) {
      int r = cnt < cap ? cnt : cap - 1;
      while (r > 0 && topd[(r - 1) * kKnnThreads + lane] > s) {
        topd[r * kKnnThreads + lane] = topd[(r - 1) * kKnnThreads + lane];
        topi[r * kKnnThreads + lane] = topi[(r - 1) * kKnnThreads + lane];
        --r;
      }
      topd[r * kKnnThreads + lane] = s;
      topi[r * kKnnThreads + lane] = j;
      if (cnt < cap) ++cnt;
      worst = cnt == cap ? topd[(cap - 1) * kKnnThreads + lane] : __builtin_inf();
    }
  }
  // 2. Merge: k rounds of a wave-wide minimum over the list heads; the lane
  //    owning the minimum (indices are distinct) advances.  Lane L's selected
  //    neighbours are then the first head_L entries of its list.
  int head = 0;
  double kth = 0.0;
  for (int r = 0; r < k; ++r) {
    const double hd = head < cnt ? topd[head * kKnnThreads + lane] : __builtin_inf();
    const int hi = head < cnt ? topi[head * kKnnThreads + lane] : 0x7fffffff;
    double md = hd;
    int mi = hi;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double od = __shfl_xor(md, off);
      const int oi = __shfl_xor(mi, off);
      if (od < md || (od == md && oi < mi)) {
        md = od;
        mi = oi;
      }
    }
    if (hi == mi) ++head;
    kth = md;
  }
  // 3. Weighted local-linear fit (multiLocLinReg): basis (1, x - q), weights
  //    exp(-(d/h)^2 / 2) / sqrt(2 pi), h = k-th neighbour distance; each lane
  //    sums its own neighbours, then a fixed butterfly adds the lanes.
  constexpr int D1 = D + 1;
  double A[D1][D1], B[D1];
#pragma unroll
  for (int i = 0; i < D1; ++i) {
    B[i] = 0.0;
#pragma unroll
    for (int j = 0; j < D1; ++j) A[i][j] = 0.0;
  }
  const double h = sqrt(kth);
  for (int r = 0; r < head; ++r) {
    const double dist = sqrt(topd[r * kKnnThreads + lane]);
    const int j = topi[r * kKnnThreads + lane];
    const double tt = h > 0.0 ? dist / h : 0.0;
    const double w = exp(-tt * tt / 2) / sqrt(2.0 * 3.14159265358979323846);
    double v[D1];
    v[0] = 1.0;
#pragma unroll
    for (int dd = 0; dd < D; ++dd) v[dd + 1] = a.u[(size_t)j * D + dd] - q[dd];
    const double yr = a.y[j];
#pragma unroll
    for (int i = 0; i < D1; ++i) {
      B[i] += (i == 0 ? yr : yr * v[i]) * w;
#pragma unroll
      for (int jj = 0; jj <= i; ++jj) A[i][jj] += (i == 0 && jj == 0 ? 1.0 : v[i] * v[jj]) * w;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int i = 0; i < D1; ++i) {
      B[i] += __shfl_xor(B[i], off);
#pragma unroll
      for (int jj = 0; jj <= i; ++jj) A[i][jj] += __shfl_xor(A[i][jj], off);
    }
  }
  if (lane != 0) return;
  // Cholesky A = L L^T; a non-positive pivot (neighbours that do not span
  // the dimensions) gives the weighted mean B0 / A00.
  double amax = 0.0;
#pragma unroll
  for (int i = 0; i < D1; ++i) amax = fmax(amax, fabs(A[i][i]));
  bool ok = true;
  double L[D1][D1];
#pragma unroll
  for (int i = 0; i < D1; ++i) {
#pragma unroll
    for (int jj = 0; jj <= i; ++jj) {
      double s = A[i][jj];
#pragma unroll
      for (int c = 0; c < jj; ++c) s -= L[i][c] * L[jj][c];
      if (i == jj) {
        ok = ok && (s > 1e-12 * amax) && __builtin_isfinite(s);
        L[i][i] = sqrt(ok ? s : 1.0);
      } else {
        L[i][jj] = s / L[jj][jj];
      }
    }
  }
  double z[D1];
#pragma unroll
  for (int i = 0; i < D1; ++i) {
    double s = B[i];
#pragma unroll
    for (int c = 0; c < i; ++c) s -= L[i][c] * z[c];
    z[i] = s / L[i][i];
  }
#pragma unroll
  for (int i = D1 - 1; i >= 0; --i) {
    double s = z[i];
#pragma unroll
    for (int c = i + 1; c < D1; ++c) s -= L[c][i] * z[c];
    z[i] = s / L[i][i];
  }
  double v = ok ? z[0] : B[0] / A[0][0];
  if (a.clamp && v > 0.0) v = 0.0;  // rovCorrection (SimulateGPU.cu:3055-3073)
  a.val[blockIdx.x] = v;
}

// Multilinear interpolation of one point in its grid cell.
template <int D>
__global__ __launch_bounds__(256) void knn_eval_kernel(KnnEvalArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.m) return;
  const int R = a.R;
  const double* grid = a.surrogate;
  const double* val = a.surrogate + (size_t)D * R;
  int li[D];
  double xd[D];
#pragma unroll
  for (int dd = 0; dd < D; ++dd) {
    const double lo = grid[dd * R], hi = grid[dd * R + R - 1];
    const double pv = a.pred[(size_t)p * D + dd];
    const double c = hi != lo ? (R - 1) * (pv - lo) / (hi - lo) : 0.0;
    const int i = c >= (double)(R - 2) ? R - 2 : (c > 0.0 ? (int)c : 0);
    li[dd] = i;
    const double x0 = grid[dd * R + i], x1 = grid[dd * R + i + 1];
    xd[dd] = (fabs(x1 - x0) < 2.220446049250313e-16 || x0 == x1) ? 0.0 : (pv - x0) / (x1 - x0);
  }
  double s = 0.0;
  for (int c = 0; c < (1 << D); ++c) {
    double w = 1.0;
    size_t idx = 0;
#pragma unroll
    for (int dd = 0; dd < D; ++dd) {
      const int bit = (c >> (D - 1 - dd)) & 1;
      w *= bit ? xd[dd] : 1.0 - xd[dd];
      idx = idx * (size_t)R + (size_t)(li[dd] + bit);
    }
    s += val[idx] * w;
  }
  a.out[p] = s;
}

template <int D>
static hipError_t fit_d(const KnnFitArgs& a, hipStream_t s) {
  const size_t lds = knn_fit_lds(a.cap);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&knn_fit_kernel<D>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(knn_fit_kernel<D>, dim3((unsigned)a.G), dim3(kKnnThreads), lds, s, a);
  return hipGetLastError();
}

template <int D>
static hipError_t eval_d(const KnnEvalArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(knn_eval_kernel<D>, dim3((a.m + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_knn_fit(const KnnFitArgs& a, hipStream_t s) {
  switch (a.d) {
    case 1: return fit_d<1>(a, s);
    case 2: return fit_d<2>(a, s);
    case 3: return fit_d<3>(a, s);
    case 4: return fit_d<4>(a, s);
    case 5: return fit_d<5>(a, s);
    case 6: return fit_d<6>(a, s);
    case 7: return fit_d<7>(a, s);
    case 8: return fit_d<8>(a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_knn_eval(const KnnEvalArgs& a, hipStream_t s) {
  switch (a.d) {
    case 1: return eval_d<1>(a, s);
    case 2: return eval_d<2>(a, s);
    case 3: return eval_d<3>(a, s);
    case 4: return eval_d<4>(a, s);
    case 5: return eval_d<5>(a, s);
    case 6: return eval_d<6>(a, s);
    case 7: return eval_d<7>(a, s);
    case 8: return eval_d<8>(a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace roadmc
