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
// Here, for up to kKnnSelectMin samples, one wave per grid point does both
// the search and the fit: lane L
// scans the samples j = L (mod 64) and keeps its best (squared distance,
// index) pairs sorted in LDS (laid out [rank][lane], conflict-free); k rounds
// of a wave-wide minimum over the list heads then select the same k
// neighbours, in the same order, as a stable sort of all samples; the
// (d+1) x (d+1) weighted normal equations are summed by one lane in that
// order and Cholesky-solved in FP64 registers (templated on d so that nothing
// spills).  For more samples, one workgroup per grid point radix-selects the
// k-th neighbour exactly (knn_select_kernel).
// Compiled without multiply-add contraction, so that the neighbour ranking
// and the sums follow the same roundings as the CPU statement.
#include <hip/hip_runtime.h>

#include "knn_kernels.h"

namespace roadmc {

// Weighted local-linear fit (multiLocLinReg, SimulateGPU.cu:2934-3053) over
// the k neighbours nbd / nbi, sorted by (squared distance, index), summed in
// that order by one thread: basis (1, x - q), weights
// exp(-(d/h)^2 / 2) / sqrt(2 pi), h = the k-th neighbour's distance; FP64
// Cholesky, the weighted mean when the system is singular.
template <int D>
__device__ double knn_regress(const double* nbd, const int* nbi, int k, const double* q,
                              const double* __restrict__ u, const double* __restrict__ y) {
  constexpr int D1 = D + 1;
  double A[D1][D1], B[D1];
#pragma unroll
  for (int i = 0; i < D1; ++i) {
    B[i] = 0.0;
#pragma unroll
    for (int j = 0; j < D1; ++j) A[i][j] = 0.0;
  }
  const double h = sqrt(nbd[k - 1]);
  for (int r = 0; r < k; ++r) {
    const double dist = sqrt(nbd[r]);
    const int j = nbi[r];
    const double tt = h > 0.0 ? dist / h : 0.0;
    const double w = exp(-tt * tt / 2) / sqrt(2.0 * 3.14159265358979323846);
    double v[D1];
    v[0] = 1.0;
#pragma unroll
    for (int dd = 0; dd < D; ++dd) v[dd + 1] = u[(size_t)j * D + dd] - q[dd];
    const double yr = y[j];
#pragma unroll
    for (int i = 0; i < D1; ++i) {
      B[i] += (i == 0 ? yr : yr * v[i]) * w;
#pragma unroll
      for (int jj = 0; jj <= i; ++jj) A[i][jj] += (i == 0 && jj == 0 ? 1.0 : v[i] * v[jj]) * w;
    }
  }
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
  return ok ? z[0] : B[0] / A[0][0];
}

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
  //    owning the minimum (indices are distinct) advances and lane 0 records
  //    it, so nbd / nbi end up in (squared distance, index) order.
  double* nbd = knn_lds + (size_t)cap * kKnnThreads + (size_t)cap * kKnnThreads / 2;  // after topi
  int* nbi = reinterpret_cast<int*>(nbd + k);
  int head = 0;
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
    if (lane == 0) {
      nbd[r] = md;
      nbi[r] = mi;
    }
  }
  if (lane != 0) return;
  // 3. The fit, summed in neighbour order by lane 0.
  double v = knn_regress<D>(nbd, nbi, k, q, a.u, a.y);
  if (a.clamp && v > 0.0) v = 0.0;  // rovCorrection (SimulateGPU.cu:3055-3073)
  a.val[blockIdx.x] = v;
}

// Many samples per grid point: one 256-thread workgroup per point finds the
// k-th smallest (squared distance, index) key exactly by an 8-pass radix
// select over the bits of the non-negative FP64 squared distance
// (distances recomputed each pass, no per-sample storage), collects the
// strictly smaller keys and, in index order, the equal ones it needs, ranks
// the k neighbours into (distance, index) order, and fits as above.
template <int D>
__global__ __launch_bounds__(256) void knn_select_kernel(KnnFitArgs a) {
  __shared__ unsigned int hist[256];
  __shared__ double cd[kKnnMaxK], nbd[kKnnMaxK];
  __shared__ int ci[kKnnMaxK], nbi[kKnnMaxK];
  __shared__ unsigned long long s_prefix;
  __shared__ int s_kk, s_cnt;
  __shared__ int wcnt[4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int n = a.n, k = a.k, R = a.R;
  double q[D];
  {
    long long rem = blockIdx.x;
#pragma unroll
    for (int dd = D - 1; dd >= 0; --dd) {
      const int i = (int)(rem % R);
      rem /= R;
      q[dd] = a.qtab[dd * R + i];
    }
  }
  auto dist2 = [&](int j) -> double {
    double s = 0.0;
#pragma unroll
    for (int dd = 0; dd < D; ++dd) {
      const double e = a.u[(size_t)j * D + dd] - q[dd];
      s += e * e;
    }
    return s;
  };
  // Radix select: after the passes, prefix is the k-th smallest squared
  // distance v and kk the number of samples at exactly v to take.
  unsigned long long prefix = 0ull;
  int kk = k;
  for (int pass = 0; pass < 8; ++pass) {
    const int shift = 56 - 8 * pass;
    const unsigned long long hmask = pass == 0 ? 0ull : (~0ull << (64 - 8 * pass));
    hist[tid] = 0u;
    __syncthreads();
    for (int j = tid; j < n; j += 256) {
      const unsigned long long key = (unsigned long long)__double_as_longlong(dist2(j));
      if ((key & hmask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      int cum = 0, b = 0;
      for (; b < 255; ++b) {
        if (cum + (int)hist[b] >= kk) break;
        cum += (int)hist[b];
      }
      s_prefix = prefix | ((unsigned long long)b << shift);
      s_kk = kk - cum;
    }
    __syncthreads();
    prefix = s_prefix;
    kk = s_kk;
    __syncthreads();
  }
  // Strictly smaller keys (any order; ranked below).
  if (tid == 0) s_cnt = 0;
  __syncthreads();
  for (int j = tid; j < n; j += 256) {
    const double d2 = dist2(j);
    if ((unsigned long long)__double_as_longlong(d2) < prefix) {
      const int pos = atomicAdd(&s_cnt, 1);
      cd[pos] = d2;
      ci[pos] = j;
    }
  }
  __syncthreads();
  const int m = s_cnt;  // == k - kk
  // Keys equal to v, lowest indices first: chunks of 256 consecutive samples.
  const double vd = __longlong_as_double((long long)prefix);
  int taken = 0;
  for (int base = 0; base < n && taken < kk; base += 256) {
    const int j = base + tid;
    const bool eq = j < n && dist2(j) == vd;
    const unsigned long long msk = __ballot(eq);
    if (lane == 0) wcnt[w] = __popcll(msk);
    __syncthreads();
    int off = 0;
    for (int ww = 0; ww < w; ++ww) off += wcnt[ww];
    const int rank = taken + off +
                     __builtin_amdgcn_mbcnt_hi((unsigned)(msk >> 32),
                                               __builtin_amdgcn_mbcnt_lo((unsigned)msk, 0));
    if (eq && rank < kk) {
      cd[m + rank] = vd;
      ci[m + rank] = j;
    }
    taken += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    __syncthreads();
  }
  // Rank into (distance, index) order.
  if (tid < k) {
    const double d = cd[tid];
    const int i = ci[tid];
    int r = 0;
    for (int j = 0; j < k; ++j) r += (cd[j] < d || (cd[j] == d && ci[j] < i)) ? 1 : 0;
    nbd[r] = d;
    nbi[r] = i;
  }
  __syncthreads();
  if (tid != 0) return;
  double v = knn_regress<D>(nbd, nbi, k, q, a.u, a.y);
  if (a.clamp && v > 0.0) v = 0.0;
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
  if (a.n > kKnnSelectMin) {
    hipLaunchKernelGGL(knn_select_kernel<D>, dim3((unsigned)a.G), dim3(256), 0, s, a);
    return hipGetLastError();
  }
  const size_t lds = knn_fit_lds(a.cap, a.k);
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
