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
// spills).  For more samples, one workgroup per grid point selects the k
// nearest by radix passes over the squared distances' bits, finishing in LDS
// once the k-th key's bucket is small (knn_select_kernel).
// Compiled without multiply-add contraction, so that the neighbour ranking
// and the sums follow the same roundings as the CPU statement.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "knn_kernels.h"

namespace roadmc {

// Weighted local-linear fit (multiLocLinReg, SimulateGPU.cu:2934-3053) over
// the k neighbours nbd / nbi, sorted by (squared distance, index), each sum
// taken in that order (by one thread, or by one thread per sum in the select
// kernel, with the same roundings): basis (1, x - q), weights
// exp(-(d/h)^2 / 2) / sqrt(2 pi), h = the k-th neighbour's distance; FP64
// Cholesky, the weighted mean when the system is singular.
// The sums of the fit are B_i = sum_r t_B,i(r) and A_ij = sum_r t_A,ij(r)
// (j <= i), neighbour r's terms written by knn_terms in the order B_0 .. B_D,
// then A row by row; every sum runs over r in neighbour order.
template <int D>
constexpr int knn_nterms() {
  return (D + 1) + (D + 1) * (D + 2) / 2;
}

template <int D>
__device__ __forceinline__ void knn_terms(const double* nbd, const int* nbi, int r, double h,
                                          const double* q, const double* __restrict__ u,
                                          const double* __restrict__ y, double* t) {
  constexpr int D1 = D + 1;
  const double dist = sqrt(nbd[r]);
  const int j = nbi[r];
  const double tt = h > 0.0 ? dist / h : 0.0;
  const double w = exp(-tt * tt / 2) / sqrt(2.0 * 3.14159265358979323846);
  double v[D1];
  v[0] = 1.0;
#pragma unroll
  for (int dd = 0; dd < D; ++dd) v[dd + 1] = u[(size_t)j * D + dd] - q[dd];
  const double yr = y[j];
  int o = 0;
#pragma unroll
  for (int i = 0; i < D1; ++i) t[o++] = (i == 0 ? yr : yr * v[i]) * w;
#pragma unroll
  for (int i = 0; i < D1; ++i)
#pragma unroll
    for (int jj = 0; jj <= i; ++jj) t[o++] = (i == 0 && jj == 0 ? 1.0 : v[i] * v[jj]) * w;
}

template <int D>
__device__ double knn_solve(const double* sums);

template <int D>
__device__ double knn_regress(const double* nbd, const int* nbi, int k, const double* q,
                              const double* __restrict__ u, const double* __restrict__ y) {
  constexpr int NT = knn_nterms<D>();
  double sums[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) sums[i] = 0.0;
  const double h = sqrt(nbd[k - 1]);
  for (int r = 0; r < k; ++r) {
    double t[NT];
    knn_terms<D>(nbd, nbi, r, h, q, u, y, t);
#pragma unroll
    for (int i = 0; i < NT; ++i) sums[i] += t[i];
  }
  return knn_solve<D>(sums);
}

// Cholesky solve of the summed system; the weighted mean B_0 / A_00 when it
// is singular.
template <int D>
__device__ double knn_solve(const double* sums) {
  constexpr int D1 = D + 1;
  double A[D1][D1], B[D1];
  {
    int o = 0;
#pragma unroll
    for (int i = 0; i < D1; ++i) B[i] = sums[o++];
#pragma unroll
    for (int i = 0; i < D1; ++i)
#pragma unroll
      for (int jj = 0; jj <= i; ++jj) A[i][jj] = sums[o++];
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
__global__ __launch_bounds__(kKnnThreads) void knn_fit_kernel(KnnFitBatch batch) {
  const KnnFitArgs& a = batch.f[blockIdx.y];
  if ((long long)blockIdx.x >= a.G) return;
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
  // 3. The fit: for D <= 3 each lane writes the terms of neighbours lane and
  //    lane + 64 (after the lists, in the launch's extra LDS), then one lane
  //    per sum adds them in neighbour order; beyond, lane 0 does all.
  if constexpr (D <= 3) {
    constexpr int NTERM = knn_nterms<D>();
    double* fterm = nbd + knn_fit_term_offset(cap, k);  // [NTERM][k]
    double* fsum = fterm + (size_t)NTERM * k;
    __syncthreads();  // nbd / nbi written by lane 0
    const double h = sqrt(nbd[k - 1]);
    for (int r = lane; r < k; r += kKnnThreads) {
      double t[NTERM];
      knn_terms<D>(nbd, nbi, r, h, q, a.u, a.y, t);
#pragma unroll
      for (int i = 0; i < NTERM; ++i) fterm[(size_t)i * k + r] = t[i];
    }
    __syncthreads();
    if (lane < NTERM) {
      double sm = 0.0;
      for (int r = 0; r < k; ++r) sm += fterm[(size_t)lane * k + r];
      fsum[lane] = sm;
    }
    __syncthreads();
    if (lane != 0) return;
    double v = knn_solve<D>(fsum);
    if (a.clamp && v > 0.0) v = 0.0;  // rovCorrection (SimulateGPU.cu:3055-3073)
    a.val[blockIdx.x] = v;
  } else {
    if (lane != 0) return;
    double v = knn_regress<D>(nbd, nbi, k, q, a.u, a.y);
    if (a.clamp && v > 0.0) v = 0.0;
    a.val[blockIdx.x] = v;
  }
}

// Many samples per grid point: one 256-thread workgroup per point.  The keys
// are the bits of the non-negative FP64 squared distances, recomputed on
// every pass (no per-sample storage), ordered with the sample index as the
// tie-break.  Radix passes of 11 bits from the top, over all samples, narrow
// down the bucket of the k-th key until it holds at most kKnnSelectCap
// samples; one more pass gathers the keys below that bucket (fewer than k)
// and the bucket itself into LDS, where the bucket is bitonic-sorted by
// (key, index) and its first entries complete the k neighbours.  If every
// bit is resolved and the bucket is still larger (that many samples at one
// distance), the equal keys are taken in index order by chunked ballots.
// Either way the k neighbours are ranked into (distance, index) order and
// fitted as above.
template <int D>
__global__ __launch_bounds__(256) void knn_select_kernel(KnnFitBatch batch) {
  const KnnFitArgs& a = batch.f[blockIdx.y];
  if ((long long)blockIdx.x >= a.G) return;
  constexpr int kBins = 2048;
  __shared__ unsigned long long ekey[kKnnSelectCap];
  // The radix passes' histogram and its per-thread partial sums live in the
  // sort buffer, which holds nothing while they run (27.8 KB of LDS in all at
  // D = 2: five workgroups per CU).
  static_assert((kBins + 256) * sizeof(unsigned int) <= sizeof(ekey),
                "the histogram must fit the sort buffer");
  unsigned int* hist = reinterpret_cast<unsigned int*>(ekey);
  unsigned int* part = hist + kBins;
  __shared__ int eidx[kKnnSelectCap];
  __shared__ double cd[kKnnMaxK], nbd[kKnnMaxK];
  __shared__ int ci[kKnnMaxK], nbi[kKnnMaxK];
  __shared__ int s_b, s_cum, s_hb, s_cnt, s_ecnt;
  __shared__ int wcnt[4];
  // The fit's per-neighbour terms (in the sort buffer, free by then) and
  // sums: for D <= 3 one thread per neighbour writes its terms, then one
  // thread per sum adds them in neighbour order; beyond, one thread does all.
  constexpr bool kParFit = D <= 3;
  constexpr int NTERM = knn_nterms<D>();
  static_assert(!kParFit || NTERM * kKnnMaxK * sizeof(double) <=
                                kKnnSelectCap * sizeof(unsigned long long),
                "the fit's terms must fit the sort buffer");
  double(*fterm)[kKnnMaxK] = reinterpret_cast<double(*)[kKnnMaxK]>(ekey);
  __shared__ double fsum[NTERM];
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
  auto keyof = [&](int j) -> unsigned long long {
    return (unsigned long long)__double_as_longlong(dist2(j));
  };
  // fn(j, key) for the samples j = j0, j0 + stride, ... below n, four at a
  // time: the four loads are issued together (indices clamped to n - 1, so
  // none is conditional), then the keys below n are used.  The callers do not
  // depend on the order of the calls.
  auto for_keys = [&](const int j0, const int stride, auto&& fn) {
    for (int jb = j0; jb < n; jb += 4 * stride) {
      unsigned long long key[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) key[u] = keyof(min(jb + u * stride, n - 1));
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (jb + u * stride < n) fn(jb + u * stride, key[u]);
    }
  };
  // Radix passes over every step-th sample: prefix holds the resolved high
  // bits of the rank-th key among them, kk its rank among the keys that share
  // them, hb their number.  At least min_passes passes, then until the
  // bucket fits LDS.
  unsigned long long prefix = 0ull, hmask = 0ull;
  int kk = k, hb = n;
  auto radix = [&](const int step, const int min_passes, const int rank) {
    prefix = 0ull;
    hmask = 0ull;
    kk = rank;
    hb = n;
    for (int pass = 0; pass < 6; ++pass) {
      const int shift = pass < 5 ? 53 - 11 * pass : 0;
      const unsigned int dmask = pass < 5 ? kBins - 1 : 511u;
      for (int i = tid; i < kBins; i += 256) hist[i] = 0u;
      __syncthreads();
      for_keys(tid * step, 256 * step, [&](int, unsigned long long key) {
        if ((key & hmask) == prefix) atomicAdd(&hist[(unsigned)(key >> shift) & dmask], 1u);
      });
      __syncthreads();
      {
        unsigned int sum = 0u;
#pragma unroll
        for (int i = 0; i < kBins / 256; ++i) sum += hist[tid * (kBins / 256) + i];
        part[tid] = sum;
      }
      __syncthreads();
      if (w == 0) {
        // First bin whose cumulative count reaches kk: an inclusive scan over
        // the lanes' four partial sums, then within the lane's 32 bins.
        const int s4 = (int)(part[4 * lane] + part[4 * lane + 1] + part[4 * lane + 2] +
                             part[4 * lane + 3]);
        int inc = s4;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const int o = __shfl_up(inc, off);
          if (lane >= off) inc += o;
        }
        const unsigned long long reach = __ballot(inc >= kk);  // hb >= kk keys counted
        if (lane == __ffsll((long long)reach) - 1) {
          int cum = inc - s4, c = 4 * lane;
          while (cum + (int)part[c] < kk) cum += (int)part[c++];
          int b = c * (kBins / 256);
          while (cum + (int)hist[b] < kk) cum += (int)hist[b++];
          s_b = b;
          s_cum = cum;
          s_hb = (int)hist[b];
        }
      }
      __syncthreads();
      prefix |= (unsigned long long)s_b << shift;
      hmask |= (unsigned long long)dmask << shift;
      kk -= s_cum;
      hb = s_hb;
      __syncthreads();
      if (pass + 1 >= min_passes && hb <= kKnnSelectCap) break;
    }
  };
  // Bitonic sort of the first h entries of ekey / eidx by (key, index).
  auto sort_bucket = [&](const int h) {
    int S2 = 1;
    while (S2 < h) S2 <<= 1;
    for (int i = h + tid; i < S2; i += 256) {
      ekey[i] = ~0ull;
      eidx[i] = 0x7fffffff;
    }
    __syncthreads();
    for (int size = 2; size <= S2; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int i = tid; i < S2 / 2; i += 256) {
          const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
          const unsigned long long kl = ekey[lo], kh = ekey[hi];
          const int il = eidx[lo], ih = eidx[hi];
          const bool gt = kl > kh || (kl == kh && il > ih);
          if (gt == ((lo & size) == 0)) {
            ekey[lo] = kh;
            ekey[hi] = kl;
            eidx[lo] = ih;
            eidx[hi] = il;
          }
        }
        __syncthreads();
      }
    }
  };
  // The fit over nbd / nbi (ready in LDS), by every thread of the workgroup.
  auto fit = [&]() {
    if constexpr (kParFit) {
      const double h = sqrt(nbd[k - 1]);
      if (tid < k) {
        double t[NTERM];
        knn_terms<D>(nbd, nbi, tid, h, q, a.u, a.y, t);
#pragma unroll
        for (int i = 0; i < NTERM; ++i) fterm[i][tid] = t[i];
      }
      __syncthreads();
      if (tid < NTERM) {
        double sm = 0.0;
        for (int r = 0; r < k; ++r) sm += fterm[tid][r];
        fsum[tid] = sm;
      }
      __syncthreads();
      if (tid != 0) return;
      double v = knn_solve<D>(fsum);
      if (a.clamp && v > 0.0) v = 0.0;  // rovCorrection (SimulateGPU.cu:3055-3073)
      a.val[blockIdx.x] = v;
    } else {
      if (tid != 0) return;
      double v = knn_regress<D>(nbd, nbi, k, q, a.u, a.y);
      if (a.clamp && v > 0.0) v = 0.0;
      a.val[blockIdx.x] = v;
    }
  };
  // 0. A bound from a subsample of every step-th sample: the top of the
  //    radix bucket of its ks-th smallest key.  When at least k samples have
  //    a key at or below it, they include the k nearest (with every tie at
  //    the bound): one pass gathers them and, when they fit LDS, sorting them
  //    by (key, index) gives the k neighbours in the order of a stable sort
  //    of all samples.  Otherwise the full radix select runs.  ks is k / step
  //    plus five standard deviations of the number of subsample members
  //    among the k nearest (binomial, p = 1 / step), so the bound holds at
  //    least k samples but in rare cases; about step x ks samples are
  //    gathered (a 256-entry sort at k = 114).  With ks = k every sample
  //    below the bound's rank in the subsample is kept and the sort is four
  //    times as long, which cost more than the subsample reads it saves.
  const int step = min(4, n / (4 * k));
  if (step >= 2) {
    const double pk = (double)k / step;
    const int ks = min(k, (int)ceil(pk + 5.0 * sqrt(pk * (1.0 - 1.0 / step)) + 1.0));
    radix(step, 2, ks);
    const unsigned long long bound = prefix | ~hmask;
    if (tid == 0) s_ecnt = 0;
    __syncthreads();
    for_keys(tid, 256, [&](int j, unsigned long long key) {
      if (key <= bound) {
        const int pos = atomicAdd(&s_ecnt, 1);
        if (pos < kKnnSelectCap) {
          ekey[pos] = key;
          eidx[pos] = j;
        }
      }
    });
    __syncthreads();
    const int cnt = s_ecnt;
    if (cnt >= k && cnt <= kKnnSelectCap) {
      sort_bucket(cnt);
      if (tid < k) {
        nbd[tid] = __longlong_as_double((long long)ekey[tid]);
        nbi[tid] = eidx[tid];
      }
      __syncthreads();
      fit();
      return;
    }
    __syncthreads();
  }
  // 1. The full radix select over all samples.
  radix(1, 1, k);
  // 2. Keys below the prefix (exactly k - kk of them) and the bucket.
  if (tid == 0) {
    s_cnt = 0;
    s_ecnt = 0;
  }
  __syncthreads();
  const bool in_lds = hb <= kKnnSelectCap;
  for (int j = tid; j < n; j += 256) {
    const unsigned long long key = keyof(j);
    if (key < prefix) {
      const int pos = atomicAdd(&s_cnt, 1);
      if (pos < kKnnMaxK) {
        cd[pos] = __longlong_as_double((long long)key);
        ci[pos] = j;
      }
    } else if (in_lds && (key & hmask) == prefix) {
      const int pos = atomicAdd(&s_ecnt, 1);
      if (pos < kKnnSelectCap) {
        ekey[pos] = key;
        eidx[pos] = j;
      }
    }
  }
  __syncthreads();
  const int nl = k - kk;  // == s_cnt
  if (in_lds) {
    sort_bucket(hb);
    if (tid < nl) {
      const double d = cd[tid];
      const int i = ci[tid];
      int r = 0;
      for (int j = 0; j < nl; ++j) r += (cd[j] < d || (cd[j] == d && ci[j] < i)) ? 1 : 0;
      nbd[r] = d;
      nbi[r] = i;
    }
    if (tid < kk) {
      nbd[nl + tid] = __longlong_as_double((long long)ekey[tid]);
      nbi[nl + tid] = eidx[tid];
    }
  } else {
    // Every bit resolved: hb samples at exactly the k-th key; the kk lowest
    // indices, in chunks of 256 consecutive samples.
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
        cd[nl + rank] = vd;
        ci[nl + rank] = j;
      }
      taken += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
      __syncthreads();
    }
    if (tid < k) {
      const double d = cd[tid];
      const int i = ci[tid];
      int r = 0;
      for (int j = 0; j < k; ++j) r += (cd[j] < d || (cd[j] == d && ci[j] < i)) ? 1 : 0;
      nbd[r] = d;
      nbi[r] = i;
    }
  }
  __syncthreads();
  fit();
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
static hipError_t fit_d(const KnnFitArgs* fits, int nfits, hipStream_t s) {
  KnnFitBatch sel{}, wave{};
  long long gsel = 0, gwave = 0;
  size_t lds = 0;
  for (int i = 0; i < nfits; ++i) {
    const KnnFitArgs& a = fits[i];
    if (a.n > kKnnSelectMin) {
      sel.f[sel.nfits++] = a;
      gsel = std::max(gsel, a.G);
    } else {
      wave.f[wave.nfits++] = a;
      gwave = std::max(gwave, a.G);
      lds = std::max(lds, knn_fit_lds(a.cap, a.k) +
                              (D <= 3 ? knn_fit_term_lds(a.k, knn_nterms<D>()) : 0));
    }
  }
  if (sel.nfits) {
    hipLaunchKernelGGL(knn_select_kernel<D>, dim3((unsigned)gsel, sel.nfits), dim3(256), 0, s,
                       sel);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (wave.nfits) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&knn_fit_kernel<D>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(knn_fit_kernel<D>, dim3((unsigned)gwave, wave.nfits), dim3(kKnnThreads),
                       lds, s, wave);
    return hipGetLastError();
  }
  return hipSuccess;
}

template <int D>
static hipError_t eval_d(const KnnEvalArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(knn_eval_kernel<D>, dim3((a.m + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_knn_fit_batch(const KnnFitArgs* fits, int nfits, hipStream_t s) {
  if (nfits < 1) return hipSuccess;
  if (nfits > kKnnMaxBatch) return hipErrorInvalidValue;
  const int d = fits[0].d;
  for (int i = 1; i < nfits; ++i)
    if (fits[i].d != d) return hipErrorInvalidValue;
  switch (d) {
    case 1: return fit_d<1>(fits, nfits, s);
    case 2: return fit_d<2>(fits, nfits, s);
    case 3: return fit_d<3>(fits, nfits, s);
    case 4: return fit_d<4>(fits, nfits, s);
    case 5: return fit_d<5>(fits, nfits, s);
    case 6: return fit_d<6>(fits, nfits, s);
    case 7: return fit_d<7>(fits, nfits, s);
    case 8: return fit_d<8>(fits, nfits, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_knn_fit(const KnnFitArgs& a, hipStream_t s) {
  return launch_knn_fit_batch(&a, 1, s);
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
