// Road -> crossing counts, transition and survival matrices for gfx950.
//
// Reference: SpeciesRoadPatches::roadCrossings, computeTransitionProbabilities
// and computeSurvivalProbabilities (include/road/SpeciesRoadPatches.cpp:
// 296-451), whose crossing counts come from Utility::lineSegmentIntersect
// (include/utilities/Utility.cpp:3-91; the reference's GPU form is
// pathAdjacencyKernel + roadCrossingsKernel, SimulateGPU.cu:454-493, in FP32
// with one thread per (link, segment) and an HBM adjacency matrix).
//
// Here, in FP64 with the CPU path's operations and order (no contractions):
//   1. road_cross_kernel: a workgroup tests 4096 matrix entries against the
//      crossing cut-off (column-major, so the zero fill coalesces), lists the
//      in-range pairs i < j in LDS and runs one thread per listed pair
//      against the road staged in LDS, so lanes stay busy however sparse the
//      in-range pairs are.  The segment test decides the two parameter
//      ranges from the signs and magnitudes of the numerators and the
//      denominator, which is exact, and divides only when that cannot
//      decide (a parameter within an ulp of 0 or 1, or a denominator that is
//      zero, subnormal or huge);
//   2. road_drop_kernel: the reference's uncounted last pair, from the
//      per-workgroup maxima (no global atomics anywhere);
//   3. road_trans_sum_kernel / road_trans_write_kernel: 64 rows x 256
//      columns per workgroup, lanes on consecutive rows so that column-major
//      stores coalesce; row sums are fixed-order sums of per-chunk partials
//      and the weights are recomputed rather than re-read;
//   4. road_surv_kernel: survival from a per-workgroup LDS table of
//      exp(-k flow (width + length) / (speed 3600)) for small counts k,
//      evaluated by the same expression as the direct form.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mte_kernels.h"
#include "road_kernels.h"

// The segment test, the distances and the cut-off comparisons must round like
// the reference's separate multiplies and adds, so this file is compiled with
// -ffp-contract=off (the __d*_rn helpers are plain operators in HIP and would
// otherwise fuse).

namespace roadmc {

constexpr int kRoadSegChunk = 512;  // segments staged per LDS pass (16 KiB)
constexpr int kSurvTable = 64;       // tabulated crossing counts per control
constexpr int kSurvMaxC = 16;        // controls the table covers

__device__ __forceinline__ double pair_dist(const double* x, const double* y, int i, int j) {
  const double dx = __dsub_rn(x[j], x[i]), dy = __dsub_rn(y[j], y[i]);
  return __dsqrt_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)));
}

// Utility::lineSegmentIntersect (Utility.cpp:45-83) for the link from
// (x1, y1) with direction (x21, y21) and the segment from (x3, y3) with
// direction (x43, y43): 1 when RN(na / de) and RN(nb / de) both lie in
// [0, 1].  The differences are the reference's, rounded the same way.
//
// For |de| in [2^-900, 2^100] the verdict is read off s = n * sign(de)
// (an exact sign flip), so that q = s / |de|:
//  - 0 <= s <= |de| (s = -0 included): 0 <= q <= 1, inside;
//  - s <= -2^-900: q <= -2^-1000, which does not round to -0, outside;
//  - s > RN(|de| (1 + 2^-51)): q > 1 + 2^-53, which rounds above 1, outside;
//  - NaN or infinite s: outside, as the division would be.
// The remaining bands (-2^-900 < s < 0 and |de| < s <= RN(|de| (1 + 2^-51)))
// and every other |de| (0, subnormal, huge, non-finite) take the division.
__device__ __forceinline__ int seg_hit(double x1, double y1, double x21, double y21, double x3,
                                       double y3, double x43, double y43) {
  const double y13 = __dsub_rn(y1, y3), x13 = __dsub_rn(x1, x3);
  const double na = __dsub_rn(__dmul_rn(x43, y13), __dmul_rn(y43, x13));
  const double nb = __dsub_rn(__dmul_rn(x21, y13), __dmul_rn(y21, x13));
  const double de = __dsub_rn(__dmul_rn(y43, x21), __dmul_rn(x43, y21));
  const double ad = fabs(de);
  // Bitwise '&' keeps the tests branch-free.
  if ((ad >= 0x1p-900) & (ad <= 0x1p100)) {
    const double sa = de < 0.0 ? -na : na, sb = de < 0.0 ? -nb : nb;
    const double hi = __dmul_rn(ad, 1.0 + 0x1p-51);
    // Outside unless both lie in [-2^-900, hi] (the common, cheap exit).
    if (!((sa >= -0x1p-900) & (sa <= hi) & (sb >= -0x1p-900) & (sb <= hi))) return 0;
    if ((sa >= 0.0) & (sa <= ad) & (sb >= 0.0) & (sb <= ad)) return 1;
    // Otherwise a parameter lies in one of the bands next to 0 or 1.
  }
  const double ua = __ddiv_rn(na, de), ub = __ddiv_rn(nb, de);
  return (ua >= 0.0 && ua <= 1.0 && ub >= 0.0 && ub <= 1.0) ? 1 : 0;
}

// Workgroup b owns chunk c = gridDim.x - 1 - b: matrix entries e in
// [c * kRoadCand, (c + 1) * kRoadCand),
// e = j * P + i (row i, column j; column-major, so the zero fill coalesces).
// Entries on the diagonal or beyond the cut-off are zeroed; in-range pairs
// with i < j are listed in LDS and counted against the road (staged in LDS
// in chunks), and both (i, j) and (j, i) receive the count.  The workgroup's
// largest in-range key i * P + j goes to blockmax[c] for the dropped link.
__global__ __launch_bounds__(kRoadThreads) void road_cross_kernel(RoadArgs a) {
  __shared__ double4 seg[kRoadSegChunk];
  __shared__ int list[kRoadCand];
  __shared__ int n_list;
  __shared__ int red[kRoadThreads / 64];
  const int P = a.P, S = a.S;
  const long long n = (long long)P * P;
  // Chunks late in column-major order hold the most pairs i < j: dispatch
  // them first so that the light ones fill the tail.
  const int chunk = gridDim.x - 1 - blockIdx.x;
  const long long e0 = (long long)chunk * kRoadCand;
  if (threadIdx.x == 0) n_list = 0;
  __syncthreads();
  int kmax = -1;
  for (int k = 0; k < kRoadCand / kRoadThreads; ++k) {
    const long long e = e0 + k * kRoadThreads + threadIdx.x;
    if (e >= n) break;
    const int j = (int)(e / P), i = (int)(e - (long long)j * P);
    const bool in_range = i != j && pair_dist(a.x, a.y, i, j) <= a.cut_cross;
    if (!in_range) {
      a.cross[e] = 0;
    } else if (i < j) {
      const int key = i * P + j;
      list[atomicAdd(&n_list, 1)] = key;
      kmax = max(kmax, key);
    }
  }
  for (int off = 32; off > 0; off >>= 1) kmax = max(kmax, __shfl_xor(kmax, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = kmax;
  __syncthreads();
  const int nl = n_list;
  if (threadIdx.x == 0) {
    int m = red[0];
    for (int w = 1; w < kRoadThreads / 64; ++w) m = max(m, red[w]);
    a.blockmax[chunk] = m;
  }
  constexpr int kPer = kRoadCand / kRoadThreads;
  int count[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) count[r] = 0;
  for (int s0 = 0; s0 < S; s0 += kRoadSegChunk) {
    const int ns = min(kRoadSegChunk, S - s0);
    __syncthreads();
    for (int q = threadIdx.x; q < ns; q += kRoadThreads) {
      const double4 g = a.seg[s0 + q];  // (x3, y3, x4, y4) -> (x3, y3, x4 - x3, y4 - y3)
      seg[q] = make_double4(g.x, g.y, __dsub_rn(g.z, g.x), __dsub_rn(g.w, g.y));
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int t = r * kRoadThreads + threadIdx.x;
      if (t < nl) {
        const int key = list[t];
        const int i = key / P, j = key - i * P;
        const double x1 = a.x[i], y1 = a.y[i];
        const double x21 = __dsub_rn(a.x[j], x1), y21 = __dsub_rn(a.y[j], y1);
        int c = 0;
        for (int q = 0; q < ns; ++q) {
          const double4 g = seg[q];
          c += seg_hit(x1, y1, x21, y21, g.x, g.y, g.z, g.w);
        }
        count[r] += c;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int t = r * kRoadThreads + threadIdx.x;
    if (t < nl) {
      const int key = list[t];
      const int i = key / P, j = key - i * P;
      a.cross[(size_t)j * P + i] = count[r];  // (i, j)
      a.cross[(size_t)i * P + j] = count[r];  // (j, i)
    }
  }
}

// roadCrossings leaves its last in-range pair (largest i * P + j) out of the
// count (SpeciesRoadPatches.cpp:369-374): one workgroup finds it among the
// per-workgroup maxima and zeroes it.
__global__ __launch_bounds__(1024) void road_drop_kernel(RoadArgs a, int n_blocks) {
  __shared__ int red[1024 / 64];
  int m = -1;
  for (int b = threadIdx.x; b < n_blocks; b += 1024) m = max(m, a.blockmax[b]);
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < 1024 / 64; ++w) m = max(m, red[w]);
  m = max(m, red[0]);
  if (m < 0) return;
  const int i = m / a.P, j = m - i * a.P;
  a.cross[(size_t)j * a.P + i] = 0;
  a.cross[(size_t)i * a.P + j] = 0;
}

// Transition weight of (i, j) before normalisation
// (SpeciesRoadPatches.cpp:403-419).
__device__ __forceinline__ double trans_weight(const RoadArgs& a, int i, int j) {
  const double d = pair_dist(a.x, a.y, i, j);
  if (!(d <= a.cut_trans)) return 0.0;
  if (i == j) return a.pw[j];
  const double e = exp(__dmul_rn(-a.lambda, d));
  return a.form == 0 ? __ddiv_rn(a.pw[j], e) : __dmul_rn(a.pw[j], e);
}

// Workgroup (rb, cb): rows rb*64 + lane, columns cb*256 + wave*64 .. +63.
// Each thread sums its 64 columns in order; the four waves' sums are added
// in wave order into part[cb][row].
__global__ __launch_bounds__(kRoadThreads) void road_trans_sum_kernel(RoadArgs a) {
  __shared__ double red[4][64];
  const int P = a.P;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = blockIdx.x * 64 + lane;
  const int j0 = blockIdx.y * kRoadCols + w * 64;
  double s = 0.0;
  if (i < P)
    for (int jj = 0; jj < 64; ++jj) {
      const int j = j0 + jj;
      if (j >= P) break;
      s = __dadd_rn(s, trans_weight(a, i, j));
    }
  red[w][lane] = s;
  __syncthreads();
  if (w == 0 && i < P)
    a.part[(size_t)blockIdx.y * P + i] =
        __dadd_rn(__dadd_rn(__dadd_rn(red[0][lane], red[1][lane]), red[2][lane]), red[3][lane]);
}

// Row sums from the partials (fixed order), weights recomputed and written
// normalised; a non-finite result raises tail->bad.
__global__ __launch_bounds__(kRoadThreads) void road_trans_write_kernel(RoadArgs a) {
  const int P = a.P;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = blockIdx.x * 64 + lane;
  if (i >= P) return;
  const int nchunk = (P + kRoadCols - 1) / kRoadCols;
  double sum = 0.0;
  for (int c = 0; c < nchunk; ++c) sum = __dadd_rn(sum, a.part[(size_t)c * P + i]);
  const int j0 = blockIdx.y * kRoadCols + w * 64;
  bool bad = false;
  for (int jj = 0; jj < 64; ++jj) {
    const int j = j0 + jj;
    if (j >= P) break;
    const double v = __ddiv_rn(trans_weight(a, i, j), sum);
    bad |= !__builtin_isfinite(v);
    a.trans[(size_t)j * P + i] = v;
  }
  if (bad) a.tail->bad = 1;
}

__device__ __forceinline__ double surv_value(double cr, double flow, const RoadArgs& a) {
  double v = __dmul_rn(-cr, flow);  // SpeciesRoadPatches.cpp:447-449, in that order
  v = __dmul_rn(v, a.width_len);
  v = __ddiv_rn(v, a.speed_3600);
  return exp(v);
}

// survProbs[c] = exp(-crossings * flow_c * (width + length) / (speed * 3600)).
__global__ __launch_bounds__(kRoadThreads) void road_surv_kernel(RoadArgs a) {
  __shared__ double table[kSurvMaxC][kSurvTable];
  const bool tab = a.C <= kSurvMaxC;
  if (tab) {
    for (int t = threadIdx.x; t < a.C * kSurvTable; t += kRoadThreads) {
      const int c = t / kSurvTable, k = t - c * kSurvTable;
      table[c][k] = surv_value((double)k, a.flow[c], a);
    }
    __syncthreads();
  }
  const size_t n = (size_t)a.P * a.P;
  for (size_t e = (size_t)blockIdx.x * kRoadThreads + threadIdx.x; e < n;
       e += (size_t)gridDim.x * kRoadThreads) {
    const int k = a.cross[e];
    for (int c = 0; c < a.C; ++c)
      a.surv[(size_t)c * n + e] =
          (tab && k < kSurvTable) ? table[c][k] : surv_value((double)k, a.flow[c], a);
  }
}

// Movement-and-mortality matrix in MFMA B order, as the host packing forms
// it (trans o surv, then dense2Sparse's drop rule on the binary32 value).
__global__ __launch_bounds__(256) void road_pack_kernel(const double* __restrict__ trans,
                                                       const double* __restrict__ surv, int P,
                                                       int Pp, int Pk, double tol, int fp64,
                                                       void* __restrict__ out) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)Pk * Pp) return;
  const int j = (int)(e / Pk), k = (int)(e - (size_t)j * Pk);
  double v = 0.0;
  if (k < P && j < P) {
    const size_t cm = (size_t)j * P + k;
    v = __dmul_rn(trans[cm], surv ? surv[cm] : 1.0);
    if (tol > 0.0 && !((float)v > (float)tol)) v = 0.0;
  }
  const size_t idx = mte_pack_index(fp64 != 0, Pk, k, j);
  if (fp64)
    static_cast<double*>(out)[idx] = v;
  else
    static_cast<float*>(out)[idx] = (float)v;
}

hipError_t launch_road_pack(bool fp64, const double* trans, const double* surv, int P, int Pp,
                            int Pk, double tol, void* out, hipStream_t s) {
  const size_t n = (size_t)Pk * Pp;
  hipLaunchKernelGGL(road_pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, trans,
                     surv, P, Pp, Pk, tol, fp64 ? 1 : 0, out);
  return hipGetLastError();
}

hipError_t launch_road_build(const RoadArgs& a, bool want_trans, bool want_surv, hipStream_t s) {
  const long long n = (long long)a.P * a.P;
  const int blocks = (int)((n + kRoadThreads - 1) / kRoadThreads);
  const int cblocks = road_cross_blocks(a.P);
  hipLaunchKernelGGL(road_cross_kernel, dim3(cblocks), dim3(kRoadThreads), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (a.drop_last) {
    hipLaunchKernelGGL(road_drop_kernel, dim3(1), dim3(1024), 0, s, a, cblocks);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (want_trans) {
    const dim3 g((a.P + 63) / 64, (a.P + kRoadCols - 1) / kRoadCols);
    hipLaunchKernelGGL(road_trans_sum_kernel, g, dim3(kRoadThreads), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(road_trans_write_kernel, g, dim3(kRoadThreads), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (want_surv && a.C > 0) {
    hipLaunchKernelGGL(road_surv_kernel, dim3(std::min(blocks, 4096)), dim3(kRoadThreads), 0, s,
                       a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace roadmc
