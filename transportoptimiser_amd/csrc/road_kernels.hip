// Road -> crossing counts, transition and survival matrices for gfx950.
//
// Reference: SpeciesRoadPatches::roadCrossings, computeTransitionProbabilities
// and computeSurvivalProbabilities (include/road/SpeciesRoadPatches.cpp:
// 296-451), whose crossing counts come from Utility::lineSegmentIntersect
// (include/utilities/Utility.cpp:3-92; the reference's GPU form is
// pathAdjacencyKernel + roadCrossingsKernel, SimulateGPU.cu:454-493, in FP32
// with one thread per (link, segment) and an HBM adjacency matrix).  Here one
// thread owns a patch pair, the road segments are staged in LDS, the test is
// FP64 with the CPU path's operation order and no fused multiply-adds, and
// counts, weights and exponentials are written straight into the caller's
// column-major layout.
#include <hip/hip_runtime.h>

#include "road_kernels.h"

namespace roadmc {

constexpr int kRoadSegChunk = 256;

// Utility::lineSegmentIntersect (Utility.cpp:53-77): link (x1,y1)-(x2,y2)
// against segment (x3,y3)-(x4,y4); intersect when both parameters lie in
// [0, 1].  Explicit round-to-nearest operations keep the CPU evaluation
// order (no contraction into FMAs), so boundary cases match the CPU path.
__device__ __forceinline__ int seg_hit(double x1, double y1, double x2, double y2, double x3,
                                       double y3, double x4, double y4) {
  const double x43 = __dsub_rn(x4, x3), y13 = __dsub_rn(y1, y3);
  const double y43 = __dsub_rn(y4, y3), x13 = __dsub_rn(x1, x3);
  const double x21 = __dsub_rn(x2, x1), y21 = __dsub_rn(y2, y1);
  const double na = __dsub_rn(__dmul_rn(x43, y13), __dmul_rn(y43, x13));
  const double nb = __dsub_rn(__dmul_rn(x21, y13), __dmul_rn(y21, x13));
  const double de = __dsub_rn(__dmul_rn(y43, x21), __dmul_rn(x43, y21));
  const double ua = __ddiv_rn(na, de), ub = __ddiv_rn(nb, de);
  return (ua >= 0.0 && ua <= 1.0 && ub >= 0.0 && ub <= 1.0) ? 1 : 0;
}

__device__ __forceinline__ double pair_dist(const RoadArgs& a, int i, int j) {
  const double dx = __dsub_rn(a.x[i], a.x[j]), dy = __dsub_rn(a.y[i], a.y[j]);
  return __dsqrt_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)));
}

// Crossing count of every pair i < j within the crossing cut-off; the
// symmetric matrix is written both ways.  For the reference's dropped link,
// the largest row-major index of an in-range pair is recorded.
__global__ __launch_bounds__(256) void road_cross_kernel(RoadArgs a) {
  __shared__ double4 seg[kRoadSegChunk];
  const int P = a.P;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(idx / P), j = (int)(idx - (long long)i * P);
  const bool mine = idx < (long long)P * P && i < j;
  bool in_range = false;
  double x1 = 0, y1 = 0, x2 = 0, y2 = 0;
  if (mine) {
    in_range = pair_dist(a, i, j) <= a.cut_cross;
    x1 = a.x[i];
    y1 = a.y[i];
    x2 = a.x[j];
    y2 = a.y[j];
  }
  int count = 0;
  for (int s0 = 0; s0 < a.S; s0 += kRoadSegChunk) {
    const int ns = min(kRoadSegChunk, a.S - s0);
    __syncthreads();
    for (int s = threadIdx.x; s < ns; s += blockDim.x) seg[s] = a.seg[s0 + s];
    __syncthreads();
    if (in_range)
      for (int s = 0; s < ns; ++s) {
        const double4 q = seg[s];
        count += seg_hit(x1, y1, x2, y2, q.x, q.y, q.z, q.w);
      }
  }
  if (mine) {
    a.cross[(size_t)j * P + i] = count;  // (i, j), column-major
    a.cross[(size_t)i * P + j] = count;  // (j, i)
    if (in_range && a.last) atomicMax(a.last, (int)idx);
  }
  if (idx < (long long)P * P && i == j) a.cross[(size_t)i * P + i] = 0;
}

// Zeroes the pair roadCrossings leaves out of its count (one thread).
__global__ void road_drop_kernel(RoadArgs a) {
  const int k = *a.last;
  if (k < 0) return;
  const int i = k / a.P, j = k - i * a.P;
  a.cross[(size_t)j * a.P + i] = 0;
  a.cross[(size_t)i * a.P + j] = 0;
}

// Transition row i (computeTransitionProbabilities, SpeciesRoadPatches.cpp:
// 401-425): weights of the patches within the transition cut-off, then the
// row normalised by its sum (fixed-order reduction).
__global__ __launch_bounds__(256) void road_trans_kernel(RoadArgs a) {
  __shared__ double red[256];
  const int i = blockIdx.x, P = a.P;
  double part = 0.0;
  for (int j = threadIdx.x; j < P; j += blockDim.x) {
    const double d = pair_dist(a, i, j);
    double w = 0.0;
    if (d <= a.cut_trans) {
      if (i == j)
        w = a.pw[j];
      else if (a.form == 0)
        w = __ddiv_rn(a.pw[j], exp(__dmul_rn(-a.lambda, d)));
      else
        w = __dmul_rn(a.pw[j], exp(__dmul_rn(-a.lambda, d)));
    }
    a.trans[(size_t)j * P + i] = w;
    part += w;
  }
  red[threadIdx.x] = part;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  const double sum = red[0];
  for (int j = threadIdx.x; j < P; j += blockDim.x)
    a.trans[(size_t)j * P + i] = __ddiv_rn(a.trans[(size_t)j * P + i], sum);
}

// survProbs[c] = exp(-crossings * flow_c * (width + length) / (speed * 3600))
// (SpeciesRoadPatches.cpp:446-450), in that order of operations.
__global__ void road_surv_kernel(RoadArgs a) {
  const size_t n = (size_t)a.P * a.P;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double cr = (double)a.cross[e];
  for (int c = 0; c < a.C; ++c) {
    double v = __dmul_rn(-cr, a.flow[c]);
    v = __dmul_rn(v, a.width_len);
    v = __ddiv_rn(v, a.speed_3600);
    a.surv[(size_t)c * n + e] = exp(v);
  }
}

hipError_t launch_road_build(const RoadArgs& a, bool drop_last, bool want_trans,
                             bool want_surv, hipStream_t s) {
  const long long pairs = (long long)a.P * a.P;
  const int blocks = (int)((pairs + 255) / 256);
  hipLaunchKernelGGL(road_cross_kernel, dim3(blocks), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (drop_last) {
    hipLaunchKernelGGL(road_drop_kernel, dim3(1), dim3(1), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (want_trans) {
    hipLaunchKernelGGL(road_trans_kernel, dim3(a.P), dim3(256), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (want_surv && a.C > 0) {
    hipLaunchKernelGGL(road_surv_kernel, dim3(blocks), dim3(256), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace roadmc
