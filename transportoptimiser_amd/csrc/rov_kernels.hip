// ROV kernels: exogenous price paths and regression Monte Carlo backward
// induction (FP64) for gfx950.
//
// Reference: forwardPathKernel (include/gpuCode/SimulateGPU.cu:789-956) for
// the price law and random controls, backwardInduction (:2333-2784) for the
// per-path decision, multiQuadReg/solveLinearSystem (:6092-6207, 6297-6358)
// for the regression, firstPeriodInduction (:2210-2330) for time 0.  There
// each backward step costs five launches, two serial <<<1,1>>> kernels and a
// host round trip with the regression solved on the CPU; here a step is three
// kernels (normal-equation assembly as per-block partial sums, a one-workgroup
// FP64 solve, the decision) that the host captures once into a HIP graph for
// the whole induction.
#include <hip/hip_runtime.h>

#include "rng.h"
#include "rov_kernels.h"

namespace roadmc {

enum { kSdeMrjd = 0, kSdeOu = 1, kSdeGbm = 2 };
enum { kRovSwitching = 0, kRovAmerican = 1 };
enum { kValueRegressed = 0, kValueRealized = 1 };

// Random stream of the ROV forward pass (kind 2 of rng.h):
//   philox({path, step, blk, tag}) with blk = 0,1 diffusion normals for
//   uncertainties 0-3, 4-7; blk = 2,3 jump-size normals; blk = 4,5 jump
//   uniforms (one per word); blk = 6 control uniform (word 0).
__device__ __forceinline__ void rov_normals(uint32_t p, uint32_t t, uint32_t blk, uint32_t tag,
                                            uint32_t k0, uint32_t k1, double z[4]) {
  rmc_normal4_d(p, t, blk, tag, k0, k1, z);
}

__global__ __launch_bounds__(256) void rov_forward_kernel(RovForwardArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.N) return;
  const int U = a.U;
  // Per-uncertainty arrays are indexed only by unrolled loop counters, so
  // they stay in registers.
  double x[kRovMaxU];
#pragma unroll
  for (int u = 0; u < kRovMaxU; ++u) {
    x[u] = 0.0;
    if (u < U) {
      x[u] = a.par[u * 7 + 1];
      a.states[(size_t)u * a.N + p] = x[u];
    }
  }
  const double dt = a.dt, sqdt = sqrt(a.dt);
  for (int t = 0; t < a.T; ++t) {
    double z[kRovMaxU], zj[kRovMaxU], uj[kRovMaxU];
#pragma unroll
    for (int b = 0; b < kRovMaxU / 4; ++b) {
      if (4 * b >= U) break;
      double w[4];
      rov_normals((uint32_t)p, (uint32_t)t, (uint32_t)b, a.tag, a.k0, a.k1, w);
#pragma unroll
      for (int i = 0; i < 4; ++i) z[4 * b + i] = w[i];
      rov_normals((uint32_t)p, (uint32_t)t, (uint32_t)(2 + b), a.tag, a.k0, a.k1, w);
#pragma unroll
      for (int i = 0; i < 4; ++i) zj[4 * b + i] = w[i];
      const rmc_u4 r = rmc_philox((uint32_t)p, (uint32_t)t, (uint32_t)(4 + b), a.tag, a.k0, a.k1);
      uj[4 * b + 0] = rmc_u01d(r.x);
      uj[4 * b + 1] = rmc_u01d(r.y);
      uj[4 * b + 2] = rmc_u01d(r.z);
      uj[4 * b + 3] = rmc_u01d(r.w);
    }
#pragma unroll
    for (int u = 0; u < kRovMaxU; ++u) {
      if (u >= U) break;
      // Correlated diffusion normal: row u of L z, L the Cholesky factor.
      double zc = 0.0;
#pragma unroll
      for (int v = 0; v <= u; ++v) zc += a.chol[u * kRovMaxU + v] * z[v];
      const double* q = a.par + u * 7;  // kind, x0, mean, sd, reversion, J, p
      const int kind = (int)q[0];
      const double xv = x[u];
      double nx;
      if (kind == kSdeOu) {
        nx = xv + q[4] * (q[2] - xv) * dt + q[3] * sqdt * zc;
      } else if (kind == kSdeGbm) {
        nx = xv * exp((q[2] - 0.5 * q[3] * q[3]) * dt + q[3] * sqdt * zc);
      } else {
        // SimulateGPU.cu:930-952: Brownian term x * sd * z, jump size
        // J' = J^2 z' - J^2 / 2 applied when the uniform is below p.
        const double jump = (uj[u] < q[6]) ? 1.0 : 0.0;
        const double jsz = zj[u] * q[5] * q[5] - q[5] * q[5] / 2.0;
        nx = xv + q[4] * (q[2] - xv) * dt + xv * q[3] * zc + (exp(jsz) - 1.0) * xv * jump;
      }
      x[u] = nx;
      a.states[((size_t)(t + 1) * U + u) * a.N + p] = nx;
    }
    if (a.controls) {
      // randControls (SimulateGPU.cu:720-730): floor(u C), clamped.
      const rmc_u4 r = rmc_philox((uint32_t)p, (uint32_t)t, 6u, a.tag, a.k0, a.k1);
      int c = (int)(rmc_u01d(r.x) * a.C);
      if (c >= a.C) c = a.C - 1;
      a.controls[(size_t)t * a.N + p] = c;
    }
  }
}

// Mean and 1/SD of every state dimension at every step (feature scaling):
// one 1024-thread block per (step, dimension), fixed-order tree reduction.
__global__ __launch_bounds__(1024) void rov_moments_kernel(const double* __restrict__ states,
                                                           int N, int U,
                                                           double* __restrict__ scale) {
  const int tu = blockIdx.x;  // t * U + u
  __shared__ double s1[1024], s2[1024];
  const double* x = states + (size_t)tu * N;
  double a = 0.0, b = 0.0;
  for (int p = threadIdx.x; p < N; p += blockDim.x) {
    const double v = x[p];
    a += v;
    b += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      s1[threadIdx.x] += s1[threadIdx.x + off];
      s2[threadIdx.x] += s2[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double m = s1[0] / N;
    const double v = s2[0] / N - m * m;
    const double sd = v > 1e-300 ? sqrt(v) : 1.0;
    scale[(size_t)tu * 2 + 0] = m;
    scale[(size_t)tu * 2 + 1] = (sd > 1e-12 * (fabs(m) + 1.0)) ? 1.0 / sd : 1.0;
  }
}

// All monomials of degree <= deg in U variables, in graded order:
// 1, x_i, x_i x_j (i <= j), x_i x_j x_k (i <= j <= k); f(k, value) is called
// for each in turn.  The loops run over the compile-time maximum with guards,
// so x stays in registers.
template <class F>
__device__ __forceinline__ int rov_basis_visit(const double (&x)[kRovMaxU], int U, int deg,
                                               F&& f) {
  int k = 0;
  f(k++, 1.0);
  if (deg >= 1) {
#pragma unroll
    for (int i = 0; i < kRovMaxU; ++i)
      if (i < U) f(k++, x[i]);
  }
  if (deg >= 2) {
#pragma unroll
    for (int i = 0; i < kRovMaxU; ++i)
#pragma unroll
      for (int j = i; j < kRovMaxU; ++j)
        if (j < U) f(k++, x[i] * x[j]);
  }
  if (deg >= 3) {
#pragma unroll
    for (int i = 0; i < kRovMaxU; ++i)
#pragma unroll
      for (int j = i; j < kRovMaxU; ++j)
#pragma unroll
        for (int l = j; l < kRovMaxU; ++l)
          if (l < U) f(k++, x[i] * x[j] * x[l]);
  }
  return k;
}

__device__ __forceinline__ void rov_features(const RovStepArgs& a, int t, int p,
                                             double (&x)[kRovMaxU]) {
#pragma unroll
  for (int u = 0; u < kRovMaxU; ++u) {
    double v = 0.0;
    if (u < a.U) {
      v = a.states[((size_t)t * a.U + u) * a.N + p];
      if (a.scale) {
        const double* sc = a.scale + ((size_t)t * a.U + u) * 2;
        v = (v - sc[0]) * sc[1];
      }
    }
    x[u] = v;
  }
}

__device__ __forceinline__ double rov_unit_profit(const RovStepArgs& a, int t, int p) {
  double up = a.unit_cost;
  for (int u = 0; u < a.U; ++u) up += a.weight[u] * a.states[((size_t)t * a.U + u) * a.N + p];
  return up;
}

__device__ __forceinline__ double rov_flow_scale(const RovStepArgs& a, int t, int p) {
  return a.traffic >= 0 ? a.states[((size_t)t * a.U + a.traffic) * a.N + p] : 1.0;
}

// Which fit a path of step t joins (-1: none).
__device__ __forceinline__ int rov_fit_group(const RovStepArgs& a, int t, int p) {
  if (a.mode == kRovAmerican) {
    const double e = a.strike - a.states[((size_t)t * a.U) * a.N + p];
    return (e > 0.0 || !a.itm_only) ? 0 : -1;
  }
  if (a.itm_only && rov_unit_profit(a, t, p) > 0.0) return -1;
  return a.pooled ? 0 : a.controls[(size_t)t * a.N + p];
}

// Per-block partial sums of the normal equations of step t for every fit
// group: A (upper triangle, row-major), b, and the point count.
//
// The block's 256 points form Phi = [phi_0 .. phi_{K-1}, y] (256 x KP, KP =
// K + 1 rounded up to 16, zero padded) in LDS; for group g the sums are the
// Gram matrix Phi_g^T Phi with Phi_g the rows of the group's points, built
// from 16 x 16 tiles on v_mfma_f64_16x16x4_f64 (4 points per MFMA, points in
// a fixed order).  Column K of the Gram matrix is b, entry (0, 0) the count.
constexpr int kRovAsmPoints = 256;

__host__ __device__ inline int rov_kp(int K) { return (K + 1 + 15) / 16 * 16; }

__device__ double rov_decide_path(const RovStepArgs& a, int t, int terminal, int p) {
  const double disc = a.disc;
  double* V = a.V;
  if (a.mode == kRovAmerican) {
    const double e = fmax(a.strike - a.states[((size_t)t * a.U) * a.N + p], 0.0);
    if (terminal) {
      V[(size_t)t * a.N + p] = e;
      a.optc[(size_t)(t - 1) * a.N + p] = e > 0.0 ? 1 : 0;
      return e;
    }
    const double later = disc * V[(size_t)(t + 1) * a.N + p];
    int ex = 0;
    if (e > 0.0 || !a.itm_only) {
      double x[kRovMaxU];
      rov_features(a, t, p, x);
      const double* b = a.beta + (size_t)t * a.G * a.K;
      double cont = 0.0;
      rov_basis_visit(x, a.U, a.deg, [&](int k, double v) { cont = fma(v, b[k], cont); });
      cont *= disc;
      ex = e > cont ? 1 : 0;
    }
    // Realized cash flows only: the host refuses ROADMC_VALUE_REGRESSED here.
    const double v = ex ? e : later;
    V[(size_t)t * a.N + p] = v;
    a.optc[(size_t)(t - 1) * a.N + p] = ex;
    return v;
  }
  // Flow control (backwardInduction, SimulateGPU.cu:2333-2784).
  const double up = rov_unit_profit(a, t, p);
  const double m = rov_flow_scale(a, t, p);
  // Continuation value of every fit group, one pass over the basis.
  double cont[kRovMaxC];
#pragma unroll
  for (int c = 0; c < kRovMaxC; ++c) cont[c] = 0.0;
  if (!terminal) {
    double x[kRovMaxU];
    rov_features(a, t, p, x);
    const double* b = a.beta + (size_t)t * a.G * a.K;
    rov_basis_visit(x, a.U, a.deg, [&](int k, double v) {
#pragma unroll
      for (int c = 0; c < kRovMaxC; ++c)
        if (c < a.G) cont[c] = fma(v, b[(size_t)c * a.K + k], cont[c]);
    });
  }
  double best = 0.0, best_cur = 0.0, best_pay = 0.0;
  int bc = 0;
#pragma unroll
  for (int c = 0; c < kRovMaxC; ++c) {
    if (c >= a.C) break;
    const double cur = a.flow[c] * m * up;
    double pay;
    if (terminal) {
      // Terminal step: payoff clamped at zero (SimulateGPU.cu:2385-2416).
      pay = fmin(cur, 0.0);
    } else {
      double cv = cont[a.pooled ? 0 : c] * disc;  // SimulateGPU.cu:2703
      if (cv > 0.0) cv = 0.0;                      // SimulateGPU.cu:2712-2714
      pay = cur + cv;
    }
    if (c == 0 || pay < a.hysteresis * best) {  // SimulateGPU.cu:2757-2771
      best = pay;
      bc = c;
      best_cur = cur;
      best_pay = pay;
    }
  }
  double v = best;
  if (a.value_mode == kValueRealized)
    v = terminal ? best_pay : best_cur + disc * V[(size_t)(t + 1) * a.N + p];
  V[(size_t)t * a.N + p] = v;
  a.optc[(size_t)t * a.N + p] = bc;
  return v;
}

__global__ __launch_bounds__(256) void rov_decide_kernel(const RovStepArgs* __restrict__ pa, int t,
                                                         int terminal) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < pa->N) rov_decide_path(*pa, t, terminal, p);
}

// With decide >= 0 the kernel first takes the decision of step t + 1 for its
// paths (rov_decide_path, terminal when decide == 2) and fits on the values
// it just wrote, so that step's separate decide launch is saved.
__global__ __launch_bounds__(256) void rov_assemble_kernel(const RovStepArgs* __restrict__ pa,
                                                           int t, int decide) {
  typedef double doublex4 __attribute__((ext_vector_type(4)));
  const RovStepArgs& a = *pa;
  extern __shared__ double sm[];
  const int K = a.K, E = a.E, KP = rov_kp(K), LD = KP + 1;  // +1: rotate banks per row
  double* phi = sm;  // [256][LD]
  int* sg = reinterpret_cast<int*>(phi + (size_t)kRovAsmPoints * LD);
  const int q = threadIdx.x;
  const int p = blockIdx.x * kRovAsmPoints + q;
  double* row = phi + (size_t)q * LD;
  int g = -1;
  for (int k = 0; k < KP; ++k) row[k] = 0.0;
  if (p < a.N) {
    const double vnext = decide >= 0 ? rov_decide_path(a, t + 1, decide == 2 ? 1 : 0, p)
                                     : a.V[(size_t)(t + 1) * a.N + p];
    g = rov_fit_group(a, t, p);
    double x[kRovMaxU];
    rov_features(a, t, p, x);
    rov_basis_visit(x, a.U, a.deg, [&](int k, double v) { row[k] = v; });
    row[K] = vnext;
  }
  sg[q] = g;
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nt = KP / 16, ntile = nt * (nt + 1) / 2, njobs = a.G * ntile;
  const int n_tri = K * (K + 1) / 2;
  const int kq = lane >> 4, li = lane & 15;
  for (int job = wave; job < njobs; job += 4) {
    const int grp = job / ntile;
    int I = 0, rem = job - grp * ntile;
    while (rem >= nt - I) {
      rem -= nt - I;
      ++I;
    }
    const int J = I + rem;
    doublex4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int q0 = 0; q0 < kRovAsmPoints; q0 += 4) {
      const double* r = phi + (size_t)(q0 + kq) * LD;
      const double av = sg[q0 + kq] == grp ? r[I * 16 + li] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, r[J * 16 + li], acc, 0, 0, 0);
    }
    // Item r of lane l is Gram entry (I*16 + l/16 + 4r, J*16 + l%16).
    const int j = J * 16 + li;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int i = I * 16 + kq + 4 * rr;
      int e = -1;
      if (i < K && j < K && i <= j)
        e = i * K - i * (i - 1) / 2 + (j - i);
      else if (i < K && j == K)
        e = n_tri + i;
      double* dst = a.partial + (size_t)grp * E * a.nblk + blockIdx.x;
      if (e >= 0) dst[(size_t)e * a.nblk] = acc[rr];
      if (i == 0 && j == 0) dst[(size_t)(n_tri + K) * a.nblk] = acc[rr];
    }
  }
}

// sums[e] = sum over blocks of partial[e][.], one wave per entry: lane l adds
// blocks l, l + 64, ... in order, then a fixed butterfly across lanes, so
// the totals do not depend on timing.
__global__ __launch_bounds__(64) void rov_reduce_kernel(const RovStepArgs* __restrict__ pa) {
  const RovStepArgs& a = *pa;
  const int e = blockIdx.x, lane = threadIdx.x;
  const double* src = a.partial + (size_t)e * a.nblk;
  double s = 0.0;
  for (int b = lane; b < a.nblk; b += 64) s += src[b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) a.sums[e] = s;
}

// Value of v in lane `src` (a compile-time constant in the unrolled loops
// that use it) broadcast to every lane.
__device__ __forceinline__ double bcast_lane(double v, int src) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, src);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), src);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Solves each group's K x K normal equations in FP64, one wave per group:
// Cholesky with the rows in registers (lane i owns row i; K is a template
// parameter, one instance per basis size the problem validation admits, so
// the loops are straight-line code), or, when the matrix is not numerically
// positive definite, Gaussian elimination with partial pivoting on
// |M[i][col]| (the pivot search the reference intends,
// SimulateGPU.cu:6307-6316).  Groups with fewer points than K keep zero
// coefficients (multiQuadReg skips them, SimulateGPU.cu:6101).
template <int K>
__global__ __launch_bounds__(64) void rov_solve_kernel(const RovStepArgs* __restrict__ pa, int t) {
  const RovStepArgs& a = *pa;
  extern __shared__ double sm[];
  const int E = a.E, grp = blockIdx.x, lane = threadIdx.x;
  double* M = sm;            // [K][K]
  double* L = M + K * K;     // [K][K]
  double* r = L + K * K;     // [K]
  double* y = r + K;         // [K]
  const double* S = a.sums + (size_t)grp * E;
  const int n_tri = K * (K + 1) / 2;
  for (int e = lane; e < n_tri; e += 64) {
    int i = 0, rem = e;
    while (rem >= K - i) {
      rem -= K - i;
      ++i;
    }
    const int j = i + rem;
    M[i * K + j] = S[e];
    M[j * K + i] = S[e];
  }
  for (int i = lane; i < K; i += 64) r[i] = S[n_tri + i];
  __syncthreads();
  double* coef = a.beta + ((size_t)t * a.G + grp) * K;
  const double count = S[n_tri + K];
  if (count < (double)K) {
    for (int i = lane; i < K; i += 64) coef[i] = 0.0;
    if (lane == 0) a.status[(size_t)t * a.G + grp] = 1;
    return;
  }
  double dmax = 0.0;
  for (int i = 0; i < K; ++i) dmax = fmax(dmax, fabs(M[i * K + i]));
  // Cholesky in registers: lane i holds row i of M (m) and of L (l).  Every
  // loop is unrolled over the compile-time bound, so the rows stay in
  // registers and each cross-lane value is a v_readlane broadcast from a
  // constant lane (no LDS round trip on the critical path).
  double m[K], l[K], invd[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    m[k] = lane < K ? M[lane * K + k] : 0.0;
    l[k] = 0.0;
    invd[k] = 0.0;
  }
  bool spd = true;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (spd) {
      const double djj = bcast_lane(m[j], j);
      if (djj > 1e-13 * dmax) {
        const double inv = 1.0 / sqrt(djj);
        invd[j] = inv;
        const double lij = m[j] * inv;  // L[lane][j] (lane j: L[j][j])
        l[j] = lij;
#pragma unroll
        for (int k = j + 1; k < K; ++k)
          m[k] = fma(-lij, bcast_lane(lij, k), m[k]);
      } else {
        spd = false;
      }
    }
  }
  if (spd) {
    // L y = r: y_j from lane j, removed from the lanes below (1/L_jj = invd).
    double rv = lane < K ? r[lane] : 0.0, yv = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const double yj = bcast_lane(rv, j) * invd[j];
      if (lane == j) yv = yj;
      if (lane > j) rv = fma(-l[j], yj, rv);
    }
    // L^T coef = y: coef_j from lane j, removed from the lanes above; lane i
    // needs L[j][i], a column of L: stage the rows of L in LDS.
    if (lane < K) {
#pragma unroll
      for (int k = 0; k < K; ++k) L[lane * K + k] = l[k];
    }
    __syncthreads();
    double cv = 0.0;
#pragma unroll
    for (int j = K - 1; j >= 0; --j) {
      const double cj = bcast_lane(yv, j) * invd[j];
      if (lane == j) cv = cj;
      if (lane < j) yv = fma(-L[j * K + lane], cj, yv);
    }
    if (lane < K) coef[lane] = cv;
    if (lane == 0) a.status[(size_t)t * a.G + grp] = 0;
    return;
  }
  if (lane != 0) return;
  // Rare fallback: pivoted elimination by one lane.
  int piv_ok = 1;
  for (int col = 0; col < K; ++col) {
    int best = col;
    double bv = fabs(M[col * K + col]);
    for (int i = col + 1; i < K; ++i)
      if (fabs(M[i * K + col]) > bv) {
        bv = fabs(M[i * K + col]);
        best = i;
      }
    if (!(bv > 1e-300)) {
      piv_ok = 0;
      break;
    }
    if (best != col) {
      for (int j = 0; j < K; ++j) {
        const double tmp = M[col * K + j];
        M[col * K + j] = M[best * K + j];
        M[best * K + j] = tmp;
      }
      const double tr = r[col];
      r[col] = r[best];
      r[best] = tr;
    }
    for (int i = col + 1; i < K; ++i) {
      const double f = M[i * K + col] / M[col * K + col];
      for (int j = col; j < K; ++j) M[i * K + j] -= f * M[col * K + j];
      r[i] -= f * r[col];
    }
  }
  if (piv_ok) {
    for (int i = K - 1; i >= 0; --i) {
      double s = r[i];
      for (int j = i + 1; j < K; ++j) s -= M[i * K + j] * y[j];
      y[i] = s / M[i * K + i];
    }
    for (int i = 0; i < K; ++i) coef[i] = y[i];
    a.status[(size_t)t * a.G + grp] = 2;
  } else {
    for (int i = 0; i < K; ++i) coef[i] = 0.0;
    a.status[(size_t)t * a.G + grp] = 3;
  }
}

// Decision of step t for every path.
// Time 0 (firstPeriodInduction, SimulateGPU.cu:2210-2330): per initial
// control, the mean discounted value of the paths that took it, plus the
// deterministic current payoff; out = {value, control, SD}.  One 1024-thread
// block, one pass over the paths, fixed-order tree reduction.
__global__ __launch_bounds__(1024) void rov_first_kernel(const RovStepArgs* __restrict__ pa) {
  const RovStepArgs& a = *pa;
  double* out = a.out;
  __shared__ double s1[kRovMaxC][1024 / 64], s2[kRovMaxC][1024 / 64], sn[kRovMaxC][1024 / 64];
  const int G = (a.mode == kRovAmerican || a.pooled) ? 1 : a.C;
  double x1[kRovMaxC], x2[kRovMaxC], n[kRovMaxC];
#pragma unroll
  for (int c = 0; c < kRovMaxC; ++c) x1[c] = x2[c] = n[c] = 0.0;
  for (int p = threadIdx.x; p < a.N; p += blockDim.x) {
    const int g = (G == 1) ? 0 : a.controls[p];
    const double v = a.disc * a.V[(size_t)a.N + p];
#pragma unroll
    for (int c = 0; c < kRovMaxC; ++c) {
      const bool in = g == c;
      x1[c] += in ? v : 0.0;
      x2[c] += in ? v * v : 0.0;
      n[c] += in ? 1.0 : 0.0;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kRovMaxC; ++c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      x1[c] += __shfl_xor(x1[c], off, 64);
      x2[c] += __shfl_xor(x2[c], off, 64);
      n[c] += __shfl_xor(n[c], off, 64);
    }
    if (lane == 0) {
      s1[c][wave] = x1[c];
      s2[c][wave] = x2[c];
      sn[c][wave] = n[c];
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int nw = blockDim.x / 64;
  double mean[kRovMaxC], sd[kRovMaxC], cnt[kRovMaxC];
  for (int c = 0; c < G; ++c) {
    double y1 = 0.0, y2 = 0.0, m = 0.0;
    for (int w = 0; w < nw; ++w) {
      y1 += s1[c][w];
      y2 += s2[c][w];
      m += sn[c][w];
    }
    cnt[c] = m;
    mean[c] = m > 0.0 ? y1 / m : 0.0;
    sd[c] = m > 0.0 ? sqrt(fmax(y2 / m - mean[c] * mean[c], 0.0)) : 0.0;
  }
  if (a.mode == kRovAmerican) {
    out[0] = mean[0];
    out[1] = 0.0;
    out[2] = sd[0];
    return;
  }
  // States at time 0 are the same on every path: use path 0.
  const double up = rov_unit_profit(a, 0, 0);
  const double m = rov_flow_scale(a, 0, 0);
  double best = 0.0;
  int bc = -1;
  for (int c = 0; c < a.C; ++c) {
    const int g = (G == 1) ? 0 : c;
    // A control no path took has no estimate; it is skipped rather than
    // ending the search (SimulateGPU.cu:2255-2260 stops there).
    if (cnt[g] <= 0.0) continue;
    double pay = mean[g] + a.flow[c] * m * up;
    if (pay > 0.0) pay = 0.0;
    if (bc < 0 || pay < best) {
      best = pay;
      bc = c;
    }
  }
  out[0] = best;
  out[1] = (double)(bc < 0 ? 0 : bc);
  out[2] = bc < 0 ? 0.0 : sd[(G == 1) ? 0 : bc];
}

// expPVPath (SimulateGPU.cu:256-279) with per-path noise indices: each path
// reads its own draws (the reference indexes [idx + ii], so paths overlap).
__global__ void exp_pv_kernel(ExpPvArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.N) return;
  double curr = a.x0, value = 0.0, grow = 1.0, disc = 1.0;
  for (int t = 0; t < a.T; ++t) {
    double z[4];
    rmc_normal4_d((uint32_t)p, (uint32_t)t, 0u, a.tag, a.k0, a.k1, z);
    const rmc_u4 r = rmc_philox((uint32_t)p, (uint32_t)t, 4u, a.tag, a.k0, a.k1);
    const double brown = z[0] * a.sd * a.dt;                       // N(0, sd dt)
    const double jsz = -a.jump * a.jump / 2.0 + z[1] * a.jump * a.jump;  // N(-J^2/2, J^2)
    const double jumped = rmc_u01d(r.x) < a.jump_prob ? 1.0 : 0.0;
    curr += a.reversion * (a.mean - curr) * a.dt + curr * brown + (exp(jsz) - 1.0) * curr * jumped;
    value += grow * curr / disc;
    grow *= 1.0 + a.growth;
    disc *= 1.0 + a.rrr;
  }
  a.out[p] = value;
}

hipError_t rov_set_lds_limits() {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&rov_assemble_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)rov_assemble_lds(kRovMaxK));
  if (e != hipSuccess) return e;
  return hipSuccess;
}

hipError_t launch_rov_forward(const RovForwardArgs& a, hipStream_t s) {
  if (a.N <= 0) return hipSuccess;
  hipLaunchKernelGGL(rov_forward_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_rov_moments(const double* states, int N, int U, int T1, double* scale,
                              hipStream_t s) {
  hipLaunchKernelGGL(rov_moments_kernel, dim3(T1 * U), dim3(1024), 0, s, states, N, U, scale);
  return hipGetLastError();
}

size_t rov_assemble_lds(int K) {
  return (size_t)kRovAsmPoints * (rov_kp(K) + 1) * sizeof(double) + kRovAsmPoints * sizeof(int);
}

hipError_t launch_rov_fit(const RovStepArgs& a, const RovStepArgs* pa, int t, int decide_next,
                          hipStream_t s) {
  {
    hipLaunchKernelGGL(rov_assemble_kernel, dim3(a.nblk), dim3(256), rov_assemble_lds(a.K), s,
                       pa, t, decide_next);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(rov_reduce_kernel, dim3(a.G * a.E), dim3(64), 0, s, pa);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t lds = (size_t)(2 * a.K * a.K + 2 * a.K) * sizeof(double);
    switch (a.K) {
#define ROV_SOLVE_CASE(KK)                                                                  \
  case KK:                                                                                  \
    hipLaunchKernelGGL(rov_solve_kernel<KK>, dim3(a.G), dim3(64), lds, s, pa, t);           \
    break;
      ROV_SOLVE_CASE(2) ROV_SOLVE_CASE(3) ROV_SOLVE_CASE(4) ROV_SOLVE_CASE(5) ROV_SOLVE_CASE(6)
      ROV_SOLVE_CASE(7) ROV_SOLVE_CASE(8) ROV_SOLVE_CASE(9) ROV_SOLVE_CASE(10) ROV_SOLVE_CASE(15)
      ROV_SOLVE_CASE(20) ROV_SOLVE_CASE(21) ROV_SOLVE_CASE(28) ROV_SOLVE_CASE(35)
#undef ROV_SOLVE_CASE
      default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
}

hipError_t launch_rov_decide(const RovStepArgs& a, const RovStepArgs* pa, int t, bool terminal,
                             hipStream_t s) {
  hipLaunchKernelGGL(rov_decide_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, pa, t,
                     terminal ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_rov_first(const RovStepArgs* pa, hipStream_t s) {
  hipLaunchKernelGGL(rov_first_kernel, dim3(1), dim3(1024), 0, s, pa);
  return hipGetLastError();
}

hipError_t launch_exp_pv(const ExpPvArgs& a, hipStream_t s) {
  if (a.N <= 0) return hipSuccess;
  hipLaunchKernelGGL(exp_pv_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace roadmc
