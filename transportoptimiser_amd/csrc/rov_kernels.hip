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

#include <utility>

#include "gmem.h"
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

// One step of uncertainty u's price law (kind uniform across the grid, so
// each law is a real branch rather than computed for every path and
// selected).
__device__ __forceinline__ double rov_price_step(const double* q, double xv, double zc, double zj,
                                                 double uj, double dt, double sqdt) {
  const int kind = (int)q[0];  // q: kind, x0, mean, sd, reversion, J, p
  if (kind == kSdeOu) return xv + q[4] * (q[2] - xv) * dt + q[3] * sqdt * zc;
  if (kind == kSdeGbm) return xv * exp((q[2] - 0.5 * q[3] * q[3]) * dt + q[3] * sqdt * zc);
  // SimulateGPU.cu:930-952: Brownian term x * sd * z, jump size
  // J' = J^2 z' - J^2 / 2 applied when the uniform is below p.
  const double jump = (uj < q[6]) ? 1.0 : 0.0;
  const double jsz = zj * q[5] * q[5] - q[5] * q[5] / 2.0;
  return xv + q[4] * (q[2] - xv) * dt + xv * q[3] * zc + (exp(jsz) - 1.0) * xv * jump;
}

// Price paths of U uncertainties, one path per thread.  The jump draws
// (blocks 2-5 of the stream) are generated only when some uncertainty follows
// the jump law; the stream is counter-based, so skipping them changes no other
// draw.
template <int U>
__global__ __launch_bounds__(256) void rov_forward_kernel(RovForwardArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.N) return;
  constexpr int NB = (U + 3) / 4;  // Philox blocks of four normals
  bool jumps = false;
#pragma unroll
  for (int u = 0; u < U; ++u) jumps = jumps || (int)a.par[u * 7] == kSdeMrjd;
  double x[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    x[u] = a.par[u * 7 + 1];
    a.states[(size_t)u * a.N + p] = x[u];
  }
  const double dt = a.dt, sqdt = sqrt(a.dt);
  for (int t = 0; t < a.T; ++t) {
    double z[4 * NB], zj[4 * NB], uj[4 * NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      double w[4];
      rov_normals((uint32_t)p, (uint32_t)t, (uint32_t)b, a.tag, a.k0, a.k1, w);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        z[4 * b + i] = w[i];
        zj[4 * b + i] = 0.0;
        uj[4 * b + i] = 1.0;
      }
    }
    if (jumps) {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        double w[4];
        rov_normals((uint32_t)p, (uint32_t)t, (uint32_t)(2 + b), a.tag, a.k0, a.k1, w);
        const rmc_u4 r = rmc_philox((uint32_t)p, (uint32_t)t, (uint32_t)(4 + b), a.tag, a.k0, a.k1);
        const uint32_t rw[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          zj[4 * b + i] = w[i];
          uj[4 * b + i] = rmc_u01d(rw[i]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      // Correlated diffusion normal: row u of L z, L the Cholesky factor.
      double zc = 0.0;
#pragma unroll
      for (int v = 0; v <= u; ++v) zc += a.chol[u * kRovMaxU + v] * z[v];
      x[u] = rov_price_step(a.par + u * 7, x[u], zc, zj[u], uj[u], dt, sqdt);
      a.states[((size_t)(t + 1) * U + u) * a.N + p] = x[u];
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

// Mean and 1/SD of every state dimension at every step (feature scaling),
// in two launches.  rov_moments_kernel: block (tu, c), tu = t U + u, sums
// chunk c of the paths (each thread its paths in order, loads issued four at
// a time, then a fixed-shape tree) into part [tu][kRovMomentChunks][2];
// rov_moments_final_kernel adds the chunks in order.
constexpr int kRovMomentChunks = 8;

__global__ __launch_bounds__(256) void rov_moments_kernel(const double* __restrict__ states,
                                                          int N, double* __restrict__ part) {
  const int tu = blockIdx.x, c = blockIdx.y, tid = threadIdx.x;
  __shared__ double s1[256], s2[256];
  const int lo = (int)((long long)N * c / kRovMomentChunks);
  const int hi = (int)((long long)N * (c + 1) / kRovMomentChunks);
  const double* x = states + (size_t)tu * N;
  double a = 0.0, b = 0.0;
  for (int p0 = lo + tid; p0 < hi; p0 += 4 * 256) {
    double v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = p0 + i * 256 < hi ? x[p0 + i * 256] : 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a += v[i];
      b += v[i] * v[i];
    }
  }
  s1[tid] = a;
  s2[tid] = b;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off) {
      s1[tid] += s1[tid + off];
      s2[tid] += s2[tid + off];
    }
    __syncthreads();
  }
  if (tid == 0) {
    part[((size_t)tu * kRovMomentChunks + c) * 2] = s1[0];
    part[((size_t)tu * kRovMomentChunks + c) * 2 + 1] = s2[0];
  }
}

__global__ __launch_bounds__(64) void rov_moments_final_kernel(const double* __restrict__ part,
                                                               int N, int n_tu,
                                                               double* __restrict__ scale) {
  const int tu = blockIdx.x * 64 + threadIdx.x;
  if (tu >= n_tu) return;
  double a = 0.0, b = 0.0;
  for (int c = 0; c < kRovMomentChunks; ++c) {
    a += part[((size_t)tu * kRovMomentChunks + c) * 2];
    b += part[((size_t)tu * kRovMomentChunks + c) * 2 + 1];
  }
  const double m = a / N;
  const double v = b / N - m * m;
  const double sd = v > 1e-300 ? sqrt(v) : 1.0;
  scale[(size_t)tu * 2 + 0] = m;
  scale[(size_t)tu * 2 + 1] = (sd > 1e-12 * (fabs(m) + 1.0)) ? 1.0 / sd : 1.0;
}

__device__ __forceinline__ double rov_unit_profit(const RovStepArgs& a, int t, int p) {
  double up = a.unit_cost;
  for (int u = 0; u < a.U; ++u) up += a.weight[u] * gmem(a.states)[((size_t)t * a.U + u) * a.N + p];
  return up;
}

__device__ __forceinline__ double rov_flow_scale(const RovStepArgs& a, int t, int p) {
  return a.traffic >= 0 ? gmem(a.states)[((size_t)t * a.U + a.traffic) * a.N + p] : 1.0;
}

// The polynomial basis of the regression, fixed at compile time: all
// monomials of degree <= DEG in U variables in graded order 1, x_i,
// x_i x_j (i <= j), x_i x_j x_l (i <= j <= l).  Monomial k is
// (v[i0] v[i1]) v[i2] with v = (1, x_0, .., x_{U-1}), so pairs and triples
// round exactly as x_i x_j and (x_i x_j) x_l.
template <int U, int DEG>
struct RovBasis {
  static constexpr int count() {
    int k = 1;
    if (DEG >= 1) k += U;
    if (DEG >= 2) k += U * (U + 1) / 2;
    if (DEG >= 3) k += U * (U + 1) * (U + 2) / 6;
    return k;
  }
  static constexpr int K = count();
  static constexpr int E = K * (K + 1) / 2 + K + 1;  // triangle, b, count
  struct Table {
    int i0[K], i1[K], i2[K];
  };
  static constexpr Table make() {
    Table tb{};
    int k = 0;
    tb.i0[k] = tb.i1[k] = tb.i2[k] = 0;
    ++k;
    if (DEG >= 1)
      for (int i = 0; i < U; ++i) {
        tb.i0[k] = 1 + i;
        tb.i1[k] = tb.i2[k] = 0;
        ++k;
      }
    if (DEG >= 2)
      for (int i = 0; i < U; ++i)
        for (int j = i; j < U; ++j) {
          tb.i0[k] = 1 + i;
          tb.i1[k] = 1 + j;
          tb.i2[k] = 0;
          ++k;
        }
    if (DEG >= 3)
      for (int i = 0; i < U; ++i)
        for (int j = i; j < U; ++j)
          for (int l = j; l < U; ++l) {
            tb.i0[k] = 1 + i;
            tb.i1[k] = 1 + j;
            tb.i2[k] = 1 + l;
            ++k;
          }
    return tb;
  }
  static constexpr Table tab = make();
  // Columns in blocks of four: lane c (0..3) of a quad of lanes computes
  // columns 4 F + c, F < NF, of one path, whose values v = (1, x_0, ..) are
  // spread over the quad (lane c holds v[c + 4 r] in register r < NR).
  // idx[w][F][c] is factor w of column 4 F + c (0, i.e. v[0] = 1, past K).
  static constexpr int NF = (K + 3) / 4;
  static constexpr int NQ = NF * (NF + 1) / 2;  // block pairs I <= J of the Gram matrix
  static constexpr int NR = (U + 4) / 4;        // ceil((U + 1) / 4)
  struct Quad {
    int idx[3][NF][4];
  };
  static constexpr Quad make_quad() {
    Quad q{};
    for (int F = 0; F < NF; ++F)
      for (int c = 0; c < 4; ++c) {
        const int col = 4 * F + c;
        q.idx[0][F][c] = col < K ? tab.i0[col] : 0;
        q.idx[1][F][c] = col < K ? tab.i1[col] : 0;
        q.idx[2][F][c] = col < K ? tab.i2[col] : 0;
      }
    return q;
  }
  static constexpr Quad quad = make_quad();
};

// x of quad lane CTRL-selected (DPP quad_perm: lane c reads lane (CTRL >> 2c) & 3
// of its quad).  Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_quad(double x) {
  if constexpr (CTRL == 0xE4) {
    return x;
  } else {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xf, 0xf, false);
    const uint32_t hi =
        (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
  }
}

// The sum over the four lanes j + 4 b (b < 4) of x, within each row of 16
// lanes: x plus x of the lane 4 positions along, then the same with 8 (each
// lane ends with all four; the callers read lanes b = 0).
__device__ __forceinline__ double dpp_row_sum4(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, 0x124, 0xf, 0xf, false);
  uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), 0x124, 0xf, 0xf, false);
  const double y = x + __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
  const uint64_t v = __builtin_bit_cast(uint64_t, y);
  lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)v, 0x128, 0xf, 0xf, false);
  hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(v >> 32), 0x128, 0xf, 0xf, false);
  return y + __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Register R's share of factor m (per quad lane): lanes whose index lives in
// register R take it from the quad lane that holds it.
template <int R, int NR, int M0, int M1, int M2, int M3>
__device__ __forceinline__ void rov_quad_pick(const double (&vq)[NR], int c, double& r) {
  constexpr bool u0 = (M0 >> 2) == R, u1 = (M1 >> 2) == R, u2 = (M2 >> 2) == R,
                 u3 = (M3 >> 2) == R;
  if constexpr (R < NR && (u0 || u1 || u2 || u3)) {
    constexpr int ctrl = (u0 ? (M0 & 3) : 0) | ((u1 ? (M1 & 3) : 1) << 2) |
                         ((u2 ? (M2 & 3) : 2) << 4) | ((u3 ? (M3 & 3) : 3) << 6);
    const double x = dpp_quad<ctrl>(vq[R]);
    if constexpr (u0 && u1 && u2 && u3) {
      r = x;
    } else {
      constexpr int msk = (u0 ? 1 : 0) | (u1 ? 2 : 0) | (u2 ? 4 : 0) | (u3 ? 8 : 0);
      r = ((msk >> c) & 1) ? x : r;
    }
  }
}

// Factor W of column 4 F + c.
template <int U, int DEG, int F, int W>
__device__ __forceinline__ double rov_quad_factor(const double (&vq)[RovBasis<U, DEG>::NR], int c) {
  using B = RovBasis<U, DEG>;
  constexpr int m0 = B::quad.idx[W][F][0], m1 = B::quad.idx[W][F][1],
                m2 = B::quad.idx[W][F][2], m3 = B::quad.idx[W][F][3];
  if constexpr (m0 == 0 && m1 == 0 && m2 == 0 && m3 == 0) {
    return 1.0;
  } else {
    constexpr int NR = B::NR;
    double r = 0.0;
    rov_quad_pick<0, NR, m0, m1, m2, m3>(vq, c, r);
    rov_quad_pick<1, NR, m0, m1, m2, m3>(vq, c, r);
    rov_quad_pick<2, NR, m0, m1, m2, m3>(vq, c, r);
    return r;
  }
}

// Column 4 F + c of the quad's path: (v[i0] v[i1]) v[i2] (zero past K).
template <int U, int DEG, int F>
__device__ __forceinline__ double rov_quad_column(const double (&vq)[RovBasis<U, DEG>::NR], int c) {
  using B = RovBasis<U, DEG>;
  const double f0 = rov_quad_factor<U, DEG, F, 0>(vq, c);
  const double f1 = rov_quad_factor<U, DEG, F, 1>(vq, c);
  const double f2 = rov_quad_factor<U, DEG, F, 2>(vq, c);
  double f = (f0 * f1) * f2;
  if constexpr (4 * F + 3 >= B::K) f = c < B::K - 4 * F ? f : 0.0;
  return f;
}

template <int U, int DEG, int... Fs>
__device__ __forceinline__ void rov_quad_columns_impl(const double (&vq)[RovBasis<U, DEG>::NR],
                                                      int c, double (&f)[RovBasis<U, DEG>::NF],
                                                      std::integer_sequence<int, Fs...>) {
  ((f[Fs] = rov_quad_column<U, DEG, Fs>(vq, c)), ...);
}

// f[F] = column 4 F + c of the path whose values the quad holds in vq.
template <int U, int DEG>
__device__ __forceinline__ void rov_quad_columns(const double (&vq)[RovBasis<U, DEG>::NR], int c,
                                                 double (&f)[RovBasis<U, DEG>::NF]) {
  rov_quad_columns_impl<U, DEG>(vq, c, f, std::make_integer_sequence<int, RovBasis<U, DEG>::NF>{});
}

// This lane's values of path row `row` of xs (value m of the row at
// xs[row rs + m cs], value 0 = 1) for rov_quad_columns: v[c + 4 r] (zero past U).
template <int U, int NR>
__device__ __forceinline__ void rov_quad_load(const double* xs, int row, int c, double (&vq)[NR],
                                              int rs, int cs) {
  constexpr int U1 = U + 1;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int m = c + 4 * r;
    vq[r] = m < U1 ? xs[(size_t)row * rs + (size_t)(m < U1 ? m : 0) * cs] : 0.0;
  }
}

// Step t's feature scaling (mean, 1/SD per uncertainty), loaded once.
template <int U>
__device__ __forceinline__ void rov_load_scale(const RovStepArgs& a, int t, double (&sc)[U][2]) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    sc[u][0] = a.scale ? gmem(a.scale)[((size_t)t * U + u) * 2] : 0.0;
    sc[u][1] = a.scale ? gmem(a.scale)[((size_t)t * U + u) * 2 + 1] : 1.0;
  }
}

// Raw states of step t of path p (all U loads issued together).
template <int U>
__device__ __forceinline__ void rov_load_states(const RovStepArgs& a, int t, int p,
                                                double (&x)[U]) {
#pragma unroll
  for (int u = 0; u < U; ++u) x[u] = gmem(a.states)[((size_t)t * U + u) * a.N + p];
}

// v = (1, features): the raw states standardised with step t's moments
// (sc from rov_load_scale; mean 0 and 1/SD 1 without standardisation).
template <int U>
__device__ __forceinline__ void rov_features(const double (&sc)[U][2], const double (&xr)[U],
                                             double (&v)[U + 1]) {
  v[0] = 1.0;
#pragma unroll
  for (int u = 0; u < U; ++u) v[1 + u] = (xr[u] - sc[u][0]) * sc[u][1];
}

template <int U>
__device__ __forceinline__ double rov_unit_profit_of(const RovStepArgs& a, const double (&xr)[U]) {
  double up = a.unit_cost;
#pragma unroll
  for (int u = 0; u < U; ++u) up += a.weight[u] * xr[u];
  return up;
}

// The fit a path of step t joins (-1: none): in the money (keepValidPoints,
// SimulateGPU.cu:5999-6043) and, per control, the control it drew.
template <int U>
__device__ __forceinline__ int rov_group_of(const RovStepArgs& a, const double (&xr)[U], int ctl) {
  if (a.mode == kRovAmerican) return (a.strike - xr[0] > 0.0 || !a.itm_only) ? 0 : -1;
  if (a.itm_only && rov_unit_profit_of<U>(a, xr) > 0.0) return -1;
  return a.pooled ? 0 : ctl;
}

// Whether the decision of step t reads V[t + 1] (realized values, American
// continuation).
__device__ __forceinline__ bool rov_decide_reads_later(const RovStepArgs& a, int terminal) {
  if (terminal) return false;
  return a.mode == kRovAmerican || a.value_mode == kValueRealized;
}

// The decision of step t for one path from its raw states xr at t, step t's
// feature scaling sc, the coefficients beta_t [G][K] of step t (unused at the
// terminal step) and,
// where rov_decide_reads_later, later = V[t + 1] of the path; returns the
// value and the control (American: 1 = exercise) in *bc.
template <int U, int DEG>
__device__ __forceinline__ double rov_decide(const RovStepArgs& a, int terminal,
                                             const double (&xr)[U], const double (&sc)[U][2],
                                             const double* beta_t, double later, int* bc_out) {
  using B = RovBasis<U, DEG>;
  constexpr int K = B::K;
  const double disc = a.disc;
  if (a.mode == kRovAmerican) {
    const double e = fmax(a.strike - xr[0], 0.0);
    if (terminal) {
      *bc_out = e > 0.0 ? 1 : 0;
      return e;
    }
    int ex = 0;
    if (e > 0.0 || !a.itm_only) {
      double v[U + 1];
      rov_features<U>(sc, xr, v);
      double cont = 0.0;
#pragma unroll
      for (int k = 0; k < K; ++k)
        cont = fma((v[B::tab.i0[k]] * v[B::tab.i1[k]]) * v[B::tab.i2[k]], beta_t[k], cont);
      cont *= disc;
      ex = e > cont ? 1 : 0;
    }
    // Realized cash flows only: the host refuses ROADMC_VALUE_REGRESSED here.
    *bc_out = ex;
    return ex ? e : disc * later;
  }
  // Flow control (backwardInduction, SimulateGPU.cu:2333-2784).
  const double up = rov_unit_profit_of<U>(a, xr);
  double m = 1.0;
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (u == a.traffic) m = xr[u];
  double phi[K];
  if (!terminal) {
    double v[U + 1];
    rov_features<U>(sc, xr, v);
#pragma unroll
    for (int k = 0; k < K; ++k) phi[k] = (v[B::tab.i0[k]] * v[B::tab.i1[k]]) * v[B::tab.i2[k]];
  }
  // One pass over the controls (a run-time loop: each control costs K
  // multiply-adds of its fit group's continuation, not kRovMaxC of them).
  double best = 0.0, best_cur = 0.0, best_pay = 0.0;
  int bc = 0;
  for (int c = 0; c < a.C; ++c) {
    const double cur = a.flow[c] * m * up;
    double pay;
    if (terminal) {
      // Terminal step: payoff clamped at zero (SimulateGPU.cu:2385-2416).
      pay = fmin(cur, 0.0);
    } else {
      const double* b = beta_t + (a.pooled ? 0 : c) * K;
      double cont = 0.0;
#pragma unroll
      for (int k = 0; k < K; ++k) cont = fma(phi[k], b[k], cont);
      double cv = cont * disc;  // SimulateGPU.cu:2703
      if (cv > 0.0) cv = 0.0;   // SimulateGPU.cu:2712-2714
      pay = cur + cv;
    }
    if (c == 0 || pay < a.hysteresis * best) {  // SimulateGPU.cu:2757-2771
      best = pay;
      bc = c;
      best_cur = cur;
      best_pay = pay;
    }
  }
  *bc_out = bc;
  if (a.value_mode == kValueRealized) return terminal ? best_pay : best_cur + disc * later;
  return best;
}

// Value of v in lane `src` (a compile-time constant in the unrolled loops
// that use it) broadcast to every lane.
__device__ __forceinline__ double bcast_lane(double v, int src) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, src);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), src);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Completes this wave's LDS accesses before its lanes read each other's
// values (the factorisation below runs on single waves).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Workgroup barrier ordering LDS only: unlike __syncthreads it does not wait
// for this thread's global stores, which no other thread of the block reads.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// s + sum_{m < n} src[m * stride], added in order m = 0, 1, ..., with the
// loads of each run of BATCH issued together (BATCH doubles of registers).
template <int BATCH>
__device__ __forceinline__ double rov_sum_rows(const RMC_GLOBAL double* src, size_t stride, int n,
                                               double s = 0.0) {
  for (int m0 = 0; m0 < n; m0 += BATCH) {
    double v[BATCH];
#pragma unroll
    for (int i = 0; i < BATCH; ++i) v[i] = m0 + i < n ? src[(size_t)(m0 + i) * stride] : 0.0;
#pragma unroll
    for (int i = 0; i < BATCH; ++i)
      if (m0 + i < n) s += v[i];
  }
  return s;
}

// ---- the regression of the backward induction ------------------------------
//
// Step t's fit (multiQuadReg, SimulateGPU.cu:6092-6207) solves A_{t,g} beta =
// b_{t,g} for each fit group g, A = sum phi phi^T over the group's paths and
// b = sum phi V_{t+1}.  Group membership (in the money, drawn control) and the
// basis depend only on the simulated states, so every A is formed and
// inverted before the induction starts, all steps at once (rov_gram_kernel,
// rov_invert_kernel); the induction's serial chain is then one launch per
// step (rov_step_kernel): the reduction of b of the step before, beta =
// A^{-1} b, the decisions, and the partial sums of the next b.

// LDS of rov_gram_kernel: values [U+1][kRovGramPaths + 1] (component-major,
// the odd row length spreading a quad's four components over distinct
// banks), the waves' sums [4][NQ][16]; then ints: groups and the group's path
// list, kRovGramPaths each, and 4 wave counts.
constexpr int kRovGramXs = kRovGramPaths + 1;

__host__ __device__ inline size_t rov_gram_lds_doubles(int K, int U) {
  const int nf = (K + 3) / 4, nq = nf * (nf + 1) / 2;
  return (size_t)kRovGramXs * (U + 1) + (size_t)4 * nq * 16;
}

size_t rov_gram_lds(int K, int U) {
  return rov_gram_lds_doubles(K, U) * sizeof(double) + (size_t)(2 * kRovGramPaths + 4) * sizeof(int);
}

// Entry e of a group's sums -> Gram coordinates (i, j), i <= j: the upper
// triangle row-major, then column K (b), then (0, 0) again for the count.
__device__ __forceinline__ void rov_entry_ij(int e, int K, int* i_out, int* j_out) {
  const int n_tri = K * (K + 1) / 2;
  if (e < n_tri) {
    int i = 0, rem = e;
    while (rem >= K - i) {
      rem -= K - i;
      ++i;
    }
    *i_out = i;
    *j_out = i + rem;
  } else if (e < n_tri + K) {
    *i_out = e - n_tri;
    *j_out = K;
  } else {
    *i_out = 0;
    *j_out = 0;
  }
}

// The matrix A of step t = t_lo + blockIdx.y over paths [kRovGramPaths
// blockIdx.x, +kRovGramPaths): block partial sums into partial [t - 1][blk]
// [G][E] (the b entries zero).  Per group, the block lists its paths of the
// group (in path order) and the waves take them sixteen at a time, round
// robin, on v_mfma_f64_4x4x4f64: lane l = c + 4 b + 16 k holds, for block b
// of the instruction, path k of that block's four paths and basis columns
// 4 F + c (rov_quad_columns), and the NQ products of column blocks I <= J
// accumulate sum phi_{4I+i} phi_{4J+j} in lane j + 4 b + 16 i.  The four
// instruction blocks (different paths) are summed across lanes, then the
// waves in wave order.
template <int U, int DEG>
__global__ __launch_bounds__(256) void rov_gram_kernel(const RovStepArgs* __restrict__ pa,
                                                       int t_lo) {
  using B = RovBasis<U, DEG>;
  constexpr int K = B::K, NF = B::NF, NQ = B::NQ, NR = B::NR, E = B::E, U1 = U + 1;
  constexpr int PB = kRovGramPaths, H = PB / 256;
  const RovStepArgs& a = *pa;
  const int G = a.G, GE = G * E, N = a.N;
  extern __shared__ double sm[];
  double* xs = sm;                       // [U1][kRovGramXs], row 0 = 1
  double* tiles = xs + kRovGramXs * U1;  // [4][NQ][16]
  int* gs = reinterpret_cast<int*>(sm + rov_gram_lds_doubles(K, U));  // [PB]
  int* lst = gs + PB;                    // [PB]
  int* wcnt = lst + PB;                  // [4]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, c = lane & 3;
  const int t = t_lo + (int)blockIdx.y;
  const int p0 = blockIdx.x * PB + tid;
  {
    double xr[H][U], sc[U][2];
    int ctl[H];
    rov_load_scale<U>(a, t, sc);
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const int p = p0 + 256 * h;
      ctl[h] = 0;
#pragma unroll
      for (int u = 0; u < U; ++u) xr[h][u] = 0.0;
      if (p < N) {
        rov_load_states<U>(a, t, p, xr[h]);
        ctl[h] = a.pooled ? 0 : gmem(a.controls)[(size_t)t * N + p];
      }
    }
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const int p = p0 + 256 * h, q = tid + 256 * h;
      double v[U1];
      rov_features<U>(sc, xr[h], v);
#pragma unroll
      for (int u = 0; u < U1; ++u) xs[(size_t)u * kRovGramXs + q] = p < N ? v[u] : 0.0;
      gs[q] = p < N ? rov_group_of<U>(a, xr[h], ctl[h]) : -1;
    }
  }
  // The tile slot of each entry this thread writes (entries e = tid + 256 m,
  // m < ceil(E / 256)); -1: a b entry, which stays zero.
  constexpr int ME = (E + 255) / 256;
  int slot[ME];
#pragma unroll
  for (int m = 0; m < ME; ++m) {
    const int e = tid + 256 * m;
    slot[m] = -1;
    if (e < E) {
      int i, jj;
      rov_entry_ij(e, K, &i, &jj);
      if (jj < K) {
        const int I = i >> 2, J = jj >> 2;
        const int qi = I * NF - I * (I - 1) / 2 + (J - I);
        slot[m] = qi * 16 + (i & 3) * 4 + (jj & 3);
      }
    }
  }
  lds_barrier();
  RMC_GLOBAL double* part = gmem(a.partial) + ((size_t)(t - 1) * gridDim.x + blockIdx.x) * GE;
  for (int g = 0; g < G; ++g) {
    // The block's paths of group g, in path order: wave w lists its rows
    // 256 h + 64 w + lane, h = 0 .. H - 1, after the lower waves' rows.
    uint64_t mk[H];
    int before[H], nw = 0;
#pragma unroll
    for (int h = 0; h < H; ++h) {
      mk[h] = __ballot(gs[256 * h + 64 * w + lane] == g);
      before[h] = nw;
      nw += __popcll(mk[h]);
    }
    if (lane == 0) wcnt[w] = nw;
    lds_barrier();
    int off = 0, total = 0;
#pragma unroll
    for (int ww = 0; ww < 4; ++ww) {
      off += ww < w ? wcnt[ww] : 0;
      total += wcnt[ww];
    }
#pragma unroll
    for (int h = 0; h < H; ++h)
      if ((mk[h] >> lane) & 1ull)
        lst[off + before[h] + __popcll(mk[h] & ((1ull << lane) - 1ull))] = 256 * h + 64 * w + lane;
    lds_barrier();
    double acc[NQ];
#pragma unroll
    for (int qi = 0; qi < NQ; ++qi) acc[qi] = 0.0;
    for (int r0 = 16 * w; r0 < total; r0 += 64) {
      const int idx = r0 + (lane >> 2);
      const bool ok = idx < total;
      double vq[NR], f[NF];
      rov_quad_load<U, NR>(xs, lst[ok ? idx : 0], c, vq, 1, kRovGramXs);
      rov_quad_columns<U, DEG>(vq, c, f);
#pragma unroll
      for (int F = 0; F < NF; ++F) f[F] = ok ? f[F] : 0.0;
      int qi = 0;
#pragma unroll
      for (int I = 0; I < NF; ++I)
#pragma unroll
        for (int J = I; J < NF; ++J) {
          acc[qi] = __builtin_amdgcn_mfma_f64_4x4x4f64(f[I], f[J], acc[qi], 0, 0, 0);
          ++qi;
        }
    }
    // Lane j + 16 i (b = 0) of each row receives the row's four blocks.
#pragma unroll
    for (int qi = 0; qi < NQ; ++qi) {
      const double s = dpp_row_sum4(acc[qi]);
      if ((lane & 12) == 0) tiles[((size_t)w * NQ + qi) * 16 + (lane >> 4) * 4 + (lane & 3)] = s;
    }
    lds_barrier();
#pragma unroll
    for (int m = 0; m < ME; ++m) {
      const int e = tid + 256 * m;
      if (e < E) {
        double s = 0.0;
        if (slot[m] >= 0) {
#pragma unroll
          for (int ww = 0; ww < 4; ++ww) s += tiles[(size_t)ww * NQ * 16 + slot[m]];
        }
        part[(size_t)g * E + e] = s;
      }
    }
    lds_barrier();
  }
}

// A^{-1} of one group from its sums S (upper triangle, b, count), on one
// wave: Cholesky with the rows in registers (lane i owns row i; K is a
// template parameter, one instance per basis size the problem validation
// admits, so the loops are straight-line code), then X = L^{-1} column by
// column (lane c solves L x = e_c) and A^{-1} = X^T X; when the matrix is not
// numerically positive definite, Gauss-Jordan elimination with partial
// pivoting on |M[i][col]| (the pivot search the reference intends,
// SimulateGPU.cu:6307-6316).  status: 0 Cholesky, 1 fewer points than K (A^{-1}
// = 0: the coefficients stay zero, as multiQuadReg skips such groups,
// SimulateGPU.cu:6101), 2 pivoted elimination, 3 singular (A^{-1} = 0).  work
// holds 2 K^2 + 2 K doubles of LDS.
template <int K>
__device__ void rov_invert_wave(const double* S, double* work, RMC_GLOBAL double* inv, RMC_GLOBAL int* status,
                                int lane) {
  double* M = work;       // [K][K]
  double* L = M + K * K;  // [K][K]
  constexpr int n_tri = K * (K + 1) / 2;
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
  wave_lds_sync();
  const double count = S[n_tri + K];
  if (count < (double)K) {
    for (int e = lane; e < K * K; e += 64) inv[e] = 0.0;
    if (lane == 0) *status = 1;
    return;
  }
  double dmax = 0.0;
  for (int i = 0; i < K; ++i) dmax = fmax(dmax, fabs(M[i * K + i]));
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
        const double iv = 1.0 / sqrt(djj);
        invd[j] = iv;
        const double lij = m[j] * iv;  // L[lane][j] (lane j: L[j][j])
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
    if (lane < K) {
#pragma unroll
      for (int k = 0; k < K; ++k) L[lane * K + k] = l[k];
    }
    wave_lds_sync();
    // Column `lane` of X = L^{-1} by forward substitution (zero above the
    // diagonal), into M (no longer needed) as X[i][lane].
    double x[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      double s = (i == lane) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) s = fma(-L[i * K + k], x[k], s);
      x[i] = s * invd[i];
    }
    wave_lds_sync();
    if (lane < K) {
#pragma unroll
      for (int i = 0; i < K; ++i) M[i * K + lane] = x[i];
    }
    wave_lds_sync();
    // A^{-1}[i][lane] = sum_k X[k][i] X[k][lane].
    if (lane < K) {
#pragma unroll
      for (int i = 0; i < K; ++i) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) s = fma(M[k * K + i], x[k], s);
        inv[i * K + lane] = s;
      }
    }
    if (lane == 0) *status = 0;
    return;
  }
  if (lane != 0) return;
  // Rare fallback, one lane: Gauss-Jordan on [M | I] with partial pivoting;
  // L holds the right-hand block.
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j) L[i * K + j] = i == j ? 1.0 : 0.0;
  int ok = 1;
  for (int col = 0; col < K && ok; ++col) {
    int best = col;
    double bv = fabs(M[col * K + col]);
    for (int i = col + 1; i < K; ++i)
      if (fabs(M[i * K + col]) > bv) {
        bv = fabs(M[i * K + col]);
        best = i;
      }
    if (!(bv > 1e-300)) {
      ok = 0;
      break;
    }
    if (best != col)
      for (int j = 0; j < K; ++j) {
        double tmp = M[col * K + j];
        M[col * K + j] = M[best * K + j];
        M[best * K + j] = tmp;
        tmp = L[col * K + j];
        L[col * K + j] = L[best * K + j];
        L[best * K + j] = tmp;
      }
    const double pv = M[col * K + col];
    for (int i = 0; i < K; ++i) {
      if (i == col) continue;
      const double f = M[i * K + col] / pv;
      for (int j = 0; j < K; ++j) {
        M[i * K + j] -= f * M[col * K + j];
        L[i * K + j] -= f * L[col * K + j];
      }
    }
  }
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j) inv[i * K + j] = ok ? L[i * K + j] / M[i * K + i] : 0.0;
  *status = ok ? 2 : 3;
}

// Block (step t = t_lo + blockIdx.x, group blockIdx.y): the fixed-order sum of
// the step's block partials (thread (e, h) adds its 1/H of the blocks in
// order, then the H pieces are added in order), then A^{-1} on wave 0 into
// ainv [step][G][K][K] and status [step][G].
template <int K>
__global__ __launch_bounds__(1024) void rov_invert_kernel(const RovStepArgs* __restrict__ pa,
                                                          int t_lo, int nblk_gram) {
  constexpr int E = K * (K + 1) / 2 + K + 1;  // triangle, b, count
  const RovStepArgs& a = *pa;
  const int G = a.G, GE = G * E, g = blockIdx.y, t = t_lo + (int)blockIdx.x;
  extern __shared__ double sm[];
  constexpr int H = 1024 / E > 0 ? (1024 / E < 8 ? 1024 / E : 8) : 1;
  double* piece = sm;                 // [H][E]
  double* S = piece + (size_t)H * E;  // [E]
  double* work = S + E;               // [WORK]
  const RMC_GLOBAL double* src = gmem(a.partial) + (size_t)(t - 1) * nblk_gram * GE + (size_t)g * E;
  for (int q = threadIdx.x; q < H * E; q += 1024) {
    const int h = q / E, e = q % E;
    const int lo = (int)((long long)nblk_gram * h / H), hi = (int)((long long)nblk_gram * (h + 1) / H);
    piece[(size_t)h * E + e] = rov_sum_rows<16>(src + (size_t)lo * GE + e, GE, hi - lo);
  }
  lds_barrier();
  for (int e = threadIdx.x; e < E; e += 1024) {
    double s = 0.0;
    for (int h = 0; h < H; ++h) s += piece[(size_t)h * E + e];
    S[e] = s;
  }
  lds_barrier();
  if (threadIdx.x < 64)
    rov_invert_wave<K>(S, work, gmem(a.ainv) + ((size_t)t * G + g) * K * K, gmem(a.status) + (size_t)t * G + g,
                       threadIdx.x);
}

// The inverse rows of step t that thread tid applies in rov_solve_finish
// (entry q = tid of [G][K]; threads past G K hold zeros), and the groups'
// solve status; loaded at kernel start so that their latency overlaps.
template <int K>
__device__ __forceinline__ void rov_load_inverse_row(const RovStepArgs& a, int t, double (&row)[K],
                                                     int* st) {
  const int q = threadIdx.x, GK = a.G * K;
  *st = 1;
#pragma unroll
  for (int k = 0; k < K; ++k) row[k] = 0.0;
  if (q < GK) {
    const int g = q / K, i = q % K;
    *st = gmem(a.status)[(size_t)t * a.G + g];
    const RMC_GLOBAL double* r = gmem(a.ainv) + (((size_t)t * a.G + g) * K + i) * K;
#pragma unroll
    for (int k = 0; k < K; ++k) row[k] = r[k];
  }
}

// The right-hand sides b of a step (sum over the step's paths of group g of
// phi V[t+1]) from the step kernels' block partials bpart [nblk][G][K], in a
// fixed order: thread (e, h) adds its 1/H of the blocks in block order into
// seg (entries past the first 256 go whole into bsum); rov_solve_finish,
// after a barrier, adds the H pieces in order.  The first kRovPrefetch rows
// of the thread's share are loaded by rov_partials_issue at kernel start and
// added by rov_partials_sum, so that the kernel's other loads are issued
// while they are in flight.
constexpr int kRovPrefetch = 40;

struct RovPartials {
  double v[kRovPrefetch];
  const RMC_GLOBAL double* src;  // the thread's first row, or null
  int rows;
};

template <int K>
__device__ __forceinline__ void rov_partials_issue(int G, const RMC_GLOBAL double* bpart, int nblk,
                                                   RovPartials& pf) {
  const int GK = G * K, tid = threadIdx.x;
  const int H = GK <= 256 ? 256 / GK : 1;
  pf.src = nullptr;
  pf.rows = 0;
  if (tid < GK * H) {
    const int e = tid % GK, h = tid / GK;
    const int lo = (int)((long long)nblk * h / H), hi = (int)((long long)nblk * (h + 1) / H);
    pf.src = bpart + (size_t)lo * GK + e;
    pf.rows = hi - lo;
  }
#pragma unroll
  for (int i = 0; i < kRovPrefetch; ++i) pf.v[i] = i < pf.rows ? pf.src[(size_t)i * GK] : 0.0;
}

template <int K>
__device__ __forceinline__ void rov_partials_sum(int G, const RMC_GLOBAL double* bpart, int nblk,
                                                 const RovPartials& pf, double* seg, double* bsum) {
  const int GK = G * K, tid = threadIdx.x;
  if (pf.src) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kRovPrefetch; ++i)
      if (i < pf.rows) s += pf.v[i];
    if (pf.rows > kRovPrefetch)
      s = rov_sum_rows<16>(pf.src + (size_t)kRovPrefetch * GK, GK, pf.rows - kRovPrefetch, s);
    seg[tid] = s;
  }
  for (int e = 256 + tid; e < GK; e += 256) bsum[e] = rov_sum_rows<16>(bpart + e, GK, nblk);
}

// b from rov_partials_sum's pieces (seg and bsum visible to the block), then
// beta_{t,g} = A^{-1}_{t,g} b_{t,g} into beta_l [G][K] (LDS); block 0 also
// stores it in beta [t].  row and st come from rov_load_inverse_row.
template <int K>
__device__ void rov_solve_finish(const RovStepArgs& a, int t, const double* seg, double* bsum,
                                 double* beta_l, const double (&row)[K], int st) {
  const int GK = a.G * K, tid = threadIdx.x;
  const int H = GK <= 256 ? 256 / GK : 1;
  if (tid < GK && tid < 256) {
    double s = 0.0;
    for (int h = 0; h < H; ++h) s += seg[h * GK + tid];
    bsum[tid] = s;
  }
  lds_barrier();
  for (int q = tid; q < GK; q += 256) {
    const int g = q / K;
    double s = 0.0;
    if (q == tid) {
      if (st == 0 || st == 2) {
#pragma unroll
        for (int k = 0; k < K; ++k) s = fma(row[k], bsum[g * K + k], s);
      }
    } else {  // G K > 256: rows past the first 256 read here
      const int i = q % K;
      const int sq = gmem(a.status)[(size_t)t * a.G + g];
      if (sq == 0 || sq == 2) {
        const RMC_GLOBAL double* r = gmem(a.ainv) + (((size_t)t * a.G + g) * K + i) * K;
        for (int k = 0; k < K; ++k) s = fma(r[k], bsum[g * K + k], s);
      }
    }
    beta_l[q] = s;
    if (blockIdx.x == 0) gmem(a.beta)[(size_t)t * GK + q] = s;
  }
  lds_barrier();
}

// One step of the backward induction (t >= 1), paths [512 b, 512 b + 512) on
// block b, two per thread.  With decide == 1: beta_{t+1} from the previous
// launch's partials and the decisions of step t + 1; with decide == 2: the
// terminal decisions of step t + 1.  Then the partial right-hand sides of
// step t for every group at once on v_mfma_f64_16x16x4_f64: operand A holds
// basis columns 16 I + (l & 15) of path l >> 4 of four paths, operand B the
// path's value V[t+1] in the column of its group (zero elsewhere), so column g
// of tile I accumulates sum_{paths of g} phi_{16I+i} V[t+1]; written to
// bpart [t & 1][b][G][K].
template <int U, int DEG>
__global__ __launch_bounds__(256) void rov_step_kernel(const RovStepArgs* __restrict__ pa, int t,
                                                       int decide) {
  using B = RovBasis<U, DEG>;
  constexpr int K = B::K, NF = B::NF, NR = B::NR, U1 = U + 1;
  // The parameter block is copied to LDS (one word per thread, stored after
  // the kernel's other loads are issued), so that the decisions read it
  // there instead of stalling on scalar-cache misses.
  __shared__ RovStepArgs sa;
  constexpr int kWords = (int)(sizeof(RovStepArgs) / 4);
  static_assert(sizeof(RovStepArgs) % 4 == 0 && kWords <= 256, "one word per thread");
  const uint32_t pword =
      threadIdx.x < kWords ? gmem(reinterpret_cast<const uint32_t*>(pa))[threadIdx.x] : 0u;
  const RovStepArgs& ga = *pa;
  const int G = ga.G, GK = G * K, N = ga.N;
  __shared__ double xs[kRovStepPaths * U1];
  __shared__ double ys[kRovStepPaths];
  __shared__ double seg[256];
  __shared__ double bsum[kRovMaxC * K];
  __shared__ double beta_l[kRovMaxC * K];
  __shared__ double tiles[4][2][NF][16];
  __shared__ int gs[kRovStepPaths];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int term = decide == 2 ? 1 : 0;
  // The previous launch's partials (the critical path), then the inverse
  // rows, then both paths' states.
  const RMC_GLOBAL double* bprev = gmem(ga.bpart) + (size_t)((t + 1) & 1) * ga.nblk * GK;
  RovPartials pf;
  pf.src = nullptr;
  if (!term) rov_partials_issue<K>(G, bprev, ga.nblk, pf);
  double arow[K];
  int ast = 1;
  if (!term) rov_load_inverse_row<K>(ga, t + 1, arow, &ast);
  const int p0 = blockIdx.x * kRovStepPaths + tid;
  double xn[2][U], xc[2][U], later[2] = {0.0, 0.0}, sc_n[U][2], sc_c[U][2];
  int ctl[2] = {0, 0};
  const bool need_later = rov_decide_reads_later(ga, term);
  rov_load_scale<U>(ga, t + 1, sc_n);
  rov_load_scale<U>(ga, t, sc_c);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int p = p0 + h * 256;
    if (p < N) {
      rov_load_states<U>(ga, t + 1, p, xn[h]);
      rov_load_states<U>(ga, t, p, xc[h]);
      ctl[h] = ga.pooled ? 0 : gmem(ga.controls)[(size_t)t * N + p];
      if (need_later) later[h] = gmem(ga.V)[(size_t)(t + 2) * N + p];
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) xn[h][u] = xc[h][u] = 0.0;
    }
  }
  if (tid < kWords) reinterpret_cast<uint32_t*>(&sa)[tid] = pword;
  if (!term) rov_partials_sum<K>(G, bprev, ga.nblk, pf, seg, bsum);
  lds_barrier();
  const RovStepArgs& a = sa;
  if (!term) rov_solve_finish<K>(a, t + 1, seg, bsum, beta_l, arow, ast);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int p = p0 + h * 256, q = tid + h * 256;
    int g = -1;
    double y = 0.0, v[U1];
    if (p < N) {
      int bc;
      y = rov_decide<U, DEG>(a, term, xn[h], sc_n, beta_l, later[h], &bc);
      gmem(a.V)[(size_t)(t + 1) * N + p] = y;
      gmem(a.optc)[(size_t)(a.mode == kRovAmerican ? t : t + 1) * N + p] = bc;
      g = rov_group_of<U>(a, xc[h], ctl[h]);
      rov_features<U>(sc_c, xc[h], v);
    } else {
#pragma unroll
      for (int u = 0; u < U1; ++u) v[u] = 0.0;
    }
#pragma unroll
    for (int u = 0; u < U1; ++u) xs[(size_t)q * U1 + u] = v[u];
    ys[q] = y;
    gs[q] = g;
  }
  lds_barrier();
  // Wave w takes rows 256 h + 64 w + 16 r + (lane >> 2) (h < 2, r < 4): on
  // v_mfma_f64_4x4x4f64 lane l = c + 4 b + 16 k holds basis columns 4 F + c
  // (operand A) and the row's value in the column of its group, 4 P + c
  // (operand B), of path k of block b's four; pass P = 1 only with more
  // than four groups.
  const int c = lane & 3;
  const bool two = G > 4;
  double acc[2][NF];
#pragma unroll
  for (int P = 0; P < 2; ++P)
#pragma unroll
    for (int F = 0; F < NF; ++F) acc[P][F] = 0.0;
#pragma unroll 2
  for (int r = 0; r < 8; ++r) {
    const int row = (r >> 2) * 256 + w * 64 + (r & 3) * 16 + (lane >> 2);
    double vq[NR], f[NF];
    rov_quad_load<U, NR>(xs, row, c, vq, U1, 1);
    rov_quad_columns<U, DEG>(vq, c, f);
    const int gr = gs[row];
    const double yr = ys[row];
    const double b0 = gr == c ? yr : 0.0;
#pragma unroll
    for (int F = 0; F < NF; ++F) acc[0][F] = __builtin_amdgcn_mfma_f64_4x4x4f64(f[F], b0, acc[0][F], 0, 0, 0);
    if (two) {
      const double b1 = gr == 4 + c ? yr : 0.0;
#pragma unroll
      for (int F = 0; F < NF; ++F)
        acc[1][F] = __builtin_amdgcn_mfma_f64_4x4x4f64(f[F], b1, acc[1][F], 0, 0, 0);
    }
  }
  // Lane j + 16 i of each row receives the sum of its row's four blocks:
  // column 4 F + i, group 4 P + j.
#pragma unroll
  for (int P = 0; P < 2; ++P) {
    if (P == 1 && !two) break;
#pragma unroll
    for (int F = 0; F < NF; ++F) {
      const double sum = dpp_row_sum4(acc[P][F]);
      if ((lane & 12) == 0) tiles[w][P][F][(lane >> 4) * 4 + c] = sum;
    }
  }
  lds_barrier();
  RMC_GLOBAL double* out = gmem(a.bpart) + (size_t)(t & 1) * a.nblk * GK + (size_t)blockIdx.x * GK;
  for (int q = tid; q < GK; q += 256) {
    const int g = q / K, k = q % K;
    const int P = g >> 2, F = k >> 2, l = (k & 3) * 4 + (g & 3);
    out[q] = ((tiles[0][P][F][l] + tiles[1][P][F][l]) + tiles[2][P][F][l]) + tiles[3][P][F][l];
  }
}

// beta_1 (from the last step's partials, unless step 1 is terminal), the
// decision of step 1 for every path, then time 0 (firstPeriodInduction,
// SimulateGPU.cu:2210-2330): per initial control, the mean discounted value of
// the paths that took it plus the deterministic current payoff;
// out = {value, control, SD}.  Each block sums its paths per control (count,
// sum, sum of squares) into fpart in a fixed order; the last block to arrive
// (an arrival ticket, thread 0 fencing at device scope around it) sums the
// blocks in block order and decides.
template <int U, int DEG>
__global__ __launch_bounds__(256) void rov_final_kernel(const RovStepArgs* __restrict__ pa,
                                                        int terminal) {
  constexpr int K = RovBasis<U, DEG>::K;
  const RovStepArgs& a = *pa;
  const int G = (a.mode == kRovAmerican || a.pooled) ? 1 : a.C;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  __shared__ double seg[256];
  __shared__ double bsum[kRovMaxC * K];
  __shared__ double beta_l[kRovMaxC * K];
  __shared__ double red[3 * kRovMaxC][4];
  __shared__ int s_last;
  const int p = blockIdx.x * 256 + tid;
  double xr[U], later = 0.0, sc1[U][2];
  rov_load_scale<U>(a, 1, sc1);
#pragma unroll
  for (int u = 0; u < U; ++u) xr[u] = 0.0;
  if (p < a.N) {
    rov_load_states<U>(a, 1, p, xr);
    if (rov_decide_reads_later(a, terminal)) later = gmem(a.V)[(size_t)2 * a.N + p];
  }
  if (!terminal) {
    double arow[K];
    int ast;
    rov_load_inverse_row<K>(a, 1, arow, &ast);
    const RMC_GLOBAL double* b1 = gmem(a.bpart) + (size_t)(1 & 1) * a.nblk * a.G * K;
    RovPartials pf;
    rov_partials_issue<K>(a.G, b1, a.nblk, pf);
    rov_partials_sum<K>(a.G, b1, a.nblk, pf, seg, bsum);
    lds_barrier();
    rov_solve_finish<K>(a, 1, seg, bsum, beta_l, arow, ast);
  }
  double x1[kRovMaxC], x2[kRovMaxC], n[kRovMaxC];
#pragma unroll
  for (int c = 0; c < kRovMaxC; ++c) x1[c] = x2[c] = n[c] = 0.0;
  if (p < a.N) {
    int bc;
    const double v1 = rov_decide<U, DEG>(a, terminal, xr, sc1, beta_l, later, &bc);
    gmem(a.V)[(size_t)a.N + p] = v1;
    gmem(a.optc)[(size_t)(a.mode == kRovAmerican ? 0 : 1) * a.N + p] = bc;
    const int g = (G == 1) ? 0 : gmem(a.controls)[p];
    const double v = a.disc * v1;
#pragma unroll
    for (int c = 0; c < kRovMaxC; ++c) {
      const bool in = g == c;
      x1[c] = in ? v : 0.0;
      x2[c] = in ? v * v : 0.0;
      n[c] = in ? 1.0 : 0.0;
    }
  }
#pragma unroll
  for (int c = 0; c < kRovMaxC; ++c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      x1[c] += __shfl_xor(x1[c], off, 64);
      x2[c] += __shfl_xor(x2[c], off, 64);
      n[c] += __shfl_xor(n[c], off, 64);
    }
    if (lane == 0) {
      red[c][w] = x1[c];
      red[kRovMaxC + c][w] = x2[c];
      red[2 * kRovMaxC + c][w] = n[c];
    }
  }
  __syncthreads();
  const int F = 3 * kRovMaxC;
  if (tid < F) gmem(a.fpart)[(size_t)blockIdx.x * F + tid] = ((red[tid][0] + red[tid][1]) + red[tid][2]) +
                                                       red[tid][3];
  __syncthreads();
  if (tid == 0) {
    __threadfence();
    s_last = atomicAdd(&a.ticket[0], 1) == (int)gridDim.x - 1 ? 1 : 0;
    if (s_last) __threadfence();
  }
  __syncthreads();
  if (!s_last) return;
  // Block sums in block order: thread (f, h) adds its tenth of the blocks of
  // quantity f in order, then the ten pieces are added in order.
  __shared__ double part10[3 * kRovMaxC][10];
  __shared__ double tot[3 * kRovMaxC];
  if (tid < F * 10) {
    const int f = tid / 10, h = tid % 10, nb = (int)gridDim.x;
    const int lo = (int)((long long)nb * h / 10), hi = (int)((long long)nb * (h + 1) / 10);
    part10[f][h] = rov_sum_rows<16>(gmem(a.fpart) + (size_t)lo * F + f, F, hi - lo);
  }
  __syncthreads();
  if (tid < F) {
    double s = 0.0;
    for (int h = 0; h < 10; ++h) s += part10[tid][h];
    tot[tid] = s;
  }
  if (tid == 0) atomicExch(&a.ticket[0], 0);
  __syncthreads();
  if (tid != 0) return;
  RMC_GLOBAL double* out = gmem(a.out);
  double mean[kRovMaxC], sd[kRovMaxC], cnt[kRovMaxC];
  for (int c = 0; c < G; ++c) {
    const double y1 = tot[c], y2 = tot[kRovMaxC + c], m = tot[2 * kRovMaxC + c];
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

// Basis shapes the problem validation admits (K <= 35): degree 1 in 1..8
// variables, degree 2 in 1..6, degree 3 in 1..4.
#define ROV_UD_CASES(X)                                                                     \
  X(1, 1) X(2, 1) X(3, 1) X(4, 1) X(5, 1) X(6, 1) X(7, 1) X(8, 1) X(1, 2) X(2, 2) X(3, 2)   \
  X(4, 2) X(5, 2) X(6, 2) X(1, 3) X(2, 3) X(3, 3) X(4, 3)
#define ROV_K_CASES(X)                                                                     \
  X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(15) X(20) X(21) X(28) X(35)

size_t rov_invert_lds(int K) {
  const int E = K * (K + 1) / 2 + K + 1;
  int H = 1024 / E;
  H = H < 1 ? 1 : (H > 8 ? 8 : H);
  return ((size_t)H * E + E + 2 * K * K + 2 * K) * sizeof(double);
}

hipError_t rov_set_lds_limits() {
#define ROV_LDS_UD(UU, DD)                                                                  \
  {                                                                                         \
    hipError_t e = hipFuncSetAttribute(                                                     \
        reinterpret_cast<const void*>(&rov_gram_kernel<UU, DD>),                            \
        hipFuncAttributeMaxDynamicSharedMemorySize,                                         \
        (int)rov_gram_lds(RovBasis<UU, DD>::K, UU));                                         \
    if (e != hipSuccess) return e;                                                          \
  }
  ROV_UD_CASES(ROV_LDS_UD)
#undef ROV_LDS_UD
#define ROV_LDS_K(KK)                                                                       \
  {                                                                                         \
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&rov_invert_kernel<KK>), \
                                       hipFuncAttributeMaxDynamicSharedMemorySize,            \
                                       (int)rov_invert_lds(KK));                              \
    if (e != hipSuccess) return e;                                                          \
  }
  ROV_K_CASES(ROV_LDS_K)
#undef ROV_LDS_K
  return hipSuccess;
}

hipError_t launch_rov_forward(const RovForwardArgs& a, hipStream_t s) {
  if (a.N <= 0) return hipSuccess;
  const dim3 grid((a.N + 255) / 256);
  switch (a.U) {
#define ROV_FWD_U(UU)                                                                       \
  case UU:                                                                                  \
    hipLaunchKernelGGL(rov_forward_kernel<UU>, grid, dim3(256), 0, s, a);                   \
    break;
    ROV_FWD_U(1) ROV_FWD_U(2) ROV_FWD_U(3) ROV_FWD_U(4) ROV_FWD_U(5) ROV_FWD_U(6) ROV_FWD_U(7)
    ROV_FWD_U(8)
#undef ROV_FWD_U
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rov_moments(const double* states, int N, int U, int T1, double* scale,
                              hipStream_t s) {
  // The chunk sums go after the T1 U (mean, 1/SD) pairs of scale.
  double* part = scale + (size_t)T1 * U * 2;
  hipLaunchKernelGGL(rov_moments_kernel, dim3(T1 * U, kRovMomentChunks), dim3(256), 0, s, states,
                     N, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(rov_moments_final_kernel, dim3((T1 * U + 63) / 64), dim3(64), 0, s, part, N,
                     T1 * U, scale);
  return hipGetLastError();
}

hipError_t launch_rov_gram(const RovStepArgs& a, const RovStepArgs* pa, int t_lo, int n_steps,
                           hipStream_t s) {
  if (n_steps <= 0) return hipSuccess;
  const dim3 grid((a.N + kRovGramPaths - 1) / kRovGramPaths, n_steps), inv_grid(n_steps, a.G);
  bool found = false;
#define ROV_GRAM_UD(UU, DD)                                                                 \
  if (a.U == UU && a.deg == DD) {                                                           \
    found = true;                                                                           \
    hipLaunchKernelGGL((rov_gram_kernel<UU, DD>), grid, dim3(256),                          \
                       rov_gram_lds(RovBasis<UU, DD>::K, UU), s, pa, t_lo);                 \
  }
  ROV_UD_CASES(ROV_GRAM_UD)
#undef ROV_GRAM_UD
  if (!found) return hipErrorInvalidValue;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  switch (a.K) {
#define ROV_INV_K(KK)                                                                       \
  case KK:                                                                                  \
    hipLaunchKernelGGL(rov_invert_kernel<KK>, inv_grid, dim3(1024), rov_invert_lds(KK), s,   \
                       pa, t_lo, (int)grid.x);                                              \
    break;
    ROV_K_CASES(ROV_INV_K)
#undef ROV_INV_K
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rov_step(const RovStepArgs& a, const RovStepArgs* pa, int t, int decide_next,
                           hipStream_t s) {
  bool found = false;
#define ROV_STEP_UD(UU, DD)                                                                 \
  if (a.U == UU && a.deg == DD) {                                                           \
    found = true;                                                                           \
    hipLaunchKernelGGL((rov_step_kernel<UU, DD>), dim3(a.nblk), dim3(256), 0, s, pa, t,     \
                       decide_next);                                                        \
  }
  ROV_UD_CASES(ROV_STEP_UD)
#undef ROV_STEP_UD
  return found ? hipGetLastError() : hipErrorInvalidValue;
}

hipError_t launch_rov_final(const RovStepArgs& a, const RovStepArgs* pa, bool terminal,
                            hipStream_t s) {
  const dim3 grid((a.N + 255) / 256);
  bool found = false;
#define ROV_FINAL_UD(UU, DD)                                                                \
  if (a.U == UU && a.deg == DD) {                                                           \
    found = true;                                                                           \
    hipLaunchKernelGGL((rov_final_kernel<UU, DD>), grid, dim3(256), 0, s, pa,               \
                       terminal ? 1 : 0);                                                   \
  }
  ROV_UD_CASES(ROV_FINAL_UD)
#undef ROV_FINAL_UD
  return found ? hipGetLastError() : hipErrorInvalidValue;
}

hipError_t launch_exp_pv(const ExpPvArgs& a, hipStream_t s) {
  if (a.N <= 0) return hipSuccess;
  hipLaunchKernelGGL(exp_pv_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace roadmc
