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

__global__ void rov_forward_kernel(RovForwardArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.N) return;
  const int U = a.U;
  double x[kRovMaxU];
  for (int u = 0; u < U; ++u) {
    x[u] = a.par[u * 7 + 1];
    a.states[(size_t)u * a.N + p] = x[u];
  }
  const double dt = a.dt, sqdt = sqrt(a.dt);
  for (int t = 0; t < a.T; ++t) {
    double z[kRovMaxU], zj[kRovMaxU], uj[kRovMaxU];
    for (int b = 0; b < (U + 3) / 4; ++b) {
      double w[4];
      rov_normals((uint32_t)p, (uint32_t)t, (uint32_t)b, a.tag, a.k0, a.k1, w);
      for (int i = 0; i < 4 && 4 * b + i < U; ++i) z[4 * b + i] = w[i];
      rov_normals((uint32_t)p, (uint32_t)t, (uint32_t)(2 + b), a.tag, a.k0, a.k1, w);
      for (int i = 0; i < 4 && 4 * b + i < U; ++i) zj[4 * b + i] = w[i];
      const rmc_u4 r = rmc_philox((uint32_t)p, (uint32_t)t, (uint32_t)(4 + b), a.tag, a.k0, a.k1);
      const uint32_t rw[4] = {r.x, r.y, r.z, r.w};
      for (int i = 0; i < 4 && 4 * b + i < U; ++i) uj[4 * b + i] = rmc_u01d(rw[i]);
    }
    // Correlated diffusion normals: L z with L the Cholesky factor.
    double zc[kRovMaxU];
    for (int u = 0; u < U; ++u) {
      double s = 0.0;
      for (int v = 0; v <= u; ++v) s += a.chol[u * kRovMaxU + v] * z[v];
      zc[u] = s;
    }
    for (int u = 0; u < U; ++u) {
      const double* q = a.par + u * 7;  // kind, x0, mean, sd, reversion, J, p
      const int kind = (int)q[0];
      const double xv = x[u];
      double nx;
      if (kind == kSdeOu) {
        nx = xv + q[4] * (q[2] - xv) * dt + q[3] * sqdt * zc[u];
      } else if (kind == kSdeGbm) {
        nx = xv * exp((q[2] - 0.5 * q[3] * q[3]) * dt + q[3] * sqdt * zc[u]);
      } else {
        // SimulateGPU.cu:930-952: Brownian term x * sd * z, jump size
        // J' = J^2 z' - J^2 / 2 applied when the uniform is below p.
        const double jump = (uj[u] < q[6]) ? 1.0 : 0.0;
        const double jsz = zj[u] * q[5] * q[5] - q[5] * q[5] / 2.0;
        nx = xv + q[4] * (q[2] - xv) * dt + xv * q[3] * zc[u] + (exp(jsz) - 1.0) * xv * jump;
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

// Mean and 1/SD of every state dimension at every step (feature scaling).
__global__ void rov_moments_kernel(const double* __restrict__ states, int N, int U,
                                   double* __restrict__ scale) {
  const int t = blockIdx.x;
  __shared__ double s1[256], s2[256];
  for (int u = 0; u < U; ++u) {
    const double* x = states + ((size_t)t * U + u) * N;
    double a = 0.0, b = 0.0;
    for (int p = threadIdx.x; p < N; p += blockDim.x) {
      a += x[p];
      b += x[p] * x[p];
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
      scale[((size_t)t * U + u) * 2 + 0] = m;
      scale[((size_t)t * U + u) * 2 + 1] = (sd > 1e-12 * (fabs(m) + 1.0)) ? 1.0 / sd : 1.0;
    }
    __syncthreads();
  }
}

// All monomials of degree <= deg in U variables, in graded order:
// 1, x_i, x_i x_j (i <= j), x_i x_j x_k (i <= j <= k).
__device__ __forceinline__ int rov_basis(const double* x, int U, int deg, double* phi) {
  int k = 0;
  phi[k++] = 1.0;
  if (deg >= 1)
    for (int i = 0; i < U; ++i) phi[k++] = x[i];
  if (deg >= 2)
    for (int i = 0; i < U; ++i)
      for (int j = i; j < U; ++j) phi[k++] = x[i] * x[j];
  if (deg >= 3)
    for (int i = 0; i < U; ++i)
      for (int j = i; j < U; ++j)
        for (int l = j; l < U; ++l) phi[k++] = x[i] * x[j] * x[l];
  return k;
}

__device__ __forceinline__ void rov_features(const RovStepArgs& a, int t, int p, double* x) {
  for (int u = 0; u < a.U; ++u) {
    double v = a.states[((size_t)t * a.U + u) * a.N + p];
    if (a.scale) {
      const double* sc = a.scale + ((size_t)t * a.U + u) * 2;
      v = (v - sc[0]) * sc[1];
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
__global__ __launch_bounds__(256) void rov_assemble_kernel(RovStepArgs a, int t) {
  extern __shared__ double sm[];
  const int K = a.K, E = a.E, B = blockDim.x;
  double* sphi = sm;               // [K][B]
  double* sy = sphi + (size_t)K * B;
  int* sg = reinterpret_cast<int*>(sy + B);
  const int p = blockIdx.x * B + threadIdx.x;
  int g = -1;
  if (p < a.N) {
    g = rov_fit_group(a, t, p);
    double x[kRovMaxU], phi[kRovMaxK];
    rov_features(a, t, p, x);
    rov_basis(x, a.U, a.deg, phi);
    for (int k = 0; k < K; ++k) sphi[(size_t)k * B + threadIdx.x] = phi[k];
    sy[threadIdx.x] = a.V[(size_t)(t + 1) * a.N + p];
  } else {
    for (int k = 0; k < K; ++k) sphi[(size_t)k * B + threadIdx.x] = 0.0;
    sy[threadIdx.x] = 0.0;
  }
  sg[threadIdx.x] = g;
  __syncthreads();
  const int n_tri = K * (K + 1) / 2;
  for (int e = threadIdx.x; e < a.G * E; e += B) {
    const int grp = e / E, ee = e - grp * E;
    double acc = 0.0;
    if (ee < n_tri) {
      // (i, j), i <= j, of the upper triangle in row-major order.
      int i = 0, rem = ee;
      while (rem >= K - i) {
        rem -= K - i;
        ++i;
      }
      const int j = i + rem;
      const double* pi = sphi + (size_t)i * B;
      const double* pj = sphi + (size_t)j * B;
      for (int q = 0; q < B; ++q)
        if (sg[q] == grp) acc += pi[q] * pj[q];
    } else if (ee < n_tri + K) {
      const double* pi = sphi + (size_t)(ee - n_tri) * B;
      for (int q = 0; q < B; ++q)
        if (sg[q] == grp) acc += pi[q] * sy[q];
    } else {
      for (int q = 0; q < B; ++q)
        if (sg[q] == grp) acc += 1.0;
    }
    a.partial[((size_t)blockIdx.x * a.G + grp) * E + ee] = acc;
  }
}

// Reduces the block partials in a fixed order and solves each group's K x K
// system in FP64: Cholesky, or LU with partial pivoting when the matrix is not
// numerically positive definite.  Groups with fewer points than K keep zero
// coefficients (multiQuadReg skips them, SimulateGPU.cu:6101).
__global__ __launch_bounds__(256) void rov_solve_kernel(RovStepArgs a, int t, int nblk) {
  extern __shared__ double sm[];
  const int K = a.K, E = a.E, G = a.G;
  double* sums = sm;  // [G][E]
  for (int e = threadIdx.x; e < G * E; e += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += a.partial[(size_t)b * G * E + e];
    sums[e] = s;
  }
  __syncthreads();
  if (threadIdx.x < G) {
    const int grp = threadIdx.x;
    double* coef = a.beta + ((size_t)t * G + grp) * K;
    const double* S = sums + (size_t)grp * E;
    const int n_tri = K * (K + 1) / 2;
    double M[kRovMaxK * kRovMaxK], r[kRovMaxK];
    for (int i = 0, e = 0; i < K; ++i)
      for (int j = i; j < K; ++j, ++e) {
        M[i * K + j] = S[e];
        M[j * K + i] = S[e];
      }
    for (int i = 0; i < K; ++i) r[i] = S[n_tri + i];
    const double count = S[n_tri + K];
    int status = 0;
    if (count < (double)K) {
      for (int i = 0; i < K; ++i) coef[i] = 0.0;
      status = 1;
    } else {
      // Cholesky M = L L^T in place (lower triangle).
      double L[kRovMaxK * kRovMaxK];
      bool spd = true;
      double dmax = 0.0;
      for (int i = 0; i < K; ++i) dmax = fmax(dmax, fabs(M[i * K + i]));
      for (int j = 0; j < K && spd; ++j) {
        double d = M[j * K + j];
        for (int k = 0; k < j; ++k) d -= L[j * K + k] * L[j * K + k];
        if (!(d > 1e-13 * dmax)) {
          spd = false;
          break;
        }
        const double ljj = sqrt(d);
        L[j * K + j] = ljj;
        for (int i = j + 1; i < K; ++i) {
          double s = M[i * K + j];
          for (int k = 0; k < j; ++k) s -= L[i * K + k] * L[j * K + k];
          L[i * K + j] = s / ljj;
        }
      }
      if (spd) {
        double y[kRovMaxK];
        for (int i = 0; i < K; ++i) {
          double s = r[i];
          for (int k = 0; k < i; ++k) s -= L[i * K + k] * y[k];
          y[i] = s / L[i * K + i];
        }
        for (int i = K - 1; i >= 0; --i) {
          double s = y[i];
          for (int k = i + 1; k < K; ++k) s -= L[k * K + i] * coef[k];
          coef[i] = s / L[i * K + i];
        }
      } else {
        // Gaussian elimination with partial pivoting on |M[i][col]| (the
        // pivot search the reference intends, SimulateGPU.cu:6307-6316).
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
            for (int j = i + 1; j < K; ++j) s -= M[i * K + j] * coef[j];
            coef[i] = s / M[i * K + i];
          }
          status = 2;
        } else {
          for (int i = 0; i < K; ++i) coef[i] = 0.0;
          status = 3;
        }
      }
    }
    a.status[(size_t)t * G + grp] = status;
  }
}

// Decision of step t for every path.
__global__ void rov_decide_kernel(RovStepArgs a, int t, int terminal) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.N) return;
  const double disc = a.disc;
  double* V = a.V;
  if (a.mode == kRovAmerican) {
    const double e = fmax(a.strike - a.states[((size_t)t * a.U) * a.N + p], 0.0);
    if (terminal) {
      V[(size_t)t * a.N + p] = e;
      a.optc[(size_t)(t - 1) * a.N + p] = e > 0.0 ? 1 : 0;
      return;
    }
    const double later = disc * V[(size_t)(t + 1) * a.N + p];
    int ex = 0;
    if (e > 0.0 || !a.itm_only) {
      double x[kRovMaxU], phi[kRovMaxK];
      rov_features(a, t, p, x);
      const int K = rov_basis(x, a.U, a.deg, phi);
      const double* b = a.beta + (size_t)t * a.G * K;
      double cont = 0.0;
      for (int k = 0; k < K; ++k) cont += phi[k] * b[k];
      cont *= disc;
      ex = e > cont ? 1 : 0;
    }
    double v;
    if (a.value_mode == kValueRealized) {
      v = ex ? e : later;
    } else {
      v = ex ? e : later;
    }
    V[(size_t)t * a.N + p] = v;
    a.optc[(size_t)(t - 1) * a.N + p] = ex;
    return;
  }
  // Flow control (backwardInduction, SimulateGPU.cu:2333-2784).
  const double up = rov_unit_profit(a, t, p);
  const double m = rov_flow_scale(a, t, p);
  double pay[kRovMaxC], cur[kRovMaxC];
  double x[kRovMaxU], phi[kRovMaxK];
  int K = 0;
  if (!terminal) {
    rov_features(a, t, p, x);
    K = rov_basis(x, a.U, a.deg, phi);
  }
  for (int c = 0; c < a.C; ++c) {
    cur[c] = a.flow[c] * m * up;
    if (terminal) {
      // Terminal step: payoff clamped at zero (SimulateGPU.cu:2385-2416).
      pay[c] = fmin(cur[c], 0.0);
    } else {
      const double* b = a.beta + ((size_t)t * a.G + (a.pooled ? 0 : c)) * K;
      double cont = 0.0;
      for (int k = 0; k < K; ++k) cont += phi[k] * b[k];
      cont *= disc;                 // SimulateGPU.cu:2703
      if (cont > 0.0) cont = 0.0;   // SimulateGPU.cu:2712-2714
      pay[c] = cur[c] + cont;
    }
  }
  double best = pay[0];
  int bc = 0;
  for (int c = 1; c < a.C; ++c) {
    if (pay[c] < a.hysteresis * best) {  // SimulateGPU.cu:2757-2771
      best = pay[c];
      bc = c;
    }
  }
  double v = best;
  if (a.value_mode == kValueRealized)
    v = terminal ? pay[bc] : cur[bc] + disc * V[(size_t)(t + 1) * a.N + p];
  V[(size_t)t * a.N + p] = v;
  a.optc[(size_t)t * a.N + p] = bc;
}

// Time 0 (firstPeriodInduction, SimulateGPU.cu:2210-2330): per initial
// control, the mean discounted value of the paths that took it, plus the
// deterministic current payoff; out = {value, control, SD}.
__global__ void rov_first_kernel(RovStepArgs a, double* out) {
  __shared__ double s1[kRovMaxC][256], s2[kRovMaxC][256], sn[kRovMaxC][256];
  const int G = (a.mode == kRovAmerican || a.pooled) ? 1 : a.C;
  for (int c = 0; c < G; ++c) {
    double x1 = 0.0, x2 = 0.0, n = 0.0;
    for (int p = threadIdx.x; p < a.N; p += blockDim.x) {
      const int g = (G == 1) ? 0 : a.controls[p];
      if (g != c) continue;
      const double v = a.disc * a.V[(size_t)a.N + p];
      x1 += v;
      x2 += v * v;
      n += 1.0;
    }
    s1[c][threadIdx.x] = x1;
    s2[c][threadIdx.x] = x2;
    sn[c][threadIdx.x] = n;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double mean[kRovMaxC], sd[kRovMaxC], cnt[kRovMaxC];
  for (int c = 0; c < G; ++c) {
    double x1 = 0.0, x2 = 0.0, n = 0.0;
    for (int i = 0; i < (int)blockDim.x; ++i) {
      x1 += s1[c][i];
      x2 += s2[c][i];
      n += sn[c][i];
    }
    cnt[c] = n;
    mean[c] = n > 0.0 ? x1 / n : 0.0;
    sd[c] = n > 0.0 ? sqrt(fmax(x2 / n - mean[c] * mean[c], 0.0)) : 0.0;
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
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&rov_solve_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             kRovMaxC * (kRovMaxK * (kRovMaxK + 1) / 2 + kRovMaxK + 1) * 8);
}

hipError_t launch_rov_forward(const RovForwardArgs& a, hipStream_t s) {
  if (a.N <= 0) return hipSuccess;
  hipLaunchKernelGGL(rov_forward_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_rov_moments(const double* states, int N, int U, int T1, double* scale,
                              hipStream_t s) {
  hipLaunchKernelGGL(rov_moments_kernel, dim3(T1), dim3(256), 0, s, states, N, U, scale);
  return hipGetLastError();
}

size_t rov_assemble_lds(int K) { return (size_t)K * 256 * 8 + 256 * 8 + 256 * 4; }

hipError_t launch_rov_step(const RovStepArgs& a, int t, bool fit, bool terminal, int nblk,
                           hipStream_t s) {
  if (fit) {
    hipLaunchKernelGGL(rov_assemble_kernel, dim3(nblk), dim3(256), rov_assemble_lds(a.K), s, a,
                       t);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(rov_solve_kernel, dim3(1), dim3(256), (size_t)a.G * a.E * 8, s, a, t,
                       nblk);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(rov_decide_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a, t,
                     terminal ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_rov_first(const RovStepArgs& a, double* out, hipStream_t s) {
  hipLaunchKernelGGL(rov_first_kernel, dim3(1), dim3(256), 0, s, a, out);
  return hipGetLastError();
}

hipError_t launch_exp_pv(const ExpPvArgs& a, hipStream_t s) {
  if (a.N <= 0) return hipSuccess;
  hipLaunchKernelGGL(exp_pv_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace roadmc
