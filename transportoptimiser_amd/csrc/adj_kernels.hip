// Adjoint of the MTE simulation (LOGNORMAL law) for gfx950.
//
// Objective J = dobj * (1/N) sum_paths sum_patches n_T.  With
// n'_t = n_{t-1} M (M = transProbs o survProbs, origin rows) and
// n_t = G(n'_t) = n'_t (1 + r (1 - n'_t/K)) e_t, the reverse sweep is
//   lambda_T = dobj / N,
//   mu_t = lambda_t o G'(n'_t),         G'(x) = (1 + r (1 - 2x/K)) e_t,
//   dJ/dr += sum_paths lambda_t o n'_t (1 - n'_t/K) e_t,
//   dJ/dM += n_{t-1}^T mu_t,
//   lambda_{t-1} = mu_t M^T.
// The forward pass (mte_kernels.hip with a checkpoint buffer) stores the state
// every L years.  Each segment of L years is then handled by one launch of
// adj_segment_kernel: every workgroup reloads its tile's checkpoint,
// recomputes the segment (states and pre-growth states to an HBM staging
// area), and sweeps the segment backwards with the same MFMA product loop as
// the forward pass, multiplying by a packed M^T.  adj_gradm_kernel then adds
// the segment's n^T mu products (a split-K GEMM over paths x years).
#include <hip/hip_runtime.h>

#include "adj_kernels.h"
#include "mte_common.h"

namespace roadmc {

template <class TR>
__device__ __forceinline__ void adj_noise(const AdjJob& jb, uint32_t path, int year, int col,
                                          typename TR::T* e, typename TR::T sig,
                                          typename TR::T half_var) {
  using T = typename TR::T;
  T z[4];
  TR::normal4(path >> 2, (uint32_t)year, (uint32_t)col, jb.tag0, jb.key0, jb.key1, z);
  if constexpr (sizeof(T) == 4)
    *e = __builtin_amdgcn_exp2f(fmaf(sig, z[path & 3u], -half_var));
  else
    *e = exp(sig * z[path & 3u] - half_var);
}

template <class TR>
__global__ __launch_bounds__(kMteThreads) void adj_segment_kernel(AdjJob jb) {
  using T = typename TR::T;
  using VecT = typename TR::VecT;
  using AccT = typename TR::AccT;
  constexpr int TM = TR::TM, TN = TR::TN, KV = TR::KV, NACC = TR::NACC, NT = TR::NT;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nthreads = blockDim.x, nwaves = nthreads >> 6;
  const int BM = jb.BM, Pp = jb.Pp, LDA = Pp + TR::PAD, N = jb.N;
  const int tile = blockIdx.x, p0 = tile * BM;
  T* A = reinterpret_cast<T*>(smem);
  // Growth-rate gradient: gr[c] is this workgroup's running sum; grb holds
  // one row block's column sums per year, so that every sum is formed in a
  // fixed order (no floating-point atomics, reproducible results).
  double* gr = reinterpret_cast<double*>(smem + align16((size_t)BM * LDA * sizeof(T)));
  double* grb = gr + Pp;
  const int nrb = BM / TM;

  const T* cols = static_cast<const T*>(jb.cols);
  const T* capK = cols;
  const T* capKinv = cols + Pp;
  const T* rate = cols + 2 * Pp;
  const T l2e = sizeof(T) == 4 ? (T)1.4426950408889634 : (T)1;
  const T sig = (T)(jb.sigma * (sizeof(T) == 4 ? 1.4426950408889634 : 1.0));
  const T half_var = (T)(0.5 * jb.sigma * jb.sigma) * l2e;

  T* ns = static_cast<T*>(jb.ns);
  T* np = static_cast<T*>(jb.np);
  T* mu = static_cast<T*>(jb.mu);
  T* lam = static_cast<T*>(jb.lam);
  const size_t plane = (size_t)N * Pp;  // one year of staging

  for (int c = tid; c < Pp; c += nthreads) gr[c] = 0.0;
  // Checkpoint of the segment (state entering year y0).
  const T* ck = static_cast<const T*>(jb.ckpt) + (size_t)(jb.y0 / jb.L) * plane;
  for (int idx = tid; idx < BM * Pp; idx += nthreads) {
    const int r = idx / Pp, c = idx - r * Pp;
    A[r * LDA + TR::perm(c)] = (p0 + r < N) ? ck[(size_t)(p0 + r) * Pp + c] : (T)0;
  }
  __syncthreads();

  const int ntc = Pp / TN;
  const int ntiles = (BM / TM) * ntc;
  const int base = ntiles / nwaves, extra = ntiles % nwaves;
  const int my_n = base + (wave < extra ? 1 : 0);
  const int my_t0 = wave * base + (wave < extra ? wave : extra);
  int trow[NT], tcol[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int ti = my_t0 + t;
    trow[t] = ti / ntc;
    tcol[t] = ti - (ti / ntc) * ntc;
  }
  const int nkg = Pp >> 3;
  const int arow = TR::a_row(lane), akoff = KV * TR::k_lane(lane);
  AccT acc[NT];
  VecT b0[NT];

  auto product = [&](const PackedB& Bm) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      acc[t] = AccT{};
      if (t < my_n) b0[t] = load_b<VecT>(Bm, b_voff(Bm, tcol[t], lane), 0);
    }
    switch (my_n) {
      case 1: year_product<TR, 1>(acc, b0, Bm, A, trow, tcol, LDA, arow, akoff, lane); break;
      case 2: year_product<TR, 2>(acc, b0, Bm, A, trow, tcol, LDA, arow, akoff, lane); break;
      case 3: year_product<TR, 3>(acc, b0, Bm, A, trow, tcol, LDA, arow, akoff, lane); break;
      case 4: year_product<TR, 4>(acc, b0, Bm, A, trow, tcol, LDA, arow, akoff, lane); break;
      default:
        if constexpr (NT > 4) {
          switch (my_n) {
            case 5: year_product<TR, 5>(acc, b0, Bm, A, trow, tcol, LDA, arow, akoff, lane); break;
            case 6: year_product<TR, 6>(acc, b0, Bm, A, trow, tcol, LDA, arow, akoff, lane); break;
            case 7: year_product<TR, 7>(acc, b0, Bm, A, trow, tcol, LDA, arow, akoff, lane); break;
            case 8: year_product<TR, 8>(acc, b0, Bm, A, trow, tcol, LDA, arow, akoff, lane); break;
            default: break;
          }
        }
        break;
    }
  };

  const PackedB Mp = packed_b(jb.mpack, nkg, (size_t)Pp * Pp * sizeof(T));
  const PackedB Mtp = packed_b(jb.mtpack, nkg, (size_t)Pp * Pp * sizeof(T));

  // Recompute the segment: ns[i] = n_{y0+i}, np[i] = n'_{y0+i+1}.
  for (int yr = jb.y0; yr <= jb.y1; ++yr) {
    const int i = yr - jb.y0;
    for (int idx = tid; idx < BM * Pp; idx += nthreads) {
      const int r = idx / Pp, c = idx - r * Pp;
      if (p0 + r < N) ns[i * plane + (size_t)(p0 + r) * Pp + c] = A[r * LDA + TR::perm(c)];
    }
    product(Mp);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t < my_n) {
        const int col = tcol[t] * TN + TR::acc_col(lane);
        const T Ki = capKinv[col], rr = rate[col];
#pragma unroll
        for (int r = 0; r < NACC; ++r) {
          const int row = trow[t] * TM + TR::acc_row(lane, r);
          const uint32_t path = (uint32_t)(p0 + row);
          T e;
          adj_noise<TR>(jb, path, yr, col, &e, sig, half_var);
          const T x = acc[t][r];
          if ((int)path < N) np[i * plane + (size_t)path * Pp + col] = x;
          A[row * LDA + TR::perm(col)] = x * ((T)1 + rr * ((T)1 - x * Ki)) * e;
        }
      }
    }
    __syncthreads();
  }

  // Top of the segment: mu_{y1+1} from lambda_{y1+1}; one column per thread.
  {
    const int i = jb.y1 - jb.y0;
    const bool last = (jb.y1 == jb.T - 1);
    for (int c = tid; c < Pp; c += nthreads) {
      const T Ki = capKinv[c], rr = rate[c];
      double gsum = 0.0;
      for (int r = 0; r < BM; ++r) {
        const uint32_t path = (uint32_t)(p0 + r);
        T m = (T)0;
        if ((int)path < N && c < jb.P) {
          const T l = last ? (T)(jb.dobj / N) : lam[(size_t)path * Pp + c];
          const T x = np[i * plane + (size_t)path * Pp + c];
          T e;
          adj_noise<TR>(jb, path, jb.y1, c, &e, sig, half_var);
          m = l * ((T)1 + rr * ((T)1 - (T)2 * x * Ki)) * e;
          gsum += (double)(l * x * ((T)1 - x * Ki) * e);
        }
        if ((int)path < N) mu[i * plane + (size_t)path * Pp + c] = m;
        A[r * LDA + TR::perm(c)] = m;
      }
      gr[c] = gsum;
    }
    __syncthreads();
  }

  // Backward sweep: lambda_yr = mu_{yr+1} M^T, then mu_yr.
  for (int yr = jb.y1; yr >= jb.y0; --yr) {
    product(Mtp);
    __syncthreads();
    const int i = yr - 1 - jb.y0;  // staging slot of n'_yr and mu_yr
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t < my_n) {
        const int col = tcol[t] * TN + TR::acc_col(lane);
        const T Ki = capKinv[col], rr = rate[col];
        double gsum = 0.0;
#pragma unroll
        for (int r = 0; r < NACC; ++r) {
          const int row = trow[t] * TM + TR::acc_row(lane, r);
          const uint32_t path = (uint32_t)(p0 + row);
          const T l = acc[t][r];
          if (yr == jb.y0) {
            if ((int)path < N) lam[(size_t)path * Pp + col] = l;
          } else {
            T m = (T)0;
            if ((int)path < N && col < jb.P) {
              const T x = np[i * plane + (size_t)path * Pp + col];
              T e;
              adj_noise<TR>(jb, path, yr - 1, col, &e, sig, half_var);
              m = l * ((T)1 + rr * ((T)1 - (T)2 * x * Ki)) * e;
              gsum += (double)(l * x * ((T)1 - x * Ki) * e);
            }
            if ((int)path < N) mu[i * plane + (size_t)path * Pp + col] = m;
            A[row * LDA + TR::perm(col)] = m;
          }
        }
        if (yr != jb.y0) {
          // Lanes holding the same column combine in a fixed order; one lane
          // per column writes the row block's sum.
#pragma unroll
          for (int off = TN; off < 64; off <<= 1) gsum += __shfl_xor(gsum, off, 64);
          if (lane < TN) grb[trow[t] * Pp + col] = gsum;
        }
      }
    }
    __syncthreads();
    if (yr != jb.y0) {
      for (int c = tid; c < Pp; c += nthreads) {
        double v = gr[c];
        for (int rb = 0; rb < nrb; ++rb) v += grb[rb * Pp + c];
        gr[c] = v;
      }
    }
  }
  __syncthreads();
  for (int c = tid; c < Pp; c += nthreads) jb.grad_part[(size_t)tile * Pp + c] = gr[c];
}

// grad_r[c] += sum over tiles of the segment's partials, in tile order.
__global__ void adj_reduce_r_kernel(const double* __restrict__ part, int n_tiles, int Pp, int P,
                                    double* __restrict__ grad_r) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P) return;
  double s = 0.0;
  for (int t = 0; t < n_tiles; ++t) s += part[(size_t)t * Pp + c];
  grad_r[c] += s;
}

// dJ/dM += sum_{i, paths} ns[i]^T mu[i] over the segment (K = n_year * N rows
// of [Pp] each).  Grid (Pp/TB)^2 output blocks x nsplit chunks of K; each
// workgroup (4 waves) forms one TB x TB block of one chunk on MFMA and writes
// it to its partial slot; adj_reduce_kernel sums the slots in order (FP64).
template <class TR>
__global__ __launch_bounds__(256) void adj_gradm_kernel(const typename TR::T* __restrict__ A,
                                                        const typename TR::T* __restrict__ B,
                                                        int Ktot, int Pp, int kchunk,
                                                        double* __restrict__ partial) {
  using T = typename TR::T;
  using AccT = typename TR::AccT;
  constexpr int TM = TR::TM, NACC = TR::NACC;
  constexpr int KS = 64 / TM;           // k per MFMA: 2 (f32) or 4 (f64)
  constexpr int TPW = (64 / TM) * (64 / TM) / 4;  // tiles per wave in a 64x64 block
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nb = Pp / 64;
  const int ob = blockIdx.x % (nb * nb), split = blockIdx.x / (nb * nb);
  const int i0 = (ob / nb) * 64, j0 = (ob % nb) * 64;
  const int kb = split * kchunk, ke = min(Ktot, kb + kchunk);
  AccT acc[TPW];
  int ti[TPW], tj[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tt = wave * TPW + t;
    ti[t] = i0 + (tt / (64 / TM)) * TM;
    tj[t] = j0 + (tt % (64 / TM)) * TM;
    acc[t] = AccT{};
  }
  const int kl = TR::k_lane(lane), al = TR::a_row(lane);
  for (int k = kb; k < ke; k += KS) {
    const int kk = k + kl;
    const bool ok = kk < ke;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const T a = ok ? A[(size_t)kk * Pp + ti[t] + al] : (T)0;
      const T b = ok ? B[(size_t)kk * Pp + tj[t] + al] : (T)0;
      acc[t] = TR::mfma(a, b, acc[t]);
    }
  }
  double* out = partial + (size_t)split * Pp * Pp;
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int r = 0; r < NACC; ++r)
      out[(size_t)(ti[t] + TR::acc_row(lane, r)) * Pp + tj[t] + TR::acc_col(lane)] =
          (double)acc[t][r];
}

__global__ void adj_reduce_kernel(const double* __restrict__ partial, int nsplit, int n,
                                  double* __restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += partial[(size_t)k * n + e];
  grad[e] += s;
}

size_t adj_lds_bytes(bool fp64, int BM, int Pp) {
  const size_t a = fp64 ? align16((size_t)BM * (Pp + TraitsF64::PAD) * 8)
                        : align16((size_t)BM * (Pp + TraitsF32::PAD) * 4);
  const int rows = fp64 ? TraitsF64::TM : TraitsF32::TM;
  return a + (size_t)Pp * 8 * (1 + BM / rows);
}

hipError_t adj_set_lds_limit(bool fp64, size_t bytes) {
  return fp64 ? hipFuncSetAttribute(reinterpret_cast<const void*>(&adj_segment_kernel<TraitsF64>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)
              : hipFuncSetAttribute(reinterpret_cast<const void*>(&adj_segment_kernel<TraitsF32>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t launch_adj_segment(bool fp64, const AdjJob& jb, int n_tiles, size_t lds,
                              hipStream_t s) {
  if (fp64)
    hipLaunchKernelGGL(adj_segment_kernel<TraitsF64>, dim3(n_tiles), dim3(kMteThreads), lds, s,
                       jb);
  else
    hipLaunchKernelGGL(adj_segment_kernel<TraitsF32>, dim3(n_tiles), dim3(kMteThreads), lds, s,
                       jb);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(adj_reduce_r_kernel, dim3((jb.P + 255) / 256), dim3(256), 0, s,
                     jb.grad_part, n_tiles, jb.Pp, jb.P, jb.grad_r);
  return hipGetLastError();
}

hipError_t launch_adj_gradm(bool fp64, const void* A, const void* B, int Ktot, int Pp,
                            int nsplit, double* partial, double* grad, hipStream_t s) {
  const int kchunk = ((Ktot + nsplit - 1) / nsplit + 3) / 4 * 4;
  const int nb = Pp / 64;
  const dim3 grid(nb * nb * nsplit);
  if (fp64)
    hipLaunchKernelGGL(adj_gradm_kernel<TraitsF64>, grid, dim3(256), 0, s,
                       static_cast<const double*>(A), static_cast<const double*>(B), Ktot, Pp,
                       kchunk, partial);
  else
    hipLaunchKernelGGL(adj_gradm_kernel<TraitsF32>, grid, dim3(256), 0, s,
                       static_cast<const float*>(A), static_cast<const float*>(B), Ktot, Pp,
                       kchunk, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int n = Pp * Pp;
  hipLaunchKernelGGL(adj_reduce_kernel, dim3((n + 255) / 256), dim3(256), 0, s, partial, nsplit,
                     n, grad);
  return hipGetLastError();
}

}  // namespace roadmc
