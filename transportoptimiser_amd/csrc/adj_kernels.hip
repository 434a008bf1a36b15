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
#include "gmem.h"
#include "mte_common.h"

namespace roadmc {

// Lognormal year factors e = exp(s z - s^2/2) of the four paths of quad q
// (paths 4q .. 4q+3) in patch `col`: one Philox call, as in the forward pass.
template <class TR>
__device__ __forceinline__ void adj_noise4(const AdjJob& jb, uint32_t quad, int year, int col,
                                           typename TR::T e[4], typename TR::T sig,
                                           typename TR::T half_var) {
  using T = typename TR::T;
  T z[4];
  TR::normal4(quad, (uint32_t)year, (uint32_t)col, jb.tag0, jb.key0, jb.key1, z);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if constexpr (sizeof(T) == 4)
      e[s] = __builtin_amdgcn_exp2f(fmaf(sig, z[s], -half_var));
    else
      e[s] = exp(sig * z[s] - half_var);
  }
}

// Row r of an accumulator and the quad layout: with kQuadItems (FP32), items
// 4q .. 4q+3 of a lane are 4 consecutive paths; otherwise one Philox call per
// item, whose element path & 3 is used.
template <class TR>
__device__ __forceinline__ void adj_noise_items(const AdjJob& jb, int p0, int trow, int lane,
                                                int r4, int year, int col,
                                                typename TR::T e[4], typename TR::T sig,
                                                typename TR::T half_var) {
  using T = typename TR::T;
  constexpr int TM = TR::TM;
  if constexpr (TR::kQuadItems) {
    const int row0 = trow * TM + TR::acc_row(lane, r4);
    adj_noise4<TR>(jb, (uint32_t)(p0 + row0) >> 2, year, col, e, sig, half_var);
  } else {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const uint32_t path = (uint32_t)(p0 + trow * TM + TR::acc_row(lane, r4 + s));
      T ee[4];
      adj_noise4<TR>(jb, path >> 2, year, col, ee, sig, half_var);
      e[s] = ee[path & 3u];
    }
  }
}

// NT bounds the output tiles per wave (the launch picks the smallest form
// that holds ceil(tiles / waves)): the accumulators and operands of NT tiles
// are allocated whether or not a wave uses them, so a smaller NT leaves
// registers for more resident workgroups.
template <class TR, int NT>
__global__ __launch_bounds__(kMteThreads) void adj_segment_kernel(AdjJob jb) {
  using T = typename TR::T;
  using VecT = typename TR::VecT;
  using AccT = typename TR::AccT;
  constexpr int TM = TR::TM, TN = TR::TN, KV = TR::KV, NACC = TR::NACC;
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

  const RMC_GLOBAL T* cols = gmem(static_cast<const T*>(jb.cols));
  const RMC_GLOBAL T* capKinv = cols + Pp;
  const RMC_GLOBAL T* rate = cols + 2 * Pp;
  const T l2e = sizeof(T) == 4 ? (T)1.4426950408889634 : (T)1;
  const T sig = (T)(jb.sigma * (sizeof(T) == 4 ? 1.4426950408889634 : 1.0));
  const T half_var = (T)(0.5 * jb.sigma * jb.sigma) * l2e;

  RMC_GLOBAL T* ns = gmem(static_cast<T*>(jb.ns));
  RMC_GLOBAL T* np = gmem(static_cast<T*>(jb.np));
  RMC_GLOBAL T* mu = gmem(static_cast<T*>(jb.mu));
  RMC_GLOBAL T* lam = gmem(static_cast<T*>(jb.lam));
  const size_t plane = (size_t)N * Pp;  // one year of staging

  for (int c = tid; c < Pp; c += nthreads) gr[c] = 0.0;
  // Checkpoint of the segment (state entering year y0), which is also the
  // segment's first staged state ns[0].
  const RMC_GLOBAL T* ck = gmem(static_cast<const T*>(jb.ckpt)) + (size_t)(jb.y0 / jb.L) * plane;
  for (int idx = tid; idx < BM * Pp; idx += nthreads) {
    const int r = idx / Pp, c = idx - r * Pp;
    const bool in = p0 + r < N;
    const T v = in ? ck[(size_t)(p0 + r) * Pp + c] : (T)0;
    A[r * LDA + TR::perm(c)] = v;
    if (in) ns[(size_t)(p0 + r) * Pp + c] = v;
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
      case 2:
        if constexpr (NT >= 2)
          year_product<TR, 2>(acc, b0, Bm, A, trow, tcol, LDA, arow, akoff, lane);
        break;
      case 3:
        if constexpr (NT >= 3)
          year_product<TR, 3>(acc, b0, Bm, A, trow, tcol, LDA, arow, akoff, lane);
        break;
      case 4:
        if constexpr (NT >= 4)
          year_product<TR, 4>(acc, b0, Bm, A, trow, tcol, LDA, arow, akoff, lane);
        break;
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

  // Recompute the segment: ns[i] = n_{y0+i}, np[i] = n'_{y0+i+1}; the
  // epilogue of year yr < y1 stores the next state ns[i + 1] from registers.
  for (int yr = jb.y0; yr <= jb.y1; ++yr) {
    const int i = yr - jb.y0;
    const bool stage_next = yr < jb.y1;
    product(Mp);
    __syncthreads();
    {
      // Opaque copies: the epilogue's addresses are formed here, after the
      // product, instead of being hoisted into registers live across it.
      int lanee = lane, p0e = p0;
      const RMC_GLOBAL T* colse = cols;
      RMC_GLOBAL T* npe = np;
      RMC_GLOBAL T* nse = ns;
      asm volatile("" : "+v"(lanee), "+s"(p0e), "+s"(colse), "+s"(npe), "+s"(nse));
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        if (t < my_n) {
          const int col = tcol[t] * TN + TR::acc_col(lanee);
          const T Ki = colse[Pp + col], rr = colse[2 * Pp + col];
#pragma unroll
          for (int r4 = 0; r4 < NACC; r4 += 4) {
            T e[4];
            adj_noise_items<TR>(jb, p0e, trow[t], lanee, r4, yr, col, e, sig, half_var);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
              const int row = trow[t] * TM + TR::acc_row(lanee, r4 + s);
              const uint32_t path = (uint32_t)(p0e + row);
              const T x = acc[t][r4 + s];
              const T nv = x * ((T)1 + rr * ((T)1 - x * Ki)) * e[s];
              if ((int)path < N) {
                npe[i * plane + (size_t)path * Pp + col] = x;
                if (stage_next) nse[(i + 1) * plane + (size_t)path * Pp + col] = nv;
              }
              A[row * LDA + TR::perm(col)] = nv;
            }
          }
        }
      }
    }
    __syncthreads();
  }

  // Top of the segment: mu_{y1+1} from lambda_{y1+1}.  Thread w handles
  // column w % Pp over row group w / Pp (S groups of BM / S rows, S a power
  // of two that fits the threads and the grb scratch and keeps every group a
  // whole number of path quads); the groups' growth-rate sums are then added
  // in group order.
  {
    const int i = jb.y1 - jb.y0;
    const bool last = (jb.y1 == jb.T - 1);
    int S = 1;
    while (S * 2 <= 8 && S * 2 * Pp <= nthreads && S * 2 <= nrb && BM % (S * 2 * 4) == 0) S *= 2;
    const int rows = BM / S;  // a multiple of 4
    for (int w = tid; w < Pp * S; w += nthreads) {
      const int c = w % Pp, grp = w / Pp;
      const T Ki = capKinv[c], rr = rate[c];
      double gsum = 0.0;
      for (int r0 = grp * rows; r0 < (grp + 1) * rows; r0 += 4) {  // p0 is a multiple of 4
        T e[4];
        adj_noise4<TR>(jb, (uint32_t)(p0 + r0) >> 2, jb.y1, c, e, sig, half_var);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int r = r0 + s;
          const uint32_t path = (uint32_t)(p0 + r);
          T m = (T)0;
          if ((int)path < N && c < jb.P) {
            const T l = last ? (T)(jb.dobj / N) : lam[(size_t)path * Pp + c];
            const T x = np[i * plane + (size_t)path * Pp + c];
            m = l * ((T)1 + rr * ((T)1 - (T)2 * x * Ki)) * e[s];
            gsum += (double)(l * x * ((T)1 - x * Ki) * e[s]);
          }
          if ((int)path < N) mu[i * plane + (size_t)path * Pp + c] = m;
          A[r * LDA + TR::perm(c)] = m;
        }
      }
      grb[grp * Pp + c] = gsum;
    }
    __syncthreads();
    for (int c = tid; c < Pp; c += nthreads) {
      double v = 0.0;
      for (int g = 0; g < S; ++g) v += grb[g * Pp + c];
      gr[c] = v;
    }
    __syncthreads();
  }

  // Backward sweep: lambda_yr = mu_{yr+1} M^T, then mu_yr.
  for (int yr = jb.y1; yr >= jb.y0; --yr) {
    product(Mtp);
    __syncthreads();
    const int i = yr - 1 - jb.y0;  // staging slot of n'_yr and mu_yr
    int lanee = lane, p0e = p0;
    const RMC_GLOBAL T* colse = cols;
    RMC_GLOBAL T* npe = np;
    RMC_GLOBAL T* mue = mu;
    RMC_GLOBAL T* lame = lam;
    asm volatile("" : "+v"(lanee), "+s"(p0e), "+s"(colse), "+s"(npe), "+s"(mue), "+s"(lame));
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t < my_n) {
        const int col = tcol[t] * TN + TR::acc_col(lanee);
        const T Ki = colse[Pp + col], rr = colse[2 * Pp + col];
        double gsum = 0.0;
#pragma unroll
        for (int r4 = 0; r4 < NACC; r4 += 4) {
          T e[4] = {(T)0, (T)0, (T)0, (T)0};
          if (yr != jb.y0) adj_noise_items<TR>(jb, p0e, trow[t], lanee, r4, yr - 1, col, e, sig,
                                               half_var);
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int row = trow[t] * TM + TR::acc_row(lanee, r4 + s);
            const uint32_t path = (uint32_t)(p0e + row);
            const T l = acc[t][r4 + s];
            if (yr == jb.y0) {
              if ((int)path < N) lame[(size_t)path * Pp + col] = l;
            } else {
              T m = (T)0;
              if ((int)path < N && col < jb.P) {
                const T x = npe[i * plane + (size_t)path * Pp + col];
                m = l * ((T)1 + rr * ((T)1 - (T)2 * x * Ki)) * e[s];
                gsum += (double)(l * x * ((T)1 - x * Ki) * e[s]);
              }
              if ((int)path < N) mue[i * plane + (size_t)path * Pp + col] = m;
              A[row * LDA + TR::perm(col)] = m;
            }
          }
        }
        if (yr != jb.y0) {
          // Lanes holding the same column combine in a fixed order; one lane
          // per column writes the row block's sum.
#pragma unroll
          for (int off = TN; off < 64; off <<= 1) gsum += __shfl_xor(gsum, off, 64);
          if (lanee < TN) grb[trow[t] * Pp + col] = gsum;
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
  for (int c = tid; c < Pp; c += nthreads) gmem(jb.grad_part)[(size_t)tile * Pp + c] = gr[c];
}

// grad_r[c] += sum over tiles of the segment's partials: one wave per
// column, lane l adding tiles l, l + 64, ... in order, then a fixed butterfly.
__global__ __launch_bounds__(64) void adj_reduce_r_kernel(const double* __restrict__ part,
                                                          int n_tiles, int Pp, int P,
                                                          double* __restrict__ grad_r) {
  const int c = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int t = lane; t < n_tiles; t += 64) s += part[(size_t)t * Pp + c];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) grad_r[c] += s;
}

// dJ/dM += sum_{i, paths} ns[i]^T mu[i] over the segment: C = A^T B with
// A, B of Ktot = n_year * N rows of Pp values.  Split-K: grid (Pp/128)^2
// output blocks x nsplit chunks of K; a 256-thread workgroup forms one
// 128 x 128 block of one chunk (each wave a 64 x 64 quarter), staging
// 16-row slabs of A and B in LDS (double buffered, 16-byte loads) and
// writing the block to its partial slot; adj_reduce_kernel sums the slots in
// order (FP64).  FP32 only (the FP64 checker keeps adj_gradm_f64_kernel).
constexpr int kGmBlock = 128, kGmSlab = 16;

// GATHER: the k-th row is rows[k] (a list of row indices) instead of k.
template <bool GATHER>
__global__ __launch_bounds__(256) void adj_gradm_f32_kernel(const float* __restrict__ A,
                                                            const float* __restrict__ B,
                                                            int Ktot, int Pp, int kchunk,
                                                            double* __restrict__ partial,
                                                            const int* __restrict__ rows) {
  typedef float floatx4 __attribute__((ext_vector_type(4)));
  typedef float floatx16 __attribute__((ext_vector_type(16)));
  __shared__ __attribute__((aligned(16))) float sa[2][kGmSlab][kGmBlock];
  __shared__ __attribute__((aligned(16))) float sb[2][kGmSlab][kGmBlock];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nb = Pp / kGmBlock;
  const int ob = blockIdx.x % (nb * nb), split = blockIdx.x / (nb * nb);
  const int i0 = (ob / nb) * kGmBlock, j0 = (ob % nb) * kGmBlock;
  const int kb = split * kchunk, ke = min(Ktot, kb + kchunk);
  const int wi = (wave >> 1) * 64, wj = (wave & 1) * 64;  // wave's quarter
  floatx16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = floatx16{};
  // Slab loads: 16 rows x 128 floats = 512 float4 per matrix, 2 per thread.
  auto load_slab = [&](int k0, floatx4 (&ra)[2], floatx4 (&rb)[2]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int v = tid + h * 256;           // float4 index in the slab
      const int r = v >> 5, c4 = (v & 31) * 4;
      const int k = k0 + r;
      const bool ok = k < ke;
      if constexpr (GATHER) {
        const size_t row = ok ? (size_t)rows[k] : 0;
        ra[h] = ok ? *reinterpret_cast<const floatx4*>(A + row * Pp + i0 + c4) : floatx4{};
        rb[h] = ok ? *reinterpret_cast<const floatx4*>(B + row * Pp + j0 + c4) : floatx4{};
      } else {
        ra[h] = ok ? *reinterpret_cast<const floatx4*>(A + (size_t)k * Pp + i0 + c4) : floatx4{};
        rb[h] = ok ? *reinterpret_cast<const floatx4*>(B + (size_t)k * Pp + j0 + c4) : floatx4{};
      }
    }
  };
  auto store_slab = [&](int buf, const floatx4 (&ra)[2], const floatx4 (&rb)[2]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int v = tid + h * 256;
      const int r = v >> 5, c4 = (v & 31) * 4;
      *reinterpret_cast<floatx4*>(&sa[buf][r][c4]) = ra[h];
      *reinterpret_cast<floatx4*>(&sb[buf][r][c4]) = rb[h];
    }
  };
  floatx4 ra[2], rb[2];
  load_slab(kb, ra, rb);
  store_slab(0, ra, rb);
  __syncthreads();
  int buf = 0;
  const int kl = lane >> 5, il = lane & 31;
  for (int k0 = kb; k0 < ke; k0 += kGmSlab) {
    const bool more = k0 + kGmSlab < ke;
    if (more) load_slab(k0 + kGmSlab, ra, rb);  // in flight during the MFMAs
#pragma unroll
    for (int kk = 0; kk < kGmSlab; kk += 2) {
      float a[2], b[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        a[x] = sa[buf][kk + kl][wi + x * 32 + il];
        b[x] = sb[buf][kk + kl][wj + x * 32 + il];
      }
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[x], b[y], acc[x][y], 0, 0, 0);
    }
    if (more) store_slab(buf ^ 1, ra, rb);
    __syncthreads();
    buf ^= 1;
  }
  double* out = partial + (size_t)split * Pp * Pp;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = i0 + wi + x * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int col = j0 + wj + y * 32 + (lane & 31);
        out[(size_t)row * Pp + col] = (double)acc[x][y][r];
      }
}

// FP64 checker form: 64 x 64 output blocks, operands read directly.
template <class TR>
__global__ __launch_bounds__(256) void adj_gradm_kernel(const typename TR::T* __restrict__ A,
                                                        const typename TR::T* __restrict__ B,
                                                        int Ktot, int Pp, int kchunk,
                                                        double* __restrict__ partial,
                                                        const int* __restrict__ rows) {
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
    const size_t row = ok ? (rows ? (size_t)rows[kk] : (size_t)kk) : 0;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const T a = ok ? A[row * Pp + ti[t] + al] : (T)0;
      const T b = ok ? B[row * Pp + tj[t] + al] : (T)0;
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
  int k = 0;
  // Eight partials loaded before they are added (in split order, as one at a
  // time would): their loads overlap instead of each waiting on the last.
  for (; k + 8 <= nsplit; k += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = partial[(size_t)(k + u) * n + e];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; k < nsplit; ++k) s += partial[(size_t)k * n + e];
  grad[e] += s;
}

size_t adj_lds_bytes(bool fp64, int BM, int Pp) {
  const size_t a = fp64 ? align16((size_t)BM * (Pp + TraitsF64::PAD) * 8)
                        : align16((size_t)BM * (Pp + TraitsF32::PAD) * 4);
  const int rows = fp64 ? TraitsF64::TM : TraitsF32::TM;
  return a + (size_t)Pp * 8 * (1 + BM / rows);
}

namespace {

// The segment kernel form for a job: FP64 at its traits' tile bound; FP32 at
// the smallest of 1, 2 and 4 tiles per wave that holds the job's tiles.
const void* adj_segment_fn(bool fp64, int BM, int Pp) {
  if (fp64) return reinterpret_cast<const void*>(&adj_segment_kernel<TraitsF64, TraitsF64::NT>);
  const int nwaves = kMteThreads / 64;
  const int per_wave = ((BM / TraitsF32::TM) * (Pp / TraitsF32::TN) + nwaves - 1) / nwaves;
  if (per_wave <= 1) return reinterpret_cast<const void*>(&adj_segment_kernel<TraitsF32, 1>);
  if (per_wave <= 2) return reinterpret_cast<const void*>(&adj_segment_kernel<TraitsF32, 2>);
  return reinterpret_cast<const void*>(&adj_segment_kernel<TraitsF32, TraitsF32::NT>);
}

}  // namespace

hipError_t adj_set_lds_limit(bool fp64, int BM, int Pp, size_t bytes) {
  return hipFuncSetAttribute(adj_segment_fn(fp64, BM, Pp),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t launch_adj_segment(bool fp64, const AdjJob& jb, int n_tiles, size_t lds,
                              hipStream_t s) {
  const int nwaves = kMteThreads / 64;
  const int ntiles_f32 = (jb.BM / TraitsF32::TM) * (jb.Pp / TraitsF32::TN);
  const int ntiles_f64 = (jb.BM / TraitsF64::TM) * (jb.Pp / TraitsF64::TN);
  if ((fp64 ? ntiles_f64 : ntiles_f32) > (fp64 ? TraitsF64::NT : TraitsF32::NT) * nwaves)
    return hipErrorInvalidValue;  // more tiles than the waves can hold
  void* args[] = {const_cast<AdjJob*>(&jb)};
  hipError_t e = hipLaunchKernel(adj_segment_fn(fp64, jb.BM, jb.Pp), dim3(n_tiles),
                                 dim3(kMteThreads), args, lds, s);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(adj_reduce_r_kernel, dim3(jb.P), dim3(64), 0, s, jb.grad_part, n_tiles,
                     jb.Pp, jb.P, jb.grad_r);
  return hipGetLastError();
}

hipError_t launch_adj_gradm(bool fp64, const void* A, const void* B, int Ktot, int Pp,
                            int nsplit, double* partial, double* grad, hipStream_t s,
                            const int* rows) {
  if (fp64) {
    const int kchunk = ((Ktot + nsplit - 1) / nsplit + 3) / 4 * 4;
    const int nb = Pp / 64;
    hipLaunchKernelGGL(adj_gradm_kernel<TraitsF64>, dim3(nb * nb * nsplit), dim3(256), 0, s,
                       static_cast<const double*>(A), static_cast<const double*>(B), Ktot, Pp,
                       kchunk, partial, rows);
  } else if (Pp % kGmBlock == 0) {
    const int kchunk = ((Ktot + nsplit - 1) / nsplit + kGmSlab - 1) / kGmSlab * kGmSlab;
    const int nb = Pp / kGmBlock;
    if (rows)
      hipLaunchKernelGGL(adj_gradm_f32_kernel<true>, dim3(nb * nb * nsplit), dim3(256), 0, s,
                         static_cast<const float*>(A), static_cast<const float*>(B), Ktot, Pp,
                         kchunk, partial, rows);
    else
      hipLaunchKernelGGL(adj_gradm_f32_kernel<false>, dim3(nb * nb * nsplit), dim3(256), 0, s,
                         static_cast<const float*>(A), static_cast<const float*>(B), Ktot, Pp,
                         kchunk, partial, rows);
  } else {
    const int kchunk = ((Ktot + nsplit - 1) / nsplit + 3) / 4 * 4;
    const int nb = Pp / 64;
    hipLaunchKernelGGL(adj_gradm_kernel<TraitsF32>, dim3(nb * nb * nsplit), dim3(256), 0, s,
                       static_cast<const float*>(A), static_cast<const float*>(B), Ktot, Pp,
                       kchunk, partial, rows);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int n = Pp * Pp;
  hipLaunchKernelGGL(adj_reduce_kernel, dim3((n + 255) / 256), dim3(256), 0, s, partial, nsplit,
                     n, grad);
  return hipGetLastError();
}

}  // namespace roadmc
