// MTE simulation beyond one workgroup's LDS tile (more than 1024 patches),
// for gfx950.
//
// The tiled kernels (mte_kernels.hip) keep a tile of paths' populations in
// LDS for all years, which caps P at 1024 (32 rows x 1024 FP32).  Here the
// state of every path lives in HBM, path-major [N][Pp] with Pp = P rounded
// up to 128, in two buffers, and each year is one launch: the product
// S_in M as a 128 x 128-blocked GEMM on v_mfma_f32_32x32x2_f32 (contraction
// slabs of 16 staged in LDS, double-buffered), then per element the growth
// law with the same Philox draws as the tiled kernels (one normal quad per
// four paths, patch and year), written to S_out.  M (the movement matrix
// T o S) is stored destination-major so that both operands stage the same
// way: mt[j Pp + k] = M[k][j].  The per-path scalars of the REF_GPU law
// (SimulateGPU.cu:686-701) are a separate per-year pass before the product.
// The tallies of the final state use the tiled kernels' partial layout, so
// mte_reduce_kernel and the host result code are shared.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gmem.h"
#include "mte_common.h"
#include "mte_wide.h"

namespace roadmc {

namespace {

constexpr int kWB = kWideBlock;  // rows (paths) and columns (patches) of a block
constexpr int kWK = 16;          // contraction slab
constexpr int kWLd = kWK + 4;    // LDS row: 80 bytes, 16-byte aligned, banks spread

// Slab position of contraction index k (< 16): k = 8 g + 2 s + h goes to
// 8 g + 4 h + s, so a lane's four MFMA steps s of k-group g (lane half h)
// are one 16-byte read.  Eight consecutive k of a group land as the even
// ones, then the odd ones.
__device__ __forceinline__ void wide_stage(float* row, int g, const floatx4& lo, const floatx4& hi) {
  // lo = k 8g..8g+3, hi = k 8g+4..8g+7
  *reinterpret_cast<floatx4*>(row + 8 * g) = floatx4{lo[0], lo[2], hi[0], hi[2]};
  *reinterpret_cast<floatx4*>(row + 8 * g + 4) = floatx4{lo[1], lo[3], hi[1], hi[3]};
}

}  // namespace

// LDS of one block's product: two contraction slabs of each operand.
struct WideSlabs {
  float a[2][kWB][kWLd];
  float b[2][kWB][kWLd];
};

// The 128 x 128 block (paths row0.., patches col0..) of S M with the state
// S path-major [N][Pp] and M destination-major (mt[j Pp + k] = M[k][j]), on
// v_mfma_f32_32x32x2_f32: wave w = 2 wr + wc forms rows wr 64 .. +64 and
// columns wc 64 .. +64 as acc[i][j], the 32 x 32 tile (i, j) of its quarter.
// Item r of acc[i][j] in lane l = 32 h + l32 is row 8 (r / 4) + 4 h + r % 4
// of the tile, column l32.
__device__ __forceinline__ void wide_product(const RMC_GLOBAL float* S, const RMC_GLOBAL float* Mt,
                                             int Pp, int N, int row0, int col0, WideSlabs& sl,
                                             floatx16 (&acc)[2][2]) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1, h = lane >> 5, l32 = lane & 31;
  // Staging: thread t moves row t >> 1, k-group t & 1 (8 floats) of both slabs.
  const int sr = tid >> 1, sg = tid & 1;
  const bool a_ok = row0 + sr < N;
  const RMC_GLOBAL float* a_src = S + (size_t)(a_ok ? row0 + sr : 0) * Pp + 8 * sg;
  const RMC_GLOBAL float* b_src = Mt + (size_t)(col0 + sr) * Pp + 8 * sg;
  floatx4 ra0, ra1, rb0, rb1;
  auto fetch = [&](int k0) {
    const floatx4 z = floatx4{0.f, 0.f, 0.f, 0.f};
    ra0 = a_ok ? *reinterpret_cast<const RMC_GLOBAL floatx4*>(a_src + k0) : z;
    ra1 = a_ok ? *reinterpret_cast<const RMC_GLOBAL floatx4*>(a_src + k0 + 4) : z;
    rb0 = *reinterpret_cast<const RMC_GLOBAL floatx4*>(b_src + k0);
    rb1 = *reinterpret_cast<const RMC_GLOBAL floatx4*>(b_src + k0 + 4);
  };
  auto stage = [&](int buf) {
    wide_stage(&sl.a[buf][sr][0], sg, ra0, ra1);
    wide_stage(&sl.b[buf][sr][0], sg, rb0, rb1);
  };
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx16{};
  const int nslab = Pp / kWK;
  fetch(0);
  stage(0);
  __syncthreads();
  for (int k = 0; k < nslab; ++k) {
    const int buf = k & 1;
    if (k + 1 < nslab) fetch((k + 1) * kWK);
#pragma unroll
    for (int g = 0; g < kWK / 8; ++g) {
      floatx4 av[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        av[i] = *reinterpret_cast<const floatx4*>(&sl.a[buf][wr * 64 + i * 32 + l32][8 * g + 4 * h]);
        bv[i] = *reinterpret_cast<const floatx4*>(&sl.b[buf][wc * 64 + i * 32 + l32][8 * g + 4 * h]);
      }
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i][s], bv[j][s], acc[i][j], 0, 0,
                                                              0);
    }
    if (k + 1 < nslab) {
      stage(buf ^ 1);
      __syncthreads();
    }
  }
}

template <int LAW>
__global__ __launch_bounds__(256) void mte_wide_year_kernel(MteWideArgs a) {
  __shared__ __attribute__((aligned(16))) WideSlabs sl;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1, h = lane >> 5, l32 = lane & 31;
  const int row0 = blockIdx.x * kWB, col0 = blockIdx.y * kWB;
  const int Pp = a.Pp, N = a.N;
  floatx16 acc[2][2];
  wide_product(gmem(a.s_in), gmem(a.mt), Pp, N, row0, col0, sl, acc);

  // Growth and write-back.  Item r of a lane's accumulator is row
  // 8 (r / 4) + 4 h + r % 4 of its 32 x 32 block, column l32; items
  // 4q .. 4q + 3 are four consecutive paths, one Philox quad.
  const RMC_GLOBAL float* cols = gmem(a.cols);
  RMC_GLOBAL float* out = gmem(a.s_out);
  const float ga_scale = a.keep_pre ? 0.f : 1.f;
  const float sig_y = a.keep_pre ? 0.f : a.sig, hv_y = a.keep_pre ? 0.f : a.half_var;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int col = col0 + wc * 64 + j * 32 + l32;
    const float K = cols[col], Ki = cols[Pp + col], rr = cols[2 * Pp + col];
    const float ga = 1.f + rr * ga_scale, gb = rr * ga_scale * Ki;
    const bool live = col < a.P;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int n0 = row0 + wr * 64 + i * 32 + 8 * q + 4 * h;
        float z[4];
        rmc_normal4_f((uint32_t)n0 >> 2, (uint32_t)a.year, (uint32_t)col, a.tag0, a.key0, a.key1,
                      z);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int n = n0 + s;
          if (n >= N) continue;
          const float x = acc[i][j][4 * q + s];
          float v;
          if constexpr (LAW == kLawLognormal) {
            const float e = __builtin_amdgcn_exp2f(fmaf(sig_y, z[s], -hv_y));
            v = x * fmaf(-gb, x, ga) * e;
          } else {
            float g = 0.f;
            if constexpr (LAW == kLawRefGpu) g = gmem(a.grm)[n];
            v = a.keep_pre ? x
                           : grow<float, LAW>(x, z[s], K, Ki, rr, a.sig, a.half_var, a.dt, a.mu,
                                              a.sd, g, a.lv);
          }
          out[(size_t)n * Pp + col] = live ? v : 0.f;
        }
      }
    }
  }
}

// Sum over the 32 lanes of a half wave (one accumulator column each), in a
// fixed butterfly order.
__device__ __forceinline__ float wide_half_sum(float v) {
#pragma unroll
  for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__global__ __launch_bounds__(256) void rovs_wide_year_kernel(RovsWideArgs a) {
  __shared__ __attribute__((aligned(16))) WideSlabs sl;
  __shared__ float rsum[2][kWB], gsum[2][kWB];
  __shared__ int ctl[kWB];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1, h = lane >> 5, l32 = lane & 31;
  const int row0 = blockIdx.x * kWB, col0 = blockIdx.y * kWB, c = blockIdx.z;
  const int Pp = a.Pp, N = a.N;
  if (tid < kWB) ctl[tid] = row0 + tid < N ? gmem(a.controls)[row0 + tid] : -1;
  floatx16 acc[2][2];
  wide_product(gmem(a.s_in), gmem(a.mt) + (size_t)c * Pp * Pp, Pp, N, row0, col0, sl, acc);
  // Growth of the paths that drew this control: g = grMean + z lv per patch,
  // n = n' (1 + g (K - n') / K / 100); every block's row sums of the product
  // and of the grown values.
  const RMC_GLOBAL float* cols = gmem(a.cols);
  const RMC_GLOBAL float* grm = gmem(a.grm);
  RMC_GLOBAL float* out = gmem(a.s_out);
  float rs[2][16], gs[2][16];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 16; ++r) rs[i][r] = gs[i][r] = 0.f;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int col = col0 + wc * 64 + j * 32 + l32;
    const float K = cols[col], Ki = cols[Pp + col];
    const bool live = col < a.P;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rl = wr * 64 + i * 32 + 8 * q + 4 * h;  // block row of item 4q
        float z[4];
        rmc_normal4_f((uint32_t)(row0 + rl) >> 2, (uint32_t)a.year, (uint32_t)col, a.tag0, a.key0,
                      a.key1, z);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int n = row0 + rl + s;
          const float x = acc[i][j][4 * q + s];
          rs[i][4 * q + s] += x;
          if (ctl[rl + s] == c) {
            const float g = grm[n] + z[s] * a.lv;
            const float v = live ? x * (1.f + g * (K - x) * Ki * 0.01f) : 0.f;
            gs[i][4 * q + s] += v;
            out[(size_t)n * Pp + col] = v;
          }
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const float vr = wide_half_sum(rs[i][r]), vg = wide_half_sum(gs[i][r]);
      const int rl = wr * 64 + i * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
      if (l32 == 0) {
        rsum[wc][rl] = vr;
        gsum[wc][rl] = vg;
      }
    }
  __syncthreads();
  if (tid < kWB && row0 + tid < N) {
    const int n = row0 + tid;
    gmem(a.rs_part)[((size_t)blockIdx.y * a.C + c) * N + n] = rsum[0][tid] + rsum[1][tid];
    if (ctl[tid] == c) gmem(a.tot_part)[(size_t)blockIdx.y * N + n] = gsum[0][tid] + gsum[1][tid];
  }
}

__global__ __launch_bounds__(256) void rovs_wide_tally_kernel(const float* __restrict__ rs_part,
                                                              const float* __restrict__ tot_part,
                                                              int nblk, int C, int N,
                                                              const float* __restrict__ tot_prev,
                                                              float* __restrict__ tot_next,
                                                              float* __restrict__ aars) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  const float tot = gmem(tot_prev)[n];
  for (int c = 0; c < C; ++c) {
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += gmem(rs_part)[((size_t)b * C + c) * N + n];
    gmem(aars)[(size_t)c * N + n] = s / tot;
  }
  float t = 0.f;
  for (int b = 0; b < nblk; ++b) t += gmem(tot_part)[(size_t)b * N + n];
  gmem(tot_next)[n] = t;
}

__global__ __launch_bounds__(256) void mte_wide_init_kernel(MteWideArgs a, float g0, float* s0) {
  const RMC_GLOBAL float* n0 = gmem(a.cols) + 3 * (size_t)a.Pp;
  const size_t total = (size_t)a.N * a.Pp;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (size_t)gridDim.x * 256) {
    const int c = (int)(e % a.Pp);
    gmem(s0)[e] = c < a.P ? n0[c] : 0.f;
    if (c == 0 && a.law == kLawRefGpu) gmem(a.grm)[e / a.Pp] = g0;
  }
}

// The growth-rate SDE of every path for year a.year (the per-row step of the
// tiled kernels).
__global__ __launch_bounds__(256) void mte_wide_grm_kernel(MteWideArgs a) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= a.N) return;
  float zb, zj, uj;
  path_scalars<float>((uint32_t)n, (uint32_t)a.year, a.tag1, a.key0, a.key1, &zb, &zj, &uj);
  RMC_GLOBAL float* grm = gmem(a.grm);
  const float g = grm[n];
  const float jump = (uj < a.jprob) ? 1.f : 0.f;
  const float brown = zb * a.sd;
  const float jsz = zj * a.jmp * a.jmp - a.jmp * a.jmp / 2.f;
  grm[n] = g + a.rev * (a.mu - g) * a.dt + g * brown + (__expf(jsz) - 1.f) * g * jump;
}

// One tile of kWB paths: each wave sums 32 paths' rows (lanes over columns,
// then a fixed butterfly), the per-path tallies are added in path order, the
// patch sums in row order.
__global__ __launch_bounds__(256) void mte_wide_tally_kernel(const float* __restrict__ s, int N,
                                                             int Pp, double thr,
                                                             double* __restrict__ partial,
                                                             double* __restrict__ end_pops) {
  __shared__ double red[4][3];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int row0 = blockIdx.x * kWB, nvalid = min(kWB, N - row0);
  const RMC_GLOBAL float* S = gmem(s);
  double cnt = 0.0, sum = 0.0, sq = 0.0;
  for (int r = w * 32; r < w * 32 + 32 && r < nvalid; ++r) {
    const RMC_GLOBAL float* row = S + (size_t)(row0 + r) * Pp;
    double t = 0.0;
    for (int c = lane; c < Pp; c += 64) t += (double)row[c];
    t = wave_sum(t);
    if (end_pops && lane == 0) gmem(end_pops)[row0 + r] = t;
    cnt += t >= thr ? 1.0 : 0.0;
    sum += t;
    sq += t * t;
  }
  if (lane == 0) {
    red[w][0] = cnt;
    red[w][1] = sum;
    red[w][2] = sq;
  }
  __syncthreads();
  RMC_GLOBAL double* part = gmem(partial) + (size_t)blockIdx.x * (3 + Pp);
  if (tid < 3) part[tid] = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
  for (int c = tid; c < Pp; c += 256) {
    double v = 0.0;
    for (int r = 0; r < nvalid; ++r) v += (double)S[(size_t)(row0 + r) * Pp + c];
    part[3 + c] = v;
  }
}

__global__ __launch_bounds__(256) void mte_wide_pack_kernel(const double* __restrict__ trans,
                                                            const double* __restrict__ surv,
                                                            int P, int Pp, double tol,
                                                            float* __restrict__ mt) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)Pp * Pp) return;
  const int j = (int)(e / Pp), k = (int)(e - (size_t)j * Pp);
  double v = 0.0;
  if (k < P && j < P) {
    const size_t cm = (size_t)j * P + k;
    v = __dmul_rn(gmem(trans)[cm], surv ? gmem(surv)[cm] : 1.0);
    if (tol > 0.0 && !((float)v > (float)tol)) v = 0.0;
  }
  gmem(mt)[e] = (float)v;
}

hipError_t launch_mte_wide_init(const MteWideArgs& a, float g0, float* s0, hipStream_t s) {
  const size_t total = (size_t)a.N * a.Pp;
  const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(mte_wide_init_kernel, dim3(blocks), dim3(256), 0, s, a, g0, s0);
  return hipGetLastError();
}

hipError_t launch_mte_wide_grm(const MteWideArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(mte_wide_grm_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_mte_wide_year(const MteWideArgs& a, hipStream_t s) {
  if (a.Pp % kWB != 0) return hipErrorInvalidValue;
  const dim3 grid((a.N + kWB - 1) / kWB, a.Pp / kWB);
  switch (a.law) {
    case kLawRefCpu:
      hipLaunchKernelGGL(mte_wide_year_kernel<kLawRefCpu>, grid, dim3(256), 0, s, a);
      break;
    case kLawRefGpu:
      hipLaunchKernelGGL(mte_wide_year_kernel<kLawRefGpu>, grid, dim3(256), 0, s, a);
      break;
    default:
      hipLaunchKernelGGL(mte_wide_year_kernel<kLawLognormal>, grid, dim3(256), 0, s, a);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_mte_wide_tally(const float* s, int N, int Pp, double thr, double* partial,
                                 double* end_pops, hipStream_t st) {
  hipLaunchKernelGGL(mte_wide_tally_kernel, dim3((N + kWB - 1) / kWB), dim3(256), 0, st, s, N, Pp,
                     thr, partial, end_pops);
  return hipGetLastError();
}

hipError_t launch_rovs_wide_year(const RovsWideArgs& a, hipStream_t s) {
  if (a.Pp % kWB != 0 || a.C < 1) return hipErrorInvalidValue;
  const dim3 grid((a.N + kWB - 1) / kWB, a.Pp / kWB, a.C);
  hipLaunchKernelGGL(rovs_wide_year_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_rovs_wide_tally(const float* rs_part, const float* tot_part, int nblk, int C,
                                  int N, const float* tot_prev, float* tot_next, float* aars,
                                  hipStream_t s) {
  hipLaunchKernelGGL(rovs_wide_tally_kernel, dim3((N + 255) / 256), dim3(256), 0, s, rs_part,
                     tot_part, nblk, C, N, tot_prev, tot_next, aars);
  return hipGetLastError();
}

hipError_t launch_mte_wide_pack(const double* trans, const double* surv, int P, int Pp,
                                double tol, float* mt, hipStream_t s) {
  const size_t n = (size_t)Pp * Pp;
  hipLaunchKernelGGL(mte_wide_pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     trans, surv, P, Pp, tol, mt);
  return hipGetLastError();
}

}  // namespace roadmc
