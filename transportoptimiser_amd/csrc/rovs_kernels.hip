// Species-state ROV (the reference's ALGO4) for gfx950.
//
// Reference: forwardPathKernel (include/gpuCode/SimulateGPU.cu:789-956)
// moves each path's species populations under a randomly drawn control per
// year and records the surviving fraction (AAR) under every control;
// computePathStates / allocateXYRegressionData / computeStateMinMax
// (:959-1148) form the state (adjusted populations under the highest flow,
// unit profit) and the per-control regression data; the ALGO4 loop
// (:4401-4998) fits a kNN local-linear grid per control and step and
// backwardInduction (:2333-2784) decides by interpolating it;
// firstPeriodInduction (:2210-2330) values time 0.  There the forward pass is
// one thread per path walking CSC matrices (every control's product computed
// serially), and every step makes two <<<1,1>>> kernels and a host kNN round
// trip.
//
// Here:
//   - rovs_forward_kernel: 32 paths x all patches per workgroup (one wave
//     per two 32-column tiles, at least four; P <= 512), the state
//     tile in LDS; per year, one FP32 MFMA product per control
//     (v_mfma_f32_32x32x2_f32 on the packed trans o surv_c), row sums of each
//     product for the AARs (lane shuffles, then a fixed-order sum over the
//     waves), each path's own control kept by a register select, then the
//     growth-rate SDE and the logistic growth with the reference's /100 in
//     the epilogue, noise from Philox in registers.  Under controls fixed by
//     a policy (the smoothed gradient), a second form takes the AARs as dots
//     of the year-start tile with each control's row sums and forms only the
//     products of the controls the tile's paths take;
//   - per backward step: one kernel forms the states and counts each
//     control's included paths per block, one scans the counts, one
//     scatters the samples in path order (so that kNN ties resolve as in a
//     serial pass), one reduces the state ranges, one scales them; the grid
//     fits are the knn_fit_kernel of the grid surrogate, and one kernel
//     decides every path by multilinear interpolation of the grids.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mte_common.h"
#include "rng.h"
#include "rovs_kernels.h"

namespace roadmc {

using TF = TraitsF32;

// Waves of a forward workgroup: at most kRovsNT column tiles each, at least 4.
static int rovs_waves(int Pp) {
  const int ntc = Pp / TF::TN;
  return std::max(4, (ntc + kRovsNT - 1) / kRovsNT);
}

// Population tile of 32 rows in LDS, through the arithmetic of the products:
// FP32 (TraitsF32: floats, k permuted within groups of 8) or split bf16
// (TraitsB3: a bf16 hi array then a bf16 lo array, natural k order).  Both
// take 4 bytes per entry; LDA is Pp + PAD entries.  Positions c are stored
// column positions (TR::perm of the patch).
template <class TR>
struct RovsTile {
  float* A;
  int LDA;
  __device__ uint16_t* hi() const { return reinterpret_cast<uint16_t*>(A); }
  __device__ void put(int r, int c, float v) const {
    if constexpr (TR::kSplit) {
      const uint32_t hl = TR::split(v);
      hi()[r * LDA + c] = (uint16_t)hl;
      hi()[kRovsRows * LDA + r * LDA + c] = (uint16_t)(hl >> 16);
    } else {
      A[r * LDA + c] = v;
    }
  }
  __device__ float get(int r, int c) const {
    if constexpr (TR::kSplit) {
      return __builtin_bit_cast(float, (uint32_t)hi()[r * LDA + c] << 16) +
             __builtin_bit_cast(float, (uint32_t)hi()[kRovsRows * LDA + r * LDA + c] << 16);
    } else {
      return A[r * LDA + c];
    }
  }
  // The stored bits of an entry (checkpoints): the float, or hi | lo << 16.
  __device__ uint32_t raw(int r, int c) const {
    if constexpr (TR::kSplit)
      return (uint32_t)hi()[r * LDA + c] | ((uint32_t)hi()[kRovsRows * LDA + r * LDA + c] << 16);
    else
      return __builtin_bit_cast(uint32_t, A[r * LDA + c]);
  }
  __device__ void put_raw(int r, int c, uint32_t u) const {
    if constexpr (TR::kSplit) {
      hi()[r * LDA + c] = (uint16_t)u;
      hi()[kRovsRows * LDA + r * LDA + c] = (uint16_t)(u >> 16);
    } else {
      A[r * LDA + c] = __builtin_bit_cast(float, u);
    }
  }
};

// Waves per SIMD the forward and recompute kernels are compiled for: three
// for FP32 products (167 of 168 VGPRs, a few spilled in the epilogue), two
// for split bf16, which would spill its product loop at three.
template <class TR>
constexpr int rovs_min_waves() {
  return TR::kSplit ? 2 : 3;
}

static int rovs_pad(bool split) { return split ? TraitsB3::PAD : TF::PAD; }

size_t rovs_forward_lds(int Pp, int C, bool split, bool dots) {
  const int LDA = Pp + rovs_pad(split);
  return (size_t)kRovsRows * LDA * sizeof(float)      // state tile
         + kRovsRows * sizeof(float)                  // growth-rate means
         + (size_t)rovs_waves(Pp) * kRovsRows * sizeof(float)  // per-wave row sums
         + kRovsRows * sizeof(float)                  // row totals
         + (size_t)C * kRovsRows * sizeof(float)      // AARs
         + kRovsRows * sizeof(int)                    // controls
         // the dots' partial sums: 2 strides of columns per wave
         + (dots ? (size_t)2 * rovs_waves(Pp) * C * kRovsRows * sizeof(float) : 0);
}

// Sum over the 32 columns held by the 32 lanes of a half-wave.
__device__ __forceinline__ float half_sum(float v) {
#pragma unroll
  for (int off = 1; off < 32; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <class TR, int NT>
__device__ __forceinline__ void rovs_product(floatx16* acc, const PackedB& B, const float* A,
                                             const int* tcol, int LDA, int arow, int akoff,
                                             int lane) {
  int trow[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) trow[t] = 0;
  if constexpr (TR::kSplit) {
    year_product_b3<NT, true>(acc, B, reinterpret_cast<const uint16_t*>(A), kRovsRows * LDA,
                              trow, tcol, LDA, arow, akoff, lane);
  } else {
    floatx4 b0[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) b0[t] = load_b<floatx4>(B, b_voff(B, tcol[t], lane), 0);
    year_product<TF, NT, true>(acc, b0, B, A, trow, tcol, LDA, arow, akoff, lane);
  }
}

// DOTS: the surviving fractions from dots with a.rvec and products only for
// the controls the tile's paths take (a.rvec given); otherwise from the row
// sums of every control's product.
template <class TR, bool DOTS>
__global__ __launch_bounds__(kRovsMaxThreads)
__attribute__((amdgpu_waves_per_eu(rovs_min_waves<TR>()))) void rovs_forward_kernel(
    RovsForwardArgs a) {
  extern __shared__ char smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nthreads = blockDim.x, NW = nthreads >> 6;
  const int Pp = a.Pp, LDA = Pp + TR::PAD, C = a.C;
  const int p0 = blockIdx.x * kRovsRows;
  float* A = reinterpret_cast<float*>(smem);
  float* grm = A + kRovsRows * LDA;
  float* rsum = grm + kRovsRows;    // [NW][32]
  float* tot = rsum + NW * kRovsRows;
  float* aar = tot + kRovsRows;     // [C][32]
  int* ctl = reinterpret_cast<int*>(aar + C * kRovsRows);
  const int nq = nthreads / kRovsRows;                   // column strides of the dots
  float* spart = reinterpret_cast<float*>(ctl + kRovsRows);  // [nq][C][32] (DOTS)
  const float* K = a.cols;
  const float* Kinv = a.cols + Pp;
  const float* init = a.cols + 2 * Pp;
  const float mu = (float)a.params[1], sd = (float)a.params[2], rev = (float)a.params[4];
  const float jmp = (float)a.params[5], jprob = (float)a.params[6], lv = (float)a.params[7];
  const float dt = (float)a.dt;

  const RovsTile<TR> tile{A, LDA};
  for (int idx = tid; idx < kRovsRows * Pp; idx += nthreads) {
    const int r = idx / Pp, c = idx - r * Pp;
    tile.put(r, TR::perm(c), init[c]);
  }
  if (tid < kRovsRows) {
    grm[tid] = (float)a.params[0];
    tot[tid] = a.init_total;
    if (p0 + tid < a.N) a.total_pops[p0 + tid] = a.init_total;
  }
  // Column tiles of this wave: a contiguous run of at most kRovsNT tiles.
  const int ntc = Pp / TF::TN;
  const int base = ntc / NW, extra = ntc % NW;
  const int my_n = base + (wave < extra ? 1 : 0);
  const int my_t0 = wave * base + (wave < extra ? wave : extra);
  int tcol[kRovsNT];
#pragma unroll
  for (int t = 0; t < kRovsNT; ++t) tcol[t] = my_t0 + t;
  const int arow = TR::a_row(lane), akoff = (TR::kSplit ? 8 : TR::KV) * TR::k_lane(lane);
  const int nkg = a.Pk >> 3;
  const size_t mbytes = (size_t)Pp * a.Pk * sizeof(float);
  __syncthreads();

  for (int year = 0; year < a.T; ++year) {
    // Opaque per-year copies of loop-invariant scalars: the addresses derived
    // from them are recomputed each year instead of being hoisted out of the
    // year loop into registers that would spill.
    int LDAy = LDA, p0y = p0, laney = lane;
    asm volatile("" : "+s"(LDAy), "+s"(p0y), "+v"(laney));
    if (tid < kRovsRows)
      ctl[tid] = (p0 + tid < a.N) ? a.controls[(size_t)year * a.N + p0 + tid] : 0;
    // Checkpoint of the tile at the start of the year (read-only here; the
    // epilogue's writes come after the barriers of the control loop).
    if (a.ckpt && year % a.ck == 0) {
      uint32_t* dst = reinterpret_cast<uint32_t*>(a.ckpt) + (size_t)(year / a.ck) * a.N * Pp;
      const RovsTile<TR> ty{A, LDAy};
      for (int idx = tid; idx < kRovsRows * Pp; idx += nthreads) {
        const int r = idx / Pp, c = idx - r * Pp;
        if (p0y + r < a.N) dst[(size_t)(p0y + r) * Pp + c] = ty.raw(r, c);
      }
    }
    __syncthreads();
    unsigned used = ~0u;
    if constexpr (DOTS) {
      // The surviving fractions' numerators as dots of the year-start tile
      // with each control's row sums (a stride of columns per thread, summed
      // over the strides in order after the products), and the controls some
      // path of the tile takes (the same mask in every wave).
      const int r = tid & (kRovsRows - 1), q = tid / kRovsRows;
      const RovsTile<TR> ts{A, LDAy};
      float accs[kRovsMaxC];
#pragma unroll
      for (int c = 0; c < kRovsMaxC; ++c) accs[c] = 0.f;
      for (int i = q; i < Pp; i += nq) {
        const float v = ts.get(r, i);
#pragma unroll
        for (int c = 0; c < kRovsMaxC; ++c)
          if (c < C) accs[c] += v * a.rvec[(size_t)c * Pp + i];
      }
#pragma unroll
      for (int c = 0; c < kRovsMaxC; ++c)
        if (c < C) spart[(q * C + c) * kRovsRows + r] = accs[c];
      used = 0u;
      const int cl = laney < kRovsRows ? ctl[laney] : -1;
      for (int c = 0; c < C; ++c)
        if (__ballot(cl == c) != 0ull) used |= 1u << c;
    }
    floatx16 sel[kRovsNT], acc[kRovsNT];
#pragma unroll
    for (int t = 0; t < kRovsNT; ++t) sel[t] = floatx16{};
    for (int c = 0; c < C; ++c) {
      if (!((used >> c) & 1u)) continue;
      const PackedB B = packed_b(a.mpack + (size_t)c * Pp * a.Pk, nkg, mbytes);
      if (my_n == 1)
        rovs_product<TR, 1>(acc, B, A, tcol, LDAy, arow, akoff, laney);
      else if (my_n == 2)
        rovs_product<TR, 2>(acc, B, A, tcol, LDAy, arow, akoff, laney);
      if constexpr (DOTS) {
        // The select of each path's own control only.
#pragma unroll
        for (int r = 0; r < TF::NACC; ++r)
          if (ctl[TF::acc_row(laney, r)] == c) {
#pragma unroll
            for (int t = 0; t < kRovsNT; ++t)
              if (t < my_n) sel[t][r] = acc[t][r];
          }
        continue;
      }
      // Row sums of this control's product over the wave's columns, and the
      // select of each path's own control.
#pragma unroll
      for (int r = 0; r < TF::NACC; ++r) {
        float v = 0.f;
        const int row = TF::acc_row(laney, r);
        const bool mine = ctl[row] == c;
#pragma unroll
        for (int t = 0; t < kRovsNT; ++t)
          if (t < my_n) {
            v += acc[t][r];
            if (mine) sel[t][r] = acc[t][r];
          }
        v = half_sum(v);
        if ((laney & 31) == 0) rsum[wave * kRovsRows + row] = v;
      }
      __syncthreads();
      if (tid < kRovsRows) {
        float s = 0.f;
        for (int w = 0; w < NW; ++w) s += rsum[w * kRovsRows + tid];
        aar[c * kRovsRows + tid] = s / tot[tid];
      }
      __syncthreads();
    }
    // Growth-rate SDE per path (SimulateGPU.cu:839-851), jump probability
    // from its own field.
    if (tid < kRovsRows) {
      float zb, zj, uj;
      path_scalars<float>((uint32_t)(p0 + tid), (uint32_t)year, a.tag1, a.key0, a.key1, &zb,
                          &zj, &uj);
      const float g = grm[tid];
      const float jump = (uj < jprob) ? 1.f : 0.f;
      const float jsz = zj * jmp * jmp - jmp * jmp / 2.f;
      grm[tid] = g + rev * (mu - g) * dt + g * (zb * sd) + (__expf(jsz) - 1.f) * g * jump;
      if (a.grm_out && p0 + tid < a.N) a.grm_out[(size_t)year * a.N + p0 + tid] = grm[tid];
    }
    __syncthreads();  // (DOTS: also every wave's products have read the tile)
    if constexpr (DOTS) {
      if (tid < C * kRovsRows) {
        const int c = tid / kRovsRows, r = tid - c * kRovsRows;
        float sm = 0.f;
        for (int qq = 0; qq < nq; ++qq) sm += spart[(qq * C + c) * kRovsRows + r];
        aar[c * kRovsRows + r] = sm / tot[r];
      }
    }
    // Growth (SimulateGPU.cu:902-916): g = grMean + z lv per patch,
    // n = n' (1 + g (K - n') / K / 100), written back to the tile.
    int LDAe = LDA, p0e = p0, lanee = lane;
    asm volatile("" : "+s"(LDAe), "+s"(p0e), "+v"(lanee));
    const RovsTile<TR> te{A, LDAe};
    float rowv[TF::NACC];
#pragma unroll
    for (int r = 0; r < TF::NACC; ++r) rowv[r] = 0.f;
#pragma unroll
    for (int t = 0; t < kRovsNT; ++t) {
      if (t >= my_n) break;
      const int col = tcol[t] * TF::TN + TF::acc_col(lanee);
      const int pcol = TR::perm(col);
      const float Kc = K[col], Ki = Kinv[col];
#pragma unroll
      for (int r4 = 0; r4 < TF::NACC; r4 += 4) {
        float z[4];
        const int row0 = TF::acc_row(lanee, r4);
        TF::normal4((uint32_t)(p0e + row0) >> 2, (uint32_t)year, (uint32_t)col, a.tag0, a.key0,
                    a.key1, z);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int row = TF::acc_row(lanee, r4 + s);
          const float x = sel[t][r4 + s];
          const float g = grm[row] + z[s] * lv;
          const float n = x * (1.f + g * (Kc - x) * Ki * 0.01f);
          te.put(row, pcol, n);
          rowv[r4 + s] += n;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < TF::NACC; ++r) {
      const float v = half_sum(rowv[r]);
      if ((lanee & 31) == 0) rsum[wave * kRovsRows + TF::acc_row(lanee, r)] = v;
    }
    __syncthreads();
    if (tid < kRovsRows) {
      float s = 0.f;
      for (int w = 0; w < NW; ++w) s += rsum[w * kRovsRows + tid];
      tot[tid] = s;
      const int p = p0 + tid;
      if (p < a.N) {
        a.total_pops[(size_t)(year + 1) * a.N + p] = s;
        for (int c = 0; c < C; ++c)
          a.aars[((size_t)year * C + c) * a.N + p] = aar[c * kRovsRows + tid];
      }
    }
    __syncthreads();
  }
}

template <class TR, bool DOTS>
static hipError_t launch_rovs_forward_t(const RovsForwardArgs& a, hipStream_t s) {
  const size_t lds = rovs_forward_lds(a.Pp, a.C, TR::kSplit, DOTS);
  hipError_t e =
      hipFuncSetAttribute(reinterpret_cast<const void*>(&rovs_forward_kernel<TR, DOTS>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const int blocks = (a.N + kRovsRows - 1) / kRovsRows;
  hipLaunchKernelGGL((rovs_forward_kernel<TR, DOTS>), dim3(blocks), dim3(64 * rovs_waves(a.Pp)),
                     lds, s, a);
  return hipGetLastError();
}

hipError_t launch_rovs_forward(const RovsForwardArgs& a, hipStream_t s) {
  // Split-bf16 products contract over Pp (an even count of 16-k chunks).
  if (a.split && a.Pk != a.Pp) return hipErrorInvalidValue;
  if (a.rvec) {
    // The dots' form: FP32 products, at most kRovsMaxC controls (C x 32 <= 256
    // threads finish the sums).
    if (a.split || a.C < 1 || a.C > kRovsMaxC) return hipErrorInvalidValue;
    return launch_rovs_forward_t<TF, true>(a, s);
  }
  return a.split ? launch_rovs_forward_t<TraitsB3, false>(a, s)
                 : launch_rovs_forward_t<TF, false>(a, s);
}

// Unit profit (computePathStates, SimulateGPU.cu:985-1002) with the ore
// composition drawn per path, year and commodity (kind 3 of the stream
// contract): N(ore_mean, ore_sd).
__global__ __launch_bounds__(256) void rovs_profit_kernel(RovsProfitArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.N) return;
  for (int t = 0; t < a.T; ++t) {
    const double* x = a.prices + (size_t)t * a.U * a.N;
    double fuel = 0.0;
    for (int v = 0; v < a.n_vehicles; ++v) fuel += a.fuel_cost[v] * x[(size_t)a.fuel_idx[v] * a.N + p];
    double ore = 0.0;
    for (int b = 0; 4 * b < a.n_comm; ++b) {
      double z[4];
      rmc_normal4_d((uint32_t)p, (uint32_t)t, (uint32_t)b, a.tag, a.key0, a.key1, z);
      for (int i = 0; i < 4 && 4 * b + i < a.n_comm; ++i) {
        const int c = 4 * b + i;
        const double comp = a.ore_mean[c] + a.ore_sd[c] * z[i];
        if (a.composition) a.composition[((size_t)t * a.n_comm + c) * a.N + p] = comp;
        ore += comp * x[(size_t)(a.n_fuels + c) * a.N + p];
      }
    }
    a.profit[(size_t)t * a.N + p] = a.unit_cost + fuel - a.unit_revenue * ore;
  }
}

hipError_t launch_rovs_profit(const RovsProfitArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(rovs_profit_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

// Adjusted population of species sp at step t under control c.
__device__ __forceinline__ double rovs_adj(const RovsStepArgs* a, int sp, int t, int c, int p) {
  const int N = a->N, T = a->T, C = a->C;
  return (double)a->total_pops[((size_t)sp * (T + 1) + t) * N + p] *
         (double)a->aars[(((size_t)sp * T + t) * C + c) * N + p];
}

// Last step (backwardInduction with start = T - 1, SimulateGPU.cu:2375-2452):
// payoff Q_c x unit profit (<= 0) for the valid controls, the minimum wins.
__global__ __launch_bounds__(256) void rovs_terminal_kernel(const RovsStepArgs* __restrict__ a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a->N) return;
  const int t = a->T - 1, C = a->C, S = a->S, N = a->N;
  const double up = a->profit[(size_t)t * N + p];
  double best = 0.0;
  int bc = 0;
  for (int c = 0; c < C; ++c) {
    bool valid = true;
    for (int sp = 0; sp < S && c > 0; ++sp) valid = valid && !(rovs_adj(a, sp, t, c, p) < a->thr[sp]);
    double pay = 0.0;
    if (valid) pay = fmin(a->flow[c] * up, 0.0);
    if (c == 0) {
      best = pay;
    } else if (valid && pay < best) {
      best = pay;
      bc = c;
    }
  }
  a->V[(size_t)t * N + p] = best;
  a->optc[(size_t)t * N + p] = bc;
  if (a->adj_pops)
    for (int sp = 0; sp < S; ++sp)
      a->adj_pops[((size_t)t * S + sp) * N + p] = rovs_adj(a, sp, t, C - 1, p);
}

hipError_t launch_rovs_terminal(const RovsStepArgs* a, int N, hipStream_t s) {
  hipLaunchKernelGGL(rovs_terminal_kernel, dim3((N + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

// Regression sample of path p at step t (allocateXYRegressionData,
// SimulateGPU.cu:1010-1068): included when every highest-flow adjusted
// population reaches its threshold, or when the drawn control is 0.
__device__ __forceinline__ int rovs_include(const RovsStepArgs* a, int t, int p, double* x) {
  const int S = a->S, C = a->C;
  bool valid = true;
  for (int sp = 0; sp < S; ++sp) {
    x[sp] = rovs_adj(a, sp, t, C - 1, p);
    valid = valid && !(x[sp] < a->thr[sp]);
  }
  x[S] = a->profit[(size_t)t * a->N + p];
  const int c = a->controls[(size_t)t * a->N + p];
  return (valid || c == 0) ? c : -1;
}

// Per block of kRovsScanBlock paths: each control's included count.
__global__ __launch_bounds__(kRovsScanBlock) void rovs_count_kernel(const RovsStepArgs* __restrict__ a,
                                                                    int t) {
  __shared__ int cnt[kRovsMaxC];
  const int p = blockIdx.x * kRovsScanBlock + threadIdx.x;
  if (threadIdx.x < kRovsMaxC) cnt[threadIdx.x] = 0;
  __syncthreads();
  int c = -1;
  if (p < a->N) {
    double x[kRovsMaxS + 1];
    c = rovs_include(a, t, p, x);
    a->flag[p] = c;
  }
  if (c >= 0) atomicAdd(&cnt[c], 1);
  __syncthreads();
  if (threadIdx.x < a->C) a->blk_count[blockIdx.x * a->C + threadIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of each control's block counts: one wave per control, 64
// blocks per round (a wave prefix sum, then the round's total carried on).
__global__ __launch_bounds__(64 * kRovsMaxC) void rovs_scan_kernel(const RovsStepArgs* __restrict__ a,
                                                                  int nblk) {
  const int c = threadIdx.x >> 6, lane = threadIdx.x & 63, C = a->C;
  if (c >= C) return;
  int run = 0;
  for (int b0 = 0; b0 < nblk; b0 += 64) {
    const int b = b0 + lane;
    const int v = b < nblk ? a->blk_count[b * C + c] : 0;
    int inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(inc, off);
      if (lane >= off) inc += o;
    }
    if (b < nblk) a->blk_off[b * C + c] = run + inc - v;
    run += __shfl(inc, 63);
  }
  if (lane == 0) a->count[c] = run;
}

// Scatter in path order: a path's position is its block's offset plus the
// number of earlier paths of its block with the same control.
__global__ __launch_bounds__(kRovsScanBlock) void rovs_scatter_kernel(const RovsStepArgs* __restrict__ a,
                                                                      int t) {
  constexpr int NW = kRovsScanBlock / 64;
  __shared__ int wcnt[kRovsMaxC][NW];
  const int p = blockIdx.x * kRovsScanBlock + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = p < a->N ? a->flag[p] : -1;
  int rank = 0;
  for (int cc = 0; cc < a->C; ++cc) {
    const unsigned long long m = __ballot(c == cc);
    if (lane == 0) wcnt[cc][w] = __popcll(m);
    if (c == cc)
      rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
  }
  __syncthreads();
  if (c < 0) return;
  for (int ww = 0; ww < w; ++ww) rank += wcnt[c][ww];
  const int pos = a->blk_off[blockIdx.x * a->C + c] + rank;
  double x[kRovsMaxS + 1];
  rovs_include(a, t, p, x);
  const int D = a->D, N = a->N;
  for (int d = 0; d < D; ++d) a->xs[((size_t)c * N + pos) * D + d] = x[d];
  a->ys[(size_t)c * N + pos] = a->V[(size_t)(t + 1) * N + p];
}

hipError_t launch_rovs_collect(const RovsStepArgs* a, int N, int t, hipStream_t s) {
  const int nblk = (N + kRovsScanBlock - 1) / kRovsScanBlock;
  hipLaunchKernelGGL(rovs_count_kernel, dim3(nblk), dim3(kRovsScanBlock), 0, s, a, t);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(rovs_scan_kernel, dim3(1), dim3(64 * kRovsMaxC), 0, s, a, nblk);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(rovs_scatter_kernel, dim3(nblk), dim3(kRovsScanBlock), 0, s, a, t);
  return hipGetLastError();
}

// State ranges over the samples of the controls that computeStateMinMax
// (SimulateGPU.cu:1073-1147) takes its own range from: control 0 and every
// control with more than D + 1 samples.  Block partials, then one block.
__global__ __launch_bounds__(256) void rovs_minmax_kernel(const RovsStepArgs* __restrict__ a) {
  __shared__ double smin[256], smax[256];
  const int d = blockIdx.y, D = a->D, N = a->N;
  double lo = __builtin_inf(), hi = -__builtin_inf();
  for (int c = 0; c < a->C; ++c) {
    const int n = a->count[c];
    if (c > 0 && n <= D + 1) continue;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
      const double v = a->xs[((size_t)c * N + j) * D + d];
      lo = fmin(lo, v);
      hi = fmax(hi, v);
    }
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + off]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a->mm[((size_t)blockIdx.x * D + d) * 2 + 0] = smin[0];
    a->mm[((size_t)blockIdx.x * D + d) * 2 + 1] = smax[0];
  }
}

// One wave per state dimension over the block partials (minimum and maximum
// are exact, so the order of the reduction does not matter).
__global__ __launch_bounds__(64 * (kRovsMaxS + 1)) void rovs_minmax_final_kernel(
    const RovsStepArgs* __restrict__ a, int nb) {
  const int d = threadIdx.x >> 6, lane = threadIdx.x & 63, D = a->D;
  if (d >= D) return;
  double lo = __builtin_inf(), hi = -__builtin_inf();
  for (int b = lane; b < nb; b += 64) {
    lo = fmin(lo, a->mm[((size_t)b * D + d) * 2 + 0]);
    hi = fmax(hi, a->mm[((size_t)b * D + d) * 2 + 1]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, off));
    hi = fmax(hi, __shfl_xor(hi, off));
  }
  if (lane != 0) return;
  if (!(lo <= hi)) lo = hi = 0.0;  // no samples at all
  a->range[d * 2 + 0] = lo;
  a->range[d * 2 + 1] = hi;
}

constexpr int kRovsMinMaxBlocks = 64;

hipError_t launch_rovs_ranges(const RovsStepArgs* a, int N, int D, hipStream_t s) {
  (void)N;
  hipLaunchKernelGGL(rovs_minmax_kernel, dim3(kRovsMinMaxBlocks, D), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(rovs_minmax_final_kernel, dim3(1), dim3(64 * (kRovsMaxS + 1)), 0, s, a,
                     kRovsMinMaxBlocks);
  return hipGetLastError();
}

// Scaled coordinates of the kNN search and fit (SimulateGPU.cu:4582-4607):
// species dimensions (x - min) / (max - min), unit profit
// (x - min) / (scaling (max - min)); grid coordinates x_min + i (max - min) /
// (R - 1) (createQueryPoints, :1176-1190), the query points scaled the same
// way.
__device__ __forceinline__ double rovs_scale(const RovsStepArgs* a, int d, double v) {
  const double lo = a->range[2 * d], span = a->range[2 * d + 1] - a->range[2 * d];
  if (!(span > 0.0)) return 0.0;
  return d == a->D - 1 ? (v - lo) / (a->scaling * span) : (v - lo) / span;
}

__global__ __launch_bounds__(256) void rovs_normalise_kernel(const RovsStepArgs* __restrict__ a) {
  const int D = a->D, N = a->N, R = a->R, C = a->C;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < (size_t)C * N) {
    const int c = (int)(i / N), j = (int)(i - (size_t)c * N);
    if (j < a->count[c])
      for (int d = 0; d < D; ++d)
        a->u[((size_t)c * N + j) * D + d] = rovs_scale(a, d, a->xs[((size_t)c * N + j) * D + d]);
  }
  if (i < (size_t)D * R) {
    const int d = (int)(i / R), k = (int)(i - (size_t)d * R);
    const double lo = a->range[2 * d], span = a->range[2 * d + 1] - a->range[2 * d];
    const double q = (double)k * span / (double)(R - 1) + lo;
    a->qtab[i] = rovs_scale(a, d, q);
    size_t G = 1;
    for (int dd = 0; dd < D; ++dd) G *= (size_t)R;
    for (int c = 0; c < C; ++c) a->grid[(size_t)c * ((size_t)D * R + G) + i] = q;
  }
}

hipError_t launch_rovs_normalise(const RovsStepArgs* a, int N, int D, int R, hipStream_t s) {
  const size_t n = std::max((size_t)kRovsMaxC * N, (size_t)D * R);
  hipLaunchKernelGGL(rovs_normalise_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

// Multilinear interpolation of one control's grid at state x.
template <int D>
__device__ __forceinline__ double rovs_interp(const double* grid, int R, const double* x) {
  const double* val = grid + (size_t)D * R;
  int li[D];
  double xd[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double lo = grid[d * R], hi = grid[d * R + R - 1];
    const double c = hi != lo ? (R - 1) * (x[d] - lo) / (hi - lo) : 0.0;
    const int i = c >= (double)(R - 2) ? R - 2 : (c > 0.0 ? (int)c : 0);
    li[d] = i;
    const double x0 = grid[d * R + i], x1 = grid[d * R + i + 1];
    xd[d] = (fabs(x1 - x0) < 2.220446049250313e-16 || x0 == x1) ? 0.0 : (x[d] - x0) / (x1 - x0);
  }
  double s = 0.0;
  for (int c = 0; c < (1 << D); ++c) {
    double w = 1.0;
    size_t idx = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int bit = (c >> (D - 1 - d)) & 1;
      w *= bit ? xd[d] : 1.0 - xd[d];
      idx = idx * (size_t)R + (size_t)(li[d] + bit);
    }
    s += val[idx] * w;
  }
  return s;
}

// backwardInduction at an interior step (SimulateGPU.cu:2453-2771): for each
// valid control, Q_c x unit profit plus the discounted continuation
// interpolated from that control's grid at the highest-flow state (clamped to
// <= 0); switch from the running best only below hysteresis x best.
template <int D>
__global__ __launch_bounds__(256) void rovs_decide_kernel(const RovsStepArgs* __restrict__ a, int t) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a->N) return;
  const int C = a->C, S = a->S, N = a->N, R = a->R;
  double x[D];
#pragma unroll
  for (int d = 0; d < D - 1; ++d) x[d] = rovs_adj(a, d, t, C - 1, p);
  x[D - 1] = a->profit[(size_t)t * N + p];
  size_t G = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) G *= (size_t)R;
  const size_t gsz = (size_t)D * R + G;
  double best = 0.0;
  int bc = 0;
  for (int c = 0; c < C; ++c) {
    bool valid = true;
    for (int sp = 0; sp < S && c > 0; ++sp) valid = valid && !(rovs_adj(a, sp, t, c, p) < a->thr[sp]);
    double pay = 0.0;
    if (valid) {
      const double cont = fmin(rovs_interp<D>(a->grid + (size_t)c * gsz, R, x) * a->disc, 0.0);
      pay = a->flow[c] * x[D - 1] + cont;
    }
    if (c == 0) {
      best = pay;
    } else if (valid && pay < a->hysteresis * best) {
      best = pay;
      bc = c;
    }
  }
  a->V[(size_t)t * N + p] = best;
  a->optc[(size_t)t * N + p] = bc;
  if (a->adj_pops)
    for (int sp = 0; sp < S; ++sp) a->adj_pops[((size_t)t * S + sp) * N + p] = x[sp];
}

hipError_t launch_rovs_decide(const RovsStepArgs* a, int N, int D, int t, hipStream_t s) {
  const dim3 g((N + 255) / 256), b(256);
  switch (D) {
    case 2: hipLaunchKernelGGL(rovs_decide_kernel<2>, g, b, 0, s, a, t); break;
    case 3: hipLaunchKernelGGL(rovs_decide_kernel<3>, g, b, 0, s, a, t); break;
    case 4: hipLaunchKernelGGL(rovs_decide_kernel<4>, g, b, 0, s, a, t); break;
    case 5: hipLaunchKernelGGL(rovs_decide_kernel<5>, g, b, 0, s, a, t); break;
    case 6: hipLaunchKernelGGL(rovs_decide_kernel<6>, g, b, 0, s, a, t); break;
    case 7: hipLaunchKernelGGL(rovs_decide_kernel<7>, g, b, 0, s, a, t); break;
    case 8: hipLaunchKernelGGL(rovs_decide_kernel<8>, g, b, 0, s, a, t); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---- The recomputation (ALGO6) ---------------------------------------------

int rovs_tile_col(int k, bool split) { return split ? TraitsB3::perm(k) : TF::perm(k); }

size_t rovs_recompute_lds(int Pp, int C, int S, bool split) {
  const int LDA = Pp + rovs_pad(split);
  const int nq = rovs_waves(Pp) * 64 / kRovsRows;
  return (size_t)S * kRovsRows * LDA * sizeof(float)  // the species' population tiles
         + (size_t)nq * C * kRovsRows * sizeof(float) // partial state sums
         + (size_t)S * C * kRovsRows * sizeof(float)  // states per species and control
         + kRovsRows * sizeof(float)                  // growth-rate means of the year
         + kRovsRows * sizeof(int);                   // controls of the year
}

// One path's decision at year u on the recomputed states st [S][C][32]
// (backwardInduction's rule, the terminal rule at T - 1) and its cash flow:
// the chosen control's Q_c x unit profit, or the terminal payoff.
template <int D>
__device__ double rovs_policy_step(const RovsRecomputeArgs& a, const float* st, int row, int p,
                                   int u, int* bc_out) {
  constexpr int S = D - 1;
  const int C = a.C, N = a.N, R = a.R;
  double x[D];
#pragma unroll
  for (int sp = 0; sp < S; ++sp) x[sp] = (double)st[(sp * C + C - 1) * kRovsRows + row];
  x[S] = a.profit[(size_t)u * N + p];
  size_t G = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) G *= (size_t)R;
  const size_t gsz = (size_t)D * R + G;
  const bool last = u == a.T - 1;
  const double hyst = last ? 1.0 : a.hysteresis;
  double best = 0.0;
  int bc = 0;
  for (int c = 0; c < C; ++c) {
    bool valid = true;
#pragma unroll
    for (int sp = 0; sp < S; ++sp)
      valid = valid && (c == 0 || !((double)st[(sp * C + c) * kRovsRows + row] < a.thr[sp]));
    double pay = 0.0;
    if (valid) {
      if (last) {
        pay = fmin(a.flow[c] * x[S], 0.0);
      } else {
        const double* g = a.grids + ((size_t)u * C + c) * gsz;
        pay = a.flow[c] * x[S] + fmin(rovs_interp<D>(g, R, x) * a.disc, 0.0);
      }
    }
    if (c == 0) {
      best = pay;
    } else if (valid && pay < hyst * best) {
      best = pay;
      bc = c;
    }
  }
  if (u == a.t) {
    a.optc[(size_t)u * N + p] = bc;
    if (a.adj_pops)
#pragma unroll
      for (int sp = 0; sp < S; ++sp) a.adj_pops[((size_t)u * S + sp) * N + p] = x[sp];
  }
  *bc_out = bc;
  return last ? best : a.flow[bc] * x[S];
}

// Every path of a 32-path tile from the checkpoint year y0: the drawn
// controls up to t, then the fitted policy from t to T - 1 (optimalForwardPaths,
// SimulateGPU.cu:1199-2206, with the value of the path the realized, discounted
// cash flows).  The migration and growth are the forward kernel's: one MFMA
// product per control on the species' tile, each path's own control kept by a
// register select, the epilogue with the stored growth-rate mean of the year
// and the same Philox noise, so the replay reproduces the forward pass.  The
// states of the decision, every control's total after one migration, are
// n . r_c with r_c the row sums of trans o surv_c (fixed-order sums).
template <class TR>
__global__ __launch_bounds__(kRovsMaxThreads)
__attribute__((amdgpu_waves_per_eu(rovs_min_waves<TR>()))) void rovs_recompute_kernel(
    RovsRecomputeArgs a) {
  extern __shared__ char smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nthreads = blockDim.x, NW = nthreads >> 6, nq = nthreads / kRovsRows;
  const int Pp = a.Pp, LDA = Pp + TR::PAD, C = a.C, S = a.S, N = a.N;
  const int p0 = blockIdx.x * kRovsRows;
  float* A0 = reinterpret_cast<float*>(smem);
  float* part = A0 + (size_t)S * kRovsRows * LDA;  // [nq][C][32]
  float* st = part + nq * C * kRovsRows;           // [S][C][32]
  float* gm = st + S * C * kRovsRows;              // [32]
  int* ctl = reinterpret_cast<int*>(gm + kRovsRows);

  for (int sp = 0; sp < S; ++sp) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(a.ckpt) + sp * a.ckpt_species +
                          (size_t)(a.y0 / a.ck) * N * Pp;
    const RovsTile<TR> tile{A0 + (size_t)sp * kRovsRows * LDA, LDA};
    for (int idx = tid; idx < kRovsRows * Pp; idx += nthreads) {
      const int r = idx / Pp, c = idx - r * Pp;
      tile.put_raw(r, c, (p0 + r < N) ? src[(size_t)(p0 + r) * Pp + c] : 0u);
    }
  }
  const int ntc = Pp / TF::TN;
  const int base = ntc / NW, extra = ntc % NW;
  const int my_n = base + (wave < extra ? 1 : 0);
  const int my_t0 = wave * base + (wave < extra ? wave : extra);
  int tcol[kRovsNT];
#pragma unroll
  for (int t = 0; t < kRovsNT; ++t) tcol[t] = my_t0 + t;
  const int arow = TR::a_row(lane), akoff = (TR::kSplit ? 8 : TR::KV) * TR::k_lane(lane);
  const int nkg = a.Pk >> 3;
  const size_t mbytes = (size_t)Pp * a.Pk * sizeof(float);
  double vacc = 0.0, fac = 1.0;  // the realized value, threads tid < 32
  __syncthreads();

  for (int u = a.y0; u < a.T; ++u) {
    if (u >= a.t) {
      // States: partial sums over a stride of columns, then a fixed-order
      // sum over the strides.
      const int r = tid & (kRovsRows - 1), q = tid / kRovsRows;
      for (int sp = 0; sp < S; ++sp) {
        const RovsTile<TR> tile{A0 + (size_t)sp * kRovsRows * LDA, LDA};
        const float* rv = a.rvec + (size_t)sp * C * Pp;
        float accs[kRovsMaxC];
#pragma unroll
        for (int c = 0; c < kRovsMaxC; ++c) accs[c] = 0.f;
        for (int i = q; i < Pp; i += nq) {
          const float v = tile.get(r, i);
#pragma unroll
          for (int c = 0; c < kRovsMaxC; ++c)
            if (c < C) accs[c] += v * rv[(size_t)c * Pp + i];
        }
#pragma unroll
        for (int c = 0; c < kRovsMaxC; ++c)
          if (c < C) part[(q * C + c) * kRovsRows + r] = accs[c];
        __syncthreads();
        if (tid < C * kRovsRows) {
          const int c = tid / kRovsRows, rr = tid - c * kRovsRows;
          float sum = 0.f;
          for (int qq = 0; qq < nq; ++qq) sum += part[(qq * C + c) * kRovsRows + rr];
          st[(sp * C + c) * kRovsRows + rr] = sum;
        }
        __syncthreads();
      }
      if (tid < kRovsRows) {
        const int p = p0 + tid;
        int bc = 0;
        if (p < N) {
          double cash = 0.0;
          switch (a.D) {
            case 2: cash = rovs_policy_step<2>(a, st, tid, p, u, &bc); break;
            case 3: cash = rovs_policy_step<3>(a, st, tid, p, u, &bc); break;
            case 4: cash = rovs_policy_step<4>(a, st, tid, p, u, &bc); break;
            case 5: cash = rovs_policy_step<5>(a, st, tid, p, u, &bc); break;
            case 6: cash = rovs_policy_step<6>(a, st, tid, p, u, &bc); break;
            case 7: cash = rovs_policy_step<7>(a, st, tid, p, u, &bc); break;
            default: cash = rovs_policy_step<8>(a, st, tid, p, u, &bc); break;
          }
          vacc += fac * cash;
          fac *= a.disc;
        }
        ctl[tid] = bc;
      }
      __syncthreads();
      if (u == a.T - 1) break;
    } else {
      if (tid < kRovsRows) ctl[tid] = (p0 + tid < N) ? a.controls[(size_t)u * N + p0 + tid] : 0;
      __syncthreads();
    }
    // The controls some path of the tile takes this year (the same mask in
    // every wave): only their products are formed.  The surviving fractions
    // come from the row-sum dots above, so a product no path takes is unused.
    unsigned used = 0u;
    {
      const int cl = lane < kRovsRows ? ctl[lane] : -1;
      for (int c = 0; c < C; ++c)
        if (__ballot(cl == c) != 0ull) used |= 1u << c;
    }
    // Migration under each path's control and growth, species by species.
    for (int sp = 0; sp < S; ++sp) {
      float* A = A0 + (size_t)sp * kRovsRows * LDA;
      if (tid < kRovsRows)
        gm[tid] = (p0 + tid < N) ? a.grm[((size_t)sp * a.T + u) * N + p0 + tid] : 0.f;
      int LDAy = LDA, laney = lane;
      asm volatile("" : "+s"(LDAy), "+v"(laney));
      floatx16 sel[kRovsNT], acc[kRovsNT];
#pragma unroll
      for (int t = 0; t < kRovsNT; ++t) sel[t] = floatx16{};
      for (int c = 0; c < C; ++c) {
        if (!((used >> c) & 1u)) continue;
        const PackedB B = packed_b(a.mpack + ((size_t)sp * C + c) * Pp * a.Pk, nkg, mbytes);
        if (my_n == 1)
          rovs_product<TR, 1>(acc, B, A, tcol, LDAy, arow, akoff, laney);
        else if (my_n == 2)
          rovs_product<TR, 2>(acc, B, A, tcol, LDAy, arow, akoff, laney);
#pragma unroll
        for (int r = 0; r < TF::NACC; ++r) {
          const int row = TF::acc_row(laney, r);
          if (ctl[row] == c) {
#pragma unroll
            for (int t = 0; t < kRovsNT; ++t)
              if (t < my_n) sel[t][r] = acc[t][r];
          }
        }
      }
      // Every wave has read the tile (all columns feed every product) before
      // the epilogue overwrites it.
      __syncthreads();
      int LDAe = LDA, p0e = p0, lanee = lane;
      asm volatile("" : "+s"(LDAe), "+s"(p0e), "+v"(lanee));
      const float* K = a.cols + (size_t)sp * 3 * Pp;
      const float* Kinv = K + Pp;
      const float lv = a.lv[sp];
      const RovsTile<TR> te{A, LDAe};
#pragma unroll
      for (int t = 0; t < kRovsNT; ++t) {
        if (t >= my_n) break;
        const int col = tcol[t] * TF::TN + TF::acc_col(lanee);
        const int pcol = TR::perm(col);
        const float Kc = K[col], Ki = Kinv[col];
#pragma unroll
        for (int r4 = 0; r4 < TF::NACC; r4 += 4) {
          float z[4];
          const int row0 = TF::acc_row(lanee, r4);
          TF::normal4((uint32_t)(p0e + row0) >> 2, (uint32_t)u, (uint32_t)col, a.tag0[sp], a.key0,
                      a.key1, z);
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) {
            const int row = TF::acc_row(lanee, r4 + s4);
            const float x = sel[t][r4 + s4];
            const float g = gm[row] + z[s4] * lv;
            te.put(row, pcol, x * (1.f + g * (Kc - x) * Ki * 0.01f));
          }
        }
      }
      __syncthreads();
    }
  }
  if (tid < kRovsRows && p0 + tid < N) a.V[(size_t)a.t * N + p0 + tid] = vacc;
}

template <class TR>
static hipError_t launch_rovs_recompute_t(const RovsRecomputeArgs& a, hipStream_t s) {
  const size_t lds = rovs_recompute_lds(a.Pp, a.C, a.S, TR::kSplit);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&rovs_recompute_kernel<TR>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const int blocks = (a.N + kRovsRows - 1) / kRovsRows;
  hipLaunchKernelGGL(rovs_recompute_kernel<TR>, dim3(blocks), dim3(64 * rovs_waves(a.Pp)), lds,
                     s, a);
  return hipGetLastError();
}

// Smoothed-validity value of each path under the policy and the
// coefficients dJ_p / dA_s,t (RovsSmoothArgs); one thread per path, the
// block's values summed in a fixed tree order.
__global__ __launch_bounds__(256) void rovs_smooth_value_kernel(RovsSmoothArgs a) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const int n = blockIdx.x * 256 + tid;
  double jp = 0.0;
  if (n < a.N) {
    double f = 1.0;
    for (int t = 0; t < a.T; ++t) {
      const int u = a.policy[(size_t)t * a.N + n];
      const double up = a.profit[(size_t)t * a.N + n];
      double sg[kRovsMaxS];
      double v = 1.0;
      for (int sp = 0; sp < a.S; ++sp) {
        const double A = (double)a.total_pops[((size_t)sp * (a.T + 1) + t) * a.N + n] *
                         (double)a.aars[(((size_t)sp * a.T + t) * a.C + u) * a.N + n];
        sg[sp] = 1.0 / (1.0 + exp(-(A - a.thr[sp]) / a.tau));
        v *= sg[sp];
      }
      const double dq = a.flow[u] - a.flow[0];
      jp += f * up * (a.flow[0] + dq * v);
      for (int sp = 0; sp < a.S; ++sp)
        a.coef[((size_t)sp * a.T + t) * a.N + n] = (float)(f * up * dq * v * (1.0 - sg[sp]) / a.tau);
      f *= a.disc;
    }
  }
  red[tid] = jp;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) a.jpart[blockIdx.x] = red[0];
}

hipError_t launch_rovs_smooth_value(const RovsSmoothArgs& a, hipStream_t s) {
  if (a.N < 1 || a.S < 1 || a.S > kRovsMaxS || a.C < 1 || a.C > kRovsMaxC || !(a.tau > 0))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(rovs_smooth_value_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

size_t rovs_adj_lds(int Pp) {
  return (size_t)kRovsRows * (Pp + TF::PAD) * sizeof(float)  // population, then w, tile
         + 2 * kRovsRows * sizeof(float)                     // growth-rate means, coefficients
         + kRovsRows * sizeof(int);                          // controls of the year
}

// Reverse sweep of one species under the policy (RovsAdjArgs).  lambda stays
// in the accumulator layout of the wave's column tiles from year to year: the
// tile assignment of the two products is the same, so each lane keeps the
// entries it needs for the next year's w.
// Three waves per SIMD, as the FP32 forward pass (168 VGPRs, a few spilled).
__global__ __launch_bounds__(kRovsMaxThreads) __attribute__((amdgpu_waves_per_eu(3))) void
rovs_adj_kernel(RovsAdjArgs a) {
  constexpr int NT = kRovsNT;
  extern __shared__ char smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nthreads = blockDim.x, NW = nthreads >> 6;
  const int Pp = a.Pp, LDA = Pp + TF::PAD, C = a.C, N = a.N;
  const int p0 = blockIdx.x * kRovsRows;
  float* A = reinterpret_cast<float*>(smem);
  float* gm = A + kRovsRows * LDA;
  float* cf = gm + kRovsRows;
  int* ctl = reinterpret_cast<int*>(cf + kRovsRows);
  const float* K = a.cols;
  const float* Kinv = a.cols + Pp;
  const int ntc = Pp / TF::TN;
  const int base = ntc / NW, extra = ntc % NW;
  const int my_n = base + (wave < extra ? 1 : 0);
  const int my_t0 = wave * base + (wave < extra ? wave : extra);
  int tcol[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) tcol[t] = my_t0 + t;
  const int arow = TF::a_row(lane), akoff = TF::KV * TF::k_lane(lane);
  const int nkg = a.Pk >> 3;
  const size_t mbytes = (size_t)Pp * a.Pk * sizeof(float);
  floatx16 lam[NT], sel[NT], acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) lam[t] = floatx16{};

  for (int year = a.T - 1; year >= 0; --year) {
    // The tile entering the year (stored positions, as the forward pass
    // wrote it), the year's controls, coefficients and growth-rate means.
    const float* src = a.ckpt + (size_t)year * N * Pp;
    for (int idx = tid; idx < kRovsRows * Pp; idx += nthreads) {
      const int r = idx / Pp, c = idx - r * Pp;
      A[r * LDA + c] = (p0 + r < N) ? src[(size_t)(p0 + r) * Pp + c] : 0.f;
    }
    if (tid < kRovsRows) {
      const bool in = p0 + tid < N;
      const size_t at = (size_t)year * N + p0 + tid;
      ctl[tid] = in ? a.policy[at] : 0;
      cf[tid] = in ? a.coef[at] : 0.f;
      gm[tid] = in ? a.grm[at] : 0.f;
    }
    __syncthreads();
    int LDAy = LDA, laney = lane;
    asm volatile("" : "+s"(LDAy), "+v"(laney));
    // The controls some path of the tile takes this year (the same mask in
    // every wave): only their products are formed in both sweeps below.
    unsigned used = 0u;
    {
      const int cl = lane < kRovsRows ? ctl[lane] : -1;
      for (int c = 0; c < C; ++c)
        if (__ballot(cl == c) != 0ull) used |= 1u << c;
    }
    // x_t: the product under the path's control.
#pragma unroll
    for (int t = 0; t < NT; ++t) sel[t] = floatx16{};
    for (int c = 0; c < C; ++c) {
      if (!((used >> c) & 1u)) continue;
      const PackedB B = packed_b(a.mpack + (size_t)c * Pp * a.Pk, nkg, mbytes);
      if (my_n == 1)
        rovs_product<TF, 1>(acc, B, A, tcol, LDAy, arow, akoff, laney);
      else if (my_n == 2)
        rovs_product<TF, 2>(acc, B, A, tcol, LDAy, arow, akoff, laney);
#pragma unroll
      for (int r = 0; r < TF::NACC; ++r) {
        if (ctl[TF::acc_row(laney, r)] == c) {
#pragma unroll
          for (int t = 0; t < NT; ++t)
            if (t < my_n) sel[t][r] = acc[t][r];
        }
      }
    }
    __syncthreads();  // every wave has read n_t before w_t overwrites it
    // w_t = lambda_t+1 o G'(x_t) + coef_t (the growth noise drawn as the
    // forward pass drew it), into the tile and the per-control planes.
    int LDAe = LDA, p0e = p0, lanee = lane;
    asm volatile("" : "+s"(LDAe), "+s"(p0e), "+v"(lanee));
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t >= my_n) break;
      const int col = tcol[t] * TF::TN + TF::acc_col(lanee);
      const int pcol = TF::perm(col);
      const float Kc = K[col], Ki = Kinv[col];
      const bool real = col < a.P;
#pragma unroll
      for (int r4 = 0; r4 < TF::NACC; r4 += 4) {
        float z[4];
        const int row0 = TF::acc_row(lanee, r4);
        TF::normal4((uint32_t)(p0e + row0) >> 2, (uint32_t)year, (uint32_t)col, a.tag0, a.key0,
                    a.key1, z);
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const int row = TF::acc_row(lanee, r4 + s4);
          const float x = sel[t][r4 + s4];
          const float g = gm[row] + z[s4] * a.lv;
          const float w =
              real ? lam[t][r4 + s4] * (1.f + g * (Kc - 2.f * x) * Ki * 0.01f) + cf[row] : 0.f;
          A[row * LDAe + pcol] = w;
          const int path = p0e + row;
          if (path < N) a.w_out[((size_t)year * N + path) * Pp + col] = w;
        }
      }
    }
    __syncthreads();
    // lambda_t = w_t M_u^T.
#pragma unroll
    for (int t = 0; t < NT; ++t) lam[t] = floatx16{};
    for (int c = 0; c < C; ++c) {
      if (!((used >> c) & 1u)) continue;
      const PackedB B = packed_b(a.mtpack + (size_t)c * Pp * a.Pk, nkg, mbytes);
      if (my_n == 1)
        rovs_product<TF, 1>(acc, B, A, tcol, LDAy, arow, akoff, laney);
      else if (my_n == 2)
        rovs_product<TF, 2>(acc, B, A, tcol, LDAy, arow, akoff, laney);
#pragma unroll
      for (int r = 0; r < TF::NACC; ++r) {
        if (ctl[TF::acc_row(laney, r)] == c) {
#pragma unroll
          for (int t = 0; t < NT; ++t)
            if (t < my_n) lam[t][r] = acc[t][r];
        }
      }
    }
    __syncthreads();  // every wave has read w_t before the next year's tile
  }
}

// Stable lists of the rows of each control (rows of one control in
// increasing order, the controls' lists at the caller's offsets): per block
// of kCtlBlock rows the counts per control, then per control the blocks'
// offsets (one workgroup per control: thread sums over contiguous runs of
// blocks, then a scan of the thread sums), then each row's position from the
// ballots of its wave and the counts of the waves before it.
constexpr int kCtlBlock = 1024;

__global__ __launch_bounds__(kCtlBlock) void ctl_count_kernel(const int* __restrict__ ctl, int n,
                                                              int C, int* __restrict__ blk_count) {
  __shared__ int cnt[kRovsMaxC];
  const int tid = threadIdx.x;
  if (tid < C) cnt[tid] = 0;
  __syncthreads();
  const long long i = (long long)blockIdx.x * kCtlBlock + tid;
  if (i < n) atomicAdd(&cnt[ctl[i]], 1);
  __syncthreads();
  if (tid < C) blk_count[(size_t)blockIdx.x * C + tid] = cnt[tid];
}

__global__ __launch_bounds__(1024) void ctl_scan_kernel(const int* __restrict__ blk_count,
                                                        int nblk, int C,
                                                        int* __restrict__ blk_off,
                                                        int* __restrict__ total) {
  __shared__ int part[1024];
  const int c = blockIdx.x, tid = threadIdx.x;
  const int per = (nblk + 1023) / 1024;
  const int b0 = min(nblk, tid * per), b1 = min(nblk, b0 + per);
  int sum = 0;
  for (int b = b0; b < b1; ++b) sum += blk_count[(size_t)b * C + c];
  part[tid] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
    const int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - sum;
  for (int b = b0; b < b1; ++b) {
    blk_off[(size_t)b * C + c] = run;
    run += blk_count[(size_t)b * C + c];
  }
  if (tid == 1023) total[c] = part[1023];
}

struct CtlOffsets {
  int c[kRovsMaxC];
};

__global__ __launch_bounds__(kCtlBlock) void ctl_scatter_kernel(const int* __restrict__ ctl, int n,
                                                                int C,
                                                                const int* __restrict__ blk_off,
                                                                CtlOffsets off,
                                                                int* __restrict__ rows) {
  __shared__ int wc[kCtlBlock / 64][kRovsMaxC];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long long i = (long long)blockIdx.x * kCtlBlock + tid;
  const int u = i < n ? ctl[i] : -1;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int rank = 0;
  for (int c = 0; c < C; ++c) {
    const unsigned long long m = __ballot(u == c);
    if (lane == 0) wc[wave][c] = __popcll(m);
    if (u == c) rank = __popcll(m & below);
  }
  __syncthreads();
  if (u >= 0) {
    int base = 0;
    for (int w = 0; w < wave; ++w) base += wc[w][u];
    rows[(size_t)off.c[u] + blk_off[(size_t)blockIdx.x * C + u] + base + rank] = (int)i;
  }
}

hipError_t launch_ctl_rows(const int* ctl, int n, int C, int* blk_count, int* blk_off,
                           int* total, const int* host_off, int* rows, hipStream_t s,
                           bool scatter) {
  if (n < 1 || C < 1 || C > kRovsMaxC) return hipErrorInvalidValue;
  const int nblk = (n + kCtlBlock - 1) / kCtlBlock;
  if (!scatter) {
    hipLaunchKernelGGL(ctl_count_kernel, dim3(nblk), dim3(kCtlBlock), 0, s, ctl, n, C, blk_count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ctl_scan_kernel, dim3(C), dim3(1024), 0, s, blk_count, nblk, C, blk_off,
                       total);
    return hipGetLastError();
  }
  CtlOffsets off{};
  for (int c = 0; c < C; ++c) off.c[c] = host_off[c];
  hipLaunchKernelGGL(ctl_scatter_kernel, dim3(nblk), dim3(kCtlBlock), 0, s, ctl, n, C, blk_off,
                     off, rows);
  return hipGetLastError();
}

hipError_t launch_rovs_adj(const RovsAdjArgs& a, hipStream_t s) {
  if (a.N < 1 || a.T < 1 || a.C < 1 || a.C > kRovsMaxC || a.Pp % TF::TN || a.Pp > kRovsMaxPatches ||
      a.P > a.Pp || a.Pk % 16 || a.Pk > a.Pp || a.Pk < a.P)
    return hipErrorInvalidValue;
  const size_t lds = rovs_adj_lds(a.Pp);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&rovs_adj_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const int blocks = (a.N + kRovsRows - 1) / kRovsRows;
  hipLaunchKernelGGL(rovs_adj_kernel, dim3(blocks), dim3(64 * rovs_waves(a.Pp)), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_rovs_recompute(const RovsRecomputeArgs& a, hipStream_t s) {
  if (a.D < 2 || a.D > 8 || a.S != a.D - 1 || a.C < 1 || a.C > kRovsMaxC || a.Pp % TF::TN ||
      a.Pp > kRovsMaxPatches || a.t < 1 || a.t > a.T - 2 || a.y0 > a.t || a.y0 % a.ck ||
      (a.split && a.Pk != a.Pp))
    return hipErrorInvalidValue;
  return a.split ? launch_rovs_recompute_t<TraitsB3>(a, s) : launch_rovs_recompute_t<TF>(a, s);
}

// Time 0 (firstPeriodInduction, SimulateGPU.cu:2210-2330): per control, the
// discounted mean of the step-1 values of the paths that drew it, plus
// Q_c x the mean time-0 unit profit when valid (<= 0); the minimum over the
// controls with paths wins.  Every path gets that value and control at time
// 0; the SD is over the paths that drew the chosen control.  One workgroup,
// fixed-order reductions.  out = {value, SD, control}.
__global__ __launch_bounds__(1024) void rovs_first_kernel(const RovsStepArgs* __restrict__ a,
                                                          double* out) {
  __shared__ double red[1024];
  __shared__ double sums[kRovsMaxC + 1], cnts[kRovsMaxC];
  __shared__ double res[3];
  const int N = a->N, C = a->C, S = a->S, tid = threadIdx.x;
  const double* V1 = a->V + (size_t)N;
  const int* c0 = a->controls;
  auto block_sum = [&](double v) -> double {
    red[tid] = v;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
      if (tid < off) red[tid] += red[tid + off];
      __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
  };
  // One pass for every control's sum and count, the unit profits and the
  // adjusted populations: each thread adds the same values in the same order
  // (p = tid, tid + 1024, ...) as one pass per sum would.
  double cs[kRovsMaxC], cn[kRovsMaxC];
#pragma unroll
  for (int c = 0; c < kRovsMaxC; ++c) cs[c] = cn[c] = 0.0;
  double up = 0.0;
  for (int p = tid; p < N; p += 1024) {
    const int cp = c0[p];
    const double v = V1[p];
#pragma unroll
    for (int c = 0; c < kRovsMaxC; ++c)
      if (cp == c) {
        cs[c] += v;
        cn[c] += 1.0;
      }
    up += a->profit[p];
    if (a->adj_pops)
      for (int sp = 0; sp < S; ++sp) a->adj_pops[(size_t)sp * N + p] = rovs_adj(a, sp, 0, C - 1, p);
  }
#pragma unroll
  for (int c = 0; c < kRovsMaxC; ++c) {
    if (c >= C) break;
    const double sc = block_sum(cs[c]);
    const double nc = block_sum(cn[c]);
    if (tid == 0) {
      sums[c] = sc;
      cnts[c] = nc;
    }
  }
  up = block_sum(up) / N;
  if (tid == 0) {
    double best = 0.0;
    int bc = -1;
    for (int c = 0; c < C; ++c) {
      if (cnts[c] == 0.0) continue;  // no path drew it (the reference breaks here)
      bool valid = true;
      for (int sp = 0; sp < S && c > 0; ++sp) valid = valid && !(rovs_adj(a, sp, 0, c, 0) < a->thr[sp]);
      if (!valid) continue;
      const double pay = fmin(sums[c] / cnts[c] * a->disc + a->flow[c] * up, 0.0);
      if (bc < 0 || pay < best) {
        best = pay;
        bc = c;
      }
    }
    if (bc < 0) bc = 0;
    res[0] = best;
    res[2] = (double)bc;
    sums[kRovsMaxC] = a->flow[bc] * up;
  }
  __syncthreads();
  const int bc = (int)res[2];
  const double base = sums[kRovsMaxC];
  double m = 0.0, q = 0.0;
  for (int p = tid; p < N; p += 1024)
    if (c0[p] == bc) m += V1[p] * a->disc + base;
  m = block_sum(m);
  const double n = cnts[bc] > 0.0 ? cnts[bc] : 1.0;
  const double mean = m / n;
  for (int p = tid; p < N; p += 1024) {
    if (c0[p] == bc) {
      const double d = V1[p] * a->disc + base - mean;
      q += d * d;
    }
    a->V[p] = res[0];  // time 0's row of V, apart from the V1 row read here
    a->optc[p] = bc;
  }
  q = block_sum(q);
  if (tid == 0) {
    out[0] = res[0];
    out[1] = sqrt(q / n);
    out[2] = (double)bc;
  }
}

hipError_t launch_rovs_first(const RovsStepArgs* a, double* out, hipStream_t s) {
  hipLaunchKernelGGL(rovs_first_kernel, dim3(1), dim3(1024), 0, s, a, out);
  return hipGetLastError();
}

}  // namespace roadmc
