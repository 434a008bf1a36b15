// MTE Monte Carlo kernels for gfx950 (MI355X).
//
// What the reference computes (include/gpuCode/SimulateGPU.cu:595-717,
// include/main/Simulator.cpp:686-775): for every path and year, the patch
// population vector n is moved by the movement-and-mortality matrix
// R = transProbs o survProbs (n' = R^T n, i.e. n'_j = sum_k n_k R[k][j]) and
// then grown by a stochastic logistic law.  The reference runs one thread per
// path with a serial CSC loop and reads pre-generated cuRAND arrays.
//
// How this file computes it: the paths of one species form an [N x P] state
// matrix and a year is the product [N x P] . [P x P] followed by an
// elementwise epilogue.  One workgroup (8 waves) owns BM paths for all T
// years:
//   - the BM x P population tile stays in LDS for the whole simulation, with
//     the k index permuted inside groups of 8 so that a lane's MFMA A operands
//     for 4 consecutive k-steps are one 16-byte LDS read;
//   - R is pre-packed in MFMA B-operand order, so each wave's B load is one
//     coalesced 1 KB read that L2 serves to every workgroup of the road;
//   - products run on v_mfma_f32_32x32x2_f32 (FP32 accumulate) or, for the
//     FP64 checker path, v_mfma_f64_16x16x4_f64;
//   - the epilogue draws the year noise from counter-based Philox in
//     registers (no random arrays in HBM), applies the growth law, and writes
//     the new population back to LDS; after the last year it reduces the
//     tallies in FP64.
// Growth laws: LOGNORMAL n = n'(1 + r(1 - n'/K)) exp(s z - s^2/2);
// REF_CPU  n = n'(1 + g(K - n')/K), g = dt (mu + sd z) / 100
//          (Simulator.cpp:712-728);
// REF_GPU  the growth-rate SDE of SimulateGPU.cu:686-709.
#include <hip/hip_runtime.h>

#include "gmem.h"
#include "mte_common.h"
#include "mte_kernels.h"
#include "rng.h"

namespace roadmc {

size_t mte_lds_bytes(int prec, int BM, int Pp, int n_rem_split) {
  switch (prec) {
    case kPrecF64: return lds_bytes_t<TraitsF64>(BM, Pp, n_rem_split);
    case kPrecB3: return lds_bytes_t<TraitsB3>(BM, Pp, n_rem_split);
    default: return lds_bytes_t<TraitsF32>(BM, Pp, n_rem_split);
  }
}


// Simulates one tile (BM paths of one job) for all years and writes the
// tile's tallies.  Every thread of the workgroup calls it.  The growth law is
// a template parameter so that the epilogue has no per-element branches.
// NT bounds the output tiles per wave: the accumulators and operands of NT
// tiles are allocated whether or not a wave uses them, so the launch picks the
// smallest form that holds its jobs' tiles, which leaves registers for more
// resident workgroups.
template <class TR, int LAW, int NT>
__device__ __forceinline__ void mte_tile_law(const MteJob& jb, const int tile,
                                             char* smem) {
  using T = typename TR::T;
  using VecT = typename TR::VecT;
  using AccT = typename TR::AccT;
  constexpr int TM = TR::TM, TN = TR::TN, KV = TR::KV, NACC = TR::NACC;
  const int tid = threadIdx.x, lane = tid & 63;
  // Wave index as a scalar: the per-wave tile loops below then branch on
  // SGPRs instead of exec masks.
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nthreads = blockDim.x, nwaves = nthreads >> 6;
  const int BM = jb.BM, Pp = jb.Pp, LDA = Pp + TR::PAD;
  const int p0 = tile * BM;

  T* A = reinterpret_cast<T*>(smem);
  // Split-bf16 state (TraitsB3): the hi array, then the lo array, LDA bf16
  // slots per row.
  uint16_t* const Ah = reinterpret_cast<uint16_t*>(smem);
  const int lo_off = BM * LDA;
  T* grm = reinterpret_cast<T*>(smem + align16((size_t)BM * LDA * sizeof(T)));
  double* red = reinterpret_cast<double*>(
      smem + align16((size_t)BM * LDA * sizeof(T)) +
      align16((size_t)BM * sizeof(T)));

  const RMC_GLOBAL T* cols = gmem(reinterpret_cast<const T*>(jb.cols));
  // cols holds K, 1/K, r and the initial population, Pp values each.
  const RMC_GLOBAL T* init = cols + 3 * Pp;

  const T mu = (T)jb.params[1], sd = (T)jb.params[2], rev = (T)jb.params[4];
  const T jmp = (T)jb.params[5], jprob = (T)jb.params[6], lv = (T)jb.params[7];
  const T dt = (T)jb.dt;
  // Lognormal factor exp(s z - s^2/2): the FP32 path evaluates it as
  // exp2(s' z - h') with s' = s log2(e), h' = (s^2/2) log2(e).
  const T l2e = sizeof(T) == 4 ? (T)1.4426950408889634 : (T)1;
  const T sig = (T)(jb.sigma * (sizeof(T) == 4 ? 1.4426950408889634 : 1.0));
  const T half_var = (T)(0.5 * jb.sigma * jb.sigma) * l2e;

  for (int idx = tid; idx < BM * Pp; idx += nthreads) {
    const int r = idx / Pp, c = idx - r * Pp;
    if constexpr (TR::kSplit) {
      const uint32_t hl = TR::split(init[c]);
      Ah[r * LDA + c] = (uint16_t)hl;
      Ah[lo_off + r * LDA + c] = (uint16_t)(hl >> 16);
    } else {
      A[r * LDA + TR::perm(c)] = init[c];
    }
  }
  for (int r = tid; r < BM; r += nthreads) grm[r] = (T)jb.params[0];
  __syncthreads();

  // Output tiles (TM x TN blocks of the BM x Pp result) and who computes
  // them.  Waves w and w + 4 run on the same SIMD, so the work that has to be
  // even is each SIMD's, not each wave's.
  //  - Without splitting: wave w takes a contiguous run of ceil or floor of
  //    ntiles / 8 tiles.
  //  - With splitting (jb.split): each SIMD takes q = ntiles / 4 whole tiles
  //    (ceil(q/2) in wave s, the rest in wave s + 4), and each of the
  //    ntiles % 4 remaining tiles is cut into four k-quarters, quarter s going
  //    to wave s + 4.  The quarters meet in LDS, where wave i + 4 sums the four
  //    partial products of remainder tile i in a fixed order and applies its
  //    epilogue.
  const int ntc = Pp / TN;
  const int ntiles = (BM / TM) * ntc;
  const int simd = wave & 3, upper = wave >> 2;
  int my_n, my_t0, n_rem = 0, q4 = 0;
  if (jb.split) {
    q4 = ntiles >> 2;
    n_rem = ntiles & 3;
    const int qa = (q4 + 1) >> 1;
    my_n = upper ? q4 - qa : qa;
    my_t0 = simd * q4 + (upper ? qa : 0);
  } else {
    const int base = ntiles / nwaves, extra = ntiles % nwaves;
    my_n = base + (wave < extra ? 1 : 0);
    my_t0 = wave * base + (wave < extra ? wave : extra);
  }
  const int my_np = upper ? n_rem : 0;  // k-quarter pieces of this wave
  T* qpart = reinterpret_cast<T*>(reinterpret_cast<char*>(red) + 8 * 4 * sizeof(double));
  int trow[NT], tcol[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int ti = my_t0 + t;
    trow[t] = ti / ntc;
    tcol[t] = ti - (ti / ntc) * ntc;
  }

  // The contraction runs over Pk = P rounded up to 16 (the packed matrix has
  // Pk / 8 k-groups); output columns run over Pp = P rounded up to 32.
  const int nkg = jb.Pk >> 3;
  const PackedB Bm = packed_b(jb.mpack, nkg, (size_t)Pp * jb.Pk * sizeof(T));
  // A operand offset of the lane half: KV floats (8 bf16 in the split form).
  const int arow = TR::a_row(lane), akoff = (TR::kSplit ? 8 : KV) * TR::k_lane(lane);
  const int nch = nkg >> 1;  // split form: 16-k chunks

  // All of this wave's tiles in one row block (contiguous runs usually are):
  // they can share one A fragment per k-group.
  const bool shared_a = my_n > 1 && my_t0 / ntc == (my_t0 + my_n - 1) / ntc;

  AccT acc[NT];
  VecT b0[NT];
  if constexpr (!TR::kSplit) {
#pragma unroll
    for (int t = 0; t < NT; ++t)
      if (t < my_n) b0[t] = load_b<VecT>(Bm, b_voff(Bm, tcol[t], lane), 0);
  }
  for (int year = 0; year < jb.T; ++year) {
    // Opaque per-year copies of loop-invariant scalars.  Addresses derived
    // from them (the 64 LDS write addresses and 16 Philox counters of a lane,
    // the B and A operand pointers) are recomputed each year by a few integer
    // ops instead of being hoisted out of the year loop into registers that
    // would then spill.
    int LDAy = LDA, p0y = p0;
    const RMC_GLOBAL T* colsy = cols;
    asm volatile("" : "+s"(LDAy), "+s"(p0y), "+s"(colsy));
    if (!TR::kSplit && jb.ckpt && year % jb.ckpt_every == 0) {
      // Adjoint checkpoint: the state entering this year, natural column order.
      RMC_GLOBAL T* ck = gmem(static_cast<T*>(jb.ckpt)) + ((size_t)(year / jb.ckpt_every) * jb.N) * Pp;
      for (int idx = tid; idx < BM * Pp; idx += nthreads) {
        const int r = idx / Pp, c = idx - r * Pp;
        if (p0 + r < jb.N) ck[(size_t)(p0 + r) * Pp + c] = A[r * LDA + TR::perm(c)];
      }
    }
    if constexpr (TR::kSplit) {
      // The same schedule in the split form: quarters of the 16-k chunks,
      // then the whole tiles.
      for (int i = 0; i < my_np; ++i) {
        const int ti = 4 * q4 + i;
        const int tr = ti / ntc, tc = ti - (ti / ntc) * ntc;
        AccT pacc = AccT{};
        range_product_b3(pacc, Bm, Ah, lo_off, tr, tc, (simd * nch) >> 2,
                         ((simd + 1) * nch) >> 2, LDAy, arow, akoff, lane);
        T* slot = qpart + (size_t)((i * 4 + simd) * NACC) * 64 + lane;
#pragma unroll
        for (int r = 0; r < NACC; ++r) slot[r * 64] = pacc[r];
      }
#define RMC_YP3(n, sh, i) \
  year_product_b3<n, sh>(acc + (i), Bm, Ah, lo_off, trow + (i), tcol + (i), LDAy, arow, akoff, lane)
      // Three or four tiles over two row blocks would need their A operands
      // in registers at once (spills): they run as passes of one or two.
      auto pair = [&](const int i) {
        if (trow[i] == trow[i + 1]) RMC_YP3(2, true, i);
        else RMC_YP3(2, false, i);
      };
      switch (my_n * 2 + (shared_a ? 1 : 0)) {
        case 2: case 3: RMC_YP3(1, false, 0); break;
        case 4: RMC_YP3(2, false, 0); break;
        case 5: RMC_YP3(2, true, 0); break;
        case 6: pair(0); RMC_YP3(1, false, 2); break;
        case 7: RMC_YP3(3, true, 0); break;
        case 8: pair(0); pair(2); break;
        case 9: RMC_YP3(4, true, 0); break;
        default: break;
      }
#undef RMC_YP3
    } else {
      // k-quarter pieces of the remainder tiles first (their partial products
      // go to LDS): the wave with pieces then overlaps its whole tiles with the
      // other wave on its SIMD instead of finishing alone on a one-tile chain.
      for (int i = 0; i < my_np; ++i) {
        const int ti = 4 * q4 + i;
        const int tr = ti / ntc, tc = ti - (ti / ntc) * ntc;
        AccT pacc = AccT{};
        range_product<TR>(pacc, Bm, A, tr, tc, (simd * nkg) >> 2, ((simd + 1) * nkg) >> 2,
                          LDAy, arow, akoff, lane);
        T* slot = qpart + (size_t)((i * 4 + simd) * NACC) * 64 + lane;
  #pragma unroll
        for (int r = 0; r < NACC; ++r) slot[r * 64] = pacc[r];
      }
      // One branch per year on the (wave-uniform) tile count and row sharing;
      // the product loops themselves are straight-line code.
      switch (my_n * 2 + (shared_a ? 1 : 0)) {
#define RMC_YP(n, sh)                                                                    \
  if constexpr ((n) <= NT)                                                               \
    year_product<TR, n, sh>(acc, b0, Bm, A, trow, tcol, LDAy, arow, akoff, lane);        \
  break
        case 2: case 3: RMC_YP(1, false);
        case 4: RMC_YP(2, false);
        case 5: RMC_YP(2, true);
        case 6: RMC_YP(3, false);
        case 7: RMC_YP(3, true);
        case 8: RMC_YP(4, false);
        case 9: RMC_YP(4, true);
#undef RMC_YP
        default:
          if constexpr (NT > 4) {
            switch (my_n) {
              case 5: year_product<TR, 5>(acc, b0, Bm, A, trow, tcol, LDAy, arow, akoff, lane); break;
              case 6: year_product<TR, 6>(acc, b0, Bm, A, trow, tcol, LDAy, arow, akoff, lane); break;
              case 7: year_product<TR, 7>(acc, b0, Bm, A, trow, tcol, LDAy, arow, akoff, lane); break;
              case 8: year_product<TR, 8>(acc, b0, Bm, A, trow, tcol, LDAy, arow, akoff, lane); break;
              default: break;
            }
          }
          break;
      }
    }
    __syncthreads();  // every wave has finished reading this year's state

    if constexpr (LAW == kLawRefGpu) {
      // Growth-rate SDE per path (SimulateGPU.cu:686-701).
      for (int r = tid; r < BM; r += nthreads) {
        T zb, zj, uj;
        path_scalars<T>((uint32_t)(p0 + r), (uint32_t)year, jb.tag1, jb.key0,
                        jb.key1, &zb, &zj, &uj);
        const T g = grm[r];
        const T jump = (uj < jprob) ? (T)1 : (T)0;
        const T brown = zb * sd;
        const T jsz = zj * jmp * jmp - jmp * jmp / (T)2;
        grm[r] = g + rev * (mu - g) * dt + g * brown +
                 (TR::expo(jsz) - (T)1) * g * jump;
      }
      __syncthreads();
    }

    // Fresh opaque copies for the epilogue: nothing it derives from them can
    // be computed before (and kept live across) the product loop.
    int LDAe = LDA, p0e = p0, lanee = lane;
    const RMC_GLOBAL T* colse = cols;
    asm volatile("" : "+s"(LDAe), "+s"(p0e), "+s"(colse), "+v"(lanee));
    const bool keep_pre = (year == jb.T - 1) && (jb.endpoint == kEndPre);
    // FP32 LOGNORMAL: n = x (a - b x) exp2(s' z - h') with a = 1 + r and
    // b = r / K; the pre-growth endpoint year uses s' = h' = 0, a = 1, b = 0,
    // which returns x exactly, instead of a per-element select.
    constexpr bool kFastLn = LAW == kLawLognormal && sizeof(T) == 4;
    const T sig_y = (kFastLn && keep_pre) ? (T)0 : sig;
    const T hv_y = (kFastLn && keep_pre) ? (T)0 : half_var;
    const T rscale = (kFastLn && keep_pre) ? (T)0 : (T)1;
    auto store = [&](const int row, const int pc, const T v) {
      if constexpr (TR::kSplit) {
        const uint32_t hl = TR::split(v);
        Ah[row * LDAe + pc] = (uint16_t)hl;
        Ah[lo_off + row * LDAe + pc] = (uint16_t)(hl >> 16);
      } else {
        A[row * LDAe + pc] = v;
      }
    };
    // Noise, growth and write-back of one output tile.
    auto epilogue = [&](const AccT& a, const int tr, const int tc) {
      const int col = tc * TN + TR::acc_col(lanee);
      const int pcol = TR::perm(col);
      const T K = colse[col], Ki = colse[Pp + col], rr = colse[2 * Pp + col];
      const T ga = (T)1 + rr * rscale, gb = rr * rscale * Ki;
#pragma unroll
      for (int r4 = 0; r4 < NACC; r4 += 4) {
        T z[4];
        if constexpr (TR::kQuadItems) {
          const int row0 = tr * TM + TR::acc_row(lanee, r4);
          TR::normal4((uint32_t)(p0e + row0) >> 2, (uint32_t)year,
                      (uint32_t)col, jb.tag0, jb.key0, jb.key1, z);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int row = tr * TM + TR::acc_row(lanee, r4 + s);
          T zz;
          if constexpr (TR::kQuadItems) {
            zz = z[s];
          } else {
            const uint32_t path = (uint32_t)(p0e + row);
            TR::normal4(path >> 2, (uint32_t)year, (uint32_t)col, jb.tag0,
                        jb.key0, jb.key1, z);
            zz = z[path & 3u];
          }
          const T x = a[r4 + s];
          if constexpr (kFastLn) {
            const T e = __builtin_amdgcn_exp2f(fmaf(sig_y, zz, -hv_y));
            store(row, pcol, x * fmaf(-gb, x, ga) * e);
          } else {
            T g = (T)0;
            if constexpr (LAW == kLawRefGpu) g = grm[row];
            const T n = grow<T, LAW>(x, zz, K, Ki, rr, sig, half_var, dt, mu,
                                     sd, g, lv);
            store(row, pcol, keep_pre ? x : n);
          }
        }
      }
    };
#pragma unroll
    for (int t = 0; t < NT; ++t)
      if (t < my_n) epilogue(acc[t], trow[t], tcol[t]);
    if (upper && simd < n_rem) {
      // Remainder tile `simd`: the four quarters in order, then its epilogue.
      AccT sum = AccT{};
      const T* slot = qpart + (size_t)(simd * 4 * NACC) * 64 + lanee;
#pragma unroll
      for (int qq = 0; qq < 4; ++qq)
#pragma unroll
        for (int r = 0; r < NACC; ++r) sum[r] += slot[(qq * NACC + r) * 64];
      const int ti = 4 * q4 + simd;
      epilogue(sum, ti / ntc, ti - (ti / ntc) * ntc);
    }
    __syncthreads();
  }

  // Tallies in FP64: per-path totals, exceedance count, sum, sum of squares.
  // State entry (r, stored column c), hi + lo in the split form.
  auto state_at = [&](const int r, const int c) -> double {
    if constexpr (TR::kSplit) {
      const uint16_t h = Ah[r * LDA + c], l = Ah[lo_off + r * LDA + c];
      return (double)__builtin_bit_cast(float, (uint32_t)h << 16) +
             (double)__builtin_bit_cast(float, (uint32_t)l << 16);
    } else {
      return (double)A[r * LDA + c];
    }
  };
  double cnt = 0.0, sum = 0.0, sq = 0.0;
  for (int r = tid; r < BM; r += nthreads) {
    const int path = p0 + r;
    if (path < jb.N) {
      double tot = 0.0;
      for (int c = 0; c < Pp; ++c) tot += state_at(r, c);
      if (jb.end_pops) gmem(jb.end_pops)[path] = tot;
      cnt += (tot >= jb.thr_total) ? 1.0 : 0.0;
      sum += tot;
      sq += tot * tot;
    }
  }
  cnt = wave_sum(cnt);
  sum = wave_sum(sum);
  sq = wave_sum(sq);
  if (lane == 0) {
    red[wave * 3 + 0] = cnt;
    red[wave * 3 + 1] = sum;
    red[wave * 3 + 2] = sq;
  }
  __syncthreads();
  RMC_GLOBAL double* part = gmem(jb.partial) + (size_t)tile * (3 + Pp);
  if (tid < 3) {
    double v = 0.0;
    for (int w = 0; w < nwaves; ++w) v += red[w * 3 + tid];
    part[tid] = v;
  }
  const int nvalid = min(BM, jb.N - p0);
  for (int c = tid; c < Pp; c += nthreads) {
    const int pc = TR::perm(c);
    double s = 0.0;
    for (int r = 0; r < nvalid; ++r) s += state_at(r, pc);
    part[3 + c] = s;
  }
  __syncthreads();  // LDS is reused by the next tile
}

template <class TR, int NT>
__device__ __forceinline__ void mte_tile(const MteJob& jb, const int tile, char* smem) {
  switch (jb.law) {
    case kLawRefCpu: mte_tile_law<TR, kLawRefCpu, NT>(jb, tile, smem); break;
    case kLawRefGpu: mte_tile_law<TR, kLawRefGpu, NT>(jb, tile, smem); break;
    default: mte_tile_law<TR, kLawLognormal, NT>(jb, tile, smem); break;
  }
}

template <class TR, int NT>
__global__ __launch_bounds__(kMteThreads) void mte_direct_kernel(
    const MteJob* __restrict__ jobs, const int2* __restrict__ work) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int2 w = work[blockIdx.x];
  mte_tile<TR, NT>(jobs[w.x], w.y, smem);
}

// Persistent form.  The work items are split into nq queues, one per XCD,
// each holding whole jobs with their tiles contiguous, so that the CUs of an
// XCD share one movement matrix in their L2.  Workgroups are dispatched to
// XCDs round robin (workgroup b runs on XCD b % 8), so workgroup b first
// drains queue b % nq and then takes what is left in the others.  Every
// workgroup leaves each queue once its counter passes the queue length.
template <class TR, int NT>
__global__ __launch_bounds__(kMteThreads) void mte_persistent_kernel(
    const MteJob* __restrict__ jobs, const int2* __restrict__ work,
    const int* __restrict__ qbeg, int nq, int* __restrict__ counters) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  __shared__ int next;
  const int home = (int)(blockIdx.x % (unsigned)nq);
  for (int qi = 0; qi < nq; ++qi) {
    const int q = (home + qi) % nq;
    const int base = qbeg[q], len = qbeg[q + 1] - base;
    for (;;) {
      if (threadIdx.x == 0) next = atomicAdd(&counters[q], 1);
      __syncthreads();
      const int w = next;
      __syncthreads();
      if (w >= len) break;
      const int2 it = work[base + w];
      mte_tile<TR, NT>(jobs[it.x], it.y, smem);
    }
  }
}

// Sums the tile partials of each job in a fixed order (deterministic).
__global__ void mte_reduce_kernel(const MteJob* __restrict__ jobs,
                                  double* __restrict__ out,
                                  const int64_t* __restrict__ out_off) {
  const MteJob& jb = jobs[blockIdx.x];
  const int W = 3 + jb.Pp;
  double* o = out + out_off[blockIdx.x];
  for (int e = threadIdx.x; e < W; e += blockDim.x) {
    double s = 0.0;
    // Unrolled so that eight tiles' loads are in flight; the sum keeps its
    // tile order.
#pragma unroll 8
    for (int t = 0; t < jb.n_tiles; ++t) s += gmem(jb.partial)[(size_t)t * W + e];
    o[e] = s;
  }
}

__global__ void fill_i32_kernel(int* p, int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

namespace {

// Kernel forms: FP64 and split-bf16 at their traits' tile bound; FP32 at 1, 2
// or 4 tiles per wave.  direct selects mte_direct_kernel, else the persistent
// kernel.
template <class TR, int NT>
const void* mte_fn(bool direct) {
  return direct ? reinterpret_cast<const void*>(&mte_direct_kernel<TR, NT>)
                : reinterpret_cast<const void*>(&mte_persistent_kernel<TR, NT>);
}

const void* mte_form(int prec, int tiles_per_wave, bool direct) {
  switch (prec) {
    case kPrecF64: return mte_fn<TraitsF64, TraitsF64::NT>(direct);
    case kPrecB3: return mte_fn<TraitsB3, TraitsB3::NT>(direct);
    default:
      if (tiles_per_wave <= 1) return mte_fn<TraitsF32, 1>(direct);
      if (tiles_per_wave <= 2) return mte_fn<TraitsF32, 2>(direct);
      return mte_fn<TraitsF32, TraitsF32::NT>(direct);
  }
}

int mte_max_tiles_per_wave(int prec) {
  return prec == kPrecF64 ? TraitsF64::NT : prec == kPrecB3 ? TraitsB3::NT : TraitsF32::NT;
}

}  // namespace

int mte_tiles_per_wave(int prec, int BM, int Pp) {
  const int tm = prec == kPrecF64 ? TraitsF64::TM : TraitsF32::TM;
  const int tn = prec == kPrecF64 ? TraitsF64::TN : TraitsF32::TN;
  const int nwaves = kMteThreads / 64;
  return ((BM / tm) * (Pp / tn) + nwaves - 1) / nwaves;
}

hipError_t mte_set_lds_limit(int prec, size_t bytes) {
  for (int tpw : {1, 2, 4})
    for (bool direct : {true, false}) {
      const hipError_t e = hipFuncSetAttribute(mte_form(prec, tpw, direct),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)bytes);
      if (e != hipSuccess) return e;
    }
  return hipSuccess;
}

hipError_t launch_mte_direct(int prec, int tiles_per_wave, const MteJob* jobs, const int2* work,
                             int n_work, size_t lds, hipStream_t s) {
  if (n_work <= 0) return hipSuccess;
  if (tiles_per_wave > mte_max_tiles_per_wave(prec)) return hipErrorInvalidValue;
  void* args[] = {&jobs, &work};
  const hipError_t e = hipLaunchKernel(mte_form(prec, tiles_per_wave, true), dim3(n_work),
                                       dim3(kMteThreads), args, lds, s);
  return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launch_mte_persistent(int prec, int tiles_per_wave, const MteJob* jobs,
                                 const int2* work, const int* qbeg, int nq,
                                 int* counters, int grid, size_t lds,
                                 hipStream_t s) {
  if (nq <= 0 || grid <= 0) return hipSuccess;
  if (tiles_per_wave > mte_max_tiles_per_wave(prec)) return hipErrorInvalidValue;
  void* args[] = {&jobs, &work, &qbeg, &nq, &counters};
  const hipError_t e = hipLaunchKernel(mte_form(prec, tiles_per_wave, false), dim3(grid),
                                       dim3(kMteThreads), args, lds, s);
  return e != hipSuccess ? e : hipGetLastError();
}

// In-place conversion of FP32-packed movement matrices to the split-bf16
// operand order.  Each 512-float chunk (one column block, k-groups 2c and
// 2c + 1: 16 k x 32 columns) is converted on its own.  In the FP32 order,
// k = 16c + 8g + 2s + h of column n sits at float (64 g + 32 h + n) * 4 + s
// of the chunk; in the split order lane l = n + 32g holds k = 16c + 8g + i,
// i = 0..7, as 8 bf16 hi values at bytes 16 l and 8 bf16 lo values at bytes
// 1024 + 16 l.  One wave per chunk reads all of it before writing.
__global__ __launch_bounds__(64) void mte_split_b3_kernel(float* __restrict__ m) {
  float* c = m + (size_t)blockIdx.x * 512;
  const int l = threadIdx.x, n = l & 31, g = l >> 5;
  float v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = c[(64 * g + 32 * (i & 1) + n) * 4 + (i >> 1)];
  __syncthreads();
  uint32_t hi[4], lo[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t a = TraitsB3::split(v[2 * j]), b = TraitsB3::split(v[2 * j + 1]);
    hi[j] = (a & 0xffffu) | (b << 16);
    lo[j] = (a >> 16) | (b & 0xffff0000u);
  }
  uint4* out = reinterpret_cast<uint4*>(c);
  out[l] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
  out[64 + l] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
}

hipError_t launch_mte_split_b3(float* m, size_t n_floats, hipStream_t s) {
  if (n_floats % 512 != 0) return hipErrorInvalidValue;
  if (n_floats == 0) return hipSuccess;
  hipLaunchKernelGGL(mte_split_b3_kernel, dim3((unsigned)(n_floats / 512)), dim3(64), 0, s, m);
  return hipGetLastError();
}

hipError_t launch_mte_reduce(const MteJob* jobs, int n_jobs, double* out,
                             const int64_t* out_off, hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(mte_reduce_kernel, dim3(n_jobs), dim3(256), 0, s, jobs,
                     out, out_off);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int* p, int value, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                     p, value, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Reference-design baseline.  These kernels keep the structure of mteKernel
// (SimulateGPU.cu:595-717) and of its host driver (SimulateGPU.cu:3720-3843):
// the random streams are written to HBM first (there by curandGenerate*), then
// one thread per path runs the years with its state in a path-major
// [N][2][P] ping-pong buffer and a serial CSC loop.  They draw the same
// Philox numbers as the tiled kernels.

template <typename T>
__global__ void gen_patch_normals_kernel(T* __restrict__ out, int N, int Ty,
                                         int P, uint32_t tag, uint32_t k0,
                                         uint32_t k1) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nq = (N + 3) / 4;
  if (idx >= nq * Ty * P) return;
  const int j = (int)(idx % P);
  const int64_t rest = idx / P;
  const int t = (int)(rest % Ty);
  const int q = (int)(rest / Ty);
  T z[4];
  if constexpr (sizeof(T) == 4)
    rmc_normal4_f((uint32_t)q, (uint32_t)t, (uint32_t)j, tag, k0, k1, z);
  else
    rmc_normal4_d((uint32_t)q, (uint32_t)t, (uint32_t)j, tag, k0, k1, z);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int path = 4 * q + s;
    if (path < N) out[((int64_t)path * Ty + t) * P + j] = z[s];
  }
}

template <typename T>
__global__ void gen_path_scalars_kernel(T* __restrict__ brown,
                                        T* __restrict__ jsize,
                                        T* __restrict__ jump, int N, int Ty,
                                        uint32_t tag, uint32_t k0,
                                        uint32_t k1) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)N * Ty) return;
  const int t = (int)(idx % Ty);
  const int path = (int)(idx / Ty);
  T zb, zj, uj;
  path_scalars<T>((uint32_t)path, (uint32_t)t, tag, k0, k1, &zb, &zj, &uj);
  brown[idx] = zb;
  jsize[idx] = zj;
  jump[idx] = uj;
}

template <typename T>
__global__ void mte_plain_kernel(PlainArgs a) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  if (ii >= a.N) return;
  const int P = a.P, Ty = a.T;
  const T* rgr = static_cast<const T*>(a.rgr);
  const T* prm = static_cast<const T*>(a.params);
  const T* init = static_cast<const T*>(a.init);
  const T* caps = static_cast<const T*>(a.caps);
  const T* rates = static_cast<const T*>(a.rates);
  const T* val = static_cast<const T*>(a.val);
  T* cur = static_cast<T*>(a.path_pops) + (size_t)ii * 2 * P;
  T* nxt = cur + P;
  const T dt = (T)a.dt;
  const T l2e = sizeof(T) == 4 ? (T)1.4426950408889634 : (T)1;
  const T sig = (T)(a.sigma * (sizeof(T) == 4 ? 1.4426950408889634 : 1.0));
  const T half_var = (T)(0.5 * a.sigma * a.sigma) * l2e;
  for (int j = 0; j < P; ++j) cur[j] = init[j];
  T grm = prm[0];
  for (int t = 0; t < Ty; ++t) {
    for (int k = 0; k < P; ++k) nxt[k] = (T)0;
    int it = 0;
    for (int k = 0; k < P; ++k) {
      for (int l = 0; l < a.per_col[k]; ++l) {
        nxt[k] += cur[a.row_idx[it]] * val[it];
        ++it;
      }
    }
    if (a.law == kLawRefGpu) {
      const size_t ix = (size_t)ii * Ty + t;
      const T u = static_cast<const T*>(a.jump)[ix];
      const T jump = (u < prm[6]) ? (T)1 : (T)0;
      const T brown = static_cast<const T*>(a.brown)[ix] * prm[2];
      const T jsz = static_cast<const T*>(a.jsize)[ix] * prm[5] * prm[5] -
                    prm[5] * prm[5] / (T)2;
      T e;
      if constexpr (sizeof(T) == 4)
        e = __expf(jsz);
      else
        e = exp(jsz);
      grm = grm + prm[4] * (prm[1] - grm) * dt + grm * brown +
            (e - (T)1) * grm * jump;
    }
    for (int k = 0; k < P; ++k) {
      const T z = rgr[((size_t)ii * Ty + t) * P + k];
      const T x = nxt[k];
      const T Ki = (T)1 / caps[k];
      if (a.law == kLawRefCpu)
        cur[k] = grow<T, kLawRefCpu>(x, z, caps[k], Ki, rates[k], sig, half_var, dt, prm[1],
                                     prm[2], grm, prm[7]);
      else if (a.law == kLawRefGpu)
        cur[k] = grow<T, kLawRefGpu>(x, z, caps[k], Ki, rates[k], sig, half_var, dt, prm[1],
                                     prm[2], grm, prm[7]);
      else
        cur[k] = grow<T, kLawLognormal>(x, z, caps[k], Ki, rates[k], sig, half_var, dt,
                                        prm[1], prm[2], grm, prm[7]);
    }
  }
  const T* v = (a.endpoint == kEndPre) ? nxt : cur;
  T tot = (T)0;
  for (int j = 0; j < P; ++j) tot += v[j];
  static_cast<T*>(a.eps)[ii] = tot;
}

hipError_t launch_gen_patch_normals(bool fp64, void* out, int N, int T, int P,
                                    uint32_t tag, uint32_t k0, uint32_t k1,
                                    hipStream_t s) {
  const int64_t total = (int64_t)((N + 3) / 4) * T * P;
  if (total <= 0) return hipSuccess;
  const int64_t blocks = (total + 255) / 256;
  if (fp64)
    hipLaunchKernelGGL(gen_patch_normals_kernel<double>, dim3((unsigned)blocks),
                       dim3(256), 0, s, static_cast<double*>(out), N, T, P,
                       tag, k0, k1);
  else
    hipLaunchKernelGGL(gen_patch_normals_kernel<float>, dim3((unsigned)blocks),
                       dim3(256), 0, s, static_cast<float*>(out), N, T, P, tag,
                       k0, k1);
  return hipGetLastError();
}

hipError_t launch_gen_path_scalars(bool fp64, void* brown, void* jsize,
                                   void* jump, int N, int T, uint32_t tag,
                                   uint32_t k0, uint32_t k1, hipStream_t s) {
  const int64_t total = (int64_t)N * T;
  if (total <= 0) return hipSuccess;
  const int64_t blocks = (total + 255) / 256;
  if (fp64)
    hipLaunchKernelGGL(gen_path_scalars_kernel<double>, dim3((unsigned)blocks),
                       dim3(256), 0, s, static_cast<double*>(brown),
                       static_cast<double*>(jsize), static_cast<double*>(jump),
                       N, T, tag, k0, k1);
  else
    hipLaunchKernelGGL(gen_path_scalars_kernel<float>, dim3((unsigned)blocks),
                       dim3(256), 0, s, static_cast<float*>(brown),
                       static_cast<float*>(jsize), static_cast<float*>(jump),
                       N, T, tag, k0, k1);
  return hipGetLastError();
}

hipError_t launch_mte_plain(bool fp64, const PlainArgs& a, hipStream_t s) {
  if (a.N <= 0) return hipSuccess;
  // The reference launches min(N, maxThreadsPerBlock) threads per block
  // (SimulateGPU.cu:3824-3833).
  const int tpb = a.N < 1024 ? a.N : 1024;
  const int blocks = (a.N + tpb - 1) / tpb;
  if (fp64)
    hipLaunchKernelGGL(mte_plain_kernel<double>, dim3(blocks), dim3(tpb), 0, s,
                       a);
  else
    hipLaunchKernelGGL(mte_plain_kernel<float>, dim3(blocks), dim3(tpb), 0, s,
                       a);
  return hipGetLastError();
}

}  // namespace roadmc
