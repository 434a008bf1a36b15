// Device building blocks shared by the MTE and adjoint kernels: the MFMA
// traits (operand and accumulator lane layouts of the two matrix
// instructions used), the k-permuted LDS layout, the packed B-operand order,
// the growth law and the per-year product loop.
#ifndef ROADMC_MTE_COMMON_H
#define ROADMC_MTE_COMMON_H

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mte_kernels.h"
#include "rng.h"

namespace roadmc {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex2 __attribute__((ext_vector_type(2)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

enum { kLawLognormal = 0, kLawRefCpu = 1, kLawRefGpu = 2 };
enum { kEndPost = 0, kEndPre = 1 };

// v_mfma_f32_32x32x2_f32: A lane l = A[l%32][l/32]; B lane l = B[l/32][l%32];
// D item r of lane l = D[8(r/4) + 4(l/32) + r%4][l%32].
struct TraitsF32 {
  using T = float;
  using VecT = floatx4;
  using AccT = floatx16;
  static constexpr int TM = 32, TN = 32, KV = 4, NACC = 16, NT = 4, PAD = 4;
  static constexpr bool kQuadItems = true;  // items 4q..4q+3 are 4 paths
  static constexpr bool kSplit = false;
  __device__ static AccT mfma(float a, float b, AccT c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  __device__ static int acc_row(int lane, int r) {
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
  }
  __device__ static int acc_col(int lane) { return lane & 31; }
  __device__ static int a_row(int lane) { return lane & 31; }
  __device__ static int k_lane(int lane) { return lane >> 5; }
  // k = 8g + 2s + h  ->  8g + 4h + s
  __host__ __device__ static int perm(int k) {
    return (k & ~7) | ((k & 1) << 2) | ((k >> 1) & 3);
  }
  __device__ static float expo(float x) { return __expf(x); }
  __device__ static void normal4(uint32_t q, uint32_t y, uint32_t p,
                                 uint32_t tag, uint32_t k0, uint32_t k1,
                                 float z[4]) {
    rmc_normal4_f(q, y, p, tag, k0, k1, z);
  }
};

// v_mfma_f64_16x16x4_f64: A lane l = A[l%16][l/16]; B lane l = B[l/16][l%16];
// D item r of lane l = D[l/16 + 4r][l%16].
struct TraitsF64 {
  using T = double;
  using VecT = doublex2;
  using AccT = doublex4;
  static constexpr int TM = 16, TN = 16, KV = 2, NACC = 4, NT = 8, PAD = 2;
  static constexpr bool kQuadItems = false;
  static constexpr bool kSplit = false;
  __device__ static AccT mfma(double a, double b, AccT c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int acc_row(int lane, int r) { return (lane >> 4) + 4 * r; }
  __device__ static int acc_col(int lane) { return lane & 15; }
  __device__ static int a_row(int lane) { return lane & 15; }
  __device__ static int k_lane(int lane) { return lane >> 4; }
  // k = 8g + 4s + h  ->  8g + 2h + s
  __host__ __device__ static int perm(int k) {
    return (k & ~7) | ((k & 3) << 1) | ((k >> 2) & 1);
  }
  __device__ static double expo(double x) { return exp(x); }
  __device__ static void normal4(uint32_t q, uint32_t y, uint32_t p,
                                 uint32_t tag, uint32_t k0, uint32_t k1,
                                 double z[4]) {
    rmc_normal4_d(q, y, p, tag, k0, k1, z);
  }
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Split-bf16 products on v_mfma_f32_32x32x16_bf16 (FP32 accumulate).  A value
// x is held as the pair hi = bf16(x), lo = bf16(x - hi), which carries about
// 17 significant bits; a product x.m is hi.m_hi + hi.m_lo + lo.m_hi (the
// dropped lo.m_lo is below 2^-16 of it).  The population tile lives in LDS as
// two bf16 arrays (hi, then lo), k in natural order; a lane's operand is the
// 8 consecutive k of its row and lane half: A lane l = A[l%32][8(l/32) + i],
// B lane l = B[8(l/32) + i][l%32], i = 0..7 (one 16-byte vector).  The
// accumulator layout is the one of v_mfma_f32_32x32x2_f32.
struct TraitsB3 {
  using T = float;
  using VecT = floatx4;  // 16 bytes: 8 bf16 operands
  using AccT = floatx16;
  // Each bf16 row has Pp + PAD slots (16 bytes of padding), so the hi and lo
  // arrays together take BM (Pp + PAD) 4-byte units, as lds_bytes_t counts.
  static constexpr int TM = 32, TN = 32, KV = 1, NACC = 16, NT = 4, PAD = 8;
  static constexpr bool kQuadItems = true;
  static constexpr bool kSplit = true;
  __device__ static AccT mfma(floatx4 a, floatx4 b, AccT c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                    __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  }
  __device__ static int acc_row(int lane, int r) {
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
  }
  __device__ static int acc_col(int lane) { return lane & 31; }
  __device__ static int a_row(int lane) { return lane & 31; }
  __device__ static int k_lane(int lane) { return lane >> 5; }
  __host__ __device__ static int perm(int k) { return k; }
  __device__ static float expo(float x) { return __expf(x); }
  __device__ static void normal4(uint32_t q, uint32_t y, uint32_t p,
                                 uint32_t tag, uint32_t k0, uint32_t k1,
                                 float z[4]) {
    rmc_normal4_f(q, y, p, tag, k0, k1, z);
  }
  // The split pair of x, hi in the low half.
  __device__ static uint32_t split(float x) {
    const __bf16 h = (__bf16)x;
    const __bf16 l = (__bf16)(x - (float)h);
    return (uint32_t)__builtin_bit_cast(uint16_t, h) |
           ((uint32_t)__builtin_bit_cast(uint16_t, l) << 16);
  }
};

__host__ __device__ static inline size_t align16(size_t x) {
  return (x + 15) & ~(size_t)15;
}

// LDS of one MTE workgroup: the BM x (Pp + PAD) population tile, the per-row
// growth-rate means, 8 x 4 doubles for the tally reduction and, when
// remainder tiles are split into k-quarters, 4 partial accumulators per
// remainder tile.
template <class TR>
__host__ __device__ static inline size_t lds_bytes_t(int BM, int Pp, int n_rem_split) {
  using T = typename TR::T;
  return align16((size_t)BM * (Pp + TR::PAD) * sizeof(T)) +
         align16((size_t)BM * sizeof(T)) + 8 * 4 * sizeof(double) +
         (size_t)n_rem_split * 4 * TR::NACC * 64 * sizeof(T);
}

template <typename T, int LAW>
__device__ __forceinline__ T grow(T x, T z, T K, T Kinv, T r, T sig,
                                  T half_var, T dt, T mu, T sd, T grm, T lv) {
  if constexpr (LAW == kLawRefCpu) {
    const T g = dt * (mu + sd * z) * (T)0.01;
    return x * ((T)1 + g * (K - x) * Kinv);
  }
  if constexpr (LAW == kLawRefGpu) {
    const T g = lv * z * grm + grm;
    return x * ((T)1 + g * (K - x) * Kinv);
  }
  T e;
  if constexpr (sizeof(T) == 4)
    // sig and half_var arrive pre-scaled by log2(e): one FMA and v_exp_f32.
    e = __builtin_amdgcn_exp2f(fmaf(sig, z, -half_var));
  else
    e = exp(sig * z - half_var);
  return x * ((T)1 + r * ((T)1 - x * Kinv)) * e;
}

template <typename T>
__device__ __forceinline__ void path_scalars(uint32_t path, uint32_t year,
                                             uint32_t tag, uint32_t k0,
                                             uint32_t k1, T* zb, T* zj,
                                             T* uj) {
  const rmc_u4 x = rmc_philox(path, year, 0u, tag, k0, k1);
  if constexpr (sizeof(T) == 4) {
    rmc_box_muller_f(rmc_u01f(x.x), rmc_u01f(x.y), zb, zj);
    *uj = rmc_u01f(x.z);
  } else {
    rmc_box_muller_d(rmc_u01d(x.x), rmc_u01d(x.y), zb, zj);
    *uj = rmc_u01d(x.z);
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Global-address-space view of a device pointer, so that loads compile to
// global_load (not flat_load).
template <typename V>
__device__ __forceinline__ const __attribute__((address_space(1))) V* as_global(const V* p) {
  return (const __attribute__((address_space(1))) V*)(p);
}

// Movement matrix of one job in MFMA B-operand order, read through a buffer
// resource: the k-group step goes into the scalar offset and the tile's
// column block into a 32-bit vector offset, so operand loads need no vector
// address arithmetic.  Reads past `bytes` return zero instead of faulting.
struct PackedB {
  __amdgpu_buffer_rsrc_t rsrc;
  int nkg;  // k-groups per column block
};

__device__ __forceinline__ PackedB packed_b(const void* p, int nkg, size_t bytes) {
  PackedB b;
  // Word 3 of the descriptor: raw 32-bit data format of the gfx9 family.
  b.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
  b.nkg = nkg;
  return b;
}

// Byte offset of lane `lane`'s 16-byte operand in column block `tcol`.
__device__ __forceinline__ uint32_t b_voff(const PackedB& B, int tcol, int lane) {
  return ((uint32_t)tcol * (uint32_t)B.nkg * 64u + (uint32_t)lane) * 16u;
}

template <class VecT>
__device__ __forceinline__ VecT load_b(const PackedB& B, uint32_t voff, int kg) {
  static_assert(sizeof(VecT) == 16, "B operands are 16-byte vectors");
  return __builtin_bit_cast(
      VecT, __builtin_amdgcn_raw_buffer_load_b128(B.rsrc, (int)voff, kg * 64 * 16, 0));
}

// One year's product for the NTILE output tiles of a wave:
// acc[t] = A[rows of tile t, :] . M[:, cols of tile t].
// Two register stages ping-pong over the k-groups (the k-group count B.nkg is
// even): the B operands and LDS A operands of k-group kg+1 are in flight while
// the MFMAs of k-group kg run.
// On entry b0 holds the B operands of k-group 0; on exit it holds them again
// (the last stage-0 load re-fetches k-group 0, which the next year starts
// with, so that load's latency is hidden behind the epilogue).
template <class TR, int NTILE, bool SHARED_A = false>
__device__ __forceinline__ void year_product(typename TR::AccT* acc,
                                             typename TR::VecT* b0,
                                             const PackedB& B,
                                             const typename TR::T* A,
                                             const int* trow, const int* tcol,
                                             int LDA, int arow, int akoff, int lane) {
  using T = typename TR::T;
  using VecT = typename TR::VecT;
  using AccT = typename TR::AccT;
  constexpr int TM = TR::TM, KV = TR::KV;
  // With SHARED_A all NTILE tiles lie in one row block: one A fragment per
  // k-group serves every tile.
  constexpr int NA = SHARED_A ? 1 : NTILE;
  const int nkg = B.nkg;
  uint32_t boff[NTILE];
  const T* ap[NA];
#pragma unroll
  for (int t = 0; t < NTILE; ++t) boff[t] = b_voff(B, tcol[t], lane);
#pragma unroll
  for (int t = 0; t < NA; ++t) ap[t] = A + (trow[t] * TM + arow) * LDA + akoff;
  auto ldb = [&](int t, int kg) -> VecT { return load_b<VecT>(B, boff[t], kg); };
  auto ia = [](int t) { return SHARED_A ? 0 : t; };
  VecT b1[NTILE], a0[NA], a1[NA];
#pragma unroll
  for (int t = 0; t < NA; ++t) a0[t] = *reinterpret_cast<const VecT*>(ap[t]);
  // Two k-groups per step.  FIRST: the step's first MFMA per tile starts
  // from a zero accumulator (an inline constant, no register clearing).
  auto step = [&](const int kg, auto first) {
    // stage 1 loads (k-group kg+1 < nkg always: nkg is even)
#pragma unroll
    for (int t = 0; t < NTILE; ++t) b1[t] = ldb(t, kg + 1);
#pragma unroll
    for (int t = 0; t < NA; ++t) a1[t] = *reinterpret_cast<const VecT*>(ap[t] + 8);
    // Keep the loads ahead of the MFMAs that do not need them: the scheduler
    // would otherwise sink them next to their use and expose their latency.
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < NTILE; ++t) {
      if constexpr (decltype(first)::value)
        acc[t] = TR::mfma(a0[ia(t)][0], b0[t][0], AccT{});
      else
        acc[t] = TR::mfma(a0[ia(t)][0], b0[t][0], acc[t]);
#pragma unroll
      for (int s = 1; s < KV; ++s) acc[t] = TR::mfma(a0[ia(t)][s], b0[t][s], acc[t]);
    }
    // stage 0 loads for k-group kg+2.  After the last step b0 re-fetches
    // k-group 0, which the next year starts with; a0 then reads the 8 values
    // past column Pk - 1 of its row (inside the LDS allocation, never used:
    // the next call reloads a0), which keeps its address an immediate offset.
    const int k2 = (kg + 2 < nkg) ? kg + 2 : 0;
#pragma unroll
    for (int t = 0; t < NTILE; ++t) b0[t] = ldb(t, k2);
#pragma unroll
    for (int t = 0; t < NA; ++t) {
      a0[t] = *reinterpret_cast<const VecT*>(ap[t] + 16);
      ap[t] += 16;
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < NTILE; ++t) {
#pragma unroll
      for (int s = 0; s < KV; ++s) acc[t] = TR::mfma(a1[ia(t)][s], b1[t][s], acc[t]);
    }
  };
  step(0, std::true_type{});
  for (int kg = 2; kg < nkg; kg += 2) step(kg, std::false_type{});
}

// acc += A[rows of tile (trow, tcol), k-groups kg0 .. kg1) . M[those k-groups,
// cols of the tile] for one output tile, kg1 > kg0, with the same two-stage
// ping-pong as year_product (the operands of the next k-group are in flight
// while the MFMAs of the current one run) and an odd tail step.
template <class TR>
__device__ __forceinline__ void range_product(typename TR::AccT& acc, const PackedB& B,
                                              const typename TR::T* A, int trow,
                                              int tcol, int kg0, int kg1,
                                              int LDA, int arow, int akoff, int lane) {
  using T = typename TR::T;
  using VecT = typename TR::VecT;
  constexpr int TM = TR::TM, KV = TR::KV;
  const uint32_t boff = b_voff(B, tcol, lane);
  const T* ap = A + (trow * TM + arow) * LDA + akoff;
  VecT b0 = load_b<VecT>(B, boff, kg0);
  VecT a0 = *reinterpret_cast<const VecT*>(ap + kg0 * 8);
  int kg = kg0;
  for (; kg + 1 < kg1; kg += 2) {
    const VecT b1 = load_b<VecT>(B, boff, kg + 1);
    const VecT a1 = *reinterpret_cast<const VecT*>(ap + (kg + 1) * 8);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < KV; ++s) acc = TR::mfma(a0[s], b0[s], acc);
    const int k2 = min(kg + 2, kg1 - 1);
    b0 = load_b<VecT>(B, boff, k2);
    a0 = *reinterpret_cast<const VecT*>(ap + k2 * 8);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < KV; ++s) acc = TR::mfma(a1[s], b1[s], acc);
  }
  if (kg < kg1) {
#pragma unroll
    for (int s = 0; s < KV; ++s) acc = TR::mfma(a0[s], b0[s], acc);
  }
}

// Split-bf16 form of year_product (TraitsB3).  A k-chunk is 16 k: the A
// operands are 8 bf16 of the hi and of the lo array (Ah and Ah + lo_off), the
// B operands packed k-groups 2c (hi) and 2c + 1 (lo).  Each chunk issues the
// three products of every tile, one product kind at a time across the tiles,
// so that consecutive MFMAs are independent.  Stages ping-pong over the chunks
// (the chunk count is even).  Unlike year_product, no operands stay in
// registers between years: the epilogue needs them.
template <int NTILE, bool SHARED_A = false>
__device__ __forceinline__ void year_product_b3(floatx16* acc, const PackedB& B,
                                                const uint16_t* Ah,
                                                int lo_off, const int* trow, const int* tcol,
                                                int LDA, int arow, int akoff, int lane) {
  constexpr int NA = SHARED_A ? 1 : NTILE;
  const int nch = B.nkg >> 1;
  uint32_t boff[NTILE];
  const uint16_t* ap[NA];
#pragma unroll
  for (int t = 0; t < NTILE; ++t) boff[t] = b_voff(B, tcol[t], lane);
#pragma unroll
  for (int t = 0; t < NA; ++t) ap[t] = Ah + (trow[t] * 32 + arow) * LDA + akoff;
  auto ldb = [&](int t, int kg) -> floatx4 { return load_b<floatx4>(B, boff[t], kg); };
  auto lda = [&](int t, int off) -> floatx4 {
    return *reinterpret_cast<const floatx4*>(ap[t] + off);
  };
  auto ia = [](int t) { return SHARED_A ? 0 : t; };
  floatx4 bh0[NTILE], bl0[NTILE], bh1[NTILE], bl1[NTILE], ah0[NA], al0[NA], ah1[NA], al1[NA];
#pragma unroll
  for (int t = 0; t < NTILE; ++t) {
    bh0[t] = ldb(t, 0);
    bl0[t] = ldb(t, 1);
  }
#pragma unroll
  for (int t = 0; t < NA; ++t) {
    ah0[t] = lda(t, 0);
    al0[t] = lda(t, lo_off);
  }
  auto products = [&](const floatx4* ah, const floatx4* al, const floatx4* bh,
                      const floatx4* bl, auto first) {
#pragma unroll
    for (int t = 0; t < NTILE; ++t) {
      if constexpr (decltype(first)::value)
        acc[t] = TraitsB3::mfma(ah[ia(t)], bh[t], floatx16{});
      else
        acc[t] = TraitsB3::mfma(ah[ia(t)], bh[t], acc[t]);
    }
#pragma unroll
    for (int t = 0; t < NTILE; ++t) acc[t] = TraitsB3::mfma(ah[ia(t)], bl[t], acc[t]);
#pragma unroll
    for (int t = 0; t < NTILE; ++t) acc[t] = TraitsB3::mfma(al[ia(t)], bh[t], acc[t]);
  };
  auto step = [&](const int c, auto first) {
#pragma unroll
    for (int t = 0; t < NTILE; ++t) {
      bh1[t] = ldb(t, 2 * c + 2);
      bl1[t] = ldb(t, 2 * c + 3);
    }
#pragma unroll
    for (int t = 0; t < NA; ++t) {
      ah1[t] = lda(t, 16);
      al1[t] = lda(t, lo_off + 16);
    }
    __builtin_amdgcn_sched_barrier(0);
    products(ah0, al0, bh0, bl0, first);
    // The last step's stage-0 loads read chunk 0 again (B) and past the
    // row's last chunk (A: inside the LDS allocation); both are never used.
    const int c2 = (c + 2 < nch) ? c + 2 : 0;
#pragma unroll
    for (int t = 0; t < NTILE; ++t) {
      bh0[t] = ldb(t, 2 * c2);
      bl0[t] = ldb(t, 2 * c2 + 1);
    }
#pragma unroll
    for (int t = 0; t < NA; ++t) {
      ah0[t] = lda(t, 32);
      al0[t] = lda(t, lo_off + 32);
      ap[t] += 32;
    }
    __builtin_amdgcn_sched_barrier(0);
    products(ah1, al1, bh1, bl1, std::false_type{});
  };
  step(0, std::true_type{});
  for (int c = 2; c < nch; c += 2) step(c, std::false_type{});
}

// acc += the split-bf16 product over k-chunks c0 .. c1 (c1 > c0) of one tile,
// with range_product's ping-pong and odd tail.
__device__ __forceinline__ void range_product_b3(floatx16& acc, const PackedB& B,
                                                 const uint16_t* Ah, int lo_off, int trow,
                                                 int tcol, int c0, int c1, int LDA, int arow,
                                                 int akoff, int lane) {
  const uint32_t boff = b_voff(B, tcol, lane);
  const uint16_t* ap = Ah + (trow * 32 + arow) * LDA + akoff;
  auto lda = [&](int off) -> floatx4 { return *reinterpret_cast<const floatx4*>(ap + off); };
  floatx4 bh0 = load_b<floatx4>(B, boff, 2 * c0), bl0 = load_b<floatx4>(B, boff, 2 * c0 + 1);
  floatx4 ah0 = lda(16 * c0), al0 = lda(lo_off + 16 * c0);
  auto products = [&](floatx4 ah, floatx4 al, floatx4 bh, floatx4 bl) {
    acc = TraitsB3::mfma(ah, bh, acc);
    acc = TraitsB3::mfma(ah, bl, acc);
    acc = TraitsB3::mfma(al, bh, acc);
  };
  int c = c0;
  for (; c + 1 < c1; c += 2) {
    const floatx4 bh1 = load_b<floatx4>(B, boff, 2 * c + 2);
    const floatx4 bl1 = load_b<floatx4>(B, boff, 2 * c + 3);
    const floatx4 ah1 = lda(16 * (c + 1)), al1 = lda(lo_off + 16 * (c + 1));
    __builtin_amdgcn_sched_barrier(0);
    products(ah0, al0, bh0, bl0);
    const int k2 = min(c + 2, c1 - 1);
    bh0 = load_b<floatx4>(B, boff, 2 * k2);
    bl0 = load_b<floatx4>(B, boff, 2 * k2 + 1);
    ah0 = lda(16 * k2);
    al0 = lda(lo_off + 16 * k2);
    __builtin_amdgcn_sched_barrier(0);
    products(ah1, al1, bh1, bl1);
  }
  if (c < c1) products(ah0, al0, bh0, bl0);
}

}  // namespace roadmc

#endif
