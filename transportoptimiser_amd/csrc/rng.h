/*
 * Counter-based random streams shared by every host and device path of
 * libroadmc.
 *
 * The reference draws its Monte Carlo noise from cuRAND's XORWOW generator,
 * pre-generated into HBM and seeded from time(NULL)
 * (include/gpuCode/SimulateGPU.cu:3732-3751), and the CPU path draws from a
 * static std::default_random_engine (include/transportbase.h:144,
 * include/main/Simulator.cpp:712-722).  Neither stream can be reproduced, so
 * this library replaces them with Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11) keyed by an explicit
 * 64-bit seed.  Because the generator is counter based, a kernel computes the
 * number for (path, year, patch) where it needs it instead of reading a
 * pre-generated array, and every implementation (fused kernel, plain kernel,
 * stream dump) produces the same numbers.
 *
 * Stream contract (all indices unsigned 32-bit):
 *   key      = { seed & 0xffffffff, seed >> 32 }
 *   tag      = (road_id << 8) | (species << 2) | kind        road_id < 2^24
 *   kind 0   patch noise:   z(path, year, patch) =
 *            normal4(philox({path >> 2, year, patch, tag}))[path & 3]
 *   kind 1   path scalars (growth-rate SDE of SimulateGPU.cu:691-701):
 *            x = philox({path, year, 0, tag});
 *            (z_brownian, z_jumpsize) = box_muller(u(x0), u(x1)),
 *            u_jump = u(x2)
 *   kind 2   ROV streams, laid out in rov_kernels.hip.
 *   kind 3   ore composition of the species-state ROV: z(path, year, c) =
 *            normal4(philox({path, year, c >> 2, tag}))[c & 3].
 *   u(x)     = ((x >> 9) + 0.5) * 2^-23, exact in binary32, in (0, 1).
 *   normal4  = two Box-Muller pairs: (u0, u1) -> (r cos 2 pi u1, r sin 2 pi u1),
 *              r = sqrt(-2 ln u0); (u2, u3) likewise.
 *
 * The header compiles as C (gcc, for host tools) and as HIP (host and
 * device).
 */
#ifndef ROADMC_RNG_H
#define ROADMC_RNG_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RMC_HD __host__ __device__ __forceinline__
#define RMC_HD_STATIC static __host__ __device__ __forceinline__
#else
#include <math.h>
#define RMC_HD static inline
#define RMC_HD_STATIC static inline
#endif

#define RMC_PHILOX_M0 0xD2511F53u
#define RMC_PHILOX_M1 0xCD9E8D57u
#define RMC_PHILOX_W0 0x9E3779B9u
#define RMC_PHILOX_W1 0xBB67AE85u

#define RMC_KIND_PATCH 0u
#define RMC_KIND_PATH 1u
#define RMC_KIND_ROV 2u
#define RMC_KIND_ORE 3u

typedef struct {
  uint32_t x, y, z, w;
} rmc_u4;

/* a ^ b ^ c: one v_bitop3_b32 (truth table 0x96) in gfx950 device code. */
RMC_HD_STATIC uint32_t rmc_xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

/* Philox4x32 with 10 rounds: round r uses key + r*(W0, W1). */
RMC_HD_STATIC rmc_u4 rmc_philox(uint32_t c0, uint32_t c1, uint32_t c2,
                                uint32_t c3, uint32_t k0, uint32_t k1) {
#if defined(__HIPCC__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)RMC_PHILOX_M0 * (uint64_t)c0;
    const uint64_t p1 = (uint64_t)RMC_PHILOX_M1 * (uint64_t)c2;
    const uint32_t n0 = rmc_xor3((uint32_t)(p1 >> 32), c1, k0);
    const uint32_t n2 = rmc_xor3((uint32_t)(p0 >> 32), c3, k1);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += RMC_PHILOX_W0;
    k1 += RMC_PHILOX_W1;
  }
  rmc_u4 out;
  out.x = c0;
  out.y = c1;
  out.z = c2;
  out.w = c3;
  return out;
}

RMC_HD_STATIC uint32_t rmc_tag(uint32_t road_id, uint32_t species,
                               uint32_t kind) {
  return (road_id << 8) | ((species & 63u) << 2) | (kind & 3u);
}

/* Uniform in (0,1): 23 random bits plus a half-step offset, exact in float. */
RMC_HD_STATIC float rmc_u01f(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  // (m + 1/2) 2^-23 as one FMA; exact, so equal to the expression below.
  return __builtin_fmaf((float)(x >> 9), 1.0f / 8388608.0f, 1.0f / 16777216.0f);
#else
  return ((float)(x >> 9) + 0.5f) * (1.0f / 8388608.0f);
#endif
}

RMC_HD_STATIC double rmc_u01d(uint32_t x) {
  return ((double)(x >> 9) + 0.5) * (1.0 / 8388608.0);
}

#if defined(__HIPCC__)
/* Device Box-Muller in binary32: 2 normals from 2 uniforms. */
__device__ __forceinline__ void rmc_box_muller_f(float u0, float u1, float* z0,
                                                 float* z1) {
  // v_log_f32 is log2; v_sqrt_f32 is the 1-ulp hardware square root.
  const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u0));
  // v_sin_f32 / v_cos_f32 take their argument in revolutions: sin(2 pi u1)
  // for u1 in (0, 1), no range reduction needed.
  *z0 = r * __builtin_amdgcn_cosf(u1);
  *z1 = r * __builtin_amdgcn_sinf(u1);
}

__device__ __forceinline__ void rmc_box_muller_d(double u0, double u1,
                                                 double* z0, double* z1) {
  const double r = sqrt(-2.0 * log(u0));
  double s, c;
  sincospi(2.0 * u1, &s, &c);
  *z0 = r * c;
  *z1 = r * s;
}

/* Four normals for the patch-noise stream at (path quad, year, patch). */
__device__ __forceinline__ void rmc_normal4_f(uint32_t quad, uint32_t year,
                                              uint32_t patch, uint32_t tag,
                                              uint32_t k0, uint32_t k1,
                                              float z[4]) {
  const rmc_u4 x = rmc_philox(quad, year, patch, tag, k0, k1);
  rmc_box_muller_f(rmc_u01f(x.x), rmc_u01f(x.y), &z[0], &z[1]);
  rmc_box_muller_f(rmc_u01f(x.z), rmc_u01f(x.w), &z[2], &z[3]);
}

__device__ __forceinline__ void rmc_normal4_d(uint32_t quad, uint32_t year,
                                              uint32_t patch, uint32_t tag,
                                              uint32_t k0, uint32_t k1,
                                              double z[4]) {
  const rmc_u4 x = rmc_philox(quad, year, patch, tag, k0, k1);
  rmc_box_muller_d(rmc_u01d(x.x), rmc_u01d(x.y), &z[0], &z[1]);
  rmc_box_muller_d(rmc_u01d(x.z), rmc_u01d(x.w), &z[2], &z[3]);
}
#endif

#endif /* ROADMC_RNG_H */
