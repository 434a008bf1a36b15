/*
 * CPU oracle, MTE part.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Philox4x32-10 (Salmon et al., SC'11): ten rounds of two 32x32->64 products,
 * key bumped by the Weyl constants between rounds. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  int round;
  for (round = 0; round < 10; ++round) {
    uint64_t prod0, prod1;
    uint32_t t0, t1, t2, t3;
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    prod0 = (uint64_t)0xD2511F53u * c0;
    prod1 = (uint64_t)0xCD9E8D57u * c2;
    t0 = (uint32_t)(prod1 >> 32) ^ c1 ^ k0;
    t1 = (uint32_t)(prod1 & 0xffffffffu);
    t2 = (uint32_t)(prod0 >> 32) ^ c3 ^ k1;
    t3 = (uint32_t)(prod0 & 0xffffffffu);
    c0 = t0;
    c1 = t1;
    c2 = t2;
    c3 = t3;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

static double u01(uint32_t x) { return ((double)(x >> 9) + 0.5) / 8388608.0; }

static const double kTwoPi = 6.283185307179586476925286766559;

static void box_muller(double u0, double u1, int fp32, double* z0, double* z1) {
  if (fp32) {
    float r = sqrtf(-2.0f * logf((float)u0));
    float a = (float)(kTwoPi * u1);
    *z0 = (double)(r * cosf(a));
    *z1 = (double)(r * sinf(a));
  } else {
    double r = sqrt(-2.0 * log(u0));
    *z0 = r * cos(kTwoPi * u1);
    *z1 = r * sin(kTwoPi * u1);
  }
}

double orc_patch_normal(uint64_t seed, uint32_t tag, uint32_t path,
                        uint32_t year, uint32_t patch, int fp32) {
  uint32_t ctr[4], key[2], x[4];
  double z[4];
  ctr[0] = path >> 2;
  ctr[1] = year;
  ctr[2] = patch;
  ctr[3] = tag;
  key[0] = (uint32_t)(seed & 0xffffffffu);
  key[1] = (uint32_t)(seed >> 32);
  orc_philox4x32_10(ctr, key, x);
  box_muller(u01(x[0]), u01(x[1]), fp32, &z[0], &z[1]);
  box_muller(u01(x[2]), u01(x[3]), fp32, &z[2], &z[3]);
  return z[path & 3u];
}

void orc_path_scalars(uint64_t seed, uint32_t tag, uint32_t path,
                      uint32_t year, double out[3]) {
  uint32_t ctr[4], key[2], x[4];
  ctr[0] = path;
  ctr[1] = year;
  ctr[2] = 0;
  ctr[3] = tag;
  key[0] = (uint32_t)(seed & 0xffffffffu);
  key[1] = (uint32_t)(seed >> 32);
  orc_philox4x32_10(ctr, key, x);
  box_muller(u01(x[0]), u01(x[1]), 0, &out[0], &out[1]);
  out[2] = u01(x[2]);
}

static uint32_t make_tag(uint32_t road, uint32_t species, uint32_t kind) {
  return (road << 8) | ((species & 63u) << 2) | (kind & 3u);
}

/* One path, FP64 arithmetic.  R is the row-major origin x destination
 * matrix R[k*P + j] = T(k,j) * S(k,j) (Simulator.cpp:751-754). */
static void path_f64(const orc_species* sp, const double* R, uint32_t tag0,
                     uint32_t tag1, int32_t T, double dt, int32_t law,
                     int32_t endpoint, uint64_t seed, uint32_t path,
                     double* n, double* np, double* total_out) {
  const int P = sp->n_patches;
  const double* p = sp->params;
  double grm = p[0];
  int t, j, k;
  for (j = 0; j < P; ++j) n[j] = sp->init_pop[j];
  for (t = 0; t < T; ++t) {
    /* Movement and road mortality: pops = R^T pops (Simulator.cpp:763),
     * i.e. the CSC column sums of mteKernel (SimulateGPU.cu:628-635). */
    for (j = 0; j < P; ++j) np[j] = 0.0;
    for (k = 0; k < P; ++k) {
      const double nk = n[k];
      const double* row = R + (size_t)k * P;
      for (j = 0; j < P; ++j) np[j] += nk * row[j];
    }
    if (law == ORC_LAW_REF_GPU) {
      /* grMean update, SimulateGPU.cu:686-701. */
      double s[3];
      double jump, brown, jsz;
      orc_path_scalars(seed, tag1, path, (uint32_t)t, s);
      jump = (s[2] < p[6]) ? 1.0 : 0.0;
      brown = s[0] * p[2];
      jsz = s[1] * p[5] * p[5] - p[5] * p[5] / 2.0;
      grm = grm + p[4] * (p[1] - grm) * dt + grm * brown +
            (exp(jsz) - 1.0) * grm * jump;
    }
    for (j = 0; j < P; ++j) {
      const double z = orc_patch_normal(seed, tag0, path, (uint32_t)t,
                                        (uint32_t)j, 0);
      const double x = np[j];
      const double K = sp->capacity[j];
      double g;
      switch (law) {
        case ORC_LAW_REF_CPU:
          /* naturalBirthDeath, Simulator.cpp:712-728. */
          g = dt * (p[1] + p[2] * z) * 0.01;
          n[j] = x * (1.0 + g * (K - x) / K);
          break;
        case ORC_LAW_REF_GPU:
          /* SimulateGPU.cu:703-709. */
          g = p[7] * z * grm + grm;
          n[j] = x * (1.0 + g * (K - x) / K);
          break;
        default: {
          const double r = sp->growth ? sp->growth[j] : 0.0;
          const double s = sp->noise_sd;
          n[j] = x * (1.0 + r * (1.0 - x / K)) * exp(s * z - 0.5 * s * s);
        }
      }
    }
  }
  {
    const double* v = (endpoint == ORC_END_PRE) ? np : n;
    double tot = 0.0;
    for (j = 0; j < P; ++j) tot += v[j];
    *total_out = tot;
    if (endpoint == ORC_END_PRE) {
      for (j = 0; j < P; ++j) n[j] = np[j];
    }
  }
}

/* One path, binary32 arithmetic (mirrors the precision of the FP32 kernel). */
static void path_f32(const orc_species* sp, const float* R, uint32_t tag0,
                     uint32_t tag1, int32_t T, float dt, int32_t law,
                     int32_t endpoint, uint64_t seed, uint32_t path, float* n,
                     float* np, double* total_out) {
  const int P = sp->n_patches;
  float p[8];
  float grm;
  int t, j, k, i;
  for (i = 0; i < 8; ++i) p[i] = (float)sp->params[i];
  grm = p[0];
  for (j = 0; j < P; ++j) n[j] = (float)sp->init_pop[j];
  for (t = 0; t < T; ++t) {
    for (j = 0; j < P; ++j) np[j] = 0.0f;
    for (k = 0; k < P; ++k) {
      const float nk = n[k];
      const float* row = R + (size_t)k * P;
      for (j = 0; j < P; ++j) np[j] += nk * row[j];
    }
    if (law == ORC_LAW_REF_GPU) {
      double s[3];
      float jump, brown, jsz;
      orc_path_scalars(seed, tag1, path, (uint32_t)t, s);
      jump = ((float)s[2] < p[6]) ? 1.0f : 0.0f;
      brown = (float)s[0] * p[2];
      jsz = (float)s[1] * p[5] * p[5] - p[5] * p[5] / 2.0f;
      grm = grm + p[4] * (p[1] - grm) * dt + grm * brown +
            (expf(jsz) - 1.0f) * grm * jump;
    }
    for (j = 0; j < P; ++j) {
      const float z = (float)orc_patch_normal(seed, tag0, path, (uint32_t)t,
                                              (uint32_t)j, 1);
      const float x = np[j];
      const float K = (float)sp->capacity[j];
      float g;
      switch (law) {
        case ORC_LAW_REF_CPU:
          g = dt * (p[1] + p[2] * z) * 0.01f;
          n[j] = x * (1.0f + g * (K - x) / K);
          break;
        case ORC_LAW_REF_GPU:
          g = p[7] * z * grm + grm;
          n[j] = x * (1.0f + g * (K - x) / K);
          break;
        default: {
          const float r = sp->growth ? (float)sp->growth[j] : 0.0f;
          const float s = (float)sp->noise_sd;
          n[j] = x * (1.0f + r * (1.0f - x / K)) * expf(s * z - 0.5f * s * s);
        }
      }
    }
  }
  {
    const float* v = (endpoint == ORC_END_PRE) ? np : n;
    double tot = 0.0;
    for (j = 0; j < P; ++j) tot += (double)v[j];
    *total_out = tot;
    if (endpoint == ORC_END_PRE) {
      for (j = 0; j < P; ++j) n[j] = np[j];
    }
  }
}

int orc_mte_paths(const orc_species* sp, uint32_t species_index,
                  uint32_t road_id, int32_t n_years, double time_step,
                  int32_t law, int32_t endpoint, double drop_tol,
                  uint64_t seed, int64_t path_begin, int64_t path_end,
                  int fp32, int n_threads, double* end_totals,
                  double* patch_sum) {
  const int P = sp->n_patches;
  const uint32_t tag0 = make_tag(road_id, species_index, 0u);
  const uint32_t tag1 = make_tag(road_id, species_index, 1u);
  double* R;
  float* Rf;
  int64_t count;
  int k, j;
  if (P <= 0 || n_years < 0 || path_end < path_begin) return 1;
  R = (double*)malloc(sizeof(double) * (size_t)P * P);
  Rf = (float*)malloc(sizeof(float) * (size_t)P * P);
  if (!R || !Rf) {
    free(R);
    free(Rf);
    return 3;
  }
  /* R = T o S (Simulator.cpp:751-754; SimulateGPU.cu:3783-3787), stored
   * row-major by origin.  Inputs are column-major: (k, j) at j*P + k.  With a
   * drop tolerance, entries whose binary32 value is <= tol are removed, as
   * dense2Sparse does (SimulateGPU.cu:5985). */
  for (k = 0; k < P; ++k) {
    for (j = 0; j < P; ++j) {
      const size_t cm = (size_t)j * P + k;
      double v = sp->trans[cm] * (sp->surv ? sp->surv[cm] : 1.0);
      if (drop_tol > 0.0 && !((float)v > (float)drop_tol)) v = 0.0;
      R[(size_t)k * P + j] = v;
      Rf[(size_t)k * P + j] = (float)v;
    }
  }
  count = path_end - path_begin;
  if (patch_sum) {
    for (j = 0; j < P; ++j) patch_sum[j] = 0.0;
  }
#ifdef _OPENMP
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#endif
  {
    double* n = (double*)malloc(sizeof(double) * (size_t)P);
    double* np = (double*)malloc(sizeof(double) * (size_t)P);
    float* nf = (float*)malloc(sizeof(float) * (size_t)P);
    float* npf = (float*)malloc(sizeof(float) * (size_t)P);
    double* local = (double*)calloc((size_t)P, sizeof(double));
    int64_t i;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (i = 0; i < count; ++i) {
      const uint32_t path = (uint32_t)(path_begin + i);
      int jj;
      if (fp32) {
        path_f32(sp, Rf, tag0, tag1, n_years, (float)time_step, law, endpoint,
                 seed, path, nf, npf, &end_totals[i]);
        for (jj = 0; jj < P; ++jj) local[jj] += (double)nf[jj];
      } else {
        path_f64(sp, R, tag0, tag1, n_years, time_step, law, endpoint, seed,
                 path, n, np, &end_totals[i]);
        for (jj = 0; jj < P; ++jj) local[jj] += n[jj];
      }
    }
    if (patch_sum) {
#ifdef _OPENMP
#pragma omp critical
#endif
      {
        int jj;
        for (jj = 0; jj < P; ++jj) patch_sum[jj] += local[jj];
      }
    }
    free(n);
    free(np);
    free(nf);
    free(npf);
    free(local);
  }
  free(R);
  free(Rf);
  return 0;
}

int32_t orc_dense_to_csc(const float* dense, int32_t rows, int32_t cols,
                         float tol, float* val, int32_t* row_idx,
                         int32_t* per_col) {
  /* Column by column, keep entries > tol (SimulateGPU.cu:5973-5996). */
  int32_t nnz = 0, c, r;
  for (c = 0; c < cols; ++c) {
    int32_t in_col = 0;
    for (r = 0; r < rows; ++r) {
      const float v = dense[(size_t)c * rows + r];
      if (v > tol) {
        val[nnz] = v;
        row_idx[nnz] = r;
        ++nnz;
        ++in_col;
      }
    }
    per_col[c] = in_col;
  }
  return nnz;
}

static double as_rational(double t) {
  /* Abramowitz and Stegun 26.2.23, |error| < 4.5e-4. */
  const double c0 = 2.515517, c1 = 0.802853, c2 = 0.010328;
  const double d1 = 1.432788, d2 = 0.189269, d3 = 0.001308;
  return t - ((c2 * t + c1) * t + c0) / (((d3 * t + d2) * t + d1) * t + 1.0);
}

double orc_normal_cdf_inverse(double p) {
  if (!(p > 0.0 && p < 1.0)) return NAN;
  if (p < 0.5) return -as_rational(sqrt(-2.0 * log(p)));
  return as_rational(sqrt(-2.0 * log(1.0 - p)));
}

double orc_mte_penalty(double threshold, double end_mean, double end_sd,
                       double confidence_interval, double initial_population,
                       double cost_per_animal) {
  /* Simulator.cpp:197-212. */
  const double z = orc_normal_cdf_inverse(1.0 - confidence_interval);
  const double level = z * end_sd + end_mean;
  if (threshold > level)
    return (threshold - level) * initial_population * cost_per_animal;
  return 0.0;
}
