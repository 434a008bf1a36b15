/*
 * CPU oracle, adjoint part.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Reverse-mode derivative of J = (1/N) sum_paths sum_patches n_T for the
 * LOGNORMAL law on the patch-noise stream, with respect to the crossing
 * mortality m (S = 1 - m, R = T o S) and the growth rates r.  Stores the whole
 * trajectory (small problems only).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int orc_mte_adjoint(const orc_species* sp, uint32_t road_id, int32_t n_years,
                    int64_t n_paths, uint64_t seed, double* objective,
                    double* grad_m, double* grad_r) {
  const int P = sp->n_patches, T = n_years;
  const int64_t N = n_paths;
  const uint32_t tag0 = (road_id << 8) | 0u;
  const double s = sp->noise_sd;
  double *R, *gM, *n, *np, *eps, *lam, *mu;
  int64_t p;
  int t, j, k;
  R = (double*)malloc(sizeof(double) * (size_t)P * P);
  gM = (double*)calloc((size_t)P * P, sizeof(double));
  n = (double*)malloc(sizeof(double) * (size_t)(T + 1) * P);  /* n_t */
  np = (double*)malloc(sizeof(double) * (size_t)(T + 1) * P); /* n'_t */
  eps = (double*)malloc(sizeof(double) * (size_t)(T + 1) * P);
  lam = (double*)malloc(sizeof(double) * P);
  mu = (double*)malloc(sizeof(double) * P);
  for (k = 0; k < P; ++k)
    for (j = 0; j < P; ++j) {
      const size_t cm = (size_t)j * P + k;
      R[(size_t)k * P + j] = sp->trans[cm] * (sp->surv ? sp->surv[cm] : 1.0);
    }
  memset(grad_r, 0, sizeof(double) * P);
  *objective = 0.0;
  for (p = 0; p < N; ++p) {
    for (j = 0; j < P; ++j) n[j] = sp->init_pop[j];
    for (t = 1; t <= T; ++t) {
      for (j = 0; j < P; ++j) {
        double a = 0.0;
        for (k = 0; k < P; ++k) a += n[(size_t)(t - 1) * P + k] * R[(size_t)k * P + j];
        np[(size_t)t * P + j] = a;
      }
      for (j = 0; j < P; ++j) {
        const double z = orc_patch_normal(seed, tag0, (uint32_t)p, (uint32_t)(t - 1), (uint32_t)j, 0);
        const double e = exp(s * z - 0.5 * s * s);
        const double x = np[(size_t)t * P + j], K = sp->capacity[j], r = sp->growth[j];
        eps[(size_t)t * P + j] = e;
        n[(size_t)t * P + j] = x * (1.0 + r * (1.0 - x / K)) * e;
      }
    }
    for (j = 0; j < P; ++j) *objective += n[(size_t)T * P + j] / (double)N;
    for (j = 0; j < P; ++j) lam[j] = 1.0 / (double)N;
    for (t = T; t >= 1; --t) {
      for (j = 0; j < P; ++j) {
        const double x = np[(size_t)t * P + j], K = sp->capacity[j], r = sp->growth[j];
        const double e = eps[(size_t)t * P + j];
        mu[j] = lam[j] * (1.0 + r * (1.0 - 2.0 * x / K)) * e;
        grad_r[j] += lam[j] * x * (1.0 - x / K) * e;
      }
      for (k = 0; k < P; ++k) {
        double a = 0.0;
        const double nk = n[(size_t)(t - 1) * P + k];
        for (j = 0; j < P; ++j) {
          gM[(size_t)k * P + j] += nk * mu[j];
          a += mu[j] * R[(size_t)k * P + j];
        }
        lam[k] = a;
      }
    }
  }
  /* dJ/dm_kj = -T_kj dJ/dR_kj, column-major output. */
  for (k = 0; k < P; ++k)
    for (j = 0; j < P; ++j) {
      const size_t cm = (size_t)j * P + k;
      grad_m[cm] = -sp->trans[cm] * gM[(size_t)k * P + j];
    }
  free(R);
  free(gM);
  free(n);
  free(np);
  free(eps);
  free(lam);
  free(mu);
  return 0;
}
