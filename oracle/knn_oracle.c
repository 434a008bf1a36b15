/*
 * CPU oracle: kNN local-linear grid surrogate.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates the reference's
 * surrogate fit (SimulateGPU::buildSurrogateMTECUDA / buildSurrogateROVCUDA,
 * include/gpuCode/SimulateGPU.cu:5533-5917, with createQueryPoints :1152-1193,
 * multiLocLinReg :2934-3053 and rovCorrection :3055-3073) and its
 * multilinear evaluation (Utility::interpolateSurrogate,
 * include/utilities/Utility.cpp:253-363), with the decisions of the library:
 * FP64 throughout, neighbours ranked by (squared distance, sample index),
 * neighbours and fit in one coordinate system, a Cholesky solve with the
 * weighted mean as the singular fallback, and interpolation weights applied
 * to the dimension they belong to.
 */
#pragma GCC optimize("fp-contract=off")

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define KNN_PI 3.14159265358979323846

int32_t orc_knn_k(int32_t n, int32_t rule) {
  int k;
  if (rule == 0) { /* buildSurrogateMTECUDA, SimulateGPU.cu:5545-5554 */
    if (n < 50)
      k = 5;
    else if (n < 100)
      k = 7;
    else if (n < 500)
      k = (int)ceil(0.02 * (n - 100) + 7);
    else
      k = (int)ceil(log((double)n) + 8);
  } else { /* buildSurrogateROVCUDA, :5711-5724; (n-50)/50 and (n-100)/400 are
              integer divisions there, both 0 on their ranges */
    if (n < 50)
      k = 5;
    else if (n < 100)
      k = 5 + 5 * ((n - 50) / 50);
    else if (n < 500)
      k = 10 + 15 * ((n - 100) / 400);
    else
      k = (int)(25 + ceil((n - 500) * 0.02));
  }
  return k < n ? k : n;
}

typedef struct {
  double d2;
  int32_t j;
} knn_pair;

static int knn_cmp(const void* a, const void* b) {
  const knn_pair* p = (const knn_pair*)a;
  const knn_pair* q = (const knn_pair*)b;
  if (p->d2 < q->d2) return -1;
  if (p->d2 > q->d2) return 1;
  return (p->j > q->j) - (p->j < q->j);
}

/* Weighted local-linear intercept at q over the neighbours nb[0..k). */
static double knn_local_linear(int d, int k, const knn_pair* nb, const double* u,
                               const double* y, const double* q) {
  const int D1 = d + 1;
  double A[9 * 9], B[9], L[9 * 9], z[9];
  memset(A, 0, sizeof(A));
  memset(B, 0, sizeof(B));
  const double h = sqrt(nb[k - 1].d2);
  for (int r = 0; r < k; ++r) {
    const double dist = sqrt(nb[r].d2);
    const double t = h > 0.0 ? dist / h : 0.0;
    const double w = exp(-t * t / 2) / sqrt(2 * KNN_PI);
    double v[9];
    v[0] = 1.0;
    for (int a = 1; a < D1; ++a) v[a] = u[(size_t)nb[r].j * d + a - 1] - q[a - 1];
    const double yr = y[nb[r].j];
    for (int a = 0; a < D1; ++a) {
      B[a] += (a == 0 ? yr : yr * v[a]) * w;
      for (int b = 0; b <= a; ++b) A[a * 9 + b] += (a == 0 && b == 0 ? 1.0 : v[a] * v[b]) * w;
    }
  }
  /* Cholesky A = L L^T (lower triangle of A used). */
  double amax = 0.0;
  for (int a = 0; a < D1; ++a) amax = fmax(amax, fabs(A[a * 9 + a]));
  for (int a = 0; a < D1; ++a) {
    for (int b = 0; b <= a; ++b) {
      double s = A[a * 9 + b];
      for (int c = 0; c < b; ++c) s -= L[a * 9 + c] * L[b * 9 + c];
      if (a == b) {
        if (!(s > 1e-12 * amax) || !isfinite(s)) return B[0] / A[0];
        L[a * 9 + a] = sqrt(s);
      } else {
        L[a * 9 + b] = s / L[b * 9 + b];
      }
    }
  }
  for (int a = 0; a < D1; ++a) {
    double s = B[a];
    for (int c = 0; c < a; ++c) s -= L[a * 9 + c] * z[c];
    z[a] = s / L[a * 9 + a];
  }
  for (int a = D1 - 1; a >= 0; --a) {
    double s = z[a];
    for (int c = a + 1; c < D1; ++c) s -= L[c * 9 + a] * z[c];
    z[a] = s / L[a * 9 + a];
  }
  return z[0];
}

int orc_knn_fit_scaled(int32_t n, int32_t d, const double* u, const double* y, int32_t R,
                       int32_t k, const double* qtab, double* val) {
  size_t G = 1;
  for (int dd = 0; dd < d; ++dd) G *= (size_t)R;
  knn_pair* nb = (knn_pair*)malloc(sizeof(knn_pair) * (size_t)(n > 0 ? n : 1));
  if (!nb) return -2;
  if (k > n) k = n;
  for (size_t g = 0; g < G; ++g) {
    double q[8];
    size_t rem = g;
    for (int dd = d - 1; dd >= 0; --dd) {
      q[dd] = qtab[dd * R + (int)(rem % (size_t)R)];
      rem /= (size_t)R;
    }
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int dd = 0; dd < d; ++dd) {
        const double e = u[(size_t)j * d + dd] - q[dd];
        s += e * e;
      }
      nb[j].d2 = s;
      nb[j].j = j;
    }
    qsort(nb, (size_t)n, sizeof(knn_pair), knn_cmp);
    val[g] = knn_local_linear(d, k, nb, u, y, q);
  }
  free(nb);
  return 0;
}

int orc_knn_fit(int32_t n, int32_t d, const double* x, const double* y, int32_t R, int32_t k,
                int32_t flags, double* surrogate) {
  if (d < 1 || d > 8 || R < 2 || k < 1) return -1;
  size_t G = 1;
  for (int dd = 0; dd < d; ++dd) G *= (size_t)R;
  double mn[8], mx[8], span[8];
  for (int dd = 0; dd < d; ++dd) {
    mn[dd] = n > 0 ? x[dd] : 0.0;
    mx[dd] = mn[dd];
    for (int j = 1; j < n; ++j) {
      mn[dd] = fmin(mn[dd], x[(size_t)j * d + dd]);
      mx[dd] = fmax(mx[dd], x[(size_t)j * d + dd]);
    }
    span[dd] = mx[dd] - mn[dd];
    for (int i = 0; i < R; ++i) surrogate[dd * R + i] = (double)i * span[dd] / (double)(R - 1) + mn[dd];
  }
  double* val = surrogate + (size_t)d * R;
  if (n < 3) {
    for (size_t g = 0; g < G; ++g) val[g] = 0.0;
    return 0;
  }
  if (k > n) k = n;
  const int norm = flags & 1;
  double* u = (double*)malloc(sizeof(double) * (size_t)n * d);
  knn_pair* nb = (knn_pair*)malloc(sizeof(knn_pair) * (size_t)n);
  if (!u || !nb) {
    free(u);
    free(nb);
    return -2;
  }
  for (int j = 0; j < n; ++j)
    for (int dd = 0; dd < d; ++dd) {
      const double v = x[(size_t)j * d + dd];
      u[(size_t)j * d + dd] = norm ? (span[dd] > 0.0 ? (v - mn[dd]) / span[dd] : 0.0) : v;
    }
  for (size_t g = 0; g < G; ++g) {
    double q[8];
    size_t rem = g;
    for (int dd = d - 1; dd >= 0; --dd) {
      const int i = (int)(rem % (size_t)R);
      rem /= (size_t)R;
      q[dd] = norm ? (span[dd] > 0.0 ? (double)i / (double)(R - 1) : 0.0) : surrogate[dd * R + i];
    }
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int dd = 0; dd < d; ++dd) {
        const double e = u[(size_t)j * d + dd] - q[dd];
        s += e * e;
      }
      nb[j].d2 = s;
      nb[j].j = j;
    }
    qsort(nb, (size_t)n, sizeof(knn_pair), knn_cmp);
    double v = knn_local_linear(d, k, nb, u, y, q);
    if ((flags & 2) && v > 0.0) v = 0.0; /* rovCorrection */
    val[g] = v;
  }
  free(u);
  free(nb);
  return 0;
}

void orc_knn_eval(int32_t d, int32_t R, const double* surrogate, int32_t m, const double* pred,
                  double* out) {
  const double* val = surrogate + (size_t)d * R;
  for (int p = 0; p < m; ++p) {
    int li[8];
    double xd[8];
    for (int dd = 0; dd < d; ++dd) {
      const double lo = surrogate[dd * R], hi = surrogate[dd * R + R - 1];
      const double pv = pred[(size_t)p * d + dd];
      double c = hi != lo ? (R - 1) * (pv - lo) / (hi - lo) : 0.0;
      int i = c >= (double)(R - 2) ? R - 2 : (c > 0.0 ? (int)c : 0);
      li[dd] = i;
      const double x0 = surrogate[dd * R + i], x1 = surrogate[dd * R + i + 1];
      xd[dd] = (fabs(x1 - x0) < DBL_EPSILON || x0 == x1) ? 0.0 : (pv - x0) / (x1 - x0);
    }
    double s = 0.0;
    for (int c = 0; c < (1 << d); ++c) {
      double w = 1.0;
      size_t idx = 0;
      for (int dd = 0; dd < d; ++dd) {
        const int bit = (c >> (d - 1 - dd)) & 1;
        w *= bit ? xd[dd] : 1.0 - xd[dd];
        idx = idx * (size_t)R + (size_t)(li[dd] + bit);
      }
      s += val[idx] * w;
    }
    out[p] = s;
  }
}
