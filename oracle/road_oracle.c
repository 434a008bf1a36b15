/*
 * CPU oracle: road -> crossing counts, transition and survival matrices.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates, in the reference's own
 * order of operations and without fused multiply-adds:
 *   - SpeciesRoadPatches::habitatPatchDistances (include/road/
 *     SpeciesRoadPatches.cpp:270-294);
 *   - SpeciesRoadPatches::roadCrossings (:296-381) with the CPU branch of
 *     Utility::lineSegmentIntersect (include/utilities/Utility.cpp:3-91) and
 *     Utility::sliceIntoPairs (include/utilities/Utility.h:128-139);
 *   - SpeciesRoadPatches::computeTransitionProbabilities (:383-426);
 *   - SpeciesRoadPatches::computeSurvivalProbabilities (:428-451).
 * Matrices are column-major P x P (Eigen's default order), row = origin.
 */
#pragma GCC optimize("fp-contract=off")

#include <math.h>
#include <stdlib.h>

#include "oracle.h"

/* dists(i, j) = sqrt((x_i - x_j)^2 + (y_i - y_j)^2), as the repmat form
 * (SpeciesRoadPatches.cpp:288-293) evaluates it element by element. */
static double road_dist(const double* x, const double* y, int i, int j) {
  const double dx = x[j] - x[i], dy = y[j] - y[i];
  return sqrt(dx * dx + dy * dy);
}

/* One entry of INT_B (Utility.cpp:45-83): line (x1,y1)-(x2,y2) against road
 * segment (x3,y3)-(x4,y4). */
static int road_hit(double x1, double y1, double x2, double y2, double x3, double y3, double x4,
                    double y4) {
  const double x4_x3 = x4 - x3, y1_y3 = y1 - y3, y4_y3 = y4 - y3;
  const double x1_x3 = x1 - x3, x2_x1 = x2 - x1, y2_y1 = y2 - y1;
  const double num_a = x4_x3 * y1_y3 - y4_y3 * x1_x3;
  const double num_b = x2_x1 * y1_y3 - y2_y1 * x1_x3;
  const double den = y4_y3 * x2_x1 - x4_x3 * y2_y1;
  const double u_a = num_a / den, u_b = num_b / den;
  return (u_a >= 0 && u_a <= 1 && u_b >= 0 && u_b <= 1) ? 1 : 0;
}

int64_t orc_road_crossings(int32_t P, const double* x, const double* y, double lambda_mean,
                           double lambda_sd, double lambda_mult, int32_t n_points,
                           const double* road_x, const double* road_y, const int32_t* road_type,
                           int32_t drop_last, int32_t* out) {
  /* Cut-off with the raw multiplier (SpeciesRoadPatches.cpp:328-332). */
  const double max_dist = log(0.05) / (-(lambda_mean + lambda_sd * lambda_mult));
  const int n_segs = n_points > 1 ? n_points - 1 : 0;
  double* vis = (double*)malloc(sizeof(double) * 4 * (n_segs > 0 ? n_segs : 1));
  int64_t* pairs = (int64_t*)malloc(sizeof(int64_t) * ((int64_t)P * P / 2 + 1));
  if (!vis || !pairs) {
    free(vis);
    free(pairs);
    return -1;
  }
  /* Open-road segments only (:337-347). */
  int n_vis = 0;
  for (int s = 0; s < n_segs; ++s)
    if (!road_type || road_type[s] == 0) {
      vis[4 * n_vis + 0] = road_x[s];
      vis[4 * n_vis + 1] = road_y[s];
      vis[4 * n_vis + 2] = road_x[s + 1];
      vis[4 * n_vis + 3] = road_y[s + 1];
      ++n_vis;
    }
  /* Pairs i < j within the cut-off, in loop order (:352-366). */
  int64_t n_pairs = 0;
  for (int i = 0; i < P; ++i)
    for (int j = i + 1; j < P; ++j)
      if (road_dist(x, y, i, j) <= max_dist) pairs[n_pairs++] = (int64_t)i * P + j;
  /* roadCrossings passes rows [0, validCrossings) on, validCrossings being
   * the index of the last pair: the last pair is not counted (:369-374). */
  const int64_t n_used = drop_last ? (n_pairs > 0 ? n_pairs - 1 : 0) : n_pairs;
  for (int64_t e = 0; e < (int64_t)P * P; ++e) out[e] = 0;
  for (int64_t k = 0; k < n_used; ++k) {
    const int i = (int)(pairs[k] / P), j = (int)(pairs[k] % P);
    int count = 0; /* INT_B.rowwise().sum() (Utility.cpp:90) */
    for (int s = 0; s < n_vis; ++s)
      count += road_hit(x[i], y[i], x[j], y[j], vis[4 * s], vis[4 * s + 1], vis[4 * s + 2],
                        vis[4 * s + 3]);
    out[(int64_t)j * P + i] = count; /* sliceIntoPairs(orgs, dests) */
    out[(int64_t)i * P + j] = count; /* sliceIntoPairs(dests, orgs) */
  }
  free(vis);
  free(pairs);
  return n_pairs;
}

void orc_road_transitions(int32_t P, const double* x, const double* y, const double* area,
                          const double* pref_mean, const double* pref_sd, double pref_mult,
                          double lambda_mean, double lambda_sd, double lambda_mult,
                          int32_t form, double* out) {
  const double lda = lambda_mult * lambda_sd + lambda_mean;
  const double max_dist = log(0.05) / (-(lambda_mean + lambda_sd * lda));
  for (int64_t e = 0; e < (int64_t)P * P; ++e) out[e] = 0.0;
  for (int i = 0; i < P; ++i) {
    double summ = 0;
    for (int j = 0; j < P; ++j) {
      const double d = road_dist(x, y, i, j);
      if (d <= max_dist) {
        const double w = area[j] * lda * exp(pref_mean[j] + pref_sd[j] * pref_mult);
        double t;
        if (i == j)
          t = w;
        else if (form == 0)
          t = w / exp(-lda * d); /* :413-416 */
        else
          t = w * exp(-lda * d);
        out[(int64_t)j * P + i] = t;
        summ += t;
      }
    }
    for (int j = 0; j < P; ++j) out[(int64_t)j * P + i] = out[(int64_t)j * P + i] / summ;
  }
}

void orc_road_survival(int32_t P, const int32_t* crossings, int32_t n_controls,
                       const double* flow, double vehicle_width, double animal_length,
                       double animal_speed, double* out) {
  const int64_t n = (int64_t)P * P;
  for (int c = 0; c < n_controls; ++c)
    for (int64_t e = 0; e < n; ++e)
      out[c * n + e] = exp(-(double)crossings[e] * flow[c] * (vehicle_width + animal_length) /
                           (animal_speed * 3600));
}
