// Host-visible declarations of the road-matrix kernels (road_kernels.hip).
#ifndef ROADMC_ROAD_KERNELS_H
#define ROADMC_ROAD_KERNELS_H

#include <hip/hip_runtime.h>

namespace roadmc {

struct RoadArgs {
  int P, S, C;              // patches, visible road segments, controls
  const double* x;          // [P] patch centroids
  const double* y;          // [P]
  const double* pw;         // [P] area * lambda * exp(pref_mean + pref_sd * mult)
  double lambda;            // transition decay rate
  double cut_cross;         // distance cut-off of the crossing counts
  double cut_trans;         // distance cut-off of the transitions
  const double4* seg;       // [S] visible segments (x1, y1, x2, y2)
  int form;                 // 0: weight / exp(-lambda d); 1: weight * exp(-lambda d)
  const double* flow;       // [C]
  double width_len;         // vehicle width + animal length
  double speed_3600;        // animal speed * 3600
  int* cross;               // [P*P] column-major counts
  double* trans;            // [P*P] column-major
  double* surv;             // [C][P*P] column-major
  int* last;                // [1] largest row-major index of an in-range pair, or null
};

hipError_t launch_road_build(const RoadArgs& a, bool drop_last, bool want_trans,
                             bool want_surv, hipStream_t s);

}  // namespace roadmc

#endif
