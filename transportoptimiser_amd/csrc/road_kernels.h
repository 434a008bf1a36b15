// Host-visible declarations of the road-matrix kernels (road_kernels.hip).
#ifndef ROADMC_ROAD_KERNELS_H
#define ROADMC_ROAD_KERNELS_H

#include <hip/hip_runtime.h>

namespace roadmc {

// Device status word of one build, zeroed before the launch.
struct RoadTail {
  int bad;  // set when a normalised transition weight is not finite
  int pad[3];
};

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
  int drop_last;            // leave the last in-range pair uncounted
  const double* flow;       // [C]
  double width_len;         // vehicle width + animal length
  double speed_3600;        // animal speed * 3600
  int* cross;               // [P*P] column-major counts
  double* trans;            // [P*P] column-major
  double* surv;             // [C][P*P] column-major
  int* blockmax;            // [road_cross_blocks(P)] largest in-range key per workgroup
  double* part;             // [ceil(P / kRoadCols)][P] partial transition row sums
  RoadTail* tail;
};

constexpr int kRoadThreads = 256;
constexpr int kRoadCols = 256;   // columns per transition workgroup
constexpr int kRoadCand = 4096;  // matrix entries per crossing workgroup

inline int road_cross_blocks(int P) {
  return (int)(((long long)P * P + kRoadCand - 1) / kRoadCand);
}
inline size_t road_part_elems(int P) { return (size_t)((P + kRoadCols - 1) / kRoadCols) * P; }

hipError_t launch_road_build(const RoadArgs& a, bool want_trans, bool want_surv, hipStream_t s);

// Packs trans o surv (column-major P x P, FP64; surv may be null) into the
// MTE kernels' B-operand order (Pp columns, Pk contraction rows), entries
// whose binary32 value is not above tol dropped when tol > 0.
hipError_t launch_road_pack(bool fp64, const double* trans, const double* surv, int P, int Pp,
                            int Pk, double tol, void* out, hipStream_t s);

}  // namespace roadmc

#endif
