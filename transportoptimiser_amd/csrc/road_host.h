// Host side of the road -> matrix builder shared by roadmc_build_species_road
// and the MTE runs that build their matrices on the device (internal).
#ifndef ROADMC_ROAD_HOST_H
#define ROADMC_ROAD_HOST_H

#include <vector>

#include "ctx.h"
#include "road_kernels.h"

namespace roadmc {

// Host copies a build's uploads read from; kept alive until the stream has
// passed them.
struct RoadStaging {
  std::vector<double> in;  // road_stage blocks
  RoadTail tail;
};

int road_validate(roadmc_ctx* ctx, const roadmc_road_species* r);
int road_visible_segments(const roadmc_road_species* r);
// Device scratch for builds of up to P patches, S segments and C controls.
int road_reserve(roadmc_ctx* ctx, int P, int S, int C);
// Uploads one species' patches and road and launches the build on ctx's
// stream (no synchronisation).  With reset_tail the status word is cleared
// first; a non-finite transition weight sets tail.bad.  The scratch must
// already be reserved for this build.
int road_launch(roadmc_ctx* ctx, const roadmc_road_species* r, bool want_trans, bool want_surv,
                bool reset_tail, RoadStaging& st, RoadArgs* out);
// Appends one species' build inputs to a host block (patch x, y and weights,
// the open-road segments, the flows; 32-byte aligned parts) and returns the
// block's offset in doubles, so that many builds share one upload.
size_t road_stage(const roadmc_road_species* r, std::vector<double>& in);
// Device scratch of one build: crossings, transitions, survival, the
// per-workgroup maxima and the row-sum partials.
struct RoadScratch {
  int* cross;
  double* trans;
  double* surv;
  int* blockmax;
  double* part;
};
// Launches the build of r whose staged block is at d_in on the device, in
// the given scratch on stream s.
int road_launch_staged(roadmc_ctx* ctx, const roadmc_road_species* r, const double* d_in,
                       bool want_trans, bool want_surv, const RoadScratch& sc, hipStream_t s,
                       RoadArgs* out);
// The scratch of ctx's own stream (road_reserve).
RoadScratch road_scratch(roadmc_ctx* ctx);
// Side-stream lanes for overlapping builds: their streams and events, and
// scratch for builds of up to P patches and C controls.
int road_reserve_lanes(roadmc_ctx* ctx, int P, int C);
RoadTail* road_tail(roadmc_ctx* ctx);

}  // namespace roadmc

#endif
