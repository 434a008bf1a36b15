// C-ABI host code for the road -> matrix builder (road_kernels.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"
#include "road_host.h"
#include "road_kernels.h"
#include "roadmc.h"

namespace roadmc {

int road_validate(roadmc_ctx* ctx, const roadmc_road_species* r) {
  if (!r) return rmc_fail(ctx, ROADMC_ERR_INVALID, "road species is NULL");
  if (r->n_patches < 1 || r->n_patches > 16384)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_patches must be in [1, 16384]");
  if (!r->x || !r->y || !r->area || !r->pref_mean || !r->pref_sd)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "patch arrays must not be NULL");
  if (r->n_road_points < 0 || (r->n_road_points > 0 && (!r->road_x || !r->road_y)))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "road vertices invalid");
  if (r->n_controls < 0 || (r->n_controls > 0 && !r->flow))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_controls/flow invalid");
  if (r->transition_form != ROADMC_TRANS_REFERENCE && r->transition_form != ROADMC_TRANS_DECAY)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "unknown transition_form");
  const double sc[] = {r->pref_mult, r->lambda_mean, r->lambda_sd, r->lambda_mult,
                       r->vehicle_width, r->animal_length, r->animal_speed};
  for (double v : sc)
    if (!std::isfinite(v)) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite parameter");
  if (!(r->animal_speed > 0.0))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "animal_speed must be positive");
  for (int i = 0; i < r->n_patches; ++i)
    if (!std::isfinite(r->x[i]) || !std::isfinite(r->y[i]) || !std::isfinite(r->area[i]))
      return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite patch data");
  return ROADMC_OK;
}

// Only open-road segments count (roadCrossings, SpeciesRoadPatches.cpp:337-347).
int road_visible_segments(const roadmc_road_species* r) {
  int S = 0;
  for (int s = 0; s + 1 < r->n_road_points; ++s)
    if (!r->road_type || r->road_type[s] == ROADMC_SEG_ROAD) ++S;
  return S;
}

namespace {

inline size_t up4(size_t n) { return (n + 3) & ~(size_t)3; }

// Doubles of one staged block: x, y, weights (padded to 4), the segments as
// (x3, y3, x4, y4), the flows.
size_t staged_len(int P, int S, int C) { return up4(3 * (size_t)P) + 4 * (size_t)S + up4(C); }

}  // namespace

int road_reserve(roadmc_ctx* ctx, int P, int S, int C) {
  const size_t n = (size_t)P * P;
  auto& B = ctx->road_bufs;
  int rc;
  if ((rc = rmc_buf_ensure(ctx, B.a, staged_len(P, S, C) * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, B.c, n * sizeof(int32_t))) ||
      (rc = rmc_buf_ensure(ctx, B.d, n * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, B.e, std::max<size_t>((size_t)C * n, 1) * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, B.g, (size_t)road_cross_blocks(P) * sizeof(int))) ||
      (rc = rmc_buf_ensure(ctx, B.h, road_part_elems(P) * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, B.tail, sizeof(RoadTail))))
    return rc;
  return ROADMC_OK;
}

RoadTail* road_tail(roadmc_ctx* ctx) { return static_cast<RoadTail*>(ctx->road_bufs.tail.ptr); }

size_t road_stage(const roadmc_road_species* r, std::vector<double>& in) {
  const int P = r->n_patches, C = r->n_controls, S = road_visible_segments(r);
  const size_t off = in.size();
  in.resize(off + staged_len(P, S, C), 0.0);
  double* b = in.data() + off;
  // Transition rate (SpeciesRoadPatches.cpp:328-332) and the patch weights
  // A_j lambda e^(h_j) of the transition entries (:402-419).
  const double lambda = r->lambda_mult * r->lambda_sd + r->lambda_mean;
  std::memcpy(b, r->x, P * sizeof(double));
  std::memcpy(b + P, r->y, P * sizeof(double));
  for (int j = 0; j < P; ++j)
    b[2 * (size_t)P + j] = r->area[j] * lambda * std::exp(r->pref_mean[j] + r->pref_sd[j] * r->pref_mult);
  double* seg = b + up4(3 * (size_t)P);
  for (int s = 0, q = 0; s + 1 < r->n_road_points; ++s)
    if (!r->road_type || r->road_type[s] == ROADMC_SEG_ROAD) {
      seg[4 * q + 0] = r->road_x[s];
      seg[4 * q + 1] = r->road_y[s];
      seg[4 * q + 2] = r->road_x[s + 1];
      seg[4 * q + 3] = r->road_y[s + 1];
      ++q;
    }
  if (C > 0) std::memcpy(seg + 4 * (size_t)S, r->flow, C * sizeof(double));
  return off;
}

RoadScratch road_scratch(roadmc_ctx* ctx) {
  auto& B = ctx->road_bufs;
  return RoadScratch{static_cast<int*>(B.c.ptr), static_cast<double*>(B.d.ptr),
                     static_cast<double*>(B.e.ptr), static_cast<int*>(B.g.ptr),
                     static_cast<double*>(B.h.ptr)};
}

int road_reserve_lanes(roadmc_ctx* ctx, int P, int C) {
  const size_t n = (size_t)P * P;
  int rc;
  if (!ctx->road_ready)
    RMC_HIP(ctx, hipEventCreateWithFlags(&ctx->road_ready, hipEventDisableTiming));
  for (auto& l : ctx->road_lanes) {
    if (!l.stream) RMC_HIP(ctx, hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking));
    if (!l.done) RMC_HIP(ctx, hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
    if ((rc = rmc_buf_ensure(ctx, l.cross, n * sizeof(int32_t))) ||
        (rc = rmc_buf_ensure(ctx, l.trans, n * sizeof(double))) ||
        (rc = rmc_buf_ensure(ctx, l.surv, std::max<size_t>((size_t)C * n, 1) * sizeof(double))) ||
        (rc = rmc_buf_ensure(ctx, l.blockmax, (size_t)road_cross_blocks(P) * sizeof(int))) ||
        (rc = rmc_buf_ensure(ctx, l.part, road_part_elems(P) * sizeof(double))))
      return rc;
  }
  return ROADMC_OK;
}

int road_launch_staged(roadmc_ctx* ctx, const roadmc_road_species* r, const double* d_in,
                       bool want_trans, bool want_surv, const RoadScratch& sc, hipStream_t s,
                       RoadArgs* out) {
  const int P = r->n_patches, C = r->n_controls, S = road_visible_segments(r);
  // The two distance cut-offs, as the reference forms them
  // (SpeciesRoadPatches.cpp:328-332, 387-395).
  const double lambda = r->lambda_mult * r->lambda_sd + r->lambda_mean;
  const double cut_cross = std::log(0.05) / (-(r->lambda_mean + r->lambda_sd * r->lambda_mult));
  const double cut_trans = std::log(0.05) / (-(r->lambda_mean + r->lambda_sd * lambda));
  RoadArgs a{};
  a.P = P;
  a.S = S;
  a.C = C;
  a.x = d_in;
  a.y = d_in + P;
  a.pw = d_in + 2 * P;
  a.lambda = lambda;
  a.cut_cross = cut_cross;
  a.cut_trans = cut_trans;
  a.seg = reinterpret_cast<const double4*>(d_in + up4(3 * (size_t)P));
  a.form = r->transition_form == ROADMC_TRANS_DECAY ? 1 : 0;
  a.drop_last = r->drop_last_link ? 1 : 0;
  a.flow = d_in + up4(3 * (size_t)P) + 4 * (size_t)S;
  a.width_len = r->vehicle_width + r->animal_length;
  a.speed_3600 = r->animal_speed * 3600.0;
  a.cross = sc.cross;
  a.trans = sc.trans;
  a.surv = sc.surv;
  a.blockmax = sc.blockmax;
  a.part = sc.part;
  a.tail = road_tail(ctx);
  RMC_HIP(ctx, launch_road_build(a, want_trans, want_surv, s));
  *out = a;
  return ROADMC_OK;
}

int road_launch(roadmc_ctx* ctx, const roadmc_road_species* r, bool want_trans, bool want_surv,
                bool reset_tail, RoadStaging& st, RoadArgs* out) {
  st.in.clear();
  road_stage(r, st.in);
  hipStream_t s = ctx->stream;
  RMC_HIP(ctx, hipMemcpyAsync(ctx->road_bufs.a.ptr, st.in.data(), st.in.size() * sizeof(double),
                              hipMemcpyHostToDevice, s));
  if (reset_tail) {
    st.tail = RoadTail{0, {0, 0, 0}};
    RMC_HIP(ctx, hipMemcpyAsync(road_tail(ctx), &st.tail, sizeof(RoadTail),
                                hipMemcpyHostToDevice, s));
  }
  return road_launch_staged(ctx, r, static_cast<const double*>(ctx->road_bufs.a.ptr), want_trans,
                            want_surv, road_scratch(ctx), s, out);
}

}  // namespace roadmc

extern "C" int roadmc_build_species_road(roadmc_ctx* ctx, const roadmc_road_species* r,
                                         int32_t* crossings, double* trans, double* surv) {
  if (!ctx) return ROADMC_ERR_INVALID;
  int rc = roadmc::road_validate(ctx, r);
  if (rc) return rc;
  if (!crossings && !trans && !surv)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "no output requested");
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  const int P = r->n_patches, C = r->n_controls;
  if ((rc = roadmc::road_reserve(ctx, P, roadmc::road_visible_segments(r), C))) return rc;
  roadmc::RoadStaging st;
  roadmc::RoadArgs a;
  if ((rc = roadmc::road_launch(ctx, r, trans != nullptr, surv != nullptr, true, st, &a)))
    return rc;
  hipStream_t s = ctx->stream;
  roadmc::RoadTail tail;
  RMC_HIP(ctx, hipMemcpyAsync(&tail, roadmc::road_tail(ctx), sizeof(tail), hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipStreamSynchronize(s));
  if (trans && tail.bad)
    return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite transition weight");
  // The caller's buffers are written only once the build has succeeded.
  const size_t n = (size_t)P * P;
  if (crossings)
    RMC_HIP(ctx, hipMemcpyAsync(crossings, a.cross, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (trans)
    RMC_HIP(ctx, hipMemcpyAsync(trans, a.trans, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (surv && C > 0)
    RMC_HIP(ctx, hipMemcpyAsync(surv, a.surv, (size_t)C * n * sizeof(double),
                                hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipStreamSynchronize(s));
  return ROADMC_OK;
}
