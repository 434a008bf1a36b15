// libroadmc context and device-memory helpers (internal).
#ifndef ROADMC_CTX_H
#define ROADMC_CTX_H

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "roadmc.h"

struct rmc_devbuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct rmc_job_buffers {
  rmc_devbuf mpack, cols, partial, out, endpop, jobs, work, off, counter;
};

struct roadmc_ctx {
  int device = 0;
  int n_cu = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int fail_allocs = 0;  // test hook: next n allocations fail
  size_t lds_limit_f32 = 0, lds_limit_f64 = 0, lds_limit_b3 = 0;
  // Tile settings of the MTE launches per precision (0: FP32, 1: split bf16)
  // and patch-count bucket (P <= 128, 256, 512, 1024): rows of paths per
  // workgroup; whether remainder tiles may be split over k (1) or not (0);
  // workgroups per CU of the persistent launch (0: as many as LDS allows,
  // else 1 or 2).  Set by roadmc_set_tile or roadmc_autotune_ex.
  struct TileSetting {
    int rows, split, per_cu;
  };
  TileSetting tiles[2][4] = {{{256, 1, 0}, {128, 1, 0}, {64, 1, 0}, {32, 1, 0}},
                             {{256, 1, 0}, {128, 1, 0}, {64, 1, 0}, {32, 1, 0}}};
  rmc_job_buffers mte_bufs;
  struct {
    rmc_devbuf a, b, c, d, e, f, g, h;
  } rov_bufs;
  // Backward-induction graph of the last ROV shape; its kernels read their
  // parameters from rov_param (device), so one graph serves every problem of
  // that shape.
  struct {
    hipGraphExec_t exec = nullptr;
    long long key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    rmc_devbuf param;
  } rov_graph;
  // Adjoint backward-pass scratch, kept between calls: one segment's states
  // and multipliers (ns, np, mu), lambda, the gradients and their partials.
  // ckpt_spare holds the checkpoints of the last destroyed adjoint handle for
  // the next forward pass to reuse (several GB at full size).
  struct {
    rmc_devbuf ns, np, mu, lam, gradr, gradm, gpart, rpart, ckpt_spare;
  } adj_bufs;
  // MTE and the species-state forward pass beyond the LDS tile: the
  // destination-major movement matrices, the per-patch columns, the two
  // state buffers, the growth-rate state and the per-block row sums.
  struct {
    rmc_devbuf mt, cols, s0, s1, grm, rs, tot;
  } wide_bufs;
  struct {
    rmc_devbuf a, b, c, d, e, f, g, h, tail, in;
  } road_bufs;
  // Side streams (default priority) on which the road builds of one batch
  // overlap, each with its own scratch; created on first use.
  static constexpr int kRoadLanes = 4;
  struct {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    rmc_devbuf cross, trans, surv, blockmax, part;
  } road_lanes[kRoadLanes];
  hipEvent_t road_ready = nullptr;
  // Pinned host staging of the packed operands (grown on demand, kept).
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  struct {
    rmc_devbuf a, b, c, d;
  } knn_bufs;
  struct {
    rmc_devbuf mpack, cols, prices, controls, tp, aars, profit, V, optc, adj, flag, blk, xs, ys,
        u, small, grid, args, ckpt, grm, rvec, grids;
  } rovs_bufs;
  // The smoothed-validity gradient of the species-state valuation: the
  // transposed movement matrices, the policy, dJ/dA coefficients, the w
  // plane, the contraction's partials and result, the per-block values, the
  // rows of each control and the counts that place them.
  struct {
    rmc_devbuf mtpack, policy, coef, w, gpart, grad, jpart, rows, blk;
  } rovs_grad_bufs;
};

int rmc_fail(roadmc_ctx* ctx, int code, const std::string& msg);
int rmc_malloc(roadmc_ctx* ctx, void** p, size_t bytes);
int rmc_buf_ensure(roadmc_ctx* ctx, rmc_devbuf& b, size_t bytes);
// ctx->pinned holds at least bytes (hipHostMalloc); the old contents are lost.
int rmc_pinned_ensure(roadmc_ctx* ctx, size_t bytes);
int rmc_buf_alloc_exact(roadmc_ctx* ctx, rmc_devbuf& b, size_t bytes);
void rmc_buf_free(rmc_devbuf& b);

#define RMC_HIP(ctx, expr)                                                  \
  do {                                                                      \
    const hipError_t rmc_e_ = (expr);                                       \
    if (rmc_e_ != hipSuccess)                                               \
      return rmc_fail((ctx), ROADMC_ERR_HIP,                                \
                      std::string(#expr) + ": " + hipGetErrorString(rmc_e_)); \
  } while (0)

#endif
