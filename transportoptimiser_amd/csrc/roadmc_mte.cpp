// C-ABI host code for the MTE path: problem validation, packing of the
// movement matrices into MFMA operand order, launch of the tiled kernels, the
// batched/graph path and the reference-design baseline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "ctx.h"
#include "road_host.h"
#include "mte_kernels.h"
#include "rng.h"
#include "roadmc.h"
#include "mte_wide.h"

using roadmc::MteJob;

namespace {

constexpr int kBuckets = 4;
const int kBucketMax[kBuckets] = {128, 256, 512, 1024};
constexpr size_t kLdsPerCu = 160 * 1024;  // gfx950 LDS per CU

int round_up(int x, int m) { return (x + m - 1) / m * m; }

int bucket_of(int Pp) {
  for (int b = 0; b < kBuckets; ++b)
    if (Pp <= kBucketMax[b]) return b;
  return kBuckets - 1;
}

bool finite_array(const double* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

// Whether a species' host matrices (trans, and surv when given) are finite;
// true when there are none or its patch count is out of range (the other
// checks reject those first).
bool species_matrices_finite(const roadmc_species& sp, int max_patches) {
  const int P = sp.n_patches;
  if (!sp.trans || P < 1 || P > max_patches) return true;
  return finite_array(sp.trans, (size_t)P * P) &&
         (!sp.surv || finite_array(sp.surv, (size_t)P * P));
}

// fn(i0, i1) over contiguous ranges of [0, n) on up to 16 host threads
// (serially when no thread can be had).
template <class F>
void parallel_ranges(size_t n, const F& fn) {
  const size_t nth = std::min<size_t>(
      std::max<size_t>(n, 1), std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  std::vector<std::thread> pool;
  try {
    for (size_t t = 1; t < nth; ++t) pool.emplace_back(fn, n * t / nth, n * (t + 1) / nth);
    fn(0, n / nth);
  } catch (...) {
    for (auto& th : pool) th.join();
    pool.clear();
    fn(0, n);
    return;
  }
  for (auto& th : pool) th.join();
}

// species_matrices_finite for every species of roads[0 .. n_roads), in
// chunks of at most 2^20 entries over host threads: ok[off[r] + s] for
// species s of road r (off[r + 1] - off[r] species, none for a road whose
// species list is invalid).
void matrices_finite(const roadmc_mte_problem* roads, int n_roads, int max_patches,
                     std::vector<size_t>& off, std::vector<char>& ok) {
  off.assign((size_t)n_roads + 1, 0);
  for (int r = 0; r < n_roads; ++r) {
    const int ns = roads[r].species && roads[r].n_species >= 1 && roads[r].n_species <= 63
                       ? roads[r].n_species
                       : 0;
    off[r + 1] = off[r] + (size_t)ns;
  }
  ok.assign(std::max<size_t>(off[n_roads], 1), 1);
  struct Chunk {
    size_t flag;
    const double* p;
    size_t n;
  };
  constexpr size_t kChunk = (size_t)1 << 20;
  std::vector<Chunk> chunks;
  for (int r = 0; r < n_roads; ++r)
    for (size_t s = 0; s < off[r + 1] - off[r]; ++s) {
      const roadmc_species& sp = roads[r].species[s];
      const int P = sp.n_patches;
      if (!sp.trans || P < 1 || P > max_patches) continue;
      const size_t n = (size_t)P * P;
      for (const double* m : {sp.trans, sp.surv}) {
        if (!m) continue;
        for (size_t i = 0; i < n; i += kChunk)
          chunks.push_back({off[r] + s, m + i, std::min(kChunk, n - i)});
      }
    }
  std::vector<char> chunk_ok(std::max<size_t>(chunks.size(), 1), 1);
  parallel_ranges(chunks.size(), [&](size_t i0, size_t i1) {
    for (size_t i = i0; i < i1; ++i) chunk_ok[i] = finite_array(chunks[i].p, chunks[i].n) ? 1 : 0;
  });
  for (size_t i = 0; i < chunks.size(); ++i)
    if (!chunk_ok[i]) ok[chunks[i].flag] = 0;
}

// mats_finite, when given, holds species_matrices_finite for each species
// (computed beforehand, e.g. in parallel); the checks and their order are the
// same either way.
int validate_problem(roadmc_ctx* ctx, const roadmc_mte_problem* pr,
                     int max_patches = ROADMC_MAX_PATCHES, const char* mats_finite = nullptr) {
  if (!pr) return rmc_fail(ctx, ROADMC_ERR_INVALID, "problem is NULL");
  if (pr->n_species < 1 || pr->n_species > 63 || !pr->species)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_species must be in [1, 63]");
  if (pr->n_years < 1)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_years must be >= 1");
  if (pr->n_paths < 1 || pr->n_paths > 0x7fffffffLL)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "n_paths must be in [1, 2^31)");
  if (pr->growth_law < ROADMC_LAW_LOGNORMAL || pr->growth_law > ROADMC_LAW_REF_GPU)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "unknown growth_law");
  if (pr->endpoint != ROADMC_END_POST_GROWTH && pr->endpoint != ROADMC_END_PRE_GROWTH)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "unknown endpoint");
  if (pr->precision != ROADMC_PREC_FP32 && pr->precision != ROADMC_PREC_FP64 &&
      pr->precision != ROADMC_PREC_BF16X3)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "unknown precision");
  if (!(pr->drop_tol >= 0.0) || !std::isfinite(pr->time_step))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "drop_tol/time_step invalid");
  if (pr->road_id >= (1u << 24))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "road_id must be < 2^24");
  for (int s = 0; s < pr->n_species; ++s) {
    const roadmc_species& sp = pr->species[s];
    const int P = sp.n_patches;
    if (P < 1 || P > max_patches)
      return rmc_fail(ctx, ROADMC_ERR_INVALID,
                      max_patches > ROADMC_MAX_PATCHES
                          ? "n_patches must be in [1, ROADMC_MAX_PATCHES_WIDE]"
                          : "n_patches must be in [1, ROADMC_MAX_PATCHES] (roadmc_mte_run in "
                            "FP32 takes up to ROADMC_MAX_PATCHES_WIDE)");
    if (!sp.init_pop || !sp.capacity || (!sp.trans && !sp.road))
      return rmc_fail(ctx, ROADMC_ERR_INVALID,
                      "init_pop, capacity and trans (or a road) are required");
    if (!sp.trans) {
      // Matrices built on the device from the species' road.
      int rc = roadmc::road_validate(ctx, sp.road);
      if (rc) return rc;
      if (sp.road->n_patches != P)
        return rmc_fail(ctx, ROADMC_ERR_INVALID, "road n_patches differs from the species'");
      if (sp.road_control >= sp.road->n_controls)
        return rmc_fail(ctx, ROADMC_ERR_INVALID, "road_control out of range");
    }
    for (int j = 0; j < P; ++j)
      if (!(sp.capacity[j] > 0.0) || !std::isfinite(sp.capacity[j]))
        return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "capacity must be finite and > 0");
    const bool mats_ok = mats_finite ? mats_finite[s] != 0 : species_matrices_finite(sp, max_patches);
    if (!finite_array(sp.init_pop, P) || !mats_ok ||
        (sp.growth && !finite_array(sp.growth, P)) || !finite_array(sp.params, 8) ||
        !std::isfinite(sp.noise_sd) || !std::isfinite(sp.threshold_ratio))
      return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite input");
  }
  return ROADMC_OK;
}

// Movement-and-mortality entry (origin k, destination j) as the reference
// forms it: T o S (SimulateGPU.cu:3783-3787), with dense2Sparse's drop rule
// applied to the binary32 value (SimulateGPU.cu:5985).
inline double movement(const roadmc_species& sp, int P, int k, int j, double tol) {
  const size_t cm = (size_t)j * P + k;
  double v = sp.trans[cm] * (sp.surv ? sp.surv[cm] : 1.0);
  if (tol > 0.0 && !((float)v > (float)tol)) v = 0.0;
  return v;
}

// Packs the movement matrix (Pp output columns, Kp contraction rows) in MFMA
// B-operand order and the per-patch columns K, 1/K, r, n0.
template <typename T>
void pack_species(const roadmc_species& sp, int Pp, int Kp, double tol, bool fp64, T* mp,
                  T* cols) {
  const int P = sp.n_patches;
  // A species given by its road is packed on the device over the whole
  // region (padding included); the host leaves its region untouched.
  if (sp.trans) {
    std::fill(mp, mp + (size_t)Pp * Kp, (T)0);
    for (int j = 0; j < P; ++j)    // destination: one contiguous column of trans and surv
      for (int k = 0; k < P; ++k)
        mp[roadmc::mte_pack_index(fp64, Kp, k, j)] = (T)movement(sp, P, k, j, tol);
  }
  for (int j = 0; j < Pp; ++j) {
    const bool in = j < P;
    cols[j] = in ? (T)sp.capacity[j] : (T)1;
    cols[Pp + j] = in ? (T)(1.0 / sp.capacity[j]) : (T)1;
    cols[2 * Pp + j] = (in && sp.growth) ? (T)sp.growth[j] : (T)0;
    cols[3 * Pp + j] = in ? (T)sp.init_pop[j] : (T)0;
  }
}

// Row of ctx->tiles for a precision (FP64 has none: its rows follow LDS).
int tile_slot(int prec) { return prec == roadmc::kPrecB3 ? 1 : 0; }

int rows_for(const roadmc_ctx* ctx, int prec, int Pp, int64_t N) {
  if (prec == ROADMC_PREC_FP64) {
    int bm = (roadmc::kMteTileElemsF64 / Pp) / roadmc::kMteRowsF64 * roadmc::kMteRowsF64;
    bm = std::max(roadmc::kMteRowsF64, std::min(bm, 256));
    const int cap = round_up((int)std::min<int64_t>(N, 256), roadmc::kMteRowsF64);
    return std::min(bm, cap);
  }
  int bm = ctx->tiles[tile_slot(prec)][bucket_of(Pp)].rows;
  const int lim = (roadmc::kMteTileElemsF32 / Pp) / roadmc::kMteRowsF32 * roadmc::kMteRowsF32;
  bm = std::max(roadmc::kMteRowsF32, std::min(bm, lim));
  const int cap = round_up((int)std::min<int64_t>(N, 256), roadmc::kMteRowsF32);
  return std::min(bm, cap);
}

// Host-side description of a set of (road, species) jobs, shared by the
// single-road and batched entry points.
struct JobSet {
  int prec = roadmc::kPrecF32;
  bool fp64 = false;
  std::vector<MteJob> jobs;
  std::vector<int> road_of, species_of;
  std::vector<int64_t> mpack_off, cols_off, part_off, out_off, endpop_off;
  std::vector<int2> work;
  std::vector<int> qbeg;  // per-XCD queue boundaries into work (persistent)
  size_t mpack_elems = 0, cols_elems = 0, part_elems = 0, out_elems = 0,
         endpop_elems = 0;
  size_t lds = 0;
  int tpw = 1;  // most output tiles per wave over the jobs
};

void plan_jobs(const roadmc_ctx* ctx, const roadmc_mte_problem* roads, int n_roads,
               bool want_end_pops, JobSet& js) {
  js.prec = roads[0].precision;
  js.fp64 = js.prec == ROADMC_PREC_FP64;
  for (int r = 0; r < n_roads; ++r) {
    const roadmc_mte_problem& pr = roads[r];
    for (int s = 0; s < pr.n_species; ++s) {
      const roadmc_species& sp = pr.species[s];
      MteJob jb{};
      jb.P = sp.n_patches;
      jb.Pp = round_up(jb.P, 32);
      // The split-bf16 products take 16-k chunks two at a time.
      jb.Pk = js.prec == roadmc::kPrecB3 ? jb.Pp : round_up(jb.P, 16);
      jb.N = (int32_t)pr.n_paths;
      jb.T = pr.n_years;
      jb.BM = rows_for(ctx, js.prec, jb.Pp, pr.n_paths);
      jb.law = pr.growth_law;
      jb.endpoint = pr.endpoint;
      jb.n_tiles = (int)((pr.n_paths + jb.BM - 1) / jb.BM);
      for (int i = 0; i < 8; ++i) jb.params[i] = sp.params[i];
      jb.sigma = sp.noise_sd;
      jb.dt = pr.time_step;
      double tot0 = 0.0;
      for (int j = 0; j < jb.P; ++j) tot0 += sp.init_pop[j];
      jb.thr_total = sp.threshold_ratio * tot0;
      jb.tag0 = rmc_tag(pr.road_id, (uint32_t)s, RMC_KIND_PATCH);
      jb.tag1 = rmc_tag(pr.road_id, (uint32_t)s, RMC_KIND_PATH);
      jb.key0 = (uint32_t)(pr.seed & 0xffffffffu);
      jb.key1 = (uint32_t)(pr.seed >> 32);
      js.mpack_off.push_back((int64_t)js.mpack_elems);
      js.cols_off.push_back((int64_t)js.cols_elems);
      js.part_off.push_back((int64_t)js.part_elems);
      js.out_off.push_back((int64_t)js.out_elems);
      js.endpop_off.push_back(want_end_pops ? (int64_t)js.endpop_elems : -1);
      js.mpack_elems += (size_t)jb.Pp * jb.Pk;
      js.cols_elems += 4 * (size_t)jb.Pp;
      js.part_elems += (size_t)jb.n_tiles * (3 + jb.Pp);
      js.out_elems += 3 + (size_t)jb.Pp;
      if (want_end_pops) js.endpop_elems += (size_t)jb.N;
      // Remainder tiles (ntiles % 4) would leave one SIMD with more work
      // than the others; split them into k-quarters when LDS has room for
      // the partial products (FP32 kernel, at least 8 k-groups).
      const int ntiles = (jb.BM / roadmc::kMteRowsF32) * (jb.Pp / 32);
      const int rem = ntiles % 4;
      jb.split = 0;
      if (!js.fp64 && rem && ctx->tiles[tile_slot(js.prec)][bucket_of(jb.Pp)].split &&
          (jb.Pk >> 3) >= 8 &&
          roadmc::mte_lds_bytes(js.prec, jb.BM, jb.Pp, rem) <= kLdsPerCu)
        jb.split = 1;
      js.lds = std::max(js.lds, roadmc::mte_lds_bytes(js.prec, jb.BM, jb.Pp,
                                                      jb.split ? rem : 0));
      js.tpw = std::max(js.tpw, roadmc::mte_tiles_per_wave(js.prec, jb.BM, jb.Pp));
      js.jobs.push_back(jb);
      js.road_of.push_back(r);
      js.species_of.push_back(s);
    }
  }
  // Jobs in decreasing cost (cost ~ n_tiles * BM * Pp^2 * T), assigned to
  // the per-XCD queues by longest-processing-time-first; each queue lists
  // its jobs' tiles contiguously.
  std::vector<int> order(js.jobs.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  auto cost = [&](int j) {
    const MteJob& x = js.jobs[j];
    return (double)x.n_tiles * x.BM * x.Pp * x.Pp * x.T;
  };
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return cost(a) > cost(b); });
  const int nq = roadmc::kMteQueues;
  std::vector<std::vector<int>> queues(nq);
  std::vector<double> load(nq, 0.0);
  for (int j : order) {
    const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    queues[q].push_back(j);
    load[q] += cost(j);
  }
  js.qbeg.assign(1, 0);
  for (int q = 0; q < nq; ++q) {
    for (int j : queues[q])
      for (int t = 0; t < js.jobs[j].n_tiles; ++t) js.work.push_back(make_int2(j, t));
    js.qbeg.push_back((int)js.work.size());
  }
}

// Device buffers of one job set.
using JobBuffers = rmc_job_buffers;

template <typename T>
int upload_jobs(roadmc_ctx* ctx, const roadmc_mte_problem* roads, JobSet& js,
                JobBuffers& b, bool persistent_bufs) {
  const size_t es = sizeof(T);
  int rc;
  auto grow = [&](rmc_devbuf& buf, size_t bytes) {
    return persistent_bufs ? rmc_buf_alloc_exact(ctx, buf, bytes)
                           : rmc_buf_ensure(ctx, buf, bytes);
  };
  if ((rc = grow(b.mpack, js.mpack_elems * es)) ||
      (rc = grow(b.cols, js.cols_elems * es)) ||
      (rc = grow(b.partial, js.part_elems * sizeof(double))) ||
      (rc = grow(b.out, js.out_elems * sizeof(double))) ||
      (rc = grow(b.endpop, std::max<size_t>(js.endpop_elems, 1) * sizeof(double))) ||
      (rc = grow(b.jobs, js.jobs.size() * sizeof(MteJob))) ||
      (rc = grow(b.work, js.work.size() * sizeof(int2))) ||
      (rc = grow(b.off, js.jobs.size() * sizeof(int64_t))) ||
      (rc = grow(b.counter, 256)))
    return rc;
  // Host packing into the context's pinned staging (so that the upload runs
  // at DMA speed), the jobs split over threads by matrix size; every job
  // writes its own contiguous regions, which tile both arrays.
  const size_t mp_bytes = (js.mpack_elems * es + 255) / 256 * 256;
  if ((rc = rmc_pinned_ensure(ctx, mp_bytes + std::max<size_t>(js.cols_elems, 1) * es))) return rc;
  T* const mp_h = static_cast<T*>(ctx->pinned);
  T* const cols_h = reinterpret_cast<T*>(static_cast<char*>(ctx->pinned) + mp_bytes);
  const size_t nj = js.jobs.size();
  auto pack_range = [&](size_t j0, size_t j1) {
    for (size_t j = j0; j < j1; ++j) {
      const roadmc_mte_problem& pr = roads[js.road_of[j]];
      pack_species<T>(pr.species[js.species_of[j]], js.jobs[j].Pp, js.jobs[j].Pk, pr.drop_tol,
                      js.fp64, mp_h + js.mpack_off[j], cols_h + js.cols_off[j]);
    }
  };
  const int nth = (int)std::min<size_t>(
      nj, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  std::vector<size_t> bounds(nth + 1, nj);
  bounds[0] = 0;
  for (size_t j = 0, acc = 0, t = 1; j < nj && t < (size_t)nth; ++j) {
    acc += (size_t)js.jobs[j].Pp * js.jobs[j].Pk + 4096;
    if (acc * nth >= (js.mpack_elems + 4096 * nj) * t) bounds[t++] = j + 1;
  }
  {
    std::vector<std::thread> pool;
    try {
      for (int t = 1; t < nth; ++t) pool.emplace_back(pack_range, bounds[t], bounds[t + 1]);
      pack_range(bounds[0], bounds[1]);
    } catch (...) {
      for (auto& th : pool) th.join();
      pool.clear();
      pack_range(0, nj);  // no threads to be had: pack serially
    }
    for (auto& th : pool) th.join();
  }
  for (size_t j = 0; j < nj; ++j) {
    MteJob& jb = js.jobs[j];
    jb.mpack = static_cast<T*>(b.mpack.ptr) + js.mpack_off[j];
    jb.cols = static_cast<T*>(b.cols.ptr) + js.cols_off[j];
    jb.partial = static_cast<double*>(b.partial.ptr) + js.part_off[j];
    jb.end_pops = js.endpop_off[j] >= 0
                      ? static_cast<double*>(b.endpop.ptr) + js.endpop_off[j]
                      : nullptr;
  }
  hipStream_t s = ctx->stream;
  RMC_HIP(ctx, hipMemcpyAsync(b.mpack.ptr, mp_h, js.mpack_elems * es,
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(b.cols.ptr, cols_h, js.cols_elems * es,
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(b.jobs.ptr, js.jobs.data(), js.jobs.size() * sizeof(MteJob),
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(b.work.ptr, js.work.data(), js.work.size() * sizeof(int2),
                              hipMemcpyHostToDevice, s));
  RMC_HIP(ctx, hipMemcpyAsync(b.off.ptr, js.out_off.data(),
                              js.out_off.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
  // Queue boundaries live after the 16 queue counters in the counter buffer.
  RMC_HIP(ctx, hipMemcpyAsync(static_cast<int*>(b.counter.ptr) + 16, js.qbeg.data(),
                              js.qbeg.size() * sizeof(int), hipMemcpyHostToDevice, s));
  // Species given by a road: build trans and surv on the device and pack
  // them into their job's slot (after the upload above).  Every road's
  // inputs go up in one copy; the builds then run one after the other in
  // the shared scratch.
  int maxP = 0, maxC = 0;
  for (size_t j = 0; j < js.jobs.size(); ++j) {
    const roadmc_species& sp = roads[js.road_of[j]].species[js.species_of[j]];
    if (sp.trans) continue;
    maxP = std::max(maxP, sp.road->n_patches);
    maxC = std::max(maxC, sp.road->n_controls);
  }
  std::vector<double> road_in;
  std::vector<size_t> road_off(js.jobs.size(), 0);
  roadmc::RoadTail tail0{0, {0, 0, 0}};
  if (maxP > 0) {
    if ((rc = rmc_buf_ensure(ctx, ctx->road_bufs.tail, sizeof(roadmc::RoadTail)))) return rc;
    for (size_t j = 0; j < js.jobs.size(); ++j) {
      const roadmc_species& sp = roads[js.road_of[j]].species[js.species_of[j]];
      if (!sp.trans) road_off[j] = roadmc::road_stage(sp.road, road_in);
    }
    if ((rc = rmc_buf_ensure(ctx, ctx->road_bufs.in, road_in.size() * sizeof(double)))) return rc;
    const double* d_in = static_cast<const double*>(ctx->road_bufs.in.ptr);
    RMC_HIP(ctx, hipMemcpyAsync(ctx->road_bufs.in.ptr, road_in.data(),
                                road_in.size() * sizeof(double), hipMemcpyHostToDevice, s));
    RMC_HIP(ctx, hipMemcpyAsync(roadmc::road_tail(ctx), &tail0, sizeof(tail0),
                                hipMemcpyHostToDevice, s));
    // The builds run on the side lanes, round robin, each lane in its own
    // scratch and after the uploads above; ctx's stream waits for all of
    // them before anything reads the packed operands.
    if ((rc = roadmc::road_reserve_lanes(ctx, maxP, maxC))) return rc;
    RMC_HIP(ctx, hipEventRecord(ctx->road_ready, s));
    for (auto& l : ctx->road_lanes) RMC_HIP(ctx, hipStreamWaitEvent(l.stream, ctx->road_ready, 0));
    int lane = 0;
    for (size_t j = 0; j < js.jobs.size(); ++j) {
      const roadmc_mte_problem& pr = roads[js.road_of[j]];
      const roadmc_species& sp = pr.species[js.species_of[j]];
      if (sp.trans) continue;
      const roadmc_road_species* rd = sp.road;
      const int control = sp.road_control < 0 ? rd->n_controls - 1 : sp.road_control;
      auto& L = ctx->road_lanes[lane];
      lane = (lane + 1) % roadmc_ctx::kRoadLanes;
      const roadmc::RoadScratch sc{static_cast<int*>(L.cross.ptr), static_cast<double*>(L.trans.ptr),
                                   static_cast<double*>(L.surv.ptr), static_cast<int*>(L.blockmax.ptr),
                                   static_cast<double*>(L.part.ptr)};
      roadmc::RoadArgs a;
      if ((rc = roadmc::road_launch_staged(ctx, rd, d_in + road_off[j], true, control >= 0, sc,
                                           L.stream, &a)))
        return rc;
      const double* surv = control >= 0 ? a.surv + (size_t)control * rd->n_patches * rd->n_patches
                                        : nullptr;
      T* dst = static_cast<T*>(b.mpack.ptr) + js.mpack_off[j];
      RMC_HIP(ctx, roadmc::launch_road_pack(js.fp64, a.trans, surv, rd->n_patches, js.jobs[j].Pp,
                                            js.jobs[j].Pk, pr.drop_tol, dst, L.stream));
    }
    for (auto& l : ctx->road_lanes) {
      RMC_HIP(ctx, hipEventRecord(l.done, l.stream));
      RMC_HIP(ctx, hipStreamWaitEvent(s, l.done, 0));
    }
  }
  // Split-bf16 operands are converted in place from the FP32 packing.
  if (js.prec == roadmc::kPrecB3)
    RMC_HIP(ctx, roadmc::launch_mte_split_b3(static_cast<float*>(b.mpack.ptr), js.mpack_elems, s));
  roadmc::RoadTail tail{0, {0, 0, 0}};
  if (maxP > 0)
    RMC_HIP(ctx, hipMemcpyAsync(&tail, roadmc::road_tail(ctx), sizeof(tail),
                                hipMemcpyDeviceToHost, s));
  // The host vectors die at return: wait for the copies.
  RMC_HIP(ctx, hipStreamSynchronize(s));
  if (tail.bad) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite transition weight");
  return ROADMC_OK;
}

int ensure_lds_limit(roadmc_ctx* ctx, int prec, size_t lds) {
  size_t& cur = prec == roadmc::kPrecF64  ? ctx->lds_limit_f64
                : prec == roadmc::kPrecB3 ? ctx->lds_limit_b3
                                          : ctx->lds_limit_f32;
  if (lds > cur) {
    RMC_HIP(ctx, roadmc::mte_set_lds_limit(prec, lds));
    cur = lds;
  }
  return ROADMC_OK;
}

// Writes caller outputs from the reduced per-job sums; nothing is written
// before every device step has succeeded.
void write_results(const JobSet& js, const std::vector<double>& out,
                   const std::vector<double>& endpop, const roadmc_mte_problem* roads,
                   roadmc_mte_result* results) {
  std::vector<int64_t> patch_base(js.jobs.size());
  {
    int last_road = -1;
    int64_t acc = 0;
    for (size_t j = 0; j < js.jobs.size(); ++j) {
      if (js.road_of[j] != last_road) {
        acc = 0;
        last_road = js.road_of[j];
      }
      patch_base[j] = acc;
      acc += js.jobs[j].P;
    }
  }
  for (size_t j = 0; j < js.jobs.size(); ++j) {
    const MteJob& jb = js.jobs[j];
    const roadmc_mte_result& res = results[js.road_of[j]];
    const int s = js.species_of[j];
    const double* o = out.data() + js.out_off[j];
    const double N = (double)jb.N;
    const double mean = o[1] / N;
    const double var = std::max(0.0, o[2] / N - mean * mean);
    if (res.mte_prob) res.mte_prob[s] = o[0] / N;
    if (res.end_mean) res.end_mean[s] = mean;
    if (res.end_sd) res.end_sd[s] = std::sqrt(var);
    if (res.patch_mean)
      for (int p = 0; p < jb.P; ++p) res.patch_mean[patch_base[j] + p] = o[3 + p] / N;
    if (res.end_pops && js.endpop_off[j] >= 0)
      std::memcpy(res.end_pops + (size_t)s * jb.N, endpop.data() + js.endpop_off[j],
                  sizeof(double) * jb.N);
  }
  (void)roads;
}

int fetch_results(roadmc_ctx* ctx, const JobSet& js, JobBuffers& b,
                  const roadmc_mte_problem* roads, roadmc_mte_result* results) {
  std::vector<double> out(js.out_elems), endpop(js.endpop_elems);
  hipStream_t s = ctx->stream;
  RMC_HIP(ctx, hipMemcpyAsync(out.data(), b.out.ptr, out.size() * sizeof(double),
                              hipMemcpyDeviceToHost, s));
  if (!endpop.empty())
    RMC_HIP(ctx, hipMemcpyAsync(endpop.data(), b.endpop.ptr, endpop.size() * sizeof(double),
                                hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipStreamSynchronize(s));
  for (double v : out)
    if (!std::isfinite(v))
      return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite population (diverging growth law?)");
  write_results(js, out, endpop, roads, results);
  return ROADMC_OK;
}

bool wants_end_pops(const roadmc_mte_result* results, int n) {
  for (int i = 0; i < n; ++i)
    if (results[i].end_pops) return true;
  return false;
}

}  // namespace

// ---------------------------------------------------------------------------
// More patches than one workgroup's LDS tile takes: the state of every path
// in HBM, one blocked product per year (mte_wide.hip), species one after the
// other; the tallies and the results go through the tiled path's partial
// layout, reduction and result code.

namespace {

int mte_run_wide(roadmc_ctx* ctx, const roadmc_mte_problem* pr, roadmc_mte_result* result) {
  if (pr->precision != ROADMC_PREC_FP32)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED,
                    "more than ROADMC_MAX_PATCHES patches: FP32 products only");
  constexpr int kB = roadmc::kWideBlock;
  const int64_t N = pr->n_paths;
  const int S = pr->n_species;
  JobSet js;
  js.prec = ROADMC_PREC_FP32;
  js.fp64 = false;
  const bool want_end = result->end_pops != nullptr;
  int maxPp = 0;
  for (int s = 0; s < S; ++s) {
    const roadmc_species& sp = pr->species[s];
    MteJob jb{};
    jb.P = sp.n_patches;
    jb.Pp = round_up(jb.P, kB);
    jb.Pk = jb.Pp;
    jb.N = (int32_t)N;
    jb.T = pr->n_years;
    jb.BM = kB;
    jb.law = pr->growth_law;
    jb.endpoint = pr->endpoint;
    jb.n_tiles = (int)((N + kB - 1) / kB);
    for (int i = 0; i < 8; ++i) jb.params[i] = sp.params[i];
    jb.sigma = sp.noise_sd;
    jb.dt = pr->time_step;
    double tot0 = 0.0;
    for (int j = 0; j < jb.P; ++j) tot0 += sp.init_pop[j];
    jb.thr_total = sp.threshold_ratio * tot0;
    jb.tag0 = rmc_tag(pr->road_id, (uint32_t)s, RMC_KIND_PATCH);
    jb.tag1 = rmc_tag(pr->road_id, (uint32_t)s, RMC_KIND_PATH);
    jb.key0 = (uint32_t)(pr->seed & 0xffffffffu);
    jb.key1 = (uint32_t)(pr->seed >> 32);
    js.part_off.push_back((int64_t)js.part_elems);
    js.out_off.push_back((int64_t)js.out_elems);
    js.endpop_off.push_back(want_end ? (int64_t)js.endpop_elems : -1);
    js.part_elems += (size_t)jb.n_tiles * (3 + jb.Pp);
    js.out_elems += 3 + (size_t)jb.Pp;
    if (want_end) js.endpop_elems += (size_t)N;
    js.jobs.push_back(jb);
    js.road_of.push_back(0);
    js.species_of.push_back(s);
    maxPp = std::max(maxPp, jb.Pp);
  }
  JobBuffers& b = ctx->mte_bufs;
  auto& W = ctx->wide_bufs;
  const size_t plane = (size_t)N * maxPp;
  int rc;
  if ((rc = rmc_buf_ensure(ctx, b.partial, js.part_elems * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, b.out, js.out_elems * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, b.endpop, std::max<size_t>(js.endpop_elems, 1) * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, b.jobs, js.jobs.size() * sizeof(MteJob))) ||
      (rc = rmc_buf_ensure(ctx, b.off, js.jobs.size() * sizeof(int64_t))) ||
      (rc = rmc_buf_ensure(ctx, W.mt, (size_t)maxPp * maxPp * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, W.cols, 4 * (size_t)maxPp * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, W.s0, plane * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, W.s1, plane * sizeof(float))) ||
      (rc = rmc_buf_ensure(ctx, W.grm, (size_t)N * sizeof(float))))
    return rc;
  for (size_t j = 0; j < js.jobs.size(); ++j) {
    MteJob& jb = js.jobs[j];
    jb.partial = static_cast<double*>(b.partial.ptr) + js.part_off[j];
    jb.end_pops = js.endpop_off[j] >= 0 ? static_cast<double*>(b.endpop.ptr) + js.endpop_off[j]
                                        : nullptr;
  }
  hipStream_t st = ctx->stream;
  RMC_HIP(ctx, hipMemcpyAsync(b.jobs.ptr, js.jobs.data(), js.jobs.size() * sizeof(MteJob),
                              hipMemcpyHostToDevice, st));
  RMC_HIP(ctx, hipMemcpyAsync(b.off.ptr, js.out_off.data(), js.out_off.size() * sizeof(int64_t),
                              hipMemcpyHostToDevice, st));
  const double l2e = 1.4426950408889634;
  for (int s = 0; s < S; ++s) {
    const roadmc_species& sp = pr->species[s];
    const MteJob& jb = js.jobs[(size_t)s];
    const int P = jb.P, Pp = jb.Pp;
    float* mt = static_cast<float*>(W.mt.ptr);
    // The previous species' uploads leave the pinned staging before it is
    // refilled.
    RMC_HIP(ctx, hipStreamSynchronize(st));
    const size_t mt_bytes = sp.trans ? (size_t)Pp * Pp * sizeof(float) : 0;
    if ((rc = rmc_pinned_ensure(ctx, mt_bytes + 4 * (size_t)Pp * sizeof(float)))) return rc;
    float* cols_h = reinterpret_cast<float*>(static_cast<char*>(ctx->pinned) + mt_bytes);
    if (sp.trans) {
      float* mt_h = static_cast<float*>(ctx->pinned);
      // Destination rows over host threads (each writes its own rows).
      parallel_ranges((size_t)Pp, [&](size_t j0, size_t j1) {
        for (int j = (int)j0; j < (int)j1; ++j)
          for (int k = 0; k < Pp; ++k)
            mt_h[(size_t)j * Pp + k] =
                (j < P && k < P) ? (float)movement(sp, P, k, j, pr->drop_tol) : 0.f;
      });
      RMC_HIP(ctx, hipMemcpyAsync(mt, mt_h, mt_bytes, hipMemcpyHostToDevice, st));
    } else {
      const roadmc_road_species* rd = sp.road;
      const int control = sp.road_control < 0 ? rd->n_controls - 1 : sp.road_control;
      if ((rc = roadmc::road_reserve(ctx, P, roadmc::road_visible_segments(rd), rd->n_controls)))
        return rc;
      roadmc::RoadStaging stg;
      roadmc::RoadArgs ra;
      if ((rc = roadmc::road_launch(ctx, rd, true, control >= 0, true, stg, &ra))) return rc;
      const double* surv =
          control >= 0 ? ra.surv + (size_t)control * P * P : nullptr;
      RMC_HIP(ctx, roadmc::launch_mte_wide_pack(ra.trans, surv, P, Pp, pr->drop_tol, mt, st));
      roadmc::RoadTail tail;
      RMC_HIP(ctx, hipMemcpyAsync(&tail, roadmc::road_tail(ctx), sizeof(tail),
                                  hipMemcpyDeviceToHost, st));
      RMC_HIP(ctx, hipStreamSynchronize(st));
      if (tail.bad) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite transition weight");
    }
    for (int j = 0; j < Pp; ++j) {
      const bool in = j < P;
      cols_h[j] = in ? (float)sp.capacity[j] : 1.f;
      cols_h[Pp + j] = in ? (float)(1.0 / sp.capacity[j]) : 1.f;
      cols_h[2 * Pp + j] = (in && sp.growth) ? (float)sp.growth[j] : 0.f;
      cols_h[3 * Pp + j] = in ? (float)sp.init_pop[j] : 0.f;
    }
    RMC_HIP(ctx, hipMemcpyAsync(W.cols.ptr, cols_h, 4 * (size_t)Pp * sizeof(float),
                                hipMemcpyHostToDevice, st));
    roadmc::MteWideArgs a{};
    a.mt = mt;
    a.cols = static_cast<const float*>(W.cols.ptr);
    a.grm = static_cast<float*>(W.grm.ptr);
    a.P = P;
    a.Pp = Pp;
    a.N = (int32_t)N;
    a.law = pr->growth_law;
    a.sig = (float)(sp.noise_sd * l2e);
    a.half_var = (float)(0.5 * sp.noise_sd * sp.noise_sd * l2e);
    a.dt = (float)pr->time_step;
    a.mu = (float)sp.params[1];
    a.sd = (float)sp.params[2];
    a.rev = (float)sp.params[4];
    a.jmp = (float)sp.params[5];
    a.jprob = (float)sp.params[6];
    a.lv = (float)sp.params[7];
    a.tag0 = jb.tag0;
    a.tag1 = jb.tag1;
    a.key0 = jb.key0;
    a.key1 = jb.key1;
    float* cur = static_cast<float*>(W.s0.ptr);
    float* nxt = static_cast<float*>(W.s1.ptr);
    RMC_HIP(ctx, roadmc::launch_mte_wide_init(a, (float)sp.params[0], cur, st));
    for (int y = 0; y < pr->n_years; ++y) {
      a.year = y;
      a.keep_pre = (y == pr->n_years - 1 && pr->endpoint == ROADMC_END_PRE_GROWTH) ? 1 : 0;
      a.s_in = cur;
      a.s_out = nxt;
      if (pr->growth_law == ROADMC_LAW_REF_GPU) RMC_HIP(ctx, roadmc::launch_mte_wide_grm(a, st));
      RMC_HIP(ctx, roadmc::launch_mte_wide_year(a, st));
      std::swap(cur, nxt);
    }
    RMC_HIP(ctx, roadmc::launch_mte_wide_tally(cur, (int)N, Pp, jb.thr_total, jb.partial,
                                               jb.end_pops, st));
  }
  RMC_HIP(ctx, roadmc::launch_mte_reduce(static_cast<MteJob*>(b.jobs.ptr), (int)js.jobs.size(),
                                         static_cast<double*>(b.out.ptr),
                                         static_cast<int64_t*>(b.off.ptr), st));
  return fetch_results(ctx, js, b, pr, result);
}

}  // namespace

extern "C" int roadmc_mte_run(roadmc_ctx* ctx, const roadmc_mte_problem* problem,
                              roadmc_mte_result* result) {
  if (!ctx) return ROADMC_ERR_INVALID;
  if (!result) return rmc_fail(ctx, ROADMC_ERR_INVALID, "result is NULL");
  int rc;
  if (!problem) {
    rc = validate_problem(ctx, problem, ROADMC_MAX_PATCHES_WIDE);
  } else {
    std::vector<size_t> mat_off;
    std::vector<char> mat_ok;
    matrices_finite(problem, 1, ROADMC_MAX_PATCHES_WIDE, mat_off, mat_ok);
    rc = validate_problem(ctx, problem, ROADMC_MAX_PATCHES_WIDE,
                          mat_off[1] > 0 ? mat_ok.data() : nullptr);
  }
  if (rc) return rc;
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  for (int s = 0; s < problem->n_species; ++s)
    if (problem->species[s].n_patches > ROADMC_MAX_PATCHES)
      return mte_run_wide(ctx, problem, result);
  JobSet js;
  plan_jobs(ctx, problem, 1, result->end_pops != nullptr, js);
  JobBuffers& b = ctx->mte_bufs;
  rc = js.fp64 ? upload_jobs<double>(ctx, problem, js, b, false)
               : upload_jobs<float>(ctx, problem, js, b, false);
  if (rc) return rc;
  if ((rc = ensure_lds_limit(ctx, js.prec, js.lds))) return rc;
  RMC_HIP(ctx, roadmc::launch_mte_direct(js.prec, js.tpw, static_cast<MteJob*>(b.jobs.ptr),
                                         static_cast<int2*>(b.work.ptr), (int)js.work.size(),
                                         js.lds, ctx->stream));
  RMC_HIP(ctx, roadmc::launch_mte_reduce(static_cast<MteJob*>(b.jobs.ptr), (int)js.jobs.size(),
                                         static_cast<double*>(b.out.ptr),
                                         static_cast<int64_t*>(b.off.ptr), ctx->stream));
  return fetch_results(ctx, js, b, problem, result);
}

// ---------------------------------------------------------------------------
// Batched generation.

struct roadmc_batch {
  roadmc_ctx* ctx = nullptr;
  JobSet js;
  JobBuffers bufs;
  std::vector<roadmc_mte_problem> roads;  // shallow copies, used for output shapes
  std::vector<std::vector<roadmc_species>> species;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int grid = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  size_t timed_from = 0;
};

namespace {

int capture_batch_graph(roadmc_batch* bt) {
  roadmc_ctx* ctx = bt->ctx;
  hipStream_t s = ctx->stream;
  JobSet& js = bt->js;
  JobBuffers& b = bt->bufs;
  RMC_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipError_t e = roadmc::launch_fill_i32(static_cast<int*>(b.counter.ptr), 0, 16, s);
  if (e == hipSuccess)
    e = roadmc::launch_mte_persistent(
        js.prec, js.tpw, static_cast<MteJob*>(b.jobs.ptr), static_cast<int2*>(b.work.ptr),
        static_cast<int*>(b.counter.ptr) + 16, (int)js.qbeg.size() - 1,
        static_cast<int*>(b.counter.ptr), bt->grid, js.lds, s);
  if (e == hipSuccess)
    e = roadmc::launch_mte_reduce(static_cast<MteJob*>(b.jobs.ptr), (int)js.jobs.size(),
                                  static_cast<double*>(b.out.ptr),
                                  static_cast<int64_t*>(b.off.ptr), s);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(s, &g);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return rmc_fail(ctx, ROADMC_ERR_HIP, std::string("batch capture: ") + hipGetErrorString(e));
  }
  RMC_HIP(ctx, e2);
  bt->graph = g;
  RMC_HIP(ctx, hipGraphInstantiate(&bt->exec, g, nullptr, nullptr, 0));
  return ROADMC_OK;
}

}  // namespace

extern "C" int roadmc_batch_prepare(roadmc_ctx* ctx, const roadmc_mte_problem* roads,
                                    int32_t n_roads, roadmc_batch** out) {
  if (!ctx) return ROADMC_ERR_INVALID;
  if (!out || !roads || n_roads < 1)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "roads/out invalid");
  *out = nullptr;
  // The host matrices' finiteness (most of the validation's reads) over
  // threads; then the checks road by road, in order, with those flags.
  std::vector<size_t> mat_off;
  std::vector<char> mat_ok;
  matrices_finite(roads, n_roads, ROADMC_MAX_PATCHES, mat_off, mat_ok);
  for (int r = 0; r < n_roads; ++r) {
    int rc = validate_problem(ctx, &roads[r], ROADMC_MAX_PATCHES,
                              mat_off[r + 1] > mat_off[r] ? &mat_ok[mat_off[r]] : nullptr);
    if (rc) return rc;
    if (roads[r].precision != roads[0].precision)
      return rmc_fail(ctx, ROADMC_ERR_INVALID, "all roads of a batch need one precision");
  }
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  roadmc_batch* bt = new roadmc_batch();
  bt->ctx = ctx;
  plan_jobs(ctx, roads, n_roads, true, bt->js);
  int rc = bt->js.fp64 ? upload_jobs<double>(ctx, roads, bt->js, bt->bufs, true)
                       : upload_jobs<float>(ctx, roads, bt->js, bt->bufs, true);
  if (!rc) rc = ensure_lds_limit(ctx, bt->js.prec, bt->js.lds);
  if (!rc) {
    // One workgroup per CU fits the LDS of the largest tile; small tiles
    // leave room for a second one, unless a bucket of the batch asks for one.
    int want = 2;
    if (!bt->js.fp64)
      for (const MteJob& jb : bt->js.jobs) {
        const int pc = ctx->tiles[tile_slot(bt->js.prec)][bucket_of(jb.Pp)].per_cu;
        want = std::min(want, pc == 0 ? 2 : pc);
      }
    const int per_cu = (want == 2 && bt->js.lds * 2 <= kLdsPerCu) ? 2 : 1;
    bt->grid = ctx->n_cu * per_cu;
    bt->roads.assign(roads, roads + n_roads);
    rc = capture_batch_graph(bt);
  }
  if (rc) {
    roadmc_batch_destroy(bt);
    return rc;
  }
  *out = bt;
  return ROADMC_OK;
}

extern "C" int roadmc_batch_launch(roadmc_batch* bt) {
  if (!bt) return ROADMC_ERR_INVALID;
  roadmc_ctx* ctx = bt->ctx;
  hipEvent_t a = nullptr, z = nullptr;
  RMC_HIP(ctx, hipEventCreate(&a));
  RMC_HIP(ctx, hipEventCreate(&z));
  RMC_HIP(ctx, hipEventRecord(a, ctx->stream));
  RMC_HIP(ctx, hipGraphLaunch(bt->exec, ctx->stream));
  RMC_HIP(ctx, hipEventRecord(z, ctx->stream));
  bt->events.emplace_back(a, z);
  return ROADMC_OK;
}

extern "C" int roadmc_batch_sync(roadmc_batch* bt) {
  if (!bt) return ROADMC_ERR_INVALID;
  RMC_HIP(bt->ctx, hipStreamSynchronize(bt->ctx->stream));
  return ROADMC_OK;
}

extern "C" int roadmc_batch_kernel_ms(roadmc_batch* bt, double* ms) {
  if (!bt || !ms) return ROADMC_ERR_INVALID;
  roadmc_ctx* ctx = bt->ctx;
  RMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double tot = 0.0;
  size_t n = 0;
  for (size_t i = bt->timed_from; i < bt->events.size(); ++i) {
    float t = 0.f;
    RMC_HIP(ctx, hipEventElapsedTime(&t, bt->events[i].first, bt->events[i].second));
    tot += t;
    ++n;
  }
  bt->timed_from = bt->events.size();
  if (!n) return rmc_fail(ctx, ROADMC_ERR_INVALID, "no launches since the last query");
  *ms = tot / (double)n;
  return ROADMC_OK;
}

extern "C" int roadmc_batch_fetch(roadmc_batch* bt, roadmc_mte_result* results) {
  if (!bt || !results) return ROADMC_ERR_INVALID;
  return fetch_results(bt->ctx, bt->js, bt->bufs, bt->roads.data(), results);
}

extern "C" void roadmc_batch_destroy(roadmc_batch* bt) {
  if (!bt) return;
  roadmc_ctx* ctx = bt->ctx;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& ev : bt->events) {
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
  if (bt->graph) (void)hipGraphDestroy(bt->graph);
  rmc_devbuf* all[] = {&bt->bufs.mpack, &bt->bufs.cols, &bt->bufs.partial, &bt->bufs.out,
                       &bt->bufs.endpop, &bt->bufs.jobs, &bt->bufs.work, &bt->bufs.off,
                       &bt->bufs.counter};
  for (rmc_devbuf* d : all) rmc_buf_free(*d);
  delete bt;
}

extern "C" int roadmc_mte_run_batch(roadmc_ctx* ctx, const roadmc_mte_problem* roads,
                                    int32_t n_roads, roadmc_mte_result* results) {
  if (!ctx) return ROADMC_ERR_INVALID;
  if (!results) return rmc_fail(ctx, ROADMC_ERR_INVALID, "results is NULL");
  roadmc_batch* bt = nullptr;
  int rc = roadmc_batch_prepare(ctx, roads, n_roads, &bt);
  if (rc) return rc;
  rc = roadmc_batch_launch(bt);
  if (!rc) rc = roadmc_batch_fetch(bt, results);
  roadmc_batch_destroy(bt);
  return rc;
}

// ---------------------------------------------------------------------------
// Tile autotuner: for each patch-count bucket, every valid setting (rows of
// paths per workgroup, k-split of remainder tiles, workgroups per CU) is
// timed on a batch of synthetic roads whose patch counts spread over the
// bucket (so that ragged counts leave remainder tiles, as a generation's
// roads do), with the persistent kernel's own events, and the fastest is
// kept.

namespace {

bool valid_precision(int prec) { return prec == ROADMC_PREC_FP32 || prec == ROADMC_PREC_BF16X3; }

int bucket_index(int bucket_max_p) {
  for (int b = 0; b < kBuckets; ++b)
    if (kBucketMax[b] == bucket_max_p) return b;
  return -1;
}

// A road of P patches with seeded random transitions (rows summing to one),
// random crossing survival and capacities: the values do not change the
// work, only keep the arithmetic ordinary.
struct TuneRoad {
  std::vector<double> trans, surv, init, cap, gr;
  roadmc_species sp{};
  explicit TuneRoad(int P) : trans((size_t)P * P), surv((size_t)P * P), init(P), cap(P), gr(P) {
    uint64_t x = 0x9e3779b97f4a7c15ull ^ (uint64_t)P;
    auto unif = [&x]() {
      x ^= x << 13;
      x ^= x >> 7;
      x ^= x << 17;
      return (double)(x >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int i = 0; i < P; ++i) {
      double row = 0.0;
      for (int j = 0; j < P; ++j) row += trans[(size_t)j * P + i] = 0.05 + unif();
      for (int j = 0; j < P; ++j) trans[(size_t)j * P + i] /= row;
      cap[i] = 500.0 + 4500.0 * unif();
      init[i] = 0.8 * cap[i];
      gr[i] = 0.02 + 0.08 * unif();
    }
    for (double& v : surv) v = 1.0 - 0.3 * unif();
    sp.n_patches = P;
    sp.init_pop = init.data();
    sp.capacity = cap.data();
    sp.growth = gr.data();
    sp.trans = trans.data();
    sp.surv = surv.data();
    sp.noise_sd = 0.1;
    sp.threshold_ratio = 0.5;
  }
};

}  // namespace

extern "C" int roadmc_set_tile(roadmc_ctx* ctx, int32_t precision, int32_t bucket_max_p,
                               int32_t rows, int32_t split, int32_t per_cu) {
  if (!ctx) return ROADMC_ERR_INVALID;
  const int b = bucket_index(bucket_max_p);
  if (!valid_precision(precision) || b < 0)
    return rmc_fail(ctx, ROADMC_ERR_INVALID,
                    "roadmc_set_tile: precision FP32 or BF16X3, bucket 128, 256, 512 or 1024");
  if (rows < roadmc::kMteRowsF32 || rows > 256 || (rows & (rows - 1)) ||
      rows * kBucketMax[b] > roadmc::kMteTileElemsF32)
    return rmc_fail(ctx, ROADMC_ERR_INVALID,
                    "roadmc_set_tile: rows a power of two in [32, 256] with rows x P <= 32768");
  if ((split != 0 && split != 1) || per_cu < 0 || per_cu > 2)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "roadmc_set_tile: split 0 or 1, per_cu 0, 1 or 2");
  ctx->tiles[tile_slot(precision)][b] = {rows, split, per_cu};
  return ROADMC_OK;
}

extern "C" int roadmc_get_tile(const roadmc_ctx* ctx, int32_t precision, int32_t bucket_max_p,
                               int32_t* rows, int32_t* split, int32_t* per_cu) {
  if (!ctx) return ROADMC_ERR_INVALID;
  const int b = bucket_index(bucket_max_p);
  if (!valid_precision(precision) || b < 0) return ROADMC_ERR_INVALID;
  const roadmc_ctx::TileSetting& t = ctx->tiles[tile_slot(precision)][b];
  if (rows) *rows = t.rows;
  if (split) *split = t.split;
  if (per_cu) *per_cu = t.per_cu;
  return ROADMC_OK;
}

extern "C" int roadmc_autotune_ex(roadmc_ctx* ctx, int32_t precision, int32_t n_paths,
                                  int32_t n_years, roadmc_tile_trial* trials, int32_t max_trials,
                                  int32_t* n_trials) {
  if (!ctx) return ROADMC_ERR_INVALID;
  if (!valid_precision(precision))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "autotune: precision FP32 or BF16X3");
  if (n_paths < 256 || n_years < 1)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "autotune needs n_paths >= 256 and n_years >= 1");
  if (max_trials < 0 || (max_trials > 0 && !trials))
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "autotune: trials is NULL");
  const int slot = tile_slot(precision);
  roadmc_ctx::TileSetting saved[kBuckets], best[kBuckets];
  for (int b = 0; b < kBuckets; ++b) saved[b] = best[b] = ctx->tiles[slot][b];
  auto restore = [&]() {
    for (int b = 0; b < kBuckets; ++b) ctx->tiles[slot][b] = saved[b];
  };
  int count = 0;
  constexpr int kRoads = 8;
  for (int b = 0; b < kBuckets; ++b) {
    const int P = kBucketMax[b], lo = b ? kBucketMax[b - 1] : 64;
    // Patch counts lo + (i + 1) (P - lo) / kRoads, rounded up to 8: ragged,
    // the last one the bucket's largest.
    std::vector<TuneRoad> roads;
    roads.reserve(kRoads);
    for (int i = 0; i < kRoads; ++i)
      roads.emplace_back(std::min(P, round_up(lo + (i + 1) * (P - lo) / kRoads, 8)));
    std::vector<roadmc_mte_problem> prs(kRoads);
    for (int i = 0; i < kRoads; ++i) {
      roadmc_mte_problem& pr = prs[(size_t)i];
      pr = roadmc_mte_problem{};
      pr.n_species = 1;
      pr.species = &roads[(size_t)i].sp;
      pr.n_years = n_years;
      pr.time_step = 1.0;
      pr.n_paths = n_paths;
      pr.precision = precision;
      pr.seed = 12345;
      pr.road_id = (uint32_t)i;
    }
    double best_ms = 1e300;
    int best_trial = -1;
    for (int rows = roadmc::kMteRowsF32; rows <= 256; rows *= 2) {
      if (rows * P > roadmc::kMteTileElemsF32) break;
      // Two workgroups per CU only where the largest road's tile (with a
      // split's partial products) leaves room for a second.
      const size_t lds_max = roadmc::mte_lds_bytes(precision, rows, P, 3);
      for (int split = 0; split <= 1; ++split) {
        const size_t lds = split ? lds_max : roadmc::mte_lds_bytes(precision, rows, P, 0);
        for (int per_cu = 1; per_cu <= (lds * 2 <= kLdsPerCu ? 2 : 1); ++per_cu) {
          ctx->tiles[slot][b] = {rows, split, per_cu};
          roadmc_batch* bt = nullptr;
          int rc = roadmc_batch_prepare(ctx, prs.data(), kRoads, &bt);
          double ms = 0.0;
          if (!rc) rc = roadmc_batch_launch(bt);  // warm-up
          if (!rc) rc = roadmc_batch_kernel_ms(bt, &ms);
          if (!rc) rc = roadmc_batch_launch(bt);
          if (!rc) rc = roadmc_batch_launch(bt);
          if (!rc) rc = roadmc_batch_kernel_ms(bt, &ms);
          if (bt) roadmc_batch_destroy(bt);
          if (rc) {
            restore();
            return rc;
          }
          if (count < max_trials) {
            roadmc_tile_trial& tr = trials[count];
            tr.precision = precision;
            tr.bucket_max_p = P;
            tr.rows = rows;
            tr.split = split;
            tr.per_cu = per_cu;
            tr.chosen = 0;
            tr.ms = ms;
          }
          if (ms < best_ms) {
            best_ms = ms;
            best[b] = {rows, split, per_cu};
            best_trial = count;
          }
          ++count;
        }
      }
    }
    if (best_trial >= 0 && best_trial < max_trials) trials[best_trial].chosen = 1;
    ctx->tiles[slot][b] = saved[b];
  }
  for (int b = 0; b < kBuckets; ++b) ctx->tiles[slot][b] = best[b];
  if (n_trials) *n_trials = count;
  return ROADMC_OK;
}

extern "C" int roadmc_autotune(roadmc_ctx* ctx, int32_t n_paths, int32_t n_years) {
  int rc = roadmc_autotune_ex(ctx, ROADMC_PREC_FP32, n_paths, n_years, nullptr, 0, nullptr);
  if (!rc) rc = roadmc_autotune_ex(ctx, ROADMC_PREC_BF16X3, n_paths, n_years, nullptr, 0, nullptr);
  return rc;
}

extern "C" int roadmc_get_tiles(const roadmc_ctx* ctx, int32_t* bucket_max_p, int32_t* rows,
                                int32_t* n_buckets) {
  if (!ctx || !n_buckets) return ROADMC_ERR_INVALID;
  *n_buckets = kBuckets;
  for (int b = 0; b < kBuckets; ++b) {
    if (bucket_max_p) bucket_max_p[b] = kBucketMax[b];
    if (rows) rows[b] = ctx->tiles[0][b].rows;
  }
  return ROADMC_OK;
}

// ---------------------------------------------------------------------------
// Reference-design baseline (SimulateGPU.cu:3680-3843): per species, allocate,
// pre-generate the streams into HBM, build the CSC matrix on the host, run one
// thread per path, copy the per-path totals back, free.

namespace {

template <typename T>
int plain_species(roadmc_ctx* ctx, const roadmc_mte_problem& pr, int s,
                  std::vector<double>& eps_out, std::vector<double>* patch_sum) {
  const roadmc_species& sp = pr.species[s];
  const int P = sp.n_patches, N = (int)pr.n_paths, Ty = pr.n_years;
  const bool fp64 = sizeof(T) == 8;
  const uint32_t k0 = (uint32_t)(pr.seed & 0xffffffffu), k1 = (uint32_t)(pr.seed >> 32);
  // CSC of the binary32 movement matrix.  The reference keeps entries
  // > 1e-5; here entries > drop_tol (0 keeps every non-zero).
  std::vector<float> dense((size_t)P * P);
  for (int k = 0; k < P; ++k)
    for (int j = 0; j < P; ++j) dense[(size_t)j * P + k] = (float)movement(sp, P, k, j, 0.0);
  std::vector<float> valf((size_t)P * P);
  std::vector<int32_t> row_idx((size_t)P * P), per_col(P);
  int32_t nnz = 0;
  roadmc_dense_to_csc(dense.data(), P, P, (float)pr.drop_tol, valf.data(), row_idx.data(),
                      per_col.data(), &nnz);
  std::vector<T> val(nnz > 0 ? nnz : 1), prm(8), init(P), caps(P), rates(P);
  if (fp64) {
    // Double-precision values for the FP64 variant, same sparsity.
    int it = 0;
    for (int c = 0; c < P; ++c)
      for (int l = 0; l < per_col[c]; ++l, ++it)
        val[it] = (T)movement(sp, P, row_idx[it], c, 0.0);
  } else {
    for (int i = 0; i < nnz; ++i) val[i] = (T)valf[i];
  }
  for (int i = 0; i < 8; ++i) prm[i] = (T)sp.params[i];
  for (int j = 0; j < P; ++j) {
    init[j] = (T)sp.init_pop[j];
    caps[j] = (T)sp.capacity[j];
    rates[j] = sp.growth ? (T)sp.growth[j] : (T)0;
  }
  void *d_rgr = nullptr, *d_br = nullptr, *d_js = nullptr, *d_ju = nullptr, *d_prm = nullptr,
       *d_init = nullptr, *d_caps = nullptr, *d_rates = nullptr, *d_val = nullptr,
       *d_pp = nullptr, *d_eps = nullptr;
  int *d_row = nullptr, *d_col = nullptr;
  const size_t es = sizeof(T);
  int rc = ROADMC_OK;
  hipStream_t st = ctx->stream;
  auto cleanup = [&]() {
    void* all[] = {d_rgr, d_br, d_js, d_ju, d_prm, d_init, d_caps, d_rates, d_val, d_pp,
                   d_eps, d_row, d_col};
    for (void* p : all)
      if (p) (void)hipFree(p);
  };
  if ((rc = rmc_malloc(ctx, &d_rgr, (size_t)N * Ty * P * es)) ||
      (rc = rmc_malloc(ctx, &d_br, (size_t)N * Ty * es)) ||
      (rc = rmc_malloc(ctx, &d_js, (size_t)N * Ty * es)) ||
      (rc = rmc_malloc(ctx, &d_ju, (size_t)N * Ty * es)) ||
      (rc = rmc_malloc(ctx, &d_prm, 8 * es)) || (rc = rmc_malloc(ctx, &d_init, P * es)) ||
      (rc = rmc_malloc(ctx, &d_caps, P * es)) || (rc = rmc_malloc(ctx, &d_rates, P * es)) ||
      (rc = rmc_malloc(ctx, &d_val, val.size() * es)) ||
      (rc = rmc_malloc(ctx, (void**)&d_row, row_idx.size() * sizeof(int))) ||
      (rc = rmc_malloc(ctx, (void**)&d_col, per_col.size() * sizeof(int))) ||
      (rc = rmc_malloc(ctx, &d_pp, (size_t)N * 2 * P * es)) ||
      (rc = rmc_malloc(ctx, &d_eps, (size_t)N * es))) {
    cleanup();
    return rc;
  }
  hipError_t e = roadmc::launch_gen_patch_normals(
      fp64, d_rgr, N, Ty, P, rmc_tag(pr.road_id, (uint32_t)s, RMC_KIND_PATCH), k0, k1, st);
  if (e == hipSuccess)
    e = roadmc::launch_gen_path_scalars(fp64, d_br, d_js, d_ju, N, Ty,
                                        rmc_tag(pr.road_id, (uint32_t)s, RMC_KIND_PATH), k0,
                                        k1, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_prm, prm.data(), 8 * es, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_init, init.data(), P * es, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_caps, caps.data(), P * es, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_rates, rates.data(), P * es, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_val, val.data(), val.size() * es, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_row, row_idx.data(), row_idx.size() * sizeof(int),
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_col, per_col.data(), per_col.size() * sizeof(int),
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    roadmc::PlainArgs a{};
    a.N = N;
    a.T = Ty;
    a.P = P;
    a.law = pr.growth_law;
    a.endpoint = pr.endpoint;
    a.dt = pr.time_step;
    a.sigma = sp.noise_sd;
    a.rgr = d_rgr;
    a.brown = d_br;
    a.jsize = d_js;
    a.jump = d_ju;
    a.params = d_prm;
    a.init = d_init;
    a.caps = d_caps;
    a.rates = d_rates;
    a.val = d_val;
    a.row_idx = d_row;
    a.per_col = d_col;
    a.path_pops = d_pp;
    a.eps = d_eps;
    e = roadmc::launch_mte_plain(fp64, a, st);
  }
  std::vector<T> eps(N);
  if (e == hipSuccess) e = hipMemcpyAsync(eps.data(), d_eps, N * es, hipMemcpyDeviceToHost, st);
  std::vector<T> pops;
  if (e == hipSuccess && patch_sum) {
    pops.resize((size_t)N * 2 * P);
    e = hipMemcpyAsync(pops.data(), d_pp, pops.size() * es, hipMemcpyDeviceToHost, st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess)
    return rmc_fail(ctx, ROADMC_ERR_HIP, std::string("plain path: ") + hipGetErrorString(e));
  eps_out.resize(N);
  for (int i = 0; i < N; ++i) eps_out[i] = (double)eps[i];
  if (patch_sum) {
    patch_sum->assign(P, 0.0);
    const int slot = pr.endpoint == ROADMC_END_PRE_GROWTH ? 1 : 0;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < P; ++j) (*patch_sum)[j] += (double)pops[((size_t)i * 2 + slot) * P + j];
  }
  return ROADMC_OK;
}

}  // namespace

extern "C" int roadmc_mte_run_plain(roadmc_ctx* ctx, const roadmc_mte_problem* problem,
                                    roadmc_mte_result* result) {
  if (!ctx) return ROADMC_ERR_INVALID;
  if (!result) return rmc_fail(ctx, ROADMC_ERR_INVALID, "result is NULL");
  int rc = validate_problem(ctx, problem);
  if (rc) return rc;
  for (int sp = 0; sp < problem->n_species; ++sp)
    if (!problem->species[sp].trans)
      return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED, "the plain path needs host matrices");
  if (problem->precision == ROADMC_PREC_BF16X3)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED, "the plain path runs FP32 or FP64");
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  const int S = problem->n_species;
  const int64_t N = problem->n_paths;
  std::vector<std::vector<double>> eps(S), psum(S);
  for (int s = 0; s < S; ++s) {
    rc = problem->precision == ROADMC_PREC_FP64
             ? plain_species<double>(ctx, *problem, s, eps[s],
                                     result->patch_mean ? &psum[s] : nullptr)
             : plain_species<float>(ctx, *problem, s, eps[s],
                                    result->patch_mean ? &psum[s] : nullptr);
    if (rc) return rc;
  }
  // Statistics as Simulator::simulateMTE forms them (Simulator.cpp:184-191).
  int64_t pbase = 0;
  for (int s = 0; s < S; ++s) {
    const roadmc_species& sp = problem->species[s];
    double tot0 = 0.0;
    for (int j = 0; j < sp.n_patches; ++j) tot0 += sp.init_pop[j];
    const double thr = sp.threshold_ratio * tot0;
    double sum = 0.0, cnt = 0.0;
    for (double v : eps[s]) {
      sum += v;
      cnt += v >= thr ? 1.0 : 0.0;
    }
    const double mean = sum / (double)N;
    double ss = 0.0;
    for (double v : eps[s]) ss += (v - mean) * (v - mean);
    if (result->mte_prob) result->mte_prob[s] = cnt / (double)N;
    if (result->end_mean) result->end_mean[s] = mean;
    if (result->end_sd) result->end_sd[s] = std::sqrt(ss / (double)N);
    if (result->end_pops) std::memcpy(result->end_pops + s * N, eps[s].data(), sizeof(double) * N);
    if (result->patch_mean)
      for (int j = 0; j < sp.n_patches; ++j) result->patch_mean[pbase + j] = psum[s][j] / (double)N;
    pbase += sp.n_patches;
  }
  return ROADMC_OK;
}

extern "C" int roadmc_patch_normals(roadmc_ctx* ctx, uint64_t seed, uint32_t road_id,
                                    uint32_t species, int32_t n_paths, int32_t n_years,
                                    int32_t n_patches, float* out) {
  if (!ctx) return ROADMC_ERR_INVALID;
  if (!out || n_paths < 1 || n_years < 1 || n_patches < 1 || road_id >= (1u << 24) ||
      species > 63)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "roadmc_patch_normals: bad arguments");
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  const size_t n = (size_t)n_paths * n_years * n_patches;
  void* d = nullptr;
  int rc = rmc_malloc(ctx, &d, n * sizeof(float));
  if (rc) return rc;
  hipError_t e = roadmc::launch_gen_patch_normals(
      false, d, n_paths, n_years, n_patches, rmc_tag(road_id, species, RMC_KIND_PATCH),
      (uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32), ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess)
    return rmc_fail(ctx, ROADMC_ERR_HIP, std::string("patch normals: ") + hipGetErrorString(e));
  return ROADMC_OK;
}

// ---------------------------------------------------------------------------
// Adjoint: forward with checkpoints, backward by segments.

#include "adj_kernels.h"

struct roadmc_adjoint {
  roadmc_ctx* ctx = nullptr;
  bool fp64 = false;
  MteJob job{};
  int P = 0, L = 10;
  std::vector<double> trans;  // column-major, for dJ/dm = -T o dJ/dM
  rmc_job_buffers bufs;       // mpack, cols, partial, out, endpop(unused), jobs, work, off
  rmc_devbuf mtpack, ckpt;    // the backward scratch lives in ctx->adj_bufs
};

namespace {

template <typename T>
int adjoint_upload(roadmc_adjoint* adj, const roadmc_species& sp, double tol) {
  roadmc_ctx* ctx = adj->ctx;
  const int Pp = adj->job.Pp, P = adj->P;
  const bool fp64 = adj->fp64;
  std::vector<T> mp((size_t)Pp * Pp), mt((size_t)Pp * Pp, (T)0), cols(4 * (size_t)Pp);
  pack_species<T>(sp, Pp, Pp, tol, fp64, mp.data(), cols.data());
  for (int k = 0; k < P; ++k)
    for (int j = 0; j < P; ++j)
      mt[roadmc::mte_pack_index(fp64, Pp, j, k)] = (T)movement(sp, P, k, j, tol);
  int rc;
  if ((rc = rmc_buf_alloc_exact(ctx, adj->bufs.mpack, mp.size() * sizeof(T))) ||
      (rc = rmc_buf_alloc_exact(ctx, adj->mtpack, mt.size() * sizeof(T))) ||
      (rc = rmc_buf_alloc_exact(ctx, adj->bufs.cols, cols.size() * sizeof(T))))
    return rc;
  RMC_HIP(ctx, hipMemcpy(adj->bufs.mpack.ptr, mp.data(), mp.size() * sizeof(T),
                         hipMemcpyHostToDevice));
  RMC_HIP(ctx, hipMemcpy(adj->mtpack.ptr, mt.data(), mt.size() * sizeof(T),
                         hipMemcpyHostToDevice));
  RMC_HIP(ctx, hipMemcpy(adj->bufs.cols.ptr, cols.data(), cols.size() * sizeof(T),
                         hipMemcpyHostToDevice));
  return ROADMC_OK;
}

}  // namespace

extern "C" void roadmc_adjoint_destroy(roadmc_adjoint* adj) {
  if (!adj) return;
  (void)hipStreamSynchronize(adj->ctx->stream);
  // The checkpoints go to the context for the next forward pass when they
  // are larger than the spare it holds.
  rmc_devbuf& spare = adj->ctx->adj_bufs.ckpt_spare;
  if (adj->ckpt.ptr && adj->ckpt.bytes > spare.bytes) {
    rmc_buf_free(spare);
    spare = adj->ckpt;
    adj->ckpt = rmc_devbuf{};
  }
  rmc_devbuf* all[] = {&adj->bufs.mpack, &adj->bufs.cols, &adj->bufs.partial, &adj->bufs.out,
                       &adj->bufs.endpop, &adj->bufs.jobs, &adj->bufs.work, &adj->bufs.off,
                       &adj->bufs.counter, &adj->mtpack, &adj->ckpt};
  for (rmc_devbuf* d : all) rmc_buf_free(*d);
  delete adj;
}

extern "C" int roadmc_mte_adjoint_fwd(roadmc_ctx* ctx, const roadmc_mte_problem* pr,
                                      int32_t checkpoint_every, roadmc_adjoint** out,
                                      double* objective) {
  if (!ctx) return ROADMC_ERR_INVALID;
  if (!out || !objective) return rmc_fail(ctx, ROADMC_ERR_INVALID, "out/objective is NULL");
  *out = nullptr;
  int rc = validate_problem(ctx, pr);
  if (rc) return rc;
  if (!pr->species[0].trans)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED, "the adjoint needs host matrices");
  if (pr->precision == ROADMC_PREC_BF16X3)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED, "the adjoint runs FP32 or FP64");
  if (pr->n_species != 1 || pr->growth_law != ROADMC_LAW_LOGNORMAL ||
      pr->endpoint != ROADMC_END_POST_GROWTH || pr->drop_tol != 0.0)
    return rmc_fail(ctx, ROADMC_ERR_UNSUPPORTED,
                    "adjoint: one species, LOGNORMAL law, post-growth endpoint, drop_tol 0");
  if (checkpoint_every < 1) return rmc_fail(ctx, ROADMC_ERR_INVALID, "checkpoint_every < 1");
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  const roadmc_species& sp = pr->species[0];
  roadmc_adjoint* adj = new roadmc_adjoint();
  adj->ctx = ctx;
  adj->fp64 = pr->precision == ROADMC_PREC_FP64;
  adj->P = sp.n_patches;
  adj->L = checkpoint_every;
  adj->trans.assign(sp.trans, sp.trans + (size_t)adj->P * adj->P);
  MteJob& jb = adj->job;
  jb.P = sp.n_patches;
  jb.Pp = round_up(jb.P, 64);  // the gradient GEMM works in 64 x 64 blocks
  jb.Pk = jb.Pp;               // the adjoint kernels contract over Pp
  jb.split = 0;
  jb.N = (int32_t)pr->n_paths;
  jb.T = pr->n_years;
  // FP32: 32-row tiles.  The segment kernels hold one tile per workgroup, and
  // finer tiles even out the last round of workgroups: 38.9 ms against
  // 41.7 ms with 128-row tiles at 256 patches, 100 years, 50,000 paths.
  jb.BM = adj->fp64 ? rows_for(ctx, ROADMC_PREC_FP64, jb.Pp, pr->n_paths) : roadmc::kMteRowsF32;
  jb.law = ROADMC_LAW_LOGNORMAL;
  jb.endpoint = ROADMC_END_POST_GROWTH;
  jb.n_tiles = (int)((pr->n_paths + jb.BM - 1) / jb.BM);
  for (int i = 0; i < 8; ++i) jb.params[i] = sp.params[i];
  jb.sigma = sp.noise_sd;
  jb.dt = pr->time_step;
  double tot0 = 0.0;
  for (int j = 0; j < jb.P; ++j) tot0 += sp.init_pop[j];
  jb.thr_total = sp.threshold_ratio * tot0;
  jb.tag0 = rmc_tag(pr->road_id, 0u, RMC_KIND_PATCH);
  jb.tag1 = rmc_tag(pr->road_id, 0u, RMC_KIND_PATH);
  jb.key0 = (uint32_t)(pr->seed & 0xffffffffu);
  jb.key1 = (uint32_t)(pr->seed >> 32);
  jb.ckpt_every = checkpoint_every;
  const size_t es = adj->fp64 ? 8 : 4;
  const int nck = (jb.T + checkpoint_every - 1) / checkpoint_every;
  rc = adj->fp64 ? adjoint_upload<double>(adj, sp, 0.0) : adjoint_upload<float>(adj, sp, 0.0);
  const size_t ck_bytes = (size_t)nck * jb.N * jb.Pp * es;
  rmc_devbuf& spare = ctx->adj_bufs.ckpt_spare;
  if (!rc && spare.ptr && spare.bytes >= ck_bytes) {
    adj->ckpt = spare;  // a destroyed handle's checkpoints (no queued work reads them)
    spare = rmc_devbuf{};
  } else if (!rc) {
    rc = rmc_buf_alloc_exact(ctx, adj->ckpt, ck_bytes);
  }
  if (!rc) rc = rmc_buf_alloc_exact(ctx, adj->bufs.partial,
                                    (size_t)jb.n_tiles * (3 + jb.Pp) * sizeof(double));
  if (!rc) rc = rmc_buf_alloc_exact(ctx, adj->bufs.out, (3 + (size_t)jb.Pp) * sizeof(double));
  if (!rc) rc = rmc_buf_alloc_exact(ctx, adj->bufs.jobs, sizeof(MteJob));
  if (!rc) rc = rmc_buf_alloc_exact(ctx, adj->bufs.work, (size_t)jb.n_tiles * sizeof(int2));
  if (!rc) rc = rmc_buf_alloc_exact(ctx, adj->bufs.off, sizeof(int64_t));
  if (rc) {
    roadmc_adjoint_destroy(adj);
    return rc;
  }
  jb.mpack = adj->bufs.mpack.ptr;
  jb.cols = adj->bufs.cols.ptr;
  jb.partial = static_cast<double*>(adj->bufs.partial.ptr);
  jb.end_pops = nullptr;
  jb.ckpt = adj->ckpt.ptr;
  std::vector<int2> work(jb.n_tiles);
  for (int t = 0; t < jb.n_tiles; ++t) work[t] = make_int2(0, t);
  const int64_t off0 = 0;
  const size_t lds = roadmc::mte_lds_bytes(adj->fp64 ? roadmc::kPrecF64 : roadmc::kPrecF32, jb.BM, jb.Pp, 0);
  hipStream_t s = ctx->stream;
  auto fail = [&](int code) {
    roadmc_adjoint_destroy(adj);
    return code;
  };
  if ((rc = ensure_lds_limit(ctx, adj->fp64 ? roadmc::kPrecF64 : roadmc::kPrecF32, lds))) return fail(rc);
  hipError_t e = hipMemcpyAsync(adj->bufs.jobs.ptr, &jb, sizeof(MteJob), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(adj->bufs.work.ptr, work.data(), work.size() * sizeof(int2),
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(adj->bufs.off.ptr, &off0, sizeof(off0), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = roadmc::launch_mte_direct(adj->fp64 ? roadmc::kPrecF64 : roadmc::kPrecF32,
                                  roadmc::mte_tiles_per_wave(adj->fp64 ? roadmc::kPrecF64
                                                                       : roadmc::kPrecF32,
                                                             jb.BM, jb.Pp),
                                  static_cast<MteJob*>(adj->bufs.jobs.ptr),
                                  static_cast<int2*>(adj->bufs.work.ptr), jb.n_tiles, lds, s);
  if (e == hipSuccess)
    e = roadmc::launch_mte_reduce(static_cast<MteJob*>(adj->bufs.jobs.ptr), 1,
                                  static_cast<double*>(adj->bufs.out.ptr),
                                  static_cast<int64_t*>(adj->bufs.off.ptr), s);
  double sums[3] = {0, 0, 0};
  if (e == hipSuccess)
    e = hipMemcpyAsync(sums, adj->bufs.out.ptr, sizeof(sums), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    rmc_fail(ctx, ROADMC_ERR_HIP, std::string("adjoint forward: ") + hipGetErrorString(e));
    return fail(ROADMC_ERR_HIP);
  }
  if (!std::isfinite(sums[1])) {
    rmc_fail(ctx, ROADMC_ERR_NUMERIC, "adjoint forward: non-finite population");
    return fail(ROADMC_ERR_NUMERIC);
  }
  *objective = sums[1] / (double)jb.N;
  *out = adj;
  return ROADMC_OK;
}

extern "C" int roadmc_mte_adjoint_bwd(roadmc_adjoint* adj, double d_objective,
                                      double* grad_mortality, double* grad_growth) {
  if (!adj) return ROADMC_ERR_INVALID;
  roadmc_ctx* ctx = adj->ctx;
  if (!grad_mortality && !grad_growth)
    return rmc_fail(ctx, ROADMC_ERR_INVALID, "no gradient output requested");
  RMC_HIP(ctx, hipSetDevice(ctx->device));
  const MteJob& j0 = adj->job;
  const int Pp = j0.Pp, N = j0.N, L = adj->L, P = adj->P;
  const size_t es = adj->fp64 ? 8 : 4;
  const size_t plane = (size_t)N * Pp;
  const int nsplit = 128;
  // Scratch shared by every adjoint handle of the context: the calls on one
  // context are serial and this one ends with a stream synchronization.
  auto& sc = ctx->adj_bufs;
  int rc;
  if ((rc = rmc_buf_ensure(ctx, sc.ns, (size_t)L * plane * es)) ||
      (rc = rmc_buf_ensure(ctx, sc.np, (size_t)L * plane * es)) ||
      (rc = rmc_buf_ensure(ctx, sc.mu, (size_t)L * plane * es)) ||
      (rc = rmc_buf_ensure(ctx, sc.lam, plane * es)) ||
      (rc = rmc_buf_ensure(ctx, sc.gradr, (size_t)Pp * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, sc.gradm, (size_t)Pp * Pp * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, sc.gpart, (size_t)nsplit * Pp * Pp * sizeof(double))) ||
      (rc = rmc_buf_ensure(ctx, sc.rpart, (size_t)j0.n_tiles * Pp * sizeof(double))))
    return rc;
  hipStream_t s = ctx->stream;
  RMC_HIP(ctx, hipMemsetAsync(sc.gradr.ptr, 0, (size_t)Pp * sizeof(double), s));
  RMC_HIP(ctx, hipMemsetAsync(sc.gradm.ptr, 0, (size_t)Pp * Pp * sizeof(double), s));
  roadmc::AdjJob a{};
  a.mpack = adj->bufs.mpack.ptr;
  a.mtpack = adj->mtpack.ptr;
  a.cols = adj->bufs.cols.ptr;
  a.ckpt = adj->ckpt.ptr;
  a.ns = sc.ns.ptr;
  a.np = sc.np.ptr;
  a.mu = sc.mu.ptr;
  a.lam = sc.lam.ptr;
  a.grad_r = static_cast<double*>(sc.gradr.ptr);
  a.grad_part = static_cast<double*>(sc.rpart.ptr);
  a.sigma = j0.sigma;
  a.dobj = d_objective;
  a.P = P;
  a.Pp = Pp;
  a.N = N;
  a.T = j0.T;
  a.BM = j0.BM;
  a.L = L;
  a.tag0 = j0.tag0;
  a.key0 = j0.key0;
  a.key1 = j0.key1;
  const size_t lds = roadmc::adj_lds_bytes(adj->fp64, j0.BM, Pp);
  RMC_HIP(ctx, roadmc::adj_set_lds_limit(adj->fp64, j0.BM, Pp, lds));
  const int nseg = (j0.T + L - 1) / L;
  for (int sg = nseg - 1; sg >= 0; --sg) {
    a.y0 = sg * L;
    a.y1 = std::min(j0.T, a.y0 + L) - 1;
    RMC_HIP(ctx, roadmc::launch_adj_segment(adj->fp64, a, j0.n_tiles, lds, s));
    const int K = (a.y1 - a.y0 + 1) * N;
    RMC_HIP(ctx, roadmc::launch_adj_gradm(adj->fp64, sc.ns.ptr, sc.mu.ptr, K, Pp, nsplit,
                                          static_cast<double*>(sc.gpart.ptr),
                                          static_cast<double*>(sc.gradm.ptr), s));
  }
  std::vector<double> gm((size_t)Pp * Pp), gr(Pp);
  RMC_HIP(ctx, hipMemcpyAsync(gm.data(), sc.gradm.ptr, gm.size() * sizeof(double),
                              hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipMemcpyAsync(gr.data(), sc.gradr.ptr, gr.size() * sizeof(double),
                              hipMemcpyDeviceToHost, s));
  RMC_HIP(ctx, hipStreamSynchronize(s));
  for (double v : gr)
    if (!std::isfinite(v)) return rmc_fail(ctx, ROADMC_ERR_NUMERIC, "non-finite gradient");
  // M = T o (1 - m): dJ/dm_kj = -T_kj dJ/dM_kj; outputs column-major like trans.
  if (grad_mortality)
    for (int k = 0; k < P; ++k)
      for (int j = 0; j < P; ++j)
        grad_mortality[(size_t)j * P + k] = -adj->trans[(size_t)j * P + k] * gm[(size_t)k * Pp + j];
  if (grad_growth)
    for (int j = 0; j < P; ++j) grad_growth[j] = gr[j];
  return ROADMC_OK;
}
