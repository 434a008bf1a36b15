// Host-visible declarations of the MTE kernels (mte_kernels.hip).
#ifndef ROADMC_MTE_KERNELS_H
#define ROADMC_MTE_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace roadmc {

// Arithmetic of a launch (the values of ROADMC_PREC_*): FP32, FP64, or FP32
// accumulation of split-bf16 products (state and movement matrix held as
// bf16 hi + lo pairs).
enum { kPrecF32 = 0, kPrecF64 = 1, kPrecB3 = 2 };

// One (road, species) pair of a launch.  Device arrays are of the launch's
// element type (float for FP32, double for FP64).
struct MteJob {
  const void* mpack;   // movement matrix in MFMA B-operand order, Pp*Pp
  const void* cols;    // 4*Pp: capacity, 1/capacity, growth rate, initial pop
  double* end_pops;    // [N] per-path end totals, or nullptr
  void* ckpt;          // adjoint checkpoints [T/ckpt_every][N][Pp], or nullptr
  int32_t ckpt_every;
  double* partial;     // [n_tiles][3 + Pp]: count, sum, sum of squares,
                       // per-patch sums of one tile
  double params[8];    // packed species parameters (SimulateGPU.cu:3690-3714)
  double sigma;        // LOGNORMAL year-noise SD
  double dt;           // time step
  double thr_total;    // threshold_ratio * initial total
  int32_t P, Pp, N, T, BM, law, endpoint, n_tiles;
  int32_t Pk;          // contraction extent: P rounded up to 16 (<= Pp); Pp
                       // in the split-bf16 form (an even count of 16-k chunks)
  int32_t split;       // split the ntiles % 4 remainder tiles into k-quarters
  uint32_t tag0, tag1, key0, key1;
};

constexpr int kMteThreads = 512;  // 8 waves per workgroup

// Element budget BM*Pp of one workgroup's population tile.
constexpr int kMteTileElemsF32 = 32768;
constexpr int kMteTileElemsF64 = 16384;
// Row granularity (MFMA tile height) of the FP32 / FP64 kernels.
constexpr int kMteRowsF32 = 32;
constexpr int kMteRowsF64 = 16;

// LDS bytes of one workgroup; n_rem_split is the number of remainder tiles
// split into k-quarters (0 when the job does not split).
size_t mte_lds_bytes(int prec, int BM, int Pp, int n_rem_split);
// Packed index of movement entry (k = origin, j = destination) in a matrix
// packed with Kp / 8 k-groups per column block (Kp: the contraction extent):
// element (k, j) lands where lane l of the wave handling column tile j / TN
// reads it as B operand s of k-group k / 8.
__host__ __device__ inline size_t mte_pack_index(bool fp64, int Kp, int k, int j) {
  const int kg = k >> 3, kk = k & 7;
  const int nkg = Kp >> 3;
  if (!fp64) {
    const int s = kk >> 1, h = kk & 1;
    const int lane = h * 32 + (j & 31), cb = j >> 5;
    return (((size_t)cb * nkg + kg) * 64 + lane) * 4 + s;
  }
  const int s = kk >> 2, h = kk & 3;
  const int lane = h * 16 + (j & 15), cb = j >> 4;
  return (((size_t)cb * nkg + kg) * 64 + lane) * 2 + s;
}

// Output tiles per wave of a job with BM rows of paths over Pp patches; the
// launches take the maximum over their jobs and run the smallest kernel form
// that holds it (more resident workgroups where registers allow).
int mte_tiles_per_wave(int prec, int BM, int Pp);
// Dynamic LDS limit of every kernel form of a precision.
hipError_t mte_set_lds_limit(int prec, size_t bytes);
hipError_t launch_mte_direct(int prec, int tiles_per_wave, const MteJob* jobs,
                             const int2* work, int n_work, size_t lds, hipStream_t s);
// Converts FP32-packed matrices (n_floats, a multiple of 512, each matrix
// with an even k-group count) in place to the split-bf16 operand order.
hipError_t launch_mte_split_b3(float* m, size_t n_floats, hipStream_t s);
// Work queues: items work[qbeg[q] .. qbeg[q+1]) form queue q; counters[q]
// must be zero at launch.
constexpr int kMteQueues = 8;  // one per XCD of an MI355X
hipError_t launch_mte_persistent(int prec, int tiles_per_wave, const MteJob* jobs,
                                 const int2* work, const int* qbeg, int nq,
                                 int* counters, int grid, size_t lds,
                                 hipStream_t s);
hipError_t launch_mte_reduce(const MteJob* jobs, int n_jobs, double* out,
                             const int64_t* out_off, hipStream_t s);
hipError_t launch_fill_i32(int* p, int value, int n, hipStream_t s);

// Reference-design baseline (one thread per path, pre-generated streams).
hipError_t launch_gen_patch_normals(bool fp64, void* out, int N, int T, int P,
                                    uint32_t tag, uint32_t k0, uint32_t k1,
                                    hipStream_t s);
hipError_t launch_gen_path_scalars(bool fp64, void* brown, void* jsize,
                                   void* jump, int N, int T, uint32_t tag,
                                   uint32_t k0, uint32_t k1, hipStream_t s);
struct PlainArgs {
  int N, T, P, law, endpoint;
  double dt, sigma;
  const void* rgr;
  const void* brown;
  const void* jsize;
  const void* jump;
  const void* params;  // 8
  const void* init;
  const void* caps;
  const void* rates;
  const void* val;
  const int* row_idx;
  const int* per_col;
  void* path_pops;  // [N][2][P]
  void* eps;        // [N]
};
hipError_t launch_mte_plain(bool fp64, const PlainArgs& a, hipStream_t s);

}  // namespace roadmc

#endif
