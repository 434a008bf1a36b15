// Host-visible declarations of the adjoint kernels (adj_kernels.hip).
#ifndef ROADMC_ADJ_KERNELS_H
#define ROADMC_ADJ_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace roadmc {

struct AdjJob {
  const void* mpack;   // M in MFMA B-operand order
  const void* mtpack;  // M^T in MFMA B-operand order
  const void* cols;    // 4*Pp: capacity, 1/capacity, growth rate, initial pop
  const void* ckpt;    // [T/L][N][Pp] states entering years 0, L, 2L, ...
  void* ns;            // [L][N][Pp] states entering each year of the segment
  void* np;            // [L][N][Pp] pre-growth states leaving each year
  void* mu;            // [L][N][Pp] adjoints mu of each year
  void* lam;           // [N][Pp] lambda at the segment's upper end (in/out)
  double* grad_r;      // [P] accumulated dJ/dr
  double* grad_part;   // [n_tiles][Pp] per-tile dJ/dr of the current segment
  double sigma, dobj;
  int32_t P, Pp, N, T, BM, L, y0, y1;
  uint32_t tag0, key0, key1;
};

size_t adj_lds_bytes(bool fp64, int BM, int Pp);
// Dynamic LDS limit of the segment kernel form that launch_adj_segment runs
// for tiles of BM paths over Pp patches.
hipError_t adj_set_lds_limit(bool fp64, int BM, int Pp, size_t bytes);
hipError_t launch_adj_segment(bool fp64, const AdjJob& jb, int n_tiles, size_t lds,
                              hipStream_t s);
// grad += sum_k A[r_k]^T B[r_k] for K rows of Pp (Pp a multiple of 64), r_k = k,
// or rows[k] when rows is given (a list of K row indices).
hipError_t launch_adj_gradm(bool fp64, const void* A, const void* B, int Ktot, int Pp,
                            int nsplit, double* partial, double* grad, hipStream_t s,
                            const int* rows = nullptr);

}  // namespace roadmc

#endif
