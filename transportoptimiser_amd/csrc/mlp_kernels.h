// Host-visible declarations of the surrogate MLP kernels (mlp_kernels.hip).
#ifndef ROADMC_MLP_KERNELS_H
#define ROADMC_MLP_KERNELS_H

#include <hip/hip_runtime.h>

namespace roadmc {

constexpr int kMlpMaxIn = 16;
constexpr int kMlpMaxOut = 4;
constexpr int kMlpMaxMembers = 8;

struct MlpArgs {
  int n_in, n_out, E;
  float* params;  // [E][n_params]
  float* m;       // Adam first moments
  float* v;       // Adam second moments
};

int mlp_param_count(int n_in, int n_out);
size_t mlp_grad_lds(int n_in, int n_out);
hipError_t mlp_set_lds_limit(int n_in, int n_out);
hipError_t launch_mlp_grad(const MlpArgs& a, const float* X, const float* Y, int n, float* partial,
                           float* loss_part, hipStream_t s);
hipError_t launch_mlp_update(const MlpArgs& a, const float* partial, int nchunk, float* grad_out,
                             float lr, float bc1, float bc2, const float* loss_part, float* loss,
                             hipStream_t s);
hipError_t launch_mlp_predict(const MlpArgs& a, const float* X, int n, float* out, hipStream_t s);

}  // namespace roadmc

#endif
