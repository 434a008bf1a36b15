// Surrogate MLP kernels (FP32): an ensemble of E networks
// n_in -> H -> H -> n_out with tanh hidden layers, trained by Adam on the
// mean squared error.  The reference fits its GA surrogate with kNN local
// regression on a grid (SimulateGPU.cu:2934-3155, 5533-5968); this network
// replaces that fit and full Monte Carlo stays as the fallback.
//
// Parameter layout of one member: W1[n_in][H], b1[H], W2[H][H], b2[H],
// W3[H][n_out], b3[n_out] (row-major, input index first).
#include <hip/hip_runtime.h>

#include "mlp_kernels.h"

namespace roadmc {

constexpr int kMlpH = 64;
constexpr int kMlpChunk = 64;  // samples per gradient block

__host__ __device__ inline int mlp_n_params(int n_in, int n_out) {
  return n_in * kMlpH + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH * n_out + n_out;
}

// One block: member e = blockIdx.y, samples [chunk*64, chunk*64 + 64).
__global__ __launch_bounds__(256) void mlp_grad_kernel(MlpArgs a, const float* __restrict__ X,
                                                       const float* __restrict__ Y, int n,
                                                       float* __restrict__ partial,
                                                       float* __restrict__ loss_part) {
  constexpr int H = kMlpH, S = kMlpChunk;
  const int NI = a.n_in, NO = a.n_out, NP = mlp_n_params(NI, NO);
  const int e = blockIdx.y, chunk = blockIdx.x, s0 = chunk * S;
  const float* w = a.params + (size_t)e * NP;
  extern __shared__ float sm[];
  float* W = sm;                 // NP
  float* x = W + NP;             // [S][NI]
  float* h1 = x + S * NI;        // [S][H]
  float* h2 = h1 + S * H;        // [S][H]
  float* d1 = h2 + S * H;        // [S][H]
  float* d2 = d1 + S * H;        // [S][H]
  float* dy = d2 + S * H;        // [S][NO]
  float* red = dy + S * NO;      // [256]
  const float* W1 = W;
  const float* b1 = W1 + NI * H;
  const float* W2 = b1 + H;
  const float* b2 = W2 + H * H;
  const float* W3 = b2 + H;
  const float* b3 = W3 + H * NO;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < NP; i += nt) W[i] = w[i];
  for (int i = tid; i < S * NI; i += nt) {
    const int s = i / NI;
    x[i] = (s0 + s < n) ? X[(size_t)(s0 + s) * NI + (i - s * NI)] : 0.0f;
  }
  __syncthreads();
  // h1 = tanh(x W1 + b1)
  for (int i = tid; i < S * H; i += nt) {
    const int s = i / H, j = i - s * H;
    float v = b1[j];
    for (int k = 0; k < NI; ++k) v += x[s * NI + k] * W1[k * H + j];
    h1[i] = tanhf(v);
  }
  __syncthreads();
  for (int i = tid; i < S * H; i += nt) {
    const int s = i / H, j = i - s * H;
    float v = b2[j];
    for (int k = 0; k < H; ++k) v += h1[s * H + k] * W2[k * H + j];
    h2[i] = tanhf(v);
  }
  __syncthreads();
  // y = h2 W3 + b3; MSE over n samples and NO outputs.
  float lsum = 0.0f;
  const float scale = 2.0f / ((float)n * (float)NO);
  for (int i = tid; i < S * NO; i += nt) {
    const int s = i / NO, k = i - s * NO;
    float v = b3[k];
    for (int j = 0; j < H; ++j) v += h2[s * H + j] * W3[j * NO + k];
    float g = 0.0f;
    if (s0 + s < n) {
      const float r = v - Y[(size_t)(s0 + s) * NO + k];
      lsum += r * r;
      g = scale * r;
    }
    dy[i] = g;
  }
  red[tid] = lsum;
  __syncthreads();
  // d2 = (dy W3^T) o (1 - h2^2)
  for (int i = tid; i < S * H; i += nt) {
    const int s = i / H, j = i - s * H;
    float v = 0.0f;
    for (int k = 0; k < NO; ++k) v += dy[s * NO + k] * W3[j * NO + k];
    d2[i] = v * (1.0f - h2[i] * h2[i]);
  }
  __syncthreads();
  // d1 = (d2 W2^T) o (1 - h1^2)
  for (int i = tid; i < S * H; i += nt) {
    const int s = i / H, j = i - s * H;
    float v = 0.0f;
    for (int k = 0; k < H; ++k) v += d2[s * H + k] * W2[j * H + k];
    d1[i] = v * (1.0f - h1[i] * h1[i]);
  }
  __syncthreads();
  // Weight gradients summed over the block's samples (fixed order).
  float* g = partial + ((size_t)e * gridDim.x + chunk) * NP;
  for (int p = tid; p < NP; p += nt) {
    float v = 0.0f;
    int q = p;
    if (q < NI * H) {
      const int i = q / H, j = q - i * H;
      for (int s = 0; s < S; ++s) v += x[s * NI + i] * d1[s * H + j];
    } else if ((q -= NI * H) < H) {
      for (int s = 0; s < S; ++s) v += d1[s * H + q];
    } else if ((q -= H) < H * H) {
      const int i = q / H, j = q - i * H;
      for (int s = 0; s < S; ++s) v += h1[s * H + i] * d2[s * H + j];
    } else if ((q -= H * H) < H) {
      for (int s = 0; s < S; ++s) v += d2[s * H + q];
    } else if ((q -= H) < H * NO) {
      const int i = q / NO, k = q - i * NO;
      for (int s = 0; s < S; ++s) v += h2[s * H + i] * dy[s * NO + k];
    } else {
      q -= H * NO;
      for (int s = 0; s < S; ++s) v += dy[s * NO + q];
    }
    g[p] = v;
  }
  if (tid == 0) {
    float l = 0.0f;
    for (int i = 0; i < nt; ++i) l += red[i];
    loss_part[(size_t)e * gridDim.x + chunk] = l / ((float)n * (float)NO);
  }
}

// Sums the block partials (fixed order) and, when lr > 0, applies one Adam
// step; grad_out (optional) receives the summed gradient.
__global__ void mlp_adam_kernel(MlpArgs a, const float* __restrict__ partial, int nchunk,
                                float* __restrict__ grad_out, float lr, float bc1, float bc2) {
  const int NP = mlp_n_params(a.n_in, a.n_out);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NP * a.E) return;
  const int e = i / NP, p = i - e * NP;
  float g = 0.0f;
  for (int c = 0; c < nchunk; ++c) g += partial[((size_t)e * nchunk + c) * NP + p];
  if (grad_out) grad_out[i] = g;
  if (lr > 0.0f) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float m = b1 * a.m[i] + (1.0f - b1) * g;
    const float v = b2 * a.v[i] + (1.0f - b2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    a.params[i] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
  }
}

__global__ void mlp_loss_kernel(const float* __restrict__ loss_part, int nchunk, int E,
                                float* __restrict__ loss) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  float l = 0.0f;
  for (int c = 0; c < nchunk; ++c) l += loss_part[(size_t)e * nchunk + c];
  loss[e] = l;
}

// out[e][s][k]: one thread per (sample, member), weights read through L1/L2.
__global__ void mlp_predict_kernel(MlpArgs a, const float* __restrict__ X, int n,
                                   float* __restrict__ out) {
  constexpr int H = kMlpH;
  const int NI = a.n_in, NO = a.n_out, NP = mlp_n_params(NI, NO);
  const int s = blockIdx.x * blockDim.x + threadIdx.x, e = blockIdx.y;
  if (s >= n) return;
  const float* w = a.params + (size_t)e * NP;
  const float* W1 = w;
  const float* b1 = W1 + NI * H;
  const float* W2 = b1 + H;
  const float* b2 = W2 + H * H;
  const float* W3 = b2 + H;
  const float* b3 = W3 + H * NO;
  float xv[kMlpMaxIn], h1[H];
  for (int k = 0; k < NI; ++k) xv[k] = X[(size_t)s * NI + k];
  for (int j = 0; j < H; ++j) {
    float v = b1[j];
    for (int k = 0; k < NI; ++k) v += xv[k] * W1[k * H + j];
    h1[j] = tanhf(v);
  }
  float y[kMlpMaxOut];
  for (int k = 0; k < NO; ++k) y[k] = b3[k];
  for (int j = 0; j < H; ++j) {
    float v = b2[j];
    for (int k = 0; k < H; ++k) v += h1[k] * W2[k * H + j];
    const float h = tanhf(v);
    for (int k = 0; k < NO; ++k) y[k] += h * W3[j * NO + k];
  }
  for (int k = 0; k < NO; ++k) out[((size_t)e * n + s) * NO + k] = y[k];
}

int mlp_param_count(int n_in, int n_out) { return mlp_n_params(n_in, n_out); }

size_t mlp_grad_lds(int n_in, int n_out) {
  const int S = kMlpChunk, H = kMlpH;
  return sizeof(float) * ((size_t)mlp_n_params(n_in, n_out) + S * n_in + 4 * S * H + S * n_out + 256);
}

hipError_t mlp_set_lds_limit(int n_in, int n_out) {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_grad_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)mlp_grad_lds(n_in, n_out));
}

hipError_t launch_mlp_grad(const MlpArgs& a, const float* X, const float* Y, int n, float* partial,
                           float* loss_part, hipStream_t s) {
  const dim3 grid((n + kMlpChunk - 1) / kMlpChunk, a.E);
  hipLaunchKernelGGL(mlp_grad_kernel, grid, dim3(256), mlp_grad_lds(a.n_in, a.n_out), s, a, X, Y,
                     n, partial, loss_part);
  return hipGetLastError();
}

hipError_t launch_mlp_update(const MlpArgs& a, const float* partial, int nchunk, float* grad_out,
                             float lr, float bc1, float bc2, const float* loss_part, float* loss,
                             hipStream_t s) {
  const int total = mlp_n_params(a.n_in, a.n_out) * a.E;
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((total + 255) / 256), dim3(256), 0, s, a, partial, nchunk,
                     grad_out, lr, bc1, bc2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !loss) return e;
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(64), 0, s, loss_part, nchunk, a.E, loss);
  return hipGetLastError();
}

hipError_t launch_mlp_predict(const MlpArgs& a, const float* X, int n, float* out, hipStream_t s) {
  const dim3 grid((n + 255) / 256, a.E);
  hipLaunchKernelGGL(mlp_predict_kernel, grid, dim3(256), 0, s, a, X, n, out);
  return hipGetLastError();
}

}  // namespace roadmc
