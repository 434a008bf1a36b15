/*
 * CPU oracle, surrogate MLP.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Mean squared error of the network n_in -> H -> H -> n_out (tanh hidden
 * layers) and its gradient, by reverse mode in FP64.  Parameter layout as in
 * transportoptimiser_amd/csrc/mlp_kernels.hip.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

double orc_mlp_loss_grad(int n_in, int H, int n_out, const double* w, const double* X,
                         const double* Y, int n, double* grad) {
  const double* W1 = w;
  const double* b1 = W1 + n_in * H;
  const double* W2 = b1 + H;
  const double* b2 = W2 + H * H;
  const double* W3 = b2 + H;
  const double* b3 = W3 + H * n_out;
  const int np = n_in * H + H + H * H + H + H * n_out + n_out;
  double *gW1 = grad, *gb1 = gW1 + n_in * H, *gW2 = gb1 + H, *gb2 = gW2 + H * H,
         *gW3 = gb2 + H, *gb3 = gW3 + H * n_out;
  double* h1 = (double*)malloc(sizeof(double) * H);
  double* h2 = (double*)malloc(sizeof(double) * H);
  double* d1 = (double*)malloc(sizeof(double) * H);
  double* d2 = (double*)malloc(sizeof(double) * H);
  double dy[16], loss = 0.0;
  int s, i, j, k;
  if (grad) memset(grad, 0, sizeof(double) * np);
  for (s = 0; s < n; ++s) {
    const double* x = X + (size_t)s * n_in;
    for (j = 0; j < H; ++j) {
      double v = b1[j];
      for (k = 0; k < n_in; ++k) v += x[k] * W1[k * H + j];
      h1[j] = tanh(v);
    }
    for (j = 0; j < H; ++j) {
      double v = b2[j];
      for (k = 0; k < H; ++k) v += h1[k] * W2[k * H + j];
      h2[j] = tanh(v);
    }
    for (k = 0; k < n_out; ++k) {
      double v = b3[k];
      for (j = 0; j < H; ++j) v += h2[j] * W3[j * n_out + k];
      v -= Y[(size_t)s * n_out + k];
      loss += v * v / ((double)n * n_out);
      dy[k] = 2.0 * v / ((double)n * n_out);
    }
    if (!grad) continue;
    for (j = 0; j < H; ++j) {
      double v = 0.0;
      for (k = 0; k < n_out; ++k) {
        v += dy[k] * W3[j * n_out + k];
        gW3[j * n_out + k] += h2[j] * dy[k];
      }
      d2[j] = v * (1.0 - h2[j] * h2[j]);
    }
    for (k = 0; k < n_out; ++k) gb3[k] += dy[k];
    for (i = 0; i < H; ++i) {
      double v = 0.0;
      for (j = 0; j < H; ++j) {
        v += d2[j] * W2[i * H + j];
        gW2[i * H + j] += h1[i] * d2[j];
      }
      d1[i] = v * (1.0 - h1[i] * h1[i]);
    }
    for (j = 0; j < H; ++j) gb2[j] += d2[j];
    for (i = 0; i < n_in; ++i)
      for (j = 0; j < H; ++j) gW1[i * H + j] += x[i] * d1[j];
    for (j = 0; j < H; ++j) gb1[j] += d1[j];
  }
  free(h1);
  free(h2);
  free(d1);
  free(d2);
  return loss;
}
