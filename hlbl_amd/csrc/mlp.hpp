// The MLP surrogate (host-side interface).
#pragma once

#include <hip/hip_runtime.h>

#include "context.hpp"
#include "lattice.hpp"

namespace hlbl {

// offsets of W[out][in] and b[out] of the five layers in the packed parameters
struct MlpLayout {
  int in[5], out[5], w[5], b[5], total;
};
HLBL_HD MlpLayout mlp_layout() {
  MlpLayout L{};
  int off = 0;
  for (int l = 0; l < 5; ++l) {
    L.in[l] = l == 0 ? kMlpIn : kMlpHidden;
    L.out[l] = l == 4 ? kMlpOut : kMlpHidden;
    L.w[l] = off;
    off += L.in[l] * L.out[l];
    L.b[l] = off;
    off += L.out[l];
  }
  L.total = off;
  return L;
}
constexpr int kMlpParamsPadded = (kMlpParams + 7) / 8 * 8;

void mlp_init_params(hlbl_qed* q, uint64_t seed);
void mlp_requantise(hlbl_qed* q, hipStream_t st);
void mlp_train(hlbl_qed* q, int n_samples, int epochs, int batch, float lr, uint64_t seed,
               float* loss_hist, hipStream_t st);
// forward of n rows with the FP32 master weights (training kernels)
struct MlpWork;
void mlp_forward_fp32(hlbl_qed* q, const float* x, int n, float* y, hipStream_t st);
// forward with the BF16 weights on the inference kernel
void mlp_infer(hlbl_qed* q, const float* x, int n, float* y, hipStream_t st);
// MSE loss and gradient at the BF16-rounded weights
double mlp_loss_grad(hlbl_qed* q, const float* x, const float* t, int n, float* grad, hipStream_t st);

}  // namespace hlbl
