// augur_mlp.h -- the gradient-surrogate MLP (D -> H -> H -> D, GELU) trained
// and evaluated on the device by the runtime's own kernels.
//
// Parameters live in one FP32 vector: W1 (H x D), b1 (H), W2 (H x H), b2 (H),
// W3 (D x H), b3 (D).  Inputs and targets are standardised per column; the
// network predicts the standardised target.  Training products take BF16
// operands (rounded to nearest even) with FP32 accumulation; biases,
// activations, losses and the Adam state stay FP32.
#pragma once
#include "augur_common.h"

namespace augur {

struct MlpShape {
  int d, h;
  size_t w1() const { return 0; }
  size_t b1() const { return (size_t)h * d; }
  size_t w2() const { return b1() + h; }
  size_t b2() const { return w2() + (size_t)h * h; }
  size_t w3() const { return b2() + h; }
  size_t b3() const { return w3() + (size_t)d * h; }
  size_t count() const { return b3() + d; }
};

// Training state and scratch for batches of `batch` rows.
struct Mlp {
  MlpShape s{};
  int batch = 0;
  DevBuf<float> params, grads, m1, m2;
  DevBuf<float> z1, h1, z2, h2, out, dout, dh, dz;
  DevBuf<float> kpart;  // split-k partial tiles of the training GEMMs
  DevBuf<double> loss;  // one value per launch of the loss kernel
  int adam_t = 0;

  void init(int d, int h, int batch_rows, uint32_t seed, hipStream_t st);
  // out = f(x) for `rows` <= batch rows of x (rows x d, standardised)
  void forward(const float* x, int rows, hipStream_t st);
  // forward + MSE loss against t + gradients into `grads`; loss in loss.p[0]
  void loss_grad(const float* x, const float* t, int rows, hipStream_t st);
  void adam(double lr, hipStream_t st);
  // C = A B (+ bias, GELU) with BF16 operands and FP32 accumulation on the
  // matrix cores; strides as gemm_f32
  void mm(hipStream_t st, int M, int N, int K, const float* A, long sam, long sak, const float* B, long sbk,
          long sbn, float* C, int ldc, const float* bias = nullptr, float* zout = nullptr, int act = 0);
};

// C (M x N, row stride ldc) = A B with A(m, k) = A[m sam + k sak] and
// B(k, n) = B[k sbk + n sbn]; FP32, fixed summation order.
void gemm_f32(hipStream_t st, int M, int N, int K, const float* A, long sam, long sak, const float* B,
              long sbk, long sbn, float* C, int ldc);

// stats = [mean x | 1/sd x | mean t | 1/sd t] over n rows of x and t (n x d)
void mlp_colstats(const float* x, const float* t, int n, int d, float* stats, hipStream_t st);
// `rows` rows drawn uniformly from the n pairs (Philox, step), standardised
void mlp_gather(const float* x, const float* t, int n, int d, const float* stats, int rows,
                uint32_t seed, uint32_t step, float* xb, float* tb, hipStream_t st);

}  // namespace augur
