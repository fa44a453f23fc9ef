// augur_surrogate.h -- surrogate-gradient HMC for the theta block of the
// hierarchical logistic regression sampler.
//
// The data term's gradient is replaced inside the leapfrog by a fitted map
// of theta: a least-squares linear term plus an MLP trained on the linear
// term's residual, both fitted on (theta, gradient) pairs that exact HMC
// produced; the prior term stays exact.  A leapfrog map is volume preserving
// and reversible for any position-dependent force, so the Metropolis test
// with the exact log-density keeps the target exact.  The proposal's
// log-density comes from a data pass in FP64 arithmetic.
#pragma once
#include "augur_common.h"
#include "augur_hmc.h"

namespace augur {

// Inference form of the fitted gradient (all scalings folded in):
//   g(theta) = lin(theta - lin_mean) + W3 h2 + b3,
//   h2 = gelu(W2 h1 + b2), h1 = gelu(W1 xin + b1),
//   xin = (theta - in_mean) * in_scale.
struct SurrTrajArgs {
  int d, h, P, ds, c_real;
  HmcStep hs;                  // per-chain step size and leapfrog count
  uint32_t seed, sweep, tag;   // momentum stream: tag + 0, as exact HMC
  const uint16_t* w1t;  // d x h (BF16), W1 transposed
  const float* b1;      // h
  const uint16_t* w2t;  // h x h (BF16), W2 transposed
  const float* b2;      // h
  const uint16_t* w3t;  // h x d (BF16), W3 transposed
  const float* b3;      // d
  const float* in_mean;   // d
  const float* in_scale;  // d
  const float* lin_mean;  // d
  const float* lin_t;     // d x d (FP32), lin_t[i * d + o] = coefficient of input i in output o
  const double* state;  // C x P
  double* prop;         // C x P
  double* p;            // C x ds
  double* k0;           // C
  float* beta32;        // C x ds
};

struct Ll64Args {
  const float* x;
  const int32_t* y;
  int n, d, cp, has_b, P;
  const double* point;  // cp x P: (theta, b) evaluated in FP64
  double* part;         // [cp][G]
  int G;
  double* out;          // cp
};

// Device buffers of the inference form.
struct SurrInfer {
  DevBuf<uint16_t> w1t, w2t, w3t;
  DevBuf<float> b3, in_mean, in_scale, lin_mean, lin_t;
  void alloc(int d, int h) {
    w1t.alloc((size_t)d * h);
    w2t.alloc((size_t)h * h);
    w3t.alloc((size_t)h * d);
    b3.alloc(d);
    in_mean.alloc(d);
    in_scale.alloc(d);
    lin_mean.alloc(d);
    lin_t.alloc((size_t)d * d);
  }
};

void launch_surr_traj(const SurrTrajArgs& a, int chains, hipStream_t st);
// Grid of launch_ll64 for these shapes (generic_grid when the matrix-core
// form does not apply); part must hold cp x that many doubles.
bool ll64_mfma_form(int d, int cp, int has_b);
int ll64_grid(int n, int d, int cp, int has_b, int generic_grid, int cus);
void launch_ll64(const Ll64Args& a, hipStream_t st);

// In place: x = (x - mx) * ix, t = (t - mt) * it, stats = [mx | ix | mt | it].
void surr_standardise(float* x, float* t, int n, int d, const float* stats, hipStream_t st);
// Least-squares fit t ~ x B over n rows (B: d x d, FP64, B[i * d + o]) by the
// normal equations with ridge * n added to the diagonal, then t -= x B in
// place.  Returns the FP64 solution in B.
void surr_linear_fit(const float* x, float* t, int n, int d, double ridge, double* B, hipStream_t st);
// Folds the two standardisations (stats1 of the raw pairs, stats2 of the
// standardised inputs and the residual targets), the linear map B and the
// network parameters into the inference form.
void surr_fold(int d, int h, const float* stats1, const float* stats2, const double* B,
               const float* params, SurrInfer& out, hipStream_t st);

}  // namespace augur
