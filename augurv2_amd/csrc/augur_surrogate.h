// augur_surrogate.h -- surrogate-gradient HMC for the theta block of the
// hierarchical logistic regression sampler.
//
// The data term's gradient is replaced inside the leapfrog by an MLP trained
// on (theta, gradient) pairs that exact HMC produced; the prior term stays
// exact.  A leapfrog map is volume preserving and reversible for any
// position-dependent force, so the Metropolis test with the exact
// log-density keeps the target exact.  The proposal's log-density comes from
// a data pass in FP64 arithmetic.
#pragma once
#include "augur_common.h"

namespace augur {

struct SurrTrajArgs {
  int d, h, P, ds, c_real;
  double step;
  int steps;
  uint32_t seed, sweep, tag;  // momentum stream: tag + 0, as exact HMC
  const uint16_t* w1t;  // d x h (BF16), W1 transposed
  const float* b1;      // h
  const uint16_t* w2t;  // h x h (BF16), W2 transposed
  const float* b2;      // h
  const uint16_t* w3t;  // h x d (BF16), W3 transposed
  const float* b3;      // d
  const float* stats;   // [mean x | 1/sd x | mean g | 1/sd g], d each
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

void launch_surr_traj(const SurrTrajArgs& a, int chains, hipStream_t st);
void launch_ll64(const Ll64Args& a, hipStream_t st);
// Transposed BF16 inference copies from the FP32 parameter vector.
void make_inference_weights(const float* params, int d, int h, uint16_t* w1t, uint16_t* w2t,
                            uint16_t* w3t, hipStream_t st);

}  // namespace augur
