// augur_hmc.h -- device helpers shared by the HMC samplers (plain and batched
// hierarchical logistic regression).
#pragma once
#include "augur_common.h"

namespace augur {

constexpr int kMaxChains = 16;

struct HmcStep {
  double step[kMaxChains];
  int steps[kMaxChains];
};

// eps' = eps U(0.8, 1.2); L = max(1, round(U(0.9, 1.1) simLen / eps'))
// (cbits/augur_mcmclib.h:401-405), uniforms from the block's jitter stream.
__host__ __device__ inline void hmc_jitter(const Stream& s, double sim_len, double eps,
                                          int fixed_steps, double* step, int* steps) {
#pragma clang fp contract(off)
  if (fixed_steps > 0) {
    *step = eps;
    *steps = fixed_steps;
    return;
  }
  const u32x4 w = s.draw(0, 0);
  if (fixed_steps < 0) {  // the reference's step jitter, a fixed step count
    *step = (0.8 + 0.4 * u01(w.x)) * eps;
    *steps = -fixed_steps;
    return;
  }
  const double st = (0.8 + 0.4 * u01(w.x)) * eps;
  const double sc = 0.9 + 0.2 * u01(w.y);
  const double r = round(sc * sim_len / st);
  *step = st;
  *steps = r < 1.0 ? 1 : (int)r;
}

// cbits/augur_dist.c:372-381 (ref_compat: -log var; else -0.5 log var)
__device__ inline double normal_ll0(double x, double var, int ref_compat) {
  if (var > 0.0) return -(ref_compat ? 1.0 : 0.5) * log(var) - 0.5 * x * x / var;
  return -kLargePosNum;
}

// cbits/augur_dist.c:391-394 (ref_compat keeps -0.5 * var)
__device__ inline double normal_dotarg2_0(double x, double var, int ref_compat) {
  return (ref_compat ? -0.5 * var : -0.5 / var) + 0.5 * (x * x) / (var * var);
}

// The v block's full conditional depends on the K coefficients only through
// K and S = sum x_j^2: sum_j normal_ll(x_j; 0, v) and its v-derivative
// (normal_dotarg2 summed) in closed form, so a v trajectory needs no
// per-step reduction.
__device__ inline double normal_sum_ll(int K, double S, double v, int ref_compat) {
  if (v > 0.0) return (double)K * (-(ref_compat ? 1.0 : 0.5) * log(v)) - 0.5 * S / v;
  return (double)K * -kLargePosNum;
}
__device__ inline double normal_sum_dv(int K, double S, double v, int ref_compat) {
  return (double)K * (ref_compat ? -0.5 * v : -0.5 / v) + 0.5 * S / (v * v);
}

// Prior of the variance: InvGamma(a, b) (cbits/augur_dist.c:345-351) or
// Exponential(a) (:231-247).
__device__ inline double v_prior_ll0(int kind, double a, double b, double v) {
  if (kind == 0) return a * log(b) - lgamma(a) - (a + 1.0) * log(v) - b / v;
  if (a > 0.0 && v > 0.0) return log(a) - a * v;
  return -kLargePosNum;
}
__device__ inline double v_prior_dot0(int kind, double a, double b, double v) {
  if (kind == 0) return -(a + 1.0) / v + b / (v * v);
  return (a > 0.0 && v > 0.0) ? -a : 0.0;
}

// One update of the variance v given K terms with S = sum x_j^2, for the
// block whose streams are (seed, chain, sweep, tag + 0 | 1 | 2):
//   mwg = 0: HMC (cbits/augur_mcmclib.h:389-424), jitter from tag + 1,
//            momentum tag + 0, accept uniform tag + 2;
//   mwg = 1: Metropolis-within-Gibbs (cbits/augur_mcmclib.h:178-193) with the
//            proposal v' ~ Normal(v, eps) (eps is the proposal variance, as in
//            the schedule's `MWG [v] ~ Normal(v, eps)`), normal from tag + 0,
//            accept uniform tag + 2.  The Hastings term of the reference's
//            generated proposal (Core/CgUserProp.hs:130-141) evaluates both
//            proposal densities with the current point's parameters:
//            ll(v | v, eps) - ll(v' | v, eps) = (v' - v)^2 / (2 eps), kept with
//            ref_compat; otherwise 0, the correct term of a symmetric proposal.
// Returns whether the move was accepted; *v_out is the new value.
__device__ inline bool v_update(int K, double S, double v0, int ref_compat, int prior_kind, double prior_a,
                                double prior_b, int mwg, double sim_len, double eps, int fixed_steps,
                                uint32_t seed, uint32_t chain, uint32_t sweep, uint32_t tag, double* v_out) {
  auto ll_of = [&](double v) { return normal_sum_ll(K, S, v, ref_compat) + v_prior_ll0(prior_kind, prior_a, prior_b, v); };
  const double u = u01(Stream{seed, chain, sweep, tag + 2}.draw(0, 0).x);
  double v = v0, extra;
  if (mwg) {
    v = v0 + sqrt(eps) * normal_at(Stream{seed, chain, sweep, tag + 0}, 0, 0);
    extra = ref_compat ? 0.5 * (v - v0) * (v - v0) / eps : 0.0;
  } else {
    auto grad_of = [&](double w) {
      return normal_sum_dv(K, S, w, ref_compat) + v_prior_dot0(prior_kind, prior_a, prior_b, w);
    };
    double step;
    int steps;
    hmc_jitter(Stream{seed, chain, sweep, tag + 1}, sim_len, eps, fixed_steps, &step, &steps);
    const double p0 = normal_at(Stream{seed, chain, sweep, tag + 0}, 0, 0);
    double p = p0 + 0.5 * step * grad_of(v);
    for (int l = 0; l < steps; ++l) {
      v += step * p;
      if (l != steps - 1) p += step * grad_of(v);
    }
    p += 0.5 * step * grad_of(v);
    extra = 0.5 * p0 * p0 - 0.5 * p * p;
  }
  const bool acc = log(u) < (ll_of(v) - ll_of(v0)) + extra;
  *v_out = acc ? v : v0;
  return acc;
}

// Data term of one row: e = exp(-|t|), u = 1 + e; sigmoid(|t|) = 1/u and
// sigmoid(-|t|) = e/u; a = y - sigmoid(t) (= bernoulli_dotarg1 *
// expit_grad_0, cbits/augur_dist.c:68-71, cbits/augur_math.c:56-59), taken
// as sigmoid(-t) for y = 1 and -sigmoid(t) for y = 0 so that it carries no
// cancellation; returns log sigmoid(+-t) = min(+-t, 0) - log1p(e).
// FP32 on the hardware exp2 / log2 / reciprocal (about one ulp each).  The
// exponent -|t| log2(e) carries its rounding error and the tail of log2(e)
// as a first-order correction (2^(h + l) = 2^h (1 + l ln 2)), so e keeps
// about one ulp for any |t|; log1p(e) = log(u) + (e - (u - 1)) / u: u - 1
// is exact (u in [1, 2]) and so is e - (u - 1), the rounding error of 1 + e,
// so the first-order term ln(1 + r/u) ~ r/u restores what the rounding of u
// took (the error left is O(r^2), below an ulp); it reuses 1/u, and when u
// rounds to 1 it gives e itself.  (Goldberg 1991, Thm. 4's e log(u) / (u - 1)
// needs one more reciprocal.)
struct SigParts {
  float e, u, inv;  // exp(-|t|), 1 + e, 1 / u
};

__device__ inline SigParts sig_parts(float t) {
  constexpr float kL2eHi = 1.44269502162933349609375f;  // log2(e) rounded to FP32
  constexpr float kL2eLo = 1.92596299e-08f;             // log2(e) - kL2eHi
  const float at = fabsf(t);
  const float h = -at * kL2eHi;
  const float l = fmaf(-at, kL2eHi, -h) - at * kL2eLo;
  const float e0 = __builtin_amdgcn_exp2f(h);
  const float e = fmaf(e0, l * 0.693147180559945309f, e0);
  const float u = 1.0f + e;
  return {e, u, __builtin_amdgcn_rcpf(u)};
}

// y = 1: sigmoid(-t); y = 0: -sigmoid(t).  The magnitude is the small value
// e / u exactly when y = 1 and t >= 0 or y = 0 and t < 0, else 1 / u.
__device__ inline float sig_adjoint(float t, int y, const SigParts& p) {
  const float se = p.e * p.inv;
  const float mag = ((y != 0) == (t >= 0.0f)) ? se : p.inv;
  return y ? mag : -mag;
}

__device__ inline float log_sigmoid_pm(float t, int y, float* a_out) {
  const SigParts p = sig_parts(t);
  *a_out = sig_adjoint(t, y, p);
  const float l1p = fmaf(p.e - (p.u - 1.0f), p.inv, __builtin_amdgcn_logf(p.u) * 0.693147180559945309f);
  const float st = y ? t : -t;
  return fminf(st, 0.0f) - l1p;
}

// The adjoint alone, the same value log_sigmoid_pm stores in *a_out: a
// trajectory's leapfrog passes before its last need the gradient only.
__device__ inline float sigmoid_adjoint_pm(float t, int y) { return sig_adjoint(t, y, sig_parts(t)); }

template <int CTRL>
__device__ inline float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the LPR lanes of an aligned group (LPR = 2, 4, 8 or 16), every
// lane of the group receiving it: DPP quad permutations for lanes i ^ 1 and
// i ^ 2, then the half-row mirror (quad 0 <-> quad 1 within 8 lanes) and the
// row mirror (the two halves of a 16-lane row).
template <int LPR>
__device__ inline float row_sum(float v) {
  v += dpp_f<0xB1>(v);                            // quad_perm 1,0,3,2
  if constexpr (LPR >= 4) v += dpp_f<0x4E>(v);    // quad_perm 2,3,0,1
  if constexpr (LPR >= 8) v += dpp_f<0x141>(v);   // row_half_mirror
  if constexpr (LPR >= 16) v += dpp_f<0x140>(v);  // row_mirror
  return v;
}

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

}  // namespace augur
