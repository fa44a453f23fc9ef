/*
 * augur_oracle.c -- CPU FP64 restatement of the AugurV2 sampler core.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP runtime, and the CPU
 * baseline timed beside it).  Nothing in the product links this file.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math, so
 * FP32 expressions round exactly as written and match the GPU bit for bit).
 *
 * Reference files restated here (paths relative to the AugurV2 tree):
 *   cbits/augur_dist.c      densities, derivatives, samplers, conjugate draws
 *   cbits/augur_math.c      expit and its derivative
 *   cbits/augur_matop.c     Cholesky family
 *   cbits/augur_mcmclib.h   HMC, leapfrog, Metropolis accept test
 *   compiler/augur/src/Core/CgDisc.hs, CgConj.hs, CgGrad.hs, CgLike.hs
 *   compiler/augur/src/CudaC/CgKern.hs  sweep order
 *
 * Deliberate, documented differences from the reference (same target
 * distribution in every case):
 *   - RNG: Philox4x32-10 counter streams instead of GSL MT19937 / curand
 *     XORWOW (the reference streams cannot be reproduced without GSL/CUDA).
 *   - chi-squared(m) is drawn as 2 * Gamma(m / 2) instead of a sum of m
 *     squared normals (cbits/augur_dist.c:215-222), which is O(m).
 *   - the GMM categorical step works in log space with max subtraction
 *     instead of multiplying raw pdfs (cbits/augur_dist.c:798-813), which
 *     underflows to an all-zero pmf at D = 16.
 */
#include "augur_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                        */
/* ------------------------------------------------------------------------ */

static void or_mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t hi0, lo0, hi1, lo1;
    or_mulhilo(0xD2511F53u, c0, &hi0, &lo0);
    or_mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void or_draw4(uint32_t seed, uint32_t chain, uint32_t site, uint32_t sub,
              uint32_t sweep, uint32_t tag, uint32_t out[4]) {
  uint32_t ctr[4] = {site, sub, sweep, tag};
  uint32_t key[2] = {seed, chain};
  or_philox4x32_10(ctr, key, out);
}

double or_u01(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }

float or_u01f(uint32_t w) { return (float)(w >> 8) * 5.9604644775390625e-8f; }

double or_normal_from(uint32_t w0, uint32_t w1) {
  double u1 = or_u01(w0), u2 = or_u01(w1);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * OR_PI * u2);
}

double or_normal(uint32_t seed, uint32_t chain, uint32_t site, uint32_t sub,
                 uint32_t sweep, uint32_t tag) {
  uint32_t w[4];
  or_draw4(seed, chain, site, sub, sweep, tag, w);
  return or_normal_from(w[0], w[1]);
}

/* Marsaglia-Tsang, cbits/augur_dist.c:287-312 (shape < 1 boosted by U^(1/a)). */
double or_gamma(uint32_t seed, uint32_t chain, uint32_t site, uint32_t sweep,
                uint32_t tag, double shape) {
  if (!(shape > 0.0) || !isfinite(shape)) return NAN;
  double a = shape < 1.0 ? shape + 1.0 : shape;
  double d = a - 1.0 / 3.0;
  double c = 1.0 / sqrt(9.0 * d);
  for (uint32_t j = 0; j < OR_GAMMA_MAX_ATTEMPTS; j++) {
    uint32_t w[4];
    or_draw4(seed, chain, site, j, sweep, tag, w);
    double z = or_normal_from(w[0], w[1]);
    double u = or_u01(w[2]);
    double t = 1.0 + c * z;
    double v = t * t * t;
    if (z > -1.0 / c && log(u) < 0.5 * z * z + d - d * v + d * log(v)) {
      double x = d * v;
      if (shape < 1.0) x *= pow(or_u01(w[3]), 1.0 / shape);
      return x;
    }
  }
  return NAN;
}

/* ------------------------------------------------------------------------ */
/* Densities and derivatives.                                                */
/* ------------------------------------------------------------------------ */

/* cbits/augur_math.c:52-54 */
double or_expit(double x) { return 1.0 / (1.0 + exp(-x)); }

/* cbits/augur_dist.c:55-66 */
double or_bernoulli_ll(int y, double bias) {
  if (bias >= 0.0 && bias <= 1.0) return log(y * bias + (1 - y) * (1.0 - bias));
  return -OR_LARGE_POS_NUM;
}

/* cbits/augur_dist.c:68-71 */
double or_bernoulli_dotarg1(int y, double bias) {
  double denom = y * bias + (1 - y) * (1.0 - bias);
  return (double)(y + y - 1) / denom;
}

/* cbits/augur_dist.c:372-381.  The reference uses -log(var); the textbook
 * density has -0.5 log(var) (and -0.5 log(2 pi)).  ref_compat selects. */
double or_normal_ll(double pt, double mean, double var, int ref_compat) {
  if (var > 0.0) {
    double t = pt - mean;
    double c = ref_compat ? 1.0 : 0.5;
    return -c * log(var) - 0.5 * t * t / var;
  }
  return -OR_LARGE_POS_NUM;
}

/* cbits/augur_dist.c:383-385 */
double or_normal_dotpt(double pt, double mean, double var) { return -(pt - mean) / var; }

/* cbits/augur_dist.c:391-394 (reference: -0.5 * var; textbook: -0.5 / var) */
double or_normal_dotarg2(double pt, double mean, double var, int ref_compat) {
  double t = pt - mean;
  double first = ref_compat ? -0.5 * var : -0.5 / var;
  return first + 0.5 * (t * t) / (var * var);
}

/* cbits/augur_dist.c:345-351 */
double or_invgamma_ll(double pt, double shape, double scale) {
  return shape * log(scale) - lgamma(shape) - (shape + 1.0) * log(pt) - scale / pt;
}
double or_invgamma_dotpt(double pt, double shape, double scale) {
  return -(shape + 1.0) / pt + scale / (pt * pt);
}

/* cbits/augur_dist.c:231-247 */
double or_exponential_ll(double pt, double rate) {
  if (rate > 0.0 && pt > 0.0) return log(rate) - rate * pt;
  return -OR_LARGE_POS_NUM;
}
double or_exponential_dotpt(double pt, double rate) {
  if (rate > 0.0 && pt > 0.0) return -rate;
  return 0.0;
}

/* cbits/augur_dist.c:798-813 augur_unnorm_disc_samp, with the uniform given. */
int or_categorical_ref(const double* pmf, int k, double u) {
  double norm = 0.0;
  for (int i = 0; i < k; i++) norm += pmf[i];
  double acc = pmf[0] / norm;
  for (int i = 1; i < k; i++) {
    if (u < acc) return i - 1;
    acc += pmf[i] / norm;
  }
  return k - 1;
}

/* The categorical draw exactly as the HIP LDA kernel performs it in FP32:
 * pmf_i = a_i * b_i, running sum c_i, draw the first i with u * c_{k-1} < c_i
 * (i < k - 1), else k - 1.  Same choice as augur_unnorm_disc_samp with the
 * normalisation moved to the uniform side. */
int or_categorical_f32(const float* a, const float* b, int k, float u) {
  float cum[1024];
  float c = 0.0f;
  for (int i = 0; i < k; i++) {
    float p = a[i] * b[i];
    c = c + p;
    cum[i] = c;
  }
  float t = u * c;
  for (int i = 0; i < k - 1; i++)
    if (t < cum[i]) return i;
  return k - 1;
}

/* Log-space categorical draw (max subtraction), FP64. */
int or_categorical_logits(const double* logits, int k, double u) {
  double mx = -INFINITY;
  for (int i = 0; i < k; i++)
    if (logits[i] > mx) mx = logits[i];
  double cum[1024];
  double c = 0.0;
  for (int i = 0; i < k; i++) {
    c += exp(logits[i] - mx);
    cum[i] = c;
  }
  double t = u * c;
  for (int i = 0; i < k - 1; i++)
    if (t < cum[i]) return i;
  return k - 1;
}

/* cbits/augur_dist.c:526-538 augur_dirichlet_categorical_conj.  Gamma draw
 * for element i uses site site0 + i. */
void or_dirichlet_conj(uint32_t seed, uint32_t chain, uint32_t sweep, uint32_t tag,
                       uint32_t site0, const double* alpha, const int* cnts,
                       int k, double* out) {
  double acc = 0.0;
  for (int i = 0; i < k; i++) {
    double g = or_gamma(seed, chain, site0 + (uint32_t)i, sweep, tag,
                        alpha[i] + (double)(cnts ? cnts[i] : 0));
    out[i] = g;
    acc += g;
  }
  for (int i = 0; i < k; i++) out[i] /= acc;
}

/* ------------------------------------------------------------------------ */
/* Small dense FP64 linear algebra, row-major n x n.                         */
/* ------------------------------------------------------------------------ */
#define M(a, i, j) (a)[(i) * n + (j)]

/* cbits/augur_matop.c:312-328 (lower factor in place, upper part untouched) */
void or_mat_cholesky(double* a, int n) {
  for (int i = 0; i < n; i++) {
    for (int j = 0; j <= i; j++) {
      double s = 0.0;
      for (int k = 0; k < j; k++) s += M(a, i, k) * M(a, j, k);
      if (i == j)
        M(a, i, i) = sqrt(M(a, i, i) - s);
      else
        M(a, i, j) = (M(a, i, j) - s) / M(a, j, j);
    }
  }
}

/* cbits/augur_matop.c:335-345: inverse of the lower factor in place; the
 * strict upper part becomes zero. */
void or_mat_lower_trmi(double* l, int n) {
  for (int c = 0; c < n; c++) {
    for (int m = 0; m < n; m++) {
      double b = (c == m) ? 1.0 : 0.0;
      for (int i = 0; i < m; i++) b -= M(l, m, i) * M(l, i, c);
      M(l, m, c) = b / M(l, m, m);
    }
  }
}

/* cbits/augur_matop.c:363-375: from a Cholesky factor to the full inverse. */
void or_mat_cholesky_invert(double* llt, int n) {
  or_mat_lower_trmi(llt, n);
  for (int c = 0; c < n; c++) {
    for (int r = c; r < n; r++) {
      double acc = 0.0;
      for (int k = c; k < n; k++) acc += M(llt, k, r) * M(llt, k, c);
      M(llt, r, c) = acc;
    }
  }
  for (int c = 0; c < n; c++)
    for (int r = c + 1; r < n; r++) M(llt, c, r) = M(llt, r, c);
}

/* cbits/augur_matop.c:381-392: l2 = l1 * l2 for lower-triangular operands. */
void or_mat_trtrmm(const double* l1, double* l2, int n) {
  for (int i = n - 1; i >= 0; i--) {
    for (int j = 0; j < n; j++) {
      double t = 0.0;
      for (int k = 0; k < i; k++) t += M(l1, i, k) * M(l2, k, j);
      t += M(l1, i, i) * M(l2, i, j);
      M(l2, i, j) = t;
    }
  }
}

/* cbits/augur_matop.c:398-406: v = L v */
void or_mat_trvm(const double* l, double* v, int n) {
  for (int i = n - 1; i >= 0; i--) {
    double t = 0.0;
    for (int j = 0; j <= i; j++) t += M(l, i, j) * v[j];
    v[i] = t;
  }
}

/* cbits/augur_dist.c:548-568 */
double or_mvnormal_ll(const double* pt, const double* mean, const double* cov, int n) {
  double l[32 * 32], w[32];
  for (int i = 0; i < n; i++) w[i] = pt[i] - mean[i];
  memcpy(l, cov, sizeof(double) * n * n);
  or_mat_cholesky(l, n);
  double det = 1.0;
  for (int i = 0; i < n; i++) det *= M(l, i, i);
  det *= det;
  or_mat_lower_trmi(l, n);
  or_mat_trvm(l, w, n);
  double q = 0.0;
  for (int i = 0; i < n; i++) q += w[i] * w[i];
  double c = -0.5 * (double)n * log(2.0 * OR_PI);
  return c - 0.5 * log(det) - 0.5 * q;
}

/* cbits/augur_math.c:80-86: log of the multivariate Gamma function, with
 * the reference's constant log(3.1415) in place of log(pi). */
double or_multi_lgamma(int p, double x) {
  double r = (p * (p - 1) * 0.25) * log(3.1415);
  for (int i = 1; i < p + 1; i++) r += lgamma(x - (i - 1.0) / 2.0);
  return r;
}

/* cbits/augur_dist.c:488-496: sum (alpha_i - 1) log pt_i (the reference
 * leaves out the normalising constant). */
double or_dirichlet_ll(const double* pt, const double* alpha, int k) {
  double acc = 0.0;
  for (int i = 0; i < k; i++) acc += (alpha[i] - 1.0) * log(pt[i]);
  return acc;
}

/* cbits/augur_dist.c:679-711: determinants from the Cholesky factors'
 * diagonals, tr(scale X^-1) through the Cholesky inverse. */
double or_invwishart_ll(const double* pt, int df, const double* scale, int d) {
  double l[32 * 32], w[32 * 32];
  memcpy(l, scale, sizeof(double) * d * d);
  or_mat_cholesky(l, d);
  double det_psi = 1.0;
  for (int i = 0; i < d; i++) det_psi *= l[i * d + i];
  det_psi *= det_psi;
  memcpy(w, pt, sizeof(double) * d * d);
  or_mat_cholesky(w, d);
  double det_x = 1.0;
  for (int i = 0; i < d; i++) det_x *= w[i * d + i];
  det_x *= det_x;
  or_mat_cholesky_invert(w, d);
  double tr = 0.0;
  for (int i = 0; i < d; i++) {
    double t = 0.0;
    for (int j = 0; j < d; j++) t += scale[i * d + j] * w[j * d + i];
    tr += t;
  }
  const double c = -df * d / 2.0 * log(2.0) - or_multi_lgamma(d, df / 2.0);
  return c - (df + d + 1.0) / 2.0 * log(det_x) + df / 2.0 * log(det_psi) - 0.5 * tr;
}

/* The model's log density (the generated ll_mod behind augur_get_curr_ll,
 * CudaC/CgStInit.hs runCgCallLLMod) for hmvgmm.rv: Dirichlet(pis | alpha)
 * (sample_pi), MvNormal(mu_k | mu0, cov0), InvWishart(cov_k | df, scale0)
 * (sample_cov), Categorical(z_n | pis), MvNormal(y_n | mu_z, cov_z).  With
 * sample_pi / sample_cov = 0 (mvgmm.rv) pis and cov are hyperparameters and
 * carry no term. */
double or_gmm_ll_mod(const or_gmm_model* m, const double* pis, const double* mu, const double* cov,
                     const int* z) {
  const int d = m->d, k = m->k;
  double ll = 0.0;
  if (m->sample_pi) ll += or_dirichlet_ll(pis, m->alpha, k);
  for (int c = 0; c < k; c++) {
    ll += or_mvnormal_ll(mu + c * d, m->mu0, m->cov0, d);
    if (m->sample_cov) ll += or_invwishart_ll(cov + c * d * d, m->df, m->scale0, d);
  }
  for (int n = 0; n < m->n; n++) {
    const int zn = z[n];
    ll += log(pis[zn]) + or_mvnormal_ll(m->y + n * d, mu + zn * d, cov + zn * d * d, d);
  }
  return ll;
}

/* cbits/augur_dist.c:616-630 with the standard normals from sites site0 + i. */
static void or_mvnormal_sample(uint32_t seed, uint32_t chain, uint32_t sweep, uint32_t tag,
                               uint32_t site0, const double* mean, const double* cov,
                               int n, double* out) {
  double l[32 * 32];
  for (int i = 0; i < n; i++) out[i] = or_normal(seed, chain, site0 + (uint32_t)i, 0, sweep, tag);
  memcpy(l, cov, sizeof(double) * n * n);
  or_mat_cholesky(l, n);
  or_mat_trvm(l, out, n);
  for (int i = 0; i < n; i++) out[i] += mean[i];
}

/* cbits/augur_dist.c:633-669.  sum = sum of the assigned points (not yet
 * divided by cnt); the reference divides in place. */
void or_mvnormal_mvnormal_conj(uint32_t seed, uint32_t chain, uint32_t sweep, uint32_t site0,
                               const double* mu0, const double* cov0,
                               const double* sum, int cnt, const double* cov,
                               int n, double* out) {
  double l0[32 * 32], l[32 * 32], v[32], mbar[32], work[32], mean[32];
  memcpy(l0, cov0, sizeof(double) * n * n);
  or_mat_cholesky(l0, n);
  or_mat_cholesky_invert(l0, n);          /* cov0^-1 */
  memcpy(l, cov, sizeof(double) * n * n);
  or_mat_cholesky(l, n);
  or_mat_cholesky_invert(l, n);           /* cov^-1 */
  for (int i = 0; i < n * n; i++) l[i] *= (double)cnt; /* n cov^-1 */
  for (int i = 0; i < n; i++) {
    double t = 0.0;
    for (int j = 0; j < n; j++) t += M(l0, i, j) * mu0[j];
    v[i] = t;
  }
  double s = cnt == 0 ? 1.0 : 1.0 / (double)cnt;
  for (int i = 0; i < n; i++) mbar[i] = sum[i] * s;
  for (int i = 0; i < n; i++) {
    double t = 0.0;
    for (int j = 0; j < n; j++) t += M(l, i, j) * mbar[j];
    work[i] = t;
  }
  for (int i = 0; i < n; i++) v[i] += work[i];
  for (int i = 0; i < n * n; i++) l0[i] += l[i];
  or_mat_cholesky(l0, n);
  or_mat_cholesky_invert(l0, n);          /* A = (cov0^-1 + n cov^-1)^-1 */
  for (int i = 0; i < n; i++) {
    double t = 0.0;
    for (int j = 0; j < n; j++) t += M(l0, i, j) * v[j];
    mean[i] = t;
  }
  or_mvnormal_sample(seed, chain, sweep, OR_TAG_GMM_MU, site0, mean, l0, n, out);
}

/* cbits/augur_dist.c:722-749 (Bartlett).  Sites: element (i, j) of
 * component block site0 uses site0 + i * n + j. */
void or_invwishart_sample(uint32_t seed, uint32_t chain, uint32_t sweep, uint32_t site0,
                          int df, const double* scale, int n, double* out) {
  double l[32 * 32];
  memcpy(l, scale, sizeof(double) * n * n);
  or_mat_cholesky(l, n);
  or_mat_cholesky_invert(l, n);  /* scale^-1 */
  or_mat_cholesky(l, n);         /* chol(scale^-1) */
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) {
      uint32_t site = site0 + (uint32_t)(i * n + j);
      if (j < i)
        M(out, i, j) = or_normal(seed, chain, site, 0, sweep, OR_TAG_GMM_COV);
      else if (j > i)
        M(out, i, j) = 0.0;
      else
        M(out, i, j) = sqrt(2.0 * or_gamma(seed, chain, site, sweep, OR_TAG_GMM_COV,
                                           0.5 * (double)(df - i)));
    }
  }
  or_mat_trtrmm(l, out, n);
  or_mat_cholesky_invert(out, n);
}
#undef M

/* ------------------------------------------------------------------------ */
/* LDA.                                                                      */
/* ------------------------------------------------------------------------ */

/* Core/CgConj.hs:171-183: cnt_theta[d][z]++, cnt_phi[z][w]++ */
void or_lda_counts(int n_docs, int k, int v, const int* doc_off, const int* w,
                   const int* z, int* cnt_theta, int* cnt_phi) {
  memset(cnt_theta, 0, sizeof(int) * (size_t)n_docs * k);
  memset(cnt_phi, 0, sizeof(int) * (size_t)k * v);
  for (int d = 0; d < n_docs; d++)
    for (int t = doc_off[d]; t < doc_off[d + 1]; t++) {
      cnt_theta[(size_t)d * k + z[t]]++;
      cnt_phi[(size_t)z[t] * v + w[t]]++;
    }
}

/* Core/CgDisc.hs:122-144 (pmf_k = theta[d][k] * phi[k][w]), drawn in FP32 as
 * the kernel does; phi_t is word-major (v x k). */
void or_lda_zstep_f32(int n_docs, int k, const int* doc_off, const int* w,
                      const float* theta, const float* phi_t, uint32_t seed,
                      uint32_t chain, uint32_t sweep, int* z) {
  for (int d = 0; d < n_docs; d++)
    for (int t = doc_off[d]; t < doc_off[d + 1]; t++) {
      uint32_t r[4];
      or_draw4(seed, chain, (uint32_t)t, 0, sweep, OR_TAG_LDA_Z, r);
      z[t] = or_categorical_f32(theta + (size_t)d * k, phi_t + (size_t)w[t] * k, k,
                                or_u01f(r[0]));
    }
}

/* One LDA sweep in the reference order (CudaC/CgKern.hs:192-238 with the
 * schedule ConjGibbs [theta] (*) ConjGibbs [phi] (*) DiscGibbs [z]).  The
 * z step is the FP64 reference form, using the same uniform as the kernel. */
void or_lda_sweep(int n_docs, int k, int v, const int* doc_off, const int* w,
                  const double* alpha, const double* beta, uint32_t seed,
                  uint32_t chain, uint32_t sweep, double* theta, double* phi, int* z) {
  int* ct = (int*)malloc(sizeof(int) * (size_t)n_docs * k);
  int* cp = (int*)malloc(sizeof(int) * (size_t)k * v);
  double* pmf = (double*)malloc(sizeof(double) * k);
  or_lda_counts(n_docs, k, v, doc_off, w, z, ct, cp);
  for (int d = 0; d < n_docs; d++)
    or_dirichlet_conj(seed, chain, sweep, OR_TAG_LDA_THETA, (uint32_t)(d * k), alpha,
                      ct + (size_t)d * k, k, theta + (size_t)d * k);
  for (int j = 0; j < k; j++)
    or_dirichlet_conj(seed, chain, sweep, OR_TAG_LDA_PHI, (uint32_t)(j * v), beta,
                      cp + (size_t)j * v, v, phi + (size_t)j * v);
  for (int d = 0; d < n_docs; d++)
    for (int t = doc_off[d]; t < doc_off[d + 1]; t++) {
      for (int j = 0; j < k; j++) pmf[j] = theta[(size_t)d * k + j] * phi[(size_t)j * v + w[t]];
      uint32_t r[4];
      or_draw4(seed, chain, (uint32_t)t, 0, sweep, OR_TAG_LDA_Z, r);
      z[t] = or_categorical_ref(pmf, k, (double)or_u01f(r[0]));
    }
  free(ct);
  free(cp);
  free(pmf);
}

/* ------------------------------------------------------------------------ */
/* HLR: y[n] ~ Bernoulli(sigmoid(dotprod(theta, x[n]) + b)).                 */
/* ------------------------------------------------------------------------ */

static double or_log_sigmoid(double t) {
  /* log(1 / (1 + exp(-t))) without cancellation */
  if (t >= 0.0) return -log1p(exp(-t));
  return t - log1p(exp(t));
}

/* Core/CgLike.hs:90-136 map-reduce of augur_bernoulli_ll(y, expit(t)).
 * stable = 0: the reference expression log(y s + (1 - y)(1 - s)).
 * stable = 1: the same quantity as log sigmoid(+-t). */
double or_hlr_data_ll(const or_hlr_model* m, const double* state, int stable) {
  const int d = m->d;
  const double b = m->has_intercept ? state[d] : 0.0;
  double acc = 0.0;
  for (int n = 0; n < m->n; n++) {
    const double* xr = m->x + (size_t)n * d;
    double t = b;
    for (int j = 0; j < d; j++) t += state[j] * xr[j];
    if (stable)
      acc += m->y[n] ? or_log_sigmoid(t) : or_log_sigmoid(-t);
    else
      acc += or_bernoulli_ll(m->y[n], or_expit(t));
  }
  return acc;
}

/* Core/CgGrad.hs:165-233: adj = bernoulli_dotarg1(y, s) * expit_grad_0(t)
 * (cbits/augur_math.c:56-59), adj_theta += adj * x[n], adj_b += adj.
 * ref_form = 0 uses the identical value y - s. */
void or_hlr_data_grad(const or_hlr_model* m, const double* state, double* grad_theta,
                      double* grad_b, int ref_form) {
  const int d = m->d;
  const double b = m->has_intercept ? state[d] : 0.0;
  for (int j = 0; j < d; j++) grad_theta[j] = 0.0;
  double gb = 0.0;
  for (int n = 0; n < m->n; n++) {
    const double* xr = m->x + (size_t)n * d;
    double t = b;
    for (int j = 0; j < d; j++) t += state[j] * xr[j];
    double s = or_expit(t);
    double a = ref_form ? or_bernoulli_dotarg1(m->y[n], s) * (s * (1.0 - s))
                        : (double)m->y[n] - s;
    for (int j = 0; j < d; j++) grad_theta[j] += a * xr[j];
    gb += a;
  }
  if (grad_b) *grad_b = gb;
}

static double or_hlr_prior_v_ll(const or_hlr_model* m, double v) {
  if (m->prior_kind == 0) return or_invgamma_ll(v, m->prior_a, m->prior_b);
  return or_exponential_ll(v, m->prior_a);
}
static double or_hlr_prior_v_dot(const or_hlr_model* m, double v) {
  if (m->prior_kind == 0) return or_invgamma_dotpt(v, m->prior_a, m->prior_b);
  return or_exponential_dotpt(v, m->prior_a);
}

/* ll of the block's full conditional (Core/CgLike.hs:254-258). */
double or_hlr_block_ll(const or_hlr_model* m, const double* state, int block) {
  const int d = m->d;
  const double v = state[d + 1];
  double acc = 0.0;
  if (block == 0) {
    acc += or_hlr_prior_v_ll(m, v);
    for (int j = 0; j < d; j++) acc += or_normal_ll(state[j], 0.0, v, m->ref_compat);
    if (m->has_intercept) acc += or_normal_ll(state[d], 0.0, v, m->ref_compat);
    return acc;
  }
  if (block == 2 || block == 3)
    for (int j = 0; j < d; j++) acc += or_normal_ll(state[j], 0.0, v, m->ref_compat);
  if ((block == 1 || block == 3) && m->has_intercept)
    acc += or_normal_ll(state[d], 0.0, v, m->ref_compat);
  return acc + or_hlr_data_ll(m, state, 1);
}

/* Gradient of the block's full conditional in block order: v | b | theta |
 * [theta..., b]. */
void or_hlr_block_grad(const or_hlr_model* m, const double* state, int block, double* g) {
  const int d = m->d;
  const double v = state[d + 1];
  if (block == 0) {
    double acc = or_hlr_prior_v_dot(m, v);
    for (int j = 0; j < d; j++) acc += or_normal_dotarg2(state[j], 0.0, v, m->ref_compat);
    if (m->has_intercept) acc += or_normal_dotarg2(state[d], 0.0, v, m->ref_compat);
    g[0] = acc;
    return;
  }
  double* gt = (double*)malloc(sizeof(double) * (size_t)d);
  double gb = 0.0;
  or_hlr_data_grad(m, state, gt, &gb, 0);
  if (block == 1) {
    g[0] = gb + or_normal_dotpt(state[d], 0.0, v);
  } else {
    for (int j = 0; j < d; j++) g[j] = gt[j] + or_normal_dotpt(state[j], 0.0, v);
    if (block == 3) g[d] = gb + or_normal_dotpt(state[d], 0.0, v);
  }
  free(gt);
}

static int or_hlr_block_dim(const or_hlr_model* m, int block) {
  if (block == 0 || block == 1) return 1;
  if (block == 2) return m->d;
  return m->d + 1;
}
static int or_hlr_block_coord(const or_hlr_model* m, int block, int i) {
  if (block == 0) return m->d + 1;
  if (block == 1) return m->d;
  return i; /* theta then b (index d) */
}

/* cbits/augur_mcmclib.h:389-424 (h_augur_mcmc_hmc) with the leapfrog of
 * :343-375 and the accept test of :70-83.  fixed_steps > 0 selects a fixed
 * step size eps and fixed_steps leapfrog steps instead of the reference's
 * jitter eps * U(0.8, 1.2), L = max(1, round(U(0.9, 1.1) simLen / eps'));
 * fixed_steps < 0 keeps the reference's step jitter with -fixed_steps steps.
 * out_info (7): eps', L, K0, K1, ll0, ll1, accepted. */
int or_hmc_step(const or_hlr_model* m, double* state, int block, int block_id,
                double sim_len, double eps, int fixed_steps, uint32_t seed,
                uint32_t chain, uint32_t sweep, double* out_info) {
  const int dim = or_hlr_block_dim(m, block);
  const int sd = m->d + 2;
  const uint32_t tag = OR_TAG_HMC_BASE + 4u * (uint32_t)block_id;
  double* p0 = (double*)malloc(sizeof(double) * dim);
  double* p = (double*)malloc(sizeof(double) * dim);
  double* g = (double*)malloc(sizeof(double) * dim);
  double* prop = (double*)malloc(sizeof(double) * sd);
  for (int i = 0; i < dim; i++) p0[i] = or_normal(seed, chain, (uint32_t)i, 0, sweep, tag + 0);
  double step;
  int steps;
  if (fixed_steps > 0) {
    step = eps;
    steps = fixed_steps;
  } else if (fixed_steps < 0) {
    uint32_t w[4];
    or_draw4(seed, chain, 0, 0, sweep, tag + 1, w);
    step = (0.8 + 0.4 * or_u01(w[0])) * eps;
    steps = -fixed_steps;
  } else {
    uint32_t w[4];
    or_draw4(seed, chain, 0, 0, sweep, tag + 1, w);
    step = (0.8 + 0.4 * or_u01(w[0])) * eps;
    double s = 0.9 + 0.2 * or_u01(w[1]);
    double r = round(s * sim_len / step);
    steps = r < 1.0 ? 1 : (int)r;
  }
  memcpy(prop, state, sizeof(double) * sd);
  or_hlr_block_grad(m, prop, block, g);
  for (int i = 0; i < dim; i++) p[i] = p0[i] + 0.5 * step * g[i];
  for (int l = 0; l < steps; l++) {
    for (int i = 0; i < dim; i++) prop[or_hlr_block_coord(m, block, i)] += step * p[i];
    if (l != steps - 1) {
      or_hlr_block_grad(m, prop, block, g);
      for (int i = 0; i < dim; i++) p[i] += step * g[i];
    }
  }
  or_hlr_block_grad(m, prop, block, g);
  for (int i = 0; i < dim; i++) p[i] += 0.5 * step * g[i];
  double k0 = 0.0, k1 = 0.0;
  for (int i = 0; i < dim; i++) {
    k0 += p0[i] * p0[i];
    k1 += p[i] * p[i];
  }
  k0 *= 0.5;
  k1 *= 0.5;
  double ll0 = or_hlr_block_ll(m, state, block);
  double ll1 = or_hlr_block_ll(m, prop, block);
  uint32_t w[4];
  or_draw4(seed, chain, 0, 0, sweep, tag + 2, w);
  int acc = log(or_u01(w[0])) < ll1 - ll0 + (k0 - k1);
  if (acc) memcpy(state, prop, sizeof(double) * sd);
  if (out_info) {
    out_info[0] = step;
    out_info[1] = steps;
    out_info[2] = k0;
    out_info[3] = k1;
    out_info[4] = ll0;
    out_info[5] = ll1;
    out_info[6] = acc;
  }
  free(p0);
  free(p);
  free(g);
  free(prop);
  return acc;
}

/* Metropolis-within-Gibbs on v (cbits/augur_mcmclib.h:178-193 with the
 * accept rule of :89-105): v' = v + sqrt(var) z, z from the block's tag + 0
 * stream, the uniform from tag + 2.  The Hastings term follows the generated
 * proposal of Core/CgUserProp.hs:130-141, which evaluates both proposal
 * densities at the current point's parameters: log N(v; v, var) -
 * log N(v'; v, var) = (v' - v)^2 / (2 var) with ref_compat; without it the
 * correct term of the symmetric random walk, 0.  out_info (4): v', ll0, ll1,
 * accepted. */
int or_mwg_v_step(const or_hlr_model* m, double* state, int block_id, double var, uint32_t seed,
                  uint32_t chain, uint32_t sweep, double* out_info) {
  const int sd = m->d + 2;
  const uint32_t tag = OR_TAG_HMC_BASE + 4u * (uint32_t)block_id;
  double* prop = (double*)malloc(sizeof(double) * sd);
  memcpy(prop, state, sizeof(double) * sd);
  const double v0 = state[m->d + 1];
  const double v1 = v0 + sqrt(var) * or_normal(seed, chain, 0, 0, sweep, tag + 0);
  prop[m->d + 1] = v1;
  const double extra = m->ref_compat ? 0.5 * (v1 - v0) * (v1 - v0) / var : 0.0;
  const double ll0 = or_hlr_block_ll(m, state, 0);
  const double ll1 = or_hlr_block_ll(m, prop, 0);
  uint32_t w[4];
  or_draw4(seed, chain, 0, 0, sweep, tag + 2, w);
  const int acc = log(or_u01(w[0])) < ll1 - ll0 + extra;
  if (acc) state[m->d + 1] = v1;
  if (out_info) {
    out_info[0] = v1;
    out_info[1] = ll0;
    out_info[2] = ll1;
    out_info[3] = acc;
  }
  free(prop);
  return acc;
}

/* MWG [b] ~ Normal(b, var) (compiler/augur/test/hlr2.infer): the same rule
 * as or_mwg_v_step on the intercept's full conditional (prior Normal(0, v)
 * plus the data term). */
int or_mwg_b_step(const or_hlr_model* m, double* state, int block_id, double var, uint32_t seed,
                  uint32_t chain, uint32_t sweep, double* out_info) {
  const int sd = m->d + 2;
  const uint32_t tag = OR_TAG_HMC_BASE + 4u * (uint32_t)block_id;
  double* prop = (double*)malloc(sizeof(double) * sd);
  memcpy(prop, state, sizeof(double) * sd);
  const double b0 = state[m->d];
  const double b1 = b0 + sqrt(var) * or_normal(seed, chain, 0, 0, sweep, tag + 0);
  prop[m->d] = b1;
  const double extra = m->ref_compat ? 0.5 * (b1 - b0) * (b1 - b0) / var : 0.0;
  const double ll0 = or_hlr_block_ll(m, state, 1);
  const double ll1 = or_hlr_block_ll(m, prop, 1);
  uint32_t w[4];
  or_draw4(seed, chain, 0, 0, sweep, tag + 2, w);
  const int acc = log(or_u01(w[0])) < ll1 - ll0 + extra;
  if (acc) state[m->d] = b1;
  if (out_info) {
    out_info[0] = b1;
    out_info[1] = ll0;
    out_info[2] = ll1;
    out_info[3] = acc;
  }
  free(prop);
  return acc;
}

/* ------------------------------------------------------------------------ */
/* GMM: hmvgmm.rv, schedule ConjGibbs pis, mu, cov, DiscGibbs z.             */
/* ------------------------------------------------------------------------ */

/* Core/CgDisc.hs:122-144 with y ~ MvNormal(mu[z], cov[z]), z ~ Cat(pis);
 * logits log pis_k + mvnormal_ll (cbits/augur_dist.c:548-568). */
void or_gmm_zstep(const or_gmm_model* m, const double* pis, const double* mu,
                  const double* cov, int* z, uint32_t seed, uint32_t chain, uint32_t sweep) {
  const int d = m->d, k = m->k;
  double* lg = (double*)malloc(sizeof(double) * k);
  for (int n = 0; n < m->n; n++) {
    for (int j = 0; j < k; j++)
      lg[j] = log(pis[j]) + or_mvnormal_ll(m->y + (size_t)n * d, mu + (size_t)j * d,
                                           cov + (size_t)j * d * d, d);
    uint32_t r[4];
    or_draw4(seed, chain, (uint32_t)n, 0, sweep, OR_TAG_GMM_Z, r);
    z[n] = or_categorical_logits(lg, k, (double)or_u01f(r[0]));
  }
  free(lg);
}

/* ---- non-conjugate moves of the GMM means ---- */

double or_gmm_comp_like(const or_gmm_model* m, const int* z, int k, const double* mu_k,
                        const double* cov_k) {
  double s = 0.0;
  for (int n = 0; n < m->n; n++)
    if (z[n] == k) s += or_mvnormal_ll(m->y + (size_t)n * m->d, mu_k, cov_k, m->d);
  return s;
}

static double or_gmm_comp_fc(const or_gmm_model* m, const int* z, int k, const double* mu_k,
                             const double* cov_k) {
  return or_mvnormal_ll(mu_k, m->mu0, m->cov0, m->d) + or_gmm_comp_like(m, z, k, mu_k, cov_k);
}

double or_gmm_mu_fc(const or_gmm_model* m, const int* z, const double* mu, const double* cov) {
  double s = 0.0;
  for (int k = 0; k < m->k; k++)
    s += or_gmm_comp_fc(m, z, k, mu + (size_t)k * m->d, cov + (size_t)k * m->d * m->d);
  return s;
}

/* d/dmu_k: -cov0^-1 (mu_k - mu0) + sum_{z_n = k} cov_k^-1 (y_n - mu_k)
 * (the gradients of augur_mvnormal_ll in its point and mean) */
void or_gmm_mu_grad(const or_gmm_model* m, const int* z, const double* mu, const double* cov,
                    double* g) {
  const int d = m->d;
  double p0[32 * 32], pk[32 * 32], r[32];
  memcpy(p0, m->cov0, sizeof(double) * d * d);
  or_mat_cholesky(p0, d);
  or_mat_cholesky_invert(p0, d);
  for (int k = 0; k < m->k; k++) {
    const double* mk = mu + (size_t)k * d;
    memcpy(pk, cov + (size_t)k * d * d, sizeof(double) * d * d);
    or_mat_cholesky(pk, d);
    or_mat_cholesky_invert(pk, d);
    for (int i = 0; i < d; i++) r[i] = 0.0;
    for (int n = 0; n < m->n; n++) {
      if (z[n] != k) continue;
      for (int i = 0; i < d; i++) r[i] += m->y[(size_t)n * d + i] - mk[i];
    }
    for (int i = 0; i < d; i++) {
      double a = 0.0, b = 0.0;
      for (int j = 0; j < d; j++) {
        a += pk[i * d + j] * r[j];
        b += p0[i * d + j] * (mk[j] - m->mu0[j]);
      }
      g[(size_t)k * d + i] = a - b;
    }
  }
}

/* cbits/augur_mcmclib.h:439-485 (augur_mcmc_eslice) for each component in
 * turn (Core/CgESlice.hs): prior MvNormal(mu0, cov0), objective the
 * component's likelihood. */
static int or_gmm_mu_eslice(const or_gmm_model* m, double* mu, const double* cov, const int* z,
                            uint32_t seed, uint32_t chain, uint32_t sweep) {
  const int d = m->d;
  const uint32_t tu = OR_TAG_GMM_MU_MOVE + 2;
  int acc = 0;
  double nu[32], prop[32];
  for (int k = 0; k < m->k; k++) {
    double* mk = mu + (size_t)k * d;
    const double* ck = cov + (size_t)k * d * d;
    or_mvnormal_sample(seed, chain, sweep, OR_TAG_GMM_MU_MOVE, (uint32_t)(k * d), m->mu0, m->cov0, d, nu);
    uint32_t w[4];
    or_draw4(seed, chain, (uint32_t)k, 1, sweep, tu, w);
    const double log_y = or_gmm_comp_like(m, z, k, mk, ck) + log(or_u01(w[0]));
    or_draw4(seed, chain, (uint32_t)k, 2, sweep, tu, w);
    double theta = 2.0 * OR_PI * or_u01(w[0]);
    double tmin = theta - 2.0 * OR_PI, tmax = theta;
    for (int j = 0; j < 200; j++) {
      if (tmax - tmin < 1e-8) break; /* keep the current point */
      for (int i = 0; i < d; i++)
        prop[i] = (mk[i] - m->mu0[i]) * cos(theta) + (nu[i] - m->mu0[i]) * sin(theta) + m->mu0[i];
      if (or_gmm_comp_like(m, z, k, prop, ck) > log_y) {
        memcpy(mk, prop, sizeof(double) * d);
        acc++;
        break;
      }
      if (theta < 0) tmin = theta;
      else tmax = theta;
      or_draw4(seed, chain, (uint32_t)k, (uint32_t)(3 + j), sweep, tu, w);
      theta = tmin + (tmax - tmin) * or_u01(w[0]);
    }
  }
  return acc;
}

/* cbits/augur_mcmclib.h:178-193 with the proposal of Core/CgUserProp.hs:
 * mu'_k ~ MvNormal(mu_k, P); llCorrect = ll(mu_k | mu_k, P) - ll(mu'_k | mu_k, P)
 * (ref_compat; 0 otherwise), accepted by :89-105 on the component's full
 * conditional. */
static int or_gmm_mu_mwg(const or_gmm_model* m, double* mu, const double* cov, const int* z,
                         uint32_t seed, uint32_t chain, uint32_t sweep) {
  const int d = m->d;
  int acc = 0;
  double prop[32];
  for (int k = 0; k < m->k; k++) {
    double* mk = mu + (size_t)k * d;
    const double* ck = cov + (size_t)k * d * d;
    or_mvnormal_sample(seed, chain, sweep, OR_TAG_GMM_MU_MOVE, (uint32_t)(k * d), mk, m->mu_prop_cov, d, prop);
    double corr = 0.0;
    if (m->ref_compat)
      corr = or_mvnormal_ll(mk, mk, m->mu_prop_cov, d) - or_mvnormal_ll(prop, mk, m->mu_prop_cov, d);
    uint32_t w[4];
    or_draw4(seed, chain, (uint32_t)k, 1, sweep, OR_TAG_GMM_MU_MOVE + 2, w);
    const double diff = or_gmm_comp_fc(m, z, k, prop, ck) - or_gmm_comp_fc(m, z, k, mk, ck);
    if (log(or_u01(w[0])) < diff + corr) {
      memcpy(mk, prop, sizeof(double) * d);
      acc++;
    }
  }
  return acc;
}

/* cbits/augur_mcmclib.h:389-424 over the block of all means (HMC [mu]). */
static int or_gmm_mu_hmc(const or_gmm_model* m, double* mu, const double* cov, const int* z,
                         uint32_t seed, uint32_t chain, uint32_t sweep) {
  const int dim = m->k * m->d;
  double* p0 = (double*)malloc(sizeof(double) * dim);
  double* p = (double*)malloc(sizeof(double) * dim);
  double* g = (double*)malloc(sizeof(double) * dim);
  double* x = (double*)malloc(sizeof(double) * dim);
  for (int i = 0; i < dim; i++) p0[i] = or_normal(seed, chain, (uint32_t)i, 0, sweep, OR_TAG_GMM_MU_MOVE);
  double step;
  int steps;
  uint32_t w[4];
  or_draw4(seed, chain, 0, 0, sweep, OR_TAG_GMM_MU_MOVE + 1, w);
  if (m->mu_fixed_steps > 0) {
    step = m->mu_step_size;
    steps = m->mu_fixed_steps;
  } else if (m->mu_fixed_steps < 0) {
    step = (0.8 + 0.4 * or_u01(w[0])) * m->mu_step_size;
    steps = -m->mu_fixed_steps;
  } else {
    step = (0.8 + 0.4 * or_u01(w[0])) * m->mu_step_size;
    double r = round((0.9 + 0.2 * or_u01(w[1])) * m->mu_sim_len / step);
    steps = r < 1.0 ? 1 : (int)r;
  }
  memcpy(x, mu, sizeof(double) * dim);
  or_gmm_mu_grad(m, z, x, cov, g);
  for (int i = 0; i < dim; i++) p[i] = p0[i] + 0.5 * step * g[i];
  for (int l = 0; l < steps; l++) {
    for (int i = 0; i < dim; i++) x[i] += step * p[i];
    if (l != steps - 1) {
      or_gmm_mu_grad(m, z, x, cov, g);
      for (int i = 0; i < dim; i++) p[i] += step * g[i];
    }
  }
  or_gmm_mu_grad(m, z, x, cov, g);
  for (int i = 0; i < dim; i++) p[i] += 0.5 * step * g[i];
  double k0 = 0.0, k1 = 0.0;
  for (int i = 0; i < dim; i++) {
    k0 += p0[i] * p0[i];
    k1 += p[i] * p[i];
  }
  const double ll0 = or_gmm_mu_fc(m, z, mu, cov), ll1 = or_gmm_mu_fc(m, z, x, cov);
  or_draw4(seed, chain, 0, 1, sweep, OR_TAG_GMM_MU_MOVE + 2, w);
  const int acc = log(or_u01(w[0])) < ll1 - ll0 + (0.5 * k0 - 0.5 * k1);
  if (acc) memcpy(mu, x, sizeof(double) * dim);
  free(p0);
  free(p);
  free(g);
  free(x);
  return acc;
}

/* cbits/augur_mcmclib.h:500-543 (h_augur_mcmc_refl_slice) over the block of
 * all means: slice level ll - Exp(1), max(1, round(simLen / eps)) steps of
 * eps, the momentum reflected on the gradient whenever a step leaves the
 * slice, the end point kept if it lies in the slice. */
static int or_gmm_mu_rslice(const or_gmm_model* m, double* mu, const double* cov, const int* z,
                            uint32_t seed, uint32_t chain, uint32_t sweep) {
  const int dim = m->k * m->d;
  double* p = (double*)malloc(sizeof(double) * dim);
  double* g = (double*)malloc(sizeof(double) * dim);
  double* x = (double*)malloc(sizeof(double) * dim);
  for (int i = 0; i < dim; i++) p[i] = or_normal(seed, chain, (uint32_t)i, 0, sweep, OR_TAG_GMM_MU_MOVE);
  uint32_t w[4];
  or_draw4(seed, chain, 0, 1, sweep, OR_TAG_GMM_MU_MOVE + 2, w);
  const double level = or_gmm_mu_fc(m, z, mu, cov) - (-log(or_u01(w[0])));
  const double r = round(m->mu_sim_len / m->mu_step_size);
  const int steps = r < 1.0 ? 1 : (int)r;
  memcpy(x, mu, sizeof(double) * dim);
  int cond = 0;
  for (int l = 0; l < steps; l++) {
    for (int i = 0; i < dim; i++) x[i] += m->mu_step_size * p[i];
    cond = or_gmm_mu_fc(m, z, x, cov) < level;
    if (cond) {
      or_gmm_mu_grad(m, z, x, cov, g);
      double a = 0.0, b = 0.0;
      for (int i = 0; i < dim; i++) {
        a += p[i] * g[i];
        b += g[i] * g[i];
      }
      for (int i = 0; i < dim; i++) p[i] += -2.0 * (a / b) * g[i];
    }
  }
  if (!cond) memcpy(mu, x, sizeof(double) * dim);
  free(p);
  free(g);
  free(x);
  return !cond;
}

/* cbits/augur_mcmclib.h:548-776 (h_augur_mcmc_nuts, h_augur_nuts_build_tree,
 * h_augur_nuts_update_s) over the block of all means, step size
 * mu_step_size, with its stacks and its Bernoulli draws in the order the
 * reference makes them.  Two properties of the reference's algorithm are
 * kept as they are: inside a subtree of depth >= 2 the u-turn test compares
 * the current end with the first point of the subtree's right half (the top
 * of its theta stack), and a proposal accepted in one doubling is replaced
 * by theta_minus / theta_plus at the next, so only the last doubling's
 * proposal can be kept.  The tree depth is capped at OR_NUTS_MAX_DEPTH
 * doublings (the reference has no cap). */
#define OR_NUTS_MAX_DEPTH 10
typedef struct {
  const or_gmm_model* m;
  const double* cov;
  const int* z;
  int dim;
  double step, logu;
  double *prop, *th, *mom, *g, *work;  /* stacks of (depth + 2) vectors; scratch */
  int ptop, ttop;
  uint32_t seed, chain, sweep, nbern;
} or_nuts_ctx;

static int or_nuts_bern(or_nuts_ctx* c, double bias) {
  uint32_t w[4];
  or_draw4(c->seed, c->chain, c->nbern++, 3, c->sweep, OR_TAG_GMM_MU_MOVE + 2, w);
  return or_u01(w[0]) < bias;
}

static double or_dot(const double* a, const double* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

static int or_nuts_update_s(or_nuts_ctx* c, int s, const double* tp, const double* mp, const double* tm,
                            const double* mm) {
  if (s == 1) {
    for (int i = 0; i < c->dim; i++) c->work[i] = tp[i] - tm[i];
    if (or_dot(c->work, mp, c->dim) >= 0.0 && or_dot(c->work, mm, c->dim) >= 0.0) return 1;
  }
  return 0;
}

static void or_nuts_build(or_nuts_ctx* c, double* theta, double* mom, int j, int minus, int* rn, int* rs) {
  const int dim = c->dim;
  if (j == 0) {
    /* cbits/augur_mcmclib.h:305-325 (leapfrog1) */
    or_gmm_mu_grad(c->m, c->z, theta, c->cov, c->g);
    for (int i = 0; i < dim; i++) mom[i] += c->step / 2.0 * c->g[i];
    for (int i = 0; i < dim; i++) theta[i] += c->step * mom[i];
    or_gmm_mu_grad(c->m, c->z, theta, c->cov, c->g);
    for (int i = 0; i < dim; i++) mom[i] += c->step / 2.0 * c->g[i];
    memcpy(c->prop + (size_t)(c->ptop++) * dim, theta, sizeof(double) * dim);
    const double ll = or_gmm_mu_fc(c->m, c->z, theta, c->cov) - 0.5 * or_dot(mom, mom, dim);
    *rn = c->logu <= ll;
    *rs = c->logu < 1000.0 + ll; /* MCMC.deltaMax */
    return;
  }
  int ln, ls;
  or_nuts_build(c, theta, mom, j - 1, minus, &ln, &ls);
  if (ls == 1) {
    if (j == 1) {
      memcpy(c->th + (size_t)c->ttop * dim, theta, sizeof(double) * dim);
      memcpy(c->mom + (size_t)c->ttop * dim, mom, sizeof(double) * dim);
      c->ttop++;
    }
    int r_n, r_s;
    or_nuts_build(c, theta, mom, j - 1, minus, &r_n, &r_s);
    const double bias = (double)r_n / (double)(ln + r_n);
    if (or_nuts_bern(c, bias)) /* keep the right subtree's proposal */
      memcpy(c->prop + (size_t)(c->ptop - 2) * dim, c->prop + (size_t)(c->ptop - 1) * dim, sizeof(double) * dim);
    c->ptop--;
    ln += r_n;
    const double* pk_t = c->th + (size_t)(c->ttop - 1) * dim;
    const double* pk_m = c->mom + (size_t)(c->ttop - 1) * dim;
    if (minus) ls = or_nuts_update_s(c, r_s, pk_t, pk_m, theta, mom);
    else ls = or_nuts_update_s(c, r_s, theta, mom, pk_t, pk_m);
    if (j > 1) c->ttop--;
  }
  *rn = ln;
  *rs = ls;
}

static int or_gmm_mu_nuts(const or_gmm_model* m, double* mu, const double* cov, const int* z,
                          uint32_t seed, uint32_t chain, uint32_t sweep) {
  const int dim = m->k * m->d, slots = OR_NUTS_MAX_DEPTH + 2;
  or_nuts_ctx c;
  c.m = m; c.cov = cov; c.z = z; c.dim = dim; c.seed = seed; c.chain = chain; c.sweep = sweep; c.nbern = 0;
  c.prop = (double*)malloc(sizeof(double) * dim * slots);
  c.th = (double*)malloc(sizeof(double) * dim * slots);
  c.mom = (double*)malloc(sizeof(double) * dim * slots);
  c.g = (double*)malloc(sizeof(double) * dim);
  c.work = (double*)malloc(sizeof(double) * dim);
  double* tm = (double*)malloc(sizeof(double) * dim);
  double* tp = (double*)malloc(sizeof(double) * dim);
  double* mm = (double*)malloc(sizeof(double) * dim);
  double* mp = (double*)malloc(sizeof(double) * dim);
  double* th = (double*)malloc(sizeof(double) * dim);
  memcpy(tm, mu, sizeof(double) * dim);
  memcpy(tp, mu, sizeof(double) * dim);
  for (int i = 0; i < dim; i++) mm[i] = or_normal(seed, chain, (uint32_t)i, 0, sweep, OR_TAG_GMM_MU_MOVE);
  memcpy(mp, mm, sizeof(double) * dim);
  uint32_t w[4];
  or_draw4(seed, chain, 0, 1, sweep, OR_TAG_GMM_MU_MOVE + 2, w);
  c.logu = or_gmm_mu_fc(m, z, mu, cov) - 0.5 * or_dot(mm, mm, dim) - (-log(or_u01(w[0])));
  c.ptop = 0;
  int n = 1, j = 0, f_cpy = 0;
  while (1) {
    int rn, rs;
    c.ttop = 0;
    const int minus = or_nuts_bern(&c, 0.5);
    if (minus) {
      c.step = -m->mu_step_size;
      memcpy(th, tm, sizeof(double) * dim);
      or_nuts_build(&c, th, mm, j, 1, &rn, &rs);
      memcpy(tm, th, sizeof(double) * dim);
    } else {
      c.step = m->mu_step_size;
      memcpy(th, tp, sizeof(double) * dim);
      or_nuts_build(&c, th, mp, j, 0, &rn, &rs);
      memcpy(tp, th, sizeof(double) * dim);
    }
    const double* cand = c.prop + (size_t)(--c.ptop) * dim;
    f_cpy = 0;
    if (rs == 1) {
      const double bias = fmin(1.0, (double)rn / (double)n);
      if (or_nuts_bern(&c, bias)) {
        memcpy(th, cand, sizeof(double) * dim);
        f_cpy = 1;
      }
    }
    n += rn;
    const int s = or_nuts_update_s(&c, rs, tp, mp, tm, mm);
    j += 1;
    if (s != 1 || j >= OR_NUTS_MAX_DEPTH) break;
  }
  if (f_cpy) memcpy(mu, th, sizeof(double) * dim);
  free(c.prop); free(c.th); free(c.mom); free(c.g); free(c.work);
  free(tm); free(tp); free(mm); free(mp); free(th);
  return f_cpy;
}

int or_gmm_mu_move(const or_gmm_model* m, double* mu, const double* cov, const int* z,
                   uint32_t seed, uint32_t chain, uint32_t sweep) {
  switch (m->mu_kernel) {
    case 5: return or_gmm_mu_nuts(m, mu, cov, z, seed, chain, sweep);
    case 1: return or_gmm_mu_eslice(m, mu, cov, z, seed, chain, sweep);
    case 2: return or_gmm_mu_hmc(m, mu, cov, z, seed, chain, sweep);
    case 3: return or_gmm_mu_mwg(m, mu, cov, z, seed, chain, sweep);
    case 4: return or_gmm_mu_rslice(m, mu, cov, z, seed, chain, sweep);
    default: return 0;
  }
}

void or_gmm_sweep(const or_gmm_model* m, double* pis, double* mu, double* cov, int* z,
                  uint32_t seed, uint32_t chain, uint32_t sweep) {
  const int d = m->d, k = m->k, n = m->n;
  int* cnt = (int*)calloc((size_t)k, sizeof(int));
  double* sum = (double*)calloc((size_t)k * d, sizeof(double));
  double* obs = (double*)malloc(sizeof(double) * (size_t)d * d);
  for (int i = 0; i < n; i++) {
    cnt[z[i]]++;
    for (int j = 0; j < d; j++) sum[(size_t)z[i] * d + j] += m->y[(size_t)i * d + j];
  }
  /* ConjGibbs [pis]: Core/CgConj.hs:171-183 -> augur_dirichlet_categorical_conj */
  if (m->sample_pi) or_dirichlet_conj(seed, chain, sweep, OR_TAG_GMM_PI, 0, m->alpha, cnt, k, pis);
  /* ConjGibbs [mu]: Core/CgConj.hs:186-206 -> augur_mvnormal_mvnormal_conj,
   * or the schedule's other kernel for mu */
  if (m->mu_kernel == 0) {
    for (int j = 0; j < k; j++)
      or_mvnormal_mvnormal_conj(seed, chain, sweep, (uint32_t)(j * d), m->mu0, m->cov0,
                                sum + (size_t)j * d, cnt[j], cov + (size_t)j * d * d, d,
                                mu + (size_t)j * d);
  } else {
    or_gmm_mu_move(m, mu, cov, z, seed, chain, sweep);
  }
  /* ConjGibbs [cov]: Core/CgConj.hs:209-246 -> augur_invwishart_mvnormal_conj
   * (cbits/augur_dist.c:751-757): IW(df + cnt, scale0 + sum (y - mu)(y - mu)^T) */
  if (m->sample_cov) {
    for (int j = 0; j < k; j++) {
      for (int a = 0; a < d * d; a++) obs[a] = 0.0;
      for (int i = 0; i < n; i++) {
        if (z[i] != j) continue;
        double diff[32];
        for (int a = 0; a < d; a++) diff[a] = m->y[(size_t)i * d + a] - mu[(size_t)j * d + a];
        for (int a = 0; a < d; a++)
          for (int b = 0; b < d; b++) obs[a * d + b] += diff[a] * diff[b];
      }
      for (int a = 0; a < d * d; a++) obs[a] += m->scale0[a];
      or_invwishart_sample(seed, chain, sweep, (uint32_t)(j * d * d), m->df + cnt[j], obs, d,
                           cov + (size_t)j * d * d);
    }
  }
  /* DiscGibbs [z] */
  or_gmm_zstep(m, pis, mu, cov, z, seed, chain, sweep);
  free(cnt);
  free(sum);
  free(obs);
}
