/*
 * augur_oracle.h -- CPU FP64 restatement of the AugurV2 sampler core.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * sampler runtime: tests, the smoke check and the benchmark's CPU-baseline
 * leg may load it; the product path never does.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Random numbers follow the Philox stream contract shared with the HIP
 * kernels (see the "RNG contract" block below), so that the oracle and the
 * GPU consume identical uniforms at every sampling site.
 */
#ifndef AUGUR_ORACLE_H
#define AUGUR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * RNG contract (identical in the HIP kernels)
 *
 *   key     = { seed (low 32 bits), chain index }
 *   counter = { site, sub, sweep, tag }
 *   one Philox4x32-10 call -> four 32-bit words w0..w3
 *
 *   uniform (double) u = (w + 0.5) * 2^-32           in (0, 1)
 *   uniform (float)  u = (w >> 8)  * 2^-24           in [0, 1)
 *   normal           n = sqrt(-2 ln u(w0)) * cos(2 pi u(w1))
 *   gamma attempt j at a site uses counter sub = j:
 *                    z = normal(w0, w1), acceptance u = u(w2),
 *                    boost for shape < 1 uses u(w3) of the accepted attempt.
 * ------------------------------------------------------------------------- */
enum {
  OR_TAG_LDA_THETA = 1,
  OR_TAG_LDA_PHI = 2,
  OR_TAG_LDA_Z = 3,
  OR_TAG_GMM_PI = 4,
  OR_TAG_GMM_MU = 5,
  OR_TAG_GMM_COV = 6,
  OR_TAG_GMM_Z = 7,
  /* non-conjugate moves of the GMM means: 8 normals, 9 HMC jitter, 10 uniforms */
  OR_TAG_GMM_MU_MOVE = 8,
  /* HMC block b uses tags 16 + 4 b + {0 momentum, 1 jitter, 2 accept} */
  OR_TAG_HMC_BASE = 16
};

#define OR_GAMMA_MAX_ATTEMPTS 256
#define OR_LARGE_POS_NUM 1e8 /* cbits/augur_math.h:24 */

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_draw4(uint32_t seed, uint32_t chain, uint32_t site, uint32_t sub,
              uint32_t sweep, uint32_t tag, uint32_t out[4]);
double or_u01(uint32_t w);
float or_u01f(uint32_t w);
double or_normal_from(uint32_t w0, uint32_t w1);
double or_normal(uint32_t seed, uint32_t chain, uint32_t site, uint32_t sub,
                 uint32_t sweep, uint32_t tag);
double or_gamma(uint32_t seed, uint32_t chain, uint32_t site, uint32_t sweep,
                uint32_t tag, double shape);

/* ---- distributions (cbits/augur_dist.c, cbits/augur_math.c) ---- */
double or_expit(double x);
double or_bernoulli_ll(int y, double bias);
double or_bernoulli_dotarg1(int y, double bias);
double or_normal_ll(double pt, double mean, double var, int ref_compat);
double or_normal_dotpt(double pt, double mean, double var);
double or_normal_dotarg2(double pt, double mean, double var, int ref_compat);
double or_invgamma_ll(double pt, double shape, double scale);
double or_invgamma_dotpt(double pt, double shape, double scale);
double or_exponential_ll(double pt, double rate);
double or_exponential_dotpt(double pt, double rate);

int or_categorical_ref(const double* pmf, int k, double u);
int or_categorical_f32(const float* a, const float* b, int k, float u);
int or_categorical_logits(const double* logits, int k, double u);

void or_dirichlet_conj(uint32_t seed, uint32_t chain, uint32_t sweep, uint32_t tag,
                       uint32_t site0, const double* alpha, const int* cnts,
                       int k, double* out);

/* ---- small FP64 linear algebra (cbits/augur_matop.c:312-447) ---- */
void or_mat_cholesky(double* a, int n);
void or_mat_lower_trmi(double* l, int n);
void or_mat_cholesky_invert(double* llt, int n);
void or_mat_trtrmm(const double* l1, double* l2, int n);
void or_mat_trvm(const double* l, double* v, int n);
double or_mvnormal_ll(const double* pt, const double* mean, const double* cov, int n);
void or_mvnormal_mvnormal_conj(uint32_t seed, uint32_t chain, uint32_t sweep, uint32_t site0,
                               const double* mu0, const double* cov0,
                               const double* sum, int cnt, const double* cov,
                               int n, double* out);
void or_invwishart_sample(uint32_t seed, uint32_t chain, uint32_t sweep, uint32_t site0,
                          int df, const double* scale, int n, double* out);

/* ---- LDA (lda.rv; sweep ConjGibbs theta, ConjGibbs phi, DiscGibbs z) ---- */
void or_lda_counts(int n_docs, int k, int v, const int* doc_off, const int* w,
                   const int* z, int* cnt_theta, int* cnt_phi);
void or_lda_zstep_f32(int n_docs, int k, const int* doc_off, const int* w,
                      const float* theta, const float* phi_t, uint32_t seed,
                      uint32_t chain, uint32_t sweep, int* z);
void or_lda_sweep(int n_docs, int k, int v, const int* doc_off, const int* w,
                  const double* alpha, const double* beta, uint32_t seed,
                  uint32_t chain, uint32_t sweep, double* theta, double* phi, int* z);

/* ---- HLR (hlr.rv family; sigmoid(dotprod(theta, x[n]) + b)) ---- */
typedef struct {
  int n, d;            /* rows, features */
  const double* x;     /* n x d row-major */
  const int* y;        /* n */
  int prior_kind;      /* 0: InvGamma(a, b), 1: Exponential(a) */
  double prior_a, prior_b;
  int has_intercept;
  int ref_compat;      /* 1: reference normal_ll / normal_dotarg2 forms */
} or_hlr_model;

/* state layout: [theta_0..theta_{d-1}, b, v] */
double or_hlr_data_ll(const or_hlr_model* m, const double* state, int stable);
void or_hlr_data_grad(const or_hlr_model* m, const double* state, double* grad_theta,
                      double* grad_b, int ref_form);
double or_hlr_block_ll(const or_hlr_model* m, const double* state, int block);
void or_hlr_block_grad(const or_hlr_model* m, const double* state, int block, double* g);
/* block: 0 = v, 1 = b, 2 = theta, 3 = [b, theta] */
int or_hmc_step(const or_hlr_model* m, double* state, int block, int block_id,
                double sim_len, double eps, int fixed_steps, uint32_t seed,
                uint32_t chain, uint32_t sweep, double* out_info);
/* Metropolis-within-Gibbs on v, proposal Normal(v, var); out_info (4). */
int or_mwg_v_step(const or_hlr_model* m, double* state, int block_id, double var, uint32_t seed,
                  uint32_t chain, uint32_t sweep, double* out_info);
/* Metropolis-within-Gibbs on the intercept b, proposal Normal(b, var). */
int or_mwg_b_step(const or_hlr_model* m, double* state, int block_id, double var, uint32_t seed,
                  uint32_t chain, uint32_t sweep, double* out_info);

/* ---- GMM (hmvgmm.rv / mvgmm.rv) ---- */
typedef struct {
  int n, d, k;
  const double* y;      /* n x d */
  const double* alpha;  /* k */
  const double* mu0;    /* d */
  const double* cov0;   /* d x d */
  int df;
  const double* scale0; /* d x d */
  int sample_pi, sample_cov;
  /* the schedule's kernel for mu: 0 ConjGibbs, 1 ESlice, 2 HMC, 3 MWG, 4 RSlice, 5 NUTS */
  int mu_kernel;
  double mu_sim_len, mu_step_size;
  int mu_fixed_steps;         /* HMC: > 0 fixed, < 0 step jitter only, 0 reference jitter */
  const double* mu_prop_cov;  /* MWG proposal covariance (d x d) */
  int ref_compat;             /* MWG: the reference's Hastings term */
} or_gmm_model;

/* Log-likelihood of component k's mean: the data terms of the points with
 * z_n = k (the reference's single-index likelihood, Core/CgLike.hs
 * runLLFnOne), each by or_mvnormal_ll (cbits/augur_dist.c:548-568). */
double or_gmm_comp_like(const or_gmm_model* m, const int* z, int k, const double* mu_k,
                        const double* cov_k);
/* Log full conditional of the mean block: priors MvNormal(mu0, cov0) of
 * every component plus every data term; and its gradient. */
double or_gmm_mu_fc(const or_gmm_model* m, const int* z, const double* mu, const double* cov);
void or_gmm_mu_grad(const or_gmm_model* m, const int* z, const double* mu, const double* cov,
                    double* g);
/* The model's mu_kernel move of all means (chain, sweep streams); returns
 * the number of accepted moves (per component for ESlice / MWG). */
int or_gmm_mu_move(const or_gmm_model* m, double* mu, const double* cov, const int* z,
                   uint32_t seed, uint32_t chain, uint32_t sweep);

void or_gmm_sweep(const or_gmm_model* m, double* pis, double* mu, double* cov, int* z,
                  uint32_t seed, uint32_t chain, uint32_t sweep);
void or_gmm_zstep(const or_gmm_model* m, const double* pis, const double* mu,
                  const double* cov, int* z, uint32_t seed, uint32_t chain, uint32_t sweep);
double or_multi_lgamma(int p, double x);
double or_dirichlet_ll(const double* pt, const double* alpha, int k);
double or_invwishart_ll(const double* pt, int df, const double* scale, int d);
double or_gmm_ll_mod(const or_gmm_model* m, const double* pis, const double* mu, const double* cov,
                     const int* z);

#ifdef __cplusplus
}
#endif
#endif
