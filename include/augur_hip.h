/*
 * augur_hip.h -- C-ABI of the MI355X (gfx950) sampler runtime.
 *
 * This library replaces, for the LDA, hierarchical-logistic-regression and
 * Gaussian-mixture families, the runtime-generated sampler library that
 * AugurV2 builds and drives through ctypes, together with the cbits GPU
 * library it links (augur_blkop, augur_vecop, augur_matop, augur_dist).
 *
 * Conventions
 *   - Every function returns an augur_status; on failure a message is
 *     available from augur_last_error() (per calling thread).
 *   - Host pointers only: the caller owns every buffer passed in or out;
 *     the library owns all device memory of a sampler.
 *   - Handles are independent: several samplers may live in one process.
 *     Calls on one handle must not run concurrently; different handles may.
 *   - Parameter names are the family's canonical names:
 *       LDA : "theta" (D x K f64), "phi" (K x V f64), "z" (T i32)
 *       HLR : "theta" (D f64), "b" (1 f64), "v" (1 f64), "state" (D + 2 f64),
 *             "bias" (N f64, read-only: the deterministic sigmoid(x[n].theta + b)
 *             of compiler/augur/test/dethlr.rv, computed when read)
 *       GMM : "pis" (K f64), "mu" (K x D f64), "cov" (K x D x D f64), "z" (N i32)
 *     `chain` selects one chain's copy.
 *   - Random numbers: Philox4x32-10 keyed by (seed, chain); see augur_rng.h.
 */
#ifndef AUGUR_HIP_H
#define AUGUR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int augur_status;
#define AUGUR_OK 0
#define AUGUR_ERR_ARG 1         /* invalid argument or shape */
#define AUGUR_ERR_HIP 2         /* HIP runtime error (no device, launch failure, ...) */
#define AUGUR_ERR_STATE 3       /* call not valid in the sampler's current state */
#define AUGUR_ERR_UNSUPPORTED 4 /* model / schedule outside the supported families */

typedef struct augur_sampler augur_sampler;

/* Message of the last failed call on this thread ("" if none). */
const char* augur_last_error(void);
/* Library version string. */
const char* augur_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
augur_status augur_device_count(int* n);

/* ------------------------------------------------------------------------
 * Sampler construction.  Each replaces the generated augur_iface_init for
 * its family (compiler/augur/src/CudaC/CgStInit.hs:474-510, called from
 * pyaugur/pyaugur/augurlib.py:445-475): hyper-parameters and data are
 * copied to device memory; parameters start at zero until augur_set_pt.
 * ---------------------------------------------------------------------- */

/* lda.rv: theta[d] ~ Dirichlet(alpha); phi[k] ~ Dirichlet(beta);
 * z[d, n] ~ Categorical(theta[d]); w[d, n] ~ Categorical(phi[z[d, n]]).
 * Schedule ConjGibbs [theta] (*) ConjGibbs [phi] (*) DiscGibbs [z]. */
typedef struct {
  int n_topics;          /* K */
  int n_docs;            /* D */
  int n_vocab;           /* V */
  int chains;
  uint32_t seed;
  const int32_t* doc_off; /* D + 1 offsets into w (CSR of the ragged w[d, n]) */
  const int32_t* w;       /* T = doc_off[D] word ids in [0, V) */
  const double* alpha;    /* K */
  const double* beta;     /* V */
} augur_lda_desc;
augur_status augur_lda_create(const augur_lda_desc* desc, augur_sampler** out);

/* hlr.rv family: v ~ InvGamma(a, b) | Exponential(a); [b ~ Normal(0, v)];
 * theta[k] ~ Normal(0, v); y[n] ~ Bernoulli(sigmoid(dotprod(theta, x[n]) [+ b])).
 * A schedule of up to 4 HMC blocks run in order each sweep. */
#define AUGUR_HMC_BLOCK_V 0
#define AUGUR_HMC_BLOCK_B 1
#define AUGUR_HMC_BLOCK_THETA 2
#define AUGUR_HMC_BLOCK_B_THETA 3
/* Metropolis-within-Gibbs on v with the proposal v' ~ Normal(v, step_size)
 * (step_size is the proposal variance): the schedule item
 * `MWG [v] ~ Normal(v, s)` (cbits/augur_mcmclib.h:178-193,
 * Core/CgUserProp.hs:130-141); sim_len and fixed_steps are ignored. */
#define AUGUR_MWG_BLOCK_V 4
/* The same on the intercept b, `MWG [b] ~ Normal(b, s)`
 * (compiler/augur/test/hlr2.infer): one data pass per sweep at the proposal,
 * accepted on b's full conditional. */
#define AUGUR_MWG_BLOCK_B 5
typedef struct {
  int block;        /* AUGUR_HMC_BLOCK_* or AUGUR_MWG_BLOCK_V */
  double sim_len;   /* reference simLen (Core/KernSyn.hs:64-75 default 1.0) */
  double step_size; /* reference eps (default 0.05) */
  int fixed_steps;  /* > 0: fixed step size and this many leapfrog steps;
                       0: reference jitter (cbits/augur_mcmclib.h:401-405) */
} augur_hmc_block;
typedef struct {
  int n_rows;       /* N */
  int n_features;   /* D (K in hlr.rv), 1..256 */
  int chains;       /* 1..16 */
  uint32_t seed;
  const float* x;   /* N x D row-major FP32 */
  const int32_t* y; /* N labels in {0, 1} */
  int prior_kind;   /* 0: v ~ InvGamma(prior_a, prior_b); 1: v ~ Exponential(prior_a) */
  double prior_a, prior_b;
  int has_intercept;
  int ref_compat;   /* 1: reference normal_ll (-log var) and normal_dotarg2 forms */
  int n_blocks;
  augur_hmc_block blocks[4];
} augur_hlr_desc;
augur_status augur_hlr_create(const augur_hlr_desc* desc, augur_sampler** out);

/* No counterpart in the reference.  Surrogate-gradient HMC for one
 * HMC [theta] block of an HLR sampler without intercept:
 *   1. the first train_sweeps sweeps run exact HMC and record every leapfrog
 *      point with its data gradient (up to max_pairs rows);
 *   2. the data gradient is then fitted on the device: a least-squares
 *      linear map of theta (FP64 normal equations) plus an MLP
 *      d -> hidden -> hidden -> d (GELU) trained on the linear map's residual
 *      (Adam, epochs passes of minibatches of `batch`, rate lr);
 *   3. later sweeps run leapfrog trajectories with the fitted data gradient
 *      (BF16 network weights, FP32 arithmetic) plus the exact prior
 *      gradient: the block's own jittered schedule when steps = 0, else
 *      `steps` steps of step_size;
 *   4. the proposal is accepted or rejected with the exact log-density (FP64
 *      row arithmetic), so the target is unchanged;
 *   5. if the acceptance rate over a window of `window` sweeps falls below
 *      min_accept, the block returns to exact HMC for good. */
typedef struct {
  int block_index;   /* index in augur_hlr_desc.blocks of the HMC [theta] block */
  int hidden;        /* 1..256 */
  int train_sweeps;
  int max_pairs;
  int epochs;
  int batch;
  double lr;
  int steps;
  double step_size;
  double min_accept;
  int window;
  uint32_t seed;     /* network initialisation and minibatch draws */
  int linear_only;   /* 1: fit the least-squares linear term alone (no network,
                        epochs ignored): the ablation of the network */
} augur_surrogate_cfg;
augur_status augur_hlr_surrogate(augur_sampler* s, const augur_surrogate_cfg* cfg);
/* info[8] (info[9] when bytes allow): phase (0 off, 1 collecting,
 * 2 surrogate, 3 fell back), pairs collected, final training loss
 * (standardised MSE of the residual), training time in ms, fallbacks,
 * surrogate sweeps run, last window's acceptance, collection sweeps run,
 * [the linear map's residual variance as a fraction of the gradient's]. */
augur_status augur_hlr_surrogate_info(augur_sampler* s, double* info, size_t bytes);
/* The surrogate network's loss and gradient on host data (rows x d inputs
 * and targets, parameters in the order W1, b1, W2, b2, W3, b3), and a run of
 * Adam steps over minibatches drawn from n pairs: the kernels the sampler
 * trains with, exposed for checks against the CPU restatement. */
augur_status augur_op_mlp_loss_grad(int d, int h, int rows, const float* params, const float* x,
                                    const float* t, float* out, double* loss, float* grads);
augur_status augur_op_mlp_train(int d, int h, int n, const float* x, const float* t, int steps,
                                int batch, double lr, uint32_t seed, float* params_inout,
                                double* loss_hist);

/* Batched form of the hlr.rv family without intercept: n_datasets
 * independent regressions of uneven N_i and D_i in {16, 32, 64, 128}, `chains`
 * chains each, schedule HMC [v] (*) HMC [theta] with a fixed leapfrog count
 * for theta.  X is stored BF16 on the device (FP32 arithmetic, FP64
 * accumulation and Metropolis test).  No counterpart in the reference, which
 * runs one model per loaded library (cbits/augur_mcmclib.h:44); each dataset
 * follows the single-model semantics of augur_hlr_create.  Parameters:
 * "state" ([chain][sum_i (D_i + 2)] f64, dataset i at offset sum_{i'<i}
 * (D_i' + 2), each slice [theta, b = 0, v]) and "state:<i>".  Philox chain
 * key of (dataset i, chain c) is i * chains + c.  rows_per_block = 0 lets the
 * autotuner time candidate block sizes at creation; use_graph captures one
 * sweep in a hipGraph and replays it. */
typedef struct {
  int n_datasets;
  const int32_t* n_rows;            /* per dataset */
  const int32_t* n_features;        /* per dataset */
  const float* const* x;            /* per dataset, N_i x D_i FP32 (rounded to BF16) */
  const int32_t* const* y;          /* per dataset, N_i labels in {0, 1} */
  int chains;                       /* 1..16 */
  uint32_t seed;
  int prior_kind;                   /* as augur_hlr_desc */
  double prior_a, prior_b;
  int ref_compat;
  augur_hmc_block v_block;          /* HMC [v] */
  augur_hmc_block theta_block;      /* HMC [theta]: fixed_steps >= 1 required */
  int rows_per_block;               /* 0: autotune; else rounded up to a multiple of 256 */
  int use_graph;
} augur_bhlr_desc;
augur_status augur_bhlr_create(const augur_bhlr_desc* desc, augur_sampler** out);
/* The rows-per-block the batched sampler runs with, and the autotuner's
 * timings of its candidates (microseconds per data pass, up to `cap`). */
augur_status augur_bhlr_tuning(augur_sampler* s, int* rows_per_block, int* n_candidates,
                               int* cand_rows, float* cand_us, int cap);
/* No counterpart in the reference, which does not adapt.  Warmup of n_sweeps
 * sweeps; after each, every dataset's theta step size follows dual averaging
 * towards a mean acceptance probability target_accept over its chains
 * (Hoffman & Gelman 2014, Algorithm 5), and with an MWG v block its proposal
 * variance follows a Robbins-Monro rule towards acceptance 0.44.  At the end
 * the step sizes take their averaged values; later sweeps keep both fixed. */
augur_status augur_bhlr_adapt(augur_sampler* s, int n_sweeps, double target_accept);
/* Per-dataset theta step sizes and v proposal variances (n >= n_datasets;
 * either pointer may be null). */
augur_status augur_bhlr_step_sizes(augur_sampler* s, double* eps, double* v_var, int n);

/* No counterpart in the reference.  Linear gradient surrogate for every
 * dataset of a batched sampler (the single-dataset sampler's least-squares
 * term, without the network):
 *   1. the next train_sweeps sweeps (ordinary or warmup) run exact HMC and
 *      keep every leapfrog point with its data gradient (FP32) on the device;
 *   2. then each dataset's data gradient is fitted on the device as
 *      g(theta) ~ gbar + W (theta - m), m and gbar the pairs' means, W by
 *      least squares on standardised theta with `ridge` added to the
 *      correlation matrix's diagonal (FP64 normal equations, Cholesky);
 *   3. later sweeps move each dataset's theta by a leapfrog trajectory with
 *      that gradient plus the exact prior gradient (its own step size,
 *      jitter and count), then run one exact data pass at the proposal for
 *      the unchanged FP64 Metropolis test, so the target is unchanged;
 *   4. a dataset whose acceptance rate over a window of `window` sweeps
 *      (0: never checked) falls below min_accept returns to exact HMC for
 *      good.
 * The pairs take train_sweeps x steps x chains x sum (D_i + 1) x 8 bytes
 * (at most 32 GiB) and stay for augur_bhlr_surrogate_pairs. */
typedef struct {
  int train_sweeps;
  double ridge;       /* e.g. 1e-6 */
  double min_accept;
  int window;
} augur_bsurr_cfg;
augur_status augur_bhlr_surrogate(augur_sampler* s, const augur_bsurr_cfg* cfg);
/* info[7]: phase (0 off, 1 collecting, 2 surrogate), datasets on the
 * surrogate, datasets that fell back, fit time in ms, pairs per dataset,
 * surrogate sweeps run, collection sweeps run. */
augur_status augur_bhlr_surrogate_info(augur_sampler* s, double* info, size_t bytes);
/* Per dataset: 1 on the surrogate, 0 exact (n >= n_datasets). */
augur_status augur_bhlr_surrogate_modes(augur_sampler* s, int32_t* modes, int n);
/* Puts dataset i on exact HMC (mode 0) or back on its fitted surrogate
 * (mode 1), e.g. to keep a dataset whose posterior the linear map cannot
 * follow on exact trajectories.  Needs a fitted surrogate.  Sweeps after it
 * run the exact datasets' leapfrog passes over their row blocks only. */
augur_status augur_bhlr_surrogate_set_mode(augur_sampler* s, int dataset, int mode);
/* Dataset i's fitted map: W (D_i x D_i, row o gives output coordinate o; the
 * FP32 values the trajectories use), m and gbar (D_i each); any pointer may
 * be null. */
augur_status augur_bhlr_surrogate_map(augur_sampler* s, int dataset, double* w, double* mean_theta,
                                      double* mean_grad);
/* Dataset i's fitting pairs, row q = slot * chains + chain with slot =
 * collection sweep * steps + leapfrog pass: theta as evaluated (FP32) and
 * its data gradient (FP32), D_i values per row.  *n_pairs receives the
 * count; with both pointers null only the count is returned. */
augur_status augur_bhlr_surrogate_pairs(augur_sampler* s, int dataset, float* theta, float* grad, int cap,
                                        int* n_pairs);

/* hmvgmm.rv: pis ~ Dirichlet(alpha); mu[k] ~ MvNormal(mu0, cov0);
 * cov[k] ~ InvWishart(df, scale0); z[n] ~ Categorical(pis);
 * y[n] ~ MvNormal(mu[z[n]], cov[z[n]]).  sample_pi / sample_cov = 0 keep
 * pis / cov fixed at their set_pt values (mvgmm.rv).  mu_kernel names the
 * schedule's kernel for mu (compiler/augur/test/hmvgmm1-3.infer,
 * mvgmm1-5.infer); the others stay ConjGibbs [pis], ConjGibbs [cov] and
 * DiscGibbs [z], in the order pis, mu, cov, z. */
#define AUGUR_GMM_MU_CONJ 0   /* ConjGibbs [mu] (Core/CgConj.hs:186-206) */
#define AUGUR_GMM_MU_ESLICE 1 /* ESlice [mu], per component (cbits/augur_mcmclib.h:439-485) */
#define AUGUR_GMM_MU_HMC 2    /* HMC [mu] [sim_len, step_size], all components in one block
                                 (cbits/augur_mcmclib.h:389-424) */
#define AUGUR_GMM_MU_MWG 3    /* MWG [mu] ~ MvNormal(mu[k], mu_prop_cov) for k, per component
                                 (cbits/augur_mcmclib.h:178-193, Core/CgUserProp.hs) */
#define AUGUR_GMM_MU_RSLICE 4 /* RSlice [mu] [sim_len, step_size], one block
                                 (cbits/augur_mcmclib.h:500-543) */
#define AUGUR_GMM_MU_NUTS 5   /* NUTS [mu] [step_size], one block (cbits/augur_mcmclib.h:548-776),
                                 at most 10 doublings */
typedef struct {
  int n_points;      /* N */
  int dim;           /* D, 1..16 */
  int n_components;  /* K, 1..64 */
  int chains;
  uint32_t seed;
  const float* y;        /* N x D FP32 */
  const double* alpha;   /* K */
  const double* mu0;     /* D */
  const double* cov0;    /* D x D */
  int df;
  const double* scale0;  /* D x D */
  int sample_pi;
  int sample_cov;
  int mu_kernel;             /* AUGUR_GMM_MU_* */
  double mu_sim_len;         /* HMC / RSlice (Core/KernSyn.hs defaults 1.0, 0.05; NUTS 0.05) */
  double mu_step_size;
  int mu_fixed_steps;        /* HMC: as augur_hmc_block.fixed_steps */
  const double* mu_prop_cov; /* MWG: D x D proposal covariance */
  int ref_compat;            /* MWG: keep the reference's Hastings term (both proposal
                                densities evaluated at the current point) */
} augur_gmm_desc;
augur_status augur_gmm_create(const augur_gmm_desc* desc, augur_sampler** out);

/* No counterpart in the C runtime: the experiments harness's predictive
 * log-likelihood on the device, per chain at the current state.
 *   LDA (experiments/lda/lda.py:149-173): held-out tokens of the model's
 *   n_docs documents (CSR doc_off[n_docs + 1], w), sum over tokens of
 *   log sum_k theta[d, k] phi[k, w].
 *   GMM (experiments/hmvgmm/hmvgmm.py:143-178): n test points y_test
 *   (n x D f64), sum over points of log sum_k pis_k MvNormal(y | mu_k, cov_k).
 * FP64 throughout. */
augur_status augur_lda_predictive_ll(augur_sampler* s, int n_docs, const int32_t* doc_off,
                                     const int32_t* w, double* out_per_chain);
augur_status augur_gmm_predictive_ll(augur_sampler* s, int n, const double* y_test,
                                     double* out_per_chain);
/* The same sums for given parameters (host buffers): theta (n_docs x K),
 * phi (K x V); pis (K), mu (K x D), cov (K x D x D).  One value out. */
augur_status augur_op_lda_predictive_ll(int n_docs, int k, int v, const double* theta,
                                        const double* phi, const int32_t* doc_off, const int32_t* w,
                                        double* out);
augur_status augur_op_gmm_predictive_ll(int k, int d, const double* pis, const double* mu,
                                        const double* cov, int n, const double* y_test, double* out);

/* ------------------------------------------------------------------------
 * Generic continuous models: a generated data pass under HMC.
 *
 * Replaces, for models outside the three families, the reference's
 * compile-and-load step: hs_compile (compiler/augur/src/FFI/AugurFFI.hs:35,
 * 78-89) generating the sampler's code (CudaC/CgCuda.hs:371-516 over
 * Core/CgLike.hs:90-267 and Core/CgGrad.hs:151-376), which pyaugur builds
 * with gcc / nvcc and dlopens (pyaugur/pyaugur/augurlib.py:263-423).
 *
 * `source` is HIP source that defines
 *   extern "C" __global__ void <kernel>(const double* q, double* part,
 *                                        const double* R, const int* I,
 *                                        const int* Z)
 * for a grid of (grid, chains) blocks of 256 threads.  For every g < grid
 * and chain c, one block writes part[(c * grid + g) * (n_params + 1) + j]:
 * a share of d ll / dq_j for j < n_params and of ll for j = n_params, at
 * chain c's point q[c * n_params ..) and discrete values Z[c * n_zints ..);
 * the grid shares of chain c sum to its ll and gradient.  R and I are the
 * `reals` and `ints` buffers.  The
 * library compiles it with hipRTC for the device and runs HMC
 * (cbits/augur_mcmclib.h:343-424).  Each block may name its own pass of the
 * same signature over its full conditional (the terms that mention its
 * variables, Core/RwCore.hs:58-67); `kernel` stays the log joint, used by
 * augur_get_curr_ll and augur_gen_eval.  Block b moves the coordinates where
 * masks[b * n_params + j] = 1, with the reference's jitter of step size
 * and count (fixed_steps as augur_hmc_block.fixed_steps).  Its Philox tags
 * are 16 + 4b (momentum), + 1 (jitter) and + 2 (accept).
 * Parameter names address slices of the vector; "state" is the whole
 * vector.  augur_stats writes (accepted, proposed) per block and chain.
 * ---------------------------------------------------------------------- */
typedef struct {
  const char* source;        /* HIP source of the data pass */
  const char* kernel;        /* its extern "C" name */
  int n_params;              /* P: reals per chain */
  int grid;                  /* blocks of the pass per chain, 1..4096 */
  int chains;                /* 1..16 */
  uint32_t seed;
  const double* reals;       /* R */
  size_t n_reals;
  const int32_t* ints;       /* I */
  size_t n_ints;
  int n_names;
  const char* const* names;  /* parameter names */
  const int* offsets;        /* first coordinate of each */
  const int* sizes;          /* coordinates of each */
  int n_blocks;              /* HMC blocks per sweep, 0..16 */
  const unsigned char* masks;  /* n_blocks x n_params */
  const double* sim_len;     /* n_blocks */
  const double* step_size;   /* n_blocks */
  const int* fixed_steps;    /* n_blocks */
  const char* const* block_kernels;  /* n_blocks names of the blocks' full-conditional
                                        passes in `source`, or null: the log joint's */
  const int* block_grids;    /* their grids */
  const int* block_kinds;    /* n_blocks AUGUR_GEN_HMC | AUGUR_GEN_DISC, or null: all HMC */
  int n_zints;               /* discrete coordinates per chain (the int32 vector "zstate") */
  int n_znames;
  const char* const* znames; /* discrete parameter names */
  const int* zoffsets;
  const int* zsizes;
} augur_gen_desc;
#define AUGUR_GEN_HMC 0
/* DiscGibbs on one discrete parameter (Core/CgDisc.hs): block_kernels[b]
 * names a kernel (const double* q, int* Z, const double* R, const int* I,
 * unsigned seed, unsigned sweep, unsigned tag) that redraws every point of
 * the parameter in Z from its full conditional; tag = 16 + 4b. */
#define AUGUR_GEN_DISC 1
augur_status augur_gen_create(const augur_gen_desc* desc, augur_sampler** out);
/* ll (chains) and d ll / dq (chains x n_params) at the points q (chains x
 * n_params): one pass and its reduction; the sampler's state is unchanged.
 * (The reference's generated ll and gradient functions,
 * Core/CgLike.hs:90-136, Core/CgGrad.hs:306-342.) */
augur_status augur_gen_eval(augur_sampler* s, const double* q, double* ll, double* grad);

/* ------------------------------------------------------------------------
 * Sampler operation (generic over families).
 * ---------------------------------------------------------------------- */

/* Replaces augur_set_pt (CgStInit.hs:518-531): write one chain's value of a
 * parameter; chain = -1 writes the same value to every chain. */
augur_status augur_set_pt(augur_sampler* s, const char* name, int chain, const void* src,
                          size_t bytes);
/* Replaces augur_cpy (CgStInit.hs:535-546): read one chain's current value. */
augur_status augur_cpy(augur_sampler* s, const char* name, int chain, void* dst, size_t bytes);
/* Replaces augur_step (CudaC/CgKern.hs:232-238): run n sweeps of the
 * schedule on every chain; returns when they are complete. */
augur_status augur_step(augur_sampler* s, int n_sweeps);
/* Replaces augur_get_curr_ll (CgStInit.hs:597-602): log-density of the
 * current state, one value per chain (FP64). */
augur_status augur_get_curr_ll(augur_sampler* s, double* out_per_chain);
/* Record `name` after every sweep into a device trace of `capacity` sweeps
 * (capacity 0 disables).  Recording starts with the next augur_step. */
augur_status augur_trace(augur_sampler* s, const char* name, int capacity);
/* Copy the recorded sweeps (oldest first, all chains: [sweep][chain][...]). */
augur_status augur_trace_read(augur_sampler* s, const char* name, void* dst, size_t bytes,
                              int* n_recorded);
/* Per-chain sampler statistics: HMC blocks report accepted proposals and
 * proposals for each block ([block][chain][2] as double); Gibbs samplers
 * report the sweep count (the GMM then, when bytes allow, [chain][2]
 * accepted and proposed moves of a non-conjugate mu kernel). */
augur_status augur_stats(augur_sampler* s, double* out, size_t bytes);
/* Number of sweeps run so far (the Philox sweep counter). */
augur_status augur_sweep_count(augur_sampler* s, uint32_t* out);
/* No counterpart in the reference (its only timing is time.clock() around
 * samplen, experiments/util/timer.py:7-14).  Times the family's dominant
 * kernel (HLR: the data pass; LDA and GMM: the z step) with HIP events on the
 * sampler's stream: during every `period`-th sweep (0 disables) each launch
 * of that kernel is bracketed by an event pair. */
augur_status augur_profile(augur_sampler* s, int period);
/* Average duration in microseconds of the launches timed since the last
 * read, and their number; resets the count. */
augur_status augur_profile_read(augur_sampler* s, double* avg_us, int* n_launches);
augur_status augur_destroy(augur_sampler* s);

/* ------------------------------------------------------------------------
 * Kernel-level entry points (host buffers in and out).  Each runs one HIP
 * kernel of the sampler on device 0 and exists so the kernels can be
 * checked one by one against the CPU restatement.
 * ---------------------------------------------------------------------- */

/* No counterpart in the reference.  Attainable device bandwidth on device
 * 0: a grid-stride read and a copy of `bytes` (rounded down to 16), each
 * launched reps times back to back after one untimed launch, timed with
 * HIP events; copy counts read + write bytes.  GB/s out. */
augur_status augur_op_stream_bandwidth(size_t bytes, int reps, double* read_gbs, double* copy_gbs);
/* n Philox4x32-10 blocks: ctr[4 n], key[2 n] -> out[4 n]. */
augur_status augur_op_philox(const uint32_t* ctr, const uint32_t* key, int n, uint32_t* out);
/* out[i] = Gamma(shape[i], 1) at site i (cbits/augur_dist.c:302-312). */
augur_status augur_op_gamma(uint32_t seed, uint32_t chain, uint32_t sweep, uint32_t tag,
                            const double* shape, int n, double* out);
/* rows independent Dirichlet(alpha + cnt[r]) draws, row r at sites r * k + i
 * (cbits/augur_dist.c:526-538). */
augur_status augur_op_dirichlet_conj(uint32_t seed, uint32_t chain, uint32_t sweep,
                                     uint32_t tag, const double* alpha, const int32_t* cnt,
                                     int rows, int k, double* out);
/* The LDA DiscGibbs [z] kernel on given FP32 theta (D x K) and word-major
 * phi_t (V x K) tables (Core/CgDisc.hs:122-144); also returns the count
 * tables of the new z (Core/CgConj.hs:171-183). */
augur_status augur_op_lda_zstep(int n_docs, int k, int v, const int32_t* doc_off,
                                const int32_t* w, const float* theta, const float* phi_t,
                                uint32_t seed, uint32_t chain, uint32_t sweep, int32_t* z,
                                int32_t* cnt_theta, int32_t* cnt_phi);
/* HLR data term for `chains` parameter vectors beta[c] = (theta[0..D), b):
 * ll[c] = sum_n log Bernoulli(y_n | sigmoid(x_n . theta + b)) and its
 * gradient grad[c] = (d/dtheta, d/db) (Core/CgLike.hs:90-136,
 * Core/CgGrad.hs:165-233).  FP32 arithmetic, FP64 accumulation. */
augur_status augur_op_hlr_ll_grad(int n_rows, int d, int chains, const float* x,
                                  const int32_t* y, const double* beta, int has_intercept,
                                  double* ll, double* grad);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the data term of augur_op_hlr_ll_grad on buffers the
 * caller owns in device memory, queued on `stream` (a hipStream_t; null: the
 * default stream) of the current device, with no allocation.  x (N x D FP32,
 * 16-byte aligned), y (N labels in {0, 1}), beta (chains x (D + 1) f64) in,
 * ll (chains) and grad (chains x (D + 1)) out, as augur_op_hlr_ll_grad.  The
 * kernels' scratch (FP32 evaluation points, block partials, per-chain sums)
 * is a library-owned static device array of 4 MiB per device, kept until
 * the process ends; every entry is written before it is read, so no call
 * depends on an earlier one.  Calls that overlap on one device (other
 * streams, other threads) race on that array.  Shapes as
 * augur_op_hlr_ll_grad (D <= 256, chains <= 16); a shape whose scratch would
 * exceed the array is rejected with AUGUR_ERR_ARG.  Returns once the kernels
 * are queued; the results are ready when the stream reaches them. */
augur_status augur_op_hlr_ll_grad_noalloc(int n_rows, int d, int chains, const float* x,
                                          const int32_t* y, const double* beta, int has_intercept,
                                          double* ll, double* grad, void* stream);
/* The same data log-density in FP64 row arithmetic (FP64 matrix cores for
 * D in {16, 32, 64, 128} without intercept): the Metropolis term of
 * surrogate HMC.  Same arguments as augur_op_hlr_ll_grad, no gradient. */
augur_status augur_op_hlr_ll64(int n_rows, int d, int chains, const float* x, const int32_t* y,
                               const double* beta, int has_intercept, double* ll);
/* The GMM DiscGibbs [z] kernel for one chain given pis, mu, cov (FP64, as
 * drawn by the conjugate kernels; cbits/augur_dist.c:548-568). */
augur_status augur_op_gmm_zstep(int n, int d, int k, const float* y, const double* pis,
                                const double* mu, const double* cov, uint32_t seed,
                                uint32_t chain, uint32_t sweep, int32_t* z);
/* Conjugate draws of one GMM sweep on the device for one chain, from given
 * counts and sums (cbits/augur_dist.c:633-669 and :740-757): updates pis
 * (if sample_pi), mu and cov in place. sum_y: K x D sums of assigned points;
 * scatter: K x D x D, sum over assigned points of (y - mu_new)(y - mu_new)^T
 * is formed on the device from sum_yy (K x D x D, sum of y y^T). */
augur_status augur_op_gmm_conj(int d, int k, const double* alpha, const double* mu0,
                               const double* cov0, int df, const double* scale0,
                               const int32_t* cnt, const double* sum_y, const double* sum_yy,
                               int sample_pi, int sample_cov, uint32_t seed, uint32_t chain,
                               uint32_t sweep, double* pis, double* mu, double* cov);

#ifdef __cplusplus
}
#endif
#endif /* AUGUR_HIP_H */
