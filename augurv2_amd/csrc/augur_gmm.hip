// augur_gmm.hip -- Gibbs sampler for the multivariate Gaussian mixture
// (hmvgmm.rv; mvgmm.rv with fixed pis and covs) on gfx950.
//
// Sweep, schedule ConjGibbs [pis] (*) ConjGibbs [mu] (*) ConjGibbs [cov] (*)
// DiscGibbs [z] (experiments/experiment.py:127, order CudaC/CgKern.hs:192-238):
//
//   k_gmm_stats_sum  the statistics of every (chain, component) summed
//               over the blocks of the statistics kernel, FP64, fixed order.
//   k_gmm_conj  one wave per (component, chain): draws pis_k's Gamma
//               (Core/CgConj.hs:171-183), mu_k | cov_k
//               (cbits/augur_dist.c:633-669), cov_k | mu_k
//               (cbits/augur_dist.c:740-757, Bartlett), all in FP64 with the
//               matrices in registers (lane i: row i), and builds the
//               component's table for the z step: L_k^-1 (packed lower
//               triangle), mu_k and -sum log diag L_k, stored FP32.
//               (k_gmm_prec0 / k_gmm_tables build cov0^-1 and the tables at
//               creation and after a set of mu or cov.)
//   k_gmm_z     point-parallel categorical draw (Core/CgDisc.hs:122-144) in
//               log space with max subtraction: logit_k = log pis_k
//               - 0.5 log det cov_k - 0.5 |L_k^-1 (y - mu_k)|^2, tables read
//               through the scalar cache (FP32 vector form).
//   k_gmm_z_mfma  the same draw for a padded dimension of 16 with
//               L_k^-1 y on the matrix cores (BF16-split 16x16x32 products)
//               minus c_k = L_k^-1 mu_k.
//   k_gmm_stats_t  the statistics of the new z for the next sweep, centred
//               at the current mu_k: count, sum (y - mu_k) and
//               sum (y - mu_k)(y - mu_k)^T, as one-hot products per wave of
//               64 points on the matrix cores (one-hot rows exact in BF16,
//               entries split in three BF16 parts; no atomics: FP32 per block
//               in a fixed order).  The next sweep's scatter around the new
//               mu is S + s d^T + d s^T + n d d^T with d = mu_old - mu_new.
//
// The reference multiplies raw pdfs in linear space (cbits/augur_dist.c:
// 798-813), which underflows at D = 16; the log-space draw targets the same
// distribution.
#include "augur_common.h"
#include "augur_hmc.h"

namespace augur {

constexpr int kGmmThreads = 256;
constexpr int kMaxDim = 16;
constexpr int kMaxComp = 64;

__host__ __device__ constexpr int tri(int d) { return d * (d + 1) / 2; }

typedef float gmm_f32x4 __attribute__((ext_vector_type(4)));

struct GmmDev {
  int n, d, k, G, S;       // S = 1 + d + d(d+1)/2 statistics per component
  int Gs;                  // blocks per chain of the statistics kernel (rows of part)
  const float* y;          // n x d
  const double* alpha;     // k
  const double* mu0;       // d
  const double* cov0;      // d x d
  const double* scale0;    // d x d
  int df, sample_pi, sample_cov;
  double* gpi;     // C x k unnormalised Dirichlet gammas (or pis when fixed)
  double* pis64;   // C x k normalised
  double* mu64;    // C x k x d
  double* cov64;   // C x k x d x d
  float* tab;      // C x k x TS   (TS = tri(DM) + DM + 1, DM = padded dim)
  double* prec;    // C x k x d x d  cov_k^-1 of the current cov_k (with the table)
  double* prec0;   // d x d          cov0^-1
  int TS, DM;
  int cus;         // compute units of the device
  float* part;     // C x Gs x k x S
  int32_t* z;      // C x n
  // the mu block's move when it is not ConjGibbs (GMM_MU_*)
  int mu_mode, ref_compat, mu_fixed_steps;
  double mu_sim_len, mu_step;
  const double* cov0_inv;  // d x d
  const double* chol0;     // d x d lower, cov0 = L L^T
  const double* cholp;     // d x d lower, MWG proposal covariance
  double* mcent;  // C x k x d      centre of the statistics (the FP32 table mu)
  double* mprec;  // C x k x d x d  cov_k^-1
  double* mlin;   // C x k x d      cov_k^-1 s, s = sum over the component of (y - centre)
  double* mcnt;   // C x k          points in the component
  double* macc;   // C x 2          accepted / proposed moves of the mu block
  double* nuts;   // C x 3 kNutsSlots x 1024: the NUTS stacks, one value per thread
};

enum { GMM_MU_CONJ = 0, GMM_MU_ESLICE = 1, GMM_MU_HMC = 2, GMM_MU_MWG = 3, GMM_MU_RSLICE = 4, GMM_MU_NUTS = 5 };

// NUTS [mu]: the tree depth is capped at kNutsMaxDepth doublings; its stacks
// (proposals, theta, momenta) hold kNutsMaxDepth + 2 vectors each.
constexpr int kNutsMaxDepth = 10;
constexpr int kNutsSlots = kNutsMaxDepth + 2;

// ---- small FP64 linear algebra (cbits/augur_matop.c:312-406), row-major ----
// Wave-parallel: one 64-lane block holds the matrices in LDS and every
// element is computed with the serial definition's operations in the same
// order (terms the serial in-place loops multiply by an exact zero are left
// out), so the results equal the serial forms bit for bit.  Callers keep
// the whole block in the call; `tmp` is d x d scratch.
constexpr int kConjLanes = 64;

// Cholesky factor in the lower triangle; the upper triangle is left as is.
__device__ void chol(double* a, int n) {
  const int t = threadIdx.x;
  for (int j = 0; j < n; ++j) {
    if (t == 0) {
      double s = 0.0;
      for (int k = 0; k < j; ++k) s += a[j * n + k] * a[j * n + k];
      a[j * n + j] = sqrt(a[j * n + j] - s);
    }
    __syncthreads();
    for (int i = j + 1 + t; i < n; i += kConjLanes) {
      double s = 0.0;
      for (int k = 0; k < j; ++k) s += a[i * n + k] * a[j * n + k];
      a[i * n + j] = (a[i * n + j] - s) / a[j * n + j];
    }
    __syncthreads();
  }
}
// Inverse of the lower triangle in place (upper set to zero): column c by
// forward substitution, one lane per column.
__device__ void lower_trmi(double* l, int n, double* tmp) {
  for (int c = threadIdx.x; c < n; c += kConjLanes)
    for (int m = 0; m < n; ++m) {
      if (m < c) {
        tmp[m * n + c] = 0.0;
        continue;
      }
      double b = (c == m) ? 1.0 : 0.0;
      for (int i = c; i < m; ++i) b -= l[m * n + i] * tmp[i * n + c];
      tmp[m * n + c] = b / l[m * n + m];
    }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += kConjLanes) l[e] = tmp[e];
  __syncthreads();
}
// From a Cholesky factor in the lower triangle to the full inverse matrix.
__device__ void chol_invert(double* a, int n, double* tmp) {
  lower_trmi(a, n, tmp);
  for (int e = threadIdx.x; e < n * n; e += kConjLanes) {
    const int r = e / n, c = e % n;
    if (r < c) continue;
    double acc = 0.0;
    for (int k = c; k < n; ++k) acc += a[k * n + r] * a[k * n + c];
    tmp[r * n + c] = acc;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += kConjLanes) {
    const int r = e / n, c = e % n;
    a[e] = r >= c ? tmp[r * n + c] : tmp[c * n + r];
  }
  __syncthreads();
}
// l2 = l1 l2 for lower triangular l1 (cbits/augur_matop.c:381-392).
__device__ void trtrmm(const double* l1, double* l2, int n, double* tmp) {
  for (int e = threadIdx.x; e < n * n; e += kConjLanes) {
    const int i = e / n, j = e % n;
    double t = 0.0;
    for (int k = 0; k < i; ++k) t += l1[i * n + k] * l2[k * n + j];
    t += l1[i * n + i] * l2[i * n + j];
    tmp[e] = t;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += kConjLanes) l2[e] = tmp[e];
  __syncthreads();
}

// ---- the same algebra with the matrices in registers -----------------------
// One wave per problem: lane i < d holds row i of a d x d matrix in
// v[0..d) (the column forms: column i).  Another lane's value is read with
// v_readlane (the source lane is uniform), so a dependent step costs VALU
// latency instead of an LDS round trip; each element keeps the serial
// definition's operation order, and terms that are exact zeros are left
// out (adding them cannot change a sum that starts at +0).
__device__ inline double lane_bcast(double x, int src) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// Cholesky factor in the lower triangle of the rows (above the diagonal
// unchanged): the order of chol() above.
__device__ inline void rchol(double (&a)[kMaxDim], int d) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < kMaxDim; ++j) {
    if (j < d) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < j; ++k) s += a[k] * a[k];
      const double djj = lane_bcast(sqrt(a[j] - s), j);
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < j; ++k) t += a[k] * lane_bcast(a[k], j);
      if (lane == j) a[j] = djj;
      else if (lane > j) a[j] = (a[j] - t) / djj;
    }
  }
}

// T = L^-1 of the lower triangle of L (rows): lane c gets column c of T
// (zeros above the diagonal), the order of lower_trmi() above.
__device__ inline void rtrmi_cols(const double (&l)[kMaxDim], double (&t)[kMaxDim], int d) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int m = 0; m < kMaxDim; ++m) {
    t[m] = 0.0;
    if (m < d) {
      double b = (lane == m) ? 1.0 : 0.0;
#pragma unroll
      for (int i = 0; i < m; ++i) {
        const double lmi = lane_bcast(l[i], m);
        if (i >= lane) b -= lmi * t[i];
      }
      const double lmm = lane_bcast(l[m], m);
      if (m >= lane) t[m] = b / lmm;
    }
  }
}

// Rows from columns (lane i gets row i).
__device__ inline void rcols_to_rows(const double (&t)[kMaxDim], double (&r)[kMaxDim], int d) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < kMaxDim; ++c) r[c] = 0.0;
#pragma unroll
  for (int i = 0; i < kMaxDim; ++i)
    if (i < d)
#pragma unroll
      for (int c = 0; c < kMaxDim; ++c)
        if (c < d) {
          const double x = lane_bcast(t[i], c);
          if (lane == i) r[c] = x;
        }
}

// Rows of T^T T from the columns of T (lane c: column c): entry (i, r) is
// the sum over k >= max(i, r) of T[k][i] T[k][r] (chol_invert() order).
__device__ inline void rinv_from_cols(const double (&t)[kMaxDim], double (&a)[kMaxDim], int d) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < kMaxDim; ++r) {
    a[r] = 0.0;
    if (r < d) {
      double acc = 0.0;
#pragma unroll
      for (int k = r; k < kMaxDim; ++k)
        if (k < d) {
          const double tkr = lane_bcast(t[k], r);
          if (k >= lane) acc += t[k] * tkr;
        }
      a[r] = acc;
    }
  }
}

// (L L^T)^-1 from the Cholesky factor L (rows), in rows.
__device__ inline void rchol_invert(double (&a)[kMaxDim], int d) {
  double t[kMaxDim];
  rtrmi_cols(a, t, d);
  rinv_from_cols(t, a, d);
}

// l2 = l1 l2 for lower triangular l1 (rows; trtrmm() order: the terms below
// the diagonal in k order, then the diagonal term).
__device__ inline void rtrtrmm(const double (&l1)[kMaxDim], double (&l2)[kMaxDim], int d) {
  const int lane = threadIdx.x & 63;
  double out[kMaxDim], diag = 0.0;
#pragma unroll
  for (int j = 0; j < kMaxDim; ++j) {
    out[j] = 0.0;
    if (j == lane) diag = l1[j];
  }
#pragma unroll
  for (int k = 0; k < kMaxDim; ++k)
    if (k < d)
#pragma unroll
      for (int j = 0; j < kMaxDim; ++j)
        if (j < d) {
          const double l2kj = lane_bcast(l2[j], k);
          if (k < lane) out[j] += l1[k] * l2kj;
        }
#pragma unroll
  for (int j = 0; j < kMaxDim; ++j) l2[j] = j < d ? out[j] + diag * l2[j] : 0.0;
}

// (M x)_i on lane i for M in rows and x_j on lane j.
__device__ inline double rmatvec(const double (&mrow)[kMaxDim], double x, int d) {
  double t = 0.0;
#pragma unroll
  for (int j = 0; j < kMaxDim; ++j)
    if (j < d) t += mrow[j] * lane_bcast(x, j);
  return t;
}

// Row i (lane i < d) of a d x d row-major matrix in global memory.
__device__ inline void rload(const double* g, int d, double (&a)[kMaxDim]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < kMaxDim; ++j) a[j] = (lane < d && j < d) ? g[lane * d + j] : 0.0;
}

// z-step table of one component from (mu_k, cov_k) (cov in rows): L_k^-1
// (packed lower triangle), mu_k and -sum log diag L_k, stored FP32; and
// cov_k^-1 in FP64.  Called by the whole wave.
__device__ __forceinline__ void build_table_r(const GmmDev& m, int c, int k, const double (&cov)[kMaxDim]) {
  const int d = m.d, DM = m.DM, lane = threadIdx.x & 63;
  double L[kMaxDim], t[kMaxDim], tr[kMaxDim];
#pragma unroll
  for (int j = 0; j < kMaxDim; ++j) L[j] = cov[j];
  rchol(L, d);
  double ldet = 0.0;
#pragma unroll
  for (int i = 0; i < kMaxDim; ++i)
    if (i < d) ldet += log(lane_bcast(L[i], i));
  rtrmi_cols(L, t, d);
  rcols_to_rows(t, tr, d);
  // cov_k^-1 in FP64 for the next mu_k | cov_k draw (chol_invert() of cov_k)
  double prec[kMaxDim];
  rinv_from_cols(t, prec, d);
  if (lane < d) {
    double* pg = m.prec + ((size_t)c * m.k + k) * d * d + lane * d;
#pragma unroll
    for (int j = 0; j < kMaxDim; ++j)
      if (j < d) pg[j] = prec[j];
  }
  float* tab = m.tab + ((size_t)c * m.k + k) * m.TS;
  if (lane < DM) {
#pragma unroll
    for (int j = 0; j < kMaxDim; ++j)
      if (j <= lane && j < DM) tab[tri(lane) + j] = (lane < d && j < d) ? (float)tr[j] : 0.0f;
    const double* mu = m.mu64 + ((size_t)c * m.k + k) * d;
    tab[tri(DM) + lane] = lane < d ? (float)mu[lane] : 0.0f;
  }
  if (lane == 0) tab[tri(DM) + DM] = (float)(-ldet);  // -0.5 log det cov
}

// Per (chain, component), the statistics summed over the blocks of the
// statistics kernel in FP64: grid (values / 64, K, chains), thread (value
// vi, slice sl) adds rows g = sl, sl + 4, ... (loads coalesced over the
// values of a row), then the four slices in order.  sum[c][k][S].
constexpr int kSumSlices = 4;
__global__ __launch_bounds__(64 * kSumSlices) void k_gmm_stats_sum(GmmDev m, double* __restrict__ sum) {
  __shared__ double red[kSumSlices][64];
  const int vi = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + vi, k = blockIdx.y, c = blockIdx.z, S = m.S;
  double s = 0.0;
  if (i < S) {
    const float* col = m.part + ((size_t)c * m.Gs * m.k + k) * S + i;
    const size_t row = (size_t)m.k * S;
    int g = sl;
    for (; g + 7 * kSumSlices < m.Gs; g += 8 * kSumSlices) {
      float t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] = col[(size_t)(g + u * kSumSlices) * row];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += (double)t[u];
    }
    for (; g < m.Gs; g += kSumSlices) s += (double)col[(size_t)g * row];
  }
  red[sl][vi] = s;
  __syncthreads();
  if (sl == 0 && i < S) sum[((size_t)c * m.k + k) * S + i] = ((red[0][vi] + red[1][vi]) + red[2][vi]) + red[3][vi];
}

// One wave per (component k, chain c): the summed statistics, then pi_k
// (Gamma), mu_k | cov_k (MvNormal conjugate) and cov_k | mu_k (InvWishart)
// and the new z-step table.  Matrices are held in registers (lane i: row i),
// vectors one element per lane.
__global__ __launch_bounds__(kConjLanes) void k_gmm_conj(GmmDev m, const double* __restrict__ ssum,
                                                         uint32_t seed, uint32_t sweep) {
  const int k = blockIdx.x, c = blockIdx.y, lane = threadIdx.x;
  const int d = m.d, S = m.S;
  const size_t ck = (size_t)c * m.k + k;
  const double* st = ssum + ck * S;  // count, sum (y - centre), packed lower scatter
  const int cnt = (int)llrint(st[0]);
  const bool on = lane < d;
  const double sc = on ? st[1 + lane] : 0.0;
  double Sc[kMaxDim];
#pragma unroll
  for (int j = 0; j < kMaxDim; ++j)
    Sc[j] = (on && j < d) ? (j <= lane ? st[1 + d + tri(lane) + j] : st[1 + d + tri(j) + lane]) : 0.0;
  double* mu = m.mu64 + ck * d;  // current mu = centre
  double* cov = m.cov64 + ck * d * d;
  if (m.sample_pi && lane == 0) {
    const Stream sp{seed, (uint32_t)c, sweep, TAG_GMM_PI};
    m.gpi[ck] = gamma_at(sp, (uint32_t)k, m.alpha[k] + (double)cnt);
  }
  // the statistics' centre: the FP32 table copy of the mu they were taken at
  // (saved by k_gmm_mu_prep when another kernel has moved mu since)
  const bool mu_moved = m.mu_mode != GMM_MU_CONJ;
  const double cen = on ? (mu_moved ? m.mcent[ck * d + lane] : (double)(float)mu[lane]) : 0.0;
  double munew = on ? mu[lane] : 0.0;
  if (!mu_moved) {
    // ---- mu_k | cov_k: cbits/augur_dist.c:633-669 ----
    // cov0^-1 (k_gmm_prec0) and cov_k^-1 (stored with the table of cov_k):
    // the same operations as a Cholesky factor and its chol_invert here
    double Q[kMaxDim], P[kMaxDim];
    rload(m.prec0, d, Q);
    rload(m.prec + ck * d * d, d, P);
#pragma unroll
    for (int j = 0; j < kMaxDim; ++j) P[j] *= (double)cnt;  // n cov^-1
    // the z step centred its statistics at the FP32 table copy of mu: the
    // mean of the assigned points is centre + s / n (the sum itself when n = 0)
    const double inv = cnt == 0 ? 1.0 : 1.0 / (double)cnt;
    const double mu0 = on ? m.mu0[lane] : 0.0;
    double v = rmatvec(Q, mu0, d);
    const double mean = ((double)cnt * cen + sc) * inv;
    v += rmatvec(P, mean, d);
#pragma unroll
    for (int j = 0; j < kMaxDim; ++j) Q[j] += P[j];
    rchol(Q, d);
    rchol_invert(Q, d);  // A = (cov0^-1 + n cov^-1)^-1
    const Stream sm{seed, (uint32_t)c, sweep, TAG_GMM_MU};
    const double pmean = rmatvec(Q, v, d);
    const double z = on ? normal_at(sm, (uint32_t)(k * d + lane), 0) : 0.0;
    rchol(Q, d);
    double t = 0.0;  // trvm: (L z)_i over j <= i
#pragma unroll
    for (int j = 0; j < kMaxDim; ++j)
      if (j < d) {
        const double zj = lane_bcast(z, j);
        if (j <= lane) t += Q[j] * zj;
      }
    munew = t + pmean;
  }
  double cv[kMaxDim];
  if (m.sample_cov) {
    // ---- cov_k | mu_k: IW(df + n, scale0 + sum (y - mu)(y - mu)^T) ----
    const double dl = cen - munew;  // centre - mu_new
    double A[kMaxDim];
#pragma unroll
    for (int j = 0; j < kMaxDim; ++j) {
      const double scj = lane_bcast(sc, j), dlj = lane_bcast(dl, j);
      A[j] = (on && j < d) ? Sc[j] + sc * dlj + dl * scj + (double)cnt * dl * dlj + m.scale0[lane * d + j] : 0.0;
    }
    // augur_invwishart_sample (cbits/augur_dist.c:740-749)
    rchol(A, d);
    rchol_invert(A, d);
    rchol(A, d);
    const Stream sc_s{seed, (uint32_t)c, sweep, TAG_GMM_COV};
    const int dfn = m.df + cnt;
    // Bartlett factor: normals below the diagonal, sqrt(2 Gamma) on it
    // (drawn in a rolled loop through LDS, then read into registers)
    __shared__ double bart[kMaxDim * (kMaxDim + 1)];
    if (on) {
      const uint32_t row = (uint32_t)(k * d * d + lane * d);
      for (int j = 0; j < lane; ++j) bart[lane * (kMaxDim + 1) + j] = normal_at(sc_s, row + j, 0);
      bart[lane * (kMaxDim + 1) + lane] = sqrt(2.0 * gamma_at(sc_s, row + lane, 0.5 * (double)(dfn - lane)));
    }
    __syncthreads();
    double B[kMaxDim];
#pragma unroll
    for (int j = 0; j < kMaxDim; ++j) B[j] = (on && j <= lane && j < d) ? bart[lane * (kMaxDim + 1) + j] : 0.0;
    rtrtrmm(A, B, d);
    rchol_invert(B, d);
#pragma unroll
    for (int j = 0; j < kMaxDim; ++j) {
      if (on && j < d) cov[lane * d + j] = B[j];
      cv[j] = B[j];
    }
  } else {
    rload(cov, d, cv);
  }
  if (on) mu[lane] = munew;
  build_table_r(m, c, k, cv);
}

// ---- mu moves other than ConjGibbs --------------------------------------
// Given z and cov_k, the data term of mu_k is a quadratic in mu_k fixed by the
// component's statistics around the centre c_k (count n, s = sum (y - c),
// S = sum (y - c)(y - c)^T):
//   sum_n log N(y_n | mu, cov) = -1/2 tr(P S) + d^T P s - n/2 d^T P d + const,
//   d = mu - c, P = cov^-1,
// with gradient P s - n P d.  The moves below evaluate it from the
// statistics (O(D^2) per evaluation) where the reference loops over the
// data on every evaluation; the constant -1/2 tr(P S) - n/2 log det cov cancels
// in every comparison the kernels make (slice levels, Metropolis ratios).

// Per (component, chain): P_k, P_k s_k, n_k and c_k for k_gmm_mu_move.
__global__ __launch_bounds__(kConjLanes) void k_gmm_mu_prep(GmmDev m, const double* __restrict__ ssum) {
  const int k = blockIdx.x, c = blockIdx.y, tid = threadIdx.x, d = m.d, S = m.S;
  __shared__ double L[kMaxDim * kMaxDim], sc[kMaxDim];
  const size_t ck = (size_t)c * m.k + k;
  // cov_k^-1 as stored with the table of cov_k
  for (int e = tid; e < d * d; e += kConjLanes) L[e] = m.prec[ck * d * d + e];
  for (int i = tid; i < d; i += kConjLanes) sc[i] = ssum[ck * S + 1 + i];
  __syncthreads();
  for (int e = tid; e < d * d; e += kConjLanes) m.mprec[ck * d * d + e] = L[e];
  for (int i = tid; i < d; i += kConjLanes) {
    double t = 0.0;
    for (int j = 0; j < d; ++j) t += L[i * d + j] * sc[j];
    m.mlin[ck * d + i] = t;
    m.mcent[ck * d + i] = (double)(float)m.mu64[ck * d + i];
  }
  if (tid == 0) m.mcnt[ck] = ssum[ck * S];
}

constexpr int kMuThreads = 1024;
constexpr int kMuWaves = kMuThreads / 64;
constexpr int kESliceMaxShrinks = 200;
constexpr double kTwoPi = 6.283185307179586476925;

// Data term of component k at x (lane i < d holds x_i) minus its constant;
// with_prior adds the MvNormal(mu0, cov0) prior term, also without its
// constant.  Every lane returns the same sum (xor butterfly).
__device__ double mu_comp_ll(const GmmDev& m, int c, int k, double x, bool with_prior) {
  const int lane = threadIdx.x & 63, d = m.d;
  const size_t ck = (size_t)c * m.k + k;
  const bool on = lane < d;
  const double dl = on ? x - m.mcent[ck * d + lane] : 0.0;
  const double dm = on ? x - m.mu0[lane] : 0.0;
  double pd = 0.0, q = 0.0;
  for (int j = 0; j < d; ++j) {
    const double dj = __shfl(dl, j), mj = __shfl(dm, j);
    if (on) {
      pd += m.mprec[(ck * d + lane) * d + j] * dj;
      q += m.cov0_inv[lane * d + j] * mj;
    }
  }
  double t = on ? dl * m.mlin[ck * d + lane] - 0.5 * m.mcnt[ck] * dl * pd : 0.0;
  if (with_prior && on) t -= 0.5 * dm * q;
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
  return t;
}

// Lane i < d: (L z)_i for the lower triangular L (d x d) and z_i in lane i.
__device__ double mu_lower_times(const double* L, int d, double z) {
  const int lane = threadIdx.x & 63;
  double r = 0.0;
  for (int j = 0; j < d; ++j) {
    const double zj = __shfl(z, j);
    if (lane < d && j <= lane) r += L[lane * d + j] * zj;
  }
  return r;
}

// Block-wide FP64 sum (every thread gets it), waves in a fixed order.
__device__ double mu_block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kMuWaves; ++w) s += red[w];
  return s;
}

// Gradient of the mu block's log full conditional at th (all coordinates in
// LDS, thread t = k d + i owns coordinate i of component k) and this
// thread's share of the log density (prior and data, without constants).
__device__ void mu_block_eval(const GmmDev& m, int c, const double* th, double* g, double* lt) {
  const int t = threadIdx.x, d = m.d;
  *g = 0.0;
  *lt = 0.0;
  if (t >= m.k * d) return;
  const int k = t / d, i = t % d;
  const size_t ck = (size_t)c * m.k + k;
  double pd = 0.0, q = 0.0;
  for (int j = 0; j < d; ++j) {
    pd += m.mprec[(ck * d + i) * d + j] * (th[k * d + j] - m.mcent[ck * d + j]);
    q += m.cov0_inv[i * d + j] * (th[k * d + j] - m.mu0[j]);
  }
  const double n = m.mcnt[ck], lin = m.mlin[ck * d + i];
  const double dl = th[t] - m.mcent[ck * d + i], dm = th[t] - m.mu0[i];
  *g = lin - n * pd - q;
  *lt = dl * lin - 0.5 * n * dl * pd - 0.5 * dm * q;
}

// NUTS over the block of all means (cbits/augur_mcmclib.h:548-776), the
// recursion of h_augur_nuts_build_tree unrolled into a stack of frames that
// every thread keeps alike; thread t owns coordinate t of every vector, its
// stack entries in global scratch.  Draws: momentum at TAG_GMM_MU_MOVE site
// t; the slice's Exp(1) from uniform (0, 1) and the i-th Bernoulli draw
// from uniform (i, 3) at TAG_GMM_MU_MOVE + 2, in the reference's order.
// Returns whether the last doubling's proposal was kept.
__device__ bool mu_nuts(const GmmDev& m, int c, double* th, double* red, uint32_t seed, uint32_t sweep, double* x) {
  const int t = threadIdx.x, dim = m.k * m.d;
  const bool own = t < dim;
  const Stream sn{seed, (uint32_t)c, sweep, TAG_GMM_MU_MOVE};
  const Stream su{seed, (uint32_t)c, sweep, TAG_GMM_MU_MOVE + 2};
  double* base = m.nuts + (size_t)c * 3 * kNutsSlots * kMuThreads + t;
  double* pstk = base;                                      // proposals
  double* tstk = base + (size_t)kNutsSlots * kMuThreads;    // theta
  double* mstk = base + (size_t)2 * kNutsSlots * kMuThreads; // momenta
  auto at = [](double* s, int i) -> double& { return s[(size_t)i * kMuThreads]; };
  uint32_t nb = 0;
  auto bern = [&](double bias) { return u01(su.draw(nb++, 3).x) < bias; };
  double g, lt;
  // grad and log-density share at xv (all threads, block barriers inside)
  auto eval = [&](double xv) {
    __syncthreads();
    if (own) th[t] = xv;
    __syncthreads();
    mu_block_eval(m, c, th, &g, &lt);
  };
  const double x0 = *x;
  double tm = x0, tp = x0, mm = own ? normal_at(sn, (uint32_t)t, 0) : 0.0, mp = mm;
  eval(x0);
  const double logu = mu_block_sum(lt - 0.5 * mm * mm, red) + log(u01(su.draw(0, 1).x));
  int n = 1, j = 0, ptop = 0, ttop = 0;
  bool f_cpy = false;
  double xnew = x0;
  int fj[kNutsMaxDepth + 1], fst[kNutsMaxDepth + 1], fln[kNutsMaxDepth + 1], fls[kNutsMaxDepth + 1];
  while (true) {
    ttop = 0;
    const bool minus = bern(0.5);
    const double step = minus ? -m.mu_step : m.mu_step;
    double xw = minus ? tm : tp, mw = minus ? mm : mp;
    int top = 1, rn = 0, rs = 0;
    fj[0] = j;
    fst[0] = 0;
    while (top > 0) {
      const int f = top - 1;
      if (fj[f] == 0) {  // leaf: cbits/augur_mcmclib.h:305-325 and :603-615
        eval(xw);
        mw += step / 2.0 * g;
        xw += step * mw;
        eval(xw);
        mw += step / 2.0 * g;
        if (own) at(pstk, ptop) = xw;
        ++ptop;
        const double ll = mu_block_sum(lt - 0.5 * mw * mw, red);
        rn = logu <= ll;
        rs = logu < 1000.0 + ll;
        --top;
        continue;
      }
      if (fst[f] == 0) {  // build the left subtree
        fst[f] = 1;
        fj[top] = fj[f] - 1;
        fst[top] = 0;
        ++top;
        continue;
      }
      if (fst[f] == 1) {  // the left subtree returned (rn, rs)
        fln[f] = rn;
        fls[f] = rs;
        if (rs != 1) {
          --top;
          continue;
        }
        if (fj[f] == 1) {
          if (own) {
            at(tstk, ttop) = xw;
            at(mstk, ttop) = mw;
          }
          ++ttop;
        }
        fst[f] = 2;
        fj[top] = fj[f] - 1;
        fst[top] = 0;
        ++top;
        continue;
      }
      // the right subtree returned (rn, rs)
      const double bias = (double)rn / (double)(fln[f] + rn);
      if (bern(bias) && own) at(pstk, ptop - 2) = at(pstk, ptop - 1);
      --ptop;
      fln[f] += rn;
      int snew = 0;
      if (rs == 1) {
        const double pt = own ? at(tstk, ttop - 1) : 0.0, pm = own ? at(mstk, ttop - 1) : 0.0;
        const double tpl = minus ? pt : xw, tmi = minus ? xw : pt;
        const double mpl = minus ? pm : mw, mmi = minus ? mw : pm;
        const double a = mu_block_sum((tpl - tmi) * mpl, red), b = mu_block_sum((tpl - tmi) * mmi, red);
        snew = (a >= 0.0 && b >= 0.0) ? 1 : 0;
      }
      fls[f] = snew;
      if (fj[f] > 1) --ttop;
      rn = fln[f];
      rs = fls[f];
      --top;
    }
    if (minus) {
      tm = xw;
      mm = mw;
    } else {
      tp = xw;
      mp = mw;
    }
    --ptop;
    const double cand = own ? at(pstk, ptop) : 0.0;
    f_cpy = false;
    if (rs == 1 && bern(fmin(1.0, (double)rn / (double)n))) {
      xnew = cand;
      f_cpy = true;
    }
    n += rn;
    int s = 0;
    if (rs == 1) {
      const double a = mu_block_sum((tp - tm) * mp, red), b = mu_block_sum((tp - tm) * mm, red);
      s = (a >= 0.0 && b >= 0.0) ? 1 : 0;
    }
    ++j;
    if (s != 1 || j >= kNutsMaxDepth) break;
  }
  if (f_cpy) *x = xnew;
  return f_cpy;
}

// One block of kMuThreads per chain: the schedule's move of the mu block.
//   ESlice [mu]  per component (cbits/augur_mcmclib.h:439-485, Core/CgESlice.hs):
//                nu ~ MvNormal(mu0, cov0), log y = like(mu_k) + log u, angle
//                bracket shrunk until like(prop) > log y; one wave per component.
//   MWG [mu] ~ MvNormal(mu[k], P) per component (:178-193,
//                Core/CgUserProp.hs): mu' = mu + chol(P) z, accepted on the
//                full conditional; with ref_compat the reference's Hastings
//                term 1/2 z.z (both proposal densities taken at mu).
//   HMC [mu]     one trajectory over all components (:343-424), the
//                reference's jitter, one Metropolis test.
//   RSlice [mu]  reflective slice over all components (:500-543): slice level
//                log p(mu) - Exp(1), round(sim_len / eps) steps of eps, the
//                momentum reflected on the gradient outside the slice, the end
//                point kept if it lies in the slice.
// Draws: Philox (seed, chain) at tag TAG_GMM_MU_MOVE (normals: site k D + i),
// + 1 (HMC jitter), + 2 (uniforms: site k, sub 1 the slice level / accept,
// sub 2 + j the j-th angle; block moves site 0).
__global__ __launch_bounds__(kMuThreads) void k_gmm_mu_move(GmmDev m, uint32_t seed, uint32_t sweep) {
  const int c = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6, d = m.d, K = m.k;
  __shared__ double th[kMaxComp * kMaxDim], red[kMuWaves];
  __shared__ int acc_s;
  double* mu = m.mu64 + (size_t)c * K * d;
  const Stream sn{seed, (uint32_t)c, sweep, TAG_GMM_MU_MOVE};
  const Stream su{seed, (uint32_t)c, sweep, TAG_GMM_MU_MOVE + 2};
  if (t == 0) acc_s = 0;
  __syncthreads();
  if (m.mu_mode == GMM_MU_ESLICE || m.mu_mode == GMM_MU_MWG) {
    int acc = 0;
    for (int k = wv; k < K; k += kMuWaves) {
      const bool on = lane < d;
      const double x = on ? mu[k * d + lane] : 0.0;
      const double z = on ? normal_at(sn, (uint32_t)(k * d + lane), 0) : 0.0;
      if (m.mu_mode == GMM_MU_ESLICE) {
        const double nu = mu_lower_times(m.chol0, d, z) + (on ? m.mu0[lane] : 0.0);
        const double mu0 = on ? m.mu0[lane] : 0.0;
        const double log_y = mu_comp_ll(m, c, k, x, false) + log(u01(su.draw((uint32_t)k, 1).x));
        double ang = kTwoPi * u01(su.draw((uint32_t)k, 2).x);
        double amin = ang - kTwoPi, amax = ang;
        double out = x;
        for (int j = 0; j < kESliceMaxShrinks; ++j) {
          if (amax - amin < 1e-8) break;  // keep the current point
          const double pr = (x - mu0) * cos(ang) + (nu - mu0) * sin(ang) + mu0;
          if (mu_comp_ll(m, c, k, pr, false) > log_y) {
            out = pr;
            ++acc;
            break;
          }
          if (ang < 0.0) amin = ang;
          else amax = ang;
          ang = amin + (amax - amin) * u01(su.draw((uint32_t)k, (uint32_t)(3 + j)).x);
        }
        if (on) mu[k * d + lane] = out;
      } else {
        const double pr = x + mu_lower_times(m.cholp, d, z);
        double corr = 0.0;
        if (m.ref_compat) {
          double zz = z * z;
          for (int o = 32; o > 0; o >>= 1) zz += __shfl_xor(zz, o);
          corr = 0.5 * zz;
        }
        const double lu = log(u01(su.draw((uint32_t)k, 1).x));
        const double diff = mu_comp_ll(m, c, k, pr, true) - mu_comp_ll(m, c, k, x, true);
        if (lu < diff + corr) {
          if (on) mu[k * d + lane] = pr;
          ++acc;
        }
      }
    }
    if (lane == 0) atomicAdd(&acc_s, acc);
    __syncthreads();
    if (t == 0) {
      m.macc[2 * c] += (double)acc_s;
      m.macc[2 * c + 1] += (double)K;
    }
    return;
  }
  // ---- block moves: HMC, RSlice ----
  const bool own = t < K * d;
  const double x0 = own ? mu[t] : 0.0;
  double x = x0;
  if (own) th[t] = x0;
  const double p0 = own ? normal_at(sn, (uint32_t)t, 0) : 0.0;
  double g, lt;
  __syncthreads();
  mu_block_eval(m, c, th, &g, &lt);
  const double ll0 = mu_block_sum(lt, red);
  bool accept = false;
  if (m.mu_mode == GMM_MU_NUTS) {
    accept = mu_nuts(m, c, th, red, seed, sweep, &x);
  } else if (m.mu_mode == GMM_MU_HMC) {
    double step;
    int steps;
    hmc_jitter(Stream{seed, (uint32_t)c, sweep, TAG_GMM_MU_MOVE + 1}, m.mu_sim_len, m.mu_step, m.mu_fixed_steps,
               &step, &steps);
    double p = p0 + 0.5 * step * g;
    for (int l = 0; l < steps; ++l) {
      x += step * p;
      __syncthreads();
      if (own) th[t] = x;
      __syncthreads();
      mu_block_eval(m, c, th, &g, &lt);
      if (l != steps - 1) p += step * g;
    }
    p += 0.5 * step * g;
    const double ll1 = mu_block_sum(lt, red);
    const double kin = mu_block_sum(0.5 * p0 * p0 - 0.5 * p * p, red);
    const double lu = log(u01(su.draw(0, 1).x));
    accept = lu < (ll1 - ll0) + kin;
  } else {  // GMM_MU_RSLICE
    const double level = ll0 + log(u01(su.draw(0, 1).x));  // ll - Exp(1)
    const double rs = round(m.mu_sim_len / m.mu_step);
    const int steps = rs < 1.0 ? 1 : (int)rs;
    double p = p0;
    bool out_of_slice = false;
    for (int l = 0; l < steps; ++l) {
      x += m.mu_step * p;
      __syncthreads();
      if (own) th[t] = x;
      __syncthreads();
      mu_block_eval(m, c, th, &g, &lt);
      out_of_slice = mu_block_sum(lt, red) < level;
      if (out_of_slice) {
        const double pg = mu_block_sum(p * g, red), gg = mu_block_sum(g * g, red);
        p -= 2.0 * (pg / gg) * g;
      }
    }
    accept = !out_of_slice;
  }
  if (own && accept) mu[t] = x;
  if (t == 0) {
    m.macc[2 * c] += accept ? 1.0 : 0.0;
    m.macc[2 * c + 1] += 1.0;
  }
}

// Predictive log-likelihood of test points (experiments/hmvgmm/hmvgmm.py:
// 143-178): sum_n log sum_k pis_k N(y_n | mu_k, cov_k), FP64.  The first
// kernel forms each (component, chain)'s L^-1 and log pis_k - log det L -
// D/2 log 2 pi; the second one log-sum-exp per point, one partial per block.
__global__ __launch_bounds__(kConjLanes) void k_gmm_pred_prep(GmmDev m, double* __restrict__ pb) {
  const int k = blockIdx.x, c = blockIdx.y, d = m.d, tid = threadIdx.x;
  __shared__ double L[kMaxDim * kMaxDim], tmp[kMaxDim * kMaxDim];
  const size_t ck = (size_t)c * m.k + k;
  for (int e = tid; e < d * d; e += kConjLanes) L[e] = m.cov64[ck * d * d + e];
  __syncthreads();
  chol(L, d);
  double ldet = 0.0;
  for (int i = 0; i < d; ++i) ldet += log(L[i * d + i]);
  lower_trmi(L, d, tmp);
  double* o = pb + ck * (d * d + 1);
  for (int e = tid; e < d * d; e += kConjLanes) o[e] = L[e];
  if (tid == 0) o[d * d] = log(m.pis64[ck]) - ldet - 0.5 * d * log(6.283185307179586476925);
}

constexpr int kGmmPredBlocks = 256;
__global__ __launch_bounds__(256) void k_gmm_predll(GmmDev m, const double* __restrict__ yt, int n,
                                                    const double* __restrict__ pb, double* __restrict__ part) {
  __shared__ double red[kMuWaves];
  const int c = blockIdx.y, d = m.d, K = m.k;
  double acc = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double* y = yt + (size_t)i * d;
    double mx = -INFINITY, l[kMaxComp];
    for (int k = 0; k < K; ++k) {
      const size_t ck = (size_t)c * K + k;
      const double* Li = pb + ck * (d * d + 1);
      const double* mu = m.mu64 + ck * d;
      double q = 0.0;
      for (int a = 0; a < d; ++a) {
        double t = 0.0;
        for (int b = 0; b <= a; ++b) t += Li[a * d + b] * (y[b] - mu[b]);
        q += t * t;
      }
      l[k] = Li[d * d] - 0.5 * q;
      mx = fmax(mx, l[k]);
    }
    double se = 0.0;
    for (int k = 0; k < K; ++k) se += exp(l[k] - mx);
    acc += mx + log(se);
  }
  // block sum, waves in order
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[(size_t)c * gridDim.x + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Tables from the current parameters (after set_pt).
__global__ __launch_bounds__(kConjLanes) void k_gmm_prec0(GmmDev m) {
  double a[kMaxDim];
  rload(m.cov0, m.d, a);
  rchol(a, m.d);
  rchol_invert(a, m.d);
  const int lane = threadIdx.x;
#pragma unroll
  for (int j = 0; j < kMaxDim; ++j)
    if (lane < m.d && j < m.d) m.prec0[lane * m.d + j] = a[j];
}

__global__ __launch_bounds__(kConjLanes) void k_gmm_tables(GmmDev m) {
  const int k = blockIdx.x, c = blockIdx.y;
  double cv[kMaxDim];
  rload(m.cov64 + ((size_t)c * m.k + k) * m.d * m.d, m.d, cv);
  build_table_r(m, c, k, cv);
}

// DM: padded dimension, KM: maximum number of components held in registers.
typedef __attribute__((address_space(4))) float kconst_f32;  // constant address space

template <int DM, int KM>
__global__ __launch_bounds__(kGmmThreads) void k_gmm_z(GmmDev m, uint32_t seed, uint32_t sweep,
                                                      uint32_t chain0) {
  __shared__ float lpi[KM];
  constexpr int TS = tri(DM) + DM + 1;
  const int K = m.k, d = m.d, c = blockIdx.y, tid = threadIdx.x;
  // the table is uniform over the block: read through the scalar cache
  const float* __restrict__ gtab = m.tab + (size_t)c * K * TS;
  if (tid == 0) {
    const double* g = m.gpi + (size_t)c * K;
    double tot = 0.0;
    for (int i = 0; i < K; ++i) tot += g[i];
    for (int i = 0; i < K; ++i) {
      lpi[i] = (float)log(g[i] / tot);
      if (blockIdx.x == 0) m.pis64[(size_t)c * K + i] = g[i] / tot;
    }
  }
  __syncthreads();
  const Stream zs{seed, chain0 + (uint32_t)c, sweep, TAG_GMM_Z};
  const int stride = gridDim.x * kGmmThreads;
  for (int n = blockIdx.x * kGmmThreads + tid; n < m.n; n += stride) {
    float yv[DM];
    const float* yr = m.y + (size_t)n * d;
    if (d == DM && (DM % 4) == 0) {
#pragma unroll
      for (int i = 0; i < DM; i += 4) {
        const float4 q = *(const float4*)(yr + i);
        yv[i] = q.x; yv[i + 1] = q.y; yv[i + 2] = q.z; yv[i + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < DM; ++i) yv[i] = i < d ? yr[i] : 0.0f;
    }
    float lg[KM];
    float mx = -INFINITY;
    // The table is the same for every point, so the compiler would hoist its
    // per-component addresses and values out of the point loop; they do not
    // fit in scalar registers across it and were spilled to vector lanes
    // (about 600 v_writelane / v_readlane per point).  An empty asm statement
    // makes the pointer opaque once per point, so the table is read through
    // the scalar cache inside the loop instead; the pointer is in the
    // constant address space (the table is not written during the kernel),
    // which keeps those reads scalar loads.
    const kconst_f32* gt = (const kconst_f32*)gtab;
    asm volatile("" : "+s"(gt));
#pragma unroll
    for (int kk = 0; kk < KM; ++kk) {
      if (kk < K) {
        const kconst_f32* t = gt + kk * TS;
        float dv[DM];
#pragma unroll
        for (int i = 0; i < DM; ++i) dv[i] = yv[i] - t[tri(DM) + i];
        float q = 0.0f;
#pragma unroll
        for (int i = 0; i < DM; ++i) {
          float w = 0.0f;
#pragma unroll
          for (int j = 0; j <= i; ++j) w = fmaf(t[tri(i) + j], dv[j], w);
          q = fmaf(w, w, q);
        }
        const float l = lpi[kk] + t[tri(DM) + DM] - 0.5f * q;
        lg[kk] = l;
        mx = fmaxf(mx, l);
      }
    }
    float tot = 0.0f;
#pragma unroll
    for (int kk = 0; kk < KM; ++kk)
      if (kk < K) {
        lg[kk] = expf(lg[kk] - mx);
        tot += lg[kk];
      }
    const float u = u01f(zs.draw((uint32_t)n, 0).x);
    const float tt = u * tot;
    int zk = K - 1;
    float run = 0.0f;
#pragma unroll
    for (int kk = 0; kk < KM; ++kk)
      if (kk < K - 1) {
        run += lg[kk];
        if (zk == K - 1 && tt < run) zk = kk;
      }
    m.z[(size_t)c * m.n + n] = zk;
  }
}

typedef __bf16 gmm_bf16x8 __attribute__((ext_vector_type(8)));

// V_PERMLANE16_SWAP_B32 / V_PERMLANE32_SWAP_B32 on a register pair (rows of
// 16 lanes, checked on the device with an in-bounds probe): after
// swap_rows16, x = [x.r0, y.r0, x.r2, y.r2] and y = [x.r1, y.r1, x.r3, y.r3];
// after swap_rows32, x = [x.r0, x.r1, y.r0, y.r1] and y = [x.r2, x.r3, y.r2,
// y.r3].  Vector-unit exchanges, no LDS.
__device__ inline void swap_rows16(float& x, float& y) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
  x = __uint_as_float(r[0]);
  y = __uint_as_float(r[1]);
}
__device__ inline void swap_rows32(float& x, float& y) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
  x = __uint_as_float(r[0]);
  y = __uint_as_float(r[1]);
}

// x = h + m + l, each part rounded to BF16 to nearest even by the hardware
// conversion; the residuals are exact in FP32, so the parts carry x's 24
// significand bits.
__device__ inline void gmm_split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)x;
  const float r1 = x - (float)h;
  m = (__bf16)r1;
  l = (__bf16)(r1 - (float)m);
}

constexpr size_t gmm_zm_smem_bytes(int KM) {
  return sizeof(float) * 17 * (size_t)KM + sizeof(uint4) * 96 * (size_t)KM;  // KM * 17 floats: 16-byte multiple
}

// The z step of k_gmm_z for a padded dimension of 16, with the triangular
// products on the matrix cores.  For point x and component k,
// L^-1 (x - mu_k) = L^-1 x - c_k with c_k = L^-1 mu_k (FP64 from the table,
// rounded once).  L^-1 x for 16 points at a time is a 16 x 16 x 16 product:
// L^-1 and x are each split into three BF16 parts, and the six cross
// products above 2^-16 relative (hh, hm, mh, hl, lh, mm) are summed in FP32
// by three V_MFMA_F32_16X16X32_BF16, each pairing two parts along its
// 32-deep K (A = [L_h | L_m] with B = [x_h ; x_h], A = [L_h | L_l] with
// B = [x_m ; x_h], A = [L_h | L_m] with B = [x_l ; x_m]).  The result keeps
// FP32 accuracy in L^-1 x; the subtraction of c_k adds a rounding of
// |L^-1 x| (the FP32 form subtracts mu before the product instead).  The
// log-sum-exp and the draw are those of k_gmm_z.
// Per wave: 64 points (4 tiles of 16) at a time; lane (n, kg) supplies
// point n's features 8 (kg & 1) .. + 7 to every tile's B operands, receives
// output rows 4 kg .. + 3 of point n, and owns point n of tile kg.
template <int KM>
__global__ __launch_bounds__(kGmmThreads) void k_gmm_z_mfma(GmmDev m, uint32_t seed, uint32_t sweep,
                                                           uint32_t chain0) {
  constexpr int DM = 16, TS = tri(DM) + DM + 1;
  extern __shared__ __align__(16) unsigned char smem[];
  // the small arrays first: LDS reads take a 16-bit immediate offset, so this
  // keeps almost every per-component address a constant offset from one base
  float* cvec = reinterpret_cast<float*>(smem);         // [KM][16] -c_k
  float* cst = cvec + KM * 16;                          // [KM] log pi_k + log-det term
  uint4* aop = reinterpret_cast<uint4*>(cst + KM);      // [KM][3][32] parts of L^-1 per lane
  const int K = m.k, d = m.d, c = blockIdx.y, tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6, n16 = lane & 15, kg = lane >> 4;
  const float* gtab = m.tab + (size_t)c * K * TS;
  if (tid == 0) {
    const double* g = m.gpi + (size_t)c * K;
    double tot = 0.0;
    for (int i = 0; i < K; ++i) tot += g[i];
    for (int i = 0; i < K; ++i) {
      cst[i] = (float)log(g[i] / tot) + gtab[i * TS + tri(DM) + DM];
      if (blockIdx.x == 0) m.pis64[(size_t)c * K + i] = g[i] / tot;
    }
  }
  // A operands of component k: lane (i, g) supplies 8 BF16 at K index
  // 8 g + e, the high part of L^-1[i][8 (g & 1) + e] for g < 2 and the middle
  // (first operand) or low (second operand) part for g >= 2.  Stored once
  // per part: aop[k][part][l'] for l' = i + 16 (g & 1) (32 entries per part).
  for (int e2 = tid; e2 < K * 96; e2 += kGmmThreads) {
    const int k = e2 / 96, part = (e2 >> 5) % 3, l = e2 & 31;
    const int i = l & 15, g = l >> 4;
    const float* t = gtab + k * TS;
    gmm_bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int j = 8 * g + e;
      const float a = j <= i ? t[tri(i) + j] : 0.0f;
      __bf16 h, mm, lo;
      gmm_split3(a, h, mm, lo);
      v[e] = part == 0 ? h : (part == 1 ? mm : lo);
    }
    aop[e2] = __builtin_bit_cast(uint4, v);
  }
  for (int e2 = tid; e2 < K * 16; e2 += kGmmThreads) {
    const int k = e2 >> 4, i = e2 & 15;
    const float* t = gtab + k * TS;
    double sum = 0.0;
    for (int j = 0; j <= i; ++j) sum += (double)t[tri(i) + j] * (double)t[tri(DM) + j];
    cvec[e2] = (float)(-sum);
  }
  __syncthreads();

  const Stream zs{seed, chain0 + (uint32_t)c, sweep, TAG_GMM_Z};
  const int half = kg & 1;
  const int la = (kg < 2 ? 0 : 32) + (lane & 31);       // part 0 for g < 2, part 1 for g >= 2
  const int lb = (kg < 2 ? 0 : 64) + (lane & 31);       // part 0 for g < 2, part 2 for g >= 2
  const int groups = (m.n + 63) / 64;
  for (int grp = blockIdx.x * (kGmmThreads / 64) + wv; grp < groups; grp += gridDim.x * (kGmmThreads / 64)) {
    gmm_bf16x8 b1[4], b2[4], b3[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = 64 * grp + 16 * j + n16;
      float xv[8];
      if (p < m.n && d == DM) {
        const float4* yr = reinterpret_cast<const float4*>(m.y + (size_t)p * DM + 8 * half);
        const float4 u0 = yr[0], u1 = yr[1];
        xv[0] = u0.x; xv[1] = u0.y; xv[2] = u0.z; xv[3] = u0.w;
        xv[4] = u1.x; xv[5] = u1.y; xv[6] = u1.z; xv[7] = u1.w;
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) xv[e] = (p < m.n && 8 * half + e < d) ? m.y[(size_t)p * d + 8 * half + e] : 0.0f;
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 h, mm, lo;
        gmm_split3(xv[e], h, mm, lo);
        b1[j][e] = h;
        b2[j][e] = kg < 2 ? mm : h;
        b3[j][e] = kg < 2 ? lo : mm;
      }
    }
    float lg[KM];
    float mx = -INFINITY;
#pragma unroll
    for (int kk = 0; kk < KM; ++kk) {
      if (kk < K) {
        const gmm_bf16x8 a1 = __builtin_bit_cast(gmm_bf16x8, aop[kk * 96 + la]);
        const gmm_bf16x8 a2 = __builtin_bit_cast(gmm_bf16x8, aop[kk * 96 + lb]);
        // -c_k of this lane's 4 rows: the products accumulate onto it, so the
        // accumulator ends as L^-1 x - c_k without a subtraction per value
        const gmm_f32x4 cneg = *reinterpret_cast<const gmm_f32x4*>(cvec + kk * 16 + 4 * kg);
        // qp[j]: this lane's 4 rows of ||L^-1 x - c_k||^2 for point n of tile j
        float qp[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          gmm_f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1[j], cneg, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b2[j], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b3[j], acc, 0, 0, 0);
          qp[j] = fmaf(acc[0], acc[0], fmaf(acc[1], acc[1], fmaf(acc[2], acc[2], acc[3] * acc[3])));
        }
        // transpose-reduce over the 4 lanes of point n (row kg = 0..3 of 16
        // lanes) with the row-swapping permutes: after the two 16-row swaps
        // row kg holds tiles kg & 1 (k0) and 2 + (kg & 1) (k1) summed over
        // its row pair, after the 32-row swap tile kg over all four rows
        float t0 = qp[0], t1 = qp[1], t2 = qp[2], t3 = qp[3];
        swap_rows16(t0, t1);
        swap_rows16(t2, t3);
        float k0 = t0 + t1, k1 = t2 + t3;
        swap_rows32(k0, k1);
        const float q = k0 + k1;
        const float l = cst[kk] - 0.5f * q;
        lg[kk] = l;
        mx = fmaxf(mx, l);
      }
    }
    const int n = 64 * grp + 16 * kg + n16;  // the point this lane owns
    if (n >= m.n) continue;
    float tot = 0.0f;
#pragma unroll
    for (int kk = 0; kk < KM; ++kk)
      if (kk < K) {
        lg[kk] = expf(lg[kk] - mx);
        tot += lg[kk];
      }
    const float u = u01f(zs.draw((uint32_t)n, 0).x);
    const float tt = u * tot;
    int zk = K - 1;
    float run = 0.0f;
#pragma unroll
    for (int kk = 0; kk < KM; ++kk)
      if (kk < K - 1) {
        run += lg[kk];
        if (zk == K - 1 && tt < run) zk = kk;
      }
    m.z[(size_t)c * m.n + n] = zk;
  }
}

// Sufficient statistics of z around the table's mu (the mu of the z step):
// per (chain, block) and component, count, sum e and sum e e^T (packed lower)
// with e = y - mu_z.  Each wave takes 64 points and forms
// S[k][q] += sum_p [z_p = k] F_p[q], F_p = (1, e_p, e_p e_p^T), as one-hot
// products on the matrix core (V_MFMA_F32_16X16X32_BF16: A = [z_p = k],
// 16 components x 32 points, exact in BF16; B = F, 32 points x 16 entries,
// each entry split into three BF16 parts) from a per-wave LDS slab of the
// residuals.  No atomics: FP32 per block in a fixed order, FP64 across
// blocks in k_gmm_stats_sum.
template <int DM, int KM>
__global__ __launch_bounds__(kGmmThreads) void k_gmm_stats_t(GmmDev m) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int TS = tri(DM) + DM + 1;
  constexpr int CB = (KM + 15) / 16;                // 16-component blocks
  constexpr int EB = (1 + DM + tri(DM) + 15) / 16;  // 16-entry blocks
  constexpr int XS = DM + 3;                        // slab row: e (DM), 1, 0, pad
  const int K = m.k, d = m.d, S = m.S, c = blockIdx.y, tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  float* slab = (float*)smem + (size_t)wv * 64 * XS;            // per wave: 64 x XS
  int* zsl = (int*)((float*)smem + 4 * 64 * XS) + wv * 64;      // per wave: 64
  const float* __restrict__ gtab = m.tab + (size_t)c * K * TS;
  // entry q = 16 eb + (lane % 16) this lane supplies: q = 0 -> 1 * 1,
  // q = 1 .. d -> e[q - 1] * 1, then e[a] e[b] at q = 1 + d + tri(a) + b
  // (b <= a), beyond S -> 0 (slab column DM + 1)
  int ia[EB], ib[EB];
#pragma unroll
  for (int eb = 0; eb < EB; ++eb) {
    const int q = 16 * eb + (lane & 15);
    if (q == 0) {
      ia[eb] = DM; ib[eb] = DM;
    } else if (q <= d) {
      ia[eb] = q - 1; ib[eb] = DM;
    } else if (q < S) {
      const int r = q - 1 - d;
      int a = 0;
      while (tri(a + 1) <= r) ++a;
      ia[eb] = a; ib[eb] = r - tri(a);
    } else {
      ia[eb] = DM + 1; ib[eb] = DM + 1;
    }
  }
  gmm_f32x4 acc[CB][EB];
#pragma unroll
  for (int cb = 0; cb < CB; ++cb)
#pragma unroll
    for (int eb = 0; eb < EB; ++eb) acc[cb][eb] = gmm_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int stride = gridDim.x * kGmmThreads;
  // a wave takes 64 consecutive points per step: the trip count is uniform
  // over the wave, so every lane reaches the product
  for (int n0 = blockIdx.x * kGmmThreads + wv * 64; n0 < m.n; n0 += stride) {
    const int n = n0 + lane;
    const bool live = n < m.n;
    const int zk = live ? m.z[(size_t)c * m.n + n] : -1;
    float* row = slab + lane * XS;
    if (live) {
      const float* yr = m.y + (size_t)n * d;
      const float* mu = m.tab + ((size_t)c * K + zk) * TS + tri(DM);
      if (d == DM && (DM % 4) == 0) {
#pragma unroll
        for (int i = 0; i < DM; i += 4) {
          const float4 q = *(const float4*)(yr + i);
          row[i] = q.x - mu[i];
          row[i + 1] = q.y - mu[i + 1];
          row[i + 2] = q.z - mu[i + 2];
          row[i + 3] = q.w - mu[i + 3];
        }
      } else {
#pragma unroll
        for (int i = 0; i < DM; ++i) row[i] = i < d ? yr[i] - mu[i] : 0.0f;
      }
    } else {
#pragma unroll
      for (int i = 0; i < DM; ++i) row[i] = 0.0f;
    }
    row[DM] = 1.0f;
    row[DM + 1] = 0.0f;
    zsl[lane] = zk;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // 32 points per product (V_MFMA_F32_16X16X32_BF16): lane (l % 16, g =
    // l / 16) supplies points 32 h + 8 g .. + 7 of one component row (one-hot,
    // exact in BF16) and of one entry column, the entry split into three
    // BF16 parts whose products with the one-hot row are exact, so the sums
    // are FP32 sums of the FP32 entries
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int p0 = 32 * h + 8 * (lane >> 4);
      gmm_bf16x8 a[CB];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int zp = zsl[p0 + e];
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) a[cb][e] = zp == 16 * cb + (lane & 15) ? (__bf16)1.0f : (__bf16)0.0f;
      }
#pragma unroll
      for (int eb = 0; eb < EB; ++eb) {
        gmm_bf16x8 bh, bm, bl;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const float* rp = slab + (p0 + e) * XS;
          __bf16 x0, x1, x2;
          gmm_split3(rp[ia[eb]] * rp[ib[eb]], x0, x1, x2);
          bh[e] = x0;
          bm[e] = x1;
          bl[e] = x2;
        }
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
          acc[cb][eb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[cb], bh, acc[cb][eb], 0, 0, 0);
          acc[cb][eb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[cb], bm, acc[cb][eb], 0, 0, 0);
          acc[cb][eb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[cb], bl, acc[cb][eb], 0, 0, 0);
        }
      }
    }
    // the slab is rewritten only after every lane has read it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // lane l holds S[16 cb + 4 (l / 16) + r][16 eb + l % 16]; the waves are
  // added in order (deterministic), one row per block
  __syncthreads();
  float* red = (float*)smem;
  for (int w = 0; w < kGmmThreads / 64; ++w) {
    if (wv == w) {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int eb = 0; eb < EB; ++eb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int comp = 16 * cb + 4 * (lane >> 4) + r, q = 16 * eb + (lane & 15);
            if (comp < K && q < S) red[comp * S + q] = (w == 0 ? 0.0f : red[comp * S + q]) + acc[cb][eb][r];
          }
    }
    __syncthreads();
  }
  float* dst = m.part + ((size_t)c * m.Gs + blockIdx.x) * K * S;
  for (int i = tid; i < K * S; i += kGmmThreads) dst[i] = red[i];
}

// Sorted form of the statistics for padded dimension 16.  The one-hot
// products of k_gmm_stats_t spend 31 of 32 matrix-core products on zeros and
// most of their vector work splitting e e^T into BF16 parts.  Here each block
// takes tiles of NT points (one per thread: its y row, read coalesced, and
// its z, both requested one tile ahead), sorts them by component with a
// counting sort from wave ballots (position = component, then wave, then
// lane: the same order, so the same FP32 sums, on every run), and writes the
// residuals e = y - mu_z in that order to LDS.  Each wave then accumulates
// its components' segments with FP32 matrix-core outer products,
// V_MFMA_F32_16X16X4_F32 on four points at a time: lane l supplies e_p[l %
// 16] of point p = the (l / 16)-th of the four as both A (16 dims x 4
// points) and B (4 points x 16 dims), adding sum e e^T, and with B = 1
// (valid points) sum e.  Wave w of NW owns components w, w + NW, ...: its sums stay
// in accumulation registers over the block's tiles, one row of partials per
// block and component as k_gmm_stats_t writes them.  Grid (Gs, chains), one
// resident round: block (x, c) is dispatched at x + Gs c, on XCD x mod 8 for
// every chain when Gs is a multiple of 8, so a tile's y rows read by one
// chain are in that XCD's L2 for the others.
template <int KM, int NT>
__global__ __launch_bounds__(NT) void k_gmm_stats_s(GmmDev m) {
  constexpr int DM = 16, TS = tri(DM) + DM + 1, TP = NT, NW = NT / 64;
  constexpr int KPW = (KM + NW - 1) / NW;  // components per wave
  __shared__ __align__(16) float es[TP * DM];  // the tile's residuals, sorted by component
  __shared__ float mus[KM * DM];               // the table's mu of every component
  __shared__ int wcnt[NW][KM];                 // per wave and component: count, then offset
  __shared__ int koff[KM + 1];                 // component segments in es
  const int K = m.k, d = m.d, S = m.S, c = blockIdx.y, tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6, dim = lane & 15, quad = lane >> 4;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int i = tid; i < K * DM; i += NT) {
    const int k = i / DM, j = i % DM;
    mus[i] = j < d ? m.tab[((size_t)c * K + k) * TS + tri(DM) + j] : 0.0f;
  }
  gmm_f32x4 see[KPW], se[KPW];
  float cnt[KPW];
#pragma unroll
  for (int j = 0; j < KPW; ++j) {
    see[j] = gmm_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    se[j] = gmm_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    cnt[j] = 0.0f;
  }
  const int32_t* __restrict__ zc = m.z + (size_t)c * m.n;
  const int tiles = (m.n + TP - 1) / TP;
  // this thread's point of tile t: z (-1 past the end) and its y row
  float yr[DM];
  int zk = -1;
  auto fetch = [&](int t, float (&row)[DM], int& zo) {
    const int p = t * TP + tid;
    zo = -1;
#pragma unroll
    for (int i = 0; i < DM; ++i) row[i] = 0.0f;
    if (t < tiles && p < m.n) {
      zo = zc[p];
      const float* yp = m.y + (size_t)p * d;
      if (d == DM) {
#pragma unroll
        for (int i = 0; i < DM; i += 4) {
          const float4 q = *(const float4*)(yp + i);
          row[i] = q.x; row[i + 1] = q.y; row[i + 2] = q.z; row[i + 3] = q.w;
        }
      } else {
#pragma unroll
        for (int i = 0; i < DM; ++i) row[i] = i < d ? yp[i] : 0.0f;
      }
    }
  };
  fetch(blockIdx.x, yr, zk);
  __syncthreads();  // mus
  for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
    // counting sort: rank among this wave's lanes of the same component (a
    // value outside 0 .. K - 1 matches none and counts nowhere)
    int myrank = 0;
#pragma unroll 8
    for (int k = 0; k < KM; ++k) {
      if (k < K) {
        const uint64_t b = __ballot(zk == k);
        if (zk == k) myrank = __popcll(b & below);
        if (lane == 0) wcnt[wv][k] = __popcll(b);
      }
    }
    __syncthreads();
    if (wv == 0) {
      // component k = lane (KM <= 64): totals, an exclusive scan over the
      // components, then the waves' offsets
      const int k = lane;
      int tot = 0;
      if (k < K)
#pragma unroll
        for (int w = 0; w < NW; ++w) tot += wcnt[w][k];
      int incl = tot;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
      }
      if (k < K) {
        int off = incl - tot;
        koff[k] = off;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
          const int cw = wcnt[w][k];
          wcnt[w][k] = off;
          off += cw;
        }
      }
      if (k == K - 1) koff[K] = incl;
    }
    __syncthreads();
    if (zk >= 0 && zk < K) {
      float* dst = es + (wcnt[wv][zk] + myrank) * DM;
      const float* mk = mus + zk * DM;
#pragma unroll
      for (int i = 0; i < DM; i += 4) {
        float4 v;
        v.x = yr[i] - mk[i];
        v.y = yr[i + 1] - mk[i + 1];
        v.z = yr[i + 2] - mk[i + 2];
        v.w = yr[i + 3] - mk[i + 3];
        *(float4*)(dst + i) = v;
      }
    }
    // the next tile's point, requested before this tile's products
    fetch(t + gridDim.x, yr, zk);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < KPW; ++j) {
      const int k = wv + NW * j;
      if (k < K) {  // uniform over the wave
        const int beg = koff[k], end = koff[k + 1];
        cnt[j] += (float)(end - beg);
        for (int i0 = beg; i0 < end; i0 += 4) {
          const int pt = i0 + quad;
          const bool ok = pt < end;
          const float e = ok ? es[min(pt, TP - 1) * DM + dim] : 0.0f;
          see[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(e, e, see[j], 0, 0, 0);
          se[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(e, ok ? 1.0f : 0.0f, se[j], 0, 0, 0);
        }
      }
    }
    __syncthreads();  // es, wcnt and koff are rewritten by the next tile
  }
  // lane l holds sum over points of e[4 (l / 16) + r] e[l % 16] in see[j][r]
  // and sum e[4 (l / 16) + r] (every column alike) in se[j][r]
  float* dst = m.part + ((size_t)c * m.Gs + blockIdx.x) * K * S;
#pragma unroll
  for (int j = 0; j < KPW; ++j) {
    const int k = wv + NW * j;
    if (k >= K) continue;
    float* row = dst + (size_t)k * S;
    if (lane == 0) row[0] = cnt[j];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int a = 4 * quad + r, b = dim;
      if (a < d && b == 0) row[1 + a] = se[j][r];
      if (a < d && b <= a) row[1 + d + tri(a) + b] = see[j][r];
    }
  }
}

// Dynamic LDS of k_gmm_stats_t: the per-wave slabs, reused for the block's
// statistics at the end.
inline size_t gmm_stats_smem_bytes(int DM, int KM, int S) {
  const size_t slabs = (size_t)4 * 64 * (DM + 3) + 4 * 64;
  const size_t stats = (size_t)KM * S;
  return sizeof(float) * (slabs > stats ? slabs : stats);
}

// Data part of the log joint: sum_n log pis_z + mvnormal_ll(y_n; mu_z, cov_z),
// FP32 per point from the tables, FP64 accumulation.  Grid (G, chains): one
// partial per block in part[c G + g], summed in order by k_gmm_ll_sum.
template <int DM>
__global__ __launch_bounds__(kGmmThreads) void k_gmm_ll(GmmDev m, double* part) {
  constexpr int TS = tri(DM) + DM + 1;
  const int c = blockIdx.y, tid = threadIdx.x, d = m.d, K = m.k;
  __shared__ double red[kGmmThreads];
  double accum = 0.0;
  for (int n = blockIdx.x * kGmmThreads + tid; n < m.n; n += gridDim.x * kGmmThreads) {
    const int zk = m.z[(size_t)c * m.n + n];
    const float* t = m.tab + ((size_t)c * K + zk) * m.TS;
    float dv[DM];
    for (int i = 0; i < DM; ++i) dv[i] = (i < d ? m.y[(size_t)n * d + i] : 0.0f) - t[tri(DM) + i];
    float q = 0.0f;
    for (int i = 0; i < DM; ++i) {
      float w = 0.0f;
      for (int j = 0; j <= i; ++j) w = fmaf(t[tri(i) + j], dv[j], w);
      q = fmaf(w, w, q);
    }
    accum += log(m.pis64[(size_t)c * K + zk]) + (double)t[tri(DM) + DM] - 0.5 * (double)q -
             0.5 * d * log(2.0 * 3.14159265358979323846);
  }
  red[tid] = accum;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int i = 0; i < kGmmThreads; ++i) s += red[i];
    part[(size_t)c * gridDim.x + blockIdx.x] = s;
  }
  (void)TS;
}

// Prior part of the log density per (component k, chain c), FP64:
// MvNormal(mu_k | mu0, cov0) (cbits/augur_dist.c:548-567) and, when cov is
// sampled, InvWishart(cov_k | df, scale0) (cbits/augur_dist.c:679-711, with
// the reference's log Gamma_d, cbits/augur_math.c:80-86).  out[c K + k].
__global__ __launch_bounds__(kConjLanes) void k_gmm_prior_ll(GmmDev m, double* out) {
  const int k = blockIdx.x, c = blockIdx.y, tid = threadIdx.x, d = m.d;
  __shared__ double a[kMaxDim * kMaxDim], tmp[kMaxDim * kMaxDim], v[kMaxDim];
  __shared__ double det_s;
  for (int i = tid; i < d * d; i += kConjLanes) a[i] = m.cov0[i];
  for (int i = tid; i < d; i += kConjLanes) v[i] = m.mu64[((size_t)c * m.k + k) * d + i] - m.mu0[i];
  __syncthreads();
  chol(a, d);
  if (tid == 0) {
    double det = 1.0;
    for (int i = 0; i < d; ++i) det *= a[i * d + i];
    det_s = det * det;
  }
  lower_trmi(a, d, tmp);
  __syncthreads();
  double ll = 0.0;
  if (tid == 0) {
    double q = 0.0;
    for (int i = 0; i < d; ++i) {
      double w = 0.0;
      for (int j = 0; j <= i; ++j) w += a[i * d + j] * v[j];
      q += w * w;
    }
    ll = -0.5 * (double)d * log(2.0 * 3.14159265358979323846) - 0.5 * log(det_s) - 0.5 * q;
  }
  if (m.sample_cov) {
    __syncthreads();
    for (int i = tid; i < d * d; i += kConjLanes) a[i] = m.scale0[i];
    __syncthreads();
    chol(a, d);
    double det_psi = 1.0;
    if (tid == 0) {
      for (int i = 0; i < d; ++i) det_psi *= a[i * d + i];
      det_psi *= det_psi;
    }
    __syncthreads();
    const double* cv = m.cov64 + ((size_t)c * m.k + k) * d * d;
    for (int i = tid; i < d * d; i += kConjLanes) a[i] = cv[i];
    __syncthreads();
    chol(a, d);
    double det_x = 1.0;
    if (tid == 0) {
      for (int i = 0; i < d; ++i) det_x *= a[i * d + i];
      det_x *= det_x;
    }
    chol_invert(a, d, tmp);
    __syncthreads();
    if (tid == 0) {
      double tr = 0.0;
      for (int i = 0; i < d; ++i) {
        double t = 0.0;
        for (int j = 0; j < d; ++j) t += m.scale0[i * d + j] * a[j * d + i];
        tr += t;
      }
      double mlg = (d * (d - 1) * 0.25) * log(3.1415);
      for (int i = 1; i < d + 1; ++i) mlg += lgamma(m.df / 2.0 - (i - 1.0) / 2.0);
      const double cst = -m.df * d / 2.0 * log(2.0) - mlg;
      ll += cst - (m.df + d + 1.0) / 2.0 * log(det_x) + m.df / 2.0 * log(det_psi) - 0.5 * tr;
    }
  }
  if (tid == 0) out[(size_t)c * m.k + k] = ll;
}

// out[c] = data partials (in order) + component priors + Dirichlet(pis |
// alpha) without its normaliser (cbits/augur_dist.c:488-496) when pis is
// sampled.
__global__ void k_gmm_ll_sum(GmmDev m, const double* part, const double* prior, double* out) {
  if (threadIdx.x != 0) return;
  const int c = blockIdx.x;
  double s = 0.0;
  for (int g = 0; g < m.G; ++g) s += part[(size_t)c * m.G + g];
  for (int k = 0; k < m.k; ++k) s += prior[(size_t)c * m.k + k];
  if (m.sample_pi)
    for (int k = 0; k < m.k; ++k) s += (m.alpha[k] - 1.0) * log(m.pis64[(size_t)c * m.k + k]);
  out[c] = s;
}

static int pad_dim(int d) { return d <= 4 ? 4 : (d <= 8 ? 8 : 16); }
static int pad_comp(int k) { return k <= 8 ? 8 : (k <= 16 ? 16 : (k <= 32 ? 32 : 64)); }

typedef void (*ZKernel)(GmmDev, uint32_t, uint32_t, uint32_t);
typedef void (*StatsKernel)(GmmDev);

// Threads (and points per tile) of the sorted statistics kernel: 512 gave
// 0.727 ms a GMM sweep at N = 1M, K = 32, 4 chains, against 0.741 ms for
// 256 and for 1024 (same box).
constexpr int kStatsThreads = 512;
static StatsKernel k_stats_s(int KM) {
  return KM == 8 ? k_gmm_stats_s<8, kStatsThreads> : KM == 16 ? k_gmm_stats_s<16, kStatsThreads>
                                                   : k_gmm_stats_s<32, kStatsThreads>;
}

static ZKernel pick_z(int DM, int KM) {
#define AUGUR_Z(D_, K_) if (DM == D_ && KM == K_) return k_gmm_z<D_, K_>;
  AUGUR_Z(4, 8) AUGUR_Z(4, 16) AUGUR_Z(4, 32) AUGUR_Z(4, 64)
  AUGUR_Z(8, 8) AUGUR_Z(8, 16) AUGUR_Z(8, 32) AUGUR_Z(8, 64)
  AUGUR_Z(16, 8) AUGUR_Z(16, 16) AUGUR_Z(16, 32) AUGUR_Z(16, 64)
#undef AUGUR_Z
  fail(AUGUR_ERR_UNSUPPORTED, "no z kernel for padded dim %d, components %d", DM, KM);
}

// The z step of every chain of `m`: the matrix-core form for a padded
// dimension of 16 (AUGUR_GMM_Z_VALU=1 keeps the FP32 vector form), else
// k_gmm_z.
static void launch_gmm_z(const GmmDev& m, int chains, uint32_t seed, uint32_t sweep, uint32_t chain0,
                         hipStream_t stream, const std::pair<hipEvent_t, hipEvent_t>* ev) {
  const int KM = pad_comp(m.k);
  ZKernel k = nullptr;
  size_t smem = 0;
  const char* env = getenv("AUGUR_GMM_Z_VALU");
  int gx = m.G;
  if (m.DM == 16 && !(env && env[0] == '1')) {
    if (KM == 8) k = k_gmm_z_mfma<8>;
    else if (KM == 16) k = k_gmm_z_mfma<16>;
    else if (KM == 32) k = k_gmm_z_mfma<32>;
    else k = k_gmm_z_mfma<64>;
    smem = gmm_zm_smem_bytes(KM);
    allow_lds(k, smem);
    // every block builds the component tables in LDS first: one resident
    // round of blocks (three per CU over all chains) pays for it once
    gx = std::max(1, std::min(m.G, (3 * m.cus + chains - 1) / chains));
    if (const char* e = getenv("AUGUR_GMM_ZBLOCKS")) gx = std::max(1, atoi(e));
  } else {
    k = pick_z(m.DM, KM);
  }
  if (ev)
    hipExtLaunchKernelGGL(k, dim3(gx, chains), dim3(kGmmThreads), smem, stream, ev->first, ev->second, 0, m,
                          seed, sweep, chain0);
  else
    k<<<dim3(gx, chains), kGmmThreads, smem, stream>>>(m, seed, sweep, chain0);
  AUGUR_CHECK_LAUNCH();
}

static StatsKernel pick_stats(int DM, int KM) {
#define AUGUR_S(D_, K_) if (DM == D_ && KM == K_) return k_gmm_stats_t<D_, K_>;
  AUGUR_S(4, 8) AUGUR_S(4, 16) AUGUR_S(4, 32) AUGUR_S(4, 64)
  AUGUR_S(8, 8) AUGUR_S(8, 16) AUGUR_S(8, 32) AUGUR_S(8, 64)
  AUGUR_S(16, 8) AUGUR_S(16, 16) AUGUR_S(16, 32) AUGUR_S(16, 64)
#undef AUGUR_S
  fail(AUGUR_ERR_UNSUPPORTED, "no statistics kernel for padded dim %d, components %d", DM, KM);
}

struct GmmSampler final : augur_sampler {
  GmmDev m{};
  DevBuf<float> y, tab, part;
  DevBuf<double> alpha, mu0, cov0, scale0, gpi, pis64, mu64, cov64, ll, ssum, prec, prec0;
  DevBuf<double> cov0_inv, chol0, cholp, mcent, mprec, mlin, mcnt, macc, nuts;
  DevBuf<int32_t> z;
  std::vector<double> h_alpha, h_mu0, h_cov0, h_scale0;
  bool stats_valid = false;

  const char* family() const override { return "gmm"; }

  ParamView param(const std::string& name) override {
    const size_t K = m.k, D = m.d;
    ParamView v;
    if (name == "pis") v = {pis64.p, sizeof(double), K, K, true};
    else if (name == "mu") v = {mu64.p, sizeof(double), K * D, K * D, true};
    else if (name == "cov") v = {cov64.p, sizeof(double), K * D * D, K * D * D, true};
    else if (name == "z") v = {z.p, sizeof(int32_t), (size_t)m.n, (size_t)m.n, true};
    return v;
  }

  // z-step tables and FP64 precisions of the current covariances (and
  // cov0^-1, fixed, recomputed with them)
  void refresh_tables() {
    k_gmm_prec0<<<1, kConjLanes, 0, stream>>>(m);
    AUGUR_CHECK_LAUNCH();
    k_gmm_tables<<<dim3(m.k, chains), kConjLanes, 0, stream>>>(m);
    AUGUR_CHECK_LAUNCH();
  }

  // statistics of the current z around the tables' mu: the sorted form for
  // padded dimension 16 and up to 32 components unless
  // AUGUR_GMM_STATS_ONEHOT=1 (at 64 components its 16 per wave need more
  // registers than one wave per SIMD leaves)
  void launch_stats() {
    const int KM = pad_comp(m.k);
    const char* e = getenv("AUGUR_GMM_STATS_ONEHOT");
    if (m.DM == 16 && KM <= 32 && !(e && atoi(e))) {
      k_stats_s(KM)<<<dim3(m.Gs, chains), kStatsThreads, 0, stream>>>(m);
    } else {
      pick_stats(m.DM, KM)<<<dim3(m.Gs, chains), kGmmThreads, gmm_stats_smem_bytes(m.DM, KM, m.S), stream>>>(m);
    }
    AUGUR_CHECK_LAUNCH();
  }

  // the statistics summed over the statistics kernel's blocks, then the
  // conjugate draws of every (component, chain)
  void launch_conj() {
    k_gmm_stats_sum<<<dim3((m.S + 63) / 64, m.k, chains), 64 * kSumSlices, 0, stream>>>(m, ssum.p);
    AUGUR_CHECK_LAUNCH();
    if (m.mu_mode != GMM_MU_CONJ) {
      // the schedule's mu move from the statistics, before pis / cov's draws
      // use the new mu (pis does not depend on mu; cov follows mu)
      k_gmm_mu_prep<<<dim3(m.k, chains), kConjLanes, 0, stream>>>(m, ssum.p);
      AUGUR_CHECK_LAUNCH();
      k_gmm_mu_move<<<chains, kMuThreads, 0, stream>>>(m, seed, sweep);
      AUGUR_CHECK_LAUNCH();
    }
    k_gmm_conj<<<dim3(m.k, chains), kConjLanes, 0, stream>>>(m, ssum.p, seed, sweep);
    AUGUR_CHECK_LAUNCH();
  }

  void refresh_stats() {
    refresh_tables();  // the statistics are centred at the tables' (current) mu
    launch_stats();
    stats_valid = true;
  }

  void on_set(const std::string& name) override {
    if (name == "pis") {
      AUGUR_HIP(hipMemcpyAsync(gpi.p, pis64.p, pis64.bytes(), hipMemcpyDeviceToDevice, stream));
    } else if (name == "z") {
      std::vector<int32_t> hz(z.n);
      z.download(hz.data(), hz.size(), stream);
      AUGUR_HIP(hipStreamSynchronize(stream));
      for (int32_t v : hz)
        if (v < 0 || v >= m.k) fail(AUGUR_ERR_ARG, "z value %d outside [0, %d)", v, m.k);
    }
    if (name == "cov" || name == "mu") refresh_tables();
    stats_valid = false;
  }

  void run_sweeps(int nsw) override {
    for (int i = 0; i < nsw; ++i) {
      if (!stats_valid) refresh_stats();
      launch_conj();
      AUGUR_CHECK_LAUNCH();
      launch_gmm_z(m, chains, seed, sweep, 0u, stream, prof_slot());
      launch_stats();  // for the next sweep's conjugate draws
      ++sweep;
      record_traces();
    }
  }

  void log_density(double* out) override {
    // the model's log density (the reference's ll_mod): data terms per point
    // block, prior terms per component, summed per chain in a fixed order
    const int DM = m.DM;
    const dim3 grid(m.G, chains);
    double* lp = ll.p + chains;                  // chains x G data partials
    double* pr = lp + (size_t)chains * m.G;      // chains x K component priors
    if (DM == 4) k_gmm_ll<4><<<grid, kGmmThreads, 0, stream>>>(m, lp);
    else if (DM == 8) k_gmm_ll<8><<<grid, kGmmThreads, 0, stream>>>(m, lp);
    else k_gmm_ll<16><<<grid, kGmmThreads, 0, stream>>>(m, lp);
    AUGUR_CHECK_LAUNCH();
    k_gmm_prior_ll<<<dim3(m.k, chains), kConjLanes, 0, stream>>>(m, pr);
    AUGUR_CHECK_LAUNCH();
    k_gmm_ll_sum<<<chains, 64, 0, stream>>>(m, lp, pr, ll.p);
    AUGUR_CHECK_LAUNCH();
    ll.download(out, chains, stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
  }

  void stats(double* out, size_t bytes) override {
    if (bytes < sizeof(double)) fail(AUGUR_ERR_ARG, "stats buffer too small");
    out[0] = (double)sweep;
    // then, when the buffer holds them, the mu block's accepted and proposed
    // moves per chain (non-conjugate mu kernels)
    if (bytes >= sizeof(double) * (1 + 2 * (size_t)chains)) {
      if (macc.p) {
        macc.download(out + 1, 2 * (size_t)chains, stream);
        AUGUR_HIP(hipStreamSynchronize(stream));
      } else {
        for (int i = 0; i < 2 * chains; ++i) out[1 + i] = 0.0;
      }
    }
  }
};

static void gmm_validate(const augur_gmm_desc* D) {
  if (!D) fail(AUGUR_ERR_ARG, "null descriptor");
  if (D->n_points < 1 || D->dim < 1 || D->dim > kMaxDim || D->n_components < 1 ||
      D->n_components > kMaxComp || D->chains < 1)
    fail(AUGUR_ERR_ARG, "need N >= 1, 1 <= D <= %d, 1 <= K <= %d, chains >= 1", kMaxDim, kMaxComp);
  if (!D->y || !D->alpha || !D->mu0 || !D->cov0 || !D->scale0) fail(AUGUR_ERR_ARG, "null array");
  if (D->sample_cov && D->df <= D->dim - 1) fail(AUGUR_ERR_ARG, "InvWishart needs df > D - 1");
  if (D->mu_kernel < AUGUR_GMM_MU_CONJ || D->mu_kernel > AUGUR_GMM_MU_NUTS)
    fail(AUGUR_ERR_ARG, "unknown mu kernel %d", D->mu_kernel);
  if (D->mu_kernel == AUGUR_GMM_MU_NUTS && !(D->mu_step_size > 0.0)) fail(AUGUR_ERR_ARG, "NUTS [mu] needs step_size > 0");
  if (D->mu_kernel == AUGUR_GMM_MU_HMC || D->mu_kernel == AUGUR_GMM_MU_RSLICE) {
    if (!(D->mu_step_size > 0.0) || (D->mu_fixed_steps == 0 && !(D->mu_sim_len > 0.0)) ||
        D->mu_sim_len / D->mu_step_size > 1.0e5 || D->mu_fixed_steps > 100000 || D->mu_fixed_steps < -100000)
      fail(AUGUR_ERR_ARG, "mu kernel needs step_size > 0 and sim_len > 0 (or fixed_steps)");
  }
  if (D->mu_kernel == AUGUR_GMM_MU_MWG && !D->mu_prop_cov) fail(AUGUR_ERR_ARG, "MWG [mu] needs a proposal covariance");
}

// Lower Cholesky factor of a symmetric positive definite d x d matrix on the
// host (FP64), for the hyper-parameter matrices of the mu kernels.
static std::vector<double> host_chol(const double* a, int d, const char* what) {
  std::vector<double> l((size_t)d * d, 0.0);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = a[i * d + j];
      for (int k = 0; k < j; ++k) s -= l[i * d + k] * l[j * d + k];
      if (i == j) {
        if (!(s > 0.0)) fail(AUGUR_ERR_ARG, "%s is not positive definite", what);
        l[i * d + i] = sqrt(s);
      } else {
        l[i * d + j] = s / l[j * d + j];
      }
    }
  return l;
}

// Inverse of a d x d matrix from its lower Cholesky factor (host, FP64).
static std::vector<double> host_chol_inverse(const std::vector<double>& l, int d) {
  std::vector<double> inv((size_t)d * d, 0.0), e(d), y(d);
  for (int c = 0; c < d; ++c) {
    for (int i = 0; i < d; ++i) e[i] = i == c ? 1.0 : 0.0;
    for (int i = 0; i < d; ++i) {  // L y = e
      double s = e[i];
      for (int k = 0; k < i; ++k) s -= l[i * d + k] * y[k];
      y[i] = s / l[i * d + i];
    }
    for (int i = d - 1; i >= 0; --i) {  // L^T x = y
      double s = y[i];
      for (int k = i + 1; k < d; ++k) s -= l[k * d + i] * inv[k * d + c];
      inv[i * d + c] = s / l[i * d + i];
    }
  }
  return inv;
}

static void gmm_setup(GmmSampler* s, int n, int d, int k, int chains) {
  GmmDev& m = s->m;
  m.n = n; m.d = d; m.k = k;
  m.DM = pad_dim(d);
  m.TS = tri(m.DM) + m.DM + 1;
  m.S = 1 + d + tri(d);
  const int want = (n + kGmmThreads - 1) / kGmmThreads;
  m.G = std::max(1, std::min(want, std::max(1, 1024 / chains)));
  {
    int dev = 0;
    AUGUR_HIP(hipGetDevice(&dev));
    AUGUR_HIP(hipDeviceGetAttribute(&m.cus, hipDeviceAttributeMultiprocessorCount, dev));
    m.cus = std::max(m.cus, 1);
  }
  // the sorted statistics kernel keeps its sums in registers over its tiles:
  // one resident round of blocks over all chains (a second round would
  // double its time), a multiple of 8 per chain
  m.Gs = m.G;
  if (m.DM == 16 && pad_comp(k) <= 32) {
    int bpc = 0;
    AUGUR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, (const void*)k_stats_s(pad_comp(k)),
                                                           kStatsThreads, 0));
    const int g = std::max(1, bpc) * m.cus / chains;
    const int tiles = (n + kStatsThreads - 1) / kStatsThreads;
    m.Gs = std::max(1, std::min(tiles, g >= 8 ? g / 8 * 8 : g));
  }
  s->chains = chains;
  s->y.alloc((size_t)n * d);
  s->alpha.alloc(k); s->mu0.alloc(d); s->cov0.alloc((size_t)d * d); s->scale0.alloc((size_t)d * d);
  s->gpi.alloc((size_t)chains * k); s->pis64.alloc((size_t)chains * k);
  s->mu64.alloc((size_t)chains * k * d); s->cov64.alloc((size_t)chains * k * d * d);
  s->tab.alloc((size_t)chains * k * m.TS);
  s->prec.alloc((size_t)chains * k * d * d);
  s->prec0.alloc((size_t)d * d);
  s->part.alloc((size_t)chains * m.Gs * k * m.S);
  s->ssum.alloc((size_t)chains * k * m.S);
  s->z.alloc((size_t)chains * n);
  s->ll.alloc((size_t)chains * (1 + m.G + k));  // per-chain sums, data partials, priors
  m.y = s->y.p; m.alpha = s->alpha.p; m.mu0 = s->mu0.p; m.cov0 = s->cov0.p; m.scale0 = s->scale0.p;
  m.gpi = s->gpi.p; m.pis64 = s->pis64.p; m.mu64 = s->mu64.p; m.cov64 = s->cov64.p;
  m.tab = s->tab.p; m.part = s->part.p; m.z = s->z.p;
  m.prec = s->prec.p; m.prec0 = s->prec0.p;
}

}  // namespace augur

using namespace augur;

extern "C" {

augur_status augur_gmm_create(const augur_gmm_desc* D, augur_sampler** out) {
  return guard([&] {
    if (!out) fail(AUGUR_ERR_ARG, "null output handle");
    *out = nullptr;
    gmm_validate(D);
    require_device();
    auto* s = new GmmSampler();
    try {
      AUGUR_HIP(hipGetDevice(&s->device));
      AUGUR_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
      s->seed = D->seed;
      const int n = D->n_points, d = D->dim, k = D->n_components, C = D->chains;
      gmm_setup(s, n, d, k, C);
      GmmDev& m = s->m;
      m.df = D->df; m.sample_pi = D->sample_pi; m.sample_cov = D->sample_cov;
      m.mu_mode = D->mu_kernel;
      m.ref_compat = D->ref_compat;
      m.mu_sim_len = D->mu_sim_len;
      m.mu_step = D->mu_step_size;
      m.mu_fixed_steps = D->mu_fixed_steps;
      if (m.mu_mode != GMM_MU_CONJ) {
        const std::vector<double> l0 = host_chol(D->cov0, d, "cov0");
        const std::vector<double> i0 = host_chol_inverse(l0, d);
        const std::vector<double> lp =
            m.mu_mode == GMM_MU_MWG ? host_chol(D->mu_prop_cov, d, "the MWG proposal covariance") : l0;
        s->cov0_inv.alloc((size_t)d * d); s->chol0.alloc((size_t)d * d); s->cholp.alloc((size_t)d * d);
        s->cov0_inv.upload(i0.data(), i0.size(), s->stream);
        s->chol0.upload(l0.data(), l0.size(), s->stream);
        s->cholp.upload(lp.data(), lp.size(), s->stream);
        s->mcent.alloc((size_t)C * k * d); s->mprec.alloc((size_t)C * k * d * d);
        s->mlin.alloc((size_t)C * k * d); s->mcnt.alloc((size_t)C * k); s->macc.alloc(2 * (size_t)C);
        s->macc.zero(s->stream);
        if (m.mu_mode == GMM_MU_NUTS) {
          s->nuts.alloc((size_t)C * 3 * kNutsSlots * kMuThreads);
          m.nuts = s->nuts.p;
        }
        m.cov0_inv = s->cov0_inv.p; m.chol0 = s->chol0.p; m.cholp = s->cholp.p;
        m.mcent = s->mcent.p; m.mprec = s->mprec.p; m.mlin = s->mlin.p; m.mcnt = s->mcnt.p; m.macc = s->macc.p;
        AUGUR_HIP(hipStreamSynchronize(s->stream));  // the host vectors are released on return
      }
      s->y.upload(D->y, (size_t)n * d, s->stream);
      s->alpha.upload(D->alpha, k, s->stream);
      s->mu0.upload(D->mu0, d, s->stream);
      s->cov0.upload(D->cov0, (size_t)d * d, s->stream);
      s->scale0.upload(D->scale0, (size_t)d * d, s->stream);
      // default point (AugurDefault, pyaugur/pyaugur/augurdistlib.py:143-169):
      // uniform pis, zero means, identity covariances, all points in component 0
      std::vector<double> pis((size_t)C * k, 1.0 / k), mu((size_t)C * k * d, 0.0),
          cov((size_t)C * k * d * d, 0.0);
      for (size_t c = 0; c < (size_t)C * k; ++c)
        for (int i = 0; i < d; ++i) cov[c * d * d + i * d + i] = 1.0;
      s->pis64.upload(pis.data(), pis.size(), s->stream);
      s->gpi.upload(pis.data(), pis.size(), s->stream);
      s->mu64.upload(mu.data(), mu.size(), s->stream);
      s->cov64.upload(cov.data(), cov.size(), s->stream);
      s->z.zero(s->stream);
      s->refresh_tables();
      AUGUR_HIP(hipStreamSynchronize(s->stream));
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

augur_status augur_gmm_predictive_ll(augur_sampler* s, int n, const double* y_test, double* out_per_chain) {
  return guard([&] {
    auto* h = dynamic_cast<GmmSampler*>(s);
    if (!h || n < 0 || (n > 0 && !y_test) || !out_per_chain) fail(AUGUR_ERR_ARG, "needs a GMM sampler and arrays");
    const GmmDev& m = h->m;
    const int C = h->chains;
    std::vector<double> part((size_t)C * kGmmPredBlocks, 0.0);
    if (n > 0) {
      DevBuf<double> yt((size_t)n * m.d), pb((size_t)C * m.k * (m.d * m.d + 1)), dpart(part.size());
      yt.upload(y_test, (size_t)n * m.d, h->stream);
      k_gmm_pred_prep<<<dim3(m.k, C), kConjLanes, 0, h->stream>>>(m, pb.p);
      AUGUR_CHECK_LAUNCH();
      k_gmm_predll<<<dim3(kGmmPredBlocks, C), 256, 0, h->stream>>>(m, yt.p, n, pb.p, dpart.p);
      AUGUR_CHECK_LAUNCH();
      dpart.download(part.data(), part.size(), h->stream);
      AUGUR_HIP(hipStreamSynchronize(h->stream));
    }
    for (int c = 0; c < C; ++c) {
      double t = 0.0;
      for (int b = 0; b < kGmmPredBlocks; ++b) t += part[(size_t)c * kGmmPredBlocks + b];
      out_per_chain[c] = t;
    }
  });
}

augur_status augur_op_gmm_predictive_ll(int k, int d, const double* pis, const double* mu, const double* cov,
                                        int n, const double* y_test, double* out) {
  return guard([&] {
    if (k < 1 || k > kMaxComp || d < 1 || d > kMaxDim || n < 0 || !pis || !mu || !cov || (n > 0 && !y_test) || !out)
      fail(AUGUR_ERR_ARG, "bad arguments");
    require_device();
    GmmSampler s;
    AUGUR_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    gmm_setup(&s, 1, d, k, 1);
    s.pis64.upload(pis, k, s.stream);
    s.mu64.upload(mu, (size_t)k * d, s.stream);
    s.cov64.upload(cov, (size_t)k * d * d, s.stream);
    double r = 0.0;
    const augur_status st = augur_gmm_predictive_ll(&s, n, y_test, &r);
    if (st != AUGUR_OK) fail(st, "%s", augur_last_error());
    *out = r;
  });
}

augur_status augur_op_gmm_zstep(int n, int d, int k, const float* y, const double* pis,
                                const double* mu, const double* cov, uint32_t seed, uint32_t chain,
                                uint32_t sweep, int32_t* z) {
  return guard([&] {
    if (n < 1 || d < 1 || d > kMaxDim || k < 1 || k > kMaxComp || !y || !pis || !mu || !cov || !z)
      fail(AUGUR_ERR_ARG, "bad arguments");
    require_device();
    GmmSampler s;
    AUGUR_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    gmm_setup(&s, n, d, k, 1);
    s.y.upload(y, (size_t)n * d, s.stream);
    s.gpi.upload(pis, k, s.stream);
    s.mu64.upload(mu, (size_t)k * d, s.stream);
    s.cov64.upload(cov, (size_t)k * d * d, s.stream);
    s.refresh_tables();
    launch_gmm_z(s.m, 1, seed, sweep, chain, s.stream, nullptr);
    s.z.download(z, n, s.stream);
    AUGUR_HIP(hipStreamSynchronize(s.stream));
  });
}

augur_status augur_op_gmm_conj(int d, int k, const double* alpha, const double* mu0,
                               const double* cov0, int df, const double* scale0, const int32_t* cnt,
                               const double* sum_y, const double* sum_yy, int sample_pi,
                               int sample_cov, uint32_t seed, uint32_t chain, uint32_t sweep,
                               double* pis, double* mu, double* cov) {
  return guard([&] {
    if (d < 1 || d > kMaxDim || k < 1 || k > kMaxComp || !alpha || !mu0 || !cov0 || !scale0 || !cnt ||
        !sum_y || !sum_yy || !pis || !mu || !cov)
      fail(AUGUR_ERR_ARG, "bad arguments");
    if (chain != 0) fail(AUGUR_ERR_ARG, "the conjugate op runs chain 0");
    require_device();
    GmmSampler s;
    AUGUR_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    gmm_setup(&s, 1, d, k, 1);
    GmmDev& m = s.m;
    m.G = 1;
    m.df = df; m.sample_pi = sample_pi; m.sample_cov = sample_cov;
    s.alpha.upload(alpha, k, s.stream);
    s.mu0.upload(mu0, d, s.stream);
    s.cov0.upload(cov0, (size_t)d * d, s.stream);
    s.scale0.upload(scale0, (size_t)d * d, s.stream);
    s.gpi.upload(pis, k, s.stream);
    s.mu64.upload(mu, (size_t)k * d, s.stream);
    s.cov64.upload(cov, (size_t)k * d * d, s.stream);
    // statistics centred at the current mu, as the z kernel leaves them
    std::vector<float> st((size_t)k * m.S);
    for (int j = 0; j < k; ++j) {
      double c[kMaxDim];  // the FP32 table copy of mu, as the z kernel centres
      for (int i = 0; i < d; ++i) c[i] = (double)(float)mu[(size_t)j * d + i];
      const double n = cnt[j];
      float* o = st.data() + (size_t)j * m.S;
      o[0] = (float)n;
      for (int i = 0; i < d; ++i) o[1 + i] = (float)(sum_y[j * d + i] - n * c[i]);
      for (int i = 0; i < d; ++i)
        for (int l = 0; l <= i; ++l) {
          const double syy = sum_yy[((size_t)j * d + i) * d + l];
          o[1 + d + tri(i) + l] = (float)(syy - c[i] * sum_y[j * d + l] - sum_y[j * d + i] * c[l] + n * c[i] * c[l]);
        }
    }
    s.part.upload(st.data(), st.size(), s.stream);
    s.seed = seed;
    s.sweep = sweep;
    s.refresh_tables();  // cov0^-1 and the precisions of the given covariances
    s.launch_conj();
    AUGUR_CHECK_LAUNCH();
    std::vector<double> g(k);
    s.gpi.download(g.data(), k, s.stream);
    s.mu64.download(mu, (size_t)k * d, s.stream);
    s.cov64.download(cov, (size_t)k * d * d, s.stream);
    AUGUR_HIP(hipStreamSynchronize(s.stream));
    if (sample_pi) {
      double tot = 0.0;
      for (int j = 0; j < k; ++j) tot += g[j];
      for (int j = 0; j < k; ++j) pis[j] = g[j] / tot;
    }
  });
}

}  // extern "C"
