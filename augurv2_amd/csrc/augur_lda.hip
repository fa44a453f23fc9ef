// augur_lda.hip -- LDA Gibbs sampler (lda.rv) on gfx950.
//
// Sweep (schedule ConjGibbs [theta] (*) ConjGibbs [phi] (*) DiscGibbs [z],
// order of CudaC/CgKern.hs:192-238) as two kernels per sweep:
//   k_lda_dirichlet : theta_d ~ Dir(alpha + cnt_theta[d]), phi_k ~ Dir(beta +
//                     cnt_phi[k]) (Core/CgConj.hs:171-183, 278-306 ->
//                     cbits/augur_dist.c:526-538); each count is read once and
//                     cleared by the thread that reads it.  With K <= 16 the
//                     phi rows are left unnormalised (raw gammas and block
//                     sums) and the z kernel divides, as the normalising
//                     block would, so no block waits on another here;
//                     k_lda_phi_norm writes the normalised phi the caller
//                     reads once per step (every sweep while tracing).
//   k_lda_z         : z[d, n] ~ Cat(theta[d][k] * phi[k][w]) (Core/CgDisc.hs:
//                     122-144 -> cbits/augur_dist.c:798-813), fused with the
//                     count accumulation of the new z for the next sweep.
// Layout in HBM (per chain): w, z int32 in CSR token order; theta FP32 D x K
// (doc-major); phi_t FP32 V x K (word-major, so the K
// probabilities of a token are contiguous); FP64 copies of theta, phi for
// the caller; cnt_theta D x K, cnt_phi K x V int32; the chunk list
// (document, first token) of the z kernel.
#include "augur_common.h"

namespace augur {

// FP32 categorical draw with a fixed operation order (no contraction):
// pmf_i = a_i * b_i, running sums c_i; the first i < k-1 with u * c_{k-1} < c_i,
// else k - 1.  This is augur_unnorm_disc_samp with the normalisation moved
// to the uniform side.
__device__ __noinline__ int draw_cat_f32(const float* __restrict__ a, const float* __restrict__ b,
                                         int k, float u) {
#pragma clang fp contract(off)
  float c = 0.0f;
  for (int i = 0; i < k; ++i) c = c + a[i] * b[i];
  const float t = u * c;
  float run = 0.0f;
  for (int i = 0; i < k - 1; ++i) {
    run = run + a[i] * b[i];
    if (t < run) return i;
  }
  return k - 1;
}

// The same draw for k <= kCatRegs with both operand rows held in registers
// (loaded up front instead of once per step of the search); every sum is
// formed in the order above, so the result is identical.
constexpr int kCatRegs = 16;
__device__ inline int draw_cat_f32_regs(const float* __restrict__ a, const float* __restrict__ b, int k,
                                        float u) {
#pragma clang fp contract(off)
  float pa[kCatRegs], pb[kCatRegs];
#pragma unroll
  for (int i = 0; i < kCatRegs; ++i) {
    pa[i] = i < k ? a[i] : 0.0f;
    pb[i] = i < k ? b[i] : 0.0f;
  }
  float c = 0.0f;
#pragma unroll
  for (int i = 0; i < kCatRegs; ++i)
    if (i < k) c = c + pa[i] * pb[i];
  const float t = u * c;
  float run = 0.0f;
  int z = k - 1;
  bool found = false;
#pragma unroll
  for (int i = 0; i < kCatRegs - 1; ++i) {
    if (i < k - 1) {
      run = run + pa[i] * pb[i];
      if (!found && t < run) {
        z = i;
        found = true;
      }
    }
  }
  return z;
}

constexpr int kZThreads = 128;      // z step with word-topic counts added in HBM
constexpr int kZLdsThreads = 1024;  // z step with a per-block word-topic table in LDS
constexpr int kZLdsCells = 32768;   // largest K * V of the LDS table (128 KB)

struct LdaDev {
  int K, D, V, T, n_chunks;
  const int32_t* doc_off;
  const int32_t* chunks;  // n_chunks x (document, first token, end token, 0)
  const int32_t* w;
  const double* alpha;
  const double* beta;
  int32_t* z;          // C x T
  int32_t* cnt_theta;  // C x D x K
  int32_t* cnt_phi;    // C x K x V
  double* theta64;     // C x D x K
  float* theta32;      // C x D x K
  double* phi64;       // C x K x V
  float* phit32;       // C x V x K
  double* phi_part;    // C x K x ceil(V / 256) partial row sums
  uint32_t* phi_done;  // C x K blocks of the row finished (0 between launches)
  double* gam64;       // C x V x K raw phi gammas, word-major (folded normalisation, K <= kCatRegs)
};

__device__ inline double block_sum(double v, double* red) {
  // deterministic: wave butterfly, then wave partials summed in order
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

constexpr int kDirThreads = 256;

__host__ __device__ inline int phi_blocks_per_row(int V) { return (V + kDirThreads - 1) / kDirThreads; }

// 256-thread blocks, one Gamma draw per thread, spread over the chip:
//   blocks [0, K * pb): phi row k = b / pb, words [256 j, 256 j + 256) with
//     j = b % pb (pb = ceil(V / 256)).  Each block leaves its partial sum in
//     phi_part[k][j]; the last block of the row to finish (a counter per row,
//     reset by that block) sums the partials in order j = 0 .. pb - 1 and
//     normalises the whole row into phi64 and the word-major FP32 table.
//   blocks [K * pb, K * pb + ceil(D / 4)): theta rows, one wave per document.
// kFold: the phi blocks leave the raw gammas in gam64 and their partial
// sums, and the z kernel normalises (no handshake between blocks here).
template <bool kFold>
__global__ __launch_bounds__(kDirThreads) void k_lda_dirichlet(LdaDev m, uint32_t seed, uint32_t sweep) {
  __shared__ double red[kDirThreads / 64];
  __shared__ int last;
  const int c = blockIdx.y;
  const int b = blockIdx.x;
  const int pb = phi_blocks_per_row(m.V);
  if (b < m.K * pb) {
    const int k = b / pb, j = b - k * pb;
    const Stream s{seed, (uint32_t)c, sweep, TAG_LDA_PHI};
    const size_t r = (size_t)c * m.K + k;
    int32_t* cnt = m.cnt_phi + r * m.V;
    double* row = m.phi64 + r * m.V;
    const int v = j * kDirThreads + threadIdx.x;
    double g = 0.0;
    if (v < m.V) {
      g = gamma_at(s, (uint32_t)(k * m.V + v), m.beta[v] + (double)cnt[v]);
      cnt[v] = 0;
      if (kFold) m.gam64[((size_t)c * m.V + v) * m.K + k] = g;  // a word's K gammas contiguous
      else row[v] = g;
    }
    const double part = block_sum(g, red);
    if (kFold) {
      if (threadIdx.x == 0) m.phi_part[r * pb + j] = part;
      return;
    }
    if (threadIdx.x == 0) {
      m.phi_part[r * pb + j] = part;
      __threadfence();
      last = atomicAdd(m.phi_done + r, 1u) == (uint32_t)(pb - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();  // the other blocks' gammas and partials are visible from here
    if (threadIdx.x == 0) m.phi_done[r] = 0u;
    double tot = 0.0;
    for (int i = 0; i < pb; ++i) tot += m.phi_part[r * pb + i];
    float* pt = m.phit32 + (size_t)c * m.V * m.K;
    for (int u = threadIdx.x; u < m.V; u += kDirThreads) {
      const double p = row[u] / tot;
      row[u] = p;
      pt[(size_t)u * m.K + k] = (float)p;
    }
    return;
  }
  const int d = (b - m.K * pb) * (kDirThreads >> 6) + (threadIdx.x >> 6);
  if (d >= m.D) return;
  const int lane = threadIdx.x & 63;
  const Stream s{seed, (uint32_t)c, sweep, TAG_LDA_THETA};
  const size_t off = ((size_t)c * m.D + d) * m.K;
  double part = 0.0;
  for (int k = lane; k < m.K; k += 64) {
    const double g = gamma_at(s, (uint32_t)(d * m.K + k), m.alpha[k] + (double)m.cnt_theta[off + k]);
    m.cnt_theta[off + k] = 0;
    m.theta64[off + k] = g;
    part += g;
  }
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
  for (int k = lane; k < m.K; k += 64) {
    const double p = m.theta64[off + k] / part;
    m.theta64[off + k] = p;
    m.theta32[off + k] = (float)p;
  }
}

static dim3 dirichlet_grid(const LdaDev& m, int chains) {
  return dim3(m.K * phi_blocks_per_row(m.V) + (m.D + 3) / 4, chains);
}

// One wave per chunk of up to 64 tokens of one document (grid-stride over
// the chunk list, so a corpus of T tokens runs on ~T / 64 waves whatever
// the document lengths).  Lane j takes token j of the chunk.  Topic counts
// of the chunk are formed with ballots (lane k owns topic k) and added with
// one atomic per topic present.  Word-topic counts: with kLdsTable each block
// counts its tokens in a K x V table in LDS and adds the non-zero cells to
// HBM at the end, so a frequent (topic, word) cell takes one device-scope
// atomic per block instead of one per token (device-scope atomics on one
// address are serialised, and LDA's sparse topics concentrate tokens on few
// cells); otherwise one atomic per token.  Integer adds commute, so both
// tables are exact whatever the order.  The count tables start the sweep at
// zero (cleared by the Dirichlet kernel that read them).  The Philox key of
// table row c is chain0 + c.
//
// kFold (K <= kCatRegs): phi arrives as the raw gammas of k_lda_dirichlet<true>
// and the partial sums of its row blocks.  Every block sums a row's partials
// in block order (as the normalising block of the unfolded form does) and
// forms a token's K probabilities as (float)(gamma / sum), the unfolded
// form's FP64 division, so the draws are the same.  The normalised phi the
// caller reads is written by k_lda_phi_norm after the sweeps.  This saves
// the Dirichlet kernel's cross-block handshake.
template <bool kLdsTable, bool kFold>
__global__ __launch_bounds__(kLdsTable ? kZLdsThreads : kZThreads) void k_lda_z(LdaDev m, uint32_t seed,
                                                                              uint32_t sweep, uint32_t chain0) {
  extern __shared__ __align__(16) unsigned char zsm[];
  int32_t* tab = reinterpret_cast<int32_t*>(zsm);  // K x V, kLdsTable only
  __shared__ double rsum[kFold ? kCatRegs : 1];    // phi row sums, kFold only

  const int c = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int waves = blockDim.x >> 6;
  const Stream s{seed, chain0 + (uint32_t)c, sweep, TAG_LDA_Z};
  const float* phit = m.phit32 + (size_t)c * m.V * m.K;
  int32_t* cphi = m.cnt_phi + (size_t)c * m.K * m.V;
  int32_t* zc = m.z + (size_t)c * m.T;
  const int cells = m.K * m.V;
  const double* gam = m.gam64 + (size_t)c * m.V * m.K;
  // the wave's first chunk (descriptor and word ids) is requested before the
  // barriers below, so that memory round overlaps them
  const int qstep = gridDim.x * waves;
  int q = blockIdx.x * waves + (threadIdx.x >> 6);
  int4 ch = make_int4(0, 0, 0, 0);
  int wd = 0;
  if (q < m.n_chunks) {
    ch = *reinterpret_cast<const int4*>(m.chunks + 4 * (size_t)__builtin_amdgcn_readfirstlane(q));
    if (ch.y + lane < ch.z) wd = m.w[ch.y + lane];
  }
  if (kFold) {
    const int pb = phi_blocks_per_row(m.V);
    if (threadIdx.x < m.K) {
      const double* pp = m.phi_part + ((size_t)c * m.K + threadIdx.x) * pb;
      double tot = 0.0;
      for (int i = 0; i < pb; ++i) tot += pp[i];
      rsum[threadIdx.x] = tot;
    }
    __syncthreads();
  }
  if (kLdsTable) {
    for (int i = threadIdx.x; i < cells; i += blockDim.x) tab[i] = 0;
    __syncthreads();
  }
  for (; q < m.n_chunks; q += qstep) {
    const int du = __builtin_amdgcn_readfirstlane(ch.x), t0 = ch.y, t1 = ch.z;
    const float* th = m.theta32 + ((size_t)c * m.D + du) * m.K;
    int32_t* ct = m.cnt_theta + ((size_t)c * m.D + du) * m.K;
    const int t = t0 + lane;
    const bool valid = t < t1;
    int znew = -1;
    if (valid) {
      const float u = u01f(s.draw((uint32_t)t, 0).x);
      if (kFold) {
        float pw[kCatRegs];
#pragma unroll
        for (int k = 0; k < kCatRegs; ++k) pw[k] = k < m.K ? (float)(gam[(size_t)wd * m.K + k] / rsum[k]) : 0.0f;
        znew = draw_cat_f32_regs(th, pw, m.K, u);
      } else {
        znew = m.K <= kCatRegs ? draw_cat_f32_regs(th, phit + (size_t)wd * m.K, m.K, u)
                               : draw_cat_f32(th, phit + (size_t)wd * m.K, m.K, u);
      }
      zc[t] = znew;
      if (kLdsTable) atomicAdd(tab + (size_t)znew * m.V + wd, 1);
      else atomicAdd(cphi + (size_t)znew * m.V + wd, 1);
    }
    if (m.K <= 64) {
      int my_count = 0;  // tokens of the chunk on topic `lane`
      for (int k = 0; k < m.K; ++k) {
        const unsigned long long bal = __ballot(valid && znew == k);
        if (lane == k) my_count = __popcll(bal);
      }
      if (my_count) atomicAdd(ct + lane, my_count);
    } else if (valid) {
      atomicAdd(ct + znew, 1);
    }
    if (q + qstep < m.n_chunks) {  // the next chunk of this wave
      ch = *reinterpret_cast<const int4*>(m.chunks + 4 * (size_t)__builtin_amdgcn_readfirstlane(q + qstep));
      wd = ch.y + lane < ch.z ? m.w[ch.y + lane] : 0;
    }
  }
  if (kLdsTable) {
    __syncthreads();
    for (int i = threadIdx.x; i < cells; i += blockDim.x) {
      const int n = tab[i];
      if (n) atomicAdd(cphi + i, n);
    }
  }
}

// The normalised phi of the folded form (FP64 K x V for the caller, FP32
// V x K): one block per (row k, chain), the row's partials summed in block
// order and each gamma divided as in the z step.
__global__ __launch_bounds__(256) void k_lda_phi_norm(LdaDev m) {
  const int k = blockIdx.x, c = blockIdx.y;
  const int pb = phi_blocks_per_row(m.V);
  const size_t r = (size_t)c * m.K + k;
  double tot = 0.0;
  for (int i = 0; i < pb; ++i) tot += m.phi_part[r * pb + i];
  const double* g = m.gam64 + (size_t)c * m.V * m.K;
  double* p64 = m.phi64 + r * m.V;
  float* pt = m.phit32 + (size_t)c * m.V * m.K;
  for (int v = threadIdx.x; v < m.V; v += 256) {
    const double p = g[(size_t)v * m.K + k] / tot;
    p64[v] = p;
    pt[(size_t)v * m.K + k] = (float)p;
  }
}

// Chunk list of a CSR corpus: (document, first token, end token, 0) of each
// run of up to 64 consecutive tokens of one document; empty documents have
// no chunk.
static std::vector<int32_t> lda_chunks(const int32_t* doc_off, int D) {
  std::vector<int32_t> ch;
  for (int d = 0; d < D; ++d)
    for (int t = doc_off[d]; t < doc_off[d + 1]; t += 64) {
      ch.push_back(d);
      ch.push_back(t);
      ch.push_back(std::min(t + 64, doc_off[d + 1]));
      ch.push_back(0);
    }
  return ch;
}

// The z step over all chains of `m`: the LDS-table form when the K x V table
// fits (one 1024-thread block per CU at most), else the HBM-atomic form.
static int device_cus() {
  int dev = 0, cus = 0;
  AUGUR_HIP(hipGetDevice(&dev));
  AUGUR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return std::max(cus, 1);
}

static void launch_lda_z(const LdaDev& m, int chains, uint32_t seed, uint32_t sweep, uint32_t chain0, int cus,
                         hipStream_t stream, const std::pair<hipEvent_t, hipEvent_t>* ev, bool fold = false) {
  if ((size_t)m.K * m.V <= (size_t)kZLdsCells) {
    const int per = kZLdsThreads / 64;
    const dim3 g(std::max(1, std::min((m.n_chunks + per - 1) / per, cus)), chains);
    const size_t smem = (size_t)m.K * m.V * sizeof(int32_t);
    auto k = fold ? k_lda_z<true, true> : k_lda_z<true, false>;
    allow_lds(k, smem);
    if (ev)
      hipExtLaunchKernelGGL(k, g, dim3(kZLdsThreads), smem, stream, ev->first, ev->second, 0, m, seed, sweep,
                            chain0);
    else
      k<<<g, kZLdsThreads, smem, stream>>>(m, seed, sweep, chain0);
  } else {
    const int per = kZThreads / 64;
    const dim3 g(std::max(1, std::min((m.n_chunks + per - 1) / per, 4096)), chains);
    auto k = fold ? k_lda_z<false, true> : k_lda_z<false, false>;
    if (ev)
      hipExtLaunchKernelGGL(k, g, dim3(kZThreads), 0, stream, ev->first, ev->second, 0, m, seed, sweep, chain0);
    else
      k<<<g, kZThreads, 0, stream>>>(m, seed, sweep, chain0);
  }
  AUGUR_CHECK_LAUNCH();
}

// Count tables of the current z (after set_pt / at creation).
__global__ void k_lda_count(LdaDev m) {
  const int c = blockIdx.y;
  for (int d = blockIdx.x * blockDim.x + threadIdx.x; d < m.D; d += gridDim.x * blockDim.x) {
    for (int t = m.doc_off[d]; t < m.doc_off[d + 1]; ++t) {
      const int zz = m.z[(size_t)c * m.T + t];
      atomicAdd(m.cnt_theta + ((size_t)c * m.D + d) * m.K + zz, 1);
      atomicAdd(m.cnt_phi + ((size_t)c * m.K + zz) * m.V + m.w[t], 1);
    }
  }
}

__global__ void k_to_f32(const double* __restrict__ in, float* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

__global__ void k_phi_transpose(const double* __restrict__ phi64, float* __restrict__ phit32, int K,
                                int V, int C) {
  const size_t n = (size_t)C * K * V;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t c = i / ((size_t)K * V), r = i % ((size_t)K * V);
    const int k = (int)(r / V), v = (int)(r % V);
    phit32[(c * V + v) * K + k] = (float)phi64[i];
  }
}

// Reference ll_mod: dirichlet_ll terms (cbits/augur_dist.c:488-496, no
// normalising constant) + categorical_ll of z and w.  One block per chain.
__global__ __launch_bounds__(256) void k_lda_ll(LdaDev m, double* out) {
  __shared__ double red[16];
  const int c = blockIdx.x;
  double acc = 0.0;
  for (size_t i = threadIdx.x; i < (size_t)m.D * m.K; i += blockDim.x)
    acc += (m.alpha[i % m.K] - 1.0) * log(m.theta64[(size_t)c * m.D * m.K + i]);
  for (size_t i = threadIdx.x; i < (size_t)m.K * m.V; i += blockDim.x)
    acc += (m.beta[i % m.V] - 1.0) * log(m.phi64[(size_t)c * m.K * m.V + i]);
  for (int d = 0; d < m.D; ++d)
    for (int t = m.doc_off[d] + threadIdx.x; t < m.doc_off[d + 1]; t += blockDim.x) {
      const int zz = m.z[(size_t)c * m.T + t];
      acc += log(m.theta64[((size_t)c * m.D + d) * m.K + zz]) +
             log(m.phi64[((size_t)c * m.K + zz) * m.V + m.w[t]]);
    }
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) out[c] = tot;
}

struct LdaSampler final : augur_sampler {
  LdaDev m{};
  DevBuf<int32_t> doc_off, chunks, w, z, cnt_theta, cnt_phi;
  DevBuf<uint32_t> phi_done;
  DevBuf<double> alpha, beta, theta64, phi64, phi_part, gam64, ll;
  DevBuf<float> theta32, phit32;
  int cus = 1;  // compute units of the device (grid of the z step)

  const char* family() const override { return "lda"; }

  ParamView param(const std::string& name) override {
    ParamView v;
    if (name == "theta") v = {theta64.p, sizeof(double), (size_t)m.D * m.K, (size_t)m.D * m.K, true};
    else if (name == "phi") v = {phi64.p, sizeof(double), (size_t)m.K * m.V, (size_t)m.K * m.V, true};
    else if (name == "z") v = {z.p, sizeof(int32_t), (size_t)m.T, (size_t)m.T, true};
    return v;
  }

  void recount() {
    cnt_theta.zero(stream);
    cnt_phi.zero(stream);
    k_lda_count<<<dim3((m.D + 255) / 256, chains), 256, 0, stream>>>(m);
    AUGUR_CHECK_LAUNCH();
  }

  void on_set(const std::string& name) override {
    if (name == "z") {
      // reject out-of-range topics before they index the count tables
      std::vector<int32_t> hz((size_t)m.T * chains);
      z.download(hz.data(), hz.size(), stream);
      AUGUR_HIP(hipStreamSynchronize(stream));
      for (int32_t v : hz)
        if (v < 0 || v >= m.K) fail(AUGUR_ERR_ARG, "z value %d outside [0, %d)", v, m.K);
      recount();
    } else if (name == "theta") {
      k_to_f32<<<256, 256, 0, stream>>>(theta64.p, theta32.p, theta64.n);
      AUGUR_CHECK_LAUNCH();
    } else if (name == "phi") {
      k_phi_transpose<<<256, 256, 0, stream>>>(phi64.p, phit32.p, m.K, m.V, chains);
      AUGUR_CHECK_LAUNCH();
    }
  }

  void run_sweeps(int n) override {
    const dim3 gd = dirichlet_grid(m, chains);
    const bool fold = m.gam64 != nullptr;  // K <= kCatRegs: normalisation folded into the z step
    for (int i = 0; i < n; ++i) {
      if (fold) k_lda_dirichlet<true><<<gd, kDirThreads, 0, stream>>>(m, seed, sweep);
      else k_lda_dirichlet<false><<<gd, kDirThreads, 0, stream>>>(m, seed, sweep);
      AUGUR_CHECK_LAUNCH();
      launch_lda_z(m, chains, seed, sweep, 0u, cus, stream, prof_slot(), fold);
      ++sweep;
      // the folded form writes the caller's phi when it is read: after every
      // sweep while traces record, else after the last one
      if (fold && (!traces.empty() || i == n - 1)) {
        k_lda_phi_norm<<<dim3(m.K, chains), 256, 0, stream>>>(m);
        AUGUR_CHECK_LAUNCH();
      }
      record_traces();
    }
  }

  void log_density(double* out) override {
    k_lda_ll<<<chains, 256, 0, stream>>>(m, ll.p);
    AUGUR_CHECK_LAUNCH();
    ll.download(out, chains, stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
  }

  void stats(double* out, size_t bytes) override {
    if (bytes < sizeof(double)) fail(AUGUR_ERR_ARG, "stats buffer too small");
    out[0] = (double)sweep;
  }
};

static void lda_validate(const augur_lda_desc* d) {
  if (!d) fail(AUGUR_ERR_ARG, "null descriptor");
  if (d->n_topics < 2 || d->n_docs < 1 || d->n_vocab < 1 || d->chains < 1)
    fail(AUGUR_ERR_ARG, "need K >= 2, D >= 1, V >= 1, chains >= 1");
  if (!d->doc_off || !d->w || !d->alpha || !d->beta) fail(AUGUR_ERR_ARG, "null array");
  if (d->doc_off[0] != 0) fail(AUGUR_ERR_ARG, "doc_off[0] must be 0");
  for (int i = 0; i < d->n_docs; ++i)
    if (d->doc_off[i + 1] < d->doc_off[i]) fail(AUGUR_ERR_ARG, "doc_off must be non-decreasing");
  const int T = d->doc_off[d->n_docs];
  for (int t = 0; t < T; ++t)
    if (d->w[t] < 0 || d->w[t] >= d->n_vocab)
      fail(AUGUR_ERR_ARG, "word id %d at token %d outside [0, %d)", d->w[t], t, d->n_vocab);
  for (int k = 0; k < d->n_topics; ++k)
    if (!(d->alpha[k] > 0.0)) fail(AUGUR_ERR_ARG, "alpha must be positive");
  for (int v = 0; v < d->n_vocab; ++v)
    if (!(d->beta[v] > 0.0)) fail(AUGUR_ERR_ARG, "beta must be positive");
}

// The LDA DiscGibbs kernel on caller-supplied tables (one chain).
static void lda_zstep_op(int n_docs, int k, int v, const int32_t* doc_off, const int32_t* w,
                         const float* theta, const float* phi_t, uint32_t seed, uint32_t chain,
                         uint32_t sweep, int32_t* z, int32_t* cnt_theta, int32_t* cnt_phi) {
  require_device();
  if (n_docs < 1 || k < 2 || v < 1 || !doc_off || !w || !theta || !phi_t || !z)
    fail(AUGUR_ERR_ARG, "bad arguments");
  const int T = doc_off[n_docs];
  for (int t = 0; t < T; ++t)
    if (w[t] < 0 || w[t] >= v) fail(AUGUR_ERR_ARG, "word id out of range");
  const std::vector<int32_t> ch = lda_chunks(doc_off, n_docs);
  DevBuf<int32_t> ddo(n_docs + 1), dch(std::max<size_t>(ch.size(), 4)), dw(std::max(T, 1)),
      dz(std::max(T, 1)), dct((size_t)n_docs * k), dcp((size_t)k * v);
  DevBuf<float> dth((size_t)n_docs * k), dph((size_t)v * k);
  ddo.upload(doc_off, n_docs + 1, 0);
  dch.upload(ch.data(), ch.size(), 0);
  dw.upload(w, T, 0);
  dth.upload(theta, (size_t)n_docs * k, 0);
  dph.upload(phi_t, (size_t)v * k, 0);
  dct.zero(0);
  dcp.zero(0);
  LdaDev m{};
  m.K = k; m.D = n_docs; m.V = v; m.T = T; m.n_chunks = (int)(ch.size() / 4);
  m.doc_off = ddo.p; m.chunks = dch.p; m.w = dw.p; m.z = dz.p; m.cnt_theta = dct.p; m.cnt_phi = dcp.p;
  m.theta32 = dth.p; m.phit32 = dph.p;
  launch_lda_z(m, 1, seed, sweep, chain, device_cus(), 0, nullptr);
  dz.download(z, T, 0);
  if (cnt_theta) dct.download(cnt_theta, (size_t)n_docs * k, 0);
  if (cnt_phi) dcp.download(cnt_phi, (size_t)k * v, 0);
  AUGUR_HIP(hipDeviceSynchronize());
}

}  // namespace augur

namespace augur {

// Predictive log-likelihood of held-out tokens under each chain's current
// theta and phi (experiments/lda/lda.py:149-173): sum over tokens of
// log sum_k theta[d, k] phi[k, w], FP64 log-sum-exp with max subtraction.
// Grid (blocks, chains); one partial per block, summed in order.
constexpr int kPredBlocks = 256;
__global__ __launch_bounds__(256) void k_lda_predll(LdaDev m, const int32_t* __restrict__ doc, const int32_t* __restrict__ w,
                                                    int T, double* __restrict__ part) {
  __shared__ double red[4];
  const int c = blockIdx.y, K = m.K;
  const double* th = m.theta64 + (size_t)c * m.D * K;
  const double* ph = m.phi64 + (size_t)c * K * m.V;
  double acc = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < T; i += gridDim.x * 256) {
    const int d = doc[i], v = w[i];
    double mx = -INFINITY;
    for (int k = 0; k < K; ++k) mx = fmax(mx, log(th[(size_t)d * K + k]) + log(ph[(size_t)k * m.V + v]));
    double se = 0.0;
    for (int k = 0; k < K; ++k) se += exp(log(th[(size_t)d * K + k]) + log(ph[(size_t)k * m.V + v]) - mx);
    acc += mx + log(se);
  }
  const double bs = block_sum(acc, red);
  if (threadIdx.x == 0) part[(size_t)c * gridDim.x + blockIdx.x] = bs;
}

}  // namespace augur

using namespace augur;

extern "C" {

augur_status augur_lda_predictive_ll(augur_sampler* s, int n_docs, const int32_t* doc_off, const int32_t* w,
                                     double* out_per_chain) {
  return guard([&] {
    auto* h = dynamic_cast<LdaSampler*>(s);
    if (!h || !doc_off || !w || !out_per_chain) fail(AUGUR_ERR_ARG, "needs an LDA sampler and arrays");
    if (n_docs != h->m.D) fail(AUGUR_ERR_ARG, "held-out documents: %d, the model has %d", n_docs, h->m.D);
    const int T = doc_off[n_docs];
    if (doc_off[0] != 0 || T < 0) fail(AUGUR_ERR_ARG, "bad held-out offsets");
    std::vector<int32_t> doc((size_t)std::max(T, 1));
    for (int d = 0; d < n_docs; ++d) {
      if (doc_off[d + 1] < doc_off[d]) fail(AUGUR_ERR_ARG, "held-out offsets decrease at %d", d);
      for (int i = doc_off[d]; i < doc_off[d + 1]; ++i) doc[i] = d;
    }
    for (int i = 0; i < T; ++i)
      if (w[i] < 0 || w[i] >= h->m.V) fail(AUGUR_ERR_ARG, "held-out word %d outside [0, %d)", w[i], h->m.V);
    const int C = h->chains;
    std::vector<double> part((size_t)C * kPredBlocks);
    if (T > 0) {
      DevBuf<int32_t> ddoc(T), dw(T);
      DevBuf<double> dpart((size_t)C * kPredBlocks);
      ddoc.upload(doc.data(), T, h->stream);
      dw.upload(w, T, h->stream);
      k_lda_predll<<<dim3(kPredBlocks, C), 256, 0, h->stream>>>(h->m, ddoc.p, dw.p, T, dpart.p);
      AUGUR_CHECK_LAUNCH();
      dpart.download(part.data(), part.size(), h->stream);
      AUGUR_HIP(hipStreamSynchronize(h->stream));
    }
    for (int c = 0; c < C; ++c) {
      double t = 0.0;
      for (int b = 0; b < kPredBlocks && T > 0; ++b) t += part[(size_t)c * kPredBlocks + b];
      out_per_chain[c] = t;
    }
  });
}

augur_status augur_op_lda_predictive_ll(int n_docs, int k, int v, const double* theta, const double* phi,
                                        const int32_t* doc_off, const int32_t* w, double* out) {
  return guard([&] {
    if (n_docs < 1 || k < 1 || v < 1 || !theta || !phi || !doc_off || !w || !out) fail(AUGUR_ERR_ARG, "bad arguments");
    require_device();
    const int T = doc_off[n_docs];
    if (doc_off[0] != 0 || T < 0) fail(AUGUR_ERR_ARG, "bad held-out offsets");
    std::vector<int32_t> doc((size_t)std::max(T, 1));
    for (int d = 0; d < n_docs; ++d) {
      if (doc_off[d + 1] < doc_off[d]) fail(AUGUR_ERR_ARG, "held-out offsets decrease at %d", d);
      for (int i = doc_off[d]; i < doc_off[d + 1]; ++i) doc[i] = d;
    }
    for (int i = 0; i < T; ++i)
      if (w[i] < 0 || w[i] >= v) fail(AUGUR_ERR_ARG, "held-out word %d outside [0, %d)", w[i], v);
    *out = 0.0;
    if (T == 0) return;
    hipStream_t st;
    AUGUR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    try {
      DevBuf<double> th((size_t)n_docs * k), ph((size_t)k * v), part(kPredBlocks);
      DevBuf<int32_t> ddoc(T), dw(T);
      th.upload(theta, th.n, st);
      ph.upload(phi, ph.n, st);
      ddoc.upload(doc.data(), T, st);
      dw.upload(w, T, st);
      LdaDev m{};
      m.K = k; m.D = n_docs; m.V = v; m.theta64 = th.p; m.phi64 = ph.p;
      k_lda_predll<<<dim3(kPredBlocks, 1), 256, 0, st>>>(m, ddoc.p, dw.p, T, part.p);
      AUGUR_CHECK_LAUNCH();
      std::vector<double> hp(kPredBlocks);
      part.download(hp.data(), kPredBlocks, st);
      AUGUR_HIP(hipStreamSynchronize(st));
      double t = 0.0;
      for (double x : hp) t += x;
      *out = t;
    } catch (...) {
      (void)hipStreamDestroy(st);
      throw;
    }
    AUGUR_HIP(hipStreamDestroy(st));
  });
}

augur_status augur_lda_create(const augur_lda_desc* d, augur_sampler** out) {
  return guard([&] {
    if (!out) fail(AUGUR_ERR_ARG, "null output handle");
    *out = nullptr;
    lda_validate(d);
    require_device();
    auto* s = new LdaSampler();
    try {
      AUGUR_HIP(hipGetDevice(&s->device));
      AUGUR_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
      s->cus = device_cus();
      s->chains = d->chains;
      s->seed = d->seed;
      const int K = d->n_topics, D = d->n_docs, V = d->n_vocab, C = d->chains;
      const int T = d->doc_off[D];
      const std::vector<int32_t> ch = lda_chunks(d->doc_off, D);
      s->doc_off.alloc(D + 1);
      s->chunks.alloc(std::max<size_t>(ch.size(), 4));
      s->phi_part.alloc((size_t)C * K * phi_blocks_per_row(V));
      s->phi_done.alloc((size_t)C * K);
      s->w.alloc(std::max(T, 1));
      s->alpha.alloc(K);
      s->beta.alloc(V);
      s->z.alloc((size_t)C * std::max(T, 1));
      s->cnt_theta.alloc((size_t)C * D * K);
      s->cnt_phi.alloc((size_t)C * K * V);
      s->theta64.alloc((size_t)C * D * K);
      s->theta32.alloc((size_t)C * D * K);
      s->phi64.alloc((size_t)C * K * V);
      if (K <= kCatRegs) s->gam64.alloc((size_t)C * K * V);
      s->phit32.alloc((size_t)C * V * K);
      s->ll.alloc(C);
      s->doc_off.upload(d->doc_off, D + 1, s->stream);
      s->chunks.upload(ch.data(), ch.size(), s->stream);
      s->phi_done.zero(s->stream);
      s->w.upload(d->w, T, s->stream);
      s->alpha.upload(d->alpha, K, s->stream);
      s->beta.upload(d->beta, V, s->stream);
      s->z.zero(s->stream);
      s->theta64.zero(s->stream);
      s->theta32.zero(s->stream);
      s->phi64.zero(s->stream);
      s->phit32.zero(s->stream);
      LdaDev& m = s->m;
      m.K = K; m.D = D; m.V = V; m.T = T; m.n_chunks = (int)(ch.size() / 4);
      m.doc_off = s->doc_off.p; m.chunks = s->chunks.p; m.w = s->w.p; m.alpha = s->alpha.p; m.beta = s->beta.p;
      m.z = s->z.p; m.cnt_theta = s->cnt_theta.p; m.cnt_phi = s->cnt_phi.p;
      m.theta64 = s->theta64.p; m.theta32 = s->theta32.p; m.phi64 = s->phi64.p;
      m.phit32 = s->phit32.p; m.phi_part = s->phi_part.p; m.phi_done = s->phi_done.p;
      m.gam64 = s->gam64.p;  // null unless K <= kCatRegs
      s->recount();
      AUGUR_HIP(hipStreamSynchronize(s->stream));
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

augur_status augur_op_lda_zstep(int n_docs, int k, int v, const int32_t* doc_off,
                                const int32_t* w, const float* theta, const float* phi_t,
                                uint32_t seed, uint32_t chain, uint32_t sweep, int32_t* z,
                                int32_t* cnt_theta, int32_t* cnt_phi) {
  return guard([&] {
    lda_zstep_op(n_docs, k, v, doc_off, w, theta, phi_t, seed, chain, sweep, z, cnt_theta, cnt_phi);
  });
}

}  // extern "C"
