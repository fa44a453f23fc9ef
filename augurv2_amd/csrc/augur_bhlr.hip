// augur_bhlr.hip -- batched HMC for many independent hierarchical logistic
// regressions (datasets of uneven size and width) in one launch per leapfrog
// step, on gfx950.
//
// Each dataset i follows hlr.rv without intercept: v_i ~ InvGamma(a, b) |
// Exponential(a); theta_i[k] ~ Normal(0, v_i); y_i[n] ~ Bernoulli(sigmoid(
// dotprod(theta_i, x_i[n]))), with C chains per dataset and the schedule
// HMC [v] | MWG [v] (*) HMC [theta] (fixed leapfrog count for theta, step
// optionally jittered).  Up to 4 chains per dataset the pass is FP32 on the
// vector units (bhlr_body_r; at 3-4 chains with its a x phase on FP32
// matrix-core outer products, bhlr_body_rq); from 5 chains on it is a BF16
// matrix-core GEMM (bhlr_body_mfma).  The leapfrog passes before a
// trajectory's last run gradient-only instantiations (no log-density).
//
// Layout in HBM: every dataset's X stored BF16, row-major, back to back; y as
// uint8.  The pass kernel gives each dataset ceil(N_i / R) blocks of R rows (R
// chosen by the autotuner at creation); a block stages 64-row tiles in LDS
// (BF16 -> FP32), computes t = x . theta for all chains of its dataset,
// a = y - sigmoid(t), log sigmoid(+-t) and a x (FP32 arithmetic, FP64
// accumulation), and writes its partials; a finish kernel (one thread per
// value) sums each dataset's partials in block order (deterministic) and
// applies the leapfrog update of its chains.  One sweep (v block, momentum draw, L
// passes, Metropolis step, sweep counter) can be captured in a hipGraph and
// replayed: every kernel reads the sweep counter from device memory.
//
// Philox chain key of (dataset i, chain c) is i * C + c.
#include <chrono>
#include <type_traits>

#include "augur_hmc.h"

namespace augur {

constexpr int kBTile = 64;
constexpr int kBThreads = 256;

struct BDs {
  int n, d;
  int nblk, blk0;  // pass blocks of this dataset and the first one
  int poff;        // partials: [value][nblk] from here, value = c * (d + 2) + j
  int soff;        // offset in the per-chain state vector ([theta, b, v] per dataset)
  int goff;        // offset in the per-chain (d + 1)-wide vectors
  int voff;        // first global value index of this dataset
  int woff;        // offset of its D x D surrogate map (sum of D_i'^2 before it)
  const uint16_t* x;
  const uint8_t* y;
};

// Rows of every dataset are stored padded with zero rows (X and y) to a
// multiple of kRowPad: any tile of the register or matrix-core pass that
// starts below n then lies wholly in the buffers, so their loads carry no
// bounds checks (rows at or past the block's end are masked in the row term),
// and each dataset's X and y start 512-byte and 256-byte aligned.
constexpr int kRowPad = 256;
inline int padded_rows(int n) { return (n + kRowPad - 1) / kRowPad * kRowPad; }

struct BArgs {
  const BDs* ds;
  const int* blk_ds;
  const int* val_ds;  // global value -> dataset
  int n_ds, C, SW, GW, rows_per_blk, nval;
  float* beta32;     // [C][GW]
  double* part;
  int mode, l;
  double* cache_g;   // [C][GW]
  double* cache_ll;  // [C][n_ds]
  double* gprop;     // [C][GW]
  double* llprop;    // [C][n_ds]
  double* prop;      // [C][SW]
  double* p;         // [C][GW]
  double* p0;        // [C][GW]
  double* k0;        // [C][n_ds]
  double* state;     // [C][SW]
  double* stats;     // [block 0 v | 1 theta][C][n_ds][2]
  const uint32_t* sweep;  // device sweep counter
  uint32_t seed;
  int prior_kind, ref_compat;
  double prior_a, prior_b;
  double step;       // theta block: step size (jittered per chain and sweep
  int steps;         // when step_jitter) and fixed count
  int step_jitter;
  double v_sim_len, v_eps;  // v block: HMC settings, or the MWG proposal variance
  int v_fixed;
  int v_mwg;
  // per dataset: theta step size and v proposal variance (A.step / A.v_eps
  // until warmup adaptation changes them), and the last sweep's acceptance
  // probabilities [C][n_ds] of the theta and v moves
  double* eps_ds;
  double* vvar_ds;
  double* acc_theta;
  double* acc_v;
  int x_stream;  // nontemporal X loads (all datasets' X is far beyond the Infinity Cache)
  // linear gradient surrogate (null: off).  smode[i] = 1: dataset i runs
  // surrogate trajectories and evaluates only the last pass (its proposal).
  // cap_t / cap_g: this pass's capture slot ([C][GW]) for the fit, or null.
  // sw: dataset i's map W (D x D FP32 at ds.woff, row = output coordinate);
  // sm / sg ([GW]): mean theta and mean data gradient of the fitting pairs.
  const int* smode;
  float* cap_t;
  float* cap_g;
  const float* sw;
  const double* sm;
  const double* sg;
};

enum { BMODE_CACHE = 0, BMODE_LEAPFROG = 1 };

// A surrogate dataset's blocks skip the leapfrog passes before the last.
__device__ inline bool bsurr_skip(const BArgs& A, int di) {
  return A.smode && A.mode == BMODE_LEAPFROG && A.l < A.steps && A.smode[di];
}

// Nontemporal 16- / 8-byte loads of packed BF16 X.
typedef unsigned v4u_t __attribute__((ext_vector_type(4)));
typedef unsigned v2u_t __attribute__((ext_vector_type(2)));
// s_waitcnt immediate that waits until at most n vector-memory loads are
// outstanding (gfx9 encoding: vmcnt bits 3:0 and 15:14; expcnt and lgkmcnt
// left at their maxima, i.e. not waited on)
__host__ __device__ constexpr int vmcnt_imm(int n) {
  return (n & 15) | (7 << 4) | (15 << 8) | (((n >> 4) & 3) << 14);
}

// The datasets' X and y pointers come from the device table of BDs, so the
// compiler cannot tell their address space and would load through them with
// flat instructions, which count against both the vector-memory and the LDS
// counters (every wait on one then waits on the other).  X and y are global
// memory: these casts make the loads global ones.
#define AUGUR_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ inline const AUGUR_GLOBAL T* gmem(const T* p) {
  return (const AUGUR_GLOBAL T*)p;
}

__device__ inline uint4 ld_x(const uint4* p, int nt) {
  const AUGUR_GLOBAL v4u_t* g = gmem((const v4u_t*)p);
  const v4u_t v = nt ? __builtin_nontemporal_load(g) : *g;
  return make_uint4(v[0], v[1], v[2], v[3]);
}
__device__ inline uint2 ld_x(const uint2* p, int nt) {
  const AUGUR_GLOBAL v2u_t* g = gmem((const v2u_t*)p);
  const v2u_t v = nt ? __builtin_nontemporal_load(g) : *g;
  return make_uint2(v[0], v[1]);
}

__device__ inline float bf16f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <int D, int CP>
__device__ void bhlr_body(const BArgs& A, const BDs& ds, int di, unsigned char* smem) {
  constexpr int DS = D + 4;
  constexpr int TR = kBTile;
  constexpr int NV = TR * D / 8;  // uint4 (8 BF16) per tile
  constexpr int NVT = (NV + kBThreads - 1) / kBThreads;
  constexpr int NB = (TR * CP + kBThreads - 1) / kBThreads;
  constexpr int NP = CP * (D / 4);
  constexpr int RS = NP >= kBThreads ? 1 : kBThreads / NP;
  constexpr int NS = NP > kBThreads ? NP / kBThreads : 1;
  constexpr int VS = D + 2;
  float* xs = (float*)smem;                 // TR x DS
  float* bs = xs + TR * DS;                 // CP x DS
  float* as = bs + CP * DS;                 // TR x CP
  int* ys = (int*)(as + TR * CP);           // TR
  double* red = (double*)(((uintptr_t)(ys + TR) + 15) & ~(uintptr_t)15);  // 2 x 256
  double* red2 = red + 2 * kBThreads;       // 4 x 256
  int* flag = (int*)(red2 + 4 * kBThreads);

  const int tid = threadIdx.x;
  const int local = blockIdx.x - ds.blk0;
  const int rbeg = local * A.rows_per_blk;
  const int rend = min(ds.n, rbeg + A.rows_per_blk);
  for (int i = tid; i < CP * D; i += kBThreads) {
    const int c = i / D, j = i % D;
    bs[c * DS + j] = c < A.C ? A.beta32[(size_t)c * A.GW + ds.goff + j] : 0.0f;
  }
  __syncthreads();
  const int my_c = tid % CP;
  constexpr bool BREG = D <= 64;  // theta_c in registers, else read from LDS
  float breg[BREG ? D : 1];
  if constexpr (BREG) {
#pragma unroll
    for (int j = 0; j < D; ++j) breg[j] = bs[my_c * DS + j];
  }
  double gacc[NS][4];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int k = 0; k < 4; ++k) gacc[s][k] = 0.0;
  double llacc = 0.0, gbacc = 0.0;
  uint4 pre[NVT];
  int ypre = 0;
  auto prefetch = [&](int r0) {
    const int rows = min(TR, rend - r0);
    const AUGUR_GLOBAL v4u_t* s8 = gmem((const v4u_t*)(ds.x + (size_t)r0 * D));
#pragma unroll
    for (int k = 0; k < NVT; ++k) {
      const int i = tid + k * kBThreads;
      v4u_t v = {0u, 0u, 0u, 0u};
      if (i < NV && i < rows * (D / 8)) v = s8[i];
      pre[k] = make_uint4(v[0], v[1], v[2], v[3]);
    }
    ypre = (tid < TR && tid < rows) ? (int)*gmem(ds.y + r0 + tid) : 0;
  };
  if (rbeg < rend) prefetch(rbeg);
  for (int r0 = rbeg; r0 < rend; r0 += TR) {
    const int rows = min(TR, rend - r0);
#pragma unroll
    for (int k = 0; k < NVT; ++k) {
      const int i = tid + k * kBThreads;
      if (i < NV) {
        const int e = 8 * i;
        float* dst = xs + (e / D) * DS + e % D;
        const uint4 v = pre[k];
        *(float4*)dst = make_float4(bf16f(v.x & 0xffff), bf16f(v.x >> 16), bf16f(v.y & 0xffff),
                                    bf16f(v.y >> 16));
        *(float4*)(dst + 4) = make_float4(bf16f(v.z & 0xffff), bf16f(v.z >> 16),
                                          bf16f(v.w & 0xffff), bf16f(v.w >> 16));
      }
    }
    if (tid < TR) ys[tid] = ypre;
    __syncthreads();
    if (r0 + TR < rend) prefetch(r0 + TR);
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int q = tid + kBThreads * m;
      if (q < TR * CP) {
        const int r = q / CP;
        const float* xr = xs + r * DS;
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
        for (int j = 0; j < D; j += 4) {
          const float4 xv = *(const float4*)(xr + j);
          float4 bv;
          if constexpr (BREG) bv = make_float4(breg[j], breg[j + 1], breg[j + 2], breg[j + 3]);
          else bv = *(const float4*)(bs + my_c * DS + j);
          s0 = fmaf(xv.x, bv.x, s0);
          s1 = fmaf(xv.y, bv.y, s1);
          s2 = fmaf(xv.z, bv.z, s2);
          s3 = fmaf(xv.w, bv.w, s3);
        }
        const float t = (s0 + s1) + (s2 + s3);
        float a;
        const float lr = log_sigmoid_pm(t, ys[r], &a);
        if (r < rows) {
          llacc += (double)lr;
          gbacc += (double)a;
        } else {
          a = 0.0f;
        }
        as[r * CP + my_c] = a;
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int pair = (NS > 1) ? tid + s * kBThreads : tid % NP;
      const int rs = (NS > 1) ? 0 : tid / NP;
      const int c = pair / (D / 4), jq = pair % (D / 4);
      float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 4
      for (int r = rs * (TR / RS); r < (rs + 1) * (TR / RS); ++r) {
        const float av = as[r * CP + c];
        const float4 xv = *(const float4*)(xs + r * DS + 4 * jq);
        f.x = fmaf(av, xv.x, f.x);
        f.y = fmaf(av, xv.y, f.y);
        f.z = fmaf(av, xv.z, f.z);
        f.w = fmaf(av, xv.w, f.w);
      }
      gacc[s][0] += (double)f.x;
      gacc[s][1] += (double)f.y;
      gacc[s][2] += (double)f.z;
      gacc[s][3] += (double)f.w;
    }
    __syncthreads();
  }
  // partials of this block: part[poff + value * nblk + local]; the region
  // holds the C real chains only (chains padded up to CP write nothing)
  double* part = A.part + ds.poff;
  if constexpr (RS > 1) {
#pragma unroll
    for (int k = 0; k < 4; ++k) red2[tid * 4 + k] = gacc[0][k];
    __syncthreads();
    if (tid < NP && tid / (D / 4) < A.C) {
      const int c = tid / (D / 4), jq = tid % (D / 4);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        double s = 0.0;
        for (int rs = 0; rs < RS; ++rs) s += red2[(rs * NP + tid) * 4 + k];
        part[(size_t)(c * VS + 4 * jq + k) * ds.nblk + local] = s;
      }
    }
  } else {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int pair = tid + s * kBThreads;
      const int c = pair / (D / 4), jq = pair % (D / 4);
      if (c >= A.C) continue;
#pragma unroll
      for (int k = 0; k < 4; ++k) part[(size_t)(c * VS + 4 * jq + k) * ds.nblk + local] = gacc[s][k];
    }
  }
  red[tid] = gbacc;
  red[kBThreads + tid] = llacc;
  __syncthreads();
  if (tid < A.C) {
    double sb = 0.0, sl = 0.0;
    for (int i = tid; i < kBThreads; i += CP) {
      sb += red[i];
      sl += red[kBThreads + i];
    }
    part[(size_t)(tid * VS + D) * ds.nblk + local] = sb;
    part[(size_t)(tid * VS + D + 1) * ds.nblk + local] = sl;
  }
  (void)flag;
  (void)di;
}

// The theta block's step size of chain key gc (dataset gc / C) this sweep:
// the dataset's eps, or with step_jitter eps U(0.8, 1.2) from the block's
// jitter stream, as the reference draws it (cbits/augur_mcmclib.h:401-402).
__device__ inline double theta_step(const BArgs& A, uint32_t gc) {
  const double eps = A.eps_ds[gc / (uint32_t)A.C];
  if (!A.step_jitter) return eps;
  const Stream js{A.seed, gc, *A.sweep, TAG_HMC_BASE + 4u + 1u};
  return (0.8 + 0.4 * u01(js.draw(0, 0).x)) * eps;
}

// One wave per value (dataset i, chain c, coordinate j; j = d_i + 1: the
// log-density): sums the dataset's block partials in block order (fixed
// butterfly), then applies the leapfrog update of that coordinate.
__global__ __launch_bounds__(256) void k_bhlr_finish(BArgs A) {
  // one thread per value: its dataset's block partials summed in block order
  // (a dataset has a few dozen blocks at most, so a wave per value would
  // spend its time being launched)
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= A.nval) return;
  const int di = A.val_ds[v];
  if (bsurr_skip(A, di)) return;
  const BDs ds = A.ds[di];
  const int D = ds.d, VS = D + 2;
  const int lv = v - ds.voff;
  const int c = lv / VS, j = lv % VS;
  if (j == D) return;  // no intercept in the batched model
  const size_t ci = (size_t)c * A.n_ds + di;
  const size_t gi = (size_t)c * A.GW + ds.goff + j;
  const size_t si = (size_t)c * A.SW + ds.soff;
  // the leapfrog update's operands are requested with the partials
  const bool upd = j < D && A.mode != BMODE_CACHE && !(A.smode && A.smode[di]);
  double var = 1.0, xv = 0.0, p_old = 0.0;
  if (upd) {
    var = A.state[si + D + 1];
    xv = A.prop[si + j];
    p_old = A.p[gi];
  }
  // eight independent sums (blocks b = k mod 8), combined in a fixed order:
  // eight loads in flight instead of one chain of dependent adds
  const double* pv = A.part + ds.poff + (size_t)lv * ds.nblk;
  double s8[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b0 = 0; b0 < ds.nblk; b0 += 8) {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (b0 + k < ds.nblk) s8[k] += pv[b0 + k];
  }
  const double s = ((s8[0] + s8[1]) + (s8[2] + s8[3])) + ((s8[4] + s8[5]) + (s8[6] + s8[7]));
  if (j == D + 1) {
    if (A.mode == BMODE_CACHE) A.cache_ll[ci] = s;
    else if (A.l == A.steps) A.llprop[ci] = s;
    return;
  }
  if (A.mode == BMODE_CACHE) {
    A.cache_g[gi] = s;
    return;
  }
  const int L = A.steps;
  if (A.l == L) A.gprop[gi] = s;
  // a surrogate dataset's last pass only evaluates its proposal: the
  // trajectory kernel has already moved theta and the momentum
  if (!upd) return;
  if (A.cap_t) {  // a fitting pair: the point this pass evaluated, its data gradient
    A.cap_t[gi] = A.beta32[gi];
    A.cap_g[gi] = (float)s;
  }
  const double g = s + (-xv / var);
  const double step = theta_step(A, (uint32_t)(di * A.C + c));
  const double coef = (A.l == L) ? 0.5 * step : step;
  const double pj = p_old + coef * g;
  A.p[gi] = pj;
  if (A.l < L) {
    xv += step * pj;
    A.prop[si + j] = xv;
    A.beta32[gi] = (float)xv;
  }
}

// Register form of the pass for up to 4 chains per dataset (the form of the
// single-dataset sampler's k_hlr_pass_r, on BF16 rows): lane l owns F
// consecutive features of RPT rows per tile, LPR = max(CP, D / 8) lanes per
// row and F = D / LPR (8, or 4 for D = 16 with 4 chains), kept as packed
// BF16 until use.  Phase B: the lane's share of x . theta_c, completed over
// the row's lanes by a DPP butterfly; each (row, chain) adjoint is formed
// once (lane q takes pairs q, q + LPR, ..., all of chain q % CP) and
// gathered back by ds_bpermute; phase C accumulates a x into the lane's
// CP x F share in FP32; the block sums lanes in FP64 (LDS, fixed order).  A
// ring of PD tiles per lane streams the block's rows without barriers.
constexpr int kBrRpt = 2;   // rows per lane per tile
constexpr int kBrRing = 4;  // tiles in flight per lane

__device__ inline uint32_t xword_at(const uint4& v, int k) {
  return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}
__device__ inline uint32_t xword_at(const uint2& v, int k) { return k == 0 ? v.x : v.y; }

template <int D, int CP>
__device__ void bhlr_body_r(const BArgs& A, const BDs& ds, unsigned char* smem) {
  constexpr int LPR = (CP > D / 8) ? CP : D / 8;
  constexpr int F = D / LPR;
  constexpr int FW = F / 2;  // 32-bit words of BF16 per lane and row
  constexpr int RPT = kBrRpt, PD = kBrRing;
  constexpr int RG = kBThreads / LPR;  // rows per block per row slot
  constexpr int TRr = RG * RPT;        // rows per tile
  constexpr int NC = RPT * CP;
  constexpr int SPL = (NC + LPR - 1) / LPR;
  constexpr int GS = CP * F + 1;
  constexpr int VS = D + 2;
  static_assert(LPR >= 2 && LPR <= 16 && LPR % CP == 0 && (F == 4 || F == 8), "shape");
  typedef typename std::conditional<FW == 4, uint4, uint2>::type xword;

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int q = tid % LPR, rsub = tid / LPR, gbase = lane - q, my_c = q % CP;
  const int local = blockIdx.x - ds.blk0;
  const int rbeg = local * A.rows_per_blk;
  const int rend = min(ds.n, rbeg + A.rows_per_blk);
  const int ntile = rend > rbeg ? (rend - rbeg + TRr - 1) / TRr : 0;

  // theta of every padded chain (a real chain's value loaded for the padding,
  // then zeroed): the same loads on every path, see below
  float th[CP][F];
#pragma unroll
  for (int c = 0; c < CP; ++c)
#pragma unroll
    for (int j = 0; j < F; ++j) {
      const float v = A.beta32[(size_t)min(c, A.C - 1) * A.GW + ds.goff + F * q + j];
      th[c][j] = c < A.C ? v : 0.0f;
    }

  xword pre[PD][RPT];
  int ypre[PD][RPT];
  // tiles start at multiples of TRr below rend <= n, inside the dataset's
  // rows padded to kRowPad: no bounds checks on the loads (rows at or past
  // rend are masked in the row term)
  static_assert(kRowPad % TRr == 0, "tiles within the padding");
  auto prefetch = [&](int t, xword (&dst)[RPT], int (&yd)[RPT]) {
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int row = rbeg + t * TRr + i * RG + rsub;
      dst[i] = ld_x((const xword*)(ds.x + (size_t)row * D + F * q), A.x_stream);
      yd[i] = *gmem(ds.y + row);
    }
  };
  // Every prefetch is issued, a tile past the block's last one clamped to it
  // (valid rows, loaded again and not used), so each path through the loop
  // issues the same loads and the compiler's waits count the ring exactly:
  // theta, requested before the ring (2 loads a row), is complete once at
  // most PD * RPT * 2 loads are outstanding.  With conditional prefetches the
  // row loop waited for zero outstanding loads at every slot.
  const int last = max(ntile - 1, 0);  // rbeg < n: ntile >= 1
#pragma unroll
  for (int k = 0; k < PD; ++k) prefetch(min(k, last), pre[k], ypre[k]);
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(PD * RPT * 2));

  float g[CP][F];
#pragma unroll
  for (int c = 0; c < CP; ++c)
#pragma unroll
    for (int j = 0; j < F; ++j) g[c][j] = 0.0f;
  double llacc = 0.0, gbacc = 0.0;

  for (int base = 0; base < ntile; base += PD) {
#pragma unroll
    for (int slot = 0; slot < PD; ++slot) {
      const int t = base + slot;
      if (t >= ntile) {  // uniform: the block's remaining slots are empty
        // the refill is still issued, so every path issues the same loads
        prefetch(min(t + PD, last), pre[slot], ypre[slot]);
        continue;
      }
      float xr[RPT][F];
#pragma unroll
      for (int i = 0; i < RPT; ++i)
#pragma unroll
        for (int k = 0; k < FW; ++k) {
          const uint32_t w = xword_at(pre[slot][i], k);
          xr[i][2 * k] = __uint_as_float(w << 16);
          xr[i][2 * k + 1] = __uint_as_float(w & 0xffff0000u);
        }
      const int* yv = ypre[slot];
      float tp[NC];
#pragma unroll
      for (int i = 0; i < RPT; ++i)
#pragma unroll
        for (int c = 0; c < CP; ++c) {
          float s0 = xr[i][0] * th[c][0], s1 = xr[i][1] * th[c][1];
#pragma unroll
          for (int j = 2; j < F; j += 2) {
            s0 = fmaf(xr[i][j], th[c][j], s0);
            s1 = fmaf(xr[i][j + 1], th[c][j + 1], s1);
          }
          tp[i * CP + c] = row_sum<LPR>(s0 + s1);
        }
      float aown[SPL];
#pragma unroll
      for (int s = 0; s < SPL; ++s) {
        const int m = q + s * LPR;
        float tv = tp[0];
#pragma unroll
        for (int k = 1; k < NC; ++k) tv = (m == k) ? tp[k] : tv;
        const int i = m / CP;
        int yy = yv[0];
#pragma unroll
        for (int k = 1; k < RPT; ++k) yy = (i == k) ? yv[k] : yy;
        const bool ok = m < NC && my_c < A.C && rbeg + t * TRr + i * RG + rsub < rend;
        float a;
        const float lr = log_sigmoid_pm(tv, yy, &a);
        if (ok) {
          llacc += (double)lr;
          gbacc += (double)a;
        } else {
          a = 0.0f;
        }
        aown[s] = a;
      }
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        const float av = __shfl(aown[m / LPR], gbase + m % LPR);
        const int i = m / CP, c = m % CP;
#pragma unroll
        for (int j = 0; j < F; ++j) g[c][j] = fmaf(av, xr[i][j], g[c][j]);
      }
      prefetch(min(t + PD, last), pre[slot], ypre[slot]);
    }
  }

  // partials of this block: part[poff + value * nblk + local], real chains only
  float* gl = (float*)smem;  // [256][GS]
  double* red = (double*)(smem + ((sizeof(float) * kBThreads * GS + 15) & ~(size_t)15));  // [2][4][CP]
#pragma unroll
  for (int c = 0; c < CP; ++c)
#pragma unroll
    for (int j = 0; j < F; ++j) gl[tid * GS + c * F + j] = g[c][j];
#pragma unroll
  for (int off = 32; off >= CP; off >>= 1) {
    gbacc += __shfl_xor(gbacc, off);
    llacc += __shfl_xor(llacc, off);
  }
  if (lane < CP) {
    red[wv * CP + lane] = gbacc;
    red[4 * CP + wv * CP + lane] = llacc;
  }
  __syncthreads();
  double* part = A.part + ds.poff;
  for (int v = tid; v < A.C * D; v += kBThreads) {
    const int c = v / D, j = v % D, qq = j / F, jj = j % F;
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBThreads / 64; ++w)
#pragma unroll
      for (int k = 0; k < 64 / LPR; ++k) s += (double)gl[(w * 64 + k * LPR + qq) * GS + c * F + jj];
    part[(size_t)(c * VS + j) * ds.nblk + local] = s;
  }
  if (tid < A.C) {
    double sb = 0.0, sl = 0.0;
#pragma unroll
    for (int w = 0; w < kBThreads / 64; ++w) {
      sb += red[w * CP + tid];
      sl += red[4 * CP + w * CP + tid];
    }
    part[(size_t)(tid * VS + D) * ds.nblk + local] = sb;
    part[(size_t)(tid * VS + D + 1) * ds.nblk + local] = sl;
  }
}

template <int CP>
__global__ __launch_bounds__(kBThreads) void k_bhlr_pass_r(BArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int di = A.blk_ds[blockIdx.x];
  if (bsurr_skip(A, di)) return;
  const BDs ds = A.ds[di];
  switch (ds.d) {
    case 16: bhlr_body_r<16, CP>(A, ds, smem); break;
    case 32: bhlr_body_r<32, CP>(A, ds, smem); break;
    case 64: bhlr_body_r<64, CP>(A, ds, smem); break;
    default: bhlr_body_r<128, CP>(A, ds, smem); break;
  }
}

// The register form at 4 chains and D = 16, 32, 64 with its a x phase on the
// matrix cores.  Lanes and loads as bhlr_body_r (LPR = 4 or 8 lanes a row, F
// features each, 2 rows a tile); what changes is the arithmetic after the
// partial dot products:
//  - each lane holds theta in a lane-dependent chain order, slot s of lane q
//    being chain (q & 3) ^ s ^ x with x = 3 in the second quad of an 8-lane
//    row group, else 0.  Then the quad partner q ^ 2 holds in its slots 2, 3
//    the two chains that q keeps in its slots 0, 1, and the partner q ^ 1 in
//    slot 1 the chain that q keeps in slot 0: a reduce-scatter over the quad
//    in 3 DPP adds a row (quad_perm 2,3,0,1 then 1,0,3,2), with no selects.
//    At 8 lanes a row the half-row mirror (q <-> 7 - q) pairs lanes holding
//    the same chain, and each lane keeps one row: row 0 in the first quad,
//    row 1 in the second.  Each (row, chain) term is computed by one lane.
//  - g += a x runs on V_MFMA_F32_4X4X1_16B_F32: 16 blocks of 4 lanes, block b
//    taking A[m] from lane 4b + m and B[n] from lane 4b + n, and leaving
//    D[m][n] in register m of lane 4b + n (checked on the device).  A quad is
//    one row's lanes (or half of them), so with A = the row's adjoint of the
//    chain of quad lane m and B = the lane's feature j, register m of lane q
//    accumulates g[chain of slot m][F q + j]: 16 products a tile replace 64
//    vector FMAs and 8 lane shuffles.
// FP32 arithmetic and FP64 accumulation of the row terms as bhlr_body_r.
#ifndef AUGUR_RQ_PD
#define AUGUR_RQ_PD 4
#endif
template <int D, bool LL>
__device__ void bhlr_body_rq(const BArgs& A, const BDs& ds, unsigned char* smem) {
  constexpr int CP = 4;
  constexpr int LPR = (D / 8 > CP) ? D / 8 : CP;
  constexpr int F = D / LPR;
  constexpr int FW = F / 2;
  constexpr int RPT = kBrRpt, PD = AUGUR_RQ_PD;
  constexpr int RG = kBThreads / LPR;
  constexpr int TRr = RG * RPT;
  constexpr int GS = CP * F + 1;
  constexpr int VS = D + 2;
  static_assert((LPR == 4 || LPR == 8) && (F == 4 || F == 8) && RPT == 2, "shape");
  typedef typename std::conditional<FW == 4, uint4, uint2>::type xword;
  typedef float f4 __attribute__((ext_vector_type(4)));

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int q = tid % LPR, rsub = tid / LPR;
  const int hi = (LPR == 8) ? (q >> 2) : 0;  // second quad of the row's lanes
  const int cx = hi ? 3 : 0;
  const int p = q & 3;
  const int my_c = p ^ cx;  // chain of slot 0: the one this lane's row terms use
  const int local = blockIdx.x - ds.blk0;
  const int rbeg = local * A.rows_per_blk;
  const int rend = min(ds.n, rbeg + A.rows_per_blk);
  const int ntile = rend > rbeg ? (rend - rbeg + TRr - 1) / TRr : 0;

  float th[CP][F];  // th[s] = theta of chain p ^ s ^ cx (0 for padding chains)
#pragma unroll
  for (int s = 0; s < CP; ++s) {
    const int c = p ^ s ^ cx;
#pragma unroll
    for (int j = 0; j < F; ++j) {
      const float v = A.beta32[(size_t)min(c, A.C - 1) * A.GW + ds.goff + F * q + j];
      th[s][j] = c < A.C ? v : 0.0f;
    }
  }

  // Row addresses as a block-uniform base (scalar registers) plus a 32-bit
  // lane offset that stays fixed over the tiles, so the loads take the
  // scalar-base form and the ring costs no per-lane address arithmetic.  At 8
  // lanes a row each lane loads the y of the one row whose term it computes.
  constexpr int NY = (LPR == 8) ? 1 : RPT;  // y loads a tile
  const uint32_t xlane = (uint32_t)((rsub * D + F * q) * 2);
  const uint32_t ylane = (uint32_t)(rsub + (LPR == 8 ? hi * RG : 0));
  xword pre[PD][RPT];
  int ypre[PD][NY];
  static_assert(kRowPad % TRr == 0, "tiles within the padding");
  auto prefetch = [&](int t, xword (&dst)[RPT], int (&yd)[NY]) {
    const size_t r0 = (size_t)(rbeg + t * TRr);
#pragma unroll
    for (int i = 0; i < RPT; ++i)
      dst[i] = ld_x((const xword*)((const char*)ds.x + (r0 + (size_t)i * RG) * (D * 2) + xlane),
                    A.x_stream);
#pragma unroll
    for (int i = 0; i < NY; ++i) yd[i] = *gmem(ds.y + r0 + (size_t)i * RG + ylane);
  };
  // the same unconditional, clamped ring as bhlr_body_r
  const int last = max(ntile - 1, 0);
#pragma unroll
  for (int k = 0; k < PD; ++k) prefetch(min(k, last), pre[k], ypre[k]);
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(PD * (RPT + NY)));

  f4 acc[F];
#pragma unroll
  for (int j = 0; j < F; ++j) acc[j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  double llacc = 0.0;

  for (int base = 0; base < ntile; base += PD) {
#pragma unroll
    for (int slot = 0; slot < PD; ++slot) {
      const int t = base + slot;
      if (t >= ntile) {
        prefetch(min(t + PD, last), pre[slot], ypre[slot]);
        continue;
      }
      float xr[RPT][F];
#pragma unroll
      for (int i = 0; i < RPT; ++i)
#pragma unroll
        for (int k = 0; k < FW; ++k) {
          const uint32_t w = xword_at(pre[slot][i], k);
          xr[i][2 * k] = __uint_as_float(w << 16);
          xr[i][2 * k + 1] = __uint_as_float(w & 0xffff0000u);
        }
      // partial dot products of the lane's F features, slot order
      float tp[RPT][CP];
#pragma unroll
      for (int i = 0; i < RPT; ++i)
#pragma unroll
        for (int s = 0; s < CP; ++s) {
          float acc_t = xr[i][0] * th[s][0];
#pragma unroll
          for (int j = 1; j < F; ++j) acc_t = fmaf(xr[i][j], th[s][j], acc_t);
          tp[i][s] = acc_t;
        }
      // quad reduce-scatter: slot 0 of every lane ends with its chain's sum
      // over the quad, for both rows
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        tp[i][0] += dpp_f<0x4E>(tp[i][2]);  // quad_perm 2,3,0,1
        tp[i][1] += dpp_f<0x4E>(tp[i][3]);
        tp[i][0] += dpp_f<0xB1>(tp[i][1]);  // quad_perm 1,0,3,2
      }
      float aq[RPT];  // adjoint of (row i, chain my_c), as quad lane p supplies it
      if constexpr (LPR == 8) {
        // the other quad holds the same chain: half-row mirror, one row each
        const float t0 = tp[0][0] + dpp_f<0x141>(tp[0][0]);
        const float t1 = tp[1][0] + dpp_f<0x141>(tp[1][0]);
        const float tv = hi ? t1 : t0;
        const int yy = ypre[slot][0];
        // rows past the dataset's end are zero rows of X and the padding
        // chains' gradients are never read: only the log-density needs the mask
        const bool ok = !LL || (my_c < A.C && rbeg + t * TRr + hi * RG + rsub < rend);
        float a;
        if constexpr (LL) {
          const float lr = log_sigmoid_pm(tv, yy, &a);
          llacc += (double)(ok ? lr : 0.0f);
        } else {
          a = sigmoid_adjoint_pm(tv, yy);
        }
        a = ok ? a : 0.0f;
        const float ax = dpp_f<0x141>(a);  // the other row's, same chain
        aq[0] = hi ? ax : a;
        aq[1] = hi ? a : ax;
      } else {
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          const bool ok = !LL || (my_c < A.C && rbeg + t * TRr + i * RG + rsub < rend);
          float a;
          if constexpr (LL) {
            const float lr = log_sigmoid_pm(tp[i][0], ypre[slot][i], &a);
            llacc += (double)(ok ? lr : 0.0f);
          } else {
            a = sigmoid_adjoint_pm(tp[i][0], ypre[slot][i]);
          }
          a = ok ? a : 0.0f;
          aq[i] = a;
        }
      }
#pragma unroll
      for (int i = 0; i < RPT; ++i)
#pragma unroll
        for (int j = 0; j < F; ++j) acc[j] = __builtin_amdgcn_mfma_f32_4x4x1f32(aq[i], xr[i][j], acc[j], 0, 0, 0);
      prefetch(min(t + PD, last), pre[slot], ypre[slot]);
    }
  }

  // partials of this block, as bhlr_body_r: register m of acc[j] is the
  // chain p' ^ cx of quad lane m, i.e. chain m ^ cx
  float* gl = (float*)smem;  // [256][GS]
  double* red = (double*)(smem + ((sizeof(float) * kBThreads * GS + 15) & ~(size_t)15));  // [2][4][CP]
#pragma unroll
  for (int m = 0; m < CP; ++m)
#pragma unroll
    for (int j = 0; j < F; ++j) gl[tid * GS + (m ^ cx) * F + j] = acc[j][m];
  // lanes of one chain: equal q across the wave, then (8 lanes a row) the
  // mirrored quad lane 7 - q = q ^ 7; lane c < 4 then holds chain c's sum
  if constexpr (LL) {
#pragma unroll
    for (int off = 32; off >= LPR; off >>= 1) llacc += __shfl_xor(llacc, off);
    if constexpr (LPR == 8) llacc += __shfl_xor(llacc, 7);
    if (lane < CP) red[wv * CP + lane] = llacc;
  }
  __syncthreads();
  double* part = A.part + ds.poff;
  for (int v = tid; v < A.C * D; v += kBThreads) {
    const int c = v / D, j = v % D, qq = j / F, jj = j % F;
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBThreads / 64; ++w)
#pragma unroll
      for (int k = 0; k < 64 / LPR; ++k) s += (double)gl[(w * 64 + k * LPR + qq) * GS + c * F + jj];
    part[(size_t)(c * VS + j) * ds.nblk + local] = s;
  }
  // the model has no intercept: its slot carries 0; the log-density only
  // when this pass is asked for it (0 otherwise, and the finish ignores it)
  if (tid < A.C) {
    double sl = 0.0;
    if constexpr (LL) {
#pragma unroll
      for (int w = 0; w < kBThreads / 64; ++w) sl += red[w * CP + tid];
    }
    part[(size_t)(tid * VS + D) * ds.nblk + local] = 0.0;
    part[(size_t)(tid * VS + D + 1) * ds.nblk + local] = sl;
  }
}

// LL = false: the gradient only, for a trajectory's leapfrog passes before
// its last (the log-density is read by the cache pass and the last pass).
template <bool LL>
__global__ __launch_bounds__(kBThreads) void k_bhlr_pass_rq(BArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int di = A.blk_ds[blockIdx.x];
  if (bsurr_skip(A, di)) return;
  const BDs ds = A.ds[di];
  switch (ds.d) {
    case 16: bhlr_body_rq<16, LL>(A, ds, smem); break;
    case 32: bhlr_body_rq<32, LL>(A, ds, smem); break;
    case 64: bhlr_body_rq<64, LL>(A, ds, smem); break;
    default: bhlr_body_r<128, 4>(A, ds, smem); break;
  }
}

// LDS of the register form: the lanes' gradient shares, then the chains' sums.
size_t bhlr_r_smem_bytes(int cp) {
  int f = 8;  // F * CP is largest at F = 8
  return ((sizeof(float) * kBThreads * (cp * f + 1) + 15) & ~(size_t)15) + sizeof(double) * 8 * cp;
}

template <int CP>
__global__ __launch_bounds__(kBThreads) void k_bhlr_pass(BArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int di = A.blk_ds[blockIdx.x];
  if (bsurr_skip(A, di)) return;
  const BDs ds = A.ds[di];
  switch (ds.d) {
    case 16: bhlr_body<16, CP>(A, ds, di, smem); break;
    case 32: bhlr_body<32, CP>(A, ds, di, smem); break;
    case 64: bhlr_body<64, CP>(A, ds, di, smem); break;
    default: bhlr_body<128, CP>(A, ds, di, smem); break;
  }
}

// ---------------------------------------------------------------------------
// Matrix-core form for 5..16 chains per dataset: the data pass is two GEMMs
// per 16-row tile, T = X B^T (rows x chains) and G += X^T A (features x
// chains), on V_MFMA with BF16 operands and FP32 accumulation.
//   - X is BF16 in memory, so its operands are exact.
//   - theta (FP32) enters as three BF16 parts hi + mid + lo (24 significant
//     bits), so t carries FP32-level error like the vector form.
//   - the adjoint a = y - sigmoid(t) enters as three parts as well, so the
//     gradient matches the vector form's FP32 accuracy (with two parts, 16
//     bits, trajectories of a 130-row, 128-feature dataset drifted from the
//     FP64 restatement by 3 % within two sweeps).
// Each wave takes 16-row tiles.  Lane l holds T[rows 4 (l / 16) + r][chain
// l % 16] after phase B, which is exactly the B-operand layout of the
// 16x16x16 product of phase C, so the adjoints stay in registers; X^T comes
// back from a per-wave LDS copy of the tile.  Padded chains (>= C) get
// theta = 0 and a = 0.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// a = h + m + l: each part rounded to BF16 to nearest even by the hardware
// conversion (V_CVT_PK_BF16_F32; identical to the integer round-to-nearest-
// even for every non-NaN float); the residuals are exact in FP32, so the
// three parts carry FP32's 24 significand bits.
__device__ inline void split_bf16x3(float a, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)a;
  const float r1 = a - (float)h;
  m = (__bf16)r1;
  l = (__bf16)(r1 - (float)m);
}

// split_bf16x3 of two values at once, each part of the pair packed in one
// 32-bit word (element 0 in the low half): one V_CVT_PK_BF16_F32 per part
// for both values, the same roundings and subtractions as split_bf16x3.
__device__ inline void split_bf16x3_pair(float a0, float a1, uint32_t& h, uint32_t& m, uint32_t& l) {
  typedef float f2_t __attribute__((ext_vector_type(2)));
  typedef __bf16 b2_t __attribute__((ext_vector_type(2)));
  auto cvt = [](float x, float y) { return __builtin_bit_cast(uint32_t, __builtin_convertvector(f2_t{x, y}, b2_t)); };
  auto lo = [](uint32_t w) { return __uint_as_float(w << 16); };
  auto hi = [](uint32_t w) { return __uint_as_float(w & 0xffff0000u); };
  h = cvt(a0, a1);
  const float r0 = a0 - lo(h), r1 = a1 - hi(h);
  m = cvt(r0, r1);
  l = cvt(r0 - lo(m), r1 - hi(m));
}

constexpr int kMfmaCols = 16;
constexpr int kMfRing = 4;  // 16-row tiles in flight per wave in the matrix-core form

template <int D>
constexpr int mfma_slab_stride() { return D + 8; }  // BF16 elements per LDS row

typedef short i16x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) i16x4_t lds_i16x4;  // 4 x 16 bits in LDS

// Phase-B chunking: 32 features per V_MFMA_F32_16X16X32_BF16 when D allows,
// else 16 per V_MFMA_F32_16X16X16_BF16; lane (i, kg) supplies features
// KB s + (KB / 4) kg + e of chunk s.
template <int KB> struct MfmaB;
template <> struct MfmaB<32> {
  typedef bf16x8_t vec;
  typedef uint4 raw;
  static __device__ f32x4_t mma(vec a, vec b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <> struct MfmaB<16> {
  typedef bf16x4_t vec;
  typedef uint2 raw;
  static __device__ f32x4_t mma(vec a, vec b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
};

template <int D, bool LL>
__device__ void bhlr_body_mfma(const BArgs& A, const BDs& ds, unsigned char* smem) {
  constexpr int KB = (D % 32 == 0) ? 32 : 16;  // features per phase-B product
  constexpr int KE = KB / 4;                   // features per lane per product
  constexpr int KC = D / KB;                   // phase-B chunks
  constexpr int FB = D / 16;                   // 16-feature blocks of phase C
  constexpr int XS = mfma_slab_stride<D>();
  static_assert(XS % 4 == 0, "8-byte rows for the transposing read");
  typedef typename MfmaB<KB>::vec bvec;
  typedef typename MfmaB<KB>::raw braw;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int li = lane & 15, kg = lane >> 4;
  uint16_t* slab = (uint16_t*)smem + (size_t)wv * 16 * XS;
  float* gred = (float*)(smem + (size_t)4 * 16 * XS * sizeof(uint16_t));  // [4][D][16]
  double* lred = (double*)(gred + 4 * D * kMfmaCols);                     // [4][16]

  // theta of chain li, features KB s + KE kg + e, as hi / mid / lo parts
  const bool live = li < A.C;
  bvec bh[KC], bm[KC], bl[KC];
#pragma unroll
  for (int s = 0; s < KC; ++s)
#pragma unroll
    for (int e = 0; e < KE; ++e) {
      // a real chain's value loaded for the padding columns, then zeroed:
      // the same loads on every path
      const float bv = A.beta32[(size_t)min(li, A.C - 1) * A.GW + ds.goff + KB * s + KE * kg + e];
      const float b = live ? bv : 0.0f;
      __bf16 h, m, l;
      split_bf16x3(b, h, m, l);
      bh[s][e] = h;
      bm[s][e] = m;
      bl[s][e] = l;
    }
  f32x4_t g[FB];
#pragma unroll
  for (int f = 0; f < FB; ++f) g[f] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  double ll = 0.0;

  const int local = blockIdx.x - ds.blk0;
  const int rbeg = local * A.rows_per_blk;
  const int rend = min(ds.n, rbeg + A.rows_per_blk);
  // this wave's tiles are rows w0 + 64 t; a ring of kMfRing tiles stays in
  // flight, the slot refilled as soon as its tile is in registers and LDS
  constexpr int PD = D >= 128 ? 2 : kMfRing;  // 128 features: half the ring for registers
  const int w0 = rbeg + 16 * wv;
  const int ntw = w0 < rend ? (rend - w0 + 63) / 64 : 0;
  braw qn[PD][KC];
  uint32_t yn[PD];
  // A tile starts at a multiple of 16 rows below rend <= n, inside the
  // dataset's rows padded to kRowPad: X and y load without bounds checks, y
  // as the 4 bytes of rows t0 + 4 kg .. + 3 in one aligned word.
  static_assert(kRowPad % 16 == 0, "16-row tiles within the padding");
  auto load_tile = [&](int t, braw (&qd)[KC], uint32_t& yd) {
    const int t0 = w0 + 64 * t;
    const uint16_t* xr = ds.x + (size_t)(t0 + li) * D + KE * kg;
#pragma unroll
    for (int s = 0; s < KC; ++s) qd[s] = ld_x((const braw*)(xr + KB * s), A.x_stream);
    yd = *gmem((const uint32_t*)(ds.y + t0 + 4 * kg));
  };
  // Every tile load is issued, one past the wave's last tile clamped to it
  // (tile 0 when the wave has none: rows w0 .. w0 + 15 lie inside the padded
  // rows), so each path through the loop issues the same loads and the
  // compiler's waits count the ring exactly; theta, requested before the
  // ring, is complete once at most PD (KC + 1) loads are outstanding.
  const int lastw = max(ntw - 1, 0);
#pragma unroll
  for (int k = 0; k < PD; ++k) load_tile(min(k, lastw), qn[k], yn[k]);
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(PD * (KC + 1)));
  for (int base = 0; base < ntw; base += PD) {
#pragma unroll
  for (int slot = 0; slot < PD; ++slot) {
    const int tt = base + slot;
    if (tt >= ntw) {  // uniform over the wave: its remaining slots are empty
      // the refill is still issued, so every path issues the same loads
      load_tile(min(tt + PD, lastw), qn[slot], yn[slot]);
      continue;
    }
    const int r0 = w0 + 64 * tt;
    // phase B: A operand = row r0 + li, features KB s + KE kg .. + KE - 1
#pragma unroll
    for (int s = 0; s < KC; ++s) *(braw*)(slab + li * XS + KB * s + KE * kg) = qn[slot][s];
    const uint32_t yw = yn[slot];
    f32x4_t t = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < KC; ++s) {
      const bvec xa = __builtin_bit_cast(bvec, qn[slot][s]);
      t = MfmaB<KB>::mma(xa, bh[s], t);
      t = MfmaB<KB>::mma(xa, bm[s], t);
      t = MfmaB<KB>::mma(xa, bl[s], t);
    }
    // the slot's registers are free once the products have read them
    load_tile(min(tt + PD, lastw), qn[slot], yn[slot]);
    // rows r0 + 4 kg + r, chain li
    float av[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = r0 + 4 * kg + r;
      const bool ok = rr < rend && live;
      float a;
      if constexpr (LL) {
        const float lr = log_sigmoid_pm(t[r], (int)((yw >> (8 * r)) & 0xffu), &a);
        if (ok) ll += (double)lr;
      } else {
        a = sigmoid_adjoint_pm(t[r], (int)((yw >> (8 * r)) & 0xffu));
      }
      av[r] = ok ? a : 0.0f;
    }
    // three BF16 parts of the 4 adjoints, two at a time
    uint32_t h01, m01, l01, h23, m23, l23;
    split_bf16x3_pair(av[0], av[1], h01, m01, l01);
    split_bf16x3_pair(av[2], av[3], h23, m23, l23);
    typedef uint32_t u2_t __attribute__((ext_vector_type(2)));
    const bf16x4_t ah = __builtin_bit_cast(bf16x4_t, u2_t{h01, h23});
    const bf16x4_t am = __builtin_bit_cast(bf16x4_t, u2_t{m01, m23});
    const bf16x4_t al = __builtin_bit_cast(bf16x4_t, u2_t{l01, l23});
    // the wave's tile copy is complete before any lane reads it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // phase C: A operand = X^T[feature 16 f + li][rows 4 kg + e], e = 0..3, by
    // one transposing LDS read: in each 16-lane group, lane i receives as
    // element e the (i % 4)-th 16-bit value at the address of lane 4 e + i / 4
    // of its group, so lane (li, kg) supplies row 4 kg + li / 4, columns
    // 16 f + 4 (li % 4) .. + 3 (8-byte aligned: XS is a multiple of 4)
    lds_i16x4* trow = (lds_i16x4*)(slab + (4 * kg + (li >> 2)) * XS + 4 * (li & 3));
#pragma unroll
    for (int f = 0; f < FB; ++f) {
      const bf16x4_t xt = __builtin_bit_cast(bf16x4_t, __builtin_amdgcn_ds_read_tr16_b64_v4i16(trow + 4 * f));
      g[f] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xt, ah, g[f], 0, 0, 0);
      g[f] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xt, am, g[f], 0, 0, 0);
      g[f] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xt, al, g[f], 0, 0, 0);
    }
    // the next tile's copy overwrites the slab only after these reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  }
  // g[f][r] = G[feature 16 f + 4 kg + r][chain li]; ll per (kg, chain li)
  ll += __shfl_xor(ll, 16);
  ll += __shfl_xor(ll, 32);
#pragma unroll
  for (int f = 0; f < FB; ++f)
#pragma unroll
    for (int r = 0; r < 4; ++r) gred[((size_t)wv * D + 16 * f + 4 * kg + r) * kMfmaCols + li] = g[f][r];
  if (lane < 16) lred[wv * kMfmaCols + lane] = ll;
  __syncthreads();
  // partials of this block, waves summed in order: part[poff + value * nblk + local]
  constexpr int VS = D + 2;
  double* part = A.part + ds.poff;
  for (int e = tid; e < A.C * D; e += kBThreads) {
    const int c = e / D, j = e % D;
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < 4; ++w) s += (double)gred[((size_t)w * D + j) * kMfmaCols + c];
    part[(size_t)(c * VS + j) * ds.nblk + local] = s;
  }
  if (tid < A.C) {
    part[(size_t)(tid * VS + D) * ds.nblk + local] = 0.0;  // no intercept
    part[(size_t)(tid * VS + D + 1) * ds.nblk + local] =
        ((lred[tid] + lred[kMfmaCols + tid]) + lred[2 * kMfmaCols + tid]) + lred[3 * kMfmaCols + tid];
  }
}

// LL = false: the gradient only (leapfrog passes before a trajectory's last)
template <bool LL>
__global__ __launch_bounds__(kBThreads) void k_bhlr_pass_mfma(BArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int di = A.blk_ds[blockIdx.x];
  if (bsurr_skip(A, di)) return;
  const BDs ds = A.ds[di];
  switch (ds.d) {
    case 16: bhlr_body_mfma<16, LL>(A, ds, smem); break;
    case 32: bhlr_body_mfma<32, LL>(A, ds, smem); break;
    case 64: bhlr_body_mfma<64, LL>(A, ds, smem); break;
    default: bhlr_body_mfma<128, LL>(A, ds, smem); break;
  }
}

size_t bhlr_mfma_smem_bytes(int dmax) {
  return sizeof(uint16_t) * 4 * 16 * (size_t)(dmax + 8) + sizeof(float) * 4 * (size_t)dmax * kMfmaCols +
         sizeof(double) * 4 * kMfmaCols;
}

size_t bhlr_smem_bytes(int dmax, int cp) {
  size_t b = sizeof(float) * ((size_t)kBTile * (dmax + 4) + (size_t)cp * (dmax + 4) + (size_t)kBTile * cp) +
             sizeof(int) * kBTile;
  b = (b + 15) & ~(size_t)15;
  b += sizeof(double) * 6 * kBThreads + 16;
  const size_t tot = sizeof(double) * (size_t)cp * (dmax + 2);
  return b > tot ? b : tot;
}

// One wave per (dataset, chain): momentum draw, half step, first position step.
__global__ __launch_bounds__(64) void k_bhmc_begin(BArgs A) {
  const int di = blockIdx.x, c = blockIdx.y, lane = threadIdx.x;
  if (A.smode && A.smode[di]) return;  // k_bsurr_traj moves this dataset
  const BDs ds = A.ds[di];
  const int D = ds.d;
  const uint32_t gc = (uint32_t)(di * A.C + c);
  const Stream mom{A.seed, gc, *A.sweep, TAG_HMC_BASE + 4u + 0u};
  const size_t si = (size_t)c * A.SW + ds.soff, gi = (size_t)c * A.GW + ds.goff;
  const double v = A.state[si + D + 1];
  const double step = theta_step(A, gc);
  double kin = 0.0;
  for (int j = lane; j < D + 2; j += 64) {
    const double sv = A.state[si + j];
    double xv = sv;
    if (j < D) {
      const double p0 = normal_at(mom, (uint32_t)j, 0);
      const double g = A.cache_g[gi + j] + (-sv / v);
      const double pj = p0 + 0.5 * step * g;
      A.p0[gi + j] = p0;
      A.p[gi + j] = pj;
      kin += p0 * p0;
      xv = sv + step * pj;
      A.beta32[gi + j] = (float)xv;
    }
    A.prop[si + j] = xv;
  }
  kin = wave_sum(kin);
  if (lane == 0) A.k0[(size_t)c * A.n_ds + di] = 0.5 * kin;
}

// One wave per (dataset, chain): energies, FP64 Metropolis test, accept.
__global__ __launch_bounds__(64) void k_bhmc_end(BArgs A) {
  const int di = blockIdx.x, c = blockIdx.y, lane = threadIdx.x;
  const BDs ds = A.ds[di];
  const int D = ds.d;
  const uint32_t gc = (uint32_t)(di * A.C + c);
  const size_t si = (size_t)c * A.SW + ds.soff, gi = (size_t)c * A.GW + ds.goff;
  const size_t ci = (size_t)c * A.n_ds + di;
  const double v = A.state[si + D + 1];
  double kin = 0.0, lp0 = 0.0, lp1 = 0.0;
  for (int j = lane; j < D; j += 64) {
    const double pj = A.p[gi + j];
    kin += pj * pj;
    lp0 += normal_ll0(A.state[si + j], v, A.ref_compat);
    lp1 += normal_ll0(A.prop[si + j], v, A.ref_compat);
  }
  kin = wave_sum(kin);
  lp0 = wave_sum(lp0);
  lp1 = wave_sum(lp1);
  const double ll0 = A.cache_ll[ci] + lp0, ll1 = A.llprop[ci] + lp1;
  const Stream acc_s{A.seed, gc, *A.sweep, TAG_HMC_BASE + 4u + 2u};
  const double u = u01(acc_s.draw(0, 0).x);
  const double lr = (ll1 - ll0) + (A.k0[ci] - 0.5 * kin);
  const bool acc = log(u) < lr;
  if (lane == 0) A.acc_theta[ci] = lr >= 0.0 ? 1.0 : (lr < 0.0 ? exp(lr) : 0.0);  // NaN: 0
  if (acc) {
    for (int j = lane; j < D; j += 64) {
      A.state[si + j] = A.prop[si + j];
      A.cache_g[gi + j] = A.gprop[gi + j];
    }
    if (lane == 0) A.cache_ll[ci] = A.llprop[ci];
  }
  if (lane == 0) {
    double* st = A.stats + ((size_t)1 * A.C * A.n_ds + ci) * 2;
    st[0] += acc ? 1.0 : 0.0;
    st[1] += 1.0;
  }
}

// HMC on each dataset's variance v (block 0): one wave per (dataset, chain).
__global__ __launch_bounds__(64) void k_bhmc_vblock(BArgs A) {
  const int di = blockIdx.x, c = blockIdx.y, lane = threadIdx.x;
  const BDs ds = A.ds[di];
  const int D = ds.d;
  const uint32_t gc = (uint32_t)(di * A.C + c);
  const uint32_t sw = *A.sweep;
  double* st = A.state + (size_t)c * A.SW + ds.soff;
  double S = 0.0;
  for (int j = lane; j < D; j += 64) S += st[j] * st[j];
  S = wave_sum(S);
  double v;
  const double eps = A.v_mwg ? A.vvar_ds[di] : A.v_eps;
  const bool acc = v_update(D, S, st[D + 1], A.ref_compat, A.prior_kind, A.prior_a, A.prior_b, A.v_mwg,
                            A.v_sim_len, eps, A.v_fixed, A.seed, gc, sw, TAG_HMC_BASE, &v);
  if (lane == 0) {
    st[D + 1] = v;
    A.acc_v[(size_t)c * A.n_ds + di] = acc ? 1.0 : 0.0;
    double* s2 = A.stats + ((size_t)c * A.n_ds + di) * 2;
    s2[0] += acc ? 1.0 : 0.0;
    s2[1] += 1.0;
  }
}

__global__ void k_bump_sweep(uint32_t* sweep) { *sweep += 1u; }

// Warmup adaptation after each sweep, one thread per dataset:
//  - theta step: dual averaging of log eps towards a mean acceptance
//    probability `target` over the dataset's chains (Hoffman & Gelman 2014,
//    Algorithm 5: gamma = 0.05, t0 = 10, kappa = 0.75, mu = log(10 eps0)),
//    kept within [eps0 e^-12, eps0 e^3];
//  - MWG proposal variance: Robbins-Monro, log var += (acceptance - 0.44) m^-0.6.
// ad[6 di ..] = log eps, log eps_bar, H_bar, mu, m, log eps0.  finalize sets
// eps to exp(log eps_bar), the value sampling keeps.
__global__ void k_badapt(BArgs A, double* __restrict__ ad, double target, int finalize) {
  const int di = blockIdx.x * blockDim.x + threadIdx.x;
  if (di >= A.n_ds) return;
  double* s = ad + 6 * (size_t)di;
  if (finalize) {
    A.eps_ds[di] = exp(s[1]);
    return;
  }
  double at = 0.0, av = 0.0;
  for (int c = 0; c < A.C; ++c) {
    at += A.acc_theta[(size_t)c * A.n_ds + di];
    av += A.acc_v[(size_t)c * A.n_ds + di];
  }
  at /= A.C;
  av /= A.C;
  const double m = s[4] + 1.0;
  s[4] = m;
  const double w = 1.0 / (m + 10.0);
  s[2] = (1.0 - w) * s[2] + w * (target - at);
  s[0] = fmin(fmax(s[3] - sqrt(m) / 0.05 * s[2], s[5] - 12.0), s[5] + 3.0);
  const double k = pow(m, -0.75);
  s[1] = k * s[0] + (1.0 - k) * s[1];
  A.eps_ds[di] = exp(s[0]);
  if (A.v_mwg)
    A.vvar_ds[di] = fmin(fmax(exp(log(A.vvar_ds[di]) + (av - 0.44) * pow(m, -0.6)), 1e-10), 1e10);
}

__global__ void k_bstate_to_beta32(BArgs A) {
  const int di = blockIdx.x, c = blockIdx.y;
  const BDs ds = A.ds[di];
  for (int j = threadIdx.x; j < ds.d; j += blockDim.x)
    A.beta32[(size_t)c * A.GW + ds.goff + j] = (float)A.state[(size_t)c * A.SW + ds.soff + j];
}

// ---------------------------------------------------------------------------
// Linear gradient surrogate.  Each dataset's data gradient is fitted by least
// squares on the (theta, gradient) pairs its exact leapfrog passes produced:
// g(theta) ~ gbar + W (theta - m), m and gbar the pairs' means.  Trajectories
// then use that gradient plus the exact prior gradient, and one exact pass
// at the proposal feeds the unchanged FP64 Metropolis test (k_bhmc_end), so
// the target is the exact one whatever W is.

// One block per (dataset, chain) in surrogate mode: momentum draw (the same
// Philox stream as k_bhmc_begin), then L leapfrog steps with the surrogate
// gradient; W sits in LDS (row stride D + 1).  Thread (o, g) with o = tid /
// segs, g = tid % segs takes row o's columns g * cps .. + cps - 1 of the
// matrix-vector product; the row's segs lanes are adjacent, summed by a
// butterfly; lane g = 0 owns theta_o and p_o.
__global__ __launch_bounds__(256) void k_bsurr_traj(BArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int di = blockIdx.x, c = blockIdx.y, tid = threadIdx.x;
  if (!A.smode[di]) return;
  const BDs ds = A.ds[di];
  const int D = ds.d, segs = 256 / D, cps = D / segs;
  const int o = tid / segs, g = tid % segs;
  float* wl = (float*)smem;            // D x (D + 1)
  float* df = wl + D * (D + 1);        // D: theta - m, FP32
  double* kin = (double*)(((uintptr_t)(df + D) + 15) & ~(uintptr_t)15);  // D
  const float* w = A.sw + ds.woff;
  for (int e = tid; e < D * D; e += 256) wl[(e / D) * (D + 1) + e % D] = w[e];
  const uint32_t gc = (uint32_t)(di * A.C + c);
  const size_t si = (size_t)c * A.SW + ds.soff, gi = (size_t)c * A.GW + ds.goff;
  const double v = A.state[si + D + 1];
  const double step = theta_step(A, gc);
  const Stream mom{A.seed, gc, *A.sweep, TAG_HMC_BASE + 4u + 0u};
  const double m_o = A.sm[ds.goff + o], gb_o = A.sg[ds.goff + o];
  double th = A.state[si + o];
  const double p0 = g == 0 ? normal_at(mom, (uint32_t)o, 0) : 0.0;
  if (g == 0) kin[o] = p0 * p0;
  auto grad = [&]() {
    __syncthreads();  // W loaded / the last product's reads of df done
    if (g == 0) df[o] = (float)(th - m_o);
    __syncthreads();
    float part = 0.0f;
    const float* wr = wl + o * (D + 1) + g * cps;
    const float* dr = df + g * cps;
    for (int k = 0; k < cps; ++k) part = fmaf(wr[k], dr[k], part);
    for (int off = 1; off < segs; off <<= 1) part += __shfl_xor(part, off);
    return gb_o + (double)part + (-th / v);
  };
  const int L = A.steps;
  double p = p0 + 0.5 * step * grad();
  for (int l = 1; l <= L; ++l) {
    th += step * p;
    const double gr = grad();
    p += (l == L ? 0.5 * step : step) * gr;
  }
  if (g == 0) {
    A.prop[si + o] = th;
    A.p[gi + o] = p;
    A.beta32[gi + o] = (float)th;
  }
  __syncthreads();
  if (tid == 0) {
    double k0 = 0.0;
    for (int j = 0; j < D; ++j) k0 += kin[j];
    A.k0[(size_t)c * A.n_ds + di] = 0.5 * k0;
    A.prop[si + D] = A.state[si + D];
    A.prop[si + D + 1] = v;
  }
}

size_t bsurr_traj_smem_bytes(int dmax) {
  size_t b = sizeof(float) * ((size_t)dmax * (dmax + 1) + dmax);
  return ((b + 15) & ~(size_t)15) + sizeof(double) * dmax;
}

// Means of the fitting pairs (all chains and slots), one block per dataset:
// thread j < D sums theta_j, thread D + j sums g_j, in slot and chain order.
__global__ __launch_bounds__(256) void k_bsurr_mean(BArgs A, const float* __restrict__ ct,
                                                    const float* __restrict__ cg, int slots,
                                                    double* __restrict__ sm, double* __restrict__ sg) {
  const int di = blockIdx.x, tid = threadIdx.x;
  const BDs ds = A.ds[di];
  const int D = ds.d;
  if (tid >= 2 * D) return;
  const int j = tid % D;
  const float* src = tid < D ? ct : cg;
  double acc = 0.0;
  for (int r = 0; r < slots; ++r)
    for (int c = 0; c < A.C; ++c) acc += (double)src[((size_t)r * A.C + c) * A.GW + ds.goff + j];
  const double mean = acc / ((double)slots * A.C);
  if (tid < D) sm[ds.goff + j] = mean;
  else sg[ds.goff + j] = mean;
}

// Centred cross products of the pairs, one dataset per blockIdx.x and 4096
// entries per blockIdx.y: G[j][e] over e < D is sum (theta_j - m_j)(theta_e
// - m_e), over D <= e < 2D sum (theta_j - m_j)(g_{e-D} - gbar_{e-D}).  Pairs
// are staged 16 at a time in LDS as FP64 deviations; every thread keeps 16
// entries in registers; sums run in slot and chain order.
constexpr int kGramPairs = 16;
constexpr int kGramPer = 16;
__global__ __launch_bounds__(256) void k_bsurr_gram(BArgs A, const float* __restrict__ ct,
                                                    const float* __restrict__ cg, int slots,
                                                    const double* __restrict__ sm,
                                                    const double* __restrict__ sg, double* __restrict__ G,
                                                    const size_t* __restrict__ goffs) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int di = blockIdx.x, tid = threadIdx.x;
  const BDs ds = A.ds[di];
  const int D = ds.d, W2 = 2 * D, ne = D * W2;
  const int e0 = blockIdx.y * 256 * kGramPer;
  if (e0 >= ne) return;
  double* dv = (double*)smem;  // kGramPairs x 2D: theta deviations, then g deviations
  int ja[kGramPer], ea[kGramPer];
  double acc[kGramPer];
#pragma unroll
  for (int t = 0; t < kGramPer; ++t) {
    const int e = e0 + t * 256 + tid;
    ja[t] = e < ne ? e / W2 : 0;
    ea[t] = e < ne ? e % W2 : 0;
    acc[t] = 0.0;
  }
  const int npairs = slots * A.C;
  for (int q0 = 0; q0 < npairs; q0 += kGramPairs) {
    __syncthreads();
    for (int i = tid; i < kGramPairs * W2; i += 256) {
      const int q = q0 + i / W2, k = i % W2;
      double dvi = 0.0;
      if (q < npairs) {
        const size_t base = (size_t)q * A.GW + ds.goff;  // pair q = slot * C + chain
        dvi = k < D ? (double)ct[base + k] - sm[ds.goff + k] : (double)cg[base + k - D] - sg[ds.goff + k - D];
      }
      dv[i] = dvi;
    }
    __syncthreads();
    for (int q = 0; q < kGramPairs; ++q) {
      const double* r = dv + q * W2;
#pragma unroll
      for (int t = 0; t < kGramPer; ++t) acc[t] = fma(r[ja[t]], r[ea[t]], acc[t]);
    }
  }
  double* gd = G + goffs[di];
#pragma unroll
  for (int t = 0; t < kGramPer; ++t) {
    const int e = e0 + t * 256 + tid;
    if (e < ne) gd[e] = acc[t];
  }
}

// The least-squares map of one dataset (one block per dataset), on the
// standardised inputs z_j = (theta_j - m_j) / s_j, s_j^2 = Cxx[j][j] / n:
// (R + ridge I) B = Cxg / (n s), R the pairs' correlation matrix, by
// Cholesky in place; then W[o][j] = B[j][o] / s_j (FP32).  A coordinate that
// never moved (s_j = 0) gets s_j = 1 and a zero column.
__global__ __launch_bounds__(256) void k_bsurr_solve(BArgs A, double* __restrict__ G,
                                                     const size_t* __restrict__ goffs, int npairs,
                                                     double ridge, float* __restrict__ Wout) {
  __shared__ double sc[128];
  const int di = blockIdx.x, tid = threadIdx.x;
  const BDs ds = A.ds[di];
  const int D = ds.d, W2 = 2 * D;
  double* g = G + goffs[di];
  const double n = (double)npairs;
  for (int j = tid; j < D; j += 256) {
    const double var = g[(size_t)j * W2 + j] / n;
    sc[j] = var > 0.0 ? sqrt(var) : 1.0;
  }
  __syncthreads();
  for (int e = tid; e < D * W2; e += 256) {
    const int j = e / W2, k = e % W2;
    double val = g[e] / (n * sc[j]);
    if (k < D) {
      val /= sc[k];
      if (k == j) val = (g[e] > 0.0 ? val : 1.0) + ridge;
    }
    g[e] = val;
  }
  __syncthreads();
  for (int k = 0; k < D; ++k) {
    if (tid == 0) {
      const double a = g[(size_t)k * W2 + k];
      g[(size_t)k * W2 + k] = sqrt(a > 1e-300 ? a : 1e-300);
    }
    __syncthreads();
    const double lkk = g[(size_t)k * W2 + k];
    for (int i = k + 1 + tid; i < D; i += 256) g[(size_t)i * W2 + k] /= lkk;
    __syncthreads();
    const int m = D - k - 1;
    for (int e = tid; e < m * m; e += 256) {
      const int i = k + 1 + e / m, j = k + 1 + e % m;
      if (j <= i) g[(size_t)i * W2 + j] -= g[(size_t)i * W2 + k] * g[(size_t)j * W2 + k];
    }
    __syncthreads();
  }
  // column o of the right-hand side: forward then backward substitution in
  // place (L y = b, L^T x = y)
  for (int o = tid; o < D; o += 256) {
    double* b = g + D + o;
    for (int i = 0; i < D; ++i) {
      double t = b[(size_t)i * W2];
      for (int j = 0; j < i; ++j) t -= g[(size_t)i * W2 + j] * b[(size_t)j * W2];
      b[(size_t)i * W2] = t / g[(size_t)i * W2 + i];
    }
    for (int i = D - 1; i >= 0; --i) {
      double t = b[(size_t)i * W2];
      for (int j = i + 1; j < D; ++j) t -= g[(size_t)j * W2 + i] * b[(size_t)j * W2];
      b[(size_t)i * W2] = t / g[(size_t)i * W2 + i];
    }
    for (int j = 0; j < D; ++j) Wout[ds.woff + (size_t)o * D + j] = (float)(b[(size_t)j * W2] / sc[j]);
  }
}

static uint16_t to_bf16(float f) {  // round to nearest even
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

typedef void (*BPassKernel)(BArgs);

static BPassKernel pick_bpass(int cp) {
  switch (cp) {
    case 1: return k_bhlr_pass<1>;
    case 2: return k_bhlr_pass<2>;
    case 4: return k_bhlr_pass<4>;
    case 8: return k_bhlr_pass<8>;
    default: return k_bhlr_pass<16>;
  }
}

// At 3-4 chains the form with its a x phase on the matrix cores, unless
// AUGUR_BHLR_R_VALU=1 (all-vector register form).
static BPassKernel pick_bpass_r(int cp) {
  const char* valu = getenv("AUGUR_BHLR_R_VALU");
  switch (cp) {
    case 1: return k_bhlr_pass_r<1>;
    case 2: return k_bhlr_pass_r<2>;
    default: return (valu && atoi(valu)) ? k_bhlr_pass_r<4> : k_bhlr_pass_rq<true>;
  }
}

// The gradient-only kernel of the same form for the leapfrog passes before a
// trajectory's last, or null where the form has none.
static BPassKernel pick_bpass_r_grad(int cp) {
  const char* valu = getenv("AUGUR_BHLR_R_VALU");
  return (cp == 4 && !(valu && atoi(valu))) ? k_bhlr_pass_rq<false> : nullptr;
}

struct BhlrSampler final : augur_sampler {
  int n_ds = 0, C = 1, SW = 0, GW = 0, dmax = 16, R = 4096, nblk_total = 0;
  std::vector<BDs> hds;
  std::vector<int> nrows, dims;
  DevBuf<uint16_t> x;
  DevBuf<uint8_t> y;
  DevBuf<BDs> dds;
  DevBuf<int> blk_ds;
  DevBuf<float> beta32;
  DevBuf<double> part, cache_g, cache_ll, gprop, llprop, prop, p, p0, k0, state, stats_buf;
  DevBuf<double> eps_ds, vvar_ds, acc_theta, acc_v;
  DevBuf<int> val_ds;
  DevBuf<uint32_t> sweep_dev;
  BArgs base{};
  BPassKernel pass = nullptr;
  BPassKernel pass_grad = nullptr;  // leapfrog passes before the last (null: pass)
  size_t pass_smem = 0;
  bool mfma = false;
  bool cache_valid = false, use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<float> tune_us;
  std::vector<int> tune_rows;
  size_t WT = 0;  // sum of D_i^2: the surrogate maps

  // ---- linear gradient surrogate ----
  enum { BS_OFF = 0, BS_COLLECT = 1, BS_ACTIVE = 2 };
  int bs_phase = BS_OFF;
  int bs_train = 0, bs_collected = 0, bs_window = 0, bs_since = 0, bs_fallbacks = 0, bs_active = 0;
  long long bs_sweeps = 0;
  double bs_ridge = 1e-6, bs_min_accept = 0.0, bs_fit_ms = 0.0;
  DevBuf<float> cap_t, cap_g, surr_w;
  DevBuf<double> surr_m, surr_g, gram;
  DevBuf<size_t> gram_off;
  DevBuf<int> smode;
  std::vector<int> hsmode;
  // Layout of the leapfrog passes before the last while some datasets run
  // the surrogate: only the exact datasets' rows, in blocks of R_ex rows
  // (a few small datasets would otherwise leave most CUs idle), with their
  // own block table and partials.  The other datasets have no blocks.
  DevBuf<BDs> dds_ex;
  DevBuf<int> blk_ds_ex;
  DevBuf<double> part_ex;
  int nblk_ex = 0, R_ex = 0;
  std::vector<double> win0;  // theta-block (accepted, proposed) [C][n_ds] at the window's start
  size_t traj_smem = 0;

  ~BhlrSampler() override {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
  }

  const char* family() const override { return "bhlr"; }

  ParamView param(const std::string& name) override {
    ParamView v;
    if (name == "state") return {state.p, sizeof(double), (size_t)SW, (size_t)SW, true};
    if (name.rfind("state:", 0) == 0) {
      const int i = atoi(name.c_str() + 6);
      if (i >= 0 && i < n_ds) v = {state.p + hds[i].soff, sizeof(double), (size_t)hds[i].d + 2, (size_t)SW, true};
    }
    return v;
  }

  void on_set(const std::string&) override {
    cache_valid = false;
    drop_graph();
  }

  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }

  // Blocks of R rows per dataset; partial offsets follow.
  void layout(int rows) {
    // blocks start at multiples of kRowPad rows, so the last block's tiles
    // end inside its dataset's padding
    if (rows < kRowPad || rows % kRowPad != 0 || rows % kBTile != 0)
      fail(AUGUR_ERR_ARG, "rows per block %d is not a positive multiple of %d", rows, kRowPad);
    R = rows;
    int blk = 0, poff = 0, voff = 0;
    std::vector<int> table, vtable;
    for (int i = 0; i < n_ds; ++i) {
      BDs& d = hds[i];
      d.nblk = (d.n + R - 1) / R;
      d.blk0 = blk;
      d.poff = poff;
      d.voff = voff;
      for (int b = 0; b < d.nblk; ++b) table.push_back(i);
      for (int v = 0; v < C * (d.d + 2); ++v) vtable.push_back(i);
      blk += d.nblk;
      poff += C * (d.d + 2) * d.nblk;
      voff += C * (d.d + 2);
    }
    nblk_total = blk;
    part.alloc((size_t)poff);
    blk_ds.alloc(table.size());
    blk_ds.upload(table.data(), table.size(), stream);
    val_ds.alloc(vtable.size());
    val_ds.upload(vtable.data(), vtable.size(), stream);
    dds.upload(hds.data(), hds.size(), stream);
    AUGUR_HIP(hipStreamSynchronize(stream));  // host tables are rebuilt by the next call
    base.ds = dds.p;
    base.blk_ds = blk_ds.p;
    base.val_ds = val_ds.p;
    base.nval = voff;
    base.rows_per_blk = R;
    base.part = part.p;
    // nontemporal X loads: 15-20 % faster passes on the 256-dataset workload
    // (2 GB of BF16 X per pass, far beyond the Infinity Cache)
    base.x_stream = getenv("AUGUR_BHLR_TEMPORAL") == nullptr;
  }

  // exact_only: the exact datasets' layout (dds_ex) instead of every block
  void launch_pass(int mode, int l, bool exact_only = false) {
    BArgs A = base;
    A.mode = mode;
    A.l = l;
    if (exact_only) {
      A.ds = dds_ex.p;
      A.blk_ds = blk_ds_ex.p;
      A.rows_per_blk = R_ex;
      A.part = part_ex.p;
    }
    if (bs_phase == BS_COLLECT && mode == BMODE_LEAPFROG) {
      const size_t slot = (size_t)bs_collected * base.steps + (size_t)(l - 1);
      A.cap_t = cap_t.p + slot * C * GW;
      A.cap_g = cap_g.p + slot * C * GW;
    }
    const BPassKernel k = (mode == BMODE_LEAPFROG && l < base.steps && pass_grad) ? pass_grad : pass;
    prof_begin();
    k<<<exact_only ? nblk_ex : nblk_total, kBThreads, pass_smem, stream>>>(A);
    AUGUR_CHECK_LAUNCH();
    prof_end();
    k_bhlr_finish<<<(A.nval + 255) / 256, 256, 0, stream>>>(A);
    AUGUR_CHECK_LAUNCH();
  }

  int next_cp() const {
    int p = 1;
    while (p < C) p <<= 1;
    return p;
  }

  void refresh_cache() {
    k_bstate_to_beta32<<<dim3(n_ds, C), 64, 0, stream>>>(base);
    AUGUR_CHECK_LAUNCH();
    launch_pass(BMODE_CACHE, 0);
    cache_valid = true;
  }

  // Times the cache pass for candidate rows-per-block and keeps the fastest.
  void autotune() {
    const int cands[] = {1024, 2048, 4096, 8192, 16384};
    hipEvent_t e0, e1;
    AUGUR_HIP(hipEventCreate(&e0));
    AUGUR_HIP(hipEventCreate(&e1));
    int best = 4096;
    float best_us = 1e30f;
    bool first = true;
    for (int rows : cands) {
      layout(rows);  // sets base.ds and the block / value tables
      if (first) {
        k_bstate_to_beta32<<<dim3(n_ds, C), 64, 0, stream>>>(base);
        AUGUR_CHECK_LAUNCH();
        first = false;
      }
      BArgs A = base;
      A.mode = BMODE_CACHE;
      // a leapfrog step's cost: the pass and the finish, whose work grows
      // with the number of blocks
      auto step_once = [&]() {
        pass<<<nblk_total, kBThreads, pass_smem, stream>>>(A);
        AUGUR_CHECK_LAUNCH();
        k_bhlr_finish<<<(A.nval + 255) / 256, 256, 0, stream>>>(A);
        AUGUR_CHECK_LAUNCH();
      };
      step_once();  // warm
      AUGUR_HIP(hipEventRecord(e0, stream));
      for (int r = 0; r < 3; ++r) step_once();
      AUGUR_HIP(hipEventRecord(e1, stream));
      AUGUR_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      AUGUR_HIP(hipEventElapsedTime(&ms, e0, e1));
      const float us = 1000.0f * ms / 3.0f;
      tune_rows.push_back(rows);
      tune_us.push_back(us);
      if (us < best_us) {
        best_us = us;
        best = rows;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    layout(best);
  }

  void enqueue_sweep() {
    k_bhmc_vblock<<<dim3(n_ds, C), 64, 0, stream>>>(base);
    AUGUR_CHECK_LAUNCH();
    if (bs_phase == BS_ACTIVE && bs_active > 0) {
      // surrogate datasets: one trajectory kernel and the last pass only
      k_bsurr_traj<<<dim3(n_ds, C), 256, traj_smem, stream>>>(base);
      AUGUR_CHECK_LAUNCH();
    }
    if (bs_phase == BS_ACTIVE && bs_active == n_ds) {
      launch_pass(BMODE_LEAPFROG, base.steps);
    } else if (bs_phase == BS_ACTIVE && bs_active > 0) {
      // datasets back on exact HMC: their blocks only until the last pass,
      // which every dataset runs (the surrogate ones evaluate their proposal)
      k_bhmc_begin<<<dim3(n_ds, C), 64, 0, stream>>>(base);
      AUGUR_CHECK_LAUNCH();
      for (int l = 1; l < base.steps; ++l) launch_pass(BMODE_LEAPFROG, l, true);
      launch_pass(BMODE_LEAPFROG, base.steps);
    } else {
      k_bhmc_begin<<<dim3(n_ds, C), 64, 0, stream>>>(base);
      AUGUR_CHECK_LAUNCH();
      for (int l = 1; l <= base.steps; ++l) launch_pass(BMODE_LEAPFROG, l);
    }
    k_bhmc_end<<<dim3(n_ds, C), 64, 0, stream>>>(base);
    AUGUR_CHECK_LAUNCH();
    k_bump_sweep<<<1, 1, 0, stream>>>(sweep_dev.p);
    AUGUR_CHECK_LAUNCH();
  }

  void run_sweeps(int nsw) override {
    if (nsw <= 0) return;
    if (!cache_valid) refresh_cache();
    for (int i = 0; i < nsw; ++i) {
      // collection sweeps name their capture slots in kernel arguments, so
      // they are launched directly; a phase change drops the graph
      if (use_graph && prof_period == 0 && bs_phase != BS_COLLECT) {
        if (!graph_exec) {
          AUGUR_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
          enqueue_sweep();
          AUGUR_HIP(hipStreamEndCapture(stream, &graph));
          AUGUR_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
        }
        AUGUR_HIP(hipGraphLaunch(graph_exec, stream));
      } else {
        enqueue_sweep();
      }
      ++sweep;
      record_traces();
      bs_after_sweep();
    }
  }

  // ---- surrogate phases ----
  void bs_configure(int train, double ridge, double min_accept, int window) {
    const size_t slots = (size_t)train * base.steps;
    const size_t bytes = 2 * sizeof(float) * slots * C * GW;
    if (bytes > ((size_t)32 << 30))
      fail(AUGUR_ERR_ARG, "%zu bytes of fitting pairs (train_sweeps x steps x chains x sum (D_i + 1) x 8) exceed 32 GiB",
           bytes);
    cap_t.alloc(slots * C * GW);
    cap_g.alloc(slots * C * GW);
    surr_w.alloc(WT);
    surr_m.alloc(GW);
    surr_g.alloc(GW);
    std::vector<size_t> go(n_ds);
    size_t tot = 0;
    for (int i = 0; i < n_ds; ++i) {
      go[i] = tot;
      tot += 2 * (size_t)hds[i].d * hds[i].d;
    }
    gram.alloc(tot);
    gram_off.alloc(n_ds);
    gram_off.upload(go.data(), go.size(), stream);
    smode.alloc(n_ds);
    traj_smem = bsurr_traj_smem_bytes(dmax);
    allow_lds(k_bsurr_traj, traj_smem);
    allow_lds(k_bsurr_gram, sizeof(double) * kGramPairs * 2 * dmax);
    AUGUR_HIP(hipStreamSynchronize(stream));  // go is released on return
    bs_train = train;
    bs_ridge = ridge;
    bs_min_accept = min_accept;
    bs_window = window;
    bs_collected = 0;
    bs_phase = BS_COLLECT;
    drop_graph();
  }

  void bs_after_sweep() {
    if (bs_phase == BS_COLLECT) {
      if (++bs_collected >= bs_train) bs_fit();
    } else if (bs_phase == BS_ACTIVE && bs_active > 0) {
      ++bs_sweeps;
      if (bs_window > 0 && ++bs_since >= bs_window) bs_check_window();
    }
  }

  std::vector<double> theta_counts() {
    std::vector<double> st(2 * (size_t)C * n_ds * 2);
    stats_buf.download(st.data(), st.size(), stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
    return std::vector<double>(st.begin() + (size_t)C * n_ds * 2, st.end());
  }

  void bs_fit() {
    AUGUR_HIP(hipStreamSynchronize(stream));
    const auto t0 = std::chrono::steady_clock::now();
    const int slots = bs_collected * base.steps;
    k_bsurr_mean<<<n_ds, 256, 0, stream>>>(base, cap_t.p, cap_g.p, slots, surr_m.p, surr_g.p);
    AUGUR_CHECK_LAUNCH();
    const dim3 gg(n_ds, (unsigned)((2 * (size_t)dmax * dmax + 256 * kGramPer - 1) / (256 * kGramPer)));
    k_bsurr_gram<<<gg, 256, sizeof(double) * kGramPairs * 2 * dmax, stream>>>(
        base, cap_t.p, cap_g.p, slots, surr_m.p, surr_g.p, gram.p, gram_off.p);
    AUGUR_CHECK_LAUNCH();
    k_bsurr_solve<<<n_ds, 256, 0, stream>>>(base, gram.p, gram_off.p, slots * C, bs_ridge, surr_w.p);
    AUGUR_CHECK_LAUNCH();
    hsmode.assign(n_ds, 1);
    smode.upload(hsmode.data(), hsmode.size(), stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
    bs_fit_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    base.smode = smode.p;
    base.sw = surr_w.p;
    base.sm = surr_m.p;
    base.sg = surr_g.p;
    bs_active = n_ds;
    bs_phase = BS_ACTIVE;
    bs_since = 0;
    win0 = theta_counts();
    drop_graph();
  }

  // Datasets whose acceptance over the last window fell below min_accept
  // return to exact HMC for good.
  void bs_check_window() {
    const std::vector<double> st = theta_counts();
    bool changed = false;
    for (int i = 0; i < n_ds; ++i) {
      if (!hsmode[i]) continue;
      double a = 0.0, t = 0.0;
      for (int c = 0; c < C; ++c) {
        const size_t k = ((size_t)c * n_ds + i) * 2;
        a += st[k] - win0[k];
        t += st[k + 1] - win0[k + 1];
      }
      if (t > 0.0 && a < bs_min_accept * t) {
        hsmode[i] = 0;
        ++bs_fallbacks;
        --bs_active;
        changed = true;
      }
    }
    win0 = st;
    bs_since = 0;
    if (changed) bs_modes_changed();
  }

  // hsmode changed: upload it and rebuild the exact datasets' layout, in
  // blocks of R_ex rows (a multiple of kRowPad, at most R) chosen so that
  // their rows make about 512 blocks
  void bs_modes_changed() {
    AUGUR_HIP(hipStreamSynchronize(stream));  // no launch still reads the old layout
    drop_graph();  // the graph's passes name the old layout
    long long rows = 0;
    bs_active = 0;
    for (int i = 0; i < n_ds; ++i) {
      bs_active += hsmode[i];
      if (!hsmode[i]) rows += hds[i].n;
    }
    const long long want = (rows + 511) / 512;
    R_ex = (int)std::min<long long>(R, std::max<long long>(kRowPad, (want + kRowPad - 1) / kRowPad * kRowPad));
    std::vector<BDs> ex = hds;
    std::vector<int> table;
    int blk = 0;
    size_t poff = 0;
    for (int i = 0; i < n_ds; ++i) {
      BDs& d = ex[i];
      d.nblk = hsmode[i] ? 0 : (d.n + R_ex - 1) / R_ex;
      d.blk0 = blk;
      d.poff = (int)poff;
      for (int b = 0; b < d.nblk; ++b) table.push_back(i);
      blk += d.nblk;
      poff += (size_t)C * (d.d + 2) * d.nblk;
    }
    nblk_ex = blk;
    smode.upload(hsmode.data(), hsmode.size(), stream);
    dds_ex.alloc(ex.size());
    dds_ex.upload(ex.data(), ex.size(), stream);
    blk_ds_ex.alloc(table.size());
    blk_ds_ex.upload(table.data(), table.size(), stream);
    part_ex.alloc(poff);
    AUGUR_HIP(hipStreamSynchronize(stream));  // the host vectors go out of scope
  }

  // Warmup sweeps that adapt each dataset's theta step size (and the MWG
  // proposal variance of v) after every sweep, then freeze them.  Runs
  // without the graph; the graph reads the frozen values from device memory.
  void adapt(int nsw, double target) {
    if (nsw <= 0) return;
    if (!cache_valid) refresh_cache();
    std::vector<double> e(n_ds), ad(6 * (size_t)n_ds);
    eps_ds.download(e.data(), e.size(), stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
    for (int i = 0; i < n_ds; ++i) {
      const double l0 = log(e[i]);
      ad[6 * (size_t)i + 0] = l0;
      ad[6 * (size_t)i + 1] = 0.0;
      ad[6 * (size_t)i + 2] = 0.0;
      ad[6 * (size_t)i + 3] = log(10.0) + l0;
      ad[6 * (size_t)i + 4] = 0.0;
      ad[6 * (size_t)i + 5] = l0;
    }
    DevBuf<double> dad(ad.size());
    dad.upload(ad.data(), ad.size(), stream);
    for (int i = 0; i < nsw; ++i) {
      enqueue_sweep();
      k_badapt<<<(n_ds + 63) / 64, 64, 0, stream>>>(base, dad.p, target, 0);
      AUGUR_CHECK_LAUNCH();
      ++sweep;
      record_traces();
      bs_after_sweep();
    }
    k_badapt<<<(n_ds + 63) / 64, 64, 0, stream>>>(base, dad.p, target, 1);
    AUGUR_CHECK_LAUNCH();
    AUGUR_HIP(hipStreamSynchronize(stream));  // dad is released on return
  }

  void log_density(double* out) override {
    if (!cache_valid) refresh_cache();
    std::vector<double> ll((size_t)C * n_ds), st((size_t)C * SW);
    cache_ll.download(ll.data(), ll.size(), stream);
    state.download(st.data(), st.size(), stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
    // per chain: sum over datasets of data ll (device) + priors (host, FP64)
    for (int c = 0; c < chains; ++c) {
      double acc = 0.0;
      for (int i = 0; i < n_ds; ++i) {
        const double* s = st.data() + (size_t)c * SW + hds[i].soff;
        const int D = hds[i].d;
        const double v = s[D + 1];
        const double cc = base.ref_compat ? 1.0 : 0.5;
        for (int j = 0; j < D; ++j) acc += v > 0.0 ? -cc * log(v) - 0.5 * s[j] * s[j] / v : -kLargePosNum;
        if (base.prior_kind == 0)
          acc += base.prior_a * log(base.prior_b) - lgamma(base.prior_a) - (base.prior_a + 1.0) * log(v) - base.prior_b / v;
        else
          acc += (v > 0.0) ? log(base.prior_a) - base.prior_a * v : -kLargePosNum;
        acc += ll[(size_t)c * n_ds + i];
      }
      out[c] = acc;
    }
  }

  void stats(double* out, size_t bytes) override {
    const size_t need = sizeof(double) * 2 * (size_t)C * n_ds * 2;
    if (bytes < need) fail(AUGUR_ERR_ARG, "stats need %zu bytes", need);
    stats_buf.download(out, 2 * (size_t)C * n_ds * 2, stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
  }
};

}  // namespace augur

using namespace augur;

extern "C" {

augur_status augur_bhlr_create(const augur_bhlr_desc* D, augur_sampler** out) {
  return guard([&] {
    if (!out || !D) fail(AUGUR_ERR_ARG, "null argument");
    *out = nullptr;
    if (D->n_datasets < 1 || !D->n_rows || !D->n_features || !D->x || !D->y)
      fail(AUGUR_ERR_ARG, "need n_datasets >= 1 and per-dataset arrays");
    if (D->chains < 1 || D->chains > kMaxChains) fail(AUGUR_ERR_ARG, "chains must be in [1, %d]", kMaxChains);
    if (D->theta_block.fixed_steps == 0 || D->theta_block.fixed_steps > 100000 ||
        D->theta_block.fixed_steps < -100000 || !(D->theta_block.step_size > 0.0))
      fail(AUGUR_ERR_ARG, "the batched theta block needs fixed_steps != 0 and step_size > 0");
    if (D->v_block.block != AUGUR_HMC_BLOCK_V && D->v_block.block != AUGUR_MWG_BLOCK_V)
      fail(AUGUR_ERR_ARG, "the batched v block is HMC or MWG");
    if (!(D->v_block.step_size > 0.0) ||
        (D->v_block.block == AUGUR_HMC_BLOCK_V && D->v_block.fixed_steps == 0 && !(D->v_block.sim_len > 0.0)))
      fail(AUGUR_ERR_ARG, "bad v block settings");
    if (D->prior_kind != 0 && D->prior_kind != 1) fail(AUGUR_ERR_ARG, "prior_kind must be 0 or 1");
    for (int i = 0; i < D->n_datasets; ++i) {
      const int d = D->n_features[i];
      if (D->n_rows[i] < 1 || !(d == 16 || d == 32 || d == 64 || d == 128))
        fail(AUGUR_ERR_ARG, "dataset %d: need n >= 1 and d in {16, 32, 64, 128}", i);
      if (!D->x[i] || !D->y[i]) fail(AUGUR_ERR_ARG, "dataset %d: null data", i);
      for (int r = 0; r < D->n_rows[i]; ++r)
        if (D->y[i][r] != 0 && D->y[i][r] != 1) fail(AUGUR_ERR_ARG, "dataset %d: y must be 0 or 1", i);
    }
    require_device();
    auto* s = new BhlrSampler();
    try {
      AUGUR_HIP(hipGetDevice(&s->device));
      AUGUR_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
      s->seed = D->seed;
      s->n_ds = D->n_datasets;
      s->C = s->chains = D->chains;
      s->use_graph = D->use_graph != 0;
      size_t xtot = 0, ytot = 0;
      int soff = 0, goff = 0;
      s->hds.resize(s->n_ds);
      for (int i = 0; i < s->n_ds; ++i) {
        BDs& b = s->hds[i];
        b.n = D->n_rows[i];
        b.d = D->n_features[i];
        b.soff = soff;
        b.goff = goff;
        b.woff = (int)s->WT;
        soff += b.d + 2;
        goff += b.d + 1;
        s->WT += (size_t)b.d * b.d;
        s->dmax = std::max(s->dmax, b.d);
        xtot += (size_t)padded_rows(b.n) * b.d;
        ytot += (size_t)padded_rows(b.n);
      }
      s->SW = soff;
      s->GW = goff;
      s->x.alloc(xtot);
      s->y.alloc(ytot);
      {
        std::vector<uint16_t> hx;
        std::vector<uint8_t> hy;
        size_t xo = 0, yo = 0;
        for (int i = 0; i < s->n_ds; ++i) {
          BDs& b = s->hds[i];
          const size_t ne = (size_t)b.n * b.d, np = (size_t)padded_rows(b.n);
          hx.assign(np * b.d, 0);
          for (size_t e = 0; e < ne; ++e) hx[e] = to_bf16(D->x[i][e]);
          hy.assign(np, 0);
          for (int r = 0; r < b.n; ++r) hy[r] = (uint8_t)D->y[i][r];
          AUGUR_HIP(hipMemcpy(s->x.p + xo, hx.data(), np * b.d * 2, hipMemcpyHostToDevice));
          AUGUR_HIP(hipMemcpy(s->y.p + yo, hy.data(), np, hipMemcpyHostToDevice));
          b.x = s->x.p + xo;
          b.y = s->y.p + yo;
          xo += np * b.d;
          yo += np;
        }
      }
      const size_t C = s->C;
      s->dds.alloc(s->n_ds);
      s->beta32.alloc(C * s->GW);
      s->cache_g.alloc(C * s->GW); s->gprop.alloc(C * s->GW); s->p.alloc(C * s->GW); s->p0.alloc(C * s->GW);
      s->cache_ll.alloc(C * s->n_ds); s->llprop.alloc(C * s->n_ds); s->k0.alloc(C * s->n_ds);
      s->state.alloc(C * s->SW); s->prop.alloc(C * s->SW);
      s->stats_buf.alloc(2 * C * s->n_ds * 2);
      s->sweep_dev.alloc(1);
      s->state.zero(s->stream); s->prop.zero(s->stream); s->stats_buf.zero(s->stream);
      s->sweep_dev.zero(s->stream); s->beta32.zero(s->stream);
      BArgs& A = s->base;
      A.n_ds = s->n_ds; A.C = s->C; A.SW = s->SW; A.GW = s->GW;
      A.beta32 = s->beta32.p;
      A.cache_g = s->cache_g.p; A.cache_ll = s->cache_ll.p; A.gprop = s->gprop.p; A.llprop = s->llprop.p;
      A.prop = s->prop.p; A.p = s->p.p; A.p0 = s->p0.p; A.k0 = s->k0.p; A.state = s->state.p;
      A.stats = s->stats_buf.p; A.sweep = s->sweep_dev.p; A.seed = s->seed;
      A.prior_kind = D->prior_kind; A.ref_compat = D->ref_compat;
      A.prior_a = D->prior_a; A.prior_b = D->prior_b;
      A.step = D->theta_block.step_size; A.steps = std::abs(D->theta_block.fixed_steps);
      A.step_jitter = D->theta_block.fixed_steps < 0;
      A.v_sim_len = D->v_block.sim_len; A.v_eps = D->v_block.step_size; A.v_fixed = D->v_block.fixed_steps;
      A.v_mwg = D->v_block.block == AUGUR_MWG_BLOCK_V;
      {
        const std::vector<double> e(s->n_ds, D->theta_block.step_size), vv(s->n_ds, D->v_block.step_size);
        s->eps_ds.alloc(s->n_ds);
        s->vvar_ds.alloc(s->n_ds);
        s->eps_ds.upload(e.data(), e.size(), s->stream);
        s->vvar_ds.upload(vv.data(), vv.size(), s->stream);
        s->acc_theta.alloc((size_t)C * s->n_ds);
        s->acc_v.alloc((size_t)C * s->n_ds);
        s->acc_theta.zero(s->stream);
        s->acc_v.zero(s->stream);
        AUGUR_HIP(hipStreamSynchronize(s->stream));  // e and vv are released at the block's end
        A.eps_ds = s->eps_ds.p; A.vvar_ds = s->vvar_ds.p; A.acc_theta = s->acc_theta.p; A.acc_v = s->acc_v.p;
      }
      // matrix cores from 5 chains on (the 16 chain columns of a product are
      // then at least a third used); AUGUR_BHLR_VECTOR=1 keeps the vector form
      const char* vec = getenv("AUGUR_BHLR_VECTOR");
      // up to 4 chains the register form, AUGUR_BHLR_PASS_LDS=1 the LDS-staged one
      const char* lds = getenv("AUGUR_BHLR_PASS_LDS");
      const int cpn = s->next_cp();
      s->mfma = cpn >= 8 && !(vec && atoi(vec));
      if (s->mfma) {
        s->pass = k_bhlr_pass_mfma<true>;
        s->pass_grad = k_bhlr_pass_mfma<false>;
        s->pass_smem = bhlr_mfma_smem_bytes(s->dmax);
      } else if (cpn <= 4 && !(lds && atoi(lds))) {
        s->pass = pick_bpass_r(cpn);
        s->pass_grad = pick_bpass_r_grad(cpn);
        s->pass_smem = bhlr_r_smem_bytes(cpn);
      } else {
        s->pass = pick_bpass(cpn);
        s->pass_smem = bhlr_smem_bytes(s->dmax, cpn);
      }
      // AUGUR_BHLR_GRAD_PASS=0: every pass computes the log-density too
      if (const char* gp = getenv("AUGUR_BHLR_GRAD_PASS"))
        if (atoi(gp) == 0) s->pass_grad = nullptr;
      allow_lds(s->pass, s->pass_smem);
      if (s->pass_grad) allow_lds(s->pass_grad, s->pass_smem);
      // state starts at theta = 0, v = 1
      std::vector<double> st(C * s->SW, 0.0);
      for (size_t c = 0; c < C; ++c)
        for (int i = 0; i < s->n_ds; ++i) st[c * s->SW + s->hds[i].soff + s->hds[i].d + 1] = 1.0;
      s->state.upload(st.data(), st.size(), s->stream);
      if (D->rows_per_block > 0) s->layout(padded_rows(D->rows_per_block));
      else s->autotune();
      AUGUR_HIP(hipStreamSynchronize(s->stream));
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

augur_status augur_bhlr_tuning(augur_sampler* s, int* rows_per_block, int* n_candidates,
                               int* cand_rows, float* cand_us, int cap) {
  return guard([&] {
    auto* b = dynamic_cast<BhlrSampler*>(s);
    if (!b) fail(AUGUR_ERR_ARG, "not a batched HLR sampler");
    if (rows_per_block) *rows_per_block = b->R;
    if (n_candidates) *n_candidates = (int)b->tune_rows.size();
    for (int i = 0; i < (int)b->tune_rows.size() && i < cap; ++i) {
      if (cand_rows) cand_rows[i] = b->tune_rows[i];
      if (cand_us) cand_us[i] = b->tune_us[i];
    }
  });
}

augur_status augur_bhlr_adapt(augur_sampler* s, int n_sweeps, double target_accept) {
  return guard([&] {
    auto* b = dynamic_cast<BhlrSampler*>(s);
    if (!b) fail(AUGUR_ERR_ARG, "not a batched HLR sampler");
    if (n_sweeps < 0 || !(target_accept > 0.0 && target_accept < 1.0))
      fail(AUGUR_ERR_ARG, "need n_sweeps >= 0 and 0 < target_accept < 1");
    AUGUR_HIP(hipSetDevice(b->device));
    b->adapt(n_sweeps, target_accept);
  });
}

augur_status augur_bhlr_step_sizes(augur_sampler* s, double* eps, double* v_var, int n) {
  return guard([&] {
    auto* b = dynamic_cast<BhlrSampler*>(s);
    if (!b) fail(AUGUR_ERR_ARG, "not a batched HLR sampler");
    if (n < b->n_ds) fail(AUGUR_ERR_ARG, "need room for %d datasets", b->n_ds);
    AUGUR_HIP(hipSetDevice(b->device));
    if (eps) b->eps_ds.download(eps, b->n_ds, b->stream);
    if (v_var) b->vvar_ds.download(v_var, b->n_ds, b->stream);
    AUGUR_HIP(hipStreamSynchronize(b->stream));
  });
}


augur_status augur_bhlr_surrogate(augur_sampler* s, const augur_bsurr_cfg* cfg) {
  return guard([&] {
    auto* b = dynamic_cast<BhlrSampler*>(s);
    if (!b || !cfg) fail(AUGUR_ERR_ARG, "needs a batched HLR sampler and a configuration");
    if (b->bs_phase != BhlrSampler::BS_OFF) fail(AUGUR_ERR_STATE, "surrogate already configured");
    if (cfg->train_sweeps < 1 || cfg->window < 0 || !(cfg->ridge >= 0.0) || !(cfg->min_accept >= 0.0))
      fail(AUGUR_ERR_ARG, "need train_sweeps >= 1, window >= 0, ridge >= 0 and min_accept >= 0");
    AUGUR_HIP(hipSetDevice(b->device));
    b->bs_configure(cfg->train_sweeps, cfg->ridge, cfg->min_accept, cfg->window);
  });
}

augur_status augur_bhlr_surrogate_info(augur_sampler* s, double* info, size_t bytes) {
  return guard([&] {
    auto* b = dynamic_cast<BhlrSampler*>(s);
    if (!b || !info) fail(AUGUR_ERR_ARG, "needs a batched HLR sampler and an output");
    if (bytes < 7 * sizeof(double)) fail(AUGUR_ERR_ARG, "info needs 7 doubles");
    info[0] = b->bs_phase;
    info[1] = b->bs_active;
    info[2] = b->bs_fallbacks;
    info[3] = b->bs_fit_ms;
    info[4] = (double)b->bs_collected * b->base.steps * b->C;
    info[5] = (double)b->bs_sweeps;
    info[6] = b->bs_collected;
  });
}

augur_status augur_bhlr_surrogate_modes(augur_sampler* s, int32_t* modes, int n) {
  return guard([&] {
    auto* b = dynamic_cast<BhlrSampler*>(s);
    if (!b || !modes || n < b->n_ds) fail(AUGUR_ERR_ARG, "needs a batched HLR sampler and n >= n_datasets");
    for (int i = 0; i < b->n_ds; ++i) modes[i] = b->bs_phase == BhlrSampler::BS_ACTIVE ? b->hsmode[i] : 0;
  });
}

augur_status augur_bhlr_surrogate_set_mode(augur_sampler* s, int dataset, int mode) {
  return guard([&] {
    auto* b = dynamic_cast<BhlrSampler*>(s);
    if (!b) fail(AUGUR_ERR_ARG, "not a batched HLR sampler");
    if (b->bs_phase != BhlrSampler::BS_ACTIVE) fail(AUGUR_ERR_STATE, "no fitted surrogate");
    if (dataset < 0 || dataset >= b->n_ds) fail(AUGUR_ERR_ARG, "dataset %d out of range", dataset);
    if (mode != 0 && mode != 1) fail(AUGUR_ERR_ARG, "mode must be 0 (exact HMC) or 1 (surrogate)");
    if (b->hsmode[dataset] == mode) return;
    AUGUR_HIP(hipSetDevice(b->device));
    b->hsmode[dataset] = mode;
    b->bs_modes_changed();
  });
}

augur_status augur_bhlr_surrogate_map(augur_sampler* s, int dataset, double* w, double* mean_theta,
                                      double* mean_grad) {
  return guard([&] {
    auto* b = dynamic_cast<BhlrSampler*>(s);
    if (!b) fail(AUGUR_ERR_ARG, "not a batched HLR sampler");
    if (b->bs_phase != BhlrSampler::BS_ACTIVE) fail(AUGUR_ERR_STATE, "no fitted surrogate");
    if (dataset < 0 || dataset >= b->n_ds) fail(AUGUR_ERR_ARG, "dataset %d out of range", dataset);
    const BDs& d = b->hds[dataset];
    AUGUR_HIP(hipSetDevice(b->device));
    AUGUR_HIP(hipStreamSynchronize(b->stream));
    if (w) {
      std::vector<float> h((size_t)d.d * d.d);
      AUGUR_HIP(hipMemcpy(h.data(), b->surr_w.p + d.woff, h.size() * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t e = 0; e < h.size(); ++e) w[e] = h[e];
    }
    if (mean_theta)
      AUGUR_HIP(hipMemcpy(mean_theta, b->surr_m.p + d.goff, d.d * sizeof(double), hipMemcpyDeviceToHost));
    if (mean_grad)
      AUGUR_HIP(hipMemcpy(mean_grad, b->surr_g.p + d.goff, d.d * sizeof(double), hipMemcpyDeviceToHost));
  });
}

augur_status augur_bhlr_surrogate_pairs(augur_sampler* s, int dataset, float* theta, float* grad, int cap,
                                        int* n_pairs) {
  return guard([&] {
    auto* b = dynamic_cast<BhlrSampler*>(s);
    if (!b || !n_pairs) fail(AUGUR_ERR_ARG, "needs a batched HLR sampler and n_pairs");
    if (b->bs_phase == BhlrSampler::BS_OFF) fail(AUGUR_ERR_STATE, "no surrogate configured");
    if (dataset < 0 || dataset >= b->n_ds) fail(AUGUR_ERR_ARG, "dataset %d out of range", dataset);
    const BDs& d = b->hds[dataset];
    const int np = b->bs_collected * b->base.steps * b->C;
    *n_pairs = np;
    if (!theta && !grad) return;
    if (cap < np) fail(AUGUR_ERR_ARG, "room for %d pairs, %d collected", cap, np);
    AUGUR_HIP(hipSetDevice(b->device));
    AUGUR_HIP(hipStreamSynchronize(b->stream));
    // pair q = slot * C + chain: row q of the output, the dataset's d values
    const size_t sp = b->GW * sizeof(float), dp = (size_t)d.d * sizeof(float);
    if (theta) AUGUR_HIP(hipMemcpy2D(theta, dp, b->cap_t.p + d.goff, sp, dp, np, hipMemcpyDeviceToHost));
    if (grad) AUGUR_HIP(hipMemcpy2D(grad, dp, b->cap_g.p + d.goff, sp, dp, np, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
