// augur_hlr.hip -- HMC sampler for hierarchical logistic regression (hlr.rv
// family) on gfx950.
//
//   v ~ InvGamma(a, b) | Exponential(a);  [b ~ Normal(0, v)];
//   theta[k] ~ Normal(0, v);  y[n] ~ Bernoulli(sigmoid(dotprod(theta, x[n]) [+ b]))
//
// Reference path: the generated grad_<blk> / ll_<blk> map-reduces
// (Core/CgGrad.hs:306-376, Core/CgLike.hs:90-136, lowered by
// CudaC/CgCuda.hs:371-516 to one-thread-per-row kernels with FP64 atomics and
// thrust reductions) driven by the host HMC loop of cbits/augur_mcmclib.h:
// 343-424.  Here the whole HMC iteration stays on the device:
//
//   k_hmc_begin   momentum draw, first half step, first position step
//                 (k_hmc_vblock_begin: after the v block's update in the
//                 same launch when a v block precedes the data block)
//   k_hlr_pass    one data pass per leapfrog step for all chains at once
//                 (k_hlr_pass_r: the register form, X rows in registers and
//                 x.theta completed by DPP; k_hlr_pass_t / k_hlr_pass: the
//                 LDS-staged forms): X tile staged in LDS and shared by every chain,
//                 t = x.theta + b, a = y - sigmoid(t), ll += log sigmoid(+-t),
//                 grad += a x; FP32 arithmetic, FP64 accumulation across
//                 tiles; each block writes one row of partials
//   k_hlr_finish  sums the partials in a fixed order (deterministic) and
//                 applies the leapfrog update of every chain (momentum,
//                 position, FP32 copy of the next evaluation point)
//   k_hmc_end     kinetic energies, FP64 Metropolis test, accept / reject.
//
// The data log-density and gradient at the current point are cached (they
// are exactly what the pass computed when that point was proposed), so an
// iteration of L leapfrog steps costs L data passes instead of the
// reference's L + 3 (L + 1 gradients and 2 log-densities).
#include <climits>
#include <chrono>
#include <mutex>

#include "augur_hmc.h"
#include "augur_mlp.h"
#include "augur_surrogate.h"

namespace augur {

constexpr int kTileRows = 64;
constexpr int kPassThreads = 256;
constexpr int kMaxGradSlots = 16;  // (d) * cp <= 16 * 256
constexpr int kPassBlocksPerCU = 3;  // register budget of the specialised pass (<= 168 VGPRs)

enum { MODE_CACHE = 0, MODE_LEAPFROG = 1 };

struct PassArgs {
  const float* __restrict__ x;
  const int32_t* __restrict__ y;
  int n, d, cp, c_real, has_b;
  float* beta32;   // cp x (d + 1): evaluation point (theta, b) per chain
  double* part;    // [G][cp * (d + 2)]
  int G;           // blocks of the pass kernel
  int mode;
  int l;           // leapfrog pass index (1-based)
  double* cache_g;  // cp x (d + 1)
  double* cache_ll; // cp
  double* gprop;    // cp x (d + 1)
  double* llprop;   // cp
  double* prop;     // cp x P
  double* p;        // cp x (d + 1)
  const double* state;  // cp x P
  int blk_theta, blk_b;
  HmcStep hs;
  // training-pair capture (null: off): row cap_row + c receives chain c's
  // evaluation point (FP32) and data gradient
  float* cap_x;
  float* cap_y;
  int cap_row;
};


// Reduction of the block partials and the leapfrog epilogue for values
// (chain c, coordinate j; j = d: intercept, j = d + 1: log-density); the
// partials are part[block][value].  A 1024-thread block takes kFinVals
// consecutive values: thread (group g, value i) sums rows g, g + kFinGroups,
// ... of its column (each group reads contiguous row segments); the groups
// are then summed by a butterfly within each wave and the waves in order:
// the result is deterministic.  The update of coordinate j needs only value
// j, so every value finishes independently.  (A separate kernel instead of
// a last-block epilogue: the device-scope fence each pass block would need
// writes back the L2 and serialises across blocks.)
constexpr int kFinVals = 8;
constexpr int kFinGroups = 1024 / kFinVals;
constexpr int kFinLoads = 8;  // loads in flight per thread: one round up to G = 1024

__global__ __launch_bounds__(1024) void k_hlr_finish(PassArgs A) {
  __shared__ double red[1024 / 64][kFinVals];
  const int d = A.d, ds = d + 1, vs = d + 2, P = d + 2;
  const int nval = A.cp * vs;
  const int vi = threadIdx.x % kFinVals, grp = threadIdx.x / kFinVals;
  const int v0 = blockIdx.x * kFinVals;
  const int v = v0 + vi;
  const int c = v / vs, j = v % vs;
  // the update's operands are requested with the partials: one memory round
  const bool upd = grp == 0 && v < nval && A.mode != MODE_CACHE && c < A.c_real && j <= d;
  double var = 1.0, xv = 0.0, p_old = 0.0;
  if (upd) {
    var = A.state[c * P + d + 1];
    xv = A.prop[c * P + j];
    p_old = A.p[c * ds + j];
  }
  double acc = 0.0;
  if (v < nval) {
    // kFinLoads independent loads per round, summed in a fixed order
    const double* col = A.part + v;
    for (int g0 = grp; g0 < A.G; g0 += kFinLoads * kFinGroups) {
      double t[kFinLoads];
#pragma unroll
      for (int k = 0; k < kFinLoads; ++k) {
        const int g = g0 + k * kFinGroups;
        t[k] = g < A.G ? col[(size_t)g * nval] : 0.0;
      }
#pragma unroll
      for (int w = kFinLoads / 2; w >= 1; w >>= 1)
#pragma unroll
        for (int k = 0; k < w; ++k) t[k] += t[k + w];
      acc += t[0];
    }
  }
  // lanes l and l ^ (kFinVals 2^k) hold the same value
#pragma unroll
  for (int off = 32; off >= kFinVals; off >>= 1) acc += __shfl_xor(acc, off);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane < kFinVals) red[wv][lane] = acc;
  __syncthreads();
  if (threadIdx.x >= kFinVals) return;
  if (v >= nval) return;
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < 1024 / 64; ++w) s += red[w][threadIdx.x];
  if (A.mode == MODE_CACHE) {
    if (j <= d) A.cache_g[c * ds + j] = s;
    else A.cache_ll[c] = s;
    return;
  }
  if (c >= A.c_real) return;
  const int L = A.hs.steps[c];
  if (A.l > L) return;
  if (j == d + 1) {
    if (A.l == L) A.llprop[c] = s;
    return;
  }
  const int q = c * ds + j;
  if (A.l == L) A.gprop[q] = s;
  if (A.cap_x && j < d) {
    const size_t row = (size_t)(A.cap_row + c) * d + j;
    A.cap_x[row] = A.beta32[q];  // the point this pass evaluated
    A.cap_y[row] = (float)s;
  }
  const bool in_blk = (j < d) ? A.blk_theta : (A.blk_b && A.has_b);
  if (!in_blk) return;
  const double step = A.hs.step[c];
  const double g = s + (-xv / var);
  const double coef = (A.l == L) ? 0.5 * step : step;
  const double pj = p_old + coef * g;
  A.p[q] = pj;
  if (A.l < L) {
    xv += step * pj;
    A.prop[c * P + j] = xv;
    A.beta32[q] = (float)xv;
  }
}

__global__ __launch_bounds__(kPassThreads) void k_hlr_pass(PassArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int d = A.d, cp = A.cp, ds = d + 1;
  float* xs = (float*)smem;                          // kTileRows x ds
  float* bs = xs + kTileRows * ds;                   // cp x ds
  float* as = bs + cp * ds;                          // kTileRows x cp
  int* ys = (int*)(as + kTileRows * cp);             // kTileRows
  double* red = (double*)(((uintptr_t)(ys + kTileRows) + 15) & ~(uintptr_t)15);  // 2 x 256
  int* flag = (int*)(red + 2 * kPassThreads);

  const int tid = threadIdx.x;
  const int G = gridDim.x;
  for (int i = tid; i < cp * ds; i += kPassThreads) bs[i] = A.beta32[i];

  // phase-B ownership: combo q = tid + 256 m -> row q / cp, chain q % cp (cp | 256)
  const int my_c = tid % cp;
  const int nb = (kTileRows * cp + kPassThreads - 1) / kPassThreads;
  // phase-C ownership: combo q = tid + 256 m -> chain q / d, column q % d
  const int ngc = d * cp;
  const int ng = (ngc + kPassThreads - 1) / kPassThreads;
  double gacc[kMaxGradSlots];
#pragma unroll
  for (int m = 0; m < kMaxGradSlots; ++m) gacc[m] = 0.0;
  double llacc = 0.0, gbacc = 0.0;

  const int tiles = (A.n + kTileRows - 1) / kTileRows;
  for (int tile = blockIdx.x; tile < tiles; tile += G) {
    const int r0 = tile * kTileRows;
    const int rows = min(kTileRows, A.n - r0);
    __syncthreads();
    // stage the tile: rows x d contiguous floats
    const float* src = A.x + (size_t)r0 * d;
    if ((d & 3) == 0) {
      const int nv = rows * d / 4;
      const float4* s4 = (const float4*)src;
      for (int i = tid; i < nv; i += kPassThreads) {
        const float4 v = s4[i];
        const int e = 4 * i, r = e / d, c = e % d;
        float* dst = xs + r * ds + c;
        dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; dst[3] = v.w;
      }
    } else {
      for (int i = tid; i < rows * d; i += kPassThreads) xs[(i / d) * ds + i % d] = src[i];
    }
    for (int i = rows * d + tid; i < kTileRows * d; i += kPassThreads) xs[(i / d) * ds + i % d] = 0.0f;
    if (tid < kTileRows) ys[tid] = tid < rows ? A.y[r0 + tid] : 0;
    __syncthreads();
    // phase B: t = x . theta_c + b_c per (row, chain)
    for (int m = 0; m < nb; ++m) {
      const int q = tid + kPassThreads * m;
      if (q >= kTileRows * cp) break;
      const int r = q / cp, c = q % cp;
      const float* xr = xs + r * ds;
      const float* br = bs + c * ds;
      float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
      int j = 0;
      for (; j + 4 <= d; j += 4) {
        s0 = fmaf(xr[j], br[j], s0);
        s1 = fmaf(xr[j + 1], br[j + 1], s1);
        s2 = fmaf(xr[j + 2], br[j + 2], s2);
        s3 = fmaf(xr[j + 3], br[j + 3], s3);
      }
      for (; j < d; ++j) s0 = fmaf(xr[j], br[j], s0);
      float t = (s0 + s1) + (s2 + s3);
      if (A.has_b) t += br[d];
      float a;
      const float lr = log_sigmoid_pm(t, ys[r], &a);
      if (r < rows) {
        llacc += (double)lr;
        gbacc += (double)a;
      } else {
        a = 0.0f;
      }
      as[r * cp + c] = a;
    }
    __syncthreads();
    // phase C: grad[c][j] += sum_r a[r][c] x[r][j]
#pragma unroll
    for (int m = 0; m < kMaxGradSlots; ++m) {
      if (m >= ng) break;
      const int q = tid + kPassThreads * m;
      if (q < ngc) {
        const int c = q / d, j = q % d;
        float f0 = 0.f, f1 = 0.f;
        for (int r = 0; r < kTileRows; r += 2) {
          f0 = fmaf(as[r * cp + c], xs[r * ds + j], f0);
          f1 = fmaf(as[(r + 1) * cp + c], xs[(r + 1) * ds + j], f1);
        }
        gacc[m] += (double)(f0 + f1);
      }
    }
  }
  // block partials -> part[block][value], value = c * (d + 2) + j
  const int vs = d + 2;
  double* pb = A.part + (size_t)blockIdx.x * cp * vs;
#pragma unroll
  for (int m = 0; m < kMaxGradSlots; ++m) {
    if (m >= ng) break;
    const int q = tid + kPassThreads * m;
    if (q < ngc) {
      const int c = q / d, j = q % d;
      pb[c * vs + j] = gacc[m];
    }
  }
  __syncthreads();
  red[tid] = gbacc;
  red[kPassThreads + tid] = llacc;
  __syncthreads();
  if (tid < cp) {
    double sb = 0.0, sl = 0.0;
    for (int i = tid; i < kPassThreads; i += cp) {
      sb += red[i];
      sl += red[kPassThreads + i];
    }
    pb[tid * vs + d] = sb;
    pb[tid * vs + d + 1] = sl;
  }
  (void)my_c;
  (void)G;
  (void)flag;
}

// Specialised data pass for D % 4 == 0 and a power-of-two chain count CP.
//   - rows are staged with 16-byte-aligned stride D + 4, so tile stores and
//     row reads are 128-bit LDS operations;
//   - the next tile is prefetched into registers while the current one is
//     processed;
//   - phase B keeps each thread's theta_c in registers (D <= 64);
//   - phase C gives each thread 4 adjacent columns of one chain and a slice of
//     the tile's rows (a 128-bit x read and one broadcast a read feed 4 FMAs);
//     the row slices are combined once per block, in a fixed order.
template <int D, int CP>
__global__ __launch_bounds__(kPassThreads, kPassBlocksPerCU) void k_hlr_pass_t(PassArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int DS = D + 4;
  constexpr int TR = kTileRows;
  constexpr int NV = TR * D / 4;                              // float4 per tile
  constexpr int NVT = (NV + kPassThreads - 1) / kPassThreads;  // per thread
  constexpr int NB = (TR * CP + kPassThreads - 1) / kPassThreads;
  constexpr int NP = CP * (D / 4);                             // (chain, column quad) pairs
  constexpr int RS = NP >= kPassThreads ? 1 : kPassThreads / NP;  // row slices
  constexpr int NS = NP > kPassThreads ? NP / kPassThreads : 1;  // pairs per thread
  constexpr bool BREG = D <= 64;
  constexpr int PD = D <= 64 ? 3 : 2;                          // tiles in flight (pass_ring)
  static_assert(D % 4 == 0 && (CP & (CP - 1)) == 0 && CP <= kMaxChains, "shape");
  static_assert(TR % RS == 0, "row slices");

  float* xs = (float*)smem;                 // TR x DS
  float* bs = xs + TR * DS;                 // CP x (D + 4)
  float* as = bs + CP * DS;                 // TR x CP
  int* ys = (int*)(as + TR * CP);           // TR
  double* red = (double*)(((uintptr_t)(ys + TR) + 15) & ~(uintptr_t)15);  // 2 x 256
  double* red2 = red + 2 * kPassThreads;    // RS x NP x 4 (RS * NP <= 256)
  int* flag = (int*)(red2 + 4 * kPassThreads);

  const int tid = threadIdx.x;
  const int G = gridDim.x;
  const int n = A.n;
  const int tiles = (n + TR - 1) / TR;

  // a ring of PD tiles in flight: this block's tiles blockIdx.x + k G are
  // loaded PD ahead of their use (all of them up front when the block has
  // at most PD), so the pass waits on memory about once, not once per tile.
  // The first tiles are requested before theta, so the two reads overlap.
  float4 pre[PD][NVT];
  int ypre[PD];
  auto prefetch = [&](int slot_tile, float4 (&dst)[NVT], int& ydst) {
    const int r0 = slot_tile * TR;
    const int rows = min(TR, n - r0);
    const float4* s4 = (const float4*)(A.x + (size_t)r0 * D);
#pragma unroll
    for (int k = 0; k < NVT; ++k) {
      const int i = tid + k * kPassThreads;
      dst[k] = (i < NV && i < rows * (D / 4)) ? s4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    ydst = (tid < TR && tid < rows) ? A.y[r0 + tid] : 0;
  };
#pragma unroll
  for (int k = 0; k < PD; ++k) {
    const int t = blockIdx.x + k * G;
    if (t < tiles) prefetch(t, pre[k], ypre[k]);
  }

  for (int i = tid; i < CP * (D + 1); i += kPassThreads) bs[(i / (D + 1)) * DS + i % (D + 1)] = A.beta32[i];
  __syncthreads();

  const int my_c = tid % CP;  // phase-B chain of this thread
  float breg[BREG ? D : 1];
  if constexpr (BREG) {
#pragma unroll
    for (int j = 0; j < D; ++j) breg[j] = bs[my_c * DS + j];
  }
  const float b_c = A.has_b ? bs[my_c * DS + D] : 0.0f;

  double gacc[NS][4];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int k = 0; k < 4; ++k) gacc[s][k] = 0.0;
  double llacc = 0.0, gbacc = 0.0;

  for (int base = blockIdx.x; base < tiles; base += PD * G) {
#pragma unroll
  for (int slot = 0; slot < PD; ++slot) {
    const int tile = base + slot * G;
    if (tile >= tiles) break;
    const int rows = min(TR, n - tile * TR);
#pragma unroll
    for (int k = 0; k < NVT; ++k) {
      const int i = tid + k * kPassThreads;
      if (i < NV) {
        const int e = 4 * i;
        *(float4*)(xs + (e / D) * DS + e % D) = pre[slot][k];
      }
    }
    if (tid < TR) ys[tid] = ypre[slot];
    __syncthreads();
    if (tile + PD * G < tiles) prefetch(tile + PD * G, pre[slot], ypre[slot]);
    // phase B: t = x . theta_c + b_c
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int q = tid + kPassThreads * m;
      if (q < TR * CP) {
        const int r = q / CP;
        const float* xr = xs + r * DS;
        const float* br = bs + my_c * DS;
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
        for (int j = 0; j < D; j += 4) {
          const float4 xv = *(const float4*)(xr + j);
          float4 bv;
          if constexpr (BREG) bv = make_float4(breg[j], breg[j + 1], breg[j + 2], breg[j + 3]);
          else bv = *(const float4*)(br + j);
          s0 = fmaf(xv.x, bv.x, s0);
          s1 = fmaf(xv.y, bv.y, s1);
          s2 = fmaf(xv.z, bv.z, s2);
          s3 = fmaf(xv.w, bv.w, s3);
        }
        const float t = ((s0 + s1) + (s2 + s3)) + b_c;
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
    // phase C: grad[c][4 jq .. 4 jq + 3] += sum over the thread's rows
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int pair = (NS > 1) ? tid + s * kPassThreads : tid % NP;
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
  }
  // block partials -> part[block][value], value = c * (D + 2) + j: one
  // contiguous row per block
  constexpr int VS = D + 2;
  double* pb = A.part + (size_t)blockIdx.x * CP * VS;
  if constexpr (RS > 1) {
#pragma unroll
    for (int k = 0; k < 4; ++k) red2[((tid / NP) * NP + tid % NP) * 4 + k] = gacc[0][k];
    __syncthreads();
    if (tid < NP) {
      const int c = tid / (D / 4), jq = tid % (D / 4);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        double s = 0.0;
        for (int rs = 0; rs < RS; ++rs) s += red2[(rs * NP + tid) * 4 + k];
        pb[c * VS + 4 * jq + k] = s;
      }
    }
  } else {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int pair = tid + s * kPassThreads;
      const int c = pair / (D / 4), jq = pair % (D / 4);
#pragma unroll
      for (int k = 0; k < 4; ++k) pb[c * VS + 4 * jq + k] = gacc[s][k];
    }
  }
  // intercept gradient and ll per chain: lanes l = c (mod CP) of each wave
  // summed by a butterfly, then the four waves in order
#pragma unroll
  for (int off = 32; off >= CP; off >>= 1) {
    gbacc += __shfl_xor(gbacc, off);
    llacc += __shfl_xor(llacc, off);
  }
  const int lane = tid & 63, wv = tid >> 6;
  if (lane < CP) {
    red[wv * CP + lane] = gbacc;
    red[kPassThreads + wv * CP + lane] = llacc;
  }
  __syncthreads();
  if (tid < CP) {
    constexpr int NW = kPassThreads / 64;
    double sb = 0.0, sl = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      sb += red[w * CP + tid];
      sl += red[kPassThreads + w * CP + tid];
    }
    pb[tid * VS + D] = sb;
    pb[tid * VS + D + 1] = sl;
  }
  (void)flag;
}

// Register-resident form of the data pass for D = 16, 32, 64 and up to
// D / 8 chains: no LDS staging of X and no block barrier in the row loop.
// Lane l owns 8 consecutive features (8 (l % LPR) ..) of RPT rows per tile,
// LPR = D / 8 lanes per row, so one tile is a contiguous 16 KB of X read as
// 32 bytes per lane.
//  - phase B: each lane forms its 8-feature share of x . theta_c for its rows
//    and chains; a butterfly over the row's lanes (DPP: quad permutations and
//    the half-row mirror) completes every dot product in every lane;
//  - the RPT * CP (row, chain) adjoints of a row group are computed once
//    each: lane q takes pairs q, q + LPR, ... (its chain is q % CP) and the
//    row's lanes gather them back with ds_bpermute;
//  - phase C: a * x accumulates into the lane's CP x 8 gradient share in FP32
//    registers; the block sums the lanes in FP64 at the end (LDS, fixed order).
// theta is requested before the first tiles, so waiting for it never waits on
// X, and each wave starts on its first tile while the later ones arrive.
constexpr int kRowFeat = 8;  // features per lane in the register form
typedef float f32x4_t __attribute__((ext_vector_type(4)));
// rows per lane per tile and tiles in flight: 32 KB of X per block either way
__host__ __device__ constexpr int row_rpt(int cp) { return 2; }
__host__ __device__ constexpr int row_ring(int cp) { return 2; }

// s_waitcnt immediate that waits until at most n vector-memory loads are
// outstanding (gfx9 encoding: vmcnt bits 3:0 and 15:14; expcnt and lgkmcnt
// left at their maxima, i.e. not waited on)
__host__ __device__ constexpr int vmcnt_imm(int n) {
  return (n & 15) | (7 << 4) | (15 << 8) | (((n >> 4) & 3) << 14);
}

// Chain whose row term lane q (of a 16-lane row group) evaluates in the
// matrix-core form of the D = 128, 8-chain pass, and its inverse.  With q's
// bits b0..b3, rs_chain(q) = b1 ^ 3 b2 ^ 6 b3 (XOR): the lanes q ^ 15, q ^ 7,
// q ^ 3 and q ^ 1 differ from lane q by 4, 2, 1 and 0 in it, which is what
// the reduce-scatter's four levels need; lanes 2 h and 2 h + 1 share a chain
// (slots 0 and 1), and h -> rs_chain(2 h) is a bijection of 0..7.
__device__ inline int rs_chain(int q) {
  return (((q >> 1) & 1) * 1) ^ (((q >> 2) & 1) * 3) ^ (((q >> 3) & 1) * 6);
}
__device__ inline int rs_lane_of_chain(int c) {
  // h with rs_chain(2 h) = c: g(h) for h = 0..7 is 0 1 3 2 6 7 5 4
  constexpr unsigned kInv = 0u | 1u << 3 | 3u << 6 | 2u << 9 | 7u << 12 | 6u << 15 | 4u << 18 | 5u << 21;
  return (kInv >> (3 * c)) & 7;
}

__device__ inline float f4c(const float4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }

// X and y are stored with zero rows up to a multiple of kHlrRowPad, so a tile
// of the register-form pass that starts below n lies wholly in the buffers.
constexpr int kHlrRowPad = 256;

// Blocks per CU the register form is compiled for: at D * CP > 256 its
// theta and gradient shares (2 CP F floats) and tile ring need up to 256
// registers a lane, so two blocks (two waves per SIMD) instead of three.
//
// MB (D = 128, 8 chains; 108 -> 91 us a pass at N = 1M with the changes
// below, measured against the vector form on one box):
//  - phase C on the matrix cores.  A wave's four row groups are the K = 4
//    rows of V_MFMA_F32_16X16X4_F32: lane l supplies A[m = l % 16][k = l /
//    16] = x[row l / 16][8 m + j] (the value it already holds, product j
//    covering features 8 m + j) and B[k][n = l % 16] = the adjoint of (row
//    l / 16, chain n), zero for n >= 8.  The 16 products of a tile (2 row
//    slots x 8 feature offsets j) accumulate in FP32 as phase C does; lane l
//    ends with G[8 (4 (l / 16) + r) + j][chain l % 16] in acc[j][r].  Phase
//    C's 144 vector operations per lane and tile become 16 matrix-core
//    instructions, which run beside the vector work of phase B;
//  - phase B ends in a reduce-scatter instead of 16 all-reduces: each lane
//    keeps the chains in a lane-dependent order (rs_chain), so that four
//    DPP levels leave it the one (row, chain) sum whose row term it
//    evaluates (16 DPP adds per tile instead of 64, and no selects).
// LL = false: the gradient only, for the leapfrog passes before the shortest
// trajectory's last (the log-density is read by the cache pass and by each
// chain's last pass)
template <int D, int CP, bool MB, bool LL = true>
__global__ __launch_bounds__(kPassThreads, (D * CP > 256 ? 2 : kPassBlocksPerCU)) void k_hlr_pass_r(PassArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int F = kRowFeat, RPT = row_rpt(CP);
  constexpr int LPR = D / F;                      // lanes per row
  constexpr int RG = kPassThreads / LPR;          // rows per block per row slot
  constexpr int TRr = RG * RPT;                   // rows per tile
  constexpr int NC = RPT * CP;                    // (row slot, chain) pairs
  constexpr int SPL = (NC + LPR - 1) / LPR;       // pairs per lane
  constexpr int PD = row_ring(CP);                // tiles in flight
  constexpr int GS = CP * F + 1;                  // LDS floats per lane (padded)
  static_assert(D % F == 0 && LPR >= 2 && LPR <= 16 && LPR % CP == 0, "shape");
  static_assert(kHlrRowPad % TRr == 0, "tiles end inside the row padding");
  static_assert(!MB || (LPR == 16 && CP == 8 && RPT == 2 && F == 8), "matrix-core phase C shape");

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int q = tid % LPR;          // feature group: features F q .. F q + F - 1
  const int rsub = tid / LPR;       // row within a row slot
  const int gbase = lane - q;       // first lane of this row's group
  const int my_c = q % CP;          // chain of every pair this lane evaluates
  const int G = gridDim.x, n = A.n;
  const int tiles = (n + TRr - 1) / TRr;

  // MB: lane q's registers hold the chains in the order c ^ gq (see the
  // reduce-scatter below); it evaluates the row term of (slot q & 1, chain gq)
  const int gq = MB ? rs_chain(q) : 0;
  // theta share of this lane (features F q .., every chain) and b of its chain
  float th[CP][F];
#pragma unroll
  for (int c = 0; c < CP; ++c)
#pragma unroll
    for (int j = 0; j < F; ++j) th[c][j] = A.beta32[(MB ? c ^ gq : c) * (D + 1) + F * q + j];
  const float b_c = A.has_b ? A.beta32[(MB ? gq : my_c) * (D + 1) + D] : 0.0f;
  // MB: lanes whose pairs are (slot 0 / 1, chain q) for the B operands
  const int src0 = gbase + 2 * rs_lane_of_chain(q & 7);

  // PD tiles in flight in NB = PD + 1 register buffers: the refill of a slot
  // goes to the buffer the previous slot consumed, so the loads can be issued
  // before this slot's work without copying registers around the loop
  constexpr int NB = PD + 1;
  float4 pre[NB][RPT][2];
  int ypre[NB][RPT];
  auto prefetch = [&](int tile, float4 (&dst)[RPT][2], int (&yd)[RPT]) {
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      // rows of a tile lie below the row count padded to kHlrRowPad (zero
      // rows): no bounds check; rows at or past n are masked in the row term
      const int row = tile * TRr + i * RG + rsub;
      const float4* s4 = (const float4*)(A.x + (size_t)row * D + F * q);
      dst[i][0] = s4[0];
      dst[i][1] = s4[1];
      yd[i] = A.y[row];
    }
  };
  // Every prefetch is issued, a tile past the last one clamped to it (its
  // rows are valid memory, loaded again and not used), so each path
  // through the loop issues the same loads and the compiler's waits can
  // count the ring exactly: theta and b, requested before the ring's PD
  // tiles (3 loads a row), are complete once at most PD * RPT * 3 loads are
  // outstanding.  With conditional prefetches the waits in the row loop
  // fell back to zero outstanding loads once per pass over the ring.
  const int last = tiles - 1;  // n >= 1, so tiles >= 1
#pragma unroll
  for (int k = 0; k < PD; ++k) prefetch(min((int)blockIdx.x + k * G, last), pre[k], ypre[k]);
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(PD * RPT * 3));

  // vector form: lane gradient shares g[c][j]; matrix-core form: acc[j]
  float g[MB ? 1 : CP][MB ? 1 : F];
  f32x4_t acc[MB ? F : 1];
#pragma unroll
  for (int c = 0; c < (MB ? 1 : CP); ++c)
#pragma unroll
    for (int j = 0; j < (MB ? 1 : F); ++j) g[c][j] = 0.0f;
#pragma unroll
  for (int j = 0; j < (MB ? F : 1); ++j) acc[j] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  double llacc = 0.0, gbacc = 0.0;

  for (int base = blockIdx.x; base < tiles; base += NB * G) {
#pragma unroll
    for (int slot = 0; slot < NB; ++slot) {
      const int tile = base + slot * G;
      // issued even when this slot has no tile, so that every path through
      // the loop issues the same loads
      prefetch(min(tile + PD * G, last), pre[(slot + PD) % NB], ypre[(slot + PD) % NB]);
      if (tile >= tiles) continue;
      auto xr = [&](int i, int j) { return f4c(pre[slot][i][j >> 2], j & 3); };
      const int* yv = ypre[slot];
      // phase B: shares of t[i][c] = x_i . theta_c, completed over the row's lanes
      float tp[NC];
#pragma unroll
      for (int i = 0; i < RPT; ++i)
#pragma unroll
        for (int c = 0; c < CP; ++c) {
          float s0 = xr(i, 0) * th[c][0], s1 = xr(i, 1) * th[c][1];
          s0 = fmaf(xr(i, 2), th[c][2], s0);
          s1 = fmaf(xr(i, 3), th[c][3], s1);
          s0 = fmaf(xr(i, 4), th[c][4], s0);
          s1 = fmaf(xr(i, 5), th[c][5], s1);
          s0 = fmaf(xr(i, 6), th[c][6], s0);
          s1 = fmaf(xr(i, 7), th[c][7], s1);
          tp[i * CP + c] = MB ? s0 + s1 : row_sum<LPR>(s0 + s1);
        }
      float aown[SPL];
      if constexpr (MB) {
        // Reduce-scatter of the 16 (slot, chain) shares over the row's 16
        // lanes: register 8 s + c of lane q holds pair (s, c ^ gq), and at
        // each level the partner lane holds the pair of register k in k ^ 4
        // (row mirror, q ^ 15), k ^ 2 (half-row mirror, q ^ 7), k ^ 1
        // (q ^ 3), then the same pairs (q ^ 1): 16 DPP adds instead of 64,
        // leaving t of (slot s, chain gq) in register 8 s.
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
          for (int c = 0; c < 4; ++c) tp[8 * i + c] += dpp_f<0x140>(tp[8 * i + c + 4]);
#pragma unroll
          for (int c = 0; c < 2; ++c) tp[8 * i + c] += dpp_f<0x141>(tp[8 * i + c + 2]);
          tp[8 * i] += dpp_f<0x1B>(tp[8 * i + 1]);
          tp[8 * i] += dpp_f<0xB1>(tp[8 * i]);
        }
        // bit blends rather than selects: the compiler turns a select of two
        // array elements into a dynamically indexed load, which sends the
        // whole array to scratch memory
        const int i = q & 1, im = -i;
        const float tv = __int_as_float(__float_as_int(tp[0]) ^ ((__float_as_int(tp[0]) ^ __float_as_int(tp[8])) & im));
        const int yy = yv[0] ^ ((yv[0] ^ yv[1]) & im);
        const bool ok = tile * TRr + i * RG + rsub < n;
        float a;
        if constexpr (LL) {
          const float lr = log_sigmoid_pm(tv + b_c, yy, &a);
          if (ok) llacc += (double)lr;
        } else {
          a = sigmoid_adjoint_pm(tv + b_c, yy);
        }
        a = ok ? a : 0.0f;
        gbacc += (double)a;
        aown[0] = a;
      }
      // adjoints: pair m = q + s LPR is row slot m / CP, chain m % CP = my_c
#pragma unroll
      for (int s = 0; s < (MB ? 0 : SPL); ++s) {
        const int m = q + s * LPR;
        float tv = tp[0];
#pragma unroll
        for (int k = 1; k < NC; ++k) tv = (m == k) ? tp[k] : tv;
        const int i = m / CP;
        int yy = yv[0];
#pragma unroll
        for (int k = 1; k < RPT; ++k) yy = (i == k) ? yv[k] : yy;
        const bool ok = m < NC && tile * TRr + i * RG + rsub < n;
        float a;
        if constexpr (LL) {
          const float lr = log_sigmoid_pm(tv + b_c, yy, &a);
          if (ok) llacc += (double)lr;
        } else {
          a = sigmoid_adjoint_pm(tv + b_c, yy);
        }
        a = ok ? a : 0.0f;
        gbacc += (double)a;
        aown[s] = a;
      }
      if constexpr (MB) {
        // B operands: the adjoints of (slot 0, chain q) and (slot 1, chain q)
        const float a0 = __shfl(aown[0], src0), a1 = __shfl(aown[0], src0 + 1);
        const float b0 = q < CP ? a0 : 0.0f, b1 = q < CP ? a1 : 0.0f;
#pragma unroll
        for (int j = 0; j < F; ++j) {
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr(0, j), b0, acc[j], 0, 0, 0);
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr(1, j), b1, acc[j], 0, 0, 0);
        }
      } else {
        // phase C: g[c] += a[i][c] x_i over the lane's features
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          const float av = __shfl(aown[m / LPR], gbase + m % LPR);
          const int i = m / CP, c = m % CP;
#pragma unroll
          for (int j = 0; j < F; ++j) g[c][j] = fmaf(av, xr(i, j), g[c][j]);
        }
      }
    }
  }

  // block partials -> part[block][value], value = c (D + 2) + j
  constexpr int VS = D + 2;
  float* gl = (float*)smem;                                      // [256][GS]
  double* red = (double*)(smem + ((sizeof(float) * kPassThreads * GS + 15) & ~(size_t)15));  // [2][4][CP]
  if constexpr (MB) {
    // the wave's sums: gl[(wave * CP + c) * D + feature]
    if (q < CP) {
#pragma unroll
      for (int j = 0; j < F; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) gl[(wv * CP + q) * D + F * (4 * (lane >> 4) + r) + j] = acc[j][r];
    }
  } else {
#pragma unroll
    for (int c = 0; c < CP; ++c)
#pragma unroll
      for (int j = 0; j < F; ++j) gl[tid * GS + c * F + j] = g[c][j];
  }
  // ll and intercept gradient: lanes with the same chain (lane = c mod CP;
  // MB: lanes 2 h and 2 h + 1 of every row group hold chain rs_chain(2 h))
  if constexpr (MB) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int off = k == 0 ? 32 : k == 1 ? 16 : 1;
      gbacc += __shfl_xor(gbacc, off);
      llacc += __shfl_xor(llacc, off);
    }
    if (lane < 16 && !(lane & 1)) {
      red[wv * CP + gq] = gbacc;
      red[4 * CP + wv * CP + gq] = llacc;
    }
  } else {
#pragma unroll
    for (int off = 32; off >= CP; off >>= 1) {
      gbacc += __shfl_xor(gbacc, off);
      llacc += __shfl_xor(llacc, off);
    }
    if (lane < CP) {
      red[wv * CP + lane] = gbacc;
      red[4 * CP + wv * CP + lane] = llacc;
    }
  }
  __syncthreads();
  double* pb = A.part + (size_t)blockIdx.x * CP * VS;
  for (int v = tid; v < CP * D; v += kPassThreads) {
    const int c = v / D, j = v % D, qq = j / F, jj = j % F;
    double s = 0.0;
    if constexpr (MB) {
#pragma unroll
      for (int w = 0; w < kPassThreads / 64; ++w) s += (double)gl[(w * CP + c) * D + j];
    } else {
#pragma unroll
      for (int w = 0; w < kPassThreads / 64; ++w)
#pragma unroll
        for (int k = 0; k < 64 / LPR; ++k) s += (double)gl[(w * 64 + k * LPR + qq) * GS + c * F + jj];
    }
    pb[c * VS + j] = s;
  }
  if (tid < CP) {
    double sb = 0.0, sl = 0.0;
#pragma unroll
    for (int w = 0; w < kPassThreads / 64; ++w) {
      sb += red[w * CP + tid];
      sl += red[4 * CP + w * CP + tid];
    }
    pb[tid * VS + D] = sb;
    pb[tid * VS + D + 1] = sl;
  }
}

size_t pass_r_smem_bytes(int cp) {
  return ((sizeof(float) * kPassThreads * (cp * kRowFeat + 1) + 15) & ~(size_t)15) + sizeof(double) * 8 * cp;
}

// rows per tile of the register form
inline int pass_r_rows(int d, int cp) { return kPassThreads / (d / kRowFeat) * row_rpt(cp); }

size_t pass_t_smem_bytes(int d, int cp) {
  size_t b = sizeof(float) * ((size_t)kTileRows * (d + 4) + (size_t)cp * (d + 4) + (size_t)kTileRows * cp) +
             sizeof(int) * kTileRows;
  b = (b + 15) & ~(size_t)15;
  b += sizeof(double) * 6 * kPassThreads + 16;
  const size_t tot = sizeof(double) * (size_t)cp * (d + 2);
  return b > tot ? b : tot;
}

typedef void (*PassKernel)(PassArgs);

// Tiles the specialised pass keeps in flight per block.
inline int pass_ring(int d) { return d <= 64 ? 3 : 2; }

// Specialised pass for (d, cp) when one is compiled, else null.
PassKernel pick_pass(int d, int cp) {
#define AUGUR_PASS(D_, C_) if (d == D_ && cp == C_) return k_hlr_pass_t<D_, C_>;
  AUGUR_PASS(16, 1) AUGUR_PASS(16, 2) AUGUR_PASS(16, 4) AUGUR_PASS(16, 8) AUGUR_PASS(16, 16)
  AUGUR_PASS(32, 1) AUGUR_PASS(32, 2) AUGUR_PASS(32, 4) AUGUR_PASS(32, 8) AUGUR_PASS(32, 16)
  AUGUR_PASS(64, 1) AUGUR_PASS(64, 2) AUGUR_PASS(64, 4) AUGUR_PASS(64, 8) AUGUR_PASS(64, 16)
  AUGUR_PASS(128, 1) AUGUR_PASS(128, 2) AUGUR_PASS(128, 4) AUGUR_PASS(128, 8) AUGUR_PASS(128, 16)
#undef AUGUR_PASS
  return nullptr;
}

// Register form of the pass for (d, cp), if compiled.
// At D = 128 and 8 chains phase C runs on the matrix cores unless
// AUGUR_HLR_VALU_BACK=1.
PassKernel pick_pass_r(int d, int cp) {
#define AUGUR_PASS(D_, C_) if (d == D_ && cp == C_) return k_hlr_pass_r<D_, C_, false>;
  AUGUR_PASS(16, 1) AUGUR_PASS(16, 2)
  AUGUR_PASS(32, 1) AUGUR_PASS(32, 2) AUGUR_PASS(32, 4)
  AUGUR_PASS(64, 1) AUGUR_PASS(64, 2) AUGUR_PASS(64, 4)
  AUGUR_PASS(128, 1) AUGUR_PASS(128, 2) AUGUR_PASS(128, 4)
#undef AUGUR_PASS
  if (d == 128 && cp == 8) {
    const char* e = getenv("AUGUR_HLR_VALU_BACK");
    return (e && atoi(e)) ? k_hlr_pass_r<128, 8, false> : k_hlr_pass_r<128, 8, true>;
  }
  return nullptr;
}

// The gradient-only kernel of pick_pass_r's form, if compiled.
PassKernel pick_pass_r_grad(int d, int cp) {
#define AUGUR_PASS(D_, C_) if (d == D_ && cp == C_) return k_hlr_pass_r<D_, C_, false, false>;
  AUGUR_PASS(16, 1) AUGUR_PASS(16, 2)
  AUGUR_PASS(32, 1) AUGUR_PASS(32, 2) AUGUR_PASS(32, 4)
  AUGUR_PASS(64, 1) AUGUR_PASS(64, 2) AUGUR_PASS(64, 4)
  AUGUR_PASS(128, 1) AUGUR_PASS(128, 2) AUGUR_PASS(128, 4)
#undef AUGUR_PASS
  if (d == 128 && cp == 8) {
    const char* e = getenv("AUGUR_HLR_VALU_BACK");
    return (e && atoi(e)) ? k_hlr_pass_r<128, 8, false, false> : k_hlr_pass_r<128, 8, true, false>;
  }
  return nullptr;
}

size_t pass_smem_bytes(int d, int cp) {
  size_t b = sizeof(float) * ((size_t)kTileRows * (d + 1) + (size_t)cp * (d + 1) + (size_t)kTileRows * cp) +
             sizeof(int) * kTileRows;
  b = (b + 15) & ~(size_t)15;
  b += sizeof(double) * 2 * kPassThreads + 16;
  const size_t tot = sizeof(double) * (size_t)cp * (d + 2);
  return b > tot ? b : tot;
}

struct HmcArgs {
  int d, cp, c_real, has_b, ref_compat;
  int blk_theta, blk_b, block_kind;  // block_kind: AUGUR_HMC_BLOCK_*
  uint32_t seed, sweep, tag;
  const double* cache_g;
  const double* cache_ll;
  double* cache_g_w;
  double* cache_ll_w;
  const double* gprop;
  const double* llprop;
  double* state;
  double* prop;
  double* p0;
  double* p;
  double* k0;
  float* beta32;
  double* stats;  // [chain][2] for this block
  HmcStep hs;
};


// One wave per chain: momentum draw, half step, first position step.
__device__ inline void hmc_begin_body(const HmcArgs& A) {
  const int c = blockIdx.x, lane = threadIdx.x;
  const int d = A.d, ds = d + 1, P = d + 2;
  const Stream mom{A.seed, (uint32_t)c, A.sweep, A.tag + 0};
  const double step = A.hs.step[c];
  const double v = A.state[c * P + d + 1];
  double kin = 0.0;
  for (int j = lane; j < P; j += 64) {
    const double sv = A.state[c * P + j];
    double xv = sv;
    if (j < ds) {
      const bool in_blk = (j < d) ? A.blk_theta : (A.blk_b && A.has_b);
      if (in_blk) {
        const uint32_t site = A.block_kind == AUGUR_HMC_BLOCK_B ? 0u : (uint32_t)j;
        const double p0 = normal_at(mom, site, 0);
        const double g = A.cache_g[c * ds + j] + (-sv / v);
        const double pj = p0 + 0.5 * step * g;
        A.p0[c * ds + j] = p0;
        A.p[c * ds + j] = pj;
        kin += p0 * p0;
        xv = sv + step * pj;
      }
      A.beta32[c * ds + j] = (float)xv;
    }
    A.prop[c * P + j] = xv;
  }
  kin = wave_sum(kin);
  if (lane == 0) A.k0[c] = 0.5 * kin;
}

__global__ __launch_bounds__(64) void k_hmc_begin(HmcArgs A) { hmc_begin_body(A); }

// One wave per chain: energies, FP64 Metropolis test (cbits/augur_mcmclib.h:
// 70-83, 410-423), accept or reject.
__global__ __launch_bounds__(64) void k_hmc_end(HmcArgs A) {
  const int c = blockIdx.x, lane = threadIdx.x;
  const int d = A.d, ds = d + 1, P = d + 2;
  const double v = A.state[c * P + d + 1];
  double kin = 0.0, lp0 = 0.0, lp1 = 0.0;
  for (int j = lane; j < ds; j += 64) {
    const bool in_blk = (j < d) ? A.blk_theta : (A.blk_b && A.has_b);
    if (!in_blk) continue;
    const double pj = A.p[c * ds + j];
    kin += pj * pj;
    lp0 += normal_ll0(A.state[c * P + j], v, A.ref_compat);
    lp1 += normal_ll0(A.prop[c * P + j], v, A.ref_compat);
  }
  kin = wave_sum(kin);
  lp0 = wave_sum(lp0);
  lp1 = wave_sum(lp1);
  const double ll0 = A.cache_ll[c] + lp0;
  const double ll1 = A.llprop[c] + lp1;
  const double k1 = 0.5 * kin;
  const Stream acc_s{A.seed, (uint32_t)c, A.sweep, A.tag + 2};
  const double u = u01(acc_s.draw(0, 0).x);
  const bool acc = log(u) < (ll1 - ll0) + (A.k0[c] - k1);
  if (acc) {
    for (int j = lane; j < P; j += 64) A.state[c * P + j] = A.prop[c * P + j];
    if (A.gprop)  // null in surrogate mode: no exact gradient at the proposal
      for (int j = lane; j < ds; j += 64) A.cache_g_w[c * ds + j] = A.gprop[c * ds + j];
    if (lane == 0) A.cache_ll_w[c] = A.llprop[c];
  }
  if (lane == 0) {
    A.stats[2 * c] += acc ? 1.0 : 0.0;
    A.stats[2 * c + 1] += 1.0;
  }
}

struct VBlockArgs {
  int d, has_b, ref_compat, prior_kind;
  double prior_a, prior_b;
  double sim_len, eps;  // MWG: eps is the proposal variance
  int fixed_steps;
  int mwg;              // 0: HMC, 1: Metropolis-within-Gibbs
  uint32_t seed, sweep, tag;
  double* state;  // C x P
  double* stats;  // [chain][2]
};

__device__ inline double v_prior_ll(const VBlockArgs& A, double v) {
  return v_prior_ll0(A.prior_kind, A.prior_a, A.prior_b, v);
}

// HMC or Metropolis-within-Gibbs on the scalar variance v (the full
// conditional involves only theta and b): one wave per chain.
__device__ inline void hmc_vblock_body(const VBlockArgs& A) {
  const int c = blockIdx.x, lane = threadIdx.x;
  const int d = A.d, P = d + 2;
  double* st = A.state + (size_t)c * P;
  const int nterm = d + (A.has_b ? 1 : 0);
  double S = 0.0;
  for (int j = lane; j < nterm; j += 64) S += st[j] * st[j];
  S = wave_sum(S);
  double v;
  const bool acc = v_update(nterm, S, st[d + 1], A.ref_compat, A.prior_kind, A.prior_a, A.prior_b, A.mwg,
                            A.sim_len, A.eps, A.fixed_steps, A.seed, (uint32_t)c, A.sweep, A.tag, &v);
  if (lane == 0) {
    st[d + 1] = v;
    A.stats[2 * c] += acc ? 1.0 : 0.0;
    A.stats[2 * c + 1] += 1.0;
  }
}

__global__ __launch_bounds__(64) void k_hmc_vblock(VBlockArgs A) { hmc_vblock_body(A); }

// A v block followed by a data block in one launch: the v block's update of
// v, a barrier that publishes it to the block's lanes, then the data block's
// momentum draw and first steps, which read it.  Saves a launch per sweep.
__global__ __launch_bounds__(64) void k_hmc_vblock_begin(VBlockArgs V, HmcArgs A) {
  hmc_vblock_body(V);
  __syncthreads();
  hmc_begin_body(A);
}

// Metropolis-within-Gibbs on the intercept b, the schedule item
// `MWG [b] ~ Normal(b, s)` (compiler/augur/test/hlr2.infer;
// cbits/augur_mcmclib.h:178-193 with Core/CgUserProp.hs's proposal): the
// proposal b' = b + sqrt(s) z (normal at tag + 0) is written as the pass's
// FP32 evaluation point; k_mwg_b_end accepts on the full conditional of b,
// prior Normal(0, v) plus the data term, against the cached data term of the
// current point (accept uniform at tag + 2).  The Hastings term as for v:
// (b' - b)^2 / (2 s) with ref_compat, else 0.
struct MwgBArgs {
  int d, cp, c_real, ref_compat;
  double var;
  uint32_t seed, sweep, tag;
  double* state;       // C x P
  double* prop_b;      // C: the proposed b
  float* beta32;       // cp x (d + 1)
  double* cache_g;     // cp x (d + 1)
  double* cache_ll;    // cp
  const double* gprop; // cp x (d + 1): the pass at the proposal
  const double* llprop;
  double* stats;       // [chain][2]
};

__global__ void k_mwg_b_begin(MwgBArgs A) {
  const int ds = A.d + 1, P = A.d + 2;
  for (int q = threadIdx.x; q < A.cp * ds; q += blockDim.x) {
    const int c = q / ds, j = q % ds;
    float v = 0.0f;
    if (c < A.c_real) {
      const double* st = A.state + (size_t)c * P;
      if (j < A.d) {
        v = (float)st[j];
      } else {
        const double b1 = st[A.d] + sqrt(A.var) * normal_at(Stream{A.seed, (uint32_t)c, A.sweep, A.tag}, 0, 0);
        A.prop_b[c] = b1;
        v = (float)b1;
      }
    }
    A.beta32[q] = v;
  }
}

__global__ __launch_bounds__(64) void k_mwg_b_end(MwgBArgs A) {
  const int c = blockIdx.x, lane = threadIdx.x, d = A.d, ds = d + 1, P = d + 2;
  if (c >= A.c_real) return;
  double* st = A.state + (size_t)c * P;
  const double b0 = st[d], b1 = A.prop_b[c], v = st[d + 1];
  const double ll0 = normal_ll0(b0, v, A.ref_compat) + A.cache_ll[c];
  const double ll1 = normal_ll0(b1, v, A.ref_compat) + A.llprop[c];
  const double extra = A.ref_compat ? 0.5 * (b1 - b0) * (b1 - b0) / A.var : 0.0;
  const double u = u01(Stream{A.seed, (uint32_t)c, A.sweep, A.tag + 2}.draw(0, 0).x);
  const bool acc = log(u) < ll1 - ll0 + extra;
  if (acc) {
    for (int j = lane; j < ds; j += 64) A.cache_g[c * ds + j] = A.gprop[c * ds + j];
    if (lane == 0) {
      st[d] = b1;
      A.cache_ll[c] = A.llprop[c];
    }
  }
  if (lane == 0) {
    A.stats[2 * c] += acc ? 1.0 : 0.0;
    A.stats[2 * c + 1] += 1.0;
  }
}

// The deterministic parameter bias[n] = sigmoid(dotprod(theta, x[n]) + b) of
// compiler/augur/test/dethlr.rv for every chain (Core/CgDet.hs), in FP64
// from the FP32 rows: out[c][n].
__global__ void k_hlr_det_bias(const float* __restrict__ x, const double* __restrict__ state, int n, int d,
                               int chains, int has_b, double* __restrict__ out) {
  const int P = d + 2;
  const size_t total = (size_t)chains * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / n), r = (int)(e % n);
    const double* st = state + (size_t)c * P;
    const float* xr = x + (size_t)r * d;
    double t = has_b ? st[d] : 0.0;
    for (int j = 0; j < d; ++j) t += (double)xr[j] * st[j];
    out[e] = 1.0 / (1.0 + exp(-t));  // cbits/augur_math.c:52-54
  }
}

__global__ void k_state_to_beta32(const double* __restrict__ state, float* __restrict__ beta32, int d,
                                  int cp, int c_real, int has_b) {
  const int ds = d + 1, P = d + 2;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cp * ds; q += gridDim.x * blockDim.x) {
    const int c = q / ds, j = q % ds;
    float v = 0.0f;
    if (c < c_real && (j < d || has_b)) v = (float)state[c * P + j];
    beta32[q] = v;
  }
}

// Full log joint (reference ll_mod): v prior + theta / b priors + data.
__global__ void k_hlr_ll_mod(const double* __restrict__ state, const double* __restrict__ cache_ll,
                             VBlockArgs A, int c_real, double* out) {
  const int c = blockIdx.x, lane = threadIdx.x;
  if (c >= c_real) return;
  const int d = A.d, P = d + 2;
  const double* st = state + (size_t)c * P;
  const double v = st[d + 1];
  const int nterm = d + (A.has_b ? 1 : 0);
  double s = 0.0;
  for (int j = lane; j < nterm; j += 64) s += normal_ll0(st[j], v, A.ref_compat);
  s = wave_sum(s);
  if (lane == 0) out[c] = s + v_prior_ll(A, v) + cache_ll[c];
}

static int next_pow2(int c) {
  int p = 1;
  while (p < c) p <<= 1;
  return p;
}

struct HlrSampler final : augur_sampler {
  int n = 0, d = 0, cp = 1, has_b = 0, ref_compat = 1, prior_kind = 0;
  double prior_a = 1.0, prior_b = 1.0;
  int G = 1;
  int G64 = 1;                     // grid of the FP64 log-density (launch_ll64)
  PassKernel fast_pass = nullptr;  // specialised pass for (d, cp), if compiled
  PassKernel fast_pass_grad = nullptr;  // its gradient-only form, if any
  size_t fast_smem = 0;            // its dynamic LDS bytes
  std::vector<augur_hmc_block> sched;
  DevBuf<float> x, beta32;
  DevBuf<int32_t> y;
  DevBuf<double> state, prop, p0, p, k0, cache_g, cache_ll, gprop, llprop, part, stats_buf, ll_out;
  bool cache_valid = false;

  // ---- surrogate-gradient HMC on one theta block ----
  enum { SURR_OFF = 0, SURR_COLLECT = 1, SURR_ACTIVE = 2, SURR_FALLBACK = 3 };
  int surr_phase = SURR_OFF;
  augur_surrogate_cfg scfg{};
  int surr_sweeps_collected = 0, n_pairs = 0, max_pairs = 0, surr_iters = 0, fallbacks = 0;
  double surr_loss = 0.0, surr_train_ms = 0.0, window_accept = -1.0, win_acc0 = 0.0, win_prop0 = 0.0;
  Mlp mlp;
  DevBuf<float> cap_x, cap_y, sstats, sstats2, xb, tb;
  DevBuf<double> lin_b;
  SurrInfer inf;
  double resid_frac = -1.0;  // residual variance of the linear term / target variance
  DevBuf<double> ll64_cur, ll64_prop, ll64_part;
  DevBuf<double> prop_b;  // MWG [b]: the proposals
  DevBuf<double> det;     // chains x n: the deterministic bias[n] (dethlr.rv), computed on read
  // ll64_cur no longer holds the FP64 data log-density of the current state:
  // set_pt wrote theta / b, or another block moved them since the last
  // surrogate iteration
  bool ll64_dirty = false;

  const char* family() const override { return "hlr"; }

  ParamView param(const std::string& name) override {
    const size_t P = (size_t)d + 2;
    ParamView v;
    if (name == "state") v = {state.p, sizeof(double), P, P, true};
    else if (name == "theta") v = {state.p, sizeof(double), (size_t)d, P, true};
    else if (name == "b" && has_b) v = {state.p + d, sizeof(double), 1, P, true};
    else if (name == "v") v = {state.p + d + 1, sizeof(double), 1, P, true};
    else if (name == "bias") {
      if (!det.p) det.alloc((size_t)chains * n);
      v = {det.p, sizeof(double), (size_t)n, (size_t)n, false};
    }
    return v;
  }

  void prepare(const std::string& name) override {
    if (name != "bias") return;
    if (!det.p) det.alloc((size_t)chains * n);
    k_hlr_det_bias<<<1024, 256, 0, stream>>>(x.p, state.p, n, d, chains, has_b, det.p);
    AUGUR_CHECK_LAUNCH();
  }

  void on_set(const std::string& name) override {
    if (name != "v") {
      cache_valid = false;
      ll64_dirty = true;
    }
  }

  PassArgs pass_args(int mode, int l) {
    PassArgs A{};
    A.x = x.p; A.y = y.p; A.n = n; A.d = d; A.cp = cp; A.c_real = chains; A.has_b = has_b;
    A.beta32 = beta32.p; A.part = part.p; A.G = G; A.mode = mode; A.l = l;
    A.cache_g = cache_g.p; A.cache_ll = cache_ll.p; A.gprop = gprop.p; A.llprop = llprop.p;
    A.prop = prop.p; A.p = p.p; A.state = state.p;
    return A;
  }

  // grad_only: a leapfrog pass before every chain's last (no log-density)
  void launch_pass(const PassArgs& A, bool grad_only = false) {
    PassKernel k = (grad_only && fast_pass_grad) ? fast_pass_grad : fast_pass;
    size_t smem = fast_smem;
    if (!k) {
      k = k_hlr_pass;
      smem = pass_smem_bytes(d, cp);
      allow_lds(k_hlr_pass, smem);
    }
    if (auto* ev = prof_slot()) {
      hipExtLaunchKernelGGL(k, dim3(G), dim3(kPassThreads), smem, stream, ev->first, ev->second, 0, A);
    } else {
      k<<<G, kPassThreads, smem, stream>>>(A);
    }
    AUGUR_CHECK_LAUNCH();
    k_hlr_finish<<<(cp * (d + 2) + kFinVals - 1) / kFinVals, 1024, 0, stream>>>(A);
    AUGUR_CHECK_LAUNCH();
  }

  void refresh_cache() {
    k_state_to_beta32<<<8, 256, 0, stream>>>(state.p, beta32.p, d, cp, chains, has_b);
    AUGUR_CHECK_LAUNCH();
    launch_pass(pass_args(MODE_CACHE, 0));
    cache_valid = true;
  }

  HmcArgs end_args_surrogate(int bi) {
    HmcArgs H{};
    H.d = d; H.cp = cp; H.c_real = chains; H.has_b = has_b; H.ref_compat = ref_compat;
    H.blk_theta = 1; H.blk_b = 0; H.block_kind = AUGUR_HMC_BLOCK_THETA;
    H.seed = seed; H.sweep = sweep; H.tag = TAG_HMC_BASE + 4u * (uint32_t)bi;
    H.cache_g = cache_g.p; H.cache_ll = ll64_cur.p; H.cache_g_w = cache_g.p; H.cache_ll_w = ll64_cur.p;
    H.gprop = nullptr; H.llprop = ll64_prop.p; H.state = state.p; H.prop = prop.p; H.p0 = p0.p;
    H.p = p.p; H.k0 = k0.p; H.beta32 = beta32.p; H.stats = stats_buf.p + (size_t)bi * chains * 2;
    return H;
  }

  Ll64Args ll64_args(const double* point, double* out) {
    return Ll64Args{x.p, y.p, n, d, cp, has_b, d + 2, point, ll64_part.p, G64, out};
  }

  // Fits the surrogate to the captured pairs (standardisation, least-squares
  // linear term, MLP on the linear term's residual), folds it into the
  // inference form, then switches the block to surrogate trajectories.
  void train_surrogate() {
    AUGUR_HIP(hipStreamSynchronize(stream));
    const auto t0 = std::chrono::steady_clock::now();
    mlp_colstats(cap_x.p, cap_y.p, n_pairs, d, sstats.p, stream);
    surr_standardise(cap_x.p, cap_y.p, n_pairs, d, sstats.p, stream);
    surr_linear_fit(cap_x.p, cap_y.p, n_pairs, d, 1e-6, lin_b.p, stream);
    mlp_colstats(cap_x.p, cap_y.p, n_pairs, d, sstats2.p, stream);
    const long steps = scfg.linear_only ? 0 : std::max<long>(1, (long)scfg.epochs * n_pairs / scfg.batch);
    if (scfg.linear_only) {
      // no network: its output layer is zero, so the fitted gradient is the
      // linear term alone
      const MlpShape sh{d, scfg.hidden};
      AUGUR_HIP(hipMemsetAsync(mlp.params.p + sh.w3(), 0, sizeof(float) * (sh.count() - sh.w3()), stream));
    }
    double lsum = 0.0;
    int lcount = 0;
    for (long s = 0; s < steps; ++s) {
      mlp_gather(cap_x.p, cap_y.p, n_pairs, d, sstats2.p, scfg.batch, scfg.seed, (uint32_t)s, xb.p, tb.p, stream);
      mlp.loss_grad(xb.p, tb.p, scfg.batch, stream);
      if (s >= steps - 20) {
        double l = 0.0;
        mlp.loss.download(&l, 1, stream);
        AUGUR_HIP(hipStreamSynchronize(stream));
        lsum += l;
        ++lcount;
      }
      mlp.adam(scfg.lr, stream);
    }
    surr_loss = lcount ? lsum / lcount : 0.0;
    surr_fold(d, scfg.hidden, sstats.p, sstats2.p, lin_b.p, mlp.params.p, inf, stream);
    launch_ll64(ll64_args(state.p, ll64_cur.p), stream);
    ll64_dirty = false;
    std::vector<double> st((size_t)sched.size() * chains * 2);
    stats_buf.download(st.data(), st.size(), stream);
    std::vector<float> s2(4 * (size_t)d);
    sstats2.download(s2.data(), s2.size(), stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
    double rv = 0.0;
    for (int j = 0; j < d; ++j) rv += 1.0 / ((double)s2[3 * d + j] * s2[3 * d + j]);
    resid_frac = rv / d;
    surr_train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    win_acc0 = win_prop0 = 0.0;
    for (int c = 0; c < chains; ++c) {
      win_acc0 += st[((size_t)scfg.block_index * chains + c) * 2];
      win_prop0 += st[((size_t)scfg.block_index * chains + c) * 2 + 1];
    }
    surr_phase = SURR_ACTIVE;
  }

  void surrogate_iteration(int bi, const augur_hmc_block& blk) {
    SurrTrajArgs S{};
    S.d = d; S.h = scfg.hidden; S.P = d + 2; S.ds = d + 1; S.c_real = chains;
    S.seed = seed; S.sweep = sweep; S.tag = TAG_HMC_BASE + 4u * (uint32_t)bi;
    // the block's own (jittered) schedule, or the configured fixed one
    for (int c = 0; c < chains; ++c) {
      if (scfg.steps > 0) {
        S.hs.step[c] = scfg.step_size;
        S.hs.steps[c] = scfg.steps;
      } else {
        hmc_jitter(Stream{seed, (uint32_t)c, sweep, S.tag + 1}, blk.sim_len, blk.step_size, blk.fixed_steps,
                   &S.hs.step[c], &S.hs.steps[c]);
      }
    }
    const MlpShape sh{d, scfg.hidden};
    S.w1t = inf.w1t.p; S.b1 = mlp.params.p + sh.b1(); S.w2t = inf.w2t.p; S.b2 = mlp.params.p + sh.b2();
    S.w3t = inf.w3t.p; S.b3 = inf.b3.p;
    S.in_mean = inf.in_mean.p; S.in_scale = inf.in_scale.p; S.lin_mean = inf.lin_mean.p; S.lin_t = inf.lin_t.p;
    S.state = state.p; S.prop = prop.p; S.p = p.p; S.k0 = k0.p; S.beta32 = beta32.p;
    if (ll64_dirty) {
      launch_ll64(ll64_args(state.p, ll64_cur.p), stream);
      ll64_dirty = false;
    }
    launch_surr_traj(S, chains, stream);
    launch_ll64(ll64_args(prop.p, ll64_prop.p), stream);
    k_hmc_end<<<chains, 64, 0, stream>>>(end_args_surrogate(bi));
    AUGUR_CHECK_LAUNCH();
    cache_valid = false;  // the FP32 data cache no longer tracks the state
    if (++surr_iters % scfg.window == 0) {
      std::vector<double> st((size_t)sched.size() * chains * 2);
      stats_buf.download(st.data(), st.size(), stream);
      AUGUR_HIP(hipStreamSynchronize(stream));
      double a = 0.0, pr = 0.0;
      for (int c = 0; c < chains; ++c) {
        a += st[((size_t)bi * chains + c) * 2];
        pr += st[((size_t)bi * chains + c) * 2 + 1];
      }
      window_accept = (a - win_acc0) / std::max(1.0, pr - win_prop0);
      win_acc0 = a;
      win_prop0 = pr;
      if (window_accept < scfg.min_accept) {
        surr_phase = SURR_FALLBACK;
        ++fallbacks;
      }
    }
  }

  void mwg_b_block(int bi, const augur_hmc_block& blk) {
    if (!cache_valid) refresh_cache();
    MwgBArgs M{};
    M.d = d; M.cp = cp; M.c_real = chains; M.ref_compat = ref_compat; M.var = blk.step_size;
    M.seed = seed; M.sweep = sweep; M.tag = TAG_HMC_BASE + 4u * (uint32_t)bi;
    M.state = state.p; M.prop_b = prop_b.p; M.beta32 = beta32.p; M.cache_g = cache_g.p; M.cache_ll = cache_ll.p;
    M.gprop = gprop.p; M.llprop = llprop.p; M.stats = stats_buf.p + (size_t)bi * chains * 2;
    k_mwg_b_begin<<<1, 256, 0, stream>>>(M);
    AUGUR_CHECK_LAUNCH();
    PassArgs A = pass_args(MODE_CACHE, 0);  // the data term at the proposal, into gprop / llprop
    A.cache_g = gprop.p;
    A.cache_ll = llprop.p;
    launch_pass(A);
    k_mwg_b_end<<<chains, 64, 0, stream>>>(M);
    AUGUR_CHECK_LAUNCH();
    ll64_dirty = true;
  }

  // fuse_v: the v block scheduled just before, run in the same launch as
  // this block's first kernel
  void data_block(int bi, const augur_hmc_block& blk, const VBlockArgs* fuse_v = nullptr) {
    if (blk.block == AUGUR_MWG_BLOCK_B) {
      if (fuse_v) {
        k_hmc_vblock<<<chains, 64, 0, stream>>>(*fuse_v);
        AUGUR_CHECK_LAUNCH();
      }
      mwg_b_block(bi, blk);
      return;
    }
    const bool surr_blk = surr_phase != SURR_OFF && bi == scfg.block_index;
    if (surr_blk && surr_phase == SURR_ACTIVE) {
      if (fuse_v) {
        k_hmc_vblock<<<chains, 64, 0, stream>>>(*fuse_v);
        AUGUR_CHECK_LAUNCH();
      }
      surrogate_iteration(bi, blk);
      return;
    }
    if (!cache_valid) refresh_cache();
    const uint32_t tag = TAG_HMC_BASE + 4u * (uint32_t)bi;
    HmcArgs H{};
    H.d = d; H.cp = cp; H.c_real = chains; H.has_b = has_b; H.ref_compat = ref_compat;
    H.blk_theta = blk.block == AUGUR_HMC_BLOCK_THETA || blk.block == AUGUR_HMC_BLOCK_B_THETA;
    H.blk_b = blk.block == AUGUR_HMC_BLOCK_B || blk.block == AUGUR_HMC_BLOCK_B_THETA;
    H.block_kind = blk.block;
    H.seed = seed; H.sweep = sweep; H.tag = tag;
    H.cache_g = cache_g.p; H.cache_ll = cache_ll.p; H.cache_g_w = cache_g.p; H.cache_ll_w = cache_ll.p;
    H.gprop = gprop.p; H.llprop = llprop.p; H.state = state.p; H.prop = prop.p; H.p0 = p0.p;
    H.p = p.p; H.k0 = k0.p; H.beta32 = beta32.p; H.stats = stats_buf.p + (size_t)bi * chains * 2;
    int lmax = 0, lmin = INT_MAX;
    for (int c = 0; c < chains; ++c) {
      hmc_jitter(Stream{seed, (uint32_t)c, sweep, tag + 1}, blk.sim_len, blk.step_size, blk.fixed_steps,
                 &H.hs.step[c], &H.hs.steps[c]);
      lmax = std::max(lmax, H.hs.steps[c]);
      lmin = std::min(lmin, H.hs.steps[c]);
    }
    if (fuse_v) k_hmc_vblock_begin<<<chains, 64, 0, stream>>>(*fuse_v, H);
    else k_hmc_begin<<<chains, 64, 0, stream>>>(H);
    AUGUR_CHECK_LAUNCH();
    PassArgs A = pass_args(MODE_LEAPFROG, 0);
    A.blk_theta = H.blk_theta;
    A.blk_b = H.blk_b;
    A.hs = H.hs;
    const bool capture = surr_blk && surr_phase == SURR_COLLECT;
    for (int l = 1; l <= lmax; ++l) {
      A.l = l;
      A.cap_x = nullptr;
      if (capture && n_pairs + chains <= max_pairs) {
        A.cap_x = cap_x.p;
        A.cap_y = cap_y.p;
        A.cap_row = n_pairs;
        n_pairs += chains;
      }
      launch_pass(A, l < lmin);
    }
    k_hmc_end<<<chains, 64, 0, stream>>>(H);
    AUGUR_CHECK_LAUNCH();
    if (H.blk_theta || H.blk_b) ll64_dirty = true;  // this block may have moved theta or b
    if (capture && ++surr_sweeps_collected >= scfg.train_sweeps && n_pairs > 0) train_surrogate();
  }

  VBlockArgs vargs(int bi, const augur_hmc_block& blk) {
    VBlockArgs V{};
    V.d = d; V.has_b = has_b; V.ref_compat = ref_compat; V.prior_kind = prior_kind;
    V.prior_a = prior_a; V.prior_b = prior_b; V.sim_len = blk.sim_len; V.eps = blk.step_size;
    V.fixed_steps = blk.fixed_steps; V.mwg = blk.block == AUGUR_MWG_BLOCK_V; V.seed = seed; V.sweep = sweep;
    V.tag = TAG_HMC_BASE + 4u * (uint32_t)bi; V.state = state.p;
    V.stats = stats_buf.p + (size_t)bi * chains * 2;
    return V;
  }

  void run_sweeps(int nsw) override {
    auto is_v = [](const augur_hmc_block& b) {
      return b.block == AUGUR_HMC_BLOCK_V || b.block == AUGUR_MWG_BLOCK_V;
    };
    for (int i = 0; i < nsw; ++i) {
      for (int bi = 0; bi < (int)sched.size(); ++bi) {
        const augur_hmc_block& blk = sched[bi];
        if (is_v(blk)) {
          const VBlockArgs va = vargs(bi, blk);
          if (bi + 1 < (int)sched.size() && !is_v(sched[bi + 1])) {
            data_block(bi + 1, sched[bi + 1], &va);  // v block fused into the next block's first launch
            ++bi;
          } else {
            k_hmc_vblock<<<chains, 64, 0, stream>>>(va);
            AUGUR_CHECK_LAUNCH();
          }
        } else {
          data_block(bi, blk);
        }
      }
      ++sweep;
      record_traces();
    }
  }

  void log_density(double* out) override {
    if (!cache_valid) refresh_cache();
    augur_hmc_block dummy{};
    k_hlr_ll_mod<<<cp, 64, 0, stream>>>(state.p, cache_ll.p, vargs(0, dummy), chains, ll_out.p);
    AUGUR_CHECK_LAUNCH();
    ll_out.download(out, chains, stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
  }

  void stats(double* out, size_t bytes) override {
    const size_t need = sizeof(double) * sched.size() * chains * 2;
    if (bytes < need) fail(AUGUR_ERR_ARG, "stats need %zu bytes", need);
    stats_buf.download(out, sched.size() * chains * 2, stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
  }

  void alloc(int n_, int d_, int chains_) {
    n = n_; d = d_; chains = chains_;
    cp = next_pow2(chains);
    int tiles = (n + kTileRows - 1) / kTileRows;
    const char* gen = getenv("AUGUR_HLR_GENERIC");
    const char* lds = getenv("AUGUR_HLR_PASS_LDS");
    fast_pass = (gen && atoi(gen)) ? nullptr : pick_pass(d, cp);
    fast_smem = pass_t_smem_bytes(d, cp);
    int ring = pass_ring(d);
    int per_cu = kPassBlocksPerCU;
    if (fast_pass && !(lds && atoi(lds))) {
      if (PassKernel r = pick_pass_r(d, cp)) {
        fast_pass = r;
        // AUGUR_HLR_GRAD_PASS=0: every pass computes the log-density too
        const char* gp = getenv("AUGUR_HLR_GRAD_PASS");
        fast_pass_grad = (gp && atoi(gp) == 0) ? nullptr : pick_pass_r_grad(d, cp);
        fast_smem = pass_r_smem_bytes(cp);
        tiles = (n + pass_r_rows(d, cp) - 1) / pass_r_rows(d, cp);
        ring = row_ring(cp);
        // two blocks per CU: fewer, longer blocks amortise the epilogue and
        // leave fewer partial rows (measured at N = 100k, D = 64, 4 chains:
        // 7.4 us pass + 4.1 us finish at 512 blocks, 8.2 + 4.3 at 768)
        per_cu = 2;
      }
    }
    int want;
    if (fast_pass) {
      // every block's tiles fit in the specialised pass's prefetch ring when
      // the grid allows (one memory round per pass), and every block is
      // resident at once (a second round of blocks costs ~35% at N = 100k)
      int cus = 0;
      AUGUR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
      want = std::min(per_cu * std::max(cus, 1), (tiles + ring - 1) / ring);
    } else {
      // about 256 rows per block, between one and four blocks per CU
      want = 256;
      while (want < 1024 && (size_t)want * 256 < (size_t)n) want *= 2;
    }
    if (const char* e = getenv("AUGUR_HLR_BLOCKS")) want = std::max(1, atoi(e));
    G = std::max(1, std::min(tiles, want));
    {
      int cus = 0;
      AUGUR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
      G64 = ll64_grid(n, d, cp, has_b, G, cus);
    }
    const size_t P = (size_t)d + 2, ds = (size_t)d + 1;
    // X and y padded with zero rows to a multiple of kHlrRowPad (the data
    // overwrite the first n rows)
    const size_t np = ((size_t)n + kHlrRowPad - 1) / kHlrRowPad * kHlrRowPad;
    x.alloc(np * d);
    y.alloc(std::max<size_t>(np, 1));
    x.zero(stream);
    y.zero(stream);
    beta32.alloc(cp * ds);
    state.alloc(cp * P); prop.alloc(cp * P);
    p0.alloc(cp * ds); p.alloc(cp * ds); k0.alloc(cp);
    cache_g.alloc(cp * ds); cache_ll.alloc(cp); gprop.alloc(cp * ds); llprop.alloc(cp);
    part.alloc((size_t)cp * (d + 2) * G);
    ll_out.alloc(cp);
    stats_buf.alloc(std::max<size_t>(1, 4 * (size_t)chains * 2));
    prop_b.alloc(cp);
    state.zero(stream); prop.zero(stream); beta32.zero(stream);
    stats_buf.zero(stream); p.zero(stream); p0.zero(stream);
  }
};

static void validate_shapes(int n, int d, int chains) {
  if (n < 1 || d < 1 || d > 256) fail(AUGUR_ERR_ARG, "need n >= 1 and 1 <= d <= 256 (got %d, %d)", n, d);
  if (chains < 1 || chains > kMaxChains) fail(AUGUR_ERR_ARG, "chains must be in [1, %d]", kMaxChains);
  const int cp = next_pow2(chains);
  if (d * cp > kMaxGradSlots * kPassThreads)
    fail(AUGUR_ERR_ARG, "d * chains (padded to a power of two) must be <= %d", kMaxGradSlots * kPassThreads);
}

// ---- the data term without allocation (augur_op_hlr_ll_grad_noalloc) ----
// Its scratch: the FP32 evaluation points (cp x (d + 1)), the per-chain sums
// (cp x (d + 1) gradients, cp log-densities) and the pass's block partials
// (G x cp x (d + 2)), at 256-byte aligned offsets.  One copy per device,
// held until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
constexpr int kNoAllocMaxBlocks = 512;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// The scratch's device address on the current device, looked up once per
// device with that device current.
static unsigned char* noalloc_scratch() {
  static std::mutex mu;
  static std::map<int, unsigned char*> ptrs;
  int dev = 0;
  AUGUR_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = ptrs.find(dev);
  if (it != ptrs.end()) return it->second;
  void* ptr = nullptr;
  AUGUR_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) fail(AUGUR_ERR_HIP, "no device address for the scratch array");
  ptrs[dev] = (unsigned char*)ptr;
  return (unsigned char*)ptr;
}

// beta (chains x (d + 1), the caller's layout) -> FP32 points cp x (d + 1),
// zero for padded chains and for b without an intercept
__global__ void k_beta_to_beta32(const double* __restrict__ beta, int d, int chains, int cp, int has_b,
                                 float* __restrict__ beta32) {
  const int ds = d + 1;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cp * ds; q += gridDim.x * blockDim.x) {
    const int c = q / ds, j = q % ds;
    float v = 0.0f;
    if (c < chains && (j < d || has_b)) v = (float)beta[q];
    beta32[q] = v;
  }
}

// per-chain sums (cp rows) -> the caller's ll (chains) and grad (chains x (d + 1))
__global__ void k_llgrad_out(const double* __restrict__ cache_g, const double* __restrict__ cache_ll, int d,
                             int chains, double* __restrict__ ll, double* __restrict__ grad) {
  const int ds = d + 1;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < chains * ds; q += gridDim.x * blockDim.x) {
    grad[q] = cache_g[q];
    if (q % ds == 0) ll[q / ds] = cache_ll[q / ds];
  }
}

}  // namespace augur

using namespace augur;

extern "C" {

augur_status augur_hlr_create(const augur_hlr_desc* D, augur_sampler** out) {
  return guard([&] {
    if (!out || !D) fail(AUGUR_ERR_ARG, "null argument");
    *out = nullptr;
    validate_shapes(D->n_rows, D->n_features, D->chains);
    if (!D->x || !D->y) fail(AUGUR_ERR_ARG, "null data");
    if (D->prior_kind != 0 && D->prior_kind != 1) fail(AUGUR_ERR_ARG, "prior_kind must be 0 or 1");
    if (D->n_blocks < 1 || D->n_blocks > 4) fail(AUGUR_ERR_ARG, "1..4 HMC blocks");
    for (int i = 0; i < D->n_blocks; ++i) {
      const augur_hmc_block& b = D->blocks[i];
      if (b.block < 0 || b.block > AUGUR_MWG_BLOCK_B) fail(AUGUR_ERR_ARG, "bad block kind %d", b.block);
      if ((b.block == AUGUR_HMC_BLOCK_B || b.block == AUGUR_HMC_BLOCK_B_THETA || b.block == AUGUR_MWG_BLOCK_B) &&
          !D->has_intercept)
        fail(AUGUR_ERR_ARG, "block over b without an intercept");
      if (b.block == AUGUR_MWG_BLOCK_V || b.block == AUGUR_MWG_BLOCK_B) {
        if (!(b.step_size > 0.0)) fail(AUGUR_ERR_ARG, "MWG block %d needs a proposal variance > 0", i);
        continue;
      }
      if (!(b.step_size > 0.0) || (b.fixed_steps == 0 && !(b.sim_len > 0.0)) || b.fixed_steps > 100000 ||
          b.fixed_steps < -100000)
        fail(AUGUR_ERR_ARG, "HMC block %d needs step_size > 0 and sim_len > 0 or fixed_steps", i);
      if (b.fixed_steps == 0 && b.sim_len / (0.8 * b.step_size) > 1.0e5)
        fail(AUGUR_ERR_ARG, "HMC block %d: sim_len / step_size too large", i);
    }
    for (int i = 0; i < D->n_rows; ++i)
      if (D->y[i] != 0 && D->y[i] != 1) fail(AUGUR_ERR_ARG, "y[%d] = %d is not 0 or 1", i, D->y[i]);
    require_device();
    auto* s = new HlrSampler();
    try {
      AUGUR_HIP(hipGetDevice(&s->device));
      AUGUR_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
      s->seed = D->seed;
      s->has_b = D->has_intercept;
      s->ref_compat = D->ref_compat;
      s->prior_kind = D->prior_kind;
      s->prior_a = D->prior_a;
      s->prior_b = D->prior_b;
      s->sched.assign(D->blocks, D->blocks + D->n_blocks);
      s->alloc(D->n_rows, D->n_features, D->chains);
      s->x.upload(D->x, (size_t)D->n_rows * D->n_features, s->stream);
      s->y.upload(D->y, D->n_rows, s->stream);
      AUGUR_HIP(hipStreamSynchronize(s->stream));
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

augur_status augur_op_hlr_ll_grad(int n_rows, int d, int chains, const float* x, const int32_t* y,
                                  const double* beta, int has_intercept, double* ll, double* grad) {
  return guard([&] {
    validate_shapes(n_rows, d, chains);
    if (!x || !y || !beta || !ll || !grad) fail(AUGUR_ERR_ARG, "null argument");
    require_device();
    HlrSampler s;
    AUGUR_HIP(hipGetDevice(&s.device));
    AUGUR_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    s.has_b = has_intercept;
    s.alloc(n_rows, d, chains);
    s.x.upload(x, (size_t)n_rows * d, s.stream);
    s.y.upload(y, n_rows, s.stream);
    const int P = d + 2;
    std::vector<double> st((size_t)s.cp * P, 0.0);
    for (int c = 0; c < chains; ++c)
      for (int j = 0; j <= d; ++j) st[(size_t)c * P + j] = (j < d || has_intercept) ? beta[(size_t)c * (d + 1) + j] : 0.0;
    s.state.upload(st.data(), st.size(), s.stream);
    s.refresh_cache();
    std::vector<double> g((size_t)s.cp * (d + 1)), l(s.cp);
    s.cache_g.download(g.data(), g.size(), s.stream);
    s.cache_ll.download(l.data(), l.size(), s.stream);
    AUGUR_HIP(hipStreamSynchronize(s.stream));
    for (int c = 0; c < chains; ++c) {
      ll[c] = l[c];
      for (int j = 0; j <= d; ++j) grad[(size_t)c * (d + 1) + j] = g[(size_t)c * (d + 1) + j];
    }
  });
}

augur_status augur_op_hlr_ll_grad_noalloc(int n_rows, int d, int chains, const float* x, const int32_t* y,
                                          const double* beta, int has_intercept, double* ll, double* grad,
                                          void* stream) {
  return guard([&] {
    validate_shapes(n_rows, d, chains);
    if (!x || !y || !beta || !ll || !grad) fail(AUGUR_ERR_ARG, "null argument");
    if ((reinterpret_cast<uintptr_t>(x) & 15) != 0) fail(AUGUR_ERR_ARG, "x must be 16-byte aligned");
    const int cp = next_pow2(chains), ds = d + 1;
    const size_t o_g = align256(sizeof(float) * cp * ds);
    const size_t o_ll = o_g + align256(sizeof(double) * cp * ds);
    const size_t o_part = o_ll + align256(sizeof(double) * cp);
    const size_t per_block = sizeof(double) * cp * (d + 2);
    if (o_part + per_block > kNoAllocScratchBytes)
      fail(AUGUR_ERR_ARG, "d = %d, chains = %d need more than the %zu-byte scratch", d, chains,
           kNoAllocScratchBytes);
    unsigned char* base = noalloc_scratch();
    const int tiles = (n_rows + kTileRows - 1) / kTileRows;
    const int G = (int)std::min<size_t>(std::min(tiles, kNoAllocMaxBlocks), (kNoAllocScratchBytes - o_part) / per_block);
    PassArgs A{};
    A.x = x; A.y = y; A.n = n_rows; A.d = d; A.cp = cp; A.c_real = chains; A.has_b = has_intercept ? 1 : 0;
    A.beta32 = (float*)base; A.cache_g = (double*)(base + o_g); A.cache_ll = (double*)(base + o_ll);
    A.part = (double*)(base + o_part); A.G = G; A.mode = MODE_CACHE;
    hipStream_t st = (hipStream_t)stream;
    k_beta_to_beta32<<<8, 256, 0, st>>>(beta, d, chains, cp, A.has_b, A.beta32);
    AUGUR_CHECK_LAUNCH();
    const size_t smem = pass_smem_bytes(d, cp);
    allow_lds(k_hlr_pass, smem);
    k_hlr_pass<<<G, kPassThreads, smem, st>>>(A);
    AUGUR_CHECK_LAUNCH();
    k_hlr_finish<<<(cp * (d + 2) + kFinVals - 1) / kFinVals, 1024, 0, st>>>(A);
    AUGUR_CHECK_LAUNCH();
    k_llgrad_out<<<8, 256, 0, st>>>(A.cache_g, A.cache_ll, d, chains, ll, grad);
    AUGUR_CHECK_LAUNCH();
  });
}

augur_status augur_op_hlr_ll64(int n_rows, int d, int chains, const float* x, const int32_t* y,
                               const double* beta, int has_intercept, double* ll) {
  return guard([&] {
    validate_shapes(n_rows, d, chains);
    if (!x || !y || !beta || !ll) fail(AUGUR_ERR_ARG, "null argument");
    require_device();
    HlrSampler s;
    AUGUR_HIP(hipGetDevice(&s.device));
    AUGUR_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    s.has_b = has_intercept;
    s.alloc(n_rows, d, chains);
    s.x.upload(x, (size_t)n_rows * d, s.stream);
    s.y.upload(y, n_rows, s.stream);
    const int P = d + 2;
    std::vector<double> st((size_t)s.cp * P, 0.0);
    for (int c = 0; c < chains; ++c)
      for (int j = 0; j <= d; ++j) st[(size_t)c * P + j] = (j < d || has_intercept) ? beta[(size_t)c * (d + 1) + j] : 0.0;
    s.state.upload(st.data(), st.size(), s.stream);
    s.ll64_part.alloc((size_t)s.cp * s.G64);
    s.ll64_cur.alloc(s.cp);
    launch_ll64(s.ll64_args(s.state.p, s.ll64_cur.p), s.stream);
    std::vector<double> l(s.cp);
    s.ll64_cur.download(l.data(), l.size(), s.stream);
    AUGUR_HIP(hipStreamSynchronize(s.stream));
    for (int c = 0; c < chains; ++c) ll[c] = l[c];
  });
}

augur_status augur_hlr_surrogate(augur_sampler* s, const augur_surrogate_cfg* cfg) {
  return guard([&] {
    auto* h = dynamic_cast<HlrSampler*>(s);
    if (!h || !cfg) fail(AUGUR_ERR_ARG, "needs an HLR sampler and a configuration");
    const augur_surrogate_cfg& c = *cfg;
    if (c.block_index < 0 || c.block_index >= (int)h->sched.size() ||
        h->sched[c.block_index].block != AUGUR_HMC_BLOCK_THETA)
      fail(AUGUR_ERR_ARG, "block_index must name an HMC [theta] block");
    if (h->has_b) fail(AUGUR_ERR_UNSUPPORTED, "the surrogate covers models without an intercept");
    if (h->d > 256 || c.hidden < 1 || c.hidden > 256) fail(AUGUR_ERR_ARG, "need d <= 256 and 1 <= hidden <= 256");
    if (c.train_sweeps < 1 || (c.epochs < 1 && !c.linear_only) || c.batch < 1 || c.batch > 65536 || !(c.lr > 0.0) ||
        c.window < 1 || c.steps < 0 || (c.steps > 0 && !(c.step_size > 0.0)) || c.max_pairs < h->chains)
      fail(AUGUR_ERR_ARG, "bad surrogate settings");
    if (h->surr_phase != HlrSampler::SURR_OFF) fail(AUGUR_ERR_STATE, "surrogate already configured");
    h->scfg = c;
    h->max_pairs = c.max_pairs;
    h->cap_x.alloc((size_t)c.max_pairs * h->d);
    h->cap_y.alloc((size_t)c.max_pairs * h->d);
    h->sstats.alloc(4 * (size_t)h->d);
    h->sstats2.alloc(4 * (size_t)h->d);
    h->lin_b.alloc((size_t)h->d * h->d);
    h->xb.alloc((size_t)c.batch * h->d);
    h->tb.alloc((size_t)c.batch * h->d);
    h->inf.alloc(h->d, c.hidden);
    h->ll64_cur.alloc(h->cp);
    h->ll64_prop.alloc(h->cp);
    h->ll64_part.alloc((size_t)h->cp * h->G64);
    h->mlp.init(h->d, c.hidden, c.batch, c.seed, h->stream);
    AUGUR_HIP(hipStreamSynchronize(h->stream));
    h->surr_phase = HlrSampler::SURR_COLLECT;
  });
}

augur_status augur_hlr_surrogate_info(augur_sampler* s, double* info, size_t bytes) {
  return guard([&] {
    auto* h = dynamic_cast<HlrSampler*>(s);
    if (!h || !info || bytes < 8 * sizeof(double)) fail(AUGUR_ERR_ARG, "bad arguments");
    info[0] = h->surr_phase;
    info[1] = h->n_pairs;
    info[2] = h->surr_loss;
    info[3] = h->surr_train_ms;
    info[4] = h->fallbacks;
    info[5] = h->surr_iters;
    info[6] = h->window_accept;
    info[7] = h->surr_sweeps_collected;
    if (bytes >= 9 * sizeof(double)) info[8] = h->resid_frac;
  });
}

}  // extern "C"
