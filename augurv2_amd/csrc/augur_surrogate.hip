// augur_surrogate.hip -- the fitted gradient of surrogate HMC (linear least
// squares + residual MLP, folded into an inference form), the surrogate
// leapfrog trajectories, and the FP64 data log-density used by their
// Metropolis test.
#include "augur_surrogate.h"

#include "augur_mlp.h"

namespace augur {

__device__ inline float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ inline float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
__device__ inline float bf16_to_f(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }

__device__ inline uint16_t f_to_bf16(float f) {
  // round to nearest even by the hardware conversion (V_CVT_PK_BF16_F32)
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}

__device__ inline float gelu_s(float z) {
  return 0.5f * z * (1.0f + tanhf(0.7978845608028654f * (z + 0.044715f * z * z * z)));
}

// ---------------------------------------------------------------------------
// Surrogate trajectories.  The leapfrog follows the reference's order
// (cbits/augur_mcmclib.h:343-375): half step of momentum, then alternating
// full position / momentum steps, closing with a half momentum step; L + 1
// force evaluations.  Positions and momenta in FP64; the fitted gradient in
// FP32 with BF16 network weights.

// Generic form: one 256-thread block per chain, weights read from memory.
__global__ __launch_bounds__(256) void k_surr_traj(SurrTrajArgs A) {
  __shared__ double th[256], pp[256], gg[256], red[256];
  __shared__ float xin[256], xc[256], h1[256], h2[256], o3[256];
  const int c = blockIdx.x, tid = threadIdx.x, d = A.d, h = A.h;
  const double v = A.state[(size_t)c * A.P + d + 1];
  const double eps = A.hs.step[c];
  const int L = A.hs.steps[c];
  const Stream mom{A.seed, (uint32_t)c, A.sweep, A.tag};
  double kin = 0.0;
  if (tid < d) {
    th[tid] = A.state[(size_t)c * A.P + tid];
    const double p0 = normal_at(mom, (uint32_t)tid, 0);
    pp[tid] = p0;
    kin = p0 * p0;
  }
  red[tid] = kin;
  const int nparts = d >= 256 ? 1 : 256 / d;
  auto grad = [&]() {
    if (tid < d) {
      xin[tid] = (float)(th[tid] - (double)A.in_mean[tid]) * A.in_scale[tid];
      xc[tid] = (float)(th[tid] - (double)A.lin_mean[tid]);
    }
    __syncthreads();
    if (tid < h) {
      float s = A.b1[tid];
      for (int j = 0; j < d; ++j) s = fmaf(bf16_to_f(A.w1t[(size_t)j * h + tid]), xin[j], s);
      h1[tid] = gelu_s(s);
    }
    __syncthreads();
    if (tid < h) {
      float s = A.b2[tid];
      for (int t = 0; t < h; ++t) s = fmaf(bf16_to_f(A.w2t[(size_t)t * h + tid]), h1[t], s);
      h2[tid] = gelu_s(s);
    }
    __syncthreads();
    {
      const int j = tid % d, part = tid / d;
      float s = 0.0f;
      if (part < nparts) {
        const int t0 = part * h / nparts, t1 = (part + 1) * h / nparts;
        for (int t = t0; t < t1; ++t) s = fmaf(bf16_to_f(A.w3t[(size_t)t * d + j]), h2[t], s);
      }
      o3[tid] = s;
    }
    __syncthreads();
    if (tid < d) {
      float o = A.b3[tid];
      for (int part = 0; part < nparts; ++part) o += o3[part * d + tid];
      float lin = 0.0f;
      for (int i = 0; i < d; ++i) lin = fmaf(A.lin_t[(size_t)i * d + tid], xc[i], lin);
      gg[tid] = (double)(o + lin) + (-th[tid] / v);
    }
  };
  grad();
  if (tid < d) pp[tid] += 0.5 * eps * gg[tid];
  for (int l = 0; l < L; ++l) {
    if (tid < d) th[tid] += eps * pp[tid];
    grad();
    if (tid < d) pp[tid] += (l == L - 1 ? 0.5 : 1.0) * eps * gg[tid];
  }
  __syncthreads();
  if (tid == 0) {
    double k = 0.0;
    for (int j = 0; j < d; ++j) k += red[j];
    A.k0[c] = 0.5 * k;
  }
  if (tid < d) {
    A.prop[(size_t)c * A.P + tid] = th[tid];
    A.beta32[(size_t)c * A.ds + tid] = (float)th[tid];
    A.p[(size_t)c * A.ds + tid] = pp[tid];
  } else if (tid < d + 2) {
    A.prop[(size_t)c * A.P + tid] = A.state[(size_t)c * A.P + tid];  // b and v unchanged
  }
}

// Register-resident form for fixed (D, H): one 1024-thread block per chain;
// every weight is loaded once into registers as packed BF16 pairs and the
// four mat-vecs of a force evaluation run as partial dot products over input
// slices (V_DOT2_F32_BF16: two products per instruction, FP32 accumulation),
// reduced through LDS.  Network activations pass through LDS as BF16 pairs;
// the linear term stays FP32.  Thread t owns output t % O of a layer with O
// outputs and input slice t / O, so each wave reads its inputs as LDS
// broadcasts.  Threads t < D hold theta_t and p_t in registers.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

__device__ inline float dot2_bf16(uint32_t a, uint32_t b, float c) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2_t, a), __builtin_bit_cast(bf16x2_t, b), c,
                                         false);
}

// a thread's slice of NP packed BF16 activations against its weight pairs,
// 4 pairs per LDS read
template <int NP>
__device__ inline float dot_slice(const uint32_t (&w)[NP], const uint16_t* act) {
  const uint4* src = reinterpret_cast<const uint4*>(act);
  float a = 0.0f;
#pragma unroll
  for (int q = 0; q < NP / 4; ++q) {
    const uint4 u = src[q];
    a = dot2_bf16(w[4 * q], u.x, a);
    a = dot2_bf16(w[4 * q + 1], u.y, a);
    a = dot2_bf16(w[4 * q + 2], u.z, a);
    a = dot2_bf16(w[4 * q + 3], u.w, a);
  }
  return a;
}

// dot_slice with the weights in LDS, pair q at w[1024 q] (one column per
// thread, so the reads of a wave are consecutive words).
template <int NP>
__device__ inline float dot_slice_lds(const uint32_t* w, const uint16_t* act) {
  const uint4* src = reinterpret_cast<const uint4*>(act);
  float a = 0.0f;
#pragma unroll
  for (int q = 0; q < NP / 4; ++q) {
    const uint4 u = src[q];
    a = dot2_bf16(w[(4 * q) * 1024], u.x, a);
    a = dot2_bf16(w[(4 * q + 1) * 1024], u.y, a);
    a = dot2_bf16(w[(4 * q + 2) * 1024], u.z, a);
    a = dot2_bf16(w[(4 * q + 3) * 1024], u.w, a);
  }
  return a;
}

template <int D, int H>
__global__ __launch_bounds__(1024) void k_surr_traj_t(SurrTrajArgs A) {
  constexpr int P1 = 1024 / H, I1 = D / P1;  // layer 1: H outputs, D inputs
  constexpr int I2 = H / P1;                 // layer 2: H outputs, H inputs
  constexpr int P3 = 1024 / D, I3 = H / P3;  // layer 3: D outputs, H inputs
  constexpr int IL = D / P3;                 // linear term: D outputs, D inputs
  static_assert(1024 % H == 0 && 1024 % D == 0 && H % 64 == 0 && D % 64 == 0, "shape");
  static_assert(I1 % 8 == 0 && I2 % 8 == 0 && I3 % 8 == 0 && IL >= 1, "slices");
  __shared__ __align__(16) uint16_t xin[D], h1[H], h2[H];
  __shared__ __align__(16) float xc[D];
  __shared__ float r1[P1][H], r2[P1][H], r3[P3][D], rl[P3][D];
  __shared__ double kred[D];
  const int tid = threadIdx.x, c = blockIdx.x;
  const int o1 = tid % H, p1 = tid / H;
  const int o3 = tid % D, p3 = tid / D;
  uint32_t w1[I1 / 2], w3[I3 / 2];
  float wl[IL];
  extern __shared__ uint32_t w2s[];  // layer 2 weights: pair q of thread t at w2s[q * 1024 + t]
#pragma unroll
  for (int q = 0; q < I1 / 2; ++q) {
    const int i = p1 * I1 + 2 * q;
    w1[q] = (uint32_t)A.w1t[(size_t)i * H + o1] | ((uint32_t)A.w1t[(size_t)(i + 1) * H + o1] << 16);
  }
#pragma unroll
  for (int q = 0; q < I2 / 2; ++q) {
    const int i = p1 * I2 + 2 * q;
    w2s[q * 1024 + tid] = (uint32_t)A.w2t[(size_t)i * H + o1] | ((uint32_t)A.w2t[(size_t)(i + 1) * H + o1] << 16);
  }
#pragma unroll
  for (int q = 0; q < I3 / 2; ++q) {
    const int i = p3 * I3 + 2 * q;
    w3[q] = (uint32_t)A.w3t[(size_t)i * D + o3] | ((uint32_t)A.w3t[(size_t)(i + 1) * D + o3] << 16);
  }
#pragma unroll
  for (int q = 0; q < IL; ++q) wl[q] = A.lin_t[(size_t)(p3 * IL + q) * D + o3];
  const float b1v = tid < H ? A.b1[tid] : 0.0f;
  const float b2v = tid < H ? A.b2[tid] : 0.0f;
  const float b3v = tid < D ? A.b3[tid] : 0.0f;
  const double imean = tid < D ? (double)A.in_mean[tid] : 0.0;
  const float iscale = tid < D ? A.in_scale[tid] : 0.0f;
  const double lmean = tid < D ? (double)A.lin_mean[tid] : 0.0;

  const double v = A.state[(size_t)c * A.P + D + 1];
  const double eps = A.hs.step[c];
  const int L = A.hs.steps[c];
  double th = 0.0, pm = 0.0;
  if (tid < D) {
    th = A.state[(size_t)c * A.P + tid];
    pm = normal_at(Stream{A.seed, (uint32_t)c, A.sweep, A.tag}, (uint32_t)tid, 0);
    kred[tid] = pm * pm;
    xin[tid] = f_to_bf16((float)(th - imean) * iscale);
    xc[tid] = (float)(th - lmean);
  }
  __syncthreads();
  for (int s = 0; s <= L; ++s) {
    r1[p1][o1] = dot_slice(w1, xin + p1 * I1);
    {
      float b = 0.0f;
#pragma unroll
      for (int q = 0; q < IL; ++q) b = fmaf(wl[q], xc[p3 * IL + q], b);
      rl[p3][o3] = b;
    }
    __syncthreads();
    if (tid < H) {
      float z = b1v;
#pragma unroll
      for (int p = 0; p < P1; ++p) z += r1[p][tid];
      h1[tid] = f_to_bf16(gelu_s(z));
    }
    __syncthreads();
    r2[p1][o1] = dot_slice_lds<I2 / 2>(w2s + tid, h1 + p1 * I2);
    __syncthreads();
    if (tid < H) {
      float z = b2v;
#pragma unroll
      for (int p = 0; p < P1; ++p) z += r2[p][tid];
      h2[tid] = f_to_bf16(gelu_s(z));
    }
    __syncthreads();
    r3[p3][o3] = dot_slice(w3, h2 + p3 * I3);
    __syncthreads();
    if (tid < D) {
      float o = b3v;
#pragma unroll
      for (int p = 0; p < P3; ++p) o += r3[p][tid] + rl[p][tid];
      const double g = (double)o + (-th / v);
      pm += ((s == 0 || s == L) ? 0.5 : 1.0) * eps * g;
      if (s < L) {
        th += eps * pm;
        xin[tid] = f_to_bf16((float)(th - imean) * iscale);
        xc[tid] = (float)(th - lmean);
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    double k = 0.0;
    for (int j = 0; j < D; ++j) k += kred[j];
    A.k0[c] = 0.5 * k;
  }
  if (tid < D) {
    A.prop[(size_t)c * A.P + tid] = th;
    A.beta32[(size_t)c * A.ds + tid] = (float)th;
    A.p[(size_t)c * A.ds + tid] = pm;
  } else if (tid < D + 2) {
    A.prop[(size_t)c * A.P + tid] = A.state[(size_t)c * A.P + tid];  // b and v unchanged
  }
}

template <int D, int H>
static void launch_traj_t(const SurrTrajArgs& a, int chains, hipStream_t st) {
  const size_t smem = sizeof(uint32_t) * 1024 * (size_t)(H / (1024 / H) / 2);  // layer 2 weights
  allow_lds(k_surr_traj_t<D, H>, smem);
  k_surr_traj_t<D, H><<<chains, 1024, smem, st>>>(a);
}

void launch_surr_traj(const SurrTrajArgs& a, int chains, hipStream_t st) {
  if (a.d > 256 || a.h > 256) fail(AUGUR_ERR_UNSUPPORTED, "surrogate needs d, hidden <= 256");
  if (a.d == 128 && a.h == 256) launch_traj_t<128, 256>(a, chains, st);
  else if (a.d == 64 && a.h == 256) launch_traj_t<64, 256>(a, chains, st);
  else if (a.d == 64 && a.h == 128) launch_traj_t<64, 128>(a, chains, st);
  else k_surr_traj<<<chains, 256, 0, st>>>(a);
  AUGUR_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// Data log-density sum_n log sigmoid(+-(x_n . theta + b)) in FP64 arithmetic.

// log(1 + e^-x) for x >= 0, within 2.1e-16 absolute of the exact value (a
// 4M-point sweep of [0, 40] against long double).  The terms are summed, so
// absolute accuracy is what the sum needs; this form takes about 40 FP64
// operations where exp followed by the library's log1p takes about 100.
//   e^-x: -x = n ln2 + r with ln2 split in two parts (Cody & Waite), |r| <=
//     ln2 / 2, e^r by its Taylor series to r^12 (remainder < 2e-16), then
//     scaled by 2^n.
//   log1p(t), t in (0, 1]: log(1 + f) = 2 atanh(f / (2 + f)); for t above
//     sqrt2 - 1 take log 2 + log(1 + f) with f = (t - 1) / 2 instead, so that
//     |s| = |f / (2 + f)| <= 0.172 and the odd series 2 (s + s^3/3 + ...
//     + s^19/19) is complete to 1e-17.
// x above 746 (e^-x below the smallest subnormal) is clamped; NaN passes.
__device__ inline double log1p_exp_neg(double x) {
  x = x > 746.0 ? 746.0 : x;
  constexpr double kLog2e = 1.4426950408889634, kLn2Hi = 6.93147180369123816490e-01,
                   kLn2Lo = 1.90821492927058770002e-10;
  const double nn = rint(-x * kLog2e);
  double r = fma(nn, -kLn2Hi, -x);
  r = fma(nn, -kLn2Lo, r);
  double p = 1.0 / 479001600.0;
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  const double t = ldexp(p, (int)nn);
  const bool hi = t > 0.41421356237309503;
  const double f = hi ? 0.5 * (t - 1.0) : t;
  const double sr = f / (2.0 + f);
  const double s2 = sr * sr;
  double q = 1.0 / 19;
  q = fma(q, s2, 1.0 / 17);
  q = fma(q, s2, 1.0 / 15);
  q = fma(q, s2, 1.0 / 13);
  q = fma(q, s2, 1.0 / 11);
  q = fma(q, s2, 1.0 / 9);
  q = fma(q, s2, 1.0 / 7);
  q = fma(q, s2, 1.0 / 5);
  q = fma(q, s2, 1.0 / 3);
  const double l = 2.0 * fma(sr * s2, q, sr);
  return hi ? l + 0.69314718055994531 : l;
}

// log sigmoid(u) = min(u, 0) - log(1 + e^-|u|)
__device__ inline double log_sigmoid64(double u) { return fmin(u, 0.0) - log1p_exp_neg(fabs(u)); }

// Generic form: 64-row tiles of X staged in LDS, one thread per (row, chain).
__global__ __launch_bounds__(256) void k_ll64(Ll64Args A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int d = A.d, cp = A.cp, ds = d + 1;
  double* pts = (double*)smem;               // cp x ds
  float* xs = (float*)(pts + cp * ds);       // 64 x ds
  int* ys = (int*)(xs + 64 * ds);            // 64
  double* red = (double*)(((uintptr_t)(ys + 64) + 15) & ~(uintptr_t)15);  // 256
  const int tid = threadIdx.x;
  for (int i = tid; i < cp * ds; i += 256) {
    const int c = i / ds, j = i % ds;
    pts[i] = (j < d || A.has_b) ? A.point[(size_t)c * A.P + j] : 0.0;
  }
  double acc = 0.0;
  const int tiles = (A.n + 63) / 64;
  for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int r0 = tile * 64, rows = min(64, A.n - r0);
    __syncthreads();
    for (int i = tid; i < 64 * d; i += 256) xs[(i / d) * ds + i % d] = i < rows * d ? A.x[(size_t)r0 * d + i] : 0.0f;
    if (tid < 64) ys[tid] = tid < rows ? A.y[r0 + tid] : 0;
    __syncthreads();
    for (int q = tid; q < 64 * cp; q += 256) {
      const int r = q / cp, c = q % cp;
      if (r >= rows) continue;
      const float* xr = xs + r * ds;
      const double* pr = pts + c * ds;
      double t = A.has_b ? pr[d] : 0.0;
      for (int j = 0; j < d; ++j) t = fma((double)xr[j], pr[j], t);
      const double st = ys[r] ? t : -t;
      acc += log_sigmoid64(st);
    }
  }
  // thread tid's rows all belong to chain tid % cp (cp divides 256)
  __syncthreads();
  red[tid] = acc;
  __syncthreads();
  if (tid < cp) {
    double s = 0.0;
    for (int i = tid; i < 256; i += cp) s += red[i];
    A.part[(size_t)tid * A.G + blockIdx.x] = s;
  }
}

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float v4f_t __attribute__((ext_vector_type(4)));

// FP64 matrix-core form for D % 16 == 0 without intercept: each wave takes
// 16-row tiles and computes the 16 x 16 block (rows x chains) of
// s_n x_n . theta_c, s_n = +-1 from y_n, with V_MFMA_F64_16X16X4F64: lane l
// supplies A = row l % 16, k-slot l / 16 and B = chain l % 16, k-slot
// l / 16.  The feature order inside the k loop is permuted so each lane reads
// its row as float4s (features 4 (4 q + l / 16) + e at step (q, e)); the B
// operand follows the same permutation, so the dot product is unchanged.
// Output column l % 16 is chain l % 16, whichever rows the lane holds, and
// log sigmoid(s t) is summed per chain.  Rows past the last full tile are
// left to k_ll64_finish.
//
// The chain table sits in LDS (row stride D + 2 doubles: 16-byte aligned
// pairs, and the 16 chain rows of one read fall on distinct banks), which
// keeps the kernel near 100 VGPRs and four waves per SIMD; X is read once
// with nontemporal loads when it is far beyond the Infinity Cache.
template <int D>
__global__ __launch_bounds__(256) void k_ll64_mfma(Ll64Args A, int x_stream) {
  constexpr int Q = D / 16, LD = D + 2;
  __shared__ __align__(16) double th[16 * LD];
  __shared__ double red[4][16];
  for (int j = threadIdx.x; j < 16 * D; j += 256) {
    const int c = j / D, k = j - c * D;
    th[c * LD + k] = c < A.cp ? A.point[(size_t)c * A.P + k] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = lane & 15, kb = lane >> 4;
  const double2* tb = reinterpret_cast<const double2*>(th + i * LD + 4 * kb);
  double lsum = 0.0;
  const int tiles = A.n / 16;
  // tile t: rows 16 t + i, this lane's features 16 q + 4 kb .. + 3; the next
  // tile of the wave is requested before this one's products (one in flight)
  const int tstep = gridDim.x * 4;
  v4f_t xn[Q];
  int yn = 0;
  auto load = [&](int t) {
    const v4f_t* xr = reinterpret_cast<const v4f_t*>(A.x + (size_t)(t * 16 + i) * D) + kb;
    if (x_stream) {
#pragma unroll
      for (int q = 0; q < Q; ++q) xn[q] = __builtin_nontemporal_load(xr + 4 * q);
    } else {
#pragma unroll
      for (int q = 0; q < Q; ++q) xn[q] = xr[4 * q];
    }
    yn = A.y[t * 16 + i];
  };
  int t = blockIdx.x * 4 + wv;
  if (t < tiles) load(t);
  for (; t < tiles; t += tstep) {
    v4f_t xc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) xc[q] = xn[q];
    const float sg = yn ? 1.0f : -1.0f;
    if (t + tstep < tiles) load(t + tstep);
    // the table is read again on every tile: an opaque zero offset keeps the
    // compiler from hoisting the reads into 64 loop-invariant VGPRs
    int z = 0;
    asm volatile("" : "+s"(z));
    const double2* tz = tb + z;
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const v4f_t xv = xc[q];
      const double2 b01 = tz[8 * q], b23 = tz[8 * q + 1];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64((double)(sg * xv.x), b01.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64((double)(sg * xv.y), b01.y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64((double)(sg * xv.z), b23.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64((double)(sg * xv.w), b23.y, acc, 0, 0, 0);
    }
    if (A.cp <= 8) {
      // columns 8..15 hold no chain: lane i + 8 takes rows 2, 3 of lane i's
      // column, so each lane evaluates two FP64 row terms instead of four
      const double u2 = __shfl_xor(acc[2], 8), u3 = __shfl_xor(acc[3], 8);
      const double a0 = i < 8 ? acc[0] : u2, a1 = i < 8 ? acc[1] : u3;
      lsum += log_sigmoid64(a0);
      lsum += log_sigmoid64(a1);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        lsum += log_sigmoid64(acc[k]);
      }
    }
  }
  if (A.cp <= 8) lsum += __shfl_xor(lsum, 8);  // chain i: lanes i and i + 8
  lsum += __shfl_xor(lsum, 16);
  lsum += __shfl_xor(lsum, 32);
  if (lane < 16) red[wv][lane] = lsum;
  __syncthreads();
  if (threadIdx.x < A.cp)
    A.part[(size_t)threadIdx.x * A.G + blockIdx.x] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// One wave per chain: sum of the block partials, plus rows [tail, n) done
// here directly (one row at a time, features split over the lanes).
__global__ void k_ll64_finish(Ll64Args A, int tail) {
  const int c = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int g = lane; g < A.G; g += 64) s += A.part[(size_t)c * A.G + g];
  s = wave_sum(s);
  const double* pt = A.point + (size_t)c * A.P;
  for (int r = tail; r < A.n; ++r) {
    double t = 0.0;
    for (int j = lane; j < A.d; j += 64) t = fma((double)A.x[(size_t)r * A.d + j], pt[j], t);
    t = wave_sum(t) + (A.has_b ? pt[A.d] : 0.0);
    const double st = A.y[r] ? t : -t;
    s += log_sigmoid64(st);
  }
  if (lane == 0) A.out[c] = s;
}

bool ll64_mfma_form(int d, int cp, int has_b) {
  return !has_b && cp <= 16 && (d == 16 || d == 32 || d == 64 || d == 128);
}

int ll64_grid(int n, int d, int cp, int has_b, int generic_grid, int cus) {
  if (!ll64_mfma_form(d, cp, has_b)) return generic_grid;
  // four blocks of four waves per CU (four waves per SIMD), or one 16-row
  // tile per wave when there are fewer tiles
  const int tiles = n / 16;
  return std::max(1, std::min((tiles + 3) / 4, 4 * std::max(cus, 1)));
}

void launch_ll64(const Ll64Args& a, hipStream_t st) {
  int tail = a.n;
  if (ll64_mfma_form(a.d, a.cp, a.has_b)) {
    tail = (a.n / 16) * 16;
    // nontemporal X loads once X passes the 256 MB Infinity Cache
    const int xs = (size_t)a.n * a.d * sizeof(float) > ((size_t)256 << 20);
    switch (a.d) {
      case 16: k_ll64_mfma<16><<<a.G, 256, 0, st>>>(a, xs); break;
      case 32: k_ll64_mfma<32><<<a.G, 256, 0, st>>>(a, xs); break;
      case 64: k_ll64_mfma<64><<<a.G, 256, 0, st>>>(a, xs); break;
      default: k_ll64_mfma<128><<<a.G, 256, 0, st>>>(a, xs); break;
    }
  } else {
    const int ds = a.d + 1;
    size_t smem = sizeof(double) * (size_t)a.cp * ds + sizeof(float) * 64 * (size_t)ds + sizeof(int) * 64;
    smem = ((smem + 15) & ~(size_t)15) + sizeof(double) * 256;
    allow_lds(k_ll64, smem);
    k_ll64<<<a.G, 256, smem, st>>>(a);
  }
  AUGUR_CHECK_LAUNCH();
  k_ll64_finish<<<a.cp, 64, 0, st>>>(a, tail);
  AUGUR_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// Fitting.

__global__ void k_standardise(float* __restrict__ x, float* __restrict__ t, size_t count, int d,
                              const float* __restrict__ stats) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(e % d);
    x[e] = (x[e] - stats[j]) * stats[d + j];
    t[e] = (t[e] - stats[2 * d + j]) * stats[3 * d + j];
  }
}

void surr_standardise(float* x, float* t, int n, int d, const float* stats, hipStream_t st) {
  k_standardise<<<1024, 256, 0, st>>>(x, t, (size_t)n * d, d, stats);
  AUGUR_CHECK_LAUNCH();
}

// part[s][i][c] = sum over the rows of split s of x[r][i] z[r][c], z = [x | t]
// (c < 2d), FP64.  A 16 x 16 block of (i, c) per 256-thread block; rows pass
// through LDS 64 at a time.
__global__ __launch_bounds__(256) void k_gram(const float* __restrict__ x, const float* __restrict__ t, int n,
                                              int d, int rows_per_split, double* __restrict__ part) {
  __shared__ float xa[64][17], zb[64][17];
  const int i0 = blockIdx.x * 16, c0 = blockIdx.y * 16, sp = blockIdx.z;
  const int tid = threadIdx.x, ti = tid / 16, tc = tid % 16;
  const int r0 = sp * rows_per_split, r1 = min(n, r0 + rows_per_split);
  double acc = 0.0;
  for (int rb = r0; rb < r1; rb += 64) {
    for (int e = tid; e < 1024; e += 256) {
      const int rr = e / 16, cc = e % 16, r = rb + rr;
      const int ii = i0 + cc, zc = c0 + cc;
      xa[rr][cc] = (r < r1 && ii < d) ? x[(size_t)r * d + ii] : 0.0f;
      float zv = 0.0f;
      if (r < r1 && zc < 2 * d) zv = zc < d ? x[(size_t)r * d + zc] : t[(size_t)r * d + zc - d];
      zb[rr][cc] = zv;
    }
    __syncthreads();
    for (int rr = 0; rr < 64; ++rr) acc = fma((double)xa[rr][ti], (double)zb[rr][tc], acc);
    __syncthreads();
  }
  const int i = i0 + ti, cc = c0 + tc;
  if (i < d && cc < 2 * d) part[((size_t)sp * d + i) * 2 * d + cc] = acc;
}

__global__ void k_sum_splits(const double* __restrict__ part, int splits, size_t count, double* __restrict__ out) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < splits; ++k) s += part[(size_t)k * count + e];
    out[e] = s;
  }
}

// One block: Cholesky factorisation of M = g[:, :d] + ridge I in place
// (g is d x 2d, row stride 2d), then M B = g[:, d:] by forward and back
// substitution, one thread per right-hand side.  FP64.
__global__ __launch_bounds__(1024) void k_linsolve(double* __restrict__ g, int d, double ridge,
                                                  double* __restrict__ B) {
  const int tid = threadIdx.x, ld = 2 * d;
  for (int i = tid; i < d; i += 1024) g[(size_t)i * ld + i] += ridge;
  __syncthreads();
  for (int k = 0; k < d; ++k) {
    if (tid == 0) {
      const double a = g[(size_t)k * ld + k];
      g[(size_t)k * ld + k] = sqrt(a > 1e-300 ? a : 1e-300);
    }
    __syncthreads();
    const double lkk = g[(size_t)k * ld + k];
    for (int i = k + 1 + tid; i < d; i += 1024) g[(size_t)i * ld + k] /= lkk;
    __syncthreads();
    const int m = d - k - 1;
    for (int e = tid; e < m * m; e += 1024) {
      const int i = k + 1 + e / m, j = k + 1 + e % m;
      if (j <= i) g[(size_t)i * ld + j] -= g[(size_t)i * ld + k] * g[(size_t)j * ld + k];
    }
    __syncthreads();
  }
  for (int o = tid; o < d; o += 1024) {
    for (int i = 0; i < d; ++i) {
      double s = g[(size_t)i * ld + d + o];
      for (int j = 0; j < i; ++j) s -= g[(size_t)i * ld + j] * B[(size_t)j * d + o];
      B[(size_t)i * d + o] = s / g[(size_t)i * ld + i];
    }
    for (int i = d - 1; i >= 0; --i) {
      double s = B[(size_t)i * d + o];
      for (int j = i + 1; j < d; ++j) s -= g[(size_t)j * ld + i] * B[(size_t)j * d + o];
      B[(size_t)i * d + o] = s / g[(size_t)i * ld + i];
    }
  }
}

__global__ void k_f64_to_f32(const double* __restrict__ a, size_t count, float* __restrict__ b) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
    b[e] = (float)a[e];
}

__global__ void k_sub_inplace(float* __restrict__ t, const float* __restrict__ u, size_t count) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
    t[e] -= u[e];
}

void surr_linear_fit(const float* x, float* t, int n, int d, double ridge, double* B, hipStream_t st) {
  const int splits = std::max(1, std::min(64, n / 1024));
  const int rows_per_split = (n + splits - 1) / splits;
  const size_t count = (size_t)d * 2 * d;
  DevBuf<double> part((size_t)splits * count), gram(count);
  dim3 grid((d + 15) / 16, (2 * d + 15) / 16, splits);
  k_gram<<<grid, 256, 0, st>>>(x, t, n, d, rows_per_split, part.p);
  AUGUR_CHECK_LAUNCH();
  k_sum_splits<<<256, 256, 0, st>>>(part.p, splits, count, gram.p);
  AUGUR_CHECK_LAUNCH();
  k_linsolve<<<1, 1024, 0, st>>>(gram.p, d, ridge * n, B);
  AUGUR_CHECK_LAUNCH();
  // residual t -= x B, in row chunks through an FP32 GEMM
  DevBuf<float> bf((size_t)d * d);
  k_f64_to_f32<<<64, 256, 0, st>>>(B, (size_t)d * d, bf.p);
  AUGUR_CHECK_LAUNCH();
  const int chunk = 32768;
  DevBuf<float> tmp((size_t)std::min(chunk, n) * d);
  for (int r0 = 0; r0 < n; r0 += chunk) {
    const int rows = std::min(chunk, n - r0);
    gemm_f32(st, rows, d, d, x + (size_t)r0 * d, d, 1, bf.p, d, 1, tmp.p, d);
    k_sub_inplace<<<1024, 256, 0, st>>>(t + (size_t)r0 * d, tmp.p, (size_t)rows * d);
    AUGUR_CHECK_LAUNCH();
  }
  // the temporaries are freed on return: finish their users first
  AUGUR_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// Folding into the inference form.  With stats1 = [mx | ix | mg | ig] of the
// raw pairs and stats2 = [mx2 | ix2 | mr | ir] of (standardised inputs,
// residual targets):
//   g = mg + (xs B) / ig + (net(x2) / ir + mr) / ig,
//   xs = (theta - mx) ix,  x2 = (xs - mx2) ix2 = (theta - mx - mx2 / ix) ix ix2.

__global__ void k_fold_vec(int d, const float* __restrict__ s1, const float* __restrict__ s2,
                           const float* __restrict__ b3, float* __restrict__ b3e, float* __restrict__ in_mean,
                           float* __restrict__ in_scale, float* __restrict__ lin_mean) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d) return;
  const double mx = s1[j], ix = s1[d + j], mg = s1[2 * d + j], ig = s1[3 * d + j];
  const double mx2 = s2[j], ix2 = s2[d + j], mr = s2[2 * d + j], ir = s2[3 * d + j];
  in_mean[j] = (float)(mx + mx2 / ix);
  in_scale[j] = (float)(ix * ix2);
  lin_mean[j] = (float)mx;
  b3e[j] = (float)((double)b3[j] / (ir * ig) + mr / ig + mg);
}

__global__ void k_fold_lin(int d, const float* __restrict__ s1, const double* __restrict__ B,
                           float* __restrict__ lin_t) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d * d) return;
  const int i = e / d, o = e % d;
  lin_t[e] = (float)((double)s1[d + i] * B[e] / (double)s1[3 * d + o]);
}

// w: rows x cols -> wt: cols x rows in BF16; with s1, row r is first divided
// by ir[r] ig[r] (the output scale of the last layer).
__global__ void k_transpose_bf16(const float* __restrict__ w, int rows, int cols, const float* __restrict__ s1,
                                 const float* __restrict__ s2, uint16_t* __restrict__ wt) {
  const size_t n = (size_t)rows * cols;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i / cols, cc = i % cols;
    float v = w[i];
    if (s1) v = (float)((double)v / ((double)s2[3 * rows + r] * (double)s1[3 * rows + r]));
    wt[cc * rows + r] = f_to_bf16(v);
  }
}

void surr_fold(int d, int h, const float* stats1, const float* stats2, const double* B, const float* params,
               SurrInfer& out, hipStream_t st) {
  const MlpShape s{d, h};
  k_transpose_bf16<<<256, 256, 0, st>>>(params + s.w1(), h, d, nullptr, nullptr, out.w1t.p);
  k_transpose_bf16<<<256, 256, 0, st>>>(params + s.w2(), h, h, nullptr, nullptr, out.w2t.p);
  k_transpose_bf16<<<256, 256, 0, st>>>(params + s.w3(), d, h, stats1, stats2, out.w3t.p);
  AUGUR_CHECK_LAUNCH();
  k_fold_vec<<<(d + 63) / 64, 64, 0, st>>>(d, stats1, stats2, params + s.b3(), out.b3.p, out.in_mean.p,
                                           out.in_scale.p, out.lin_mean.p);
  AUGUR_CHECK_LAUNCH();
  k_fold_lin<<<(d * d + 255) / 256, 256, 0, st>>>(d, stats1, B, out.lin_t.p);
  AUGUR_CHECK_LAUNCH();
}

}  // namespace augur
