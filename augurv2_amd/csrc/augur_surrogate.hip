// augur_surrogate.hip -- surrogate leapfrog trajectories and the FP64 data
// log-density used by their Metropolis test.
#include "augur_surrogate.h"

#include "augur_hmc.h"
#include "augur_mlp.h"

namespace augur {

__device__ inline float bf16_to_f(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }

__device__ inline uint16_t f_to_bf16(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline float gelu_s(float z) {
  return 0.5f * z * (1.0f + tanhf(0.7978845608028654f * (z + 0.044715f * z * z * z)));
}

// One block per chain runs a whole surrogate trajectory: momentum draw (the
// exact HMC's momentum stream), leapfrog with force = MLP(theta) (data term,
// de-standardised) - theta / v (prior term, exact), in the reference's order
// (cbits/augur_mcmclib.h:343-375).  Positions and momenta in FP64; the MLP
// in FP32 with BF16 weights.
__global__ __launch_bounds__(256) void k_surr_traj(SurrTrajArgs A) {
  __shared__ double th[256], pp[256], gg[256], red[256];
  __shared__ float xin[256], h1[256], h2[256], o3[256];
  const int c = blockIdx.x, tid = threadIdx.x, d = A.d, h = A.h;
  const double v = A.state[(size_t)c * A.P + d + 1];
  const float* mx = A.stats;
  const float* ix = mx + d;
  const float* mg = mx + 2 * d;
  const float* ig = mx + 3 * d;
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
    if (tid < d) xin[tid] = ((float)th[tid] - mx[tid]) * ix[tid];
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
      gg[tid] = (double)(o / ig[tid] + mg[tid]) + (-th[tid] / v);
    }
  };
  grad();
  if (tid < d) pp[tid] += 0.5 * A.step * gg[tid];
  for (int l = 0; l < A.steps; ++l) {
    if (tid < d) th[tid] += A.step * pp[tid];
    grad();
    if (tid < d) pp[tid] += (l == A.steps - 1 ? 0.5 : 1.0) * A.step * gg[tid];
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

void launch_surr_traj(const SurrTrajArgs& a, int chains, hipStream_t st) {
  if (a.d > 256 || a.h > 256) fail(AUGUR_ERR_UNSUPPORTED, "surrogate needs d, hidden <= 256");
  k_surr_traj<<<chains, 256, 0, st>>>(a);
  AUGUR_CHECK_LAUNCH();
}

// Data log-density sum_n log sigmoid(+-(x_n . theta + b)) with FP64 row
// arithmetic: 64-row tiles of X staged in LDS, one thread per (row, chain).
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
      acc += fmin(st, 0.0) - log1p(exp(-fabs(t)));
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

__global__ void k_ll64_finish(Ll64Args A) {
  const int c = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int g = lane; g < A.G; g += 64) s += A.part[(size_t)c * A.G + g];
  s = wave_sum(s);
  if (lane == 0) A.out[c] = s;
}

void launch_ll64(const Ll64Args& a, hipStream_t st) {
  const int ds = a.d + 1;
  size_t smem = sizeof(double) * (size_t)a.cp * ds + sizeof(float) * 64 * (size_t)ds + sizeof(int) * 64;
  smem = ((smem + 15) & ~(size_t)15) + sizeof(double) * 256;
  allow_lds(k_ll64, smem);
  k_ll64<<<a.G, 256, smem, st>>>(a);
  AUGUR_CHECK_LAUNCH();
  k_ll64_finish<<<a.cp, 64, 0, st>>>(a);
  AUGUR_CHECK_LAUNCH();
}

__global__ void k_transpose_bf16(const float* __restrict__ w, int rows, int cols, uint16_t* __restrict__ wt) {
  // w: rows x cols -> wt: cols x rows
  const size_t n = (size_t)rows * cols;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i / cols, cc = i % cols;
    wt[cc * rows + r] = f_to_bf16(w[i]);
  }
}

void make_inference_weights(const float* params, int d, int h, uint16_t* w1t, uint16_t* w2t, uint16_t* w3t,
                            hipStream_t st) {
  const MlpShape s{d, h};
  k_transpose_bf16<<<256, 256, 0, st>>>(params + s.w1(), h, d, w1t);
  k_transpose_bf16<<<256, 256, 0, st>>>(params + s.w2(), h, h, w2t);
  k_transpose_bf16<<<256, 256, 0, st>>>(params + s.w3(), d, h, w3t);
  AUGUR_CHECK_LAUNCH();
}

}  // namespace augur
