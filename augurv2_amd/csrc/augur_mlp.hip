// augur_mlp.hip -- training and batched evaluation of the gradient-surrogate
// MLP: a BF16 matrix-core GEMM (FP32 accumulation) with fused bias / GELU
// epilogues for the network, a strided tiled FP32 GEMM for the linear
// term's residual, GELU backward, column sums, MSE loss and Adam, all
// hand-written.
#include "augur_mlp.h"

namespace augur {

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)

__device__ inline float gelu_f(float z) {
  return 0.5f * z * (1.0f + tanhf(kGeluC * (z + 0.044715f * z * z * z)));
}
__device__ inline float gelu_grad_f(float z) {
  const float u = kGeluC * (z + 0.044715f * z * z * z);
  const float t = tanhf(u);
  const float du = kGeluC * (1.0f + 3.0f * 0.044715f * z * z);
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * du;
}

struct GemmArgs {
  int M, N, K;
  const float* A;
  long sam, sak;  // A(m, k) = A[m * sam + k * sak]
  const float* B;
  long sbk, sbn;  // B(k, n) = B[k * sbk + n * sbn]
  float* C;
  int ldc;
  const float* bias;  // [N] or null
  float* zout;        // pre-activation copy when act != 0
  int act;            // 0: none, 1: GELU
};

// C = A B (+ bias, GELU): 64 x 64 tile per 256-thread block, 4 x 4 per
// thread, k in steps of 16 through LDS.  Fixed summation order.
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ float As[16][64 + 4];
  __shared__ float Bs[16][64 + 4];
  const int tm = blockIdx.y * 64, tn = blockIdx.x * 64;
  const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    for (int e = tid; e < 1024; e += 256) {
      const int mm = e / 16, kk = e % 16;
      const int m = tm + mm, k = k0 + kk;
      As[kk][mm] = (m < g.M && k < g.K) ? g.A[m * g.sam + k * g.sak] : 0.0f;
      const int kb = e / 64, nb = e % 64;
      const int k2 = k0 + kb, n = tn + nb;
      Bs[kb][nb] = (k2 < g.K && n < g.N) ? g.B[k2 * g.sbk + n * g.sbn] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = tm + ty * 4 + i, n = tn + tx * 4 + j;
      if (m >= g.M || n >= g.N) continue;
      float v = acc[i][j] + (g.bias ? g.bias[n] : 0.0f);
      if (g.act) {
        g.zout[(size_t)m * g.ldc + n] = v;
        v = gelu_f(v);
      }
      g.C[(size_t)m * g.ldc + n] = v;
    }
}

static void gemm(hipStream_t st, int M, int N, int K, const float* A, long sam, long sak, const float* B,
                 long sbk, long sbn, float* Cm, int ldc, const float* bias = nullptr, float* zout = nullptr,
                 int act = 0) {
  GemmArgs g{M, N, K, A, sam, sak, B, sbk, sbn, Cm, ldc, bias, zout, act};
  k_gemm<<<dim3((N + 63) / 64, (M + 63) / 64), 256, 0, st>>>(g);
  AUGUR_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// The network's training GEMMs on the matrix cores: C = A B (+ bias, GELU)
// with both operands rounded to BF16 (nearest even) as they are staged and
// FP32 accumulation in V_MFMA_F32_16X16X32_BF16.  A 256-thread block owns a
// 64 x 64 tile of C, each wave a 32 x 32 quarter (2 x 2 products of 16 x 16);
// k advances 32 at a time through LDS.  Operand layouts of the product (lane
// l, i = l % 16, g = l / 16): A supplies row i, k = 8 g .. 8 g + 7; B column
// i, the same k; the result is C[4 g + r][i] in register r.  Long-k products
// with few tiles (the weight gradients, k = batch rows) split k over
// blockIdx.z: each slice writes its partial tile and k_gemm_splitk_sum adds
// the slices in order, so the result does not depend on scheduling.
typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));
constexpr int kMmTile = 64, kMmK = 32, kMmLd = kMmK + 8;  // LDS row: 40 BF16 = 80 bytes

__device__ inline float gemm_epilogue(const GemmArgs& g, int m, int n, float v) {
  v += g.bias ? g.bias[n] : 0.0f;
  if (g.act) {
    g.zout[(size_t)m * g.ldc + n] = v;
    v = gelu_f(v);
  }
  return v;
}

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g, int kper, float* __restrict__ part) {
  __shared__ __align__(16) __bf16 As[kMmTile][kMmLd];  // [m][k]
  __shared__ __align__(16) __bf16 Bs[kMmTile][kMmLd];  // [n][k]
  const int tm = blockIdx.y * kMmTile, tn = blockIdx.x * kMmTile;
  const int k_beg = blockIdx.z * kper, k_end = min(g.K, k_beg + kper);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int li = lane & 15, kg = lane >> 4;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  // staging maps: along k when k is the contiguous index, else along m / n
  const bool a_k = g.sak == 1 || g.sam != 1;
  const bool b_k = g.sbk == 1 || g.sbn != 1;
  mlp_f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = mlp_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = k_beg; k0 < k_end; k0 += kMmK) {
    if (a_k) {  // thread: row m, k = k0 + 8 q .. + 7
      const int m = tid >> 2, q = (tid & 3) * 8;
      mlp_bf16x8 v;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = k0 + q + e;
        v[e] = (__bf16)((tm + m < g.M && k < k_end) ? g.A[(size_t)(tm + m) * g.sam + (size_t)k * g.sak] : 0.0f);
      }
      *(mlp_bf16x8*)&As[m][q] = v;
    } else {  // thread: k = k0 + kk, rows m0 .. m0 + 7 (contiguous in memory)
      const int kk = tid >> 3, m0 = (tid & 7) * 8, k = k0 + kk;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int m = tm + m0 + e;
        As[m0 + e][kk] = (__bf16)((m < g.M && k < k_end) ? g.A[(size_t)m * g.sam + (size_t)k * g.sak] : 0.0f);
      }
    }
    if (b_k) {  // thread: column n, k = k0 + 8 q .. + 7
      const int n = tid >> 2, q = (tid & 3) * 8;
      mlp_bf16x8 v;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = k0 + q + e;
        v[e] = (__bf16)((tn + n < g.N && k < k_end) ? g.B[(size_t)k * g.sbk + (size_t)(tn + n) * g.sbn] : 0.0f);
      }
      *(mlp_bf16x8*)&Bs[n][q] = v;
    } else {  // thread: k = k0 + kk, columns n0 .. n0 + 7 (contiguous in memory)
      const int kk = tid >> 3, n0 = (tid & 7) * 8, k = k0 + kk;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int n = tn + n0 + e;
        Bs[n0 + e][kk] = (__bf16)((n < g.N && k < k_end) ? g.B[(size_t)k * g.sbk + (size_t)n * g.sbn] : 0.0f);
      }
    }
    __syncthreads();
    mlp_bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *(const mlp_bf16x8*)&As[wm + 16 * i + li][8 * kg];
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *(const mlp_bf16x8*)&Bs[wn + 16 * j + li][8 * kg];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  const bool split = gridDim.z > 1;
  float* pz = split ? part + (size_t)blockIdx.z * g.M * g.N : nullptr;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = tm + wm + 16 * i + 4 * kg + r, n = tn + wn + 16 * j + li;
        if (m >= g.M || n >= g.N) continue;
        if (split) pz[(size_t)m * g.N + n] = acc[i][j][r];
        else g.C[(size_t)m * g.ldc + n] = gemm_epilogue(g, m, n, acc[i][j][r]);
      }
}

__global__ void k_gemm_splitk_sum(GemmArgs g, int slices, const float* __restrict__ part) {
  const size_t mn = (size_t)g.M * g.N;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < mn; e += (size_t)gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int z = 0; z < slices; ++z) s += part[(size_t)z * mn + e];
    const int m = (int)(e / g.N), n = (int)(e % g.N);
    g.C[(size_t)m * g.ldc + n] = gemm_epilogue(g, m, n, s);
  }
}

// Slices of k for a product of M x N x K: enough blocks to cover the chip
// when the tiles alone do not, at least 256 of k per slice, and the partial
// tiles within `cap` floats.
static int splitk_slices(int M, int N, int K, size_t cap) {
  const int tiles = ((M + kMmTile - 1) / kMmTile) * ((N + kMmTile - 1) / kMmTile);
  int s = 1;
  while (tiles * s < 256 && K / (2 * s) >= 256 && (size_t)(2 * s) * M * N <= cap) s *= 2;
  return s;
}

void Mlp::mm(hipStream_t st, int M, int N, int K, const float* A, long sam, long sak, const float* B, long sbk,
             long sbn, float* Cm, int ldc, const float* bias, float* zout, int act) {
  GemmArgs g{M, N, K, A, sam, sak, B, sbk, sbn, Cm, ldc, bias, zout, act};
  const int slices = splitk_slices(M, N, K, kpart.n);
  const int kper = ((K + slices - 1) / slices + kMmK - 1) / kMmK * kMmK;
  const dim3 grid((N + kMmTile - 1) / kMmTile, (M + kMmTile - 1) / kMmTile, slices);
  k_gemm_bf16<<<grid, 256, 0, st>>>(g, kper, kpart.p);
  AUGUR_CHECK_LAUNCH();
  if (slices > 1) {
    k_gemm_splitk_sum<<<std::min<size_t>(1024, ((size_t)M * N + 255) / 256), 256, 0, st>>>(g, slices, kpart.p);
    AUGUR_CHECK_LAUNCH();
  }
}

void gemm_f32(hipStream_t st, int M, int N, int K, const float* A, long sam, long sak, const float* B,
              long sbk, long sbn, float* C, int ldc) {
  gemm(st, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc);
}

// out[n] = sum over rows of x[m][n]
// 1024-thread block per 64 columns: thread (group g, column c) sums rows
// g, g + 16, ... with four independent accumulators, then the 16 group sums
// are added in order (deterministic).
__global__ __launch_bounds__(1024) void k_colsum(const float* __restrict__ x, int rows, int cols,
                                                 float* __restrict__ out) {
  __shared__ float red[16][65];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
  if (c < cols) {
    int m = g;
    for (; m + 48 < rows; m += 64) {
      s0 += x[(size_t)m * cols + c];
      s1 += x[(size_t)(m + 16) * cols + c];
      s2 += x[(size_t)(m + 32) * cols + c];
      s3 += x[(size_t)(m + 48) * cols + c];
    }
    for (; m < rows; m += 16) s0 += x[(size_t)m * cols + c];
  }
  red[g][threadIdx.x & 63] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (threadIdx.x < 64 && c < cols) {
    float s = 0.0f;
    for (int k = 0; k < 16; ++k) s += red[k][threadIdx.x];
    out[c] = s;
  }
}

__global__ void k_gelu_bwd(const float* __restrict__ dh, const float* __restrict__ z, float* __restrict__ dz,
                           size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dz[i] = dh[i] * gelu_grad_f(z[i]);
}

// dout = 2 (out - t) / n and loss = mean (out - t)^2: kMseBlocks blocks write
// partial sums (fixed tree in each), k_mse_fin adds them in block order.
constexpr int kMseBlocks = 128;

__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ out, const float* __restrict__ t,
                                             float* __restrict__ dout, size_t n, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)kMseBlocks * 256) {
    const float df = out[i] - t[i];
    dout[i] = 2.0f * df / (float)n;
    s += (double)df * (double)df;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_mse_fin(const double* __restrict__ part, size_t n, double* loss) {
  double s = 0.0;
  for (int k = 0; k < kMseBlocks; ++k) s += part[k];
  *loss = s / (double)n;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, size_t n, float lr, float c1, float c2) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float mi = 0.9f * m[i] + 0.1f * g[i];
    const float vi = 0.999f * v[i] + 0.001f * g[i] * g[i];
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
  }
}

void Mlp::init(int d, int h, int batch_rows, uint32_t seed, hipStream_t st) {
  s = MlpShape{d, h};
  batch = batch_rows;
  const size_t P = s.count();
  params.alloc(P);
  grads.alloc(P);
  m1.alloc(P);
  m2.alloc(P);
  z1.alloc((size_t)batch * h); h1.alloc((size_t)batch * h);
  z2.alloc((size_t)batch * h); h2.alloc((size_t)batch * h);
  out.alloc((size_t)batch * d); dout.alloc((size_t)batch * d);
  dh.alloc((size_t)batch * h); dz.alloc((size_t)batch * h);
  // split-k partial tiles of the weight gradients: up to 16 slices of the
  // largest of them (H x H, D x H)
  kpart.alloc(16 * (size_t)std::max(h * h, d * h));
  loss.alloc(1 + kMseBlocks);  // [0] the loss, then the blocks' partial sums
  // Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights from Philox, zero biases
  std::vector<float> hp(P, 0.0f);
  auto fill = [&](size_t off, int rows, int fan_in, uint32_t tag) {
    const Stream rs{seed, 0u, 0u, tag};
    const float sc = 1.0f / sqrtf((float)fan_in);
    for (size_t i = 0; i < (size_t)rows * fan_in; ++i)
      hp[off + i] = sc * (float)(2.0 * u01(rs.draw((uint32_t)i, 0).x) - 1.0);
  };
  fill(s.w1(), h, d, 101u);
  fill(s.w2(), h, h, 102u);
  fill(s.w3(), d, h, 103u);
  params.upload(hp.data(), P, st);
  m1.zero(st);
  m2.zero(st);
  AUGUR_HIP(hipStreamSynchronize(st));  // hp is released on return
  adam_t = 0;
}

void Mlp::forward(const float* x, int rows, hipStream_t st) {
  const int d = s.d, h = s.h;
  const float* P = params.p;
  mm(st, rows, h, d, x, d, 1, P + s.w1(), 1, d, h1.p, h, P + s.b1(), z1.p, 1);
  mm(st, rows, h, h, h1.p, h, 1, P + s.w2(), 1, h, h2.p, h, P + s.b2(), z2.p, 1);
  mm(st, rows, d, h, h2.p, h, 1, P + s.w3(), 1, h, out.p, d, P + s.b3());
}

void Mlp::loss_grad(const float* x, const float* t, int rows, hipStream_t st) {
  const int d = s.d, h = s.h;
  forward(x, rows, st);
  const float* P = params.p;
  float* Gp = grads.p;
  k_mse<<<kMseBlocks, 256, 0, st>>>(out.p, t, dout.p, (size_t)rows * d, loss.p + 1);
  k_mse_fin<<<1, 1, 0, st>>>(loss.p + 1, (size_t)rows * d, loss.p);
  AUGUR_CHECK_LAUNCH();
  // layer 3: dW3 = dout^T h2, db3 = colsum(dout), dh2 = dout W3
  mm(st, d, h, rows, dout.p, 1, d, h2.p, h, 1, Gp + s.w3(), h);
  k_colsum<<<(d + 63) / 64, 1024, 0, st>>>(dout.p, rows, d, Gp + s.b3());
  mm(st, rows, h, d, dout.p, d, 1, P + s.w3(), h, 1, dh.p, h);
  k_gelu_bwd<<<256, 256, 0, st>>>(dh.p, z2.p, dz.p, (size_t)rows * h);
  // layer 2
  mm(st, h, h, rows, dz.p, 1, h, h1.p, h, 1, Gp + s.w2(), h);
  k_colsum<<<(h + 63) / 64, 1024, 0, st>>>(dz.p, rows, h, Gp + s.b2());
  mm(st, rows, h, h, dz.p, h, 1, P + s.w2(), h, 1, dh.p, h);
  k_gelu_bwd<<<256, 256, 0, st>>>(dh.p, z1.p, dz.p, (size_t)rows * h);
  // layer 1
  mm(st, h, d, rows, dz.p, 1, h, x, d, 1, Gp + s.w1(), d);
  k_colsum<<<(h + 63) / 64, 1024, 0, st>>>(dz.p, rows, h, Gp + s.b1());
  AUGUR_CHECK_LAUNCH();
}

void Mlp::adam(double lr, hipStream_t st) {
  ++adam_t;
  const float c1 = (float)(1.0 / (1.0 - pow(0.9, adam_t)));
  const float c2 = (float)(1.0 / (1.0 - pow(0.999, adam_t)));
  k_adam<<<512, 256, 0, st>>>(params.p, grads.p, m1.p, m2.p, s.count(), (float)lr, c1, c2);
  AUGUR_CHECK_LAUNCH();
}

// Minibatch rows drawn uniformly (Philox, step `step`) from n pairs,
// standardised on the way: xb = (x - mx) * ix, tb = (t - mt) * it.
__global__ void k_gather_std(const float* __restrict__ x, const float* __restrict__ t, int n, int d,
                             const float* __restrict__ stats, int rows, uint32_t seed, uint32_t step,
                             float* __restrict__ xb, float* __restrict__ tb) {
  const int r = blockIdx.x;
  if (r >= rows) return;
  const Stream s{seed, 0u, step, 900u};
  const uint32_t w = s.draw((uint32_t)r, 0).x;
  const int src = (int)(((uint64_t)w * (uint64_t)n) >> 32);
  const float* mx = stats;
  const float* ix = stats + d;
  const float* mt = stats + 2 * d;
  const float* it = stats + 3 * d;
  for (int j = threadIdx.x; j < d; j += blockDim.x) {
    xb[(size_t)r * d + j] = (x[(size_t)src * d + j] - mx[j]) * ix[j];
    tb[(size_t)r * d + j] = (t[(size_t)src * d + j] - mt[j]) * it[j];
  }
}

// Column means and inverse standard deviations of x and t (n x d each):
// stats = [mean x | 1/sd x | mean t | 1/sd t], FP64 sums.  Block (column
// group, row split): 256 threads = 4 row groups x 64 columns over the rows of
// its split; partial sums per split, added in split order by the finish.
constexpr int kColSplits = 64;

__global__ __launch_bounds__(256) void k_colstats_part(const float* __restrict__ x, const float* __restrict__ t,
                                                       int n, int d, double* __restrict__ part) {
  __shared__ double red[4][64][2];
  const int j = blockIdx.x * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
  const int per = (n + kColSplits - 1) / kColSplits;
  const int r0 = blockIdx.y * per, r1 = min(n, r0 + per);
  double s = 0.0, s2 = 0.0;
  if (j < 2 * d) {
    const float* a = j < d ? x : t;
    const int col = j % d;
    for (int i = r0 + g; i < r1; i += 4) {
      const double v = a[(size_t)i * d + col];
      s += v;
      s2 += v * v;
    }
  }
  red[g][threadIdx.x & 63][0] = s;
  red[g][threadIdx.x & 63][1] = s2;
  __syncthreads();
  if (threadIdx.x < 64 && j < 2 * d) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < 4; ++k) {
      a += red[k][threadIdx.x][0];
      b += red[k][threadIdx.x][1];
    }
    part[((size_t)blockIdx.y * 2 * d + j) * 2] = a;
    part[((size_t)blockIdx.y * 2 * d + j) * 2 + 1] = b;
  }
}

__global__ void k_colstats_fin(const double* __restrict__ part, int n, int d, float* __restrict__ stats) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= 2 * d) return;
  double s = 0.0, s2 = 0.0;
  for (int k = 0; k < kColSplits; ++k) {
    s += part[((size_t)k * 2 * d + j) * 2];
    s2 += part[((size_t)k * 2 * d + j) * 2 + 1];
  }
  const int col = j % d;
  const double mean = s / n;
  const double var = fmax(s2 / n - mean * mean, 1e-30);
  const int base = j < d ? 0 : 2 * d;
  stats[base + col] = (float)mean;
  stats[base + d + col] = (float)(1.0 / sqrt(var));
}

void mlp_colstats(const float* x, const float* t, int n, int d, float* stats, hipStream_t st) {
  DevBuf<double> part((size_t)kColSplits * 2 * d * 2);
  k_colstats_part<<<dim3((2 * d + 63) / 64, kColSplits), 256, 0, st>>>(x, t, n, d, part.p);
  AUGUR_CHECK_LAUNCH();
  k_colstats_fin<<<(2 * d + 63) / 64, 64, 0, st>>>(part.p, n, d, stats);
  AUGUR_CHECK_LAUNCH();
  AUGUR_HIP(hipStreamSynchronize(st));  // part is released on return
}

void mlp_gather(const float* x, const float* t, int n, int d, const float* stats, int rows, uint32_t seed,
                uint32_t step, float* xb, float* tb, hipStream_t st) {
  k_gather_std<<<rows, 128, 0, st>>>(x, t, n, d, stats, rows, seed, step, xb, tb);
  AUGUR_CHECK_LAUNCH();
}

}  // namespace augur

using namespace augur;

extern "C" {

augur_status augur_op_mlp_loss_grad(int d, int h, int rows, const float* params, const float* x,
                                    const float* t, float* out, double* loss, float* grads) {
  return guard([&] {
    if (d < 1 || h < 1 || rows < 1 || !params || !x || !t) fail(AUGUR_ERR_ARG, "bad arguments");
    require_device();
    hipStream_t st;
    AUGUR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    try {
      Mlp m;
      m.init(d, h, rows, 0, st);
      m.params.upload(params, m.s.count(), st);
      DevBuf<float> dx((size_t)rows * d), dt((size_t)rows * d);
      dx.upload(x, (size_t)rows * d, st);
      dt.upload(t, (size_t)rows * d, st);
      m.loss_grad(dx.p, dt.p, rows, st);
      if (out) m.out.download(out, (size_t)rows * d, st);
      if (loss) m.loss.download(loss, 1, st);
      if (grads) m.grads.download(grads, m.s.count(), st);
      AUGUR_HIP(hipStreamSynchronize(st));
    } catch (...) {
      (void)hipStreamDestroy(st);
      throw;
    }
    AUGUR_HIP(hipStreamDestroy(st));
  });
}

augur_status augur_op_mlp_train(int d, int h, int n, const float* x, const float* t, int steps,
                                int batch, double lr, uint32_t seed, float* params_inout,
                                double* loss_hist) {
  return guard([&] {
    if (d < 1 || h < 1 || n < 1 || steps < 0 || batch < 1 || !x || !t || !params_inout)
      fail(AUGUR_ERR_ARG, "bad arguments");
    require_device();
    hipStream_t st;
    AUGUR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    try {
      Mlp m;
      m.init(d, h, batch, seed, st);
      m.params.upload(params_inout, m.s.count(), st);
      DevBuf<float> dx((size_t)n * d), dt((size_t)n * d), xb((size_t)batch * d), tb((size_t)batch * d),
          stats(4 * (size_t)d);
      dx.upload(x, (size_t)n * d, st);
      dt.upload(t, (size_t)n * d, st);
      // identity standardisation: the caller passes standardised data
      std::vector<float> hs(4 * (size_t)d);
      for (int j = 0; j < d; ++j) {
        hs[j] = 0.0f; hs[d + j] = 1.0f; hs[2 * d + j] = 0.0f; hs[3 * d + j] = 1.0f;
      }
      stats.upload(hs.data(), hs.size(), st);
      for (int s = 0; s < steps; ++s) {
        k_gather_std<<<batch, 128, 0, st>>>(dx.p, dt.p, n, d, stats.p, batch, seed, (uint32_t)s, xb.p, tb.p);
        AUGUR_CHECK_LAUNCH();
        m.loss_grad(xb.p, tb.p, batch, st);
        if (loss_hist) AUGUR_HIP(hipMemcpyAsync(loss_hist + s, m.loss.p, sizeof(double), hipMemcpyDeviceToHost, st));
        m.adam(lr, st);
      }
      m.params.download(params_inout, m.s.count(), st);
      AUGUR_HIP(hipStreamSynchronize(st));
    } catch (...) {
      (void)hipStreamDestroy(st);
      throw;
    }
    AUGUR_HIP(hipStreamDestroy(st));
  });
}

}  // extern "C"
