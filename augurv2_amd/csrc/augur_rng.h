// augur_rng.h -- counter-based RNG for the sampler kernels (gfx950 + host).
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Every sampling site derives its randomness from
//   key     = { seed, chain }
//   counter = { site, sub, sweep, tag }
// so draws are independent of launch geometry and of the order in which
// threads run, and a chain can be resumed exactly from (seed, sweep).
// This replaces the reference's 30000 curand XORWOW states indexed by the
// outer loop index (cbits/augur_dist.c:27-47, CudaC/CgCuda.hs:116-122) and the
// host GSL generator (cbits/augur_mcmclib.h:137-141).
//
// Conversions:
//   u01(w)  = (w + 0.5) * 2^-32          double in (0, 1)
//   u01f(w) = (w >> 8) * 2^-24           float in [0, 1)
//   normal  = sqrt(-2 ln u01(w0)) * cos(2 pi u01(w1))   (Box-Muller)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace augur {

enum : uint32_t {
  TAG_LDA_THETA = 1,
  TAG_LDA_PHI = 2,
  TAG_LDA_Z = 3,
  TAG_GMM_PI = 4,
  TAG_GMM_MU = 5,
  TAG_GMM_COV = 6,
  TAG_GMM_Z = 7,
  TAG_GMM_MU_MOVE = 8,  // ESlice / HMC / MWG / RSlice on mu: 8 normals, 9 jitter, 10 uniforms
  TAG_HMC_BASE = 16,  // HMC block b: 16 + 4 b + {0 momentum, 1 jitter, 2 accept}
};

constexpr int kGammaMaxAttempts = 256;
constexpr double kLargePosNum = 1e8;  // cbits/augur_math.h:24

struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  return {c0, c1, c2, c3};
}

struct Stream {
  uint32_t seed, chain, sweep, tag;
  __host__ __device__ u32x4 draw(uint32_t site, uint32_t sub) const {
    return philox4x32_10(site, sub, sweep, tag, seed, chain);
  }
};

__host__ __device__ inline double u01(uint32_t w) {
  return ((double)w + 0.5) * 2.3283064365386963e-10;
}

__host__ __device__ inline float u01f(uint32_t w) {
  return (float)(w >> 8) * 5.9604644775390625e-8f;
}

__host__ __device__ inline double normal_from(uint32_t w0, uint32_t w1) {
  const double u1 = u01(w0), u2 = u01(w1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

__host__ __device__ inline double normal_at(const Stream& s, uint32_t site, uint32_t sub) {
  const u32x4 w = s.draw(site, sub);
  return normal_from(w.x, w.y);
}

// One Marsaglia-Tsang attempt (counter sub = j): the candidate d v, the
// word that feeds the shape < 1 boost, and whether it is accepted.
__host__ __device__ inline bool mt_attempt(const Stream& s, uint32_t site, uint32_t j, double d, double c,
                                           double* x, uint32_t* boost_word) {
  const u32x4 w = s.draw(site, j);
  const double z = normal_from(w.x, w.y);
  const double u = u01(w.z);
  const double t = 1.0 + c * z;
  const double v = t * t * t;
  *x = d * v;
  *boost_word = w.w;
  return z > -1.0 / c && log(u) < 0.5 * z * z + d - d * v + d * log(v);
}

// Gamma(shape, 1) by Marsaglia-Tsang (cbits/augur_dist.c:287-312); attempt j
// at a site uses counter sub = j and the first accepted attempt is returned.
// Attempts are evaluated in pairs (2i, 2i + 1): the two are independent, so
// a wave with one thread per site overlaps their latencies, and in a wave of
// 64 sites some thread almost always needs the second one.  The result is
// the same as taking the attempts one at a time.  Returns NaN for an invalid
// shape or after kGammaMaxAttempts rejections (acceptance is >= 95 % per
// attempt, so the cap only bounds the loop for non-finite input).
__host__ __device__ inline double gamma_at(const Stream& s, uint32_t site, double shape) {
  static_assert(kGammaMaxAttempts % 2 == 0, "attempts are taken in pairs");
  if (!(shape > 0.0) || !(shape < 1.0e300)) return __builtin_nan("");
  const double a = shape < 1.0 ? shape + 1.0 : shape;
  const double d = a - 1.0 / 3.0;
  const double c = 1.0 / sqrt(9.0 * d);
  for (int j = 0; j < kGammaMaxAttempts; j += 2) {
    double x0, x1;
    uint32_t b0, b1;
    const bool ok0 = mt_attempt(s, site, (uint32_t)j, d, c, &x0, &b0);
    const bool ok1 = mt_attempt(s, site, (uint32_t)j + 1u, d, c, &x1, &b1);
    if (ok0 || ok1) {
      double x = ok0 ? x0 : x1;
      if (shape < 1.0) x *= pow(u01(ok0 ? b0 : b1), 1.0 / shape);
      return x;
    }
  }
  return __builtin_nan("");
}

}  // namespace augur
