// augur_gen.hip -- samplers for models outside the hand-written families.
//
// This takes the place of the code the reference's compiler generates for
// an arbitrary model: the data pass of CudaC/CgCuda.hs:371-516, together
// with Core/CgLike.hs:90-267 and Core/CgGrad.hs:151-376.  The model's
// data pass arrives as HIP source, generated by the Python lowering
// (augurv2_amd/pyaugur/lower.py).  It is compiled at run time with hipRTC
// for the device's architecture and driven by device-resident HMC
// (cbits/augur_mcmclib.h:343-424).
//
// The pass kernel is augur_gen_pass(q, part, R, I, Z).  It runs on a grid of
// (G, chains) blocks of 256 threads and evaluates every chain at its
// vector q[c * P .. + P).  Each block writes one row of P + 1 FP64
// partials: d ll / dq, then ll.
//
// Per HMC block:
//   k_gen_begin   momentum draw, first half step, first position step
//   pass + k_gen_finish, once per leapfrog step.  The finish sums the G
//                 partial rows of each value in a fixed order and applies
//                 the leapfrog update of every coordinate that the block
//                 moves.
//   k_gen_end     kinetic energies, FP64 Metropolis test, accept / reject.
// A DiscGibbs block runs its generated kernel (q, Z, R, I, seed, sweep, tag),
// which redraws every point of one discrete parameter in Z.
// Each HMC block runs the pass of its own full conditional when the source
// provides one (else the log joint's).  The log density and gradient at the
// current point are cached: they are what the pass computed when that point
// was proposed.  An iteration of L leapfrog steps therefore costs L passes,
// plus one at the block's start when the previous block used another pass.
#include <hip/hiprtc.h>

#include <algorithm>
#include <mutex>

#include "augur_hmc.h"

namespace augur {

constexpr int kGenThreads = 256;
constexpr int kGenMaxBlocks = 16;
constexpr int kGenMaxGrid = 4096;

enum { GEN_CACHE = 0, GEN_LEAPFROG = 1, GEN_EVAL = 2 };

struct GenArgs {
  int P, G, mode, l;
  const double* part;          // [chains][G][P + 1]
  const unsigned char* mask;   // P: 1 where this block moves the coordinate
  HmcStep hs;
  uint32_t seed, sweep, tag;
  double* q;                   // chains x P: the current point
  double* prop;                // chains x P: the trajectory's position
  double* p0;                  // chains x P: initial momentum
  double* p;                   // chains x P: momentum
  double* k0;                  // chains
  double* gcur;                // chains x P: gradient at q
  double* llcur;               // chains
  double* gprop;               // chains x P: gradient at the trajectory's end
  double* llprop;              // chains
  double* stats;               // chains x 2 for this block
};

__device__ inline double block_sum_256(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = red[0];
#pragma unroll
  for (int i = 1; i < kGenThreads / 64; ++i) s += red[i];
  __syncthreads();
  return s;
}

// One thread per value (chain c, index j <= P): the G partial rows are
// summed in a fixed order (8 loads in flight, then a fixed tree).
__global__ __launch_bounds__(kGenThreads) void k_gen_finish(GenArgs A) {
  const int P = A.P, vs = P + 1;
  const int v = blockIdx.x * kGenThreads + threadIdx.x;
  const int c = blockIdx.y;
  if (v >= vs) return;
  const double* col = A.part + (size_t)c * A.G * vs + v;
  double s = 0.0;
  for (int g0 = 0; g0 < A.G; g0 += 8) {
    double t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = g0 + k < A.G ? col[(size_t)(g0 + k) * vs] : 0.0;
    s += ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
  }
  const int j = v;
  if (A.mode == GEN_CACHE || A.mode == GEN_EVAL) {
    double* g = A.mode == GEN_CACHE ? A.gcur : A.gprop;
    double* ll = A.mode == GEN_CACHE ? A.llcur : A.llprop;
    if (j < P) g[(size_t)c * P + j] = s;
    else ll[c] = s;
    return;
  }
  const int L = A.hs.steps[c];
  if (A.l > L) return;
  if (j == P) {
    if (A.l == L) A.llprop[c] = s;
    return;
  }
  const size_t q = (size_t)c * P + j;
  if (A.l == L) A.gprop[q] = s;
  if (!A.mask[j]) return;
  const double step = A.hs.step[c];
  const double pj = A.p[q] + ((A.l == L) ? 0.5 * step : step) * s;
  A.p[q] = pj;
  if (A.l < L) A.prop[q] += step * pj;
}

// One block per chain: momentum N(0, 1) per moved coordinate j (site j of
// the block's momentum stream), p = p0 + step / 2 g(q), prop = q + step p.
__global__ __launch_bounds__(kGenThreads) void k_gen_begin(GenArgs A) {
  __shared__ double red[kGenThreads / 64];
  const int c = blockIdx.x, P = A.P;
  const Stream mom{A.seed, (uint32_t)c, A.sweep, A.tag + 0};
  const double step = A.hs.step[c];
  double kin = 0.0;
  for (int j = threadIdx.x; j < P; j += kGenThreads) {
    const size_t q = (size_t)c * P + j;
    double x = A.q[q];
    if (A.mask[j]) {
      const double p0 = normal_at(mom, (uint32_t)j, 0);
      const double pj = p0 + 0.5 * step * A.gcur[q];
      A.p0[q] = p0;
      A.p[q] = pj;
      kin += p0 * p0;
      x += step * pj;
    }
    A.prop[q] = x;
  }
  kin = block_sum_256(kin, red);
  if (threadIdx.x == 0) A.k0[c] = 0.5 * kin;
}

// One block per chain: log u < (ll(prop) - ll(q)) + (K0 - K1)
// (cbits/augur_mcmclib.h:70-83, 410-423); on acceptance the proposal's
// log density and gradient become the cached ones.
__global__ __launch_bounds__(kGenThreads) void k_gen_end(GenArgs A) {
  __shared__ double red[kGenThreads / 64];
  __shared__ int acc_s;
  const int c = blockIdx.x, P = A.P;
  double kin = 0.0;
  for (int j = threadIdx.x; j < P; j += kGenThreads)
    if (A.mask[j]) {
      const double pj = A.p[(size_t)c * P + j];
      kin += pj * pj;
    }
  kin = block_sum_256(kin, red);
  if (threadIdx.x == 0) {
    const Stream acc_stream{A.seed, (uint32_t)c, A.sweep, A.tag + 2};
    const double u = u01(acc_stream.draw(0, 0).x);
    const bool acc = log(u) < (A.llprop[c] - A.llcur[c]) + (A.k0[c] - 0.5 * kin);
    acc_s = acc ? 1 : 0;
    if (acc) A.llcur[c] = A.llprop[c];
    A.stats[2 * c] += acc ? 1.0 : 0.0;
    A.stats[2 * c + 1] += 1.0;
  }
  __syncthreads();
  if (!acc_s) return;
  for (int j = threadIdx.x; j < P; j += kGenThreads) {
    const size_t q = (size_t)c * P + j;
    A.q[q] = A.prop[q];
    A.gcur[q] = A.gprop[q];
  }
}

// ---- hipRTC ------------------------------------------------------------------

// Compiled code objects by source text: a model compiled twice in one
// process is built once.
static std::mutex g_rtc_mutex;
static std::map<std::string, std::vector<char>> g_rtc_cache;

static std::vector<char> rtc_compile(const std::string& src, const std::string& arch) {
  const std::string key = arch + "\n" + src;
  {
    std::lock_guard<std::mutex> lk(g_rtc_mutex);
    auto it = g_rtc_cache.find(key);
    if (it != g_rtc_cache.end()) return it->second;
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "augur_gen_model.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    fail(AUGUR_ERR_HIP, "hiprtcCreateProgram failed");
  const std::string arch_opt = "--offload-arch=" + arch;
  // packed FP32 VALU ops stay off, as in the library's own build
  const char* opts[] = {arch_opt.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"};
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    fail(AUGUR_ERR_ARG, "hipRTC could not compile the model's pass: %s\n%.400s", hiprtcGetErrorString(r),
         log.c_str());
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  std::vector<char> code(n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  std::lock_guard<std::mutex> lk(g_rtc_mutex);
  g_rtc_cache[key] = code;
  return code;
}

// ---- the sampler ---------------------------------------------------------------

struct GenBlock {
  double sim_len, eps;
  int fixed_steps;
  int kernel;  // index into GenSampler::passes
  int kind;    // AUGUR_GEN_HMC or AUGUR_GEN_DISC
};

// One compiled pass: the log joint (passes[0]) or a block's full conditional.
struct GenPass {
  hipFunction_t fn = nullptr;
  int grid = 1;
};

struct GenSampler final : augur_sampler {
  int P = 0;
  hipModule_t mod = nullptr;
  std::vector<GenPass> passes;
  DevBuf<double> q, prop, p0, p, k0, gcur, llcur, gprop, llprop, part, acc, R;
  DevBuf<int32_t> I, Z;
  int NZ = 1;  // discrete coordinates per chain
  DevBuf<unsigned char> masks;
  std::vector<GenBlock> blocks;
  std::map<std::string, std::pair<int, int>> names;   // continuous: name -> (offset, size) in q
  std::map<std::string, std::pair<int, int>> znames;  // discrete: name -> (offset, size) in Z
  // the pass whose ll and gradient at q are in llcur / gcur (-1: none);
  // a block's HMC keeps its own pass's cache valid, since an accepted
  // proposal brings the values computed for it
  int cache_owner = -1;

  ~GenSampler() override {
    if (mod) (void)hipModuleUnload(mod);
  }
  const char* family() const override { return "gen"; }

  ParamView param(const std::string& name) override {
    ParamView v;
    if (name == "state") return ParamView{q.p, sizeof(double), (size_t)P, (size_t)P, true};
    if (name == "zstate") return ParamView{Z.p, sizeof(int32_t), (size_t)NZ, (size_t)NZ, true};
    auto it = names.find(name);
    if (it != names.end())
      return ParamView{q.p + it->second.first, sizeof(double), (size_t)it->second.second, (size_t)P, true};
    auto zt = znames.find(name);
    if (zt != znames.end())
      return ParamView{Z.p + zt->second.first, sizeof(int32_t), (size_t)zt->second.second, (size_t)NZ, true};
    return v;
  }

  void on_set(const std::string&) override { cache_owner = -1; }

  GenArgs args(int mode, int k) {
    GenArgs A{};
    A.P = P;
    A.G = passes[k].grid;
    A.mode = mode;
    A.part = part.p;
    A.seed = seed;
    A.sweep = sweep;
    A.q = q.p; A.prop = prop.p; A.p0 = p0.p; A.p = p.p; A.k0 = k0.p;
    A.gcur = gcur.p; A.llcur = llcur.p; A.gprop = gprop.p; A.llprop = llprop.p;
    return A;
  }

  void pass(int k, const double* at) {
    const double* qa = at;
    double* pa = part.p;
    const double* ra = R.p;
    const int32_t* ia = I.p;
    const int32_t* za = Z.p;
    void* kargs[] = {(void*)&qa, (void*)&pa, (void*)&ra, (void*)&ia, (void*)&za};
    prof_begin();
    AUGUR_HIP(hipModuleLaunchKernel(passes[k].fn, passes[k].grid, chains, 1, kGenThreads, 1, 1, 0, stream,
                                    kargs, nullptr));
    prof_end();
  }

  void finish(const GenArgs& A) {
    k_gen_finish<<<dim3((P + 1 + kGenThreads - 1) / kGenThreads, chains), kGenThreads, 0, stream>>>(A);
    AUGUR_CHECK_LAUNCH();
  }

  void ensure_cache(int k) {
    if (cache_owner == k) return;
    pass(k, q.p);
    finish(args(GEN_CACHE, k));
    cache_owner = k;
  }

  void hmc(int b) {
    const GenBlock& B = blocks[b];
    ensure_cache(B.kernel);
    GenArgs A = args(GEN_LEAPFROG, B.kernel);
    A.tag = 16u + 4u * (uint32_t)b;
    A.mask = masks.p + (size_t)b * P;
    A.stats = acc.p + (size_t)b * 2 * chains;
    int lmax = 1;
    for (int c = 0; c < chains; ++c) {
      const Stream js{seed, (uint32_t)c, sweep, A.tag + 1};
      hmc_jitter(js, B.sim_len, B.eps, B.fixed_steps, &A.hs.step[c], &A.hs.steps[c]);
      lmax = std::max(lmax, A.hs.steps[c]);
    }
    k_gen_begin<<<chains, kGenThreads, 0, stream>>>(A);
    AUGUR_CHECK_LAUNCH();
    for (int l = 1; l <= lmax; ++l) {
      pass(B.kernel, prop.p);
      A.l = l;
      finish(A);
    }
    k_gen_end<<<chains, kGenThreads, 0, stream>>>(A);
    AUGUR_CHECK_LAUNCH();
  }

  // DiscGibbs: every point of the discrete parameter drawn from its full
  // conditional by the block's generated kernel; Z changes, so no pass's
  // cache stays valid
  void disc(int b) {
    const GenBlock& B = blocks[b];
    const double* qa = q.p;
    int32_t* za = Z.p;
    const double* ra = R.p;
    const int32_t* ia = I.p;
    uint32_t sd = seed, sw = sweep, tg = 16u + 4u * (uint32_t)b;
    void* kargs[] = {(void*)&qa, (void*)&za, (void*)&ra, (void*)&ia, (void*)&sd, (void*)&sw, (void*)&tg};
    AUGUR_HIP(hipModuleLaunchKernel(passes[B.kernel].fn, passes[B.kernel].grid, chains, 1, kGenThreads, 1, 1,
                                    0, stream, kargs, nullptr));
    cache_owner = -1;
  }

  void run_sweeps(int n) override {
    for (int i = 0; i < n; ++i) {
      for (int b = 0; b < (int)blocks.size(); ++b) {
        if (blocks[b].kind == AUGUR_GEN_DISC) disc(b);
        else hmc(b);
      }
      ++sweep;
      record_traces();
    }
    AUGUR_HIP(hipStreamSynchronize(stream));
  }

  // the log joint at the current point
  void log_density(double* out) override {
    if (cache_owner == 0) {
      llcur.download(out, chains, stream);
    } else {
      pass(0, q.p);
      finish(args(GEN_EVAL, 0));
      llprop.download(out, chains, stream);
    }
    AUGUR_HIP(hipStreamSynchronize(stream));
  }

  // per HMC block b and chain c: accepted, proposed
  void stats(double* out, size_t bytes) override {
    const size_t need = sizeof(double) * blocks.size() * chains * 2;
    if (bytes < need) fail(AUGUR_ERR_ARG, "stats need %zu bytes", need);
    if (need) acc.download(out, need / sizeof(double), stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
  }

  // log joint and its gradient at caller-given points (the pass and its
  // reduction, without touching the sampler's state)
  void eval(const double* hq, double* hll, double* hgrad) {
    prop.upload(hq, (size_t)chains * P, stream);
    pass(0, prop.p);
    finish(args(GEN_EVAL, 0));
    llprop.download(hll, chains, stream);
    gprop.download(hgrad, (size_t)chains * P, stream);
    AUGUR_HIP(hipStreamSynchronize(stream));
  }
};

static void gen_validate(const augur_gen_desc* D) {
  if (!D) fail(AUGUR_ERR_ARG, "null descriptor");
  if (!D->source || !D->kernel) fail(AUGUR_ERR_ARG, "the pass source and kernel name are required");
  if (D->n_params < 1) fail(AUGUR_ERR_ARG, "n_params must be positive");
  if (D->grid < 1 || D->grid > kGenMaxGrid) fail(AUGUR_ERR_ARG, "grid must be in [1, %d]", kGenMaxGrid);
  if (D->chains < 1 || D->chains > kMaxChains) fail(AUGUR_ERR_ARG, "chains must be in [1, %d]", kMaxChains);
  if (D->n_blocks < 0 || D->n_blocks > kGenMaxBlocks)
    fail(AUGUR_ERR_ARG, "n_blocks must be in [0, %d]", kGenMaxBlocks);
  if (D->n_blocks && (!D->masks || !D->sim_len || !D->step_size || !D->fixed_steps))
    fail(AUGUR_ERR_ARG, "HMC blocks need masks, sim_len, step_size and fixed_steps");
  for (int b = 0; b < D->n_blocks; ++b)
    if (!(D->step_size[b] > 0.0) || !(D->sim_len[b] >= 0.0))
      fail(AUGUR_ERR_ARG, "block %d: step size must be positive and sim_len non-negative", b);
  for (int b = 0; b < D->n_blocks; ++b)
    if (D->block_kinds && D->block_kinds[b] != AUGUR_GEN_HMC && D->block_kinds[b] != AUGUR_GEN_DISC)
      fail(AUGUR_ERR_ARG, "block %d: kind must be AUGUR_GEN_HMC or AUGUR_GEN_DISC", b);
  for (int b = 0; b < D->n_blocks; ++b)
    if (D->block_kinds && D->block_kinds[b] == AUGUR_GEN_DISC && !D->block_kernels)
      fail(AUGUR_ERR_ARG, "block %d: a DiscGibbs block needs its kernel", b);
  if (D->n_zints < 0 || (D->n_znames && (!D->znames || !D->zoffsets || !D->zsizes)))
    fail(AUGUR_ERR_ARG, "discrete parameter names need offsets and sizes");
  for (int i = 0; i < D->n_znames; ++i)
    if (!D->znames[i] || D->zoffsets[i] < 0 || D->zsizes[i] < 1 || D->zoffsets[i] + D->zsizes[i] > D->n_zints)
      fail(AUGUR_ERR_ARG, "discrete parameter %d lies outside the vector of %d", i, D->n_zints);
  if (D->block_kernels) {
    if (!D->block_grids) fail(AUGUR_ERR_ARG, "block kernels need their grids");
    for (int b = 0; b < D->n_blocks; ++b)
      if (!D->block_kernels[b] || D->block_grids[b] < 1 || D->block_grids[b] > kGenMaxGrid)
        fail(AUGUR_ERR_ARG, "block %d: kernel name and a grid in [1, %d] are required", b, kGenMaxGrid);
  }
  if (D->n_names < 0 || (D->n_names && (!D->names || !D->offsets || !D->sizes)))
    fail(AUGUR_ERR_ARG, "parameter names need offsets and sizes");
  for (int i = 0; i < D->n_names; ++i)
    if (!D->names[i] || D->offsets[i] < 0 || D->sizes[i] < 1 || D->offsets[i] + D->sizes[i] > D->n_params)
      fail(AUGUR_ERR_ARG, "parameter %d lies outside the vector of %d", i, D->n_params);
  if ((D->n_reals && !D->reals) || (D->n_ints && !D->ints)) fail(AUGUR_ERR_ARG, "null data buffer");
}

}  // namespace augur

using namespace augur;

extern "C" {

augur_status augur_gen_create(const augur_gen_desc* D, augur_sampler** out) {
  return guard([&] {
    if (!out) fail(AUGUR_ERR_ARG, "null output handle");
    *out = nullptr;
    gen_validate(D);
    require_device();
    auto* s = new GenSampler();
    try {
      AUGUR_HIP(hipGetDevice(&s->device));
      hipDeviceProp_t prop;
      AUGUR_HIP(hipGetDeviceProperties(&prop, s->device));
      std::string arch = prop.gcnArchName;
      arch = arch.substr(0, arch.find(':'));  // gfx950:sramecc+:xnack- -> gfx950
      const std::vector<char> code = rtc_compile(D->source, arch);
      AUGUR_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
      AUGUR_HIP(hipModuleLoadData(&s->mod, code.data()));
      GenPass joint;
      AUGUR_HIP(hipModuleGetFunction(&joint.fn, s->mod, D->kernel));
      joint.grid = D->grid;
      s->passes.push_back(joint);
      int max_grid = D->grid;
      for (int b = 0; b < D->n_blocks; ++b) {
        int k = 0;
        if (D->block_kernels) {
          GenPass bp;
          AUGUR_HIP(hipModuleGetFunction(&bp.fn, s->mod, D->block_kernels[b]));
          bp.grid = D->block_grids[b];
          max_grid = std::max(max_grid, bp.grid);
          s->passes.push_back(bp);
          k = (int)s->passes.size() - 1;
        }
        s->blocks.push_back({D->sim_len[b], D->step_size[b], D->fixed_steps[b], k,
                             D->block_kinds ? D->block_kinds[b] : AUGUR_GEN_HMC});
      }
      s->seed = D->seed;
      s->chains = D->chains;
      s->P = D->n_params;
      const size_t C = (size_t)D->chains, P = (size_t)D->n_params;
      for (auto* b : {&s->q, &s->prop, &s->p0, &s->p, &s->gcur, &s->gprop}) {
        b->alloc(C * P);
        b->zero(s->stream);
      }
      for (auto* b : {&s->k0, &s->llcur, &s->llprop}) b->alloc(C);
      s->part.alloc(C * (size_t)max_grid * (P + 1));
      s->acc.alloc(std::max<size_t>(1, 2 * C * (size_t)D->n_blocks));
      s->acc.zero(s->stream);
      s->R.alloc(std::max<size_t>(1, D->n_reals));
      s->I.alloc(std::max<size_t>(1, D->n_ints));
      s->R.upload(D->reals, D->n_reals, s->stream);
      s->I.upload(D->ints, D->n_ints, s->stream);
      s->masks.alloc(std::max<size_t>(1, (size_t)D->n_blocks * P));
      if (D->n_blocks) s->masks.upload(D->masks, (size_t)D->n_blocks * P, s->stream);
      for (int i = 0; i < D->n_names; ++i) s->names[D->names[i]] = {D->offsets[i], D->sizes[i]};
      s->NZ = std::max(1, D->n_zints);
      s->Z.alloc(C * (size_t)s->NZ);
      s->Z.zero(s->stream);
      for (int i = 0; i < D->n_znames; ++i) s->znames[D->znames[i]] = {D->zoffsets[i], D->zsizes[i]};
      AUGUR_HIP(hipStreamSynchronize(s->stream));
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

augur_status augur_gen_eval(augur_sampler* s, const double* q, double* ll, double* grad) {
  return guard([&] {
    auto* g = dynamic_cast<GenSampler*>(s);
    if (!g) fail(AUGUR_ERR_ARG, "not a generic-model sampler");
    if (!q || !ll || !grad) fail(AUGUR_ERR_ARG, "null buffer");
    g->eval(q, ll, grad);
  });
}

}  // extern "C"
