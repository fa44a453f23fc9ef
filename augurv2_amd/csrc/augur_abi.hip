// augur_abi.hip -- family-independent part of the C-ABI and the RNG kernels.
#include "augur_common.h"

namespace augur {

static thread_local std::string g_last_error;

void set_last_error(const char* msg) { g_last_error = msg; }

void require_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    fail(AUGUR_ERR_HIP, "no HIP device is visible; the sampler runtime needs an MI355X (gfx950)");
}

__global__ void k_philox(const uint32_t* __restrict__ ctr, const uint32_t* __restrict__ key, int n,
                         uint32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32x4 r = philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3],
                                key[2 * i], key[2 * i + 1]);
  out[4 * i] = r.x;
  out[4 * i + 1] = r.y;
  out[4 * i + 2] = r.z;
  out[4 * i + 3] = r.w;
}

__global__ __launch_bounds__(256) void k_gamma(Stream s, const double* __restrict__ shape, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = gamma_at(s, (uint32_t)i, shape[i]);
}

// One wave per row: Gamma(alpha_i + cnt_i) at site r * k + i, then normalise
// (cbits/augur_dist.c:526-538).  The Dirichlet operator entry point; the LDA
// sampler draws the same values in its own kernel.
__global__ __launch_bounds__(256) void k_dirichlet_rows(Stream s, const double* __restrict__ alpha,
                                 const int32_t* __restrict__ cnt, int rows, int k,
                                 double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= rows) return;
  double part = 0.0;
  for (int i = lane; i < k; i += 64) {
    const double g = gamma_at(s, (uint32_t)(r * k + i), alpha[i] + (double)cnt[(size_t)r * k + i]);
    out[(size_t)r * k + i] = g;
    part += g;
  }
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
  for (int i = lane; i < k; i += 64) out[(size_t)r * k + i] /= part;
}

}  // namespace augur

using namespace augur;

augur_sampler::augur_sampler() {}

augur_sampler::~augur_sampler() {
  for (auto& e : prof_events) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  if (stream) (void)hipStreamDestroy(stream);
}

std::pair<hipEvent_t, hipEvent_t>* augur_sampler::prof_slot() {
  if (!prof_on()) return nullptr;
  if (prof_used == prof_events.size()) {
    std::pair<hipEvent_t, hipEvent_t> e;
    AUGUR_HIP(hipEventCreate(&e.first));
    AUGUR_HIP(hipEventCreate(&e.second));
    prof_events.push_back(e);
  }
  return &prof_events[prof_used++];
}

void augur_sampler::prof_begin() {
  if (!prof_on()) return;
  if (prof_used == prof_events.size()) {
    std::pair<hipEvent_t, hipEvent_t> e;
    AUGUR_HIP(hipEventCreate(&e.first));
    AUGUR_HIP(hipEventCreate(&e.second));
    prof_events.push_back(e);
  }
  AUGUR_HIP(hipEventRecord(prof_events[prof_used].first, stream));
}

void augur_sampler::prof_end() {
  if (!prof_on()) return;
  AUGUR_HIP(hipEventRecord(prof_events[prof_used].second, stream));
  ++prof_used;
}

void augur_sampler::record_traces() {
  for (auto& kv : traces) {
    Trace& t = kv.second;
    if (t.capacity <= 0 || t.recorded >= t.capacity) continue;
    prepare(kv.first);
    ParamView v = param(kv.first);
    const size_t per_chain = v.elems * v.elem_size;
    unsigned char* dst = t.buf.p + (size_t)t.recorded * per_chain * chains;
    if (v.chain_stride == v.elems) {
      AUGUR_HIP(hipMemcpyAsync(dst, v.base, per_chain * chains, hipMemcpyDeviceToDevice, stream));
    } else {
      for (int c = 0; c < chains; ++c)
        AUGUR_HIP(hipMemcpyAsync(dst + per_chain * c,
                                 (unsigned char*)v.base + v.chain_stride * v.elem_size * c,
                                 per_chain, hipMemcpyDeviceToDevice, stream));
    }
    t.recorded++;
  }
}

extern "C" {

const char* augur_last_error(void) { return g_last_error.c_str(); }

const char* augur_version(void) { return "augurv2_amd 0.1 (gfx950)"; }

augur_status augur_device_count(int* n) {
  if (!n) return AUGUR_ERR_ARG;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return AUGUR_OK;
}

static ParamView checked_view(augur_sampler* s, const char* name, int chain, size_t bytes,
                              bool all_ok) {
  if (!s) fail(AUGUR_ERR_ARG, "null sampler");
  if (!name) fail(AUGUR_ERR_ARG, "null parameter name");
  ParamView v = s->param(name);
  if (!v.base) fail(AUGUR_ERR_ARG, "%s sampler has no parameter '%s'", s->family(), name);
  if (chain < (all_ok ? -1 : 0) || chain >= s->chains)
    fail(AUGUR_ERR_ARG, "chain %d out of range [0, %d)", chain, s->chains);
  if (bytes != v.elems * v.elem_size)
    fail(AUGUR_ERR_ARG, "parameter '%s' holds %zu bytes per chain, got %zu", name,
         v.elems * v.elem_size, bytes);
  return v;
}

augur_status augur_set_pt(augur_sampler* s, const char* name, int chain, const void* src,
                          size_t bytes) {
  return guard([&] {
    ParamView v = checked_view(s, name, chain, bytes, true);
    if (!v.writable) fail(AUGUR_ERR_ARG, "parameter '%s' is read-only", name);
    if (!src) fail(AUGUR_ERR_ARG, "null source");
    const int c0 = chain < 0 ? 0 : chain, c1 = chain < 0 ? s->chains : chain + 1;
    for (int c = c0; c < c1; ++c)
      AUGUR_HIP(hipMemcpyAsync((unsigned char*)v.base + v.chain_stride * v.elem_size * c, src,
                               bytes, hipMemcpyHostToDevice, s->stream));
    s->on_set(name);
    AUGUR_HIP(hipStreamSynchronize(s->stream));
  });
}

augur_status augur_cpy(augur_sampler* s, const char* name, int chain, void* dst, size_t bytes) {
  return guard([&] {
    ParamView v = checked_view(s, name, chain, bytes, false);
    if (!dst) fail(AUGUR_ERR_ARG, "null destination");
    s->prepare(name);
    AUGUR_HIP(hipMemcpyAsync(dst, (unsigned char*)v.base + v.chain_stride * v.elem_size * chain,
                             bytes, hipMemcpyDeviceToHost, s->stream));
    AUGUR_HIP(hipStreamSynchronize(s->stream));
  });
}

augur_status augur_step(augur_sampler* s, int n_sweeps) {
  return guard([&] {
    if (!s) fail(AUGUR_ERR_ARG, "null sampler");
    if (n_sweeps < 0) fail(AUGUR_ERR_ARG, "negative sweep count");
    AUGUR_HIP(hipSetDevice(s->device));
    s->run_sweeps(n_sweeps);
    AUGUR_HIP(hipStreamSynchronize(s->stream));
  });
}

augur_status augur_get_curr_ll(augur_sampler* s, double* out) {
  return guard([&] {
    if (!s || !out) fail(AUGUR_ERR_ARG, "null argument");
    s->log_density(out);
  });
}

augur_status augur_trace(augur_sampler* s, const char* name, int capacity) {
  return guard([&] {
    if (!s || !name) fail(AUGUR_ERR_ARG, "null argument");
    if (capacity < 0) fail(AUGUR_ERR_ARG, "negative capacity");
    ParamView v = s->param(name);
    if (!v.base) fail(AUGUR_ERR_ARG, "%s sampler has no parameter '%s'", s->family(), name);
    if (capacity == 0) {
      s->traces.erase(name);
      return;
    }
    Trace& t = s->traces[name];
    t.buf.alloc((size_t)capacity * v.elems * v.elem_size * s->chains);
    t.capacity = capacity;
    t.recorded = 0;
  });
}

augur_status augur_trace_read(augur_sampler* s, const char* name, void* dst, size_t bytes,
                              int* n_recorded) {
  return guard([&] {
    if (!s || !name) fail(AUGUR_ERR_ARG, "null argument");
    auto it = s->traces.find(name);
    if (it == s->traces.end()) fail(AUGUR_ERR_STATE, "parameter '%s' is not traced", name);
    ParamView v = s->param(name);
    const size_t need = (size_t)it->second.recorded * v.elems * v.elem_size * s->chains;
    if (n_recorded) *n_recorded = it->second.recorded;
    if (!dst) return;
    if (bytes < need) fail(AUGUR_ERR_ARG, "trace of '%s' needs %zu bytes, got %zu", name, need, bytes);
    if (need) AUGUR_HIP(hipMemcpy(dst, it->second.buf.p, need, hipMemcpyDeviceToHost));
  });
}

augur_status augur_stats(augur_sampler* s, double* out, size_t bytes) {
  return guard([&] {
    if (!s || !out) fail(AUGUR_ERR_ARG, "null argument");
    s->stats(out, bytes);
  });
}

augur_status augur_sweep_count(augur_sampler* s, uint32_t* out) {
  if (!s || !out) return AUGUR_ERR_ARG;
  *out = s->sweep;
  return AUGUR_OK;
}

augur_status augur_profile(augur_sampler* s, int period) {
  return guard([&] {
    if (!s || period < 0) fail(AUGUR_ERR_ARG, "bad arguments");
    s->prof_period = period;
    s->prof_used = 0;
  });
}

augur_status augur_profile_read(augur_sampler* s, double* avg_us, int* n_launches) {
  return guard([&] {
    if (!s || !avg_us || !n_launches) fail(AUGUR_ERR_ARG, "null argument");
    AUGUR_HIP(hipStreamSynchronize(s->stream));
    double tot = 0.0;
    for (size_t i = 0; i < s->prof_used; ++i) {
      float ms = 0.0f;
      AUGUR_HIP(hipEventElapsedTime(&ms, s->prof_events[i].first, s->prof_events[i].second));
      tot += ms;
    }
    *n_launches = (int)s->prof_used;
    *avg_us = s->prof_used ? 1000.0 * tot / (double)s->prof_used : 0.0;
    s->prof_used = 0;
  });
}

augur_status augur_destroy(augur_sampler* s) {
  return guard([&] {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    delete s;
  });
}

// Attainable-bandwidth probes: a grid-stride read of float4 with four
// independent loads in flight per thread (one float per block out, so the
// loads are not dead), and a grid-stride float4 copy.
// Plain read at the best rate found for this chip: 8 nontemporal float4 loads
// in flight per thread, four 256-thread blocks per CU (swept over 1-16 loads
// in flight, 4-32 blocks per CU, ordinary and nontemporal loads; 6.9 TB/s on
// 2 GiB against 5.3-5.7 TB/s with ordinary loads).
typedef float v4f_t __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_stream_read(const float4* __restrict__ a, size_t n4, float* __restrict__ out) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  constexpr int U = 8;
  float s[U];
#pragma unroll
  for (int u = 0; u < U; ++u) s[u] = 0.f;
  for (; i + (U - 1) * stride < n4; i += U * stride) {
    v4f_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load((const v4f_t*)(a + i + u * stride));
#pragma unroll
    for (int u = 0; u < U; ++u) s[u] += (v[u][0] + v[u][1]) + (v[u][2] + v[u][3]);
  }
  for (; i < n4; i += stride) {
    const float4 v = a[i];
    s[0] += (v.x + v.y) + (v.z + v.w);
  }
  float s0 = 0.f;
#pragma unroll
  for (int u = 0; u < U; ++u) s0 += s[u];
  __shared__ float red[256];
  red[threadIdx.x] = s0;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += stride) b[i] = a[i];
}

__global__ void k_stream_fill(float4* __restrict__ a, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += stride)
    a[i] = make_float4(1.f, 1.f, 1.f, 1.f);
}

augur_status augur_op_stream_bandwidth(size_t bytes, int reps, double* read_gbs, double* copy_gbs) {
  return guard([&] {
    require_device();
    if (bytes < 16 || reps < 1 || !read_gbs || !copy_gbs) fail(AUGUR_ERR_ARG, "bad arguments");
    const size_t n4 = bytes / 16;
    int cus = 0;
    AUGUR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    const int grid = 4 * std::max(cus, 1);
    DevBuf<float4> a(n4), b(n4);
    DevBuf<float> out(grid);
    k_stream_fill<<<grid, 256>>>(a.p, n4);
    k_stream_fill<<<grid, 256>>>(b.p, n4);
    AUGUR_CHECK_LAUNCH();
    hipEvent_t e0, e1;
    AUGUR_HIP(hipEventCreate(&e0));
    AUGUR_HIP(hipEventCreate(&e1));
    auto timed = [&](bool copy) {
      // one untimed launch, then reps timed back to back
      if (copy) k_stream_copy<<<grid, 256>>>(a.p, b.p, n4);
      else k_stream_read<<<grid, 256>>>(a.p, n4, out.p);
      AUGUR_HIP(hipEventRecord(e0, 0));
      for (int r = 0; r < reps; ++r) {
        if (copy) k_stream_copy<<<grid, 256>>>(a.p, b.p, n4);
        else k_stream_read<<<grid, 256>>>(a.p, n4, out.p);
      }
      AUGUR_HIP(hipEventRecord(e1, 0));
      AUGUR_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      AUGUR_HIP(hipEventElapsedTime(&ms, e0, e1));
      return (double)ms / reps;
    };
    const double t_read = timed(false), t_copy = timed(true);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *read_gbs = (double)(n4 * 16) / (t_read * 1e-3) / 1e9;
    *copy_gbs = 2.0 * (double)(n4 * 16) / (t_copy * 1e-3) / 1e9;
  });
}

augur_status augur_op_philox(const uint32_t* ctr, const uint32_t* key, int n, uint32_t* out) {
  return guard([&] {
    require_device();
    if (n < 0 || (n > 0 && (!ctr || !key || !out))) fail(AUGUR_ERR_ARG, "bad arguments");
    if (n == 0) return;
    DevBuf<uint32_t> dc(4 * (size_t)n), dk(2 * (size_t)n), dout(4 * (size_t)n);
    dc.upload(ctr, 4 * (size_t)n, 0);
    dk.upload(key, 2 * (size_t)n, 0);
    k_philox<<<(n + 255) / 256, 256>>>(dc.p, dk.p, n, dout.p);
    AUGUR_CHECK_LAUNCH();
    dout.download(out, 4 * (size_t)n, 0);
    AUGUR_HIP(hipDeviceSynchronize());
  });
}

augur_status augur_op_gamma(uint32_t seed, uint32_t chain, uint32_t sweep, uint32_t tag,
                            const double* shape, int n, double* out) {
  return guard([&] {
    require_device();
    if (n < 0 || (n > 0 && (!shape || !out))) fail(AUGUR_ERR_ARG, "bad arguments");
    if (n == 0) return;
    DevBuf<double> ds(n), dout(n);
    ds.upload(shape, n, 0);
    k_gamma<<<(n + 255) / 256, 256>>>(Stream{seed, chain, sweep, tag}, ds.p, n, dout.p);
    AUGUR_CHECK_LAUNCH();
    dout.download(out, n, 0);
    AUGUR_HIP(hipDeviceSynchronize());
  });
}

augur_status augur_op_dirichlet_conj(uint32_t seed, uint32_t chain, uint32_t sweep,
                                     uint32_t tag, const double* alpha, const int32_t* cnt,
                                     int rows, int k, double* out) {
  return guard([&] {
    require_device();
    if (rows < 0 || k <= 0 || !alpha || !cnt || !out) fail(AUGUR_ERR_ARG, "bad arguments");
    if (rows == 0) return;
    DevBuf<double> da(k), dout((size_t)rows * k);
    DevBuf<int32_t> dc((size_t)rows * k);
    da.upload(alpha, k, 0);
    dc.upload(cnt, (size_t)rows * k, 0);
    k_dirichlet_rows<<<(rows + 3) / 4, 256>>>(Stream{seed, chain, sweep, tag}, da.p, dc.p, rows, k,
                                              dout.p);
    AUGUR_CHECK_LAUNCH();
    dout.download(out, (size_t)rows * k, 0);
    AUGUR_HIP(hipDeviceSynchronize());
  });
}

}  // extern "C"
