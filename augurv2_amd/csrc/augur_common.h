// augur_common.h -- shared host/device plumbing of the sampler runtime.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/augur_hip.h"
#include "augur_rng.h"

namespace augur {

// Errors travel as exceptions inside the library and become an
// augur_status + thread-local message at the C boundary.
struct Error : std::runtime_error {
  augur_status code;
  Error(augur_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(augur_status code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

#define AUGUR_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      ::augur::fail(AUGUR_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                             \
  } while (0)

#define AUGUR_CHECK_LAUNCH() AUGUR_HIP(hipGetLastError())

void set_last_error(const char* msg);

// Runs f, converting exceptions to a status code.
template <class F>
augur_status guard(F&& f) {
  try {
    f();
    set_last_error("");
    return AUGUR_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return AUGUR_ERR_HIP;
  }
}

// Owning device buffer.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t count) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    n = count;
    if (count) AUGUR_HIP(hipMalloc((void**)&p, count * sizeof(T)));
  }
  size_t bytes() const { return n * sizeof(T); }
  void upload(const T* h, size_t count, hipStream_t s) {
    if (count > n) fail(AUGUR_ERR_ARG, "upload of %zu elements into buffer of %zu", count, n);
    if (count) AUGUR_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T* h, size_t count, hipStream_t s) const {
    if (count > n) fail(AUGUR_ERR_ARG, "download of %zu elements from buffer of %zu", count, n);
    if (count) AUGUR_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s) {
    if (n) AUGUR_HIP(hipMemsetAsync(p, 0, bytes(), s));
  }
};

void require_device();

// Dynamic LDS beyond 64 KB must be granted per kernel (gfx950 has 160 KB).
template <class K>
inline void allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536)
    AUGUR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// A parameter the generic ABI can address by name: one contiguous block of
// `elems` elements per chain, chain c at base + c * chain_stride.
struct ParamView {
  void* base = nullptr;
  size_t elem_size = 0;
  size_t elems = 0;
  size_t chain_stride = 0;  // in elements
  bool writable = true;
};

struct Trace {
  DevBuf<unsigned char> buf;
  int capacity = 0;
  int recorded = 0;
};

}  // namespace augur

// The opaque handle of the C-ABI.
struct augur_sampler {
  int device = 0;
  int chains = 1;
  uint32_t seed = 0;
  uint32_t sweep = 0;  // Philox sweep counter (mirrors the device copy)
  hipStream_t stream = nullptr;
  std::map<std::string, augur::Trace> traces;

  augur_sampler();
  virtual ~augur_sampler();
  virtual const char* family() const = 0;
  virtual augur::ParamView param(const std::string& name) = 0;
  virtual void run_sweeps(int n) = 0;
  virtual void log_density(double* out) = 0;
  virtual void stats(double* out, size_t bytes) = 0;
  // Called after a parameter is written through augur_set_pt.
  virtual void on_set(const std::string& name) = 0;
  // Called before a parameter is read (augur_cpy, traces): computes
  // deterministic parameters on the device.
  virtual void prepare(const std::string& name) { (void)name; }
  // Called after every sweep: copies traced parameters.
  void record_traces();

  // Event timing of the family's dominant kernel: every prof_period-th
  // sweep, each launch of it carries an event pair on `stream`.  prof_slot()
  // hands out the pair for hipExtLaunchKernelGGL, which records the events
  // with the kernel's own dispatch (kernel time, no gap to the previous
  // kernel); prof_begin / prof_end bracket a launch with separate records.
  int prof_period = 0;
  size_t prof_used = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_events;
  bool prof_on() const { return prof_period > 0 && (sweep % (uint32_t)prof_period) == 0; }
  std::pair<hipEvent_t, hipEvent_t>* prof_slot();
  void prof_begin();
  void prof_end();
};
