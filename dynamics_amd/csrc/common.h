// common.h — host-side plumbing shared by the library's translation units:
// error capture, RAII device buffers, launch-shape helpers and the optional
// per-kernel event profiler.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <atomic>
#include <chrono>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dyn_nb.h"

namespace dyn {

// Every C entry point catches dyn::Error and returns its code; the message is
// stored on the handle (or globally for create calls).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline std::string fmt(const char* f, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

#define DYN_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::dyn::Error(e_ == hipErrorOutOfMemory ? DYN_ERR_NOMEM : DYN_ERR_HIP,  \
                         ::dyn::fmt("%s failed: %s (%s:%d)", #call,                \
                                    hipGetErrorString(e_), __FILE__, __LINE__));   \
  } while (0)

#define DYN_CHECK_LAUNCH(name)                                                     \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess)                                                          \
      throw ::dyn::Error(DYN_ERR_HIP, ::dyn::fmt("launch of %s failed: %s", name,  \
                                                 hipGetErrorString(e_)));          \
  } while (0)

#define DYN_REQUIRE(cond, ...)                                                     \
  do {                                                                             \
    if (!(cond)) throw ::dyn::Error(DYN_ERR_ARG, ::dyn::fmt(__VA_ARGS__));         \
  } while (0)

// Device memory layout generation: every release of a device buffer (also
// the one that precedes a larger allocation) advances it, so a captured
// graph (whose kernels hold buffer addresses) knows whether its addresses
// are still those of live buffers.  A new buffer leaves it: no graph can
// hold an address that did not exist.
inline std::atomic<uint64_t> g_layout_gen{1};
// Why the last step-graph capture was refused (diagnostics).
inline std::string g_capture_refusal;

// Device buffer owned by a handle.  Size in elements; zero-size buffers hold
// no allocation (the pointer is null and no kernel may touch it).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      g_layout_gen.fetch_add(1, std::memory_order_relaxed);
    }
    p = nullptr;
    n = 0;
  }
  // Grow-only: keeps the allocation when it is already large enough.
  void reserve(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) return;
    DYN_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  size_t bytes() const { return n * sizeof(T); }
};

template <typename T>
struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  void reserve(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) return;
    DYN_HIP(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault));
    n = count;
  }
};

inline unsigned grid_for(size_t n, unsigned block, unsigned max_blocks = 65535u * 4u) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return (unsigned)g;
}

// What enqueueing does with work on this thread: kLaunchDirect launches it;
// kLaunchCapture launches it into a stream capture (a graph's recording);
// kLaunchDry runs only the host side of the enqueueing code (its
// bookkeeping) while a graph holding the same work is launched instead.
enum LaunchMode { kLaunchDirect = 0, kLaunchDry = 1, kLaunchCapture = 2 };
inline thread_local int g_launch_mode = kLaunchDirect;

// Per-kernel timing with HIP events on the handle's stream.  Off by default;
// when on, every launch wrapped in Profiler::Scope records a start/stop pair,
// and read() synchronizes and sums elapsed times per kernel name.  Scopes
// recorded into a graph become event nodes of the graph: their records go to
// capture_sink (kept with the graph), and every launch of the graph adds them
// to the pending ones again (their events then hold that launch's times,
// read once the work has completed).
struct Profiler {
  bool enabled = false;
  const char* only = nullptr;  // when set, only scopes of this name are timed
  hipStream_t stream = nullptr;
  struct Rec {
    hipEvent_t a, b;
    const char* name;
    double host_ms;  // host time spent issuing the scope's work
    bool keep = false;  // events owned by a graph (not returned to the pool)
  };
  std::vector<Rec> pending;
  std::vector<Rec>* capture_sink = nullptr;
  std::vector<hipEvent_t> graph_events;  // events recorded into graphs
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, uint64_t>> totals;
  std::vector<std::string> order;

  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    DYN_HIP(hipEventCreate(&e));
    return e;
  }
  struct Scope {
    Profiler* p;
    Rec r;
    std::chrono::steady_clock::time_point t0;
    // inside a capture the record becomes an event node of the graph (an
    // external record: each launch of the graph records the event)
    void record(hipEvent_t e) {
      if (g_launch_mode == kLaunchCapture)
        DYN_HIP(hipEventRecordWithFlags(e, p->stream, hipEventRecordExternal));
      else
        DYN_HIP(hipEventRecord(e, p->stream));
    }
    Scope(Profiler& prof, const char* name)
        : p(prof.enabled && g_launch_mode != kLaunchDry &&
                    (!prof.only || std::strcmp(prof.only, name) == 0)
                ? &prof
                : nullptr) {
      if (p) {
        r.name = name;
        r.a = p->ev();
        r.b = p->ev();
        record(r.a);
        t0 = std::chrono::steady_clock::now();
      }
    }
    ~Scope() noexcept(false) {
      if (p) {
        r.host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        record(r.b);
        if (g_launch_mode == kLaunchCapture && p->capture_sink) {
          r.keep = true;  // a graph holds the events: they stay until the profiler goes
          p->graph_events.push_back(r.a);
          p->graph_events.push_back(r.b);
          p->capture_sink->push_back(r);
        } else {
          p->pending.push_back(r);
          if (p->pending.size() > 4096) p->drain();
        }
      }
    }
  };
  void drain() {
    if (pending.empty()) return;
    DYN_HIP(hipStreamSynchronize(stream));
    for (auto& r : pending) {
      float ms = 0.f;
      const hipError_t te = hipEventElapsedTime(&ms, r.a, r.b);
      if (te != hipSuccess) {
        // events of a graph that the runtime cannot time: left out
        (void)hipGetLastError();
        if (r.keep) continue;
        DYN_HIP(te);
      }
      auto it = totals.find(r.name);
      if (it == totals.end()) {
        totals[r.name] = {0.0, 0};
        order.push_back(r.name);
      }
      totals[r.name].first += ms;
      totals[r.name].second += 1;
      const std::string hn = std::string("host:") + r.name;
      if (totals.find(hn) == totals.end()) {
        totals[hn] = {0.0, 0};
        order.push_back(hn);
      }
      totals[hn].first += r.host_ms;
      totals[hn].second += 1;
      if (!r.keep) {
        pool.push_back(r.a);
        pool.push_back(r.b);
      }
    }
    pending.clear();
  }
  // A host-side duration (ms) under `name` (when profiling).
  void add_host(const char* name, double ms) {
    if (!enabled) return;
    const std::string hn = std::string("host:") + name;
    if (totals.find(hn) == totals.end()) {
      totals[hn] = {0.0, 0};
      order.push_back(hn);
    }
    totals[hn].first += ms;
    totals[hn].second += 1;
  }
  void reset() {
    drain();
    totals.clear();
    order.clear();
  }
  ~Profiler() {
    for (auto& r : pending)
      if (!r.keep) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
      }
    for (auto e : pool) (void)hipEventDestroy(e);
    for (auto e : graph_events) (void)hipEventDestroy(e);
  }
};

// Makes `device` current for the lifetime of the guard.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    DYN_HIP(hipGetDevice(&prev));
    if (prev != dev) DYN_HIP(hipSetDevice(dev));
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

void set_global_error(const std::string& m);

}  // namespace dyn

// Wraps the body of a C entry point: converts exceptions into status codes
// and stores the message on `holder->err` (a std::string member).
#define DYN_API_BODY(holder, ...)                                        \
  try {                                                                  \
    __VA_ARGS__;                                                         \
    return DYN_OK;                                                       \
  } catch (const ::dyn::Error& e) {                                      \
    if (holder) (holder)->err = e.what(); else ::dyn::set_global_error(e.what()); \
    return e.code;                                                       \
  } catch (const std::exception& e) {                                    \
    if (holder) (holder)->err = e.what(); else ::dyn::set_global_error(e.what()); \
    return DYN_ERR_HIP;                                                  \
  }
