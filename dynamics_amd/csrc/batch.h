// batch.h — one launch per kernel for many independent systems.
//
// A kernel body is an ordinary __device__ function of plain arguments.  The
// wrapper kernel k_batched<F> runs it for system blockIdx.y with that
// system's argument pack, read from a device array (one pack per system),
// so a pass over n systems of uneven sizes is one launch whose grid is
// (largest per-system grid) × n.  Bodies guard their own ranges (a block
// beyond its system's work returns at once).  A single system is a batch
// of one, launched through k_single with its pack as the kernel argument.
//
// Argument packs change rarely (buffer growth, a new box or λ), so each
// launch site keeps its device array and uploads only when the packs
// differ from the last upload, from a pinned staging ring whose regions are
// reused only after their copy has completed.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "common.h"

namespace dyn {

// Plain-data tuple of kernel arguments.
template <class... T>
struct Pack;
template <>
struct Pack<> {};
template <class H, class... T>
struct Pack<H, T...> {
  H h;
  Pack<T...> t;
};

template <size_t I, class H, class... T>
__host__ __device__ constexpr const auto& pack_get(const Pack<H, T...>& p) {
  if constexpr (I == 0)
    return p.h;
  else
    return pack_get<I - 1>(p.t);
}

// Argument pack type of a body function.
template <auto F>
struct BodySig;
template <class... A, void (*F)(A...)>
struct BodySig<F> {
  using P = Pack<std::remove_cv_t<std::remove_reference_t<A>>...>;
  static constexpr size_t N = sizeof...(A);
};

template <class P>
struct PackBuild;
template <>
struct PackBuild<Pack<>> {
  static Pack<> make() { return {}; }
};
template <class H, class... T>
struct PackBuild<Pack<H, T...>> {
  template <class X0, class... Xs>
  static Pack<H, T...> make(X0&& x0, Xs&&... xs) {
    static_assert(sizeof...(Xs) == sizeof...(T), "argument count differs from the body's");
    Pack<H, T...> p;
    std::memset(static_cast<void*>(&p), 0, sizeof p);  // padding bytes compare equal
    p.h = static_cast<H>(x0);
    p.t = PackBuild<Pack<T...>>::make(std::forward<Xs>(xs)...);
    return p;
  }
};

// The argument pack of body F, converting each argument to its parameter type.
template <auto F, class... X>
typename BodySig<F>::P args(X&&... x) {
  return PackBuild<typename BodySig<F>::P>::make(std::forward<X>(x)...);
}

template <auto F, class P, size_t... I>
__device__ __forceinline__ void call_body(const P& p, std::index_sequence<I...>) {
  F(pack_get<I>(p)...);
}

// Waves per SIMD that body F asks the register allocator to leave room for
// (1: no request).  A body specialises this when latency, not issue, bounds
// it and more resident waves hide that latency.
template <auto F>
struct MinWavesPerSimd {
  static constexpr int value = 1;
};

template <auto F, int LB>
__global__ void __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(MinWavesPerSimd<F>::value)))
k_batched(const typename BodySig<F>::P* __restrict__ a) {
  call_body<F>(a[blockIdx.y], std::make_index_sequence<BodySig<F>::N>{});
}

// A single system: the pack travels as the kernel argument itself (kernarg
// segment, no dependent load at kernel start).
template <auto F, int LB>
__global__ void __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(MinWavesPerSimd<F>::value)))
k_single(const typename BodySig<F>::P a) {
  call_body<F>(a, std::make_index_sequence<BodySig<F>::N>{});
}

// Uploads of argument packs so far in this process (steady-state steps
// upload none).
inline std::atomic<uint64_t> g_arg_uploads{0};

// Raised when work cannot be recorded into a graph (the caller stops the
// capture and enqueues the work directly).
struct CaptureRefused : Error {
  explicit CaptureRefused(const std::string& m) : Error(DYN_ERR_STATE, m) {}
};

// Stream ordering between two streams (fork / join), skipped by dry runs.
inline void stream_order(hipStream_t from, hipStream_t to, hipEvent_t ev) {
  if (g_launch_mode == kLaunchDry || from == to) return;
  DYN_HIP(hipEventRecord(ev, from));
  DYN_HIP(hipStreamWaitEvent(to, ev, 0));
}

// Device copies of the argument packs of each launch site.
class ArgCache {
 public:
  ArgCache() = default;
  ArgCache(const ArgCache&) = delete;
  ArgCache& operator=(const ArgCache&) = delete;
  ~ArgCache() {
    for (auto& r : inflight_) (void)hipEventDestroy(r.ev);
    for (auto e : spare_) (void)hipEventDestroy(e);
  }

  // Device array of the packs in v (site: a string literal naming the launch
  // site, tag: a second key for sites run more than once per pass).  A site
  // keeps up to kVersions device copies (steps that alternate between packs,
  // such as the SPME cadence, find theirs uploaded); reusing a copy for other
  // packs advances the layout generation, since a captured graph may hold it.
  template <class P>
  const P* put(hipStream_t stream, const char* site, int tag, const std::vector<P>& v) {
    static_assert(std::is_trivially_copyable<P>::value, "argument packs are plain data");
    Site& st = sites_[{site, tag}];
    const size_t bytes = v.size() * sizeof(P);
    ++st.clock;
    for (Copy& c : st.copies)
      if (c.last.size() == bytes && std::memcmp(c.last.data(), v.data(), bytes) == 0) {
        c.used = st.clock;
        return reinterpret_cast<const P*>(c.dev.p);
      }
    // a graph would keep copying from a staging region later calls reuse
    if (g_launch_mode == kLaunchCapture)
      throw CaptureRefused(std::string("argument packs of ") + site + " changed during capture");
    Copy* c = nullptr;
    if (st.copies.size() < kVersions) {
      st.copies.emplace_back();
      c = &st.copies.back();
    } else {
      c = &st.copies[0];
      for (Copy& x : st.copies)
        if (x.used < c->used) c = &x;
      g_layout_gen.fetch_add(1, std::memory_order_relaxed);
    }
    c->dev.reserve(bytes);
    upload(stream, c->dev.p, v.data(), bytes);
    c->last.assign(reinterpret_cast<const char*>(v.data()),
                   reinterpret_cast<const char*>(v.data()) + bytes);
    c->used = st.clock;
    return reinterpret_cast<const P*>(c->dev.p);
  }

  // Forget every device copy (the next launches upload again).
  void invalidate() {
    for (auto& kv : sites_)
      for (Copy& c : kv.second.copies) c.last.clear();
    g_layout_gen.fetch_add(1, std::memory_order_relaxed);
  }

 private:
  // a step call uses up to ~6 packs per site (first, steady and last
  // evaluations, staged or not, times the SPME cadence)
  static constexpr size_t kVersions = 8;
  struct Copy {
    DevBuf<char> dev;
    std::vector<char> last;
    uint64_t used = 0;
  };
  struct Site {
    std::deque<Copy> copies;  // stable addresses as copies are added
    uint64_t clock = 0;
  };
  struct Region {
    size_t lo, hi;
    hipEvent_t ev;
  };
  static constexpr size_t kRing = 4u << 20;

  void upload(hipStream_t stream, char* dst, const void* src, size_t bytes) {
    g_arg_uploads.fetch_add(1, std::memory_order_relaxed);
    if (bytes > kRing) {  // larger than the ring: a plain blocking copy
      DYN_HIP(hipStreamSynchronize(stream));
      DYN_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
      return;
    }
    ring_.reserve(kRing);
    const size_t size = (bytes + 255) & ~(size_t)255;
    if (head_ + size > kRing) head_ = 0;
    const size_t lo = head_, hi = head_ + size;
    // wait for (and retire) earlier copies out of the region about to be reused
    while (!inflight_.empty()) {
      Region& r = inflight_.front();
      const bool overlap = r.lo < hi && lo < r.hi;
      const bool done = hipEventQuery(r.ev) == hipSuccess;
      if (!overlap && !done) break;
      if (!done) DYN_HIP(hipEventSynchronize(r.ev));
      spare_.push_back(r.ev);
      inflight_.pop_front();
    }
    for (const Region& r : inflight_)
      if (r.lo < hi && lo < r.hi) DYN_HIP(hipEventSynchronize(r.ev));
    std::memcpy(ring_.p + lo, src, bytes);
    DYN_HIP(hipMemcpyAsync(dst, ring_.p + lo, bytes, hipMemcpyHostToDevice, stream));
    hipEvent_t ev;
    if (!spare_.empty()) {
      ev = spare_.back();
      spare_.pop_back();
    } else {
      DYN_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    DYN_HIP(hipEventRecord(ev, stream));
    inflight_.push_back({lo, hi, ev});
    head_ = hi;
  }

  std::map<std::pair<const char*, int>, Site> sites_;
  PinnedBuf<char> ring_;
  size_t head_ = 0;
  std::deque<Region> inflight_;
  std::vector<hipEvent_t> spare_;
};

// Launch context of one batched pass: n systems on one stream.
struct BatchRun {
  hipStream_t stream = nullptr;
  ArgCache* cache = nullptr;
  uint32_t n = 0;
};

// One launch of body F over the run's systems: grid (gx, n), gx the largest
// per-system grid; fn(i) returns system i's pack (built with args<F>(...)).
template <auto F, int LB = 256, class Fn>
void launch(const BatchRun& R, const char* site, int tag, unsigned gx, unsigned block,
            size_t shm, Fn&& fn) {
  if (R.n == 0 || g_launch_mode == kLaunchDry) return;
  if (R.n == 1) {
    const typename BodySig<F>::P p = fn(0u);
    hipLaunchKernelGGL((k_single<F, LB>), dim3(std::max(gx, 1u)), dim3(block), shm, R.stream, p);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      throw Error(DYN_ERR_HIP, fmt("launch of %s: %s", site, hipGetErrorString(e)));
    return;
  }
  std::vector<typename BodySig<F>::P> v;
  v.reserve(R.n);
  for (uint32_t i = 0; i < R.n; ++i) v.push_back(fn(i));
  const auto* d = R.cache->put(R.stream, site, tag, v);
  hipLaunchKernelGGL((k_batched<F, LB>), dim3(std::max(gx, 1u), R.n), dim3(block), shm, R.stream,
                     d);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw Error(DYN_ERR_HIP, fmt("launch of %s: %s", site, hipGetErrorString(e)));
}

}  // namespace dyn
