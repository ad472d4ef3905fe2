// wave.h — wavefront-level helpers (64-lane CDNA wavefronts).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dyn {

constexpr int kWaveLanes = 64;

// A unique rank for every lane with valid set among the lanes that share its
// key: counters[key] grows by the number of such lanes.  Lanes of one key
// take consecutive ranks from one atomic, and the atomics of all the wave's
// keys are issued together (one round trip, whatever the number of keys).
// Every lane of the wave must call it.  Ranks depend on the order in which
// waves reach the counters.
__device__ __forceinline__ uint32_t wave_rank(uint32_t key, bool valid,
                                              uint32_t* __restrict__ counters) {
  const int lane = (int)(threadIdx.x & (kWaveLanes - 1));
  unsigned long long todo = __ballot(valid);
  int leader = lane;
  unsigned long long peers = 0ull;
  while (todo) {
    const int l = __ffsll((long long)todo) - 1;
    const uint32_t kl = __shfl(key, l, kWaveLanes);
    const unsigned long long p = __ballot(key == kl) & todo;
    if ((p >> lane) & 1ull) {
      leader = l;
      peers = p;
    }
    todo &= ~p;
  }
  uint32_t base = 0;
  if (peers && leader == lane) base = atomicAdd(&counters[key], (uint32_t)__popcll(peers));
  base = __shfl(base, leader, kWaveLanes);
  return base + (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
}

}  // namespace dyn
