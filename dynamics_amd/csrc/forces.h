// forces.h — the end of a force evaluation, shared by the engine's finalize
// kernel and the MD step kernel that can take its place: each site's force
// as the sum of its parts, and the evaluation's list bookkeeping in the
// engine's device flags.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.h"

namespace dyn {

// Words of an engine's device flags (dflags).
constexpr int kFlDisp = 0;       // largest squared displacement this evaluation (float bits)
constexpr int kFlBuilds = 1;     // list builds so far
// words 2 and 3 are unused
constexpr int kFlReady = 4;      // 1: the staging list holds a finished build
constexpr int kFlSel = 5;        // which copy of cj, cmask and chunks the last evaluation used
constexpr int kFlCommit = 6;     // 1: this evaluation's preparation kernel took the staged list
constexpr int kFlFillDone = 8;   // finished workgroups of the fill launch
constexpr int kFlPrevDisp = 9;   // the previous evaluation's largest displacement (float bits, Å)
constexpr int kFlGrowth = 10;    // decaying maximum of its growth per evaluation (float bits, Å)
constexpr int kFlWords = 16;

// The parts of the last evaluation's site forces (fixed point, component-
// major), as the finalize kernel sums them.
struct ForceParts {
  const int64_t* fx_site;     // correction (and compatibility) forces by site
  const int64_t* fx_recip;    // SPME forces by site, or null
  const int64_t* fitem;       // pair forces of each work item's 8 i slots, or null
  const uint32_t* chunk_ptr;  // first work item of each row
  const uint32_t* slot_of;    // per site: its slot
  uint64_t cap_chunks;
  uint32_t n_sites, n_slots;
  unsigned int* fl;           // the engine's device flags
  int forced;                 // the evaluation's list build was asked for by the host
  int unused;                 // no implicit padding: argument packs are compared byte for byte
};

// Pair forces of slot k: the sum over its row's work items (integer sums:
// the order does not change the result).
__device__ __forceinline__ void slot_item_forces(const int64_t* __restrict__ fitem,
                                                 const uint32_t* __restrict__ chunk_ptr,
                                                 uint64_t cap_chunks, uint32_t k, int64_t f[3]) {
  const uint32_t I = k / DYN_CL_SIZE, ia = k % DYN_CL_SIZE;
  const uint64_t c0 = min((uint64_t)chunk_ptr[I], cap_chunks);
  const uint64_t c1 = min((uint64_t)chunk_ptr[I + 1], cap_chunks);
  uint64_t c = c0;
  // four rows per round: their twelve loads are independent and issued
  // together
  for (; c + 4 <= c1; c += 4) {
    int64_t v[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t* o = fitem + (c + u) * (3 * DYN_CL_SIZE);
      v[u][0] = o[ia];
      v[u][1] = o[DYN_CL_SIZE + ia];
      v[u][2] = o[2 * DYN_CL_SIZE + ia];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      f[0] += v[u][0];
      f[1] += v[u][1];
      f[2] += v[u][2];
    }
  }
  for (; c < c1; ++c) {
    const int64_t* o = fitem + c * (3 * DYN_CL_SIZE);
    f[0] += o[ia];
    f[1] += o[DYN_CL_SIZE + ia];
    f[2] += o[2 * DYN_CL_SIZE + ia];
  }
}

// Site s's force (fixed point) from the parts.
__device__ __forceinline__ void site_force_sum(const ForceParts& P, uint32_t s, int64_t f[3]) {
  f[0] = P.fx_site[s];
  f[1] = P.fx_site[(size_t)P.n_sites + s];
  f[2] = P.fx_site[2 * (size_t)P.n_sites + s];
  if (P.fx_recip) {
    f[0] += P.fx_recip[s];
    f[1] += P.fx_recip[(size_t)P.n_sites + s];
    f[2] += P.fx_recip[2 * (size_t)P.n_sites + s];
  }
  if (P.fitem) slot_item_forces(P.fitem, P.chunk_ptr, P.cap_chunks, min(P.slot_of[s], P.n_slots - 1), f);
}

// The end of the evaluation's list bookkeeping (one thread, after the pair
// kernel): a staged list taken by the preparation kernel is the current one
// from now on; the growth of the largest displacement (a decaying maximum of
// its increase per evaluation; a new reference makes the increase negative,
// and a forced build's reading, against the reference it replaces, counts as
// none); the displacement reading starts again at zero.
__device__ __forceinline__ void evaluation_flags_end(unsigned int* fl, int forced) {
  if (fl[kFlCommit]) {
    fl[kFlSel] ^= 1u;
    fl[kFlReady] = 0u;
    fl[kFlCommit] = 0u;
  }
  const float d = sqrtf(__uint_as_float(fl[kFlDisp]));
  float growth = __uint_as_float(fl[kFlGrowth]);
  if (!forced) growth = fmaxf(0.97f * growth, d - __uint_as_float(fl[kFlPrevDisp]));
  fl[kFlGrowth] = __float_as_uint(growth);
  fl[kFlPrevDisp] = __float_as_uint(forced ? 0.0f : d);
  fl[kFlDisp] = 0u;
}

}  // namespace dyn
