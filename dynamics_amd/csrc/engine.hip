// engine.hip — native nonbonded engine on gfx950.
//
// One evaluation, all on the handle's stream and without host round trips:
//   bin        wrap positions into the box, find each site's x–y column of
//              its class, count and scatter the sites into column ranges (rebuild)
//   colsort    per column, bitonic sort by (z, site), cut into clusters of
//              8, write cluster slots and bounding boxes             (rebuild)
//   list       full cluster-pair list: every cluster lists every cluster
//              whose box is within the cutoff (+ buffer), both ways; rows
//              are cut into work items of ≤ item_groups groups; solute–solute
//              entries get exclusion masks                            (rebuild)
//   slots      gather current positions into cluster slots       (reuse path)
//   pairs      8×8 wave tiles over the work items: LJ + erfc Coulomb; each
//              lane keeps its i-site force in registers and a work item
//              stores its 8 i-site forces once: no atomics, no j-side sums
//   corr       1-4 pairs (scaled real space) and reciprocal corrections
//   pme        spread → FFT → convolve → inverse FFT → gather
//   finalize   site forces + the work-item rows of each site's slot →
//              site-order fixed-point forces; fixed-order energy sums
// Capacities are checked on the device; an overflow sets a status bit that
// the host reads at the next synchronization and reports as an error, after
// growing the buffers for the next call.
//
// Reference semantics: src/non_bonded.rs:258-412 (pair sum), :541-677 (pair
// set), :681-795 (SPME and the 1-4 correction), src/prep.rs:114-143
// (exclusions), src/cuda/cuda.cu:185-362 (pair physics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <type_traits>
#include <vector>

#include "common.h"
#include "wave.h"
#include "dyn_math.h"
#include "engine.h"
#include "engine_internal.h"
#include "pme.h"

#pragma clang fp contract(off)

using namespace dyn;

namespace {

constexpr int kWave = 64;
constexpr int kPairBlock = 256;       // 4 waves, one work item each at a time
constexpr uint64_t kPairGridCap = 6144;  // ~4 × the pair kernel's resident workgroups
constexpr int kColCap = 4096;         // sites per column (LDS sort)
constexpr int kMaxColsPerClass = 4096;
constexpr int kNumPart = 9;           // f64 partials per work item of the pair kernel

// status bits
constexpr unsigned kStColumn = 1u, kStRow = 2u, kStEntries = 4u, kStChunks = 8u, kStMask = 16u,
                   kStPme = 32u,   // an SPME brick list overflowed
                   kStDisp = 64u;  // a site moved half the list buffer since the list's build

// The words of the engine's device flags are defined in forces.h.

struct ClassGrid {
  int ncx, ncy;
  float wx, wy;
  int col_off;    // first global column of this class
  int n;          // sites of this class
  int site_base;  // first entry of this class in the column-sorted item array
};

struct EngGeom {
  float L[3], invL[3];
  ClassGrid cg[3];
  int n_cols;
  uint32_t n_dyn, n_water;
};

// Class cls's grid by selects (an indexed copy out of an argument pack held
// in memory would go through scratch memory).
__device__ __forceinline__ ClassGrid class_grid(const EngGeom& g, int cls) {
  ClassGrid r;
  r.ncx = cls == 0 ? g.cg[0].ncx : (cls == 1 ? g.cg[1].ncx : g.cg[2].ncx);
  r.ncy = cls == 0 ? g.cg[0].ncy : (cls == 1 ? g.cg[1].ncy : g.cg[2].ncy);
  r.wx = cls == 0 ? g.cg[0].wx : (cls == 1 ? g.cg[1].wx : g.cg[2].wx);
  r.wy = cls == 0 ? g.cg[0].wy : (cls == 1 ? g.cg[1].wy : g.cg[2].wy);
  r.col_off = cls == 0 ? g.cg[0].col_off : (cls == 1 ? g.cg[1].col_off : g.cg[2].col_off);
  r.n = cls == 0 ? g.cg[0].n : (cls == 1 ? g.cg[1].n : g.cg[2].n);
  r.site_base = cls == 0 ? g.cg[0].site_base : (cls == 1 ? g.cg[1].site_base : g.cg[2].site_base);
  return r;
}

// Whether this evaluation runs a list build.  The host asks for a build
// (force: the first evaluation, a new box, positions set from the host, no
// list buffer), or the device decides from words of the engine's flags that
// hold still during the evaluation — the largest squared displacement since
// the current list's build (complete once the preparation kernel is done),
// the recent growth of that displacement and whether the preparation kernel
// took a staged list:
// - a build before the pair kernel (staged = 0) when a site has moved half
//   the list buffer (limit: that distance squared, as float bits);
// - a staged build (staged = 1) when the largest displacement is within
//   max(margin_min, 2 × growth) of half the buffer and no staged list was
//   taken in this evaluation.
// Every kernel of a build evaluates the same words, so they all agree.
struct BuildGate {
  const unsigned int* fl;
  unsigned int limit;
  float half, margin_min;
  int staged;
  int force;
  int unused;  // no implicit padding: argument packs are compared byte for byte
};
static_assert(sizeof(BuildGate) == 32, "BuildGate has no padding");
__device__ __forceinline__ bool build_on(const BuildGate& G) {
  if (G.force) return true;
  const unsigned int d2 = G.fl[kFlDisp];
  if (!G.staged) return d2 >= G.limit;
  if (G.fl[kFlCommit]) return false;
  const float growth = __uint_as_float(G.fl[kFlGrowth]);
  return sqrtf(__uint_as_float(d2)) >= G.half - fmaxf(G.margin_min, 2.0f * growth);
}

__device__ __forceinline__ float wrap1(float x, float L, float invL) {
  float w = x - L * floorf(x * invL);
  if (!(w < L)) w = 0.0f;
  if (w < 0.0f) w = 0.0f;
  return w;
}

// Site of the t-th member of a class (sites: solute, then water O, M, H0, H1
// molecule-major).
__device__ __forceinline__ uint32_t class_site(int cls, uint32_t t, uint32_t nd) {
  if (cls == DYN_CLS_SOLUTE) return t;
  if (cls == DYN_CLS_WATER_O) return nd + 4u * t;
  const uint32_t m = t / 3u;
  return nd + 4u * m + 1u + (t - 3u * m);
}

__device__ __forceinline__ uint32_t site_column(const ClassGrid& cg, float x, float y) {
  int cx = (int)(x / cg.wx), cy = (int)(y / cg.wy);
  cx = min(max(cx, 0), cg.ncx - 1);
  cy = min(max(cy, 0), cg.ncy - 1);
  return (uint32_t)(cx * cg.ncy + cy);
}

// The last workgroup of a launch to finish: every workgroup calls this once,
// after its last write the next step reads; the one that returns true sees
// every other workgroup's writes (release fence before the count, acquire
// after) and resets the counter for the next launch.
__device__ __forceinline__ bool last_workgroup(unsigned int* __restrict__ done) {
  __shared__ unsigned int s_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(done, 1u) == gridDim.x - 1u ? 1u : 0u;
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  if (threadIdx.x == 0) *done = 0u;
  return true;
}

// Inclusive sum over the wave's lanes (DPP: shifts within rows of 16 lanes,
// then the row totals broadcast to the rows above).
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}


// Find every site's x–y column (from the positions the preparation kernel
// wrapped), take a rank in it and write the site into the column's slice of
// items (col_slice entries per column; the column sort flags a column that
// would hold more, and the call is repeated with longer slices); the
// positions become the reference of the displacement readings.  The rank
// order depends on atomic timing; the column sort that follows orders each
// column by (z, site), so the clusters do not.
constexpr int kBinBlock = 256;
__device__ __forceinline__ void k_eng_bin_count(BuildGate G, const float4* __restrict__ wpos,
                                                EngGeom g, float4* __restrict__ wpos_ref,
                                                uint32_t* __restrict__ col_count,
                                                uint32_t* __restrict__ items,
                                                unsigned int* __restrict__ n_builds,
                                                uint32_t col_slice) {
  if (!build_on(G)) return;
  const uint32_t n = g.n_dyn + 4u * g.n_water;
  for (uint32_t s0 = blockIdx.x * kBinBlock; s0 < n; s0 += gridDim.x * kBinBlock) {
    const uint32_t s = s0 + threadIdx.x;
    if (s == 0) ++*n_builds;
    const bool valid = s < n;
    uint32_t col = 0;
    if (valid) {
      const float4 w = wpos[s];
      wpos_ref[s] = w;
      const bool solute = s < g.n_dyn;
      const bool oxy = !solute && ((s - g.n_dyn) & 3u) == 0u;
      // the class grid's fields by selects (a runtime-indexed struct copy
      // would go through scratch memory)
      ClassGrid cg;
      cg.wx = solute ? g.cg[0].wx : (oxy ? g.cg[1].wx : g.cg[2].wx);
      cg.wy = solute ? g.cg[0].wy : (oxy ? g.cg[1].wy : g.cg[2].wy);
      cg.ncx = solute ? g.cg[0].ncx : (oxy ? g.cg[1].ncx : g.cg[2].ncx);
      cg.ncy = solute ? g.cg[0].ncy : (oxy ? g.cg[1].ncy : g.cg[2].ncy);
      const int col_off = solute ? g.cg[0].col_off : (oxy ? g.cg[1].col_off : g.cg[2].col_off);
      col = (uint32_t)col_off + site_column(cg, w.x, w.y);
    }
    // one atomic per distinct column of the wave
    const uint32_t rank = wave_rank(col, valid, col_count);
    if (valid && rank < col_slice) items[(size_t)col * col_slice + rank] = s;
  }
}

struct SlotArgs {
  const float4* wpos;
  const int32_t* ltype;   // per site
  const int32_t* wmol;    // per site: water molecule or −1
  float4* xyzq;
  int32_t* slot_type;
  int32_t* slot_gid;
  int32_t* slot_site;
  uint32_t* slot_of;
  int none_type;
  int unused;  // no implicit padding: argument packs are compared byte for byte
};
static_assert(sizeof(SlotArgs) == 72, "SlotArgs has no padding");

__device__ __forceinline__ int32_t site_gid(const int32_t* wmol, int32_t s) {
  const int32_t m = wmol[s];
  return m >= 0 ? m : -2 - s;
}

// One column, by one workgroup: sort the column's sites by (z, site) and cut
// them into clusters of 8 (the last padded), writing the slot arrays and each
// cluster's bounding box.  The column's count is cleared for the next
// binning pass (the binning kernel counts with atomics into zeroed counts).
constexpr int kSortBlock = 256;
constexpr unsigned kBuildGrid = 1024;  // workgroups of the grid-stride build kernels
__device__ __forceinline__ void
colsort_one(uint32_t c, uint32_t n, uint32_t cl0, const uint32_t* __restrict__ items,
            uint32_t col_slice, const int* __restrict__ col_class, const SlotArgs& S,
            uint8_t* __restrict__ cl_class, float4* __restrict__ bb_c, float4* __restrict__ bb_h,
            unsigned int* __restrict__ status, unsigned long long* key) {
  const size_t b = (size_t)c * col_slice;
  const uint32_t ncl = (n + DYN_CL_SIZE - 1) / DYN_CL_SIZE;
  if (ncl == 0) return;
  if (n > col_slice) {
    if (threadIdx.x == 0) atomicOr(status, kStColumn);
    return;
  }
  uint32_t np2 = 1;
  while (np2 < n) np2 <<= 1;
  for (uint32_t k = threadIdx.x; k < np2; k += kSortBlock) {
    unsigned long long v = ~0ull;
    if (k < n) {
      const uint32_t s = items[b + k];
      v = ((unsigned long long)__float_as_uint(S.wpos[s].z) << 32) | s;
    }
    key[k] = v;
  }
  __syncthreads();
  for (uint32_t size = 2; size <= np2; size <<= 1)
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t t = threadIdx.x; t < np2; t += kSortBlock) {
        const uint32_t p = t ^ stride;
        if (p > t) {
          const bool up = (t & size) == 0;
          const unsigned long long x = key[t], y = key[p];
          if ((x > y) == up) {
            key[t] = y;
            key[p] = x;
          }
        }
      }
      __syncthreads();
    }
  for (uint32_t k = threadIdx.x; k < ncl * DYN_CL_SIZE; k += kSortBlock) {
    const size_t slot = (size_t)cl0 * DYN_CL_SIZE + k;
    if (k < n) {
      const int32_t s = (int32_t)(key[k] & 0xffffffffu);
      S.slot_site[slot] = s;
      S.xyzq[slot] = S.wpos[s];
      S.slot_type[slot] = S.ltype[s];
      S.slot_gid[slot] = site_gid(S.wmol, s);
      S.slot_of[s] = (uint32_t)slot;
    } else {
      const int32_t s0 = (int32_t)(key[k & ~(uint32_t)(DYN_CL_SIZE - 1)] & 0xffffffffu);
      float4 p = S.wpos[s0];
      p.w = 0.0f;
      S.slot_site[slot] = -1;
      S.xyzq[slot] = p;
      S.slot_type[slot] = S.none_type;
      S.slot_gid[slot] = DYN_GID_PAD;
    }
  }
  for (uint32_t k = threadIdx.x; k < ncl; k += kSortBlock) {
    const uint32_t cl = cl0 + k;
    cl_class[cl] = (uint8_t)col_class[c];
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t m = 0; m < DYN_CL_SIZE; ++m) {
      const uint32_t r = k * DYN_CL_SIZE + m;
      if (r >= n) break;
      const float4 p = S.wpos[(int32_t)(key[r] & 0xffffffffu)];
      lo[0] = fminf(lo[0], p.x);
      lo[1] = fminf(lo[1], p.y);
      lo[2] = fminf(lo[2], p.z);
      hi[0] = fmaxf(hi[0], p.x);
      hi[1] = fmaxf(hi[1], p.y);
      hi[2] = fmaxf(hi[2], p.z);
    }
    bb_c[cl] = make_float4(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]), 0.f);
    // w: the cluster's sites (its first slots; the rest are padding)
    bb_h[cl] = make_float4(0.5f * (hi[0] - lo[0]), 0.5f * (hi[1] - lo[1]), 0.5f * (hi[2] - lo[2]),
                           __uint_as_float(min(n - k * DYN_CL_SIZE, (uint32_t)DYN_CL_SIZE)));
  }
}

__device__ __forceinline__ void
k_eng_colsort(BuildGate G, const uint32_t* __restrict__ col_count, const uint32_t* __restrict__ items,
              uint32_t col_slice, uint32_t* __restrict__ col_cl_start,
              const int* __restrict__ col_class, SlotArgs S, uint8_t* __restrict__ cl_class,
              float4* __restrict__ bb_c, float4* __restrict__ bb_h,
              unsigned int* __restrict__ status, uint32_t n_cols, uint32_t* __restrict__ n_cl) {
  __shared__ unsigned long long key[kColCap];
  __shared__ uint32_t wsum[kSortBlock / kWave];
  __shared__ uint32_t s_start;
  if (!build_on(G)) return;
  // every workgroup scans the columns' cluster counts itself (a scan launch
  // of its own would cost a dependent launch; one inside the binning launch
  // a device-scope fence per workgroup); workgroup 0 writes col_cl_start
  // and the total *n_cl
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  const uint32_t per = (n_cols + kSortBlock - 1) / kSortBlock;
  const uint32_t b0 = min(n_cols, t * per), e0 = min(n_cols, b0 + per);
  uint32_t sum = 0;
  for (uint32_t i = b0; i < e0; ++i) sum += (col_count[i] + DYN_CL_SIZE - 1) / DYN_CL_SIZE;
  const uint32_t incl = wave_incl_sum(sum);
  if (lane == kWave - 1) wsum[w] = incl;
  __syncthreads();
  uint32_t base = 0, total = 0;
  for (int k = 0; k < kSortBlock / kWave; ++k) {
    if (k < w) base += wsum[k];
    total += wsum[k];
  }
  const uint32_t excl = base + incl - sum;  // clusters of the columns before b0
  if (blockIdx.x == 0) {
    uint32_t r = excl;
    for (uint32_t i = b0; i < e0; ++i) {
      col_cl_start[i] = r;
      r += (col_count[i] + DYN_CL_SIZE - 1) / DYN_CL_SIZE;
    }
    if (t == 0) {
      col_cl_start[n_cols] = total;
      *n_cl = total;
    }
  }
  // a capped grid walks the columns (a closed gate then costs few workgroups)
  for (uint32_t c = blockIdx.x; c < n_cols; c += gridDim.x) {
    if ((uint32_t)t == c / per) {
      uint32_t r = excl;
      for (uint32_t i = b0; i < c; ++i) r += (col_count[i] + DYN_CL_SIZE - 1) / DYN_CL_SIZE;
      s_start = r;
    }
    __syncthreads();
    const uint32_t cl0 = s_start;
    colsort_one(c, col_count[c], cl0, items, col_slice, col_class, S, cl_class, bb_c, bb_h, status,
                key);
    __syncthreads();  // key[] and s_start are reused by the next column
  }
}


struct ListArgs {
  BuildGate gate;
  EngGeom g;
  float rlist2, rlist;
  const uint32_t* n_cl;
  const float4* bb_c;
  const float4* bb_h;
  const uint32_t* col_cl_start;
  const uint8_t* cl_class;
  const int32_t* slot_site;
  const float4* xyzq;        // per slot
  const int32_t* slot_gid;   // per slot
  const uint32_t* slot_of;   // per site: its slot in this build
  const uint32_t* ex_ptr;    // per solute site: exclusion CSR (1-2, 1-3 and 1-4)
  const uint32_t* ex_idx;
  uint32_t n_dyn;
  uint64_t cap_entries, cap_chunks;  // groups, work items
  uint32_t row_cap;                  // entries per unit slice of ent_tmp (multiple of 64)
  uint32_t chunk;                    // j groups per work item
  uint32_t null_slot;                // padding j slot (exclusion group: padding)
  uint32_t* longest;                 // longest row that overflowed its slice
  unsigned int* status;
  uint8_t* xb_tmp;                   // per entry of ent_tmp: its exclusion byte
  uint32_t* unit_cnt;                // per walk unit: entries written
  uint32_t* col_count;               // cleared by the walk for the next binning pass
  uint32_t n_cols;
  uint32_t unused;                   // no implicit padding: argument packs are compared byte for byte
  uint32_t* tot;                     // groups, work items, longest row (from the scan)
};

// LDS of the row walk, per wave: I's slot positions, a batch of up to 64
// columns (first cluster, range in the flattened candidate sequence, image
// seen from I), the codes of a sweep's candidates that passed the box test,
// and for a solute row the exclusion partners of I's sites as a hash table
// (linear probing, at most half full): key slot << 8 | the bits of the i
// slots the partner is excluded from.
constexpr int kExHash = 1024;
constexpr uint32_t kExEmpty = 0xFFFFFFFFu;
constexpr uint32_t kExPartners = kExHash / 2;  // partners of one cluster's sites
constexpr int kListBlock = 256;                // four waves, each on its own rows
constexpr int kListWaves = kListBlock / kWave;
constexpr unsigned kWalkGrid = 16384;          // workgroups of the walk and fill (units beyond: grid stride; a resident-sized grid with a static stride balanced the walk worse)
constexpr int kWalkPer = 1;                    // candidates per lane in a sweep and in a round
struct WalkLds {
  float4 pI[DYN_CL_SIZE];  // I's slot positions (padding slots: NaN)
  uint32_t js[kWave], beg[kWave], end[kWave];
  uint32_t img[kWave];     // the column's periodic image seen from I (cand_code bits 21-24)
  uint32_t ok[kWalkPer * kWave];  // cand_code of the sweep's passing candidates, in walk order
  uint32_t ex[kExHash];
};

// A passing candidate in one word: its cluster (bits 0-20; the slot count
// is below 2^24), its column's image in x and y (bits 21-22, 23-24: k + 1,
// the shift being k·L), in z (bits 25-26: k + 1 for the image of its
// bounding box nearest I's, used for every slot when no slot pair can be
// within rlist in another image; 3: each slot pair's own minimum image) and
// its number of sites less one (bits 27-29).
constexpr uint32_t kCandJ = (1u << 21) - 1u;
constexpr uint32_t kZSlot = 3u;
__device__ __forceinline__ float cand_shift(uint32_t code, int sh, float ext) {
  return (float)((int)((code >> sh) & 3u) - 1) * ext;
}

__device__ __forceinline__ uint32_t ex_hash(uint32_t slot) {
  return (slot * 2654435761u) >> (32 - 10);  // 10 bits: kExHash = 1024
}
static_assert(kExHash == 1024, "ex_hash takes 10 bits");

// Bits (i slot ia → bit ia) of the i slots whose site is a 1-2, 1-3 or 1-4
// partner of j slot `slot`'s site.
__device__ __forceinline__ uint32_t ex_bits(const uint32_t* ex, uint32_t slot) {
  for (uint32_t h = ex_hash(slot);; h = (h + 1) & (kExHash - 1)) {
    const uint32_t v = ex[h];
    if (v == kExEmpty) return 0u;
    if ((v >> 8) == slot) return v & 0xFFu;
  }
}

// The LDS accesses of one wave stay in program order around this point (a
// wave's lanes run in lockstep; the rows of one workgroup's waves are
// independent, so no workgroup barrier is needed).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// Pass 0, one wave per unit (i cluster I, class cls, half of the class's
// column window; unit u = kUnits·I + 2·cls + half): walk the unit's
// candidate clusters in a fixed order (column x, column y, cluster), keep
// those whose bounding box is within rlist of I's box (periodic), test
// their 8 slots against I's 8 slots (eight candidates per round, a lane per
// j slot), and append every j slot within rlist of at least one non-padding
// i slot.  Splitting rows into units keeps every wave's chain of dependent
// loads short: a row's time would otherwise be the sum of its sweeps and
// rounds.  A candidate column is seen as the periodic image the column walk
// reached it in, so x and y differences need no minimum image (z neither
// when rlist and both boxes' half-heights are short of half the box: then
// only the image of the candidate's box nearest I's can hold a slot pair
// within rlist; x and y do when the column range spans the whole box).
// Slots near only their own water molecule's sites may be listed: the pair
// kernel masks such pairs.  The unit's entries go into its fixed-capacity
// slice of ent_tmp and, for solute–solute units, each entry's exclusion byte
// (bit ia set when the j site is a 1-2, 1-3 or 1-4 partner of i slot ia's
// site) into the same place of xb_tmp; the list fill concatenates the two
// halves of each class and pads the class segment to a multiple of 8.
constexpr uint32_t kUnits = 6;  // walk units per row: 3 classes × 2 halves of the column window
__device__ __forceinline__ void
k_eng_list_walk(const ListArgs& A, uint32_t* __restrict__ ent_tmp) {
  __shared__ WalkLds Ls[kListWaves];
  if (!build_on(A.gate)) return;
  // the column sort has read the column counts: clear them for the next
  // binning pass (it counts with atomics into zeroed counts)
  for (uint32_t k = blockIdx.x * kListBlock + threadIdx.x; k < A.n_cols; k += gridDim.x * kListBlock)
    A.col_count[k] = 0;
  WalkLds& L = Ls[threadIdx.x / kWave];
  const uint32_t n_cl = *A.n_cl;
  const uint32_t n_units = n_cl * kUnits;
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int jj = lane & (DYN_CL_SIZE - 1);             // j slot of the lane in a round
  const uint32_t jc = (uint32_t)(lane / DYN_CL_SIZE);  // candidate of the lane in a round
  // a column is skipped when its x–y rectangle is beyond rlist of I's box;
  // the margin keeps the test conservative against rounding of the binning
  const float col_r2 = A.rlist2 * 1.0001f + 1e-3f;
  for (uint32_t u = blockIdx.x * kListWaves + threadIdx.x / kWave; u < n_units;
       u += gridDim.x * kListWaves) {
    const uint32_t I = u / kUnits;
    const int cls = (int)(u % kUnits) >> 1, half = (int)(u % kUnits) & 1;
    const int ci = A.cl_class[I];
    const ClassGrid cg = class_grid(A.g, cls);
    uint32_t n_out = 0;
    if (dyn_kind(ci, cls) != DYN_KIND_NONE && cg.n != 0) {
      uint32_t* ent = ent_tmp + (size_t)u * A.row_cap;
      uint8_t* xb = A.xb_tmp + (size_t)u * A.row_cap;
      const float4 c = A.bb_c[I], h = A.bb_h[I];
      wave_sync();  // the previous unit's reads of L are done
      // a padding i slot's position is NaN, so no distance to it is below rlist
      if (lane < DYN_CL_SIZE) {
        float4 p = A.xyzq[I * DYN_CL_SIZE + lane];
        if ((uint32_t)lane >= __float_as_uint(h.w)) p.x = p.y = p.z = __int_as_float(0x7fc00000);
        L.pI[lane] = p;
      }
      // solute–solute unit: the exclusion partners of I's sites, by slot
      bool exl = false;
      if (ci == DYN_CLS_SOLUTE && cls == DYN_CLS_SOLUTE) {
        uint32_t len = 0, ex0 = 0;
        if (lane < DYN_CL_SIZE) {
          const int32_t sa = A.slot_site[I * DYN_CL_SIZE + lane];
          if (sa >= 0) {
            ex0 = A.ex_ptr[sa];
            len = A.ex_ptr[sa + 1] - ex0;
          }
        }
        const uint32_t incl = wave_incl_sum(len);
        uint32_t n_ex = (uint32_t)__builtin_amdgcn_readlane((int)incl, DYN_CL_SIZE - 1);
        if (n_ex > kExPartners) {
          if (lane == 0) atomicOr(A.status, kStMask);
          n_ex = 0;
        }
        exl = n_ex != 0;
        for (uint32_t k = lane; k < (uint32_t)kExHash; k += kWave) L.ex[k] = kExEmpty;
        wave_sync();
        // lane t < n_ex inserts partner t (of i slot ia: beg[ia] ≤ t < beg[ia+1])
        const uint32_t beg_l = incl - len;
        for (uint32_t t0 = 0; t0 < n_ex; t0 += kWave) {
          const uint32_t t = t0 + lane;
          int ia = 0;
#pragma unroll
          for (int v = 1; v < DYN_CL_SIZE; ++v)
            if ((uint32_t)__builtin_amdgcn_readlane((int)beg_l, v) <= t) ia = v;
          const uint32_t e0 = (uint32_t)__shfl(ex0, ia, kWave);
          const uint32_t b0 = (uint32_t)__shfl(beg_l, ia, kWave);
          if (t < n_ex) {
            const uint32_t slot = A.slot_of[A.ex_idx[e0 + (t - b0)]];
            const uint32_t bit = 1u << ia;
            for (uint32_t hh = ex_hash(slot);; hh = (hh + 1) & (kExHash - 1)) {
              uint32_t cur = L.ex[hh];
              if (cur == kExEmpty) {
                cur = atomicCAS(&L.ex[hh], kExEmpty, (slot << 8) | bit);
                if (cur == kExEmpty) break;
              }
              if ((cur >> 8) == slot) {
                atomicOr(&L.ex[hh], bit);
                break;
              }
            }
          }
        }
      }
      wave_sync();
      const uint32_t tag = (uint32_t)cls << DYN_JCLS_SHIFT;
      int x0 = (int)floorf((c.x - h.x - A.rlist) / cg.wx), x1 = (int)floorf((c.x + h.x + A.rlist) / cg.wx);
      int y0 = (int)floorf((c.y - h.y - A.rlist) / cg.wy), y1 = (int)floorf((c.y + h.y + A.rlist) / cg.wy);
      bool full_x = false, full_y = false;
      if (x1 - x0 + 1 >= cg.ncx) {
        x0 = 0;
        x1 = cg.ncx - 1;
        full_x = true;
      }
      if (y1 - y0 + 1 >= cg.ncy) {
        y0 = 0;
        y1 = cg.ncy - 1;
        full_y = true;
      }
      const bool full_xy = full_x || full_y;
      const int ny = y1 - y0 + 1, n_colc = (x1 - x0 + 1) * ny;
      // this unit's half of the window's columns (in walk order)
      const int kmid = (n_colc + 1) / 2;
      const int kbeg = half ? kmid : 0, kend = half ? n_colc : kmid;
      for (int cb = kbeg; cb < kend; cb += kWave) {
        const int k = cb + lane;
        uint32_t js = 0, cnt = 0, img = 0;
        if (k < kend) {
          const int xx = x0 + k / ny, yy = y0 + k % ny;
          float dx = 0.f, dy = 0.f;
          if (!full_x) dx = fmaxf(0.f, fmaxf(xx * cg.wx - (c.x + h.x), (c.x - h.x) - (xx + 1) * cg.wx));
          if (!full_y) dy = fmaxf(0.f, fmaxf(yy * cg.wy - (c.y + h.y), (c.y - h.y) - (yy + 1) * cg.wy));
          if (dx * dx + dy * dy < col_r2) {
            const int cx = ((xx % cg.ncx) + cg.ncx) % cg.ncx;
            const int cy = ((yy % cg.ncy) + cg.ncy) % cg.ncy;
            const int col = cg.col_off + cx * cg.ncy + cy;
            js = A.col_cl_start[col];
            cnt = A.col_cl_start[col + 1] - js;
            img = (uint32_t)((xx - cx) / cg.ncx + 1) << 21 | (uint32_t)((yy - cy) / cg.ncy + 1) << 23;
          }
        }
        const uint32_t incl = wave_incl_sum(cnt);
        const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
        L.js[lane] = js;
        L.beg[lane] = incl - cnt;
        L.end[lane] = incl;
        L.img[lane] = img;
        wave_sync();
        // sweeps of kWalkPer·64 candidates, kWalkPer per lane (their loads
        // in flight together): candidate t is lane t mod 64 of part t / 64
        for (uint32_t t0 = 0; t0 < tot; t0 += kWalkPer * kWave) {
          bool ok[kWalkPer];
          uint32_t code[kWalkPer];
#pragma unroll
          for (int r = 0; r < kWalkPer; ++r) {
            // candidate t: its cluster (first column whose range ends after
            // t), the image of its column and its bounding box
            const uint32_t t = t0 + r * kWave + lane;
            ok[r] = t < tot;
            code[r] = 0;
            if (ok[r]) {
              int lo = 0, hi = kWave - 1;
              while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (L.end[mid] > t) hi = mid;
                else lo = mid + 1;
              }
              const uint32_t J = L.js[lo] + (t - L.beg[lo]);
              const float4 cJ = A.bb_c[J], hJ = A.bb_h[J];
              float dx = c.x - cJ.x, dy = c.y - cJ.y, dz = c.z - cJ.z;
              const float kz = rintf(dz * A.g.invL[2]);
              dx = fabsf(dx - A.g.L[0] * rintf(dx * A.g.invL[0]));
              dy = fabsf(dy - A.g.L[1] * rintf(dy * A.g.invL[1]));
              dz = fabsf(dz - A.g.L[2] * kz);
              const uint32_t zc =
                  A.rlist + h.z + hJ.z < 0.45f * A.g.L[2] ? (uint32_t)((int)kz + 1) : kZSlot;
              dx = fmaxf(0.f, dx - h.x - hJ.x);
              dy = fmaxf(0.f, dy - h.y - hJ.y);
              dz = fmaxf(0.f, dz - h.z - hJ.z);
              ok[r] = dx * dx + dy * dy + dz * dz < A.rlist2;
              code[r] = J | L.img[lo] | zc << 25 | (__float_as_uint(hJ.w) - 1u) << 27;
            }
          }
          uint32_t n_ok = 0;
#pragma unroll
          for (int r = 0; r < kWalkPer; ++r) {
            const unsigned long long bal = __ballot(ok[r]);
            if (ok[r]) L.ok[n_ok + __popcll(bal & lt)] = code[r];
            n_ok += (uint32_t)__popcll(bal);
          }
          wave_sync();
          // rounds of kWalkPer·8 candidates: lane jj + 8·jc tests j slot jj of
          // candidates k0 + 8·r + jc against all eight i slots, so the kept
          // slots come out candidate by candidate, each in slot order, as a
          // one-candidate-at-a-time walk would write them
          for (uint32_t k0 = 0; k0 < n_ok; k0 += kWalkPer * (kWave / DYN_CL_SIZE)) {
            uint32_t cd[kWalkPer];
            float4 pj[kWalkPer];
            bool jv[kWalkPer], zslot[kWalkPer], keep[kWalkPer];
            bool any_zslot = false;
#pragma unroll
            for (int r = 0; r < kWalkPer; ++r) {
              const uint32_t kk = k0 + r * (kWave / DYN_CL_SIZE) + jc;
              cd[r] = 0;
              pj[r] = make_float4(0.f, 0.f, 0.f, 0.f);
              if (kk < n_ok) {
                cd[r] = L.ok[kk];
                pj[r] = A.xyzq[(cd[r] & kCandJ) * DYN_CL_SIZE + jj];
              }
              jv[r] = kk < n_ok && (uint32_t)jj <= ((cd[r] >> 27) & 7u);
              zslot[r] = jv[r] && ((cd[r] >> 25) & 3u) == kZSlot;
              any_zslot = any_zslot || zslot[r];
            }
            if (!full_xy && __ballot(any_zslot) == 0ull) {
              float px[kWalkPer], py[kWalkPer], pz[kWalkPer];
              bool kp[kWalkPer];
#pragma unroll
              for (int r = 0; r < kWalkPer; ++r) {
                px[r] = pj[r].x + cand_shift(cd[r], 21, A.g.L[0]);
                py[r] = pj[r].y + cand_shift(cd[r], 23, A.g.L[1]);
                pz[r] = pj[r].z + cand_shift(cd[r], 25, A.g.L[2]);
                kp[r] = false;
              }
#pragma unroll
              for (int v = 0; v < DYN_CL_SIZE; ++v) {
                const float4 q = L.pI[v];
#pragma unroll
                for (int r = 0; r < kWalkPer; ++r) {
                  const float d0 = q.x - px[r], d1 = q.y - py[r], d2 = q.z - pz[r];
                  kp[r] |= dyn_fmaf(d2, d2, dyn_fmaf(d1, d1, d0 * d0)) < A.rlist2;
                }
              }
#pragma unroll
              for (int r = 0; r < kWalkPer; ++r) keep[r] = kp[r] && jv[r];
            } else {
#pragma unroll
              for (int r = 0; r < kWalkPer; ++r) {
                const float px = full_xy ? pj[r].x : pj[r].x + cand_shift(cd[r], 21, A.g.L[0]);
                const float py = full_xy ? pj[r].y : pj[r].y + cand_shift(cd[r], 23, A.g.L[1]);
                const float pz = zslot[r] ? pj[r].z : pj[r].z + cand_shift(cd[r], 25, A.g.L[2]);
                bool kp = false;
#pragma unroll
                for (int v = 0; v < DYN_CL_SIZE; ++v) {
                  const float4 q = L.pI[v];
                  float d0 = q.x - px, d1 = q.y - py, d2 = q.z - pz;
                  if (full_xy) {
                    d0 = dyn_min_image_rint(d0, A.g.L[0], A.g.invL[0]);
                    d1 = dyn_min_image_rint(d1, A.g.L[1], A.g.invL[1]);
                  }
                  if (zslot[r]) d2 = dyn_min_image_rint(d2, A.g.L[2], A.g.invL[2]);
                  kp |= dyn_fmaf(d2, d2, dyn_fmaf(d1, d1, d0 * d0)) < A.rlist2;
                }
                keep[r] = kp && jv[r];
              }
            }
#pragma unroll
            for (int r = 0; r < kWalkPer; ++r) {
              const unsigned long long kb = __ballot(keep[r]);
              if (keep[r]) {
                const uint32_t pos = n_out + (uint32_t)__popcll(kb & lt);
                const uint32_t slot = (cd[r] & kCandJ) * DYN_CL_SIZE + jj;
                if (pos < A.row_cap) {
                  ent[pos] = slot | tag;
                  if (exl) xb[pos] = (uint8_t)ex_bits(L.ex, slot);
                }
              }
              n_out += (uint32_t)__popcll(kb);
            }
          }
          wave_sync();  // L.ok is rewritten by the next sweep
        }
      }
    }
    if (lane == 0) {
      if (n_out > A.row_cap) {
        atomicOr(A.status, kStRow);
        atomicMax(A.longest, n_out);
      }
      A.unit_cnt[u] = min(n_out, A.row_cap);
    }
  }
}


}  // namespace

// The walk is bound by the latency of its dependent loads: room for six
// waves per SIMD (six workgroups per compute unit, within its LDS), so that
// one launch holds every row of a 37,000-site system at once.
template <>
struct dyn::MinWavesPerSimd<k_eng_list_walk> {
  static constexpr int value = 6;
};

namespace {

// The two copies of the per-group arrays (the pair kernel reads the one
// flags[kFlSel] names while a staged build fills the other).
struct ListCopies {
  uint32_t *cj0, *cj1;
  unsigned long long *cmask0, *cmask1;
  uint2 *chunks0, *chunks1;
  unsigned int* fl;
  int staged;  // this build fills the staging list
  int unused;  // no implicit padding: argument packs are compared byte for byte
};
static_assert(sizeof(ListCopies) == 64, "ListCopies has no padding");

// Pass 1: compact the rows into cj (groups of 8 entries; each class segment
// the two halves of its window, padded with the null slot), write the work
// items (I, first group) and, for the solute–solute groups (solute rows:
// solute clusters are numbered first; the solute segment comes first in
// every row), the 64-bit masks: bit ia + 8·jb of a group is cleared when the
// sites of i slot ia and of the group's entry jb are a 1-2, 1-3 or 1-4 pair
// (entry jb's exclusion byte, shifted into place).  One wave per class
// segment of a row; the segment's entries move four per lane at a time.
constexpr int kFillUnroll = 4;
constexpr uint32_t kFillRows = 4;  // rows per workgroup of the fill (3 class segments each)

// A row's groups of 8 entries: each class segment (the two halves' entries)
// padded to 8.
__device__ __forceinline__ uint32_t row_groups(const uint32_t* __restrict__ unit_cnt, uint32_t I) {
  static_assert(kUnits == 6, "a row's unit counts are three 8-byte words");
  const uint2* uc = reinterpret_cast<const uint2*>(unit_cnt + (size_t)I * kUnits);
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint2 h = uc[k];
    m += (h.x + h.y + DYN_CL_SIZE - 1) / DYN_CL_SIZE;
  }
  return m;
}

__device__ __forceinline__ void
k_eng_list_fill(const ListArgs& A, const uint32_t* __restrict__ ent_tmp,
                uint32_t* __restrict__ row_cnt, uint32_t* __restrict__ row_ptr,
                uint32_t* __restrict__ chunk_ptr, ListCopies C) {
  __shared__ uint32_t s_m[kFillRows], s_g[kFillRows], s_i[kFillRows];
  __shared__ uint32_t s_red[2][kListWaves];
  if (!build_on(A.gate)) return;
  // the copy this evaluation's pair kernel reads (a build before it) or the
  // other one (a staged build, current from the next evaluation on; none
  // runs in an evaluation that took a staged list)
  const bool second = ((C.fl[kFlSel] ^ C.fl[kFlCommit]) != 0u) != (C.staged != 0);
  uint32_t* __restrict__ cj = second ? C.cj1 : C.cj0;
  unsigned long long* __restrict__ cmask = second ? C.cmask1 : C.cmask0;
  uint2* __restrict__ chunks = second ? C.chunks1 : C.chunks0;
  const uint32_t n_cl = *A.n_cl;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  if (n_cl == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    row_ptr[0] = 0;
    chunk_ptr[0] = 0;
    A.tot[0] = 0;
    A.tot[1] = 0;
  }
  for (uint32_t blk = blockIdx.x; blk * kFillRows < n_cl; blk += gridDim.x) {
    const uint32_t r0 = blk * kFillRows, nr = min(kFillRows, n_cl - r0);
    // the groups and work items of the rows before this block (every
    // workgroup sums them itself: a scan launch of its own would cost a
    // dependent launch on every step, and a scan inside the walk a
    // device-scope fence per workgroup)
    uint32_t sg = 0, si = 0;
    for (uint32_t I = threadIdx.x; I < r0; I += kListBlock) {
      const uint32_t m = row_groups(A.unit_cnt, I);
      sg += m;
      si += (m + A.chunk - 1) / A.chunk;
    }
    sg = (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_sum(sg), kWave - 1);
    si = (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_sum(si), kWave - 1);
    if (lane == 0) {
      s_red[0][w] = sg;
      s_red[1][w] = si;
    }
    if (threadIdx.x < kFillRows) s_m[threadIdx.x] = threadIdx.x < nr ? row_groups(A.unit_cnt, r0 + threadIdx.x) : 0u;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t bg = 0, bi = 0;
      for (int k = 0; k < kListWaves; ++k) {
        bg += s_red[0][k];
        bi += s_red[1][k];
      }
      for (uint32_t k = 0; k < kFillRows; ++k) {
        s_g[k] = bg;
        s_i[k] = bi;
        bg += s_m[k];
        bi += (s_m[k] + A.chunk - 1) / A.chunk;
      }
      if (r0 + nr == n_cl) {  // the block with the last row: the totals
        row_ptr[n_cl] = bg;
        chunk_ptr[n_cl] = bi;
        A.tot[0] = bg;
        A.tot[1] = bi;
      }
    }
    __syncthreads();
    if (threadIdx.x < nr) {
      row_cnt[r0 + threadIdx.x] = s_m[threadIdx.x];
      row_ptr[r0 + threadIdx.x] = s_g[threadIdx.x];
      chunk_ptr[r0 + threadIdx.x] = s_i[threadIdx.x];
    }
    for (uint32_t sgi = (uint32_t)w; sgi < nr * 3u; sgi += kListWaves) {
      const uint32_t I = r0 + sgi / 3u;
      const int cls = (int)(sgi % 3u);
      const uint32_t m = s_m[sgi / 3u];
      const uint64_t base = s_g[sgi / 3u];
      if (cls == 0) {
        const uint32_t c0 = s_i[sgi / 3u], nch = (m + A.chunk - 1) / A.chunk;
        if ((uint64_t)c0 + nch <= A.cap_chunks) {
          for (uint32_t k = lane; k < nch; k += kWave)
            chunks[c0 + k] = make_uint2(I, (uint32_t)base + k * A.chunk);
        } else if (lane == 0) {
          atomicOr(A.status, kStChunks);
        }
        if (base + m > A.cap_entries && lane == 0) atomicOr(A.status, kStEntries);
      }
      if (base + m > A.cap_entries) continue;
      const uint32_t* uc = A.unit_cnt + (size_t)I * kUnits;
      uint32_t off = 0;  // the segment's first entry in the row
      for (int k = 0; k < cls; ++k) off += (uc[2 * k] + uc[2 * k + 1] + DYN_CL_SIZE - 1) & ~(uint32_t)(DYN_CL_SIZE - 1);
      const uint32_t n0 = uc[2 * cls], n = n0 + uc[2 * cls + 1];
      if (n == 0) continue;
      const uint32_t np = (n + DYN_CL_SIZE - 1) & ~(uint32_t)(DYN_CL_SIZE - 1);
      const uint32_t tag = (uint32_t)cls << DYN_JCLS_SHIFT;
      const bool ss = A.cl_class[I] == DYN_CLS_SOLUTE && cls == DYN_CLS_SOLUTE;
      const size_t u0 = (size_t)I * kUnits + 2 * cls;
      uint32_t* out = cj + (base * DYN_CL_SIZE + off);
      for (uint32_t p0 = 0; p0 < np; p0 += kFillUnroll * kWave) {
        uint32_t e[kFillUnroll];
        unsigned long long xbits[kFillUnroll];
#pragma unroll
        for (int r = 0; r < kFillUnroll; ++r) {
          // entry p of the segment: from the first half, the second, or padding
          const uint32_t p = p0 + r * kWave + lane;
          const size_t src = p < n0 ? u0 * A.row_cap + p : (u0 + 1) * A.row_cap + (p - n0);
          e[r] = A.null_slot | tag;
          xbits[r] = 0ull;
          if (p < n) {
            e[r] = ent_tmp[src];
            if (ss) xbits[r] = (unsigned long long)A.xb_tmp[src] << (DYN_CL_SIZE * (p & 7u));
          }
        }
#pragma unroll
        for (int r = 0; r < kFillUnroll; ++r) {
          const uint32_t p = p0 + r * kWave + lane;
          if (p < np) out[p] = e[r];
          if (ss) {
            // the group's 8 bytes: an OR over its 8 lanes (groups start at
            // multiples of 8 within the segment, and the segment at a
            // multiple of 8 within the row)
            unsigned long long x = xbits[r];
            x |= __shfl_xor(x, 1, kWave);
            x |= __shfl_xor(x, 2, kWave);
            x |= __shfl_xor(x, 4, kWave);
            if ((lane & 7) == 0 && p < np) cmask[base + (off + p) / DYN_CL_SIZE] = ~x;
          }
        }
      }
    }
    __syncthreads();  // s_m, s_g, s_i and s_red are reused by the next block
  }
  // a staged build is complete when the launch's last workgroup is done
  if (C.staged && last_workgroup(C.fl + kFlFillDone) && threadIdx.x == 0) C.fl[kFlReady] = 1u;
}

struct PairArgs {
  DynPairParams P;
  const float4* xyzq;
  const int32_t* slot_type;
  const int32_t* slot_gid;
  const uint8_t* cl_class;
  const uint32_t* row_ptr;   // groups
  const uint32_t* cj;        // 8 entries per group (copy 0; cj2: copy 1)
  const unsigned long long* cmask;
  const uint2* chunks;
  const uint32_t* cj2;
  const unsigned long long* cmask2;
  const uint2* chunks2;
  const unsigned int* fl;    // the engine's flags: the copy in use, the largest displacement
  int staged;                // staged list builds: a stale current list voids the evaluation
  unsigned int hard_limit;   // half the list buffer, squared (float bits)
  const uint32_t* n_chunks;
  const float2* lj_table;    // (c6, c12) per type pair
  const float2* lj_se;       // (σ, ε) per type pair, for soft-core
  const double2* lj_table64; // (c6, c12) in double, for the FP64 option
  int64_t* mol_e;            // n_mol² molecule-pair energies (fixed point), or null
  const int32_t* mol_of;     // per solute atom: its molecule
  uint32_t n_mol;
  int64_t* fitem;            // per work item: 3 components × 8 i slots, fixed point
  const int32_t* slot_site;  // per slot: its site (−1: padding)
  int64_t* fx_site;          // site forces (fixed point, component-major): the i forces are added
  uint32_t n_sites;
  uint32_t keep_items;       // write each item's row of i forces to fitem (for inspection)
  double* partials;          // kNumPart values × cap_chunks (value-major), per work item
  uint64_t cap_chunks;
  uint32_t chunk;            // j groups per work item
  unsigned int* status;      // a list-build overflow leaves the list unusable
};

__device__ __forceinline__ float xor_sum(float v, int m) { return v + __shfl_xor(v, m, kWave); }
__device__ __forceinline__ double xor_sum_d(double v, int m) { return v + __shfl_xor(v, m, kWave); }

template <class T>
struct TileAccT {
  T fi[3];
  T e_lj_s, e_lj_w, e_c_s, e_c_w, vir, dh;
  uint32_t n_ss, n_lj, n_c;  // wave-uniform counts (popcounts)
};

__device__ __forceinline__ void wave_add_fx(int64_t* __restrict__ M, int key, double e);

// Molecule-pair energy of one tile entry (src/non_bonded.rs:350-361): the
// pair energy of every solute–solute lane goes to M[m_i·n + m_j] in 2⁻³²
// kcal/mol fixed point.  The full list visits each pair from both rows, so
// each visit adds its own direction and a diagonal entry takes half.  A wave
// whose live lanes share one entry adds it once (the common case: one
// molecule, or one molecule pair per tile); otherwise lane by lane.
__device__ __forceinline__ void mol_energy_add(int64_t* __restrict__ M,
                                               const int32_t* __restrict__ mol_of, uint32_t n_mol,
                                               int gi, int gj, double e) {
  int key = -1;
  if (e != 0.0) {
    const int mi = mol_of[-2 - gi], mj = mol_of[-2 - gj];  // solute slots: gid = −2 − site
    key = mi * (int)n_mol + mj;
    if (mi == mj) e *= 0.5;
  }
  wave_add_fx(M, key, e);
}

// Adds e to M[key] (2⁻³² fixed point) for every lane with key ≥ 0: one
// atomic for the wave when all such lanes share the key.  All lanes of the
// wave must be active.
__device__ __forceinline__ void wave_add_fx(int64_t* __restrict__ M, int key, double e) {
  const unsigned long long live = __ballot(key >= 0);
  if (!live) return;
  const int kf = __shfl(key, __ffsll((long long)live) - 1, kWave);
  if (__ballot(key >= 0 && key != kf) == 0ull) {
    double x = key >= 0 ? e : 0.0;
    for (int m = 1; m < kWave; m <<= 1) x += __shfl_xor(x, m, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(M + kf),
                (unsigned long long)dyn_fx64(x));
  } else if (key >= 0) {
    atomicAdd(reinterpret_cast<unsigned long long*>(M + key),
              (unsigned long long)dyn_fx64(e));
  }
}

// ALCH: the group holds at least one lane across the decoupled molecule's
// boundary (the alchemical variant returns the plain result on the others).
// ENERGY: accumulate energies, virial, dH/dλ and pair counts (forces only
// otherwise).
// F64: FP64 pair arithmetic (precision sweep; no alchemical groups).
template <int KIND, bool ALCH, bool ENERGY, bool F64>
__device__ __forceinline__ void tile_entry(const PairArgs& A, const float pi[4], int ti,
                                           const float4& pjv, int tj, bool active,
                                           TileAccT<std::conditional_t<F64, double, float>>& T,
                                           int gi, int gj) {
  using R = std::conditional_t<F64, double, float>;
  const float pj[4] = {pjv.x, pjv.y, pjv.z, pjv.w};
  R d[3], s, elj, ec, r2;
  int hit;
  if constexpr (F64) {
    dyn_tile_pair_f64<KIND>(pi, ti, pj, tj, active, A.P, A.lj_table64, d, &s, &elj, &ec, &r2, &hit);
  } else if constexpr (ALCH) {
    float dh;
    dyn_tile_pair_alch<KIND>(pi, ti, pj, tj, active, A.P, A.lj_table, A.lj_se, d, &s, &elj, &ec,
                             &r2, &hit, &dh);
    if (ENERGY) T.dh += dh;
  } else {
    dyn_tile_pair<KIND>(pi, ti, pj, tj, active, A.P, A.lj_table, d, &s, &elj, &ec, &r2, &hit);
  }
  if constexpr (F64) {
    T.fi[0] += s * d[0];
    T.fi[1] += s * d[1];
    T.fi[2] += s * d[2];
  } else {
    T.fi[0] = dyn_fmaf(s, d[0], T.fi[0]);
    T.fi[1] = dyn_fmaf(s, d[1], T.fi[1]);
    T.fi[2] = dyn_fmaf(s, d[2], T.fi[2]);
  }
  if (!ENERGY) return;
  const bool solute = KIND == DYN_KIND_SS || KIND == DYN_KIND_SO || KIND == DYN_KIND_SQ;
  if (KIND != DYN_KIND_SQ && KIND != DYN_KIND_QQ) {
    if (solute) T.e_lj_s += elj;
    else T.e_lj_w += elj;
  }
  if (KIND != DYN_KIND_SO && KIND != DYN_KIND_OO) {
    if (solute) T.e_c_s += ec;
    else T.e_c_w += ec;
  }
  T.vir += s * r2;
  if constexpr (KIND == DYN_KIND_SS)
    if (A.mol_e) mol_energy_add(A.mol_e, A.mol_of, A.n_mol, gi, gj, (double)elj + (double)ec);
  const uint32_t h = (uint32_t)__popcll(__ballot(hit != 0));
  if (KIND == DYN_KIND_SS) T.n_ss += h;
  else if (KIND == DYN_KIND_SO || KIND == DYN_KIND_OO) T.n_lj += h;
  else T.n_c += h;
}

// LDS of the FP32 pair kernel: the erfc Coulomb table, then the LJ
// (c6, c12) table of n_types² entries.  Per-lane table reads from LDS are
// counted apart from the global loads, so the j data prefetched for the next
// group stays in flight while a group reads its table values (a global table
// read would make the wave wait for every load issued before it).
constexpr size_t kPairTabEntries = 2 * DYN_COUL_TAB_N;  // float4
__host__ __device__ constexpr size_t pair_lds_bytes(int n_types, bool energy) {
  return (energy ? kPairTabEntries : DYN_COUL_TAB_N) * 16 + (size_t)n_types * (size_t)n_types * 8;
}

// One wave per work item at a time (≤ A.chunk groups of one row); lane
// ia + 8·jb holds the pair (i slot 8I+ia, j entry jb of the group).  The next
// group's j data is loaded before the current group is evaluated.  A work
// item stores its 8 i-slot forces in its own row of fitem (no atomics) and
// its energies, virial and pair counts as one partial per value, so every
// sum keeps a fixed order whichever wave runs the item; the grid holds a few
// times as many workgroups as fit on the device at once, each copying the
// tables to LDS once and taking items in a grid stride, so the hardware's
// workgroup dispatch still balances the load.
template <bool ENERGY, bool F64>
__device__ __forceinline__ void
k_eng_pairs(PairArgs A) {
  using R = std::conditional_t<F64, double, float>;
  const int lane = threadIdx.x & (kWave - 1);
  const int ia = lane & 7, jb = lane >> 3;
  if (*A.status) return;
  // in staged mode a site may have moved half the list buffer since the
  // current list's build: the list may then miss pairs, and the evaluation
  // is void (every workgroup reads the same words)
  if (A.staged && A.fl[kFlDisp] >= A.hard_limit) {
    if (threadIdx.x == 0) atomicOr(A.status, kStDisp);
    return;
  }
  if ((A.fl[kFlSel] ^ A.fl[kFlCommit]) != 0u) {
    A.cj = A.cj2;
    A.cmask = A.cmask2;
    A.chunks = A.chunks2;
  }
  if constexpr (!F64) {
    extern __shared__ float4 pair_lds[];
    float4* tab = pair_lds;
    float2* ljt = reinterpret_cast<float2*>(pair_lds + (ENERGY ? kPairTabEntries : DYN_COUL_TAB_N));
    const int nt2 = A.P.n_types * A.P.n_types;
    // forces only: the φ half of the Coulomb table (the erfc half serves energies)
    const int ntab = ENERGY ? (int)kPairTabEntries : DYN_COUL_TAB_N;
    for (int k = threadIdx.x; k < ntab; k += kPairBlock) tab[k] = A.P.coul_tab[k];
    for (int k = threadIdx.x; k < nt2; k += kPairBlock) ljt[k] = A.lj_table[k];
    __syncthreads();
    A.P.coul_tab = tab;
    A.lj_table = ljt;
  }
  const uint32_t n_items = (uint32_t)min((uint64_t)*A.n_chunks, A.cap_chunks);
  const uint32_t wave0 = blockIdx.x * (kPairBlock / kWave) + (threadIdx.x / kWave);
  const uint32_t nwaves = gridDim.x * (kPairBlock / kWave);
  for (uint32_t item = wave0; item < n_items; item += nwaves) {
    const uint2 ch = A.chunks[item];
    const uint32_t I = __builtin_amdgcn_readfirstlane(ch.x);
    const uint32_t g0 = __builtin_amdgcn_readfirstlane(ch.y);
    const uint32_t g1 = min(g0 + A.chunk, A.row_ptr[I + 1]);
    const float4 piv = A.xyzq[I * DYN_CL_SIZE + ia];
    const float pi[4] = {piv.x, piv.y, piv.z, piv.w};
    const int ti = A.slot_type[I * DYN_CL_SIZE + ia];
    const int gi = A.slot_gid[I * DYN_CL_SIZE + ia];
    const int ci = A.cl_class[I];
    TileAccT<R> T;
    T.fi[0] = T.fi[1] = T.fi[2] = R(0);
    T.e_lj_s = T.e_lj_w = T.e_c_s = T.e_c_w = T.vir = T.dh = R(0);
    T.n_ss = T.n_lj = T.n_c = 0;
    // two-deep software pipeline: while group g is evaluated, the j data of
    // group g + 1 and the entry of group g + 2 are in flight.  The loads are
    // issued on every iteration (the last ones repeat the row's last group)
    // so the wait counts stay exact.
    const uint32_t glast = g1 - 1;
    uint32_t v = A.cj[(size_t)g0 * DYN_CL_SIZE + jb];
    uint32_t vn = A.cj[(size_t)min(g0 + 1, glast) * DYN_CL_SIZE + jb];
    float4 pj = A.xyzq[v & DYN_JSLOT_MASK];
    int tj = A.slot_type[v & DYN_JSLOT_MASK];
    int gj = A.slot_gid[v & DYN_JSLOT_MASK];
    for (uint32_t g = g0; g < g1; ++g) {
      // readfirstlane returns a signed int: shift the entry as unsigned (class
      // 2 sets the top bit)
      const int cls = (int)((uint32_t)__builtin_amdgcn_readfirstlane(v) >> DYN_JCLS_SHIFT);
      const float4 pjc = pj;
      const int tjc = tj, gjc = gj;
      {
        const uint32_t jsn = vn & DYN_JSLOT_MASK;
        pj = A.xyzq[jsn];
        tj = A.slot_type[jsn];
        gj = A.slot_gid[jsn];
        v = vn;
        vn = A.cj[(size_t)min(g + 2, glast) * DYN_CL_SIZE + jb];
      }
      const int kind = dyn_kind(ci, cls);
      bool active = dyn_slot_pair_valid(gi, gjc);
      if (kind == DYN_KIND_SS) active = active && ((A.cmask[g] >> lane) & 1ull);
      const bool alch = !F64 && A.P.alch_on && kind <= DYN_KIND_SQ &&
                        __ballot(((ti ^ tjc) & DYN_TYPE_ALCH) != 0) != 0ull;
      switch (kind) {
        case DYN_KIND_QQ: tile_entry<DYN_KIND_QQ, false, ENERGY, F64>(A, pi, ti, pjc, tjc, active, T, gi, gjc); break;
        case DYN_KIND_OO: tile_entry<DYN_KIND_OO, false, ENERGY, F64>(A, pi, ti, pjc, tjc, active, T, gi, gjc); break;
        case DYN_KIND_SS:
          if (alch) tile_entry<DYN_KIND_SS, true, ENERGY, F64>(A, pi, ti, pjc, tjc, active, T, gi, gjc);
          else tile_entry<DYN_KIND_SS, false, ENERGY, F64>(A, pi, ti, pjc, tjc, active, T, gi, gjc);
          break;
        case DYN_KIND_SO:
          if (alch) tile_entry<DYN_KIND_SO, true, ENERGY, F64>(A, pi, ti, pjc, tjc, active, T, gi, gjc);
          else tile_entry<DYN_KIND_SO, false, ENERGY, F64>(A, pi, ti, pjc, tjc, active, T, gi, gjc);
          break;
        case DYN_KIND_SQ:
          if (alch) tile_entry<DYN_KIND_SQ, true, ENERGY, F64>(A, pi, ti, pjc, tjc, active, T, gi, gjc);
          else tile_entry<DYN_KIND_SQ, false, ENERGY, F64>(A, pi, ti, pjc, tjc, active, T, gi, gjc);
          break;
        default: break;
      }
    }
    // i forces: sum over the 8 j lanes of each i slot (lanes ia, ia+8, …, ia+56)
    for (int c = 0; c < 3; ++c) {
      R f = T.fi[c];
      if constexpr (F64) {
        f = xor_sum_d(f, 8);
        f = xor_sum_d(f, 16);
        f = xor_sum_d(f, 32);
      } else {
        f = xor_sum(f, 8);
        f = xor_sum(f, 16);
        f = xor_sum(f, 32);
      }
      T.fi[c] = f;
    }
    if (jb == 0) {
      // the i forces in fixed point, added to the sites' sums (integer sums:
      // the order of the items does not change them) and, for inspection,
      // kept as the item's row
      int64_t* o = A.fitem + (size_t)item * (3 * DYN_CL_SIZE);
      int64_t q[3];
      for (int c = 0; c < 3; ++c) {
        if constexpr (F64) q[c] = dyn_fx64(T.fi[c]);
        else q[c] = dyn_fx(T.fi[c]);
        if (A.keep_items) o[c * DYN_CL_SIZE + ia] = q[c];
      }
      const int32_t site = A.slot_site[I * DYN_CL_SIZE + ia];
      if (site >= 0)
        for (int c = 0; c < 3; ++c)
          atomicAdd(reinterpret_cast<unsigned long long*>(A.fx_site + (size_t)c * A.n_sites + site),
                    (unsigned long long)q[c]);
    }
    if (!ENERGY) continue;
    double acc[6] = {T.e_lj_w, T.e_lj_s, T.e_c_w, T.e_c_s, T.vir, T.dh};
    for (int k = 0; k < 6; ++k) {
      double x = acc[k];
      for (int m = 1; m < kWave; m <<= 1) x = xor_sum_d(x, m);
      acc[k] = x;
    }
    if (lane == 0) {
      for (int k = 0; k < 5; ++k) A.partials[(size_t)k * A.cap_chunks + item] = acc[k];
      A.partials[(size_t)5 * A.cap_chunks + item] = (double)T.n_ss;
      A.partials[(size_t)6 * A.cap_chunks + item] = (double)T.n_lj;
      A.partials[(size_t)7 * A.cap_chunks + item] = (double)T.n_c;
      A.partials[(size_t)8 * A.cap_chunks + item] = acc[5];
    }
  }
}

// Correction pairs: kind 1 = 1-4 pair (scaled real-space LJ + Coulomb, plus
// the reciprocal correction of the mode), kind 0 = excluded pair (standard
// mode: remove qq·erf(βr)/r from the reciprocal sum).
struct CorrArgs {
  DynPairParams P;
  const float4* wpos;      // per site (x, y, z, q)
  const int32_t* ltype;    // per site type word (DYN_TYPE_ALCH marks the decoupled molecule)
  const float2* lj_table;
  const float2* lj_se;     // (σ, ε) per type pair, for soft-core
  const uint32_t* ci;
  const uint32_t* cj;
  const uint8_t* kind;
  uint32_t n;
  int mode;
  int recip_on;            // reciprocal corrections only accompany the SPME sum
  uint32_t nblk;           // workgroups (one partial each)
  int64_t* mol_e;          // n_mol² molecule-pair energies, or null
  const int32_t* mol_of;
  uint32_t n_mol;
  int64_t* fx;             // site forces, component-major
  uint32_t n_sites;
  double* partials;        // 6 per block: e_lj14, e_c14, e_corr, vir_sr, vir_recip, dH/dλ
};

constexpr int kCorrBlock = 256;
constexpr int kCorrVals = 7;  // per-block sums of the correction kernel
__device__ __forceinline__ void
k_eng_corr(CorrArgs A) {
  __shared__ double red[kCorrBlock];
  if (blockIdx.x >= A.nblk) return;  // partials exist for the system's own blocks
  // e_lj14, e_c14, e_corr, vir_sr, vir_recip, dH/dλ, 1-4 pairs within the cutoff
  double acc[kCorrVals] = {0, 0, 0, 0, 0, 0, 0};
  // molecule-pair matrix: a running entry per thread (1-4 pairs lie within
  // one molecule), flushed when the entry changes and once at the end
  int mkey = -1;
  double me = 0.0;
  for (uint32_t k = blockIdx.x * kCorrBlock + threadIdx.x; k < A.n; k += A.nblk * kCorrBlock) {
    if (A.kind[k] == 0 && !A.recip_on) continue;
    const uint32_t i = A.ci[k], j = A.cj[k];
    const float4 a = A.wpos[i], b = A.wpos[j];
    float d0 = dyn_min_image_rint(a.x - b.x, A.P.ext[0], A.P.inv_ext[0]);
    float d1 = dyn_min_image_rint(a.y - b.y, A.P.ext[1], A.P.inv_ext[1]);
    float d2 = dyn_min_image_rint(a.z - b.z, A.P.ext[2], A.P.inv_ext[2]);
    const float r2 = dyn_fmaf(d2, d2, dyn_fmaf(d1, d1, d0 * d0));
    if (!(r2 > 1e-12f)) continue;
    const float rinv = dyn_rsqrtf(r2);
    const float qq = a.w * b.w;
    float s_sr = 0.f, s_rc = 0.f;
    double e_pair = 0.0;  // the 1-4 pair's energy (molecule-pair matrix)
    if (A.kind[k] == 1) {
      if (r2 < A.P.rc2) acc[6] += 1.0;
      // a 1-4 pair across the decoupled molecule's boundary is also an
      // alchemical pair (src/non_bonded.rs:581-599)
      const bool cross = A.P.alch_on && ((A.ltype[i] ^ A.ltype[j]) & DYN_TYPE_ALCH) != 0;
      const int tab = (A.ltype[i] & DYN_TYPE_MASK) * A.P.n_types + (A.ltype[j] & DYN_TYPE_MASK);
      if (r2 <= A.P.lj_rc2) {
        float sl, el;
        if (cross) {
          const float2 se = A.lj_se[tab];
          float dd;
          dyn_lj_soft_core(r2, se.x, se.y, A.P.sch.lj_lambda, &sl, &el, &dd);
          acc[5] += (double)(DYN_SCALE_LJ_14 * (dd * A.P.sch.lj_dlambda_dlambda));
        } else {
          const float2 c = A.lj_table[tab];
          dyn_lj(rinv * rinv, c.x, c.y, &sl, &el);
        }
        s_sr += DYN_SCALE_LJ_14 * sl;
        acc[0] += (double)(DYN_SCALE_LJ_14 * el);
        e_pair += (double)(DYN_SCALE_LJ_14 * el);
      }
      if (r2 < A.P.rc2) {
        float sc, ec;
        dyn_coul_erfc(r2, rinv, qq, A.P.beta, A.P.beta2, A.P.two_b_sqrtpi, &sc, &ec);
        sc = 0.833333333f * sc;
        ec = 0.833333333f * ec;
        if (cross) {
          acc[5] += (double)(ec * A.P.sch.coulomb_dscale_dlambda);
          sc = sc * A.P.sch.coulomb_scale;
          ec = ec * A.P.sch.coulomb_scale;
        }
        s_sr += sc;
        acc[1] += (double)ec;
        e_pair += (double)ec;
      }
      if (!A.recip_on) {
        // no reciprocal sum, nothing to correct
      } else if (A.mode == DYN_MODE_REFERENCE) {
        // vacuum correction (1/1.2 − 1)·qq/r² along r̂, force only
        s_rc = (DYN_SCALE_COUL_14 - 1.0f) * qq * (rinv * rinv * rinv);
      } else {
        float se, ee;
        dyn_coul_erf(r2, rinv, qq, A.P.beta, A.P.beta2, A.P.two_b_sqrtpi, &se, &ee);
        s_rc = -(1.0f - DYN_SCALE_COUL_14) * se;
        acc[2] += (double)(-(1.0f - DYN_SCALE_COUL_14) * ee);
      }
    } else {
      float se, ee;
      dyn_coul_erf(r2, rinv, qq, A.P.beta, A.P.beta2, A.P.two_b_sqrtpi, &se, &ee);
      s_rc = -se;
      acc[2] += (double)(-ee);
    }
    acc[3] += (double)(s_sr * r2);
    acc[4] += (double)(s_rc * r2);
    if (A.mol_e && e_pair != 0.0) {
      const int mi = A.mol_of[i], mj = A.mol_of[j];
      if (mi != mj) {  // both orders (src/non_bonded.rs:350-361)
        const uint64_t v = (uint64_t)dyn_fx64(e_pair);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.mol_e + mi * (int)A.n_mol + mj), v);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.mol_e + mj * (int)A.n_mol + mi), v);
      } else {
        const int key = mi * (int)A.n_mol + mi;
        if (key != mkey) {
          if (mkey >= 0)
            atomicAdd(reinterpret_cast<unsigned long long*>(A.mol_e + mkey),
                      (unsigned long long)dyn_fx64(me));
          mkey = key;
          me = 0.0;
        }
        me += e_pair;
      }
    }
    const float s = s_sr + s_rc;
    const int64_t q0 = dyn_fx(s * d0), q1 = dyn_fx(s * d1), q2 = dyn_fx(s * d2);
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + i), (unsigned long long)q0);
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + A.n_sites + i), (unsigned long long)q1);
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + 2 * (size_t)A.n_sites + i), (unsigned long long)q2);
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + j), (unsigned long long)(-q0));
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + A.n_sites + j), (unsigned long long)(-q1));
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + 2 * (size_t)A.n_sites + j), (unsigned long long)(-q2));
  }
  if (A.mol_e) wave_add_fx(A.mol_e, mkey, me);
  if (!A.partials) return;  // forces only
  // wave sums by shuffles, then the block's waves in a fixed order
  for (int v = 0; v < kCorrVals; ++v)
    for (int m = kWave / 2; m >= 1; m >>= 1) acc[v] += __shfl_xor(acc[v], m, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0)
    for (int v = 0; v < kCorrVals; ++v) red[v * (kCorrBlock / kWave) + threadIdx.x / kWave] = acc[v];
  __syncthreads();
  if (threadIdx.x < kCorrVals) {
    double t = 0.0;
    for (int w = 0; w < kCorrBlock / kWave; ++w) t += red[threadIdx.x * (kCorrBlock / kWave) + w];
    A.partials[kCorrVals * blockIdx.x + threadIdx.x] = t;
  }
}

// Per-site set-up of an evaluation: clear the site force accumulators and
// gather the charged sites' (x, y, z, q) for the SPME spread.
// With a decoupled molecule (sites [alch0, alch1)), pme_b receives a second
// charge set with that molecule's charges negated.
// The first launch of every evaluation: wrapping, the displacement reading
// that decides on a list build (the largest squared displacement of any
// site since the last build, one atomic per workgroup; squared distances
// are non-negative floats, ordered like their bit patterns), the slot
// positions of the current list (a build rewrites them), the SPME charges
// and their binning, and clearing of the accumulators.  Slots and charges
// take their wrapped positions from pos directly (the same wrap1 of the
// same input as wpos).
__device__ __forceinline__ float4 wrapped(const float4* __restrict__ pos,
                                          const float* __restrict__ q, const EngGeom& g,
                                          uint32_t s) {
  const float4 p = pos[s];
  return make_float4(wrap1(p.x, g.L[0], g.invL[0]), wrap1(p.y, g.L[1], g.invL[1]),
                     wrap1(p.z, g.L[2], g.invL[2]), q[s]);
}

// The list arrays a build writes that the evaluations after it read (the
// current list's, or the staging copies an asynchronous build fills).
struct ListSlots {
  float4* wpos_ref;   // positions of the build (the displacement reference)
  int32_t* slot_site;
  int32_t* slot_type;
  int32_t* slot_gid;
  uint8_t* cl_class;
  uint32_t* slot_of;
  float4* xyzq;
  uint32_t* row_ptr;
  uint32_t* chunk_ptr;
  uint32_t* counters;  // [0] clusters, [1] groups, [2] work items, [3] longest row, [4] overflow
};

struct PrepArgs {
  const float4* pos;
  const float* q;
  EngGeom g;
  float4* wpos;
  ListSlots cur;      // the current list (slot positions refreshed here)
  ListSlots stg;      // the staging list, copied into cur when it holds a finished build
  unsigned int* fl;   // the engine's device flags (kFl*)
  int force;          // the host asks for a build before the pair kernel: a staged one is dropped
  int64_t* fx_site;
  int64_t* mol_e;
  uint32_t n_mol2;
  uint32_t unused;    // no implicit padding: argument packs are compared byte for byte
};

// The SPME charges of an evaluation that runs SPME, by the first launch of
// the second stream (off the main stream's path): their wrapped positions
// (and the second evaluation's charges), their listing under the grid
// bricks, and the cleared SPME force sums.
struct PmePrepArgs {
  const float4* pos;
  const float* q;
  EngGeom g;
  const uint32_t* qsite;
  uint32_t nq;
  uint32_t alch0, alch1;
  uint32_t unused;
  float4* pme_xyzq;
  float4* pme_b;      // null: one SPME evaluation
  int64_t* fx_recip;
  BinArgs bin;
};
__device__ __forceinline__ void k_eng_pme_prep(PmePrepArgs A) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = A.g.n_dyn + 4u * A.g.n_water;
  if (t < n) {
    A.fx_recip[t] = 0;
    A.fx_recip[(size_t)n + t] = 0;
    A.fx_recip[2 * (size_t)n + t] = 0;
  }
  // whole waves past the charges return; every lane of a wave with charges
  // takes part in the binning
  if ((t & ~(uint32_t)(kWave - 1)) >= A.nq) return;
  const bool qv = t < A.nq;
  float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
  if (qv) {
    const uint32_t site = A.qsite[t];
    p = wrapped(A.pos, A.q, A.g, site);
    A.pme_xyzq[t] = p;
    if (A.pme_b)
      A.pme_b[t] = make_float4(p.x, p.y, p.z, site >= A.alch0 && site < A.alch1 ? -p.w : p.w);
  }
  pme_bin_charge(A.bin, p, qv, t);
}

// Copies n values of a staging array into the current list's (grid stride).
template <class T>
__device__ __forceinline__ void copy_grid(T* __restrict__ dst, const T* __restrict__ src, size_t n) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

// The first launch of every evaluation.  A finished staged build (from the
// previous evaluation's second stream) becomes the current list here: its
// slot and row arrays are copied in and the other copy of cj, cmask and
// chunks becomes current.  Then: wrapping, the displacement of every site
// from the current list's build positions (block max, one atomic per
// workgroup; squared distances are non-negative floats, ordered like their
// bit patterns), the slot positions of the current list, the SPME charges
// and their binning, and clearing of the accumulators.  Slots and charges
// take their wrapped positions from pos directly (the same wrap1 of the same
// input as wpos).  The build kernels and the pair kernel decide from the
// largest displacement (BuildGate, k_eng_pairs); the finalize kernel makes
// the taken staged list the current one for good.
__device__ __forceinline__ void k_eng_prep(PrepArgs A) {
  __shared__ unsigned int red[256 / kWave];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = A.g.n_dyn + 4u * A.g.n_water;
  // every workgroup reads the same words and writes the same decision; the
  // finalize kernel clears them
  const bool commit = !A.force && A.fl[kFlReady] != 0u;
  if (t == 0) A.fl[kFlCommit] = commit ? 1u : 0u;
  const ListSlots& L = commit ? A.stg : A.cur;
  unsigned int v = 0u;
  if (t < n) {
    const float4 w = wrapped(A.pos, A.q, A.g, t);
    A.wpos[t] = w;
    const float4 b = L.wpos_ref[t];
    float dx = w.x - b.x, dy = w.y - b.y, dz = w.z - b.z;
    dx -= A.g.L[0] * rintf(dx * A.g.invL[0]);
    dy -= A.g.L[1] * rintf(dy * A.g.invL[1]);
    dz -= A.g.L[2] * rintf(dz * A.g.invL[2]);
    v = __float_as_uint(dx * dx + dy * dy + dz * dz);
    A.fx_site[t] = 0;
    A.fx_site[(size_t)n + t] = 0;
    A.fx_site[2 * (size_t)n + t] = 0;
  }
  if (A.mol_e)
    for (uint32_t k = t; k < A.n_mol2; k += gridDim.x * blockDim.x) A.mol_e[k] = 0;
  const uint32_t n_cl = L.counters[0];
  for (uint32_t k = t; k < n_cl * DYN_CL_SIZE; k += gridDim.x * blockDim.x) {
    const int32_t s = L.slot_site[k];
    if (s >= 0) {
      A.cur.xyzq[k] = wrapped(A.pos, A.q, A.g, (uint32_t)s);
    } else {
      float4 p = wrapped(A.pos, A.q, A.g,
                         (uint32_t)L.slot_site[k & ~(uint32_t)(DYN_CL_SIZE - 1)]);
      p.w = 0.0f;
      A.cur.xyzq[k] = p;
    }
  }
  if (commit) {
    const size_t ns = (size_t)n_cl * DYN_CL_SIZE;
    copy_grid(A.cur.slot_site, A.stg.slot_site, ns);
    copy_grid(A.cur.slot_type, A.stg.slot_type, ns);
    copy_grid(A.cur.slot_gid, A.stg.slot_gid, ns);
    copy_grid(A.cur.cl_class, A.stg.cl_class, n_cl);
    copy_grid(A.cur.slot_of, A.stg.slot_of, n);
    copy_grid(A.cur.wpos_ref, A.stg.wpos_ref, n);
    copy_grid(A.cur.row_ptr, A.stg.row_ptr, (size_t)n_cl + 1);
    copy_grid(A.cur.chunk_ptr, A.stg.chunk_ptr, (size_t)n_cl + 1);
    copy_grid(A.cur.counters, A.stg.counters, 4);
  }
  for (int m = kWave / 2; m >= 1; m >>= 1) v = max(v, (unsigned int)__shfl_xor((int)v, m, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int mx = 0u;
    for (int w = 0; w < 256 / kWave; ++w) mx = max(mx, red[w]);
    if (mx) atomicMax(A.fl + kFlDisp, mx);
  }
}


// Site-order forces: the site accumulators (corrections), the SPME forces
// (this evaluation's or the kept ones of an earlier one) and the pair
// kernel's work-item rows of the site's slot (when the pair kernel ran).
__device__ __forceinline__ void k_eng_finalize(const int64_t* __restrict__ fx_site,
                               const int64_t* __restrict__ fx_recip,
                               const int64_t* __restrict__ fitem,
                               const uint32_t* __restrict__ chunk_ptr, uint64_t cap_chunks,
                               const uint32_t* __restrict__ slot_of, uint32_t n_sites,
                               uint32_t n_slots, int64_t* __restrict__ out,
                               unsigned int* __restrict__ fl, int forced) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) evaluation_flags_end(fl, forced);
  if (s >= n_sites) return;
  ForceParts P;
  P.fx_site = fx_site;
  P.fx_recip = fx_recip;
  P.fitem = fitem;
  P.chunk_ptr = chunk_ptr;
  P.slot_of = slot_of;
  P.cap_chunks = cap_chunks;
  P.n_sites = n_sites;
  P.n_slots = n_slots;
  int64_t f[3];
  site_force_sum(P, s, f);
  for (int c = 0; c < 3; ++c) out[(size_t)c * n_sites + s] = f[c];
}

// Slot-order pair forces (component-major), for inspection.
__device__ __forceinline__ void k_eng_slot_forces(const int64_t* __restrict__ fitem,
                                  const uint32_t* __restrict__ chunk_ptr, uint64_t cap_chunks,
                                  const uint32_t* __restrict__ n_cl, uint32_t n_slots,
                                  int64_t* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_slots || k >= *n_cl * DYN_CL_SIZE) return;
  int64_t f[3] = {0, 0, 0};
  slot_item_forces(fitem, chunk_ptr, cap_chunks, k, f);
  for (int c = 0; c < 3; ++c) out[(size_t)c * n_slots + k] = f[c];
}

// Fixed-order sums of the partials.  out[0..8] from the pair kernel's
// per-item partials (× ½: the full list visits every pair twice): e_lj_w,
// e_lj_s, e_c_w, e_c_s, virial, n_ss, n_lj, n_c, dH/dλ; out[9..14] from the
// correction partials: e_lj14, e_c14, e_corr, vir_sr14, vir_recip_corr, dH/dλ;
// out[16..17] and out[18..19] hold the SPME energy and virial of the first
// and second SPME evaluation (written by SPME itself); those not evaluated
// are cleared here.
constexpr int kRedBlock = 1024;
constexpr int kEngOut = 20;
constexpr int kRed1Grid = 128;   // first-stage workgroups (fixed: part of the sum's order)
constexpr int kRed1Block = 256;

// First stage of the pair-partial sums: workgroup b sums a fixed contiguous
// share of the work items (kNumPart values each) into red1[b].
__device__ __forceinline__ void
k_eng_energy_reduce1(const double* __restrict__ pp, const uint32_t* __restrict__ n_items_dev,
                     uint64_t stride, double* __restrict__ red1) {
  __shared__ double red[kNumPart][kRed1Block / kWave];
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  const uint32_t n_items = (uint32_t)min((uint64_t)*n_items_dev, stride);
  const uint32_t per = (n_items + kRed1Grid - 1) / kRed1Grid;
  const uint32_t r0 = blockIdx.x * per, r1 = min(n_items, r0 + per);
  double a[kNumPart];
  for (int v = 0; v < kNumPart; ++v) a[v] = 0.0;
  for (uint32_t r = r0 + t; r < r1; r += kRed1Block)
    for (int v = 0; v < kNumPart; ++v) a[v] += pp[(size_t)v * stride + r];
  for (int v = 0; v < kNumPart; ++v) {
    double x = a[v];
    for (int m = 1; m < kWave; m <<= 1) x += __shfl_xor(x, m, kWave);
    if (lane == 0) red[v][w] = x;
  }
  __syncthreads();
  if (t < kNumPart) {
    double x = 0.0;
    for (int k = 0; k < kRed1Block / kWave; ++k) x += red[t][k];
    red1[(size_t)blockIdx.x * kNumPart + t] = x;
  }
}

__device__ __forceinline__ void
k_eng_energy_reduce(const double* __restrict__ red1, const double* __restrict__ cp,
                    uint32_t n_cblk, int pme_runs, double* __restrict__ out) {
  // pme_runs < 0: the SPME sum was kept from an earlier evaluation, with its
  // energies, virials and the reciprocal correction virial (out[13])
  constexpr int kVals = kNumPart + kCorrVals;
  __shared__ double red[kVals][kRedBlock / kWave];
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  double a[kVals];
  for (int v = 0; v < kVals; ++v) a[v] = 0.0;
  if (red1)
    for (uint32_t r = t; r < (uint32_t)kRed1Grid; r += kRedBlock)
      for (int v = 0; v < kNumPart; ++v) a[v] += red1[(size_t)r * kNumPart + v];
  for (uint32_t r = t; r < n_cblk; r += kRedBlock)
    for (int v = 0; v < kCorrVals; ++v) a[kNumPart + v] += cp[(size_t)r * kCorrVals + v];
  for (int v = 0; v < kVals; ++v) {
    double x = a[v];
    for (int m = 1; m < kWave; m <<= 1) x += __shfl_xor(x, m, kWave);
    if (lane == 0) red[v][w] = x;
  }
  __syncthreads();
  if (t < kVals) {
    double x = 0.0;
    for (int k = 0; k < kRedBlock / kWave; ++k) x += red[t][k];
    if (!(pme_runs < 0 && t == 13)) out[t] = t < kNumPart ? 0.5 * x : x;
  }
  if (pme_runs >= 0 && t >= 16 && t < kEngOut && (t - 16) / 2 >= pme_runs) out[t] = 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------
// Test hook (dyn_engine_debug_inject_status): sets status bits as a device
// check would, so the error and repeat paths run without a real failure.
__global__ void k_dbg_or_status(unsigned int* status, unsigned int bits) { atomicOr(status, bits); }

struct dyn_engine {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  Profiler prof;
  dyn_engine_params prm{};
  uint32_t n_dyn = 0, n_water = 0, n_sites = 0;
  int n_types = 0;  // table size incl. water O and "none"
  int type_o = 0, type_none = 0;
  float box[3] = {0, 0, 0};
  bool box_set = false, pos_set = false, list_valid = false, computed = false;
  bool energies_valid = false;  // the last compute evaluated energies
  double e_self = 0.0, q_total = 0.0;
  // topology
  DevBuf<float> q;
  DevBuf<int32_t> ltype;
  DevBuf<int32_t> wmol;
  DevBuf<float2> lj_table;   // (c6, c12) per type pair
  DevBuf<float2> lj_se;      // (σ, ε) per type pair (soft-core)
  DevBuf<double2> lj_table64;  // (c6, c12) in double (FP64 pair option)
  DevBuf<float4> coul_tab;     // erfc Coulomb table of the pair kernel (dyn_coul_tab_build)
  std::vector<int32_t> h_type;  // per-site type words (host copy)
  // decoupled molecule: solute sites [alch0, alch1), λ
  uint32_t alch0 = 0, alch1 = 0;
  double lambda = 0.0;
  bool last_recip_b = false;  // the last evaluation ran SPME twice
  DynSchedule last_sch{};
  DevBuf<uint32_t> ex_ptr, ex_idx;
  DevBuf<uint32_t> corr_i, corr_j;
  DevBuf<uint8_t> corr_kind;
  uint32_t n_corr = 0;
  DevBuf<uint32_t> qsite;
  uint32_t n_q = 0;
  float c6_oo = 0.f, c12_oo = 0.f;
  // molecule-pair energy matrix (n_mol² fixed-point sums; n_mol = 0: off)
  uint32_t n_mol = 0;
  bool mol_valid = false;  // mol_e holds the last compute's matrix
  DevBuf<int32_t> mol_of;
  DevBuf<int64_t> mol_e;
  PinnedBuf<int64_t> h_mol;
  // positions
  DevBuf<float4> pos_own;
  const float4* pos_dev = nullptr;
  PinnedBuf<float4> h_pos;
  DevBuf<float4> wpos, wpos_ref;
  // clustering (capacities fixed by the box geometry)
  EngGeom geom{};
  DevBuf<int> col_class;
  DevBuf<uint32_t> col_count, col_cl_start, items;
  uint32_t col_slice = 0;  // entries of items per column (a column's sites, unsorted)
  uint32_t n_cols = 0, cap_cl = 0, cap_slots = 0;
  DevBuf<int32_t> slot_site, slot_type, slot_gid;
  DevBuf<uint8_t> cl_class;
  DevBuf<uint32_t> slot_of;
  DevBuf<float4> xyzq, bb_c, bb_h;
  DevBuf<uint32_t> row_cnt, row_ptr, chunk_ptr, cj, ent_tmp, unit_cnt;
  DevBuf<uint8_t> xb_tmp;  // exclusion byte of each ent_tmp entry
  DevBuf<unsigned long long> cmask;
  DevBuf<uint2> chunks;
  uint64_t cap_entries = 0, cap_chunks = 0;  // groups of 8 j entries, work items
  uint32_t row_cap = 0;                      // j entries per walk-unit slice of ent_tmp
  // device counters: [0] n_cl, [1] n_groups, [2] n_items, [3] max row (groups),
  // [4] longest row that overflowed its slice (entries)
  DevBuf<uint32_t> counters;
  // the staging list of an asynchronous build (slot and row arrays, copied
  // into the ones above by the next evaluation's preparation kernel) and the
  // second copy of the per-group arrays (dflags[kFlSel] names the current)
  DevBuf<float4> stg_wpos_ref, stg_xyzq;
  DevBuf<int32_t> stg_slot_site, stg_slot_type, stg_slot_gid;
  DevBuf<uint8_t> stg_cl_class;
  DevBuf<uint32_t> stg_slot_of, stg_row_ptr, stg_chunk_ptr, stg_counters;
  DevBuf<uint32_t> cj2;
  DevBuf<unsigned long long> cmask2;
  DevBuf<uint2> chunks2;
  ListSlots slots_cur() {
    return ListSlots{wpos_ref.p, slot_site.p, slot_type.p, slot_gid.p, cl_class.p, slot_of.p,
                     xyzq.p, row_ptr.p, chunk_ptr.p, counters.p};
  }
  ListSlots slots_stg() {
    return ListSlots{stg_wpos_ref.p, stg_slot_site.p, stg_slot_type.p, stg_slot_gid.p,
                     stg_cl_class.p, stg_slot_of.p, stg_xyzq.p, stg_row_ptr.p, stg_chunk_ptr.p,
                     stg_counters.p};
  }
  ListCopies copies(bool staged) {
    return ListCopies{cj.p, cj2.p, cmask.p, cmask2.p, chunks.p, chunks2.p, dflags.p,
                      staged ? 1 : 0, 0};
  }
  DevBuf<unsigned int> status;
  // device flags (kFlWords words, kFl* indices)
  DevBuf<unsigned int> dflags;
  // accumulators
  DevBuf<int64_t> fitem, fx_site, fx_out, fx_slot;  // fx_slot: inspection only
  DevBuf<double> pair_part, corr_part, eng_out, red1;
  unsigned corr_blocks = 0;
  // a pass over this engine alone: argument packs of both streams
  ArgCache own_cache, side_cache;
  // PME
  bool pme_on = false;
  PmeCore pme;
  DevBuf<float4> pme_xyzq, pme_b;  // charge sets of the two SPME evaluations
  DevBuf<int64_t> fx_recip;        // SPME forces of the last evaluation that ran it
  bool recip_kept = false;         // fx_recip and its energies are those of this box
  // host copies
  PinnedBuf<int64_t> h_fx;
  PinnedBuf<double> h_eng;
  // [0] status, [1..5] counters, [6] builds, [7] current copy, [8..12] staging counters
  PinnedBuf<uint32_t> h_misc;
  bool fx_fresh = false;
  // the last evaluation's force parts; fx_stale: enqueued with
  // kComputeFusedFinalize, so fx_out does not hold its forces
  ForceParts parts{};
  bool fx_stale = false;

  hipStream_t own_stream = nullptr;  // the engine's stream (a batch may lend it another)
  hipStream_t side_stream = nullptr;  // its second stream (SPME), same priority
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // the forces, SPME forces and energies of the last evaluation, kept at the
  // start of an MD call so that a repeat or a failed call restores them
  DevBuf<int64_t> fx_out_ck, fx_recip_ck;
  DevBuf<double> eng_ck;
  struct {
    bool recip_kept, last_recip_b, computed, energies_valid;
    DynSchedule last_sch;
  } ck{};
  // test hook: after dbg_skip more evaluations, the next dbg_count ones have
  // dbg_bits ORed into the status word before the pair kernel
  unsigned dbg_bits = 0;
  uint32_t dbg_skip = 0, dbg_count = 0;

  ~dyn_engine() {
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side_stream) (void)hipStreamDestroy(side_stream);
    if (own_stream) (void)hipStreamDestroy(own_stream);
  }

  void set_geometry();
  void reserve_rows(uint32_t entries);
  void reserve_list(uint64_t groups, uint64_t items);
  void check_status();
  EngineSet self_set() {
    EngineSet S;
    S.e = {this};
    S.R.stream = stream;
    S.R.cache = &own_cache;
    S.R.n = 1;
    // fully profiled evaluations stay on one stream (their per-kernel times)
    if ((!prof.enabled || prof.only) && stream == own_stream && !prm.single_stream) {
      S.side.stream = side_stream;
      S.side.cache = &side_cache;
      S.side.n = 1;
      S.ev_fork = ev_fork;
      S.ev_join = ev_join;
    }
    S.prof = &prof;
    return S;
  }
};

void dyn_engine::set_geometry() {
  const double V = (double)box[0] * box[1] * box[2];
  const uint32_t ncls[3] = {n_dyn, n_water, 3 * n_water};
  int off = 0, site_base = 0;
  std::vector<int> h_col_class;
  for (int c = 0; c < 3; ++c) {
    ClassGrid& g = geom.cg[c];
    g.n = (int)ncls[c];
    g.site_base = site_base;
    site_base += g.n;
    g.col_off = off;
    if (ncls[c] == 0) {
      g.ncx = g.ncy = 0;
      g.wx = g.wy = 1.f;
      continue;
    }
    const double s = std::cbrt(8.0 * V / (double)ncls[c]);
    g.ncx = std::max(1, std::min(64, (int)std::floor(box[0] / s)));
    g.ncy = std::max(1, std::min(64, (int)std::floor(box[1] / s)));
    g.wx = box[0] / g.ncx;
    g.wy = box[1] / g.ncy;
    off += g.ncx * g.ncy;
    for (int k = 0; k < g.ncx * g.ncy; ++k) h_col_class.push_back(c);
  }
  geom.n_cols = off;
  geom.n_dyn = n_dyn;
  geom.n_water = n_water;
  n_cols = (uint32_t)off;
  for (int a = 0; a < 3; ++a) {
    geom.L[a] = box[a];
    geom.invL[a] = 1.0f / box[a];
  }
  // clusters: Σ ceil(count/8) ≤ n_sites/8 + n_cols
  cap_cl = n_sites / DYN_CL_SIZE + n_cols + 1;
  cap_slots = cap_cl * DYN_CL_SIZE;
  // the walk's exclusion tables key a slot in 24 bits
  DYN_REQUIRE(cap_slots + DYN_CL_SIZE < (1u << 24), "too many sites (%u cluster slots; at most "
              "16,777,208)", cap_slots);
  col_class.reserve(std::max<uint32_t>(n_cols, 1));
  col_count.reserve(n_cols + 1);
  col_cl_start.reserve(n_cols + 1);
  // a column holds 8·Lz/s sites of its class on average (s: the edge of a
  // cube of 8 of them); its slice takes three times that, at least 64
  {
    double most = 0.0;
    for (int c = 0; c < 3; ++c)
      if (geom.cg[c].n) most = std::max(most, (double)geom.cg[c].n / (geom.cg[c].ncx * geom.cg[c].ncy));
    const uint32_t want = (uint32_t)std::min<double>(kColCap, 3.0 * most + 64.0);
    col_slice = std::max(col_slice, (want + 63u) & ~63u);
    col_slice = std::min<uint32_t>(col_slice, kColCap);
  }
  items.reserve((size_t)n_cols * col_slice);
  // slot arrays hold one more cluster: the null slot (index cap_slots) pads
  // the j groups of the pair list
  slot_site.reserve(cap_slots + DYN_CL_SIZE);
  slot_type.reserve(cap_slots + DYN_CL_SIZE);
  slot_gid.reserve(cap_slots + DYN_CL_SIZE);
  cl_class.reserve(cap_cl);
  xyzq.reserve(cap_slots + DYN_CL_SIZE);
  bb_c.reserve(cap_cl);
  bb_h.reserve(cap_cl);
  row_cnt.reserve(cap_cl + 1);
  unit_cnt.reserve((size_t)cap_cl * kUnits);
  row_ptr.reserve(cap_cl + 1);
  chunk_ptr.reserve(cap_cl + 1);
  stg_slot_site.reserve(cap_slots + DYN_CL_SIZE);
  stg_slot_type.reserve(cap_slots + DYN_CL_SIZE);
  stg_slot_gid.reserve(cap_slots + DYN_CL_SIZE);
  stg_cl_class.reserve(cap_cl);
  stg_xyzq.reserve(cap_slots + DYN_CL_SIZE);
  stg_row_ptr.reserve(cap_cl + 1);
  stg_chunk_ptr.reserve(cap_cl + 1);
  {
    const int32_t none = type_none, pad = DYN_GID_PAD, nosite = -1;
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    DYN_HIP(hipMemcpyAsync(slot_site.p + cap_slots, &nosite, 4, hipMemcpyHostToDevice, stream));
    DYN_HIP(hipMemcpyAsync(slot_type.p + cap_slots, &none, 4, hipMemcpyHostToDevice, stream));
    DYN_HIP(hipMemcpyAsync(slot_gid.p + cap_slots, &pad, 4, hipMemcpyHostToDevice, stream));
    DYN_HIP(hipMemcpyAsync(xyzq.p + cap_slots, &zero, 16, hipMemcpyHostToDevice, stream));
    DYN_HIP(hipStreamSynchronize(stream));
  }
  // first guesses from the site density, grown on overflow: a row holds the
  // sites within about rlist + 3 Å of its cluster (all classes, padding); a
  // walk unit (half of one class's window) well under 0.6 of that
  const double dens = n_sites / V;
  const double rl = prm.cutoff + prm.list_buffer;
  const double row_est = dens * 4.19 * std::pow(rl + 3.0, 3) * 1.2 + 3 * DYN_CL_SIZE;
  reserve_rows((uint32_t)std::min<double>(0.6 * row_est, 1 << 20));
  const double avg_groups = dens * 4.19 * std::pow(rl + 2.0, 3) / DYN_CL_SIZE + 3;
  const uint64_t groups = (uint64_t)(cap_cl * avg_groups);
  reserve_list(groups, groups / prm.item_groups + cap_cl);
  DYN_HIP(hipMemcpyAsync(col_class.p, h_col_class.data(), sizeof(int) * n_cols,
                         hipMemcpyHostToDevice, stream));
  // the binning kernel counts into zeroed column counts; the list walk
  // clears them again after use
  DYN_HIP(hipMemsetAsync(col_count.p, 0, sizeof(uint32_t) * (n_cols + 1), stream));
  // no clusters until the next build: the preparation kernel gathers slot
  // positions only for the clusters of a built list
  if (counters.p) DYN_HIP(hipMemsetAsync(counters.p, 0, 4, stream));
  DYN_HIP(hipStreamSynchronize(stream));
  list_valid = false;
}

// Unit slices of the list walk: `entries` j entries per unit (a multiple
// of 64), kUnits units per cluster.
void dyn_engine::reserve_rows(uint32_t entries) {
  row_cap = std::max(row_cap, (entries + 63u) & ~63u);
  ent_tmp.reserve((size_t)cap_cl * kUnits * row_cap);  // grows with either factor
  xb_tmp.reserve((size_t)cap_cl * kUnits * row_cap);
}

// List capacity in groups of 8 j entries, and work items.
void dyn_engine::reserve_list(uint64_t groups, uint64_t n_items) {
  if (groups > cap_entries) {
    cj.reserve(groups * DYN_CL_SIZE);
    cmask.reserve(groups);
    cj2.reserve(groups * DYN_CL_SIZE);
    cmask2.reserve(groups);
    cap_entries = groups;
  }
  if (n_items > cap_chunks) {
    chunks.reserve(n_items);
    chunks2.reserve(n_items);
    fitem.reserve((size_t)3 * DYN_CL_SIZE * n_items);
    pair_part.reserve((size_t)kNumPart * n_items);
    cap_chunks = n_items;
  }
}

namespace dyn {

// Largest per-system grid of a launch.
template <class Fn>
static unsigned max_grid(const EngineSet& S, Fn&& fn) {
  unsigned g = 1;
  for (dyn_engine* e : S.e) g = std::max(g, (unsigned)fn(*e));
  return g;
}

// Bits of the squared displacement at which an engine rebuilds its list:
// half the list buffer (every pair now within the cutoff was then within
// the list radius), less a margin for the rounding of the f32 readings.
static unsigned int build_limit(const dyn_engine& e) {
  const float half = 0.5f * e.prm.list_buffer * (1.0f - 1e-4f);
  const float h2 = half * half;
  unsigned int b;
  std::memcpy(&b, &h2, 4);
  return b;
}

// A staged build starts when the largest displacement is at least this far
// (or twice its recent growth per evaluation, if more) below half the list
// buffer: the staged list becomes current one evaluation later.
constexpr float kStageMargin = 0.25f;

// The gate of engine e's build kernels in this evaluation.
static BuildGate build_gate(const dyn_engine& e, bool force) {
  BuildGate G;
  std::memset(&G, 0, sizeof G);
  G.fl = e.dflags.p;
  G.limit = build_limit(e);
  G.half = 0.5f * e.prm.list_buffer;
  G.margin_min = std::min(kStageMargin, 0.2f * G.half);
  G.staged = 0;
  G.force = (force || e.prm.list_buffer <= 0.f) ? 1 : 0;
  return G;
}

// Cluster-pair list build of every engine of the set; the kernels of an
// engine whose gate is closed return at once.
// The list build of every engine of the set, on the run R: into the current
// list (before the pair kernel) or, staged, into the staging list and the
// other copy of the per-group arrays (beside the pair kernel, current from
// the next evaluation on).  The kernels of an engine whose gate is closed
// return at once.
static void set_build_lists(EngineSet& S, const BatchRun& R, const std::vector<BuildGate>& gate,
                            bool staged) {
  auto& E = S.e;
  Profiler& prof = *S.prof;
  std::vector<ListSlots> T(E.size());
  for (size_t i = 0; i < E.size(); ++i) T[i] = staged ? E[i]->slots_stg() : E[i]->slots_cur();
  {
    Profiler::Scope ps(prof, "nbl_bin");
    const unsigned gs = max_grid(S, [](dyn_engine& e) {
      return std::min<uint32_t>(grid_for(e.n_sites, kBinBlock), kBuildGrid);
    });
    launch<k_eng_bin_count, kBinBlock>(R, "k_eng_bin_count", 0, gs, kBinBlock, 0, [&](uint32_t i) {
      dyn_engine& e = *E[i];
      return args<k_eng_bin_count>(gate[i], e.wpos.p, e.geom, T[i].wpos_ref, e.col_count.p,
                                   e.items.p, e.dflags.p + kFlBuilds, e.col_slice);
    });
  }
  {
    Profiler::Scope ps(prof, "nbl_colsort");
    const unsigned gc = max_grid(S, [](dyn_engine& e) { return std::min(e.n_cols, kBuildGrid); });
    launch<k_eng_colsort, kSortBlock>(R, "k_eng_colsort", 0, gc, kSortBlock, 0, [&](uint32_t i) {
      dyn_engine& e = *E[i];
      SlotArgs A;
      std::memset(&A, 0, sizeof A);
      A = SlotArgs{e.wpos.p, e.ltype.p, e.wmol.p, T[i].xyzq, T[i].slot_type,
                   T[i].slot_gid, T[i].slot_site, T[i].slot_of, e.type_none, 0};
      return args<k_eng_colsort>(gate[i], e.col_count.p, e.items.p, e.col_slice,
                                 e.col_cl_start.p, e.col_class.p, A, T[i].cl_class, e.bb_c.p,
                                 e.bb_h.p, e.status.p, e.n_cols, T[i].counters + 0);
    });
  }
  std::vector<ListArgs> LA(E.size());
  for (size_t i = 0; i < E.size(); ++i) {
    dyn_engine& e = *E[i];
    ListArgs& A = LA[i];
    std::memset(&A, 0, sizeof A);
    A.gate = gate[i];
    A.g = e.geom;
    // FP64 pair arithmetic decides the cutoff in double: 0.01 Å of list
    // margin keeps every pair the f32 distance puts just outside
    A.rlist = e.prm.cutoff + std::max(0.f, e.prm.list_buffer) +
              ((e.prm.precision & DYN_PREC_PAIR_F64) ? 0.01f : 0.f);
    A.rlist2 = A.rlist * A.rlist;
    A.n_cl = T[i].counters + 0;
    A.bb_c = e.bb_c.p;
    A.bb_h = e.bb_h.p;
    A.col_cl_start = e.col_cl_start.p;
    A.cl_class = T[i].cl_class;
    A.slot_site = T[i].slot_site;
    A.xyzq = T[i].xyzq;
    A.slot_gid = T[i].slot_gid;
    A.slot_of = T[i].slot_of;
    A.ex_ptr = e.ex_ptr.p;
    A.ex_idx = e.ex_idx.p;
    A.n_dyn = e.n_dyn;
    A.cap_entries = e.cap_entries;
    A.cap_chunks = e.cap_chunks;
    A.row_cap = e.row_cap;
    A.chunk = e.prm.item_groups;
    A.null_slot = e.cap_slots;
    A.longest = T[i].counters + 4;
    A.status = e.status.p;
    A.xb_tmp = e.xb_tmp.p;
    A.unit_cnt = e.unit_cnt.p;
    A.col_count = e.col_count.p;
    A.n_cols = e.n_cols;
    A.tot = T[i].counters + 1;
  }
  // the walk: one unit per wave; the fill: kFillRows rows per workgroup
  const unsigned walk_grid = max_grid(S, [](dyn_engine& e) {
    return std::min<uint32_t>((e.cap_cl * kUnits + kListWaves - 1) / kListWaves, kWalkGrid);
  });
  const unsigned list_grid = max_grid(S, [](dyn_engine& e) {
    return std::min<uint32_t>((e.cap_cl + kFillRows - 1) / kFillRows, kWalkGrid);
  });
  {
    Profiler::Scope ps(prof, "nbl_list_walk");
    launch<k_eng_list_walk, kListBlock>(R, "k_eng_list_walk", 0, walk_grid, kListBlock, 0,
                                        [&](uint32_t i) {
      dyn_engine& e = *E[i];
      return args<k_eng_list_walk>(LA[i], e.ent_tmp.p);
    });
  }
  {
    Profiler::Scope ps(prof, "nbl_list_fill");
    launch<k_eng_list_fill, kListBlock>(R, "k_eng_list_fill", 0, list_grid, kListBlock, 0,
                                        [&](uint32_t i) {
      dyn_engine& e = *E[i];
      return args<k_eng_list_fill>(LA[i], e.ent_tmp.p, e.row_cnt.p, T[i].row_ptr,
                                   T[i].chunk_ptr, e.copies(staged));
    });
  }
}

void engine_set_enqueue(EngineSet& S, uint32_t flags,
                        const std::function<void(const BatchRun&)>& side_work) {
  auto& E = S.e;
  const uint32_t ns = (uint32_t)E.size();
  const BatchRun& R = S.R;
  const BatchRun& RS = S.side.stream ? S.side : S.R;
  Profiler& prof = *S.prof;
  // host-forced builds: asked for, or no list (never built, a new box,
  // positions set from the host, a failed evaluation); the device decides
  // for the others from their displacement readings.  Staged mode: the
  // others build a staging list beside the pair kernel when their sites
  // have moved (buffer/2 − kStageMargin) since their list's build, and their
  // evaluation is void if a site has moved buffer/2; otherwise the build
  // runs before the pair kernel when a site has moved buffer/2.
  const bool staged_mode = (flags & kComputeStagedBuild) != 0;
  std::vector<BuildGate> gate(ns), sgate(ns);
  std::vector<char> forced(ns);
  bool any_inline = false, any_staged = false;
  for (uint32_t i = 0; i < ns; ++i) {
    dyn_engine* e = E[i];
    DYN_REQUIRE(e->box_set, "box not set");
    if (!e->pos_set) throw Error(DYN_ERR_STATE, "positions not set");
    forced[i] = (flags & DYN_COMPUTE_REBUILD) || !e->list_valid || e->prm.list_buffer <= 0.f;
    gate[i] = build_gate(*e, forced[i]);
    sgate[i] = gate[i];
    sgate[i].staged = 1;
    sgate[i].force = 0;
    if (staged_mode && !forced[i]) {
      gate[i].limit = ~0u;  // no build before the pair kernel: squared distances stay below
      any_staged = true;
    } else {
      sgate[i].half = std::numeric_limits<float>::infinity();  // closed: distances stay below
    }
    any_inline = any_inline || !staged_mode || forced[i];
  }
  const bool short_on = (flags & DYN_COMPUTE_SHORT) != 0;
  const bool energy_on = (flags & DYN_COMPUTE_NO_ENERGY) == 0;
  // per system: pair parameters and which reciprocal work it does
  struct Plan {
    DynPairParams P;
    bool alch_on, reuse, recip_on, recip_b, corr_on;
  };
  std::vector<Plan> PL(ns);
  bool any_recip = false, any_b = false, any_corr = false;
  for (uint32_t i = 0; i < ns; ++i) {
    dyn_engine& e = *E[i];
    Plan& pl = PL[i];
    DynPairParams& P = pl.P;
    std::memset(&P, 0, sizeof P);
    for (int a = 0; a < 3; ++a) {
      P.ext[a] = e.box[a];
      P.inv_ext[a] = 1.0f / e.box[a];
    }
    P.rc2 = e.prm.cutoff * e.prm.cutoff;
    P.lj_rc2 = e.prm.lj_cutoff * e.prm.lj_cutoff;
    P.beta = e.prm.beta;
    P.beta2 = e.prm.beta * e.prm.beta;
    P.two_b_sqrtpi = 2.0f * e.prm.beta * DYN_INV_SQRT_PI;
    P.c6_oo = e.c6_oo;
    P.c12_oo = e.c12_oo;
    P.n_types = e.n_types;
    pl.alch_on = e.alch1 > e.alch0;
    P.alch_on = pl.alch_on ? 1 : 0;
    P.sch = dyn_schedule((float)e.lambda);
    P.coul_tab = e.coul_tab.p;
    // the reference's SPME cadence: keep the last SPME forces, energy and
    // virial (reference mode, no decoupled molecule, same box)
    pl.reuse = e.pme_on && (flags & DYN_COMPUTE_RECIP_REUSE) && !(flags & DYN_COMPUTE_RECIP) &&
               e.recip_kept && !pl.alch_on && e.prm.mode == DYN_MODE_REFERENCE;
    pl.recip_on = e.pme_on && ((flags & DYN_COMPUTE_RECIP) ||
                               ((flags & DYN_COMPUTE_RECIP_REUSE) && !pl.reuse));
    pl.corr_on = (flags & DYN_COMPUTE_CORR) && e.n_corr;
    // with a decoupled molecule the SPME sum runs twice, with its charges as
    // they are (A) and negated (B): ((1+s)A + (1−s)B)/2 keeps the
    // environment and molecule self terms and scales their cross term by s
    pl.recip_b = pl.recip_on && pl.alch_on;
    if (!pl.reuse) {
      e.last_recip_b = pl.recip_b;
      e.last_sch = P.sch;
    }
    any_recip = any_recip || pl.recip_on;
    any_b = any_b || pl.recip_b;
    any_corr = any_corr || pl.corr_on;
  }
  {
    Profiler::Scope ps(prof, "prep");
    const unsigned gp = max_grid(S, [](dyn_engine& e) {
      return std::max(grid_for(e.n_sites, 256), grid_for(e.cap_slots, 256));
    });
    launch<k_eng_prep>(R, "k_eng_prep", 0, gp, 256, 0, [&](uint32_t i) {
      dyn_engine& e = *E[i];
      const Plan& pl = PL[i];
      PrepArgs A;
      std::memset(&A, 0, sizeof A);
      A.pos = e.pos_dev;
      A.q = e.q.p;
      A.g = e.geom;
      A.wpos = e.wpos.p;
      A.cur = e.slots_cur();
      A.stg = e.slots_stg();
      A.fl = e.dflags.p;
      A.force = forced[i] ? 1 : 0;
      A.fx_site = e.fx_site.p;
      A.mol_e = energy_on && e.n_mol ? e.mol_e.p : nullptr;
      A.n_mol2 = e.n_mol * e.n_mol;
      return args<k_eng_prep>(A);
    });
  }
  // the second stream: SPME (spread → FFT → convolution → inverse FFT →
  // gather), the exclusion and 1-4 corrections and the caller's side work,
  // overlapping the list build and the pair kernel on the main stream (a
  // fork later in the step delays the side work more than it gains: an
  // event wait between streams takes 10–40 µs to release)
  stream_order(R.stream, RS.stream, S.ev_fork);
  // SPME: beside the list build and the pair kernel on the second stream,
  // or — with staged builds, where the pair kernel follows the preparation
  // kernel at once and would take the compute units the FFT planes need —
  // before the pair kernel on the main stream
  const BatchRun& RP = staged_mode ? R : RS;
  if (any_recip) {
    {
      Profiler::Scope ps(prof, "pme_prep");
      const unsigned gq = max_grid(S, [](dyn_engine& e) {
        return std::max(grid_for(e.n_sites, 256), grid_for(e.n_q, 256));
      });
      launch<k_eng_pme_prep>(RP, "k_eng_pme_prep", 0, gq, 256, 0, [&](uint32_t i) {
        dyn_engine& e = *E[i];
        const Plan& pl = PL[i];
        PmePrepArgs A;
        std::memset(&A, 0, sizeof A);
        A.pos = e.pos_dev;
        A.q = e.q.p;
        A.g = e.geom;
        A.qsite = e.qsite.p;
        A.alch0 = e.alch0;
        A.alch1 = e.alch1;
        A.pme_xyzq = e.pme_xyzq.p;
        A.pme_b = pl.recip_b ? e.pme_b.p : nullptr;
        A.fx_recip = e.fx_recip.p;
        if (pl.recip_on) {  // the SPME charges are listed under their bricks here
          e.pme.ensure_cap(e.n_q);
          A.nq = e.n_q;
          A.bin = e.pme.bin_args(e.status.p, kStPme);
        } else {
          // a system of the batch without SPME this evaluation: nothing to do
          A.g.n_dyn = 0;
          A.g.n_water = 0;
        }
        return args<k_eng_pme_prep>(A);
      });
    }
    std::vector<PmeRun> ra(ns), rb(ns);
    for (uint32_t i = 0; i < ns; ++i) {
      dyn_engine& e = *E[i];
      const Plan& pl = PL[i];
      if (!pl.recip_on) continue;
      const float s = pl.recip_b ? pl.P.sch.coulomb_scale : 1.0f;
      PmeRun r;
      r.core = &e.pme;
      r.n = e.n_q;
      r.xyzq = e.pme_xyzq.p;
      r.site = e.qsite.p;
      r.fx = e.fx_recip.p;
      r.comp_major = 3;  // component-major, one charge per site
      r.nsites = e.n_sites;
      r.fscale = pl.recip_b ? 0.5f * (1.0f + s) : 1.0f;
      r.res = e.eng_out.p + 16;
      r.status = e.status.p;
      r.status_bit = kStPme;
      r.binned = true;  // k_eng_pme_prep has listed the charges
      ra[i] = r;
      r.binned = false;
      if (pl.recip_b) {
        r.xyzq = e.pme_b.p;
        r.fscale = 0.5f * (1.0f - s);
        r.res = e.eng_out.p + 18;
        rb[i] = r;
      }
    }
    pme_run(RP, prof, ra.data(), 0);
    if (any_b) pme_run(RP, prof, rb.data(), 1);
  }
  // the exclusion and 1-4 corrections (site accumulators; finalize reads them)
  if (any_corr) {
    Profiler::Scope ps(prof, "corr");
    const unsigned gc = max_grid(S, [](dyn_engine& e) { return e.corr_blocks; });
    launch<k_eng_corr, kCorrBlock>(RS, "k_eng_corr", 0, gc, kCorrBlock, 0, [&](uint32_t i) {
      dyn_engine& e = *E[i];
      CorrArgs C;
      std::memset(&C, 0, sizeof C);
      C.P = PL[i].P;
      C.wpos = e.wpos.p;
      C.ltype = e.ltype.p;
      C.lj_table = e.lj_table.p;
      C.lj_se = e.lj_se.p;
      C.ci = e.corr_i.p;
      C.cj = e.corr_j.p;
      C.kind = e.corr_kind.p;
      C.n = PL[i].corr_on ? e.n_corr : 0u;
      C.mode = (int)e.prm.mode;
      C.recip_on = PL[i].recip_on ? 1 : 0;
      C.nblk = e.corr_blocks;
      C.fx = e.fx_site.p;
      C.n_sites = e.n_sites;
      C.partials = energy_on ? e.corr_part.p : nullptr;
      C.mol_e = energy_on && e.n_mol ? e.mol_e.p : nullptr;
      C.mol_of = e.mol_of.p;
      C.n_mol = e.n_mol;
      return args<k_eng_corr>(C);
    });
  }
  if (side_work) side_work(RS);
  // a staged build on the second stream, beside the pair kernel (after it
  // when there is one stream)
  const bool one_stream = RS.stream == R.stream;
  if (any_staged && !one_stream) set_build_lists(S, RS, sgate, true);
  if (any_inline) set_build_lists(S, R, gate, false);
  for (uint32_t i = 0; i < ns; ++i) {
    dyn_engine* e = E[i];
    if (!e->dbg_count) continue;
    if (e->dbg_skip) {
      --e->dbg_skip;
      continue;
    }
    --e->dbg_count;
    hipLaunchKernelGGL(k_dbg_or_status, dim3(1), dim3(1), 0, R.stream, e->status.p, e->dbg_bits);
    DYN_CHECK_LAUNCH("k_dbg_or_status");
  }
  if (short_on) {
    Profiler::Scope ps(prof, "pairs");
    std::vector<PairArgs> PA(ns);
    for (uint32_t i = 0; i < ns; ++i) {
      dyn_engine& e = *E[i];
      PairArgs& A = PA[i];
      std::memset(&A, 0, sizeof A);
      A.P = PL[i].P;
      A.xyzq = e.xyzq.p;
      A.slot_type = e.slot_type.p;
      A.slot_gid = e.slot_gid.p;
      A.cl_class = e.cl_class.p;
      A.row_ptr = e.row_ptr.p;
      A.cj = e.cj.p;
      A.cmask = e.cmask.p;
      A.chunks = e.chunks.p;
      A.cj2 = e.cj2.p;
      A.cmask2 = e.cmask2.p;
      A.chunks2 = e.chunks2.p;
      A.fl = e.dflags.p;
      A.staged = staged_mode && !forced[i] ? 1 : 0;
      A.hard_limit = build_limit(e);
      A.n_chunks = e.counters.p + 2;
      A.lj_table = e.lj_table.p;
      A.lj_se = e.lj_se.p;
      A.lj_table64 = e.lj_table64.p;
      A.mol_e = energy_on && e.n_mol ? e.mol_e.p : nullptr;
      A.mol_of = e.mol_of.p;
      A.n_mol = e.n_mol;
      A.fitem = e.fitem.p;
      A.slot_site = e.slot_site.p;
      A.fx_site = e.fx_site.p;
      A.n_sites = e.n_sites;
      // a fused evaluation's caller reads site sums only
      A.keep_items = (flags & kComputeFusedFinalize) ? 0u : 1u;
      A.partials = e.pair_part.p;
      A.cap_chunks = e.cap_chunks;
      A.chunk = e.prm.item_groups;
      A.status = e.status.p;
    }
    // a wave per work item, up to kPairGridCap workgroups (the item count is
    // known on the device only; the waves take items in a grid stride)
    const unsigned gp = max_grid(S, [](dyn_engine& e) {
      return (unsigned)std::min<uint64_t>(kPairGridCap, std::max<uint64_t>(1, (e.cap_chunks + 3) / 4));
    });
    // FP64 pair arithmetic when the first system asks for it (a set shares it)
    const bool f64 = (E[0]->prm.precision & DYN_PREC_PAIR_F64) != 0;
    int nt_max = 0;
    for (dyn_engine* e : E) nt_max = std::max(nt_max, e->n_types);
    const size_t shm = pair_lds_bytes(nt_max, energy_on);
    auto pa = [&](uint32_t i) { return PA[i]; };
    if (energy_on && !f64)
      launch<k_eng_pairs<true, false>, kPairBlock>(R, "k_eng_pairs<true>", 0, gp, kPairBlock, shm,
          [&](uint32_t i) { return args<k_eng_pairs<true, false>>(pa(i)); });
    else if (!f64)
      launch<k_eng_pairs<false, false>, kPairBlock>(R, "k_eng_pairs<false>", 0, gp, kPairBlock, shm,
          [&](uint32_t i) { return args<k_eng_pairs<false, false>>(pa(i)); });
    else if (energy_on)
      launch<k_eng_pairs<true, true>, kPairBlock>(R, "k_eng_pairs<true, f64>", 0, gp, kPairBlock, 0,
          [&](uint32_t i) { return args<k_eng_pairs<true, true>>(pa(i)); });
    else
      launch<k_eng_pairs<false, true>, kPairBlock>(R, "k_eng_pairs<false, f64>", 0, gp, kPairBlock, 0,
          [&](uint32_t i) { return args<k_eng_pairs<false, true>>(pa(i)); });
  }
  if (any_staged && one_stream) set_build_lists(S, R, sgate, true);
  stream_order(RS.stream, R.stream, S.ev_join);
  const bool fused = (flags & kComputeFusedFinalize) != 0;
  for (uint32_t i = 0; i < ns; ++i) {
    dyn_engine& e = *E[i];
    const Plan& pl = PL[i];
    ForceParts& P = e.parts;
    std::memset(&P, 0, sizeof P);
    P.fx_site = e.fx_site.p;
    P.fx_recip = (pl.recip_on || pl.reuse) ? e.fx_recip.p : nullptr;
    P.fitem = nullptr;  // the pair kernel adds its i forces to fx_site
    P.chunk_ptr = e.chunk_ptr.p;
    P.slot_of = e.slot_of.p;
    P.cap_chunks = e.cap_chunks;
    P.n_sites = e.n_sites;
    P.n_slots = e.cap_slots;
    P.fl = e.dflags.p;
    P.forced = forced[i] ? 1 : 0;
    e.fx_stale = fused;
  }
  {
    Profiler::Scope ps(prof, "finalize");
    const unsigned gs = max_grid(S, [](dyn_engine& e) { return grid_for(e.n_sites, 256); });
    if (!fused) launch<k_eng_finalize>(R, "k_eng_finalize", 0, gs, 256, 0, [&](uint32_t i) {
      dyn_engine& e = *E[i];
      const Plan& pl = PL[i];
      return args<k_eng_finalize>(e.fx_site.p,
                                  (pl.recip_on || pl.reuse) ? e.fx_recip.p : (const int64_t*)nullptr,
                                  (const int64_t*)nullptr, e.chunk_ptr.p,
                                  e.cap_chunks, e.slot_of.p, e.n_sites, e.cap_slots, e.fx_out.p,
                                  e.dflags.p, forced[i] ? 1 : 0);
    });
    if (short_on && energy_on)
      launch<k_eng_energy_reduce1, kRed1Block>(R, "k_eng_energy_reduce1", 0, kRed1Grid, kRed1Block,
                                               0, [&](uint32_t i) {
        dyn_engine& e = *E[i];
        return args<k_eng_energy_reduce1>((const double*)e.pair_part.p, e.counters.p + 2,
                                          e.cap_chunks, e.red1.p);
      });
    if (energy_on)
      launch<k_eng_energy_reduce, kRedBlock>(R, "k_eng_energy_reduce", 0, 1, kRedBlock, 0,
                                             [&](uint32_t i) {
        dyn_engine& e = *E[i];
        const Plan& pl = PL[i];
        return args<k_eng_energy_reduce>(
            short_on ? (const double*)e.red1.p : (const double*)nullptr, e.corr_part.p,
            pl.corr_on ? e.corr_blocks : 0u,
            pl.reuse ? -1 : pl.recip_on ? (pl.recip_b ? 2 : 1) : 0, e.eng_out.p);
      });
  }
  for (uint32_t i = 0; i < ns; ++i) {
    dyn_engine& e = *E[i];
    if (PL[i].recip_on) e.recip_kept = true;
    else if (!PL[i].reuse) e.recip_kept = false;
    e.computed = true;
    e.list_valid = true;  // every later evaluation keeps it valid on the device
    e.energies_valid = energy_on;
    e.mol_valid = energy_on && e.n_mol;
    e.fx_fresh = false;
  }
}

void engine_set_check(EngineSet& S) {
  const Error* first = nullptr;
  Error keep(0, "");
  for (dyn_engine* e : S.e) {
    try {
      e->check_status();
    } catch (const Error& err) {
      if (!first) {
        keep = err;
        first = &keep;
      }
    }
  }
  if (first) throw keep;
}

}  // namespace dyn

void dyn_engine::check_status() {
  DYN_HIP(hipMemcpyAsync(h_misc.p, status.p, 4, hipMemcpyDeviceToHost, stream));
  DYN_HIP(hipMemcpyAsync(h_misc.p + 1, counters.p, 20, hipMemcpyDeviceToHost, stream));
  DYN_HIP(hipMemcpyAsync(h_misc.p + 6, dflags.p + kFlBuilds, 4, hipMemcpyDeviceToHost, stream));
  DYN_HIP(hipMemcpyAsync(h_misc.p + 7, dflags.p + kFlSel, 4, hipMemcpyDeviceToHost, stream));
  DYN_HIP(hipMemcpyAsync(h_misc.p + 8, stg_counters.p, 20, hipMemcpyDeviceToHost, stream));
  DYN_HIP(hipStreamSynchronize(stream));
  const uint32_t st = h_misc.p[0];
  if (!st) return;
  DYN_HIP(hipMemsetAsync(status.p, 0, 4, stream));
  DYN_HIP(hipMemsetAsync(col_count.p, 0, sizeof(uint32_t) * (n_cols + 1), stream));
  list_valid = false;
  if (st & kStDisp) {
    if ((st & ~(kStDisp | kStPme)) == 0)
      throw Error(DYN_ERR_CAPACITY, "a site moved half the list buffer since its list was built "
                                    "(the list was staged too late for the step); call again");
  }
  if (st & kStPme) {
    pme.grow();
    if (st == kStPme)
      throw Error(DYN_ERR_CAPACITY, fmt("an SPME brick list overflowed; lists enlarged to %u "
                                        "charges per brick", pme.cap_brick));
  }
  // fixed limits of the design: calling again does not help
  if ((st & kStColumn) && col_slice >= (uint32_t)kColCap)
    throw Error(DYN_ERR_LIMIT, fmt("a cluster column holds more than %d sites", kColCap));
  if (st & kStMask)
    throw Error(DYN_ERR_LIMIT, fmt("the sites of a solute cluster have more than %u exclusion "
                                   "partners (1-2, 1-3, 1-4) together", kExPartners));
  if (st & kStColumn) {
    col_slice = std::min<uint32_t>(2 * col_slice, kColCap);
    items.reserve((size_t)n_cols * col_slice);
  }
  if (st & kStRow) {
    // the longest row of either list's build (counters [4])
    const uint32_t longest = std::max(h_misc.p[5], h_misc.p[12]);
    DYN_HIP(hipMemsetAsync(counters.p + 4, 0, 4, stream));
    DYN_HIP(hipMemsetAsync(stg_counters.p + 4, 0, 4, stream));
    reserve_rows(std::max<uint32_t>(longest + longest / 4, row_cap + row_cap / 2));
  }
  const uint64_t ne = std::max(h_misc.p[2], h_misc.p[9]), nch = std::max(h_misc.p[3], h_misc.p[10]);
  reserve_list(ne + ne / 4 + 1024, nch + nch / 4 + 256);
  throw Error(DYN_ERR_CAPACITY, fmt("pair list outgrew its buffers (%llu groups, %llu work "
                                    "items, row slices of %u entries); they have been enlarged, "
                                    "call again", (unsigned long long)ne,
                                    (unsigned long long)nch, row_cap));
}

// ---------------------------------------------------------------------------
// C entry points
// ---------------------------------------------------------------------------
extern "C" int dyn_engine_create(dyn_engine** out, int device, const dyn_engine_params* params,
                                 uint32_t n_dyn, uint32_t n_water, const float* q_dyn,
                                 const int32_t* type_dyn, uint32_t n_types,
                                 const float* lj_sigma_eps, uint32_t n_excl,
                                 const uint32_t* excl_pairs, uint32_t n_14,
                                 const uint32_t* pairs_14) {
  try {
    DYN_REQUIRE(out && params, "null argument");
    *out = nullptr;
    const dyn_engine_params& p = *params;
    DYN_REQUIRE(p.cutoff > 0.f && std::isfinite(p.cutoff), "cutoff must be > 0");
    DYN_REQUIRE(p.lj_cutoff > 0.f && p.lj_cutoff <= p.cutoff,
                "lj_cutoff must be in (0, cutoff] for the native engine");
    DYN_REQUIRE(p.beta > 0.f && p.beta * p.cutoff <= 4.0f,
                "beta must be > 0 with beta·cutoff ≤ 4 (range of the erfc approximation)");
    DYN_REQUIRE(p.mode == DYN_MODE_REFERENCE || p.mode == DYN_MODE_STANDARD, "unknown mode");
    DYN_REQUIRE((p.precision & ~(DYN_PREC_PAIR_F64 | DYN_PREC_PME_F64)) == 0,
                "unknown precision bits 0x%x", p.precision);
    DYN_REQUIRE(p.list_buffer >= 0.f && p.list_buffer < 10.f, "list_buffer must be in [0, 10)");
    DYN_REQUIRE(p.mesh_spacing >= 0.f && std::isfinite(p.mesh_spacing), "bad mesh_spacing");
    DYN_REQUIRE(p.mesh_spacing == 0.f || (p.grid[0] == 0 && p.grid[1] == 0 && p.grid[2] == 0),
                "give either a PME grid or a mesh spacing");
    DYN_REQUIRE(p.item_groups <= 64, "item_groups must be in 1..64 (0: 16)");
    DYN_REQUIRE(p.single_stream <= 1, "single_stream must be 0 or 1");
    DYN_REQUIRE(n_dyn + 4ull * n_water > 0, "empty system");
    DYN_REQUIRE(n_dyn + 4ull * n_water < (1ull << 30), "system too large");
    DYN_REQUIRE(n_dyn == 0 || (q_dyn && type_dyn), "null solute arrays");
    DYN_REQUIRE(n_dyn == 0 || (n_types > 0 && lj_sigma_eps), "solute needs LJ types");
    DYN_REQUIRE(n_types <= 64, "at most 64 LJ types");
    int ndev = 0;
    DYN_HIP(hipGetDeviceCount(&ndev));
    DYN_REQUIRE(device >= 0 && device < ndev, "device %d not present (%d devices)", device, ndev);
    std::unique_ptr<dyn_engine> E(new dyn_engine());
    E->device = device;
    DeviceGuard g(device);
    DYN_HIP(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    E->own_stream = E->stream;
    DYN_HIP(hipStreamCreateWithFlags(&E->side_stream, hipStreamNonBlocking));
    DYN_HIP(hipEventCreateWithFlags(&E->ev_fork, hipEventDisableTiming));
    DYN_HIP(hipEventCreateWithFlags(&E->ev_join, hipEventDisableTiming));
    E->prof.stream = E->stream;
    E->prm = p;
    if (!E->prm.item_groups) E->prm.item_groups = DYN_CHUNK;
    E->n_dyn = n_dyn;
    E->n_water = n_water;
    const uint32_t ns = n_dyn + 4 * n_water;
    E->n_sites = ns;
    // LJ types: solute types, then water O, then "none".
    const int nt = (int)n_types + 2;
    E->n_types = nt;
    E->type_o = (int)n_types;
    E->type_none = (int)n_types + 1;
    std::vector<float> sig(nt, 0.f), eps(nt, 0.f);
    for (uint32_t t = 0; t < n_types; ++t) {
      sig[t] = lj_sigma_eps[2 * t];
      eps[t] = lj_sigma_eps[2 * t + 1];
      DYN_REQUIRE(std::isfinite(sig[t]) && std::isfinite(eps[t]) && eps[t] >= 0.f,
                  "bad LJ parameters for type %u", t);
    }
    // OPC oxygen (src/solvent/mod.rs:73-79), evaluated in f32 as the reference does.
    const float o_sigma = 1.777167268f * (2.0f / 1.122462048309373f);
    sig[E->type_o] = o_sigma;
    eps[E->type_o] = 0.2128008130f;
    std::vector<float2> tab((size_t)nt * nt), tse((size_t)nt * nt);
    std::vector<double2> tab64((size_t)nt * nt);
    for (int a = 0; a < nt; ++a)
      for (int b = 0; b < nt; ++b) {
        // Lorentz–Berthelot in f32 (src/non_bonded.rs:1081-1086), then
        // c6 = 4εσ⁶, c12 = 4εσ¹² in f64.
        const float s_ab = 0.5f * (sig[a] + sig[b]);
        const float e_ab = std::sqrt(eps[a] * eps[b]);
        const double s6 = std::pow((double)s_ab, 6.0);
        tab[(size_t)a * nt + b] = make_float2((float)(4.0 * e_ab * s6), (float)(4.0 * e_ab * s6 * s6));
        tab64[(size_t)a * nt + b] = make_double2(4.0 * e_ab * s6, 4.0 * e_ab * s6 * s6);
        tse[(size_t)a * nt + b] = make_float2(s_ab, e_ab);
      }
    E->c6_oo = tab[(size_t)E->type_o * nt + E->type_o].x;
    E->c12_oo = tab[(size_t)E->type_o * nt + E->type_o].y;
    // per-site topology
    const float qh = 0.6791f * 18.2223f, qm = -2.0f * qh;
    std::vector<float> hq(ns);
    std::vector<int32_t> ht(ns), hw(ns);
    for (uint32_t i = 0; i < n_dyn; ++i) {
      DYN_REQUIRE(type_dyn[i] >= 0 && (uint32_t)type_dyn[i] < n_types, "atom %u: bad type", i);
      DYN_REQUIRE(std::isfinite(q_dyn[i]), "atom %u: non-finite charge", i);
      hq[i] = q_dyn[i];
      ht[i] = type_dyn[i];
      hw[i] = -1;
    }
    const float wq[4] = {0.f, qm, qh, qh};
    for (uint32_t m = 0; m < n_water; ++m)
      for (int k = 0; k < 4; ++k) {
        const uint32_t s = n_dyn + 4 * m + k;
        hq[s] = wq[k];
        ht[s] = k == 0 ? E->type_o : E->type_none;
        hw[s] = (int32_t)m;
      }
    // exclusions (1-2, 1-3 and 1-4 all leave the pair kernel), CSR both ways
    std::vector<std::vector<uint32_t>> ex(n_dyn);
    std::vector<uint32_t> ci, cjv;
    std::vector<uint8_t> ck;
    auto add_pair = [&](uint32_t i, uint32_t j, const char* what, uint32_t k) {
      DYN_REQUIRE(i < n_dyn && j < n_dyn && i != j, "%s pair %u out of range", what, k);
      ex[i].push_back(j);
      ex[j].push_back(i);
    };
    for (uint32_t k = 0; k < n_excl; ++k) {
      const uint32_t i = excl_pairs[2 * k], j = excl_pairs[2 * k + 1];
      add_pair(i, j, "excluded", k);
      if (p.mode == DYN_MODE_STANDARD) {
        ci.push_back(i);
        cjv.push_back(j);
        ck.push_back(0);
      }
    }
    for (uint32_t k = 0; k < n_14; ++k) {
      const uint32_t i = pairs_14[2 * k], j = pairs_14[2 * k + 1];
      add_pair(i, j, "1-4", k);
      ci.push_back(i);
      cjv.push_back(j);
      ck.push_back(1);
    }
    if (p.mode == DYN_MODE_STANDARD)
      for (uint32_t m = 0; m < n_water; ++m) {
        const uint32_t b = n_dyn + 4 * m;
        const uint32_t pr[3][2] = {{b + 1, b + 2}, {b + 1, b + 3}, {b + 2, b + 3}};
        for (auto& pp : pr) {
          ci.push_back(pp[0]);
          cjv.push_back(pp[1]);
          ck.push_back(0);
        }
      }
    std::vector<uint32_t> hptr(n_dyn + 1, 0), hidx;
    for (uint32_t i = 0; i < n_dyn; ++i) {
      auto& v = ex[i];
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      hptr[i + 1] = hptr[i] + (uint32_t)v.size();
      hidx.insert(hidx.end(), v.begin(), v.end());
    }
    std::vector<uint32_t> hqs;
    double q2 = 0.0, qt = 0.0;
    for (uint32_t s = 0; s < ns; ++s)
      if (hq[s] != 0.f) {
        hqs.push_back(s);
        q2 += (double)hq[s] * hq[s];
        qt += hq[s];
      }
    E->n_q = (uint32_t)hqs.size();
    E->e_self = -(double)p.beta / std::sqrt(M_PI) * q2;
    E->q_total = qt;
    auto up = [&](auto& buf, const auto& v) {
      buf.reserve(std::max<size_t>(v.size(), 1));
      if (!v.empty())
        DYN_HIP(hipMemcpyAsync(buf.p, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice,
                               E->stream));
    };
    up(E->q, hq);
    up(E->ltype, ht);
    up(E->wmol, hw);
    up(E->lj_table, tab);
    up(E->lj_table64, tab64);
    up(E->lj_se, tse);
    E->h_type = ht;
    up(E->ex_ptr, hptr);
    up(E->ex_idx, hidx);
    up(E->corr_i, ci);
    up(E->corr_j, cjv);
    up(E->corr_kind, ck);
    up(E->qsite, hqs);
    E->n_corr = (uint32_t)ci.size();
    E->corr_blocks = grid_for(std::max<uint32_t>(E->n_corr, 1), kCorrBlock, 256);
    E->corr_part.reserve(kCorrVals * (size_t)E->corr_blocks);
    E->pos_own.reserve(ns);
    E->h_pos.reserve(ns);
    E->wpos.reserve(ns);
    E->wpos_ref.reserve(ns);
    E->slot_of.reserve(ns);
    E->stg_wpos_ref.reserve(ns);
    E->stg_slot_of.reserve(ns);
    E->fx_site.reserve(3 * (size_t)ns);
    E->fx_out.reserve(3 * (size_t)ns);
    E->fx_recip.reserve(3 * (size_t)ns);
    E->h_fx.reserve(3 * (size_t)ns);
    E->eng_out.reserve(kEngOut);
    E->red1.reserve((size_t)kRed1Grid * kNumPart);
    E->h_eng.reserve(kEngOut);
    E->h_misc.reserve(16);
    E->counters.reserve(8);
    E->stg_counters.reserve(8);
    E->status.reserve(1);
    E->dflags.reserve(kFlWords);
    {
      std::vector<float> tab(8 * DYN_COUL_TAB_N);
      dyn_coul_tab_build(tab.data());
      E->coul_tab.reserve(2 * DYN_COUL_TAB_N);
      DYN_HIP(hipMemcpyAsync(E->coul_tab.p, tab.data(), sizeof(float) * tab.size(),
                             hipMemcpyHostToDevice, E->stream));
      DYN_HIP(hipStreamSynchronize(E->stream));
    }
    // checkpoint copies of an MD call (allocated here: a step call allocates
    // nothing, so captured graphs of its steps stay current)
    E->fx_out_ck.reserve(3 * (size_t)ns);
    E->fx_recip_ck.reserve(3 * (size_t)ns);
    E->eng_ck.reserve(kEngOut);
    E->pme_xyzq.reserve(std::max<uint32_t>(E->n_q, 1));
    E->pme_b.reserve(std::max<uint32_t>(E->n_q, 1));
    E->pme_on = (p.grid[0] > 0 && p.grid[1] > 0 && p.grid[2] > 0) || p.mesh_spacing > 0.f;
    DYN_HIP(hipMemsetAsync(E->counters.p, 0, 32, E->stream));
    DYN_HIP(hipMemsetAsync(E->slot_of.p, 0, 4 * (size_t)ns, E->stream));
    DYN_HIP(hipMemsetAsync(E->status.p, 0, 4, E->stream));
    DYN_HIP(hipMemsetAsync(E->dflags.p, 0, 4 * kFlWords, E->stream));
    DYN_HIP(hipMemsetAsync(E->stg_counters.p, 0, 4 * 8, E->stream));
    // reference positions are read before the first build writes them (the
    // first evaluation's reading is discarded): keep them defined
    DYN_HIP(hipMemsetAsync(E->wpos_ref.p, 0, sizeof(float4) * ns, E->stream));
    DYN_HIP(hipMemsetAsync(E->stg_wpos_ref.p, 0, sizeof(float4) * ns, E->stream));
    DYN_HIP(hipStreamSynchronize(E->stream));
    *out = E.release();
    return DYN_OK;
  } catch (const Error& e) {
    set_global_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_global_error(e.what());
    return DYN_ERR_HIP;
  }
}

extern "C" void dyn_engine_destroy(dyn_engine* eng) {
  if (!eng) return;
  try {
    DeviceGuard g(eng->device);
    if (eng->stream) (void)hipStreamSynchronize(eng->stream);
    delete eng;
  } catch (...) {
  }
}

extern "C" const char* dyn_engine_last_error(const dyn_engine* eng) {
  return eng ? eng->err.c_str() : dyn_global_last_error();
}

#define ENG_CHECK(e)                          \
  if (!(e)) {                                 \
    set_global_error("null engine handle");   \
    return DYN_ERR_ARG;                       \
  }

extern "C" int dyn_engine_set_box(dyn_engine* eng, const float box[3]) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    g_layout_gen.fetch_add(1);  // captured graphs hold the old arguments
    DeviceGuard g(eng->device);
    DYN_REQUIRE(box, "box is null");
    for (int a = 0; a < 3; ++a) {
      DYN_REQUIRE(box[a] > 0.f && std::isfinite(box[a]), "box lengths must be > 0");
      DYN_REQUIRE(eng->prm.cutoff + eng->prm.list_buffer < 0.5f * box[a],
                  "cutoff + list_buffer must be below half the box (minimum image)");
    }
    uint32_t K[3] = {eng->prm.grid[0], eng->prm.grid[1], eng->prm.grid[2]};
    if (eng->prm.mesh_spacing > 0.f &&
        dyn_pme_grid_n(box[0], box[1], box[2], eng->prm.mesh_spacing, K) != DYN_OK)
      throw Error(DYN_ERR_ARG, dyn_global_last_error());
    for (int a = 0; a < 3; ++a) eng->box[a] = box[a];
    eng->box_set = true;
    eng->recip_kept = false;
    eng->set_geometry();
    if (eng->pme_on) {
      // a grid of another size is planned afresh (regen_pme); the same grid
      // only takes the new influence function
      const bool same = eng->pme.ngrid != 0 && K[0] == eng->pme.K[0] && K[1] == eng->pme.K[1] &&
                        K[2] == eng->pme.K[2];
      if (!same)
        eng->pme.init(eng->device, eng->stream, &eng->prof, K, eng->box, eng->prm.beta,
                      (int)eng->prm.order, (eng->prm.precision & DYN_PREC_PME_F64) != 0);
      else
        eng->pme.set_box(eng->box);
    }
  });
}

extern "C" int dyn_engine_set_positions(dyn_engine* eng, const float* pos_dyn,
                                        const float* pos_w_o, const float* pos_w_m,
                                        const float* pos_w_h0, const float* pos_w_h1) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    const uint32_t nd = eng->n_dyn, nw = eng->n_water;
    DYN_REQUIRE(nd == 0 || pos_dyn, "pos_dyn is null");
    DYN_REQUIRE(nw == 0 || (pos_w_o && pos_w_m && pos_w_h0 && pos_w_h1), "water positions null");
    float4* h = eng->h_pos.p;
    for (uint32_t i = 0; i < nd; ++i) h[i] = make_float4(pos_dyn[3 * i], pos_dyn[3 * i + 1], pos_dyn[3 * i + 2], 0.f);
    const float* w[4] = {pos_w_o, pos_w_m, pos_w_h0, pos_w_h1};
    for (uint32_t m = 0; m < nw; ++m)
      for (int k = 0; k < 4; ++k)
        h[nd + 4 * m + k] = make_float4(w[k][3 * m], w[k][3 * m + 1], w[k][3 * m + 2], 0.f);
    for (uint32_t s = 0; s < eng->n_sites; ++s)
      DYN_REQUIRE(std::isfinite(h[s].x) && std::isfinite(h[s].y) && std::isfinite(h[s].z),
                  "non-finite position at site %u", s);
    DYN_HIP(hipMemcpyAsync(eng->pos_own.p, h, sizeof(float4) * eng->n_sites,
                           hipMemcpyHostToDevice, eng->stream));
    DYN_HIP(hipStreamSynchronize(eng->stream));
    eng->pos_dev = eng->pos_own.p;
    eng->pos_set = true;
  });
}

extern "C" int dyn_engine_set_positions_device(dyn_engine* eng, const float* xyzw_device) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    g_layout_gen.fetch_add(1);  // captured graphs hold the old arguments
    DYN_REQUIRE(xyzw_device, "null device pointer");
    DYN_REQUIRE(((uintptr_t)xyzw_device & 15) == 0, "device positions must be 16-byte aligned");
    eng->pos_dev = reinterpret_cast<const float4*>(xyzw_device);
    eng->pos_set = true;
  });
}

extern "C" int dyn_engine_get_pme_grid(dyn_engine* eng, uint32_t out[3]) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DYN_REQUIRE(out, "out is null");
    for (int a = 0; a < 3; ++a) out[a] = eng->pme_on ? eng->pme.K[a] : 0u;
  });
}

extern "C" int dyn_engine_set_alchemical(dyn_engine* eng, uint32_t first, uint32_t count,
                                         double lambda) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    g_layout_gen.fetch_add(1);  // captured graphs hold the old arguments
    DeviceGuard g(eng->device);
    DYN_REQUIRE(std::isfinite(lambda), "lambda must be finite");
    DYN_REQUIRE((uint64_t)first + count <= eng->n_dyn, "alchemical atoms [%u, %u) beyond the %u "
                "solute atoms", first, first + count, eng->n_dyn);
    DYN_REQUIRE(count == 0 || !(eng->prm.precision & DYN_PREC_PAIR_F64),
                "alchemical decoupling runs in FP32 pair arithmetic only");
    DYN_REQUIRE(count == 0 || eng->prm.mode == DYN_MODE_REFERENCE,
                "alchemical decoupling needs DYN_MODE_REFERENCE");
    const uint32_t a0 = count ? first : 0, a1 = count ? first + count : 0;
    if (a0 != eng->alch0 || a1 != eng->alch1) {
      for (uint32_t i = 0; i < eng->n_dyn; ++i) {
        int32_t& t = eng->h_type[i];
        t = (t & DYN_TYPE_MASK) | (i >= a0 && i < a1 ? DYN_TYPE_ALCH : 0);
      }
      DYN_HIP(hipMemcpyAsync(eng->ltype.p, eng->h_type.data(), sizeof(int32_t) * eng->n_dyn,
                             hipMemcpyHostToDevice, eng->stream));
      DYN_HIP(hipStreamSynchronize(eng->stream));
      eng->alch0 = a0;
      eng->alch1 = a1;
      eng->list_valid = false;  // slot types are refreshed by a rebuild
    }
    eng->lambda = lambda;
  });
}

extern "C" int dyn_engine_set_molecules(dyn_engine* eng, uint32_t n_mol,
                                        const uint32_t* mol_start) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    g_layout_gen.fetch_add(1);  // captured graphs hold the old arguments
    DeviceGuard g(eng->device);
    DYN_REQUIRE(n_mol <= DYN_MAX_MOLECULES, "%u molecules (at most %u)", n_mol,
                (unsigned)DYN_MAX_MOLECULES);
    DYN_REQUIRE(n_mol == 0 || mol_start, "mol_start is null");
    // the reference's checks (src/lib.rs:1522-1566)
    if (n_mol) {
      DYN_REQUIRE(eng->n_dyn > 0, "molecules given, but there are no solute atoms");
      DYN_REQUIRE(mol_start[0] == 0, "first molecule starts at atom %u, but it must start at 0",
                  mol_start[0]);
      for (uint32_t m = 1; m < n_mol; ++m) {
        DYN_REQUIRE(mol_start[m] < eng->n_dyn, "molecule %u starts at atom %u, but valid atom "
                    "indices are 0..%u", m, mol_start[m], eng->n_dyn - 1);
        DYN_REQUIRE(mol_start[m] > mol_start[m - 1], "molecule starts must be strictly "
                    "increasing; molecule %u starts at %u after %u", m, mol_start[m],
                    mol_start[m - 1]);
      }
    }
    DYN_HIP(hipStreamSynchronize(eng->stream));  // no evaluation still reads the old arrays
    eng->n_mol = n_mol;
    eng->mol_valid = false;
    if (n_mol) {
      std::vector<int32_t> of(eng->n_dyn);
      for (uint32_t m = 0; m < n_mol; ++m) {
        const uint32_t end = m + 1 < n_mol ? mol_start[m + 1] : eng->n_dyn;
        for (uint32_t a = mol_start[m]; a < end; ++a) of[a] = (int32_t)m;
      }
      eng->mol_of.reserve(eng->n_dyn);
      eng->mol_e.reserve((size_t)n_mol * n_mol);
      eng->h_mol.reserve((size_t)n_mol * n_mol);
      DYN_HIP(hipMemcpy(eng->mol_of.p, of.data(), sizeof(int32_t) * eng->n_dyn,
                        hipMemcpyHostToDevice));
    }
  });
}

extern "C" int dyn_engine_get_mol_energies(dyn_engine* eng, double* out, uint32_t n_mol) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    DYN_REQUIRE(eng->n_mol, "no molecules set (dyn_engine_set_molecules)");
    DYN_REQUIRE(out && n_mol == eng->n_mol, "need the %u × %u matrix", eng->n_mol, eng->n_mol);
    if (!eng->mol_valid)
      throw Error(DYN_ERR_STATE, "no compute with energies since the molecules were set");
    eng->check_status();
    const size_t nn = (size_t)n_mol * n_mol;
    DYN_HIP(hipMemcpyAsync(eng->h_mol.p, eng->mol_e.p, sizeof(int64_t) * nn,
                           hipMemcpyDeviceToHost, eng->stream));
    DYN_HIP(hipStreamSynchronize(eng->stream));
    for (size_t k = 0; k < nn; ++k) out[k] = (double)eng->h_mol.p[k] * DYN_FX_INV_SCALE;
  });
}

extern "C" int dyn_engine_compute(dyn_engine* eng, uint32_t flags) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    EngineSet S = eng->self_set();
    if (flags & DYN_COMPUTE_ASYNC) {
      engine_set_enqueue(S, flags);
    } else {
      // a list that outgrew its buffers is rebuilt once with enlarged ones
      engine_set_enqueue(S, flags);
      try {
        eng->check_status();
      } catch (const Error& e) {
        if (e.code != DYN_ERR_CAPACITY || eng->list_valid) throw;
        engine_set_enqueue(S, flags | DYN_COMPUTE_REBUILD);
        eng->check_status();
      }
    }
  });
}

extern "C" int dyn_engine_sync(dyn_engine* eng) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    eng->check_status();
  });
}

static void fetch_fx(dyn_engine* eng) {
  if (!eng->computed) throw Error(DYN_ERR_STATE, "no compute yet");
  if (eng->fx_stale) throw Error(DYN_ERR_STATE, "the last evaluation's forces were summed by its caller");
  if (eng->fx_fresh) return;
  eng->check_status();
  DYN_HIP(hipMemcpyAsync(eng->h_fx.p, eng->fx_out.p, sizeof(int64_t) * 3 * eng->n_sites,
                         hipMemcpyDeviceToHost, eng->stream));
  DYN_HIP(hipStreamSynchronize(eng->stream));
  eng->fx_fresh = true;
}

extern "C" int dyn_engine_get_forces(dyn_engine* eng, double* f_dyn, double* f_water) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    fetch_fx(eng);
    const uint32_t ns = eng->n_sites, nd = eng->n_dyn, nw = eng->n_water;
    const int64_t* F = eng->h_fx.p;
    if (f_dyn)
      for (uint32_t i = 0; i < nd; ++i)
        for (int c = 0; c < 3; ++c) f_dyn[3 * i + c] = (double)F[(size_t)c * ns + i] * DYN_FX_INV_SCALE;
    if (f_water) {
      // ForcesOnWaterMol order f_o, f_h0, f_h1, f_m; site order O, M, H0, H1
      const int site_of_field[4] = {0, 2, 3, 1};
      for (uint32_t m = 0; m < nw; ++m)
        for (int fld = 0; fld < 4; ++fld) {
          const uint32_t s = nd + 4 * m + site_of_field[fld];
          for (int c = 0; c < 3; ++c)
            f_water[12 * (size_t)m + 3 * fld + c] = (double)F[(size_t)c * ns + s] * DYN_FX_INV_SCALE;
        }
    }
  });
}

extern "C" int dyn_engine_get_fixed_forces(dyn_engine* eng, int64_t* out, uint64_t n_values) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    DYN_REQUIRE(out && n_values >= 3ull * eng->n_sites, "need %llu values",
                (unsigned long long)(3ull * eng->n_sites));
    fetch_fx(eng);
    memcpy(out, eng->h_fx.p, sizeof(int64_t) * 3 * eng->n_sites);
  });
}

extern "C" const int64_t* dyn_engine_forces_device(const dyn_engine* eng) {
  return eng ? eng->fx_out.p : nullptr;
}

extern "C" int dyn_engine_get_energies(dyn_engine* eng, dyn_energies* out) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    DYN_REQUIRE(out, "out is null");
    if (!eng->computed) throw Error(DYN_ERR_STATE, "no compute yet");
    if (!eng->energies_valid)
      throw Error(DYN_ERR_STATE, "the last compute skipped energies (DYN_COMPUTE_NO_ENERGY)");
    eng->check_status();
    DYN_HIP(hipMemcpyAsync(eng->h_eng.p, eng->eng_out.p, kEngOut * sizeof(double),
                           hipMemcpyDeviceToHost, eng->stream));
    DYN_HIP(hipStreamSynchronize(eng->stream));
    const double* v = eng->h_eng.p;
    // v[0..8]: e_lj_w, e_lj_s, e_c_w, e_c_s, virial_pairs, pair counts by
    // kind, dH/dλ; v[9..15]: e_lj14, e_c14, e_corr, vir_sr14, vir_recip_corr, dH/dλ,
    // 1-4 pairs within the cutoff;
    // v[16..17], v[18..19]: energy and virial of the SPME evaluations
    double e_rec = v[16], w_rec = v[17], dh_rec = 0.0;
    if (eng->last_recip_b) {
      // A = E(q), B = E(q with the molecule's charges negated): the cross
      // term is (A − B)/2, the rest (A + B)/2 (src/non_bonded.rs:709-742)
      const double sc = eng->last_sch.coulomb_scale;
      e_rec = 0.5 * (1.0 + sc) * v[16] + 0.5 * (1.0 - sc) * v[18];
      w_rec = 0.5 * (1.0 + sc) * v[17] + 0.5 * (1.0 - sc) * v[19];
      dh_rec = (double)eng->last_sch.coulomb_dscale_dlambda * 0.5 * (v[16] - v[18]);
    }
    dyn_energies E{};
    E.lj = v[0] + v[1] + v[9];
    E.coul_sr = v[2] + v[3] + v[10];
    E.recip = eng->pme_on ? e_rec : 0.0;
    double self = eng->pme_on ? eng->e_self : 0.0;
    if (eng->pme_on && eng->prm.mode == DYN_MODE_STANDARD) {
      const double V = (double)eng->box[0] * eng->box[1] * eng->box[2];
      self += -M_PI * eng->q_total * eng->q_total / (2.0 * V * (double)eng->prm.beta * eng->prm.beta);
    }
    E.self = self;
    E.excl_corr = v[11];
    E.total = E.lj + E.coul_sr + E.recip + E.self + E.excl_corr;
    E.reference = v[1] + v[3] + v[9] + v[10] + E.recip + E.self;
    E.virial_sr = v[4] + v[12];
    E.virial_recip = (eng->pme_on ? w_rec : 0.0) + v[13];
    E.dh_dl = v[8] + v[14] + dh_rec;
    E.pair_count_kind[0] = (uint64_t)std::llround(v[5]);
    E.pair_count_kind[1] = (uint64_t)std::llround(v[6]);
    E.pair_count_kind[2] = (uint64_t)std::llround(v[7]);
    E.pair_count = E.pair_count_kind[0] + E.pair_count_kind[1] + E.pair_count_kind[2] +
                   (uint64_t)std::llround(v[15]);
    E.list_entries = eng->h_misc.p[2];
    E.list_builds = eng->h_misc.p[6];
    *out = E;
  });
}

extern "C" int dyn_engine_debug_export(dyn_engine* eng, uint32_t* n_slots, uint32_t* n_clusters,
                                       uint64_t* n_entries, int32_t* slot_site, float* xyzq,
                                       int32_t* slot_type, int32_t* slot_gid, uint8_t* cl_class,
                                       uint32_t* row_ptr, uint32_t* cj, uint64_t* cmask,
                                       uint32_t* items, uint64_t* n_items, int64_t* fx_slot) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    if (!eng->list_valid) throw Error(DYN_ERR_STATE, "no list built yet");
    eng->check_status();
    const size_t nc = eng->h_misc.p[1], ns = nc * DYN_CL_SIZE, ne = eng->h_misc.p[2],
                 nit = eng->h_misc.p[3];
    if (n_slots) *n_slots = (uint32_t)ns;
    if (n_clusters) *n_clusters = (uint32_t)nc;
    if (n_entries) *n_entries = ne;
    if (n_items) *n_items = nit;
    hipStream_t s = eng->stream;
    if (slot_site) DYN_HIP(hipMemcpyAsync(slot_site, eng->slot_site.p, 4 * ns, hipMemcpyDeviceToHost, s));
    if (xyzq) DYN_HIP(hipMemcpyAsync(xyzq, eng->xyzq.p, 16 * ns, hipMemcpyDeviceToHost, s));
    if (slot_type) DYN_HIP(hipMemcpyAsync(slot_type, eng->slot_type.p, 4 * ns, hipMemcpyDeviceToHost, s));
    if (slot_gid) DYN_HIP(hipMemcpyAsync(slot_gid, eng->slot_gid.p, 4 * ns, hipMemcpyDeviceToHost, s));
    if (cl_class) DYN_HIP(hipMemcpyAsync(cl_class, eng->cl_class.p, nc, hipMemcpyDeviceToHost, s));
    if (row_ptr) DYN_HIP(hipMemcpyAsync(row_ptr, eng->row_ptr.p, 4 * (nc + 1), hipMemcpyDeviceToHost, s));
    const bool second = eng->h_misc.p[7] != 0;  // the current copy of the per-group arrays
    const uint32_t* cj_d = second ? eng->cj2.p : eng->cj.p;
    const unsigned long long* cm_d = second ? eng->cmask2.p : eng->cmask.p;
    const uint2* ch_d = second ? eng->chunks2.p : eng->chunks.p;
    if (cj && ne) DYN_HIP(hipMemcpyAsync(cj, cj_d, 4 * DYN_CL_SIZE * ne, hipMemcpyDeviceToHost, s));
    if (cmask && ne) DYN_HIP(hipMemcpyAsync(cmask, cm_d, 8 * ne, hipMemcpyDeviceToHost, s));
    if (items && nit) DYN_HIP(hipMemcpyAsync(items, ch_d, 8 * nit, hipMemcpyDeviceToHost, s));
    if (fx_slot) {
      eng->fx_slot.reserve(3 * (size_t)eng->cap_slots);
      EngineSet S = eng->self_set();
      launch<k_eng_slot_forces>(S.R, "k_eng_slot_forces", 0, grid_for(eng->cap_slots, 256), 256, 0,
                                [&](uint32_t) {
        return args<k_eng_slot_forces>(eng->fitem.p, eng->chunk_ptr.p, eng->cap_chunks,
                                       eng->counters.p + 0, eng->cap_slots, eng->fx_slot.p);
      });
      for (int c = 0; c < 3; ++c)
        DYN_HIP(hipMemcpyAsync(fx_slot + c * ns, eng->fx_slot.p + (size_t)c * eng->cap_slots, 8 * ns,
                               hipMemcpyDeviceToHost, s));
    }
    DYN_HIP(hipStreamSynchronize(s));
  });
}

extern "C" int dyn_engine_debug_item_partials(dyn_engine* eng, double* out, uint64_t n_items) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    DYN_REQUIRE(out, "out is null");
    if (!eng->energies_valid)
      throw Error(DYN_ERR_STATE, "the last compute evaluated no pair energies");
    eng->check_status();
    const uint64_t nit = eng->h_misc.p[3];
    DYN_REQUIRE(n_items == nit, "n_items %llu, the list has %llu work items",
                (unsigned long long)n_items, (unsigned long long)nit);
    for (int v = 0; v < kNumPart; ++v)
      if (nit)
        DYN_HIP(hipMemcpyAsync(out + (size_t)v * nit, eng->pair_part.p + (size_t)v * eng->cap_chunks,
                               8 * nit, hipMemcpyDeviceToHost, eng->stream));
    DYN_HIP(hipStreamSynchronize(eng->stream));
  });
}

extern "C" uint64_t dyn_debug_arg_uploads(void) { return g_arg_uploads.load(); }
extern "C" uint64_t dyn_debug_layout_generation(void) { return g_layout_gen.load(); }
extern "C" const char* dyn_debug_capture_refusal(void) { return g_capture_refusal.c_str(); }

extern "C" int dyn_engine_profile(dyn_engine* eng, int enable) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    eng->prof.reset();
    const bool on = enable != 0;
    const char* only = enable == 2 ? "pairs" : nullptr;
    // captured graphs hold the timing events of their mode: a new mode makes
    // new graphs (the same mode again only clears the totals)
    if (on != eng->prof.enabled || only != eng->prof.only) g_layout_gen.fetch_add(1);
    eng->prof.enabled = on;
    eng->prof.only = only;
  });
}

extern "C" int dyn_engine_profile_read(dyn_engine* eng, uint32_t max, const char** names,
                                       double* ms, uint64_t* counts, uint32_t* n) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    eng->prof.drain();
    const uint32_t k = (uint32_t)eng->prof.order.size();
    if (n) *n = k;
    for (uint32_t i = 0; i < k && i < max; ++i) {
      const auto& nm = eng->prof.order[i];
      auto it = eng->prof.totals.find(nm);
      if (names) names[i] = it->first.c_str();
      if (ms) ms[i] = it->second.first;
      if (counts) counts[i] = it->second.second;
    }
  });
}

// ---------------------------------------------------------------------------
// Internal accessors (engine_internal.h)
// ---------------------------------------------------------------------------
namespace dyn {
Profiler* engine_profiler(dyn_engine* eng) { return &eng->prof; }
hipStream_t engine_stream(const dyn_engine* eng) { return eng->stream; }
const unsigned int* engine_status_device(const dyn_engine* eng) { return eng->status.p; }
bool engine_graph_safe(const dyn_engine* eng) {
  return !eng->dbg_count && (!eng->prof.enabled || eng->prof.only);
}
bool engine_single_stream(const dyn_engine* eng) { return eng->prm.single_stream != 0; }
void engine_set_stream(dyn_engine* eng, hipStream_t s) {
  g_layout_gen.fetch_add(1);
  eng->stream = s ? s : eng->own_stream;
  eng->prof.stream = eng->stream;
  eng->pme.stream = eng->stream;
  eng->own_cache.invalidate();  // packs are uploaded on the stream in use
  eng->pme.cache.invalidate();
}
const int64_t* engine_forces_device(const dyn_engine* eng) { return eng->fx_out.p; }
ForceParts engine_force_parts(const dyn_engine* eng) { return eng->parts; }
bool engine_forces_stale(const dyn_engine* eng) { return eng->fx_stale; }
EngineSet engine_self_set(dyn_engine* eng) { return eng->self_set(); }
int engine_device(const dyn_engine* eng) { return eng->device; }
void engine_sizes(const dyn_engine* eng, uint32_t* n_dyn, uint32_t* n_water) {
  *n_dyn = eng->n_dyn;
  *n_water = eng->n_water;
}
void engine_box(const dyn_engine* eng, float box[3]) {
  for (int a = 0; a < 3; ++a) box[a] = eng->box[a];
}
int engine_mode(const dyn_engine* eng) { return (int)eng->prm.mode; }

// Keep the last evaluation's outputs (site forces, SPME forces, energies)
// on the engine's stream; engine_restore puts them back and marks the list
// for a rebuild (it was built for positions the caller is discarding).
void engine_checkpoint(dyn_engine* e) {
  if (e->fx_stale) throw Error(DYN_ERR_STATE, "the last evaluation's forces were summed by its caller");
  const size_t nf = 3 * (size_t)e->n_sites;
  e->fx_out_ck.reserve(nf);
  e->eng_ck.reserve(kEngOut);
  DYN_HIP(hipMemcpyAsync(e->fx_out_ck.p, e->fx_out.p, 8 * nf, hipMemcpyDeviceToDevice, e->stream));
  DYN_HIP(hipMemcpyAsync(e->eng_ck.p, e->eng_out.p, 8 * kEngOut, hipMemcpyDeviceToDevice, e->stream));
  if (e->pme_on) {
    e->fx_recip_ck.reserve(nf);
    DYN_HIP(hipMemcpyAsync(e->fx_recip_ck.p, e->fx_recip.p, 8 * nf, hipMemcpyDeviceToDevice,
                           e->stream));
  }
  e->ck.recip_kept = e->recip_kept;
  e->ck.last_recip_b = e->last_recip_b;
  e->ck.last_sch = e->last_sch;
  e->ck.computed = e->computed;
  e->ck.energies_valid = e->energies_valid;
}

void engine_restore(dyn_engine* e) {
  const size_t nf = 3 * (size_t)e->n_sites;
  DYN_HIP(hipMemcpyAsync(e->fx_out.p, e->fx_out_ck.p, 8 * nf, hipMemcpyDeviceToDevice, e->stream));
  DYN_HIP(hipMemcpyAsync(e->eng_out.p, e->eng_ck.p, 8 * kEngOut, hipMemcpyDeviceToDevice, e->stream));
  if (e->pme_on)
    DYN_HIP(hipMemcpyAsync(e->fx_recip.p, e->fx_recip_ck.p, 8 * nf, hipMemcpyDeviceToDevice,
                           e->stream));
  e->recip_kept = e->ck.recip_kept;
  e->last_recip_b = e->ck.last_recip_b;
  e->last_sch = e->ck.last_sch;
  e->computed = e->ck.computed;
  e->energies_valid = e->ck.energies_valid;
  e->mol_valid = false;
  e->fx_fresh = false;
  e->fx_stale = false;
  e->list_valid = false;
}
}  // namespace dyn

extern "C" int dyn_engine_debug_inject_status(dyn_engine* eng, uint32_t bits, uint32_t skip,
                                              uint32_t count) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    eng->dbg_bits = bits;
    eng->dbg_skip = skip;
    eng->dbg_count = count;
  });
}
