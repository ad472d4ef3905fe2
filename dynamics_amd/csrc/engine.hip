// engine.hip — native nonbonded engine on gfx950.
//
// Pipeline of one evaluation (all on the handle's stream):
//   prepare    wrap positions into the box, per-class column binning (rebuild)
//   colsort    per column, bitonic sort by z, cut into clusters of 8 (rebuild)
//   bbox/list  cluster bounding boxes, cluster-pair list with 64-bit
//              exclusion masks, half list (each pair once)          (rebuild)
//   slots      gather current positions/charges/types into cluster slots
//   pairs      8×8 wave tiles over the list: LJ + erfc Coulomb, int64
//              fixed-point force sums, f64 energy/virial per wave
//   corr       1-4 pairs (scaled real-space) and reciprocal corrections
//   pme        spread → FFT → convolve → inverse FFT → gather
//   finalize   slot forces + site forces → site-order fixed-point forces,
//              fixed-order energy reduction
//
// Reference semantics: src/non_bonded.rs:258-412 (pair sum), :541-677 (pair
// set), :681-795 (SPME and the 1-4 correction), src/prep.rs:114-143
// (exclusions), src/cuda/cuda.cu:185-362 (kernel physics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "common.h"
#include "dyn_math.h"
#include "engine.h"
#include "pme.h"

#pragma clang fp contract(off)

using namespace dyn;

namespace {

constexpr int kWave = 64;
constexpr int kPairBlock = 256;       // 4 waves, one i-cluster row each
constexpr int kListCap = 2048;        // j-clusters per i-cluster row
constexpr int kColCap = 4096;         // sites per column
constexpr int kNumEnergy = 8;         // per-row partials

struct ClassGrid {
  int ncx, ncy;
  float wx, wy;
  int col_off;  // first global column of this class
};

struct EngGeom {
  float L[3], invL[3];
  ClassGrid cg[3];
  int n_cols;
};

__device__ __forceinline__ float wrap1(float x, float L, float invL) {
  float w = x - L * floorf(x * invL);
  if (!(w < L)) w = 0.0f;
  if (w < 0.0f) w = 0.0f;
  return w;
}

// Wrap all sites; when `bin`, also count sites per class column.
__global__ void k_eng_prepare(const float4* __restrict__ pos, uint32_t n,
                              const float* __restrict__ q, const uint8_t* __restrict__ cls,
                              EngGeom g, int bin, float4* __restrict__ wpos,
                              uint32_t* __restrict__ col_of, uint32_t* __restrict__ col_count) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float4 p = pos[s];
  const float x = wrap1(p.x, g.L[0], g.invL[0]);
  const float y = wrap1(p.y, g.L[1], g.invL[1]);
  const float z = wrap1(p.z, g.L[2], g.invL[2]);
  wpos[s] = make_float4(x, y, z, q[s]);
  if (bin) {
    const ClassGrid cg = g.cg[cls[s]];
    int cx = (int)(x / cg.wx), cy = (int)(y / cg.wy);
    cx = min(max(cx, 0), cg.ncx - 1);
    cy = min(max(cy, 0), cg.ncy - 1);
    const uint32_t c = (uint32_t)(cg.col_off + cx * cg.ncy + cy);
    col_of[s] = c;
    atomicAdd(&col_count[c], 1u);
  }
}

__global__ void k_eng_col_clusters(const uint32_t* __restrict__ cnt, uint32_t n,
                                   uint32_t* __restrict__ ncl) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) ncl[c] = (cnt[c] + DYN_CL_SIZE - 1) / DYN_CL_SIZE;
}

__global__ void k_eng_scatter(uint32_t n, const uint32_t* __restrict__ col_of,
                              const uint32_t* __restrict__ col_start,
                              uint32_t* __restrict__ cursor, uint32_t* __restrict__ items) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t c = col_of[s];
  items[col_start[c] + atomicAdd(&cursor[c], 1u)] = s;
}

// One workgroup per column: sort the column's sites by (z, site id) and cut
// them into clusters of 8 (the last one padded with −1).
constexpr int kSortBlock = 256;
__global__ void __launch_bounds__(kSortBlock)
k_eng_colsort(const uint32_t* __restrict__ col_start, const uint32_t* __restrict__ items,
              const uint32_t* __restrict__ col_cl_start, const float4* __restrict__ wpos,
              const int* __restrict__ col_class, int32_t* __restrict__ slot_site,
              uint8_t* __restrict__ cl_class, uint32_t* __restrict__ cl_col) {
  __shared__ unsigned long long key[kColCap];
  const int c = blockIdx.x;
  const uint32_t b = col_start[c], e = col_start[c + 1];
  const uint32_t n = e - b;
  const uint32_t cl0 = col_cl_start[c], ncl = col_cl_start[c + 1] - cl0;
  if (ncl == 0) return;
  uint32_t np2 = 1;
  while (np2 < n) np2 <<= 1;
  for (uint32_t k = threadIdx.x; k < np2; k += kSortBlock) {
    unsigned long long v = ~0ull;
    if (k < n) {
      const uint32_t s = items[b + k];
      v = ((unsigned long long)__float_as_uint(wpos[s].z) << 32) | s;
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
  for (uint32_t k = threadIdx.x; k < ncl * DYN_CL_SIZE; k += kSortBlock)
    slot_site[(size_t)cl0 * DYN_CL_SIZE + k] = k < n ? (int32_t)(key[k] & 0xffffffffu) : -1;
  for (uint32_t k = threadIdx.x; k < ncl; k += kSortBlock) {
    cl_class[cl0 + k] = (uint8_t)col_class[c];
    cl_col[cl0 + k] = (uint32_t)c;
  }
}

// Slot data from the current positions.  Padding slots copy the cluster's
// first site position with q = 0 and the no-LJ type.
__global__ void k_eng_slots(const int32_t* __restrict__ slot_site, uint32_t n_slots,
                            const float4* __restrict__ wpos, const int32_t* __restrict__ ltype,
                            int none_type, float4* __restrict__ xyzq,
                            int32_t* __restrict__ slot_type, uint32_t* __restrict__ slot_of,
                            int write_map) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_slots) return;
  const int32_t s = slot_site[k];
  if (s >= 0) {
    xyzq[k] = wpos[s];
    slot_type[k] = ltype[s];
    if (write_map) slot_of[s] = k;
  } else {
    const int32_t s0 = slot_site[k & ~(uint32_t)(DYN_CL_SIZE - 1)];
    float4 p = wpos[s0];
    p.w = 0.0f;
    xyzq[k] = p;
    slot_type[k] = none_type;
  }
}

__global__ void k_eng_bbox(const float4* __restrict__ xyzq, const int32_t* __restrict__ slot_site,
                           uint32_t n_cl, float4* __restrict__ bb_c, float4* __restrict__ bb_h) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cl) return;
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  for (int k = 0; k < DYN_CL_SIZE; ++k) {
    if (slot_site[c * DYN_CL_SIZE + k] < 0) continue;
    const float4 p = xyzq[c * DYN_CL_SIZE + k];
    lo[0] = fminf(lo[0], p.x);
    lo[1] = fminf(lo[1], p.y);
    lo[2] = fminf(lo[2], p.z);
    hi[0] = fmaxf(hi[0], p.x);
    hi[1] = fmaxf(hi[1], p.y);
    hi[2] = fmaxf(hi[2], p.z);
  }
  bb_c[c] = make_float4(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]), 0.f);
  bb_h[c] = make_float4(0.5f * (hi[0] - lo[0]), 0.5f * (hi[1] - lo[1]), 0.5f * (hi[2] - lo[2]), 0.f);
}

struct ListArgs {
  EngGeom g;
  float rlist2;
  const float4* bb_c;
  const float4* bb_h;
  const uint32_t* col_cl_start;
  const uint8_t* cl_class;
  const int32_t* slot_site;
  const int32_t* wmol;       // per site: water molecule or −1
  const uint32_t* ex_ptr;    // per solute site: exclusion CSR (1-2, 1-3 and 1-4)
  const uint32_t* ex_idx;
  uint32_t n_dyn;
};

__device__ __forceinline__ bool excluded(const ListArgs& A, int32_t sa, int32_t sb) {
  if (sa < 0 || sb < 0) return true;
  if ((uint32_t)sa < A.n_dyn && (uint32_t)sb < A.n_dyn) {
    uint32_t lo = A.ex_ptr[sa], hi = A.ex_ptr[sa + 1];
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      const uint32_t v = A.ex_idx[mid];
      if (v == (uint32_t)sb) return true;
      if (v < (uint32_t)sb) lo = mid + 1;
      else hi = mid;
    }
    return false;
  }
  if ((uint32_t)sa >= A.n_dyn && (uint32_t)sb >= A.n_dyn) return A.wmol[sa] == A.wmol[sb];
  return false;
}

__device__ __forceinline__ bool class_pair_allowed(int ci, int cj) {
  // Half list over class-major cluster numbering: a row lists classes ≥ its
  // own; O–charged pairs carry no interaction.
  if (cj < ci) return false;
  return dyn_kind(ci, cj) != DYN_KIND_NONE;
}

// One wave (one workgroup of 64) per i-cluster.  PASS 0 counts entries;
// PASS 1 collects them in LDS in enumeration order, builds each entry's mask
// and writes the row.
template <int PASS>
__global__ void __launch_bounds__(kWave)
k_eng_list(ListArgs A, uint32_t n_cl, uint32_t* __restrict__ row_cnt,
           const uint32_t* __restrict__ row_ptr, uint32_t* __restrict__ cj,
           unsigned long long* __restrict__ cmask) {
  __shared__ uint32_t ent[PASS == 1 ? kListCap : 1];
  const uint32_t I = blockIdx.x;
  if (I >= n_cl) return;
  const int lane = threadIdx.x;
  const int ci = A.cl_class[I];
  const float4 c = A.bb_c[I], h = A.bb_h[I];
  const float r = sqrtf(A.rlist2);
  uint32_t cnt = 0;
  for (int cls = 0; cls < 3; ++cls) {
    if (!class_pair_allowed(ci, cls)) continue;
    const ClassGrid cg = A.g.cg[cls];
    if (cg.ncx == 0) continue;
    int x0 = (int)floorf((c.x - h.x - r) / cg.wx), x1 = (int)floorf((c.x + h.x + r) / cg.wx);
    int y0 = (int)floorf((c.y - h.y - r) / cg.wy), y1 = (int)floorf((c.y + h.y + r) / cg.wy);
    if (x1 - x0 + 1 >= cg.ncx) {
      x0 = 0;
      x1 = cg.ncx - 1;
    }
    if (y1 - y0 + 1 >= cg.ncy) {
      y0 = 0;
      y1 = cg.ncy - 1;
    }
    for (int xx = x0; xx <= x1; ++xx) {
      const int cx = ((xx % cg.ncx) + cg.ncx) % cg.ncx;
      for (int yy = y0; yy <= y1; ++yy) {
        const int cy = ((yy % cg.ncy) + cg.ncy) % cg.ncy;
        const int col = cg.col_off + cx * cg.ncy + cy;
        const uint32_t js = A.col_cl_start[col], je = A.col_cl_start[col + 1];
        for (uint32_t base = js; base < je; base += kWave) {
          const uint32_t J = base + lane;
          bool ok = J < je;
          if (ok && cls == ci) ok = J >= I;
          if (ok) {
            const float4 cJ = A.bb_c[J], hJ = A.bb_h[J];
            float dx = c.x - cJ.x, dy = c.y - cJ.y, dz = c.z - cJ.z;
            dx = fabsf(dx - A.g.L[0] * rintf(dx * A.g.invL[0]));
            dy = fabsf(dy - A.g.L[1] * rintf(dy * A.g.invL[1]));
            dz = fabsf(dz - A.g.L[2] * rintf(dz * A.g.invL[2]));
            dx = fmaxf(0.f, dx - h.x - hJ.x);
            dy = fmaxf(0.f, dy - h.y - hJ.y);
            dz = fmaxf(0.f, dz - h.z - hJ.z);
            ok = dx * dx + dy * dy + dz * dz < A.rlist2;
          }
          const unsigned long long bal = __ballot(ok);
          if (PASS == 1 && ok) {
            const uint32_t pre = __popcll(bal & ((1ull << lane) - 1ull));
            const uint32_t slot = cnt + pre;
            if (slot < (uint32_t)kListCap) ent[slot] = J;
          }
          cnt += (uint32_t)__popcll(bal);
        }
      }
    }
  }
  if (PASS == 0) {
    if (lane == 0) row_cnt[I] = cnt;
    return;
  }
  __syncthreads();
  const uint32_t base = row_ptr[I];
  const int ia = lane & 7, jb = lane >> 3;
  const int32_t sa = A.slot_site[I * DYN_CL_SIZE + ia];
  const uint32_t m = min(cnt, (uint32_t)kListCap);
  for (uint32_t k = 0; k < m; ++k) {
    const uint32_t J = ent[k];
    const int32_t sb = A.slot_site[J * DYN_CL_SIZE + jb];
    bool ok = !(I == J && ia >= jb);
    if (ok) ok = !excluded(A, sa, sb);
    const unsigned long long mk = __ballot(ok);
    if (lane == 0) {
      cj[base + k] = J;
      cmask[base + k] = mk;
    }
  }
}

struct PairArgs {
  DynPairParams P;
  const float4* xyzq;
  const int32_t* slot_type;
  const uint8_t* cl_class;
  const uint32_t* row_ptr;
  const uint32_t* cj;
  const unsigned long long* cmask;
  const float2* lj_table;
  int64_t* fx;          // component-major, n_slots per component
  uint32_t n_slots;
  double* partials;     // kNumEnergy per row
  uint32_t n_rows;
};

__device__ __forceinline__ float xor_sum(float v, int m) { return v + __shfl_xor(v, m, kWave); }
__device__ __forceinline__ double xor_sum_d(double v, int m) { return v + __shfl_xor(v, m, kWave); }

template <int KIND>
__device__ __forceinline__ void tile_entry(const PairArgs& A, const float pi[4], int ti,
                                           uint32_t J, unsigned long long mask, int lane, int ia,
                                           int jb, float fi[3], double acc[kNumEnergy]) {
  const float4 pjv = A.xyzq[J * DYN_CL_SIZE + jb];
  const float pj[4] = {pjv.x, pjv.y, pjv.z, pjv.w};
  const int tj = (KIND == DYN_KIND_SS || KIND == DYN_KIND_SO) ? A.slot_type[J * DYN_CL_SIZE + jb] : 0;
  const bool active = (mask >> lane) & 1ull;
  float d[3], s, elj, ec, r2;
  int hit;
  dyn_tile_pair<KIND>(pi, ti, pj, tj, active, A.P, A.lj_table, d, &s, &elj, &ec, &r2, &hit);
  const float f0 = s * d[0], f1 = s * d[1], f2 = s * d[2];
  fi[0] += f0;
  fi[1] += f1;
  fi[2] += f2;
  const bool solute = KIND == DYN_KIND_SS || KIND == DYN_KIND_SO || KIND == DYN_KIND_SQ;
  if (KIND != DYN_KIND_SQ && KIND != DYN_KIND_QQ) acc[solute ? 1 : 0] += (double)elj;
  if (KIND != DYN_KIND_SO && KIND != DYN_KIND_OO) acc[solute ? 3 : 2] += (double)ec;
  acc[4] += (double)(s * r2);
  // interacting pairs by work kind: LJ + Coulomb, LJ only, Coulomb only
  acc[KIND == DYN_KIND_SS ? 5 : ((KIND == DYN_KIND_SO || KIND == DYN_KIND_OO) ? 6 : 7)] += (double)hit;
  // Force on the j site: −f summed over the 8 i lanes of this j group.
  float g0 = -f0, g1 = -f1, g2 = -f2;
  g0 = xor_sum(g0, 1);
  g1 = xor_sum(g1, 1);
  g2 = xor_sum(g2, 1);
  g0 = xor_sum(g0, 2);
  g1 = xor_sum(g1, 2);
  g2 = xor_sum(g2, 2);
  g0 = xor_sum(g0, 4);
  g1 = xor_sum(g1, 4);
  g2 = xor_sum(g2, 4);
  if (ia == 0 && (mask >> (8 * jb)) & 0xffull) {
    const uint32_t slot = J * DYN_CL_SIZE + jb;
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + slot), (unsigned long long)dyn_fx(g0));
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + A.n_slots + slot), (unsigned long long)dyn_fx(g1));
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + 2 * (size_t)A.n_slots + slot), (unsigned long long)dyn_fx(g2));
  }
}

// One wave per i-cluster row; lane = ia + 8·jb holds pair (slot 8I+ia, slot 8J+jb).
__global__ void __launch_bounds__(kPairBlock)
k_eng_pairs(PairArgs A) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t I = blockIdx.x * (kPairBlock / kWave) + (threadIdx.x / kWave);
  if (I >= A.n_rows) return;
  const int ia = lane & 7, jb = lane >> 3;
  const float4 piv = A.xyzq[I * DYN_CL_SIZE + ia];
  const float pi[4] = {piv.x, piv.y, piv.z, piv.w};
  const int ti = A.slot_type[I * DYN_CL_SIZE + ia];
  const int ci = A.cl_class[I];
  float fi[3] = {0.f, 0.f, 0.f};
  double acc[kNumEnergy];
  for (int k = 0; k < kNumEnergy; ++k) acc[k] = 0.0;
  const uint32_t rs = A.row_ptr[I], re = A.row_ptr[I + 1];
  for (uint32_t e = rs; e < re; ++e) {
    const uint32_t J = __builtin_amdgcn_readfirstlane(A.cj[e]);
    const unsigned long long mask = A.cmask[e];
    const int kind = dyn_kind(ci, A.cl_class[J]);
    switch (kind) {
      case DYN_KIND_QQ: tile_entry<DYN_KIND_QQ>(A, pi, ti, J, mask, lane, ia, jb, fi, acc); break;
      case DYN_KIND_OO: tile_entry<DYN_KIND_OO>(A, pi, ti, J, mask, lane, ia, jb, fi, acc); break;
      case DYN_KIND_SS: tile_entry<DYN_KIND_SS>(A, pi, ti, J, mask, lane, ia, jb, fi, acc); break;
      case DYN_KIND_SO: tile_entry<DYN_KIND_SO>(A, pi, ti, J, mask, lane, ia, jb, fi, acc); break;
      case DYN_KIND_SQ: tile_entry<DYN_KIND_SQ>(A, pi, ti, J, mask, lane, ia, jb, fi, acc); break;
      default: break;
    }
  }
  // i forces: sum over the 8 j groups (lanes ia, ia+8, …, ia+56).
  for (int c = 0; c < 3; ++c) {
    float v = fi[c];
    v = xor_sum(v, 8);
    v = xor_sum(v, 16);
    v = xor_sum(v, 32);
    fi[c] = v;
  }
  if (jb == 0 && re > rs) {
    const uint32_t slot = I * DYN_CL_SIZE + ia;
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + slot), (unsigned long long)dyn_fx(fi[0]));
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + A.n_slots + slot), (unsigned long long)dyn_fx(fi[1]));
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + 2 * (size_t)A.n_slots + slot), (unsigned long long)dyn_fx(fi[2]));
  }
  for (int k = 0; k < kNumEnergy; ++k) {
    double v = acc[k];
    for (int m = 1; m < kWave; m <<= 1) v = xor_sum_d(v, m);
    acc[k] = v;
  }
  if (lane == 0)
    for (int k = 0; k < kNumEnergy; ++k) A.partials[(size_t)I * kNumEnergy + k] = acc[k];
}

// Correction pairs: kind 1 = 1-4 pair (scaled real-space LJ + Coulomb, plus
// the reciprocal correction of the mode), kind 0 = excluded pair (standard
// mode: remove qq·erf(βr)/r from the reciprocal sum).
struct CorrArgs {
  DynPairParams P;
  const float4* wpos;      // per site (x, y, z, q)
  const int32_t* ltype;    // per site
  const float2* lj_table;
  const uint32_t* ci;
  const uint32_t* cj;
  const uint8_t* kind;
  uint32_t n;
  int mode;
  int recip_on;            // reciprocal corrections only accompany the SPME sum
  int64_t* fx;             // site forces, component-major
  uint32_t n_sites;
  double* partials;        // 6 per block
};

constexpr int kCorrBlock = 256;
__global__ void __launch_bounds__(kCorrBlock)
k_eng_corr(CorrArgs A) {
  __shared__ double red[kCorrBlock];
  double acc[6] = {0, 0, 0, 0, 0, 0};  // e_lj14, e_c14, e_corr, vir_sr, vir_recip, unused
  for (uint32_t k = blockIdx.x * kCorrBlock + threadIdx.x; k < A.n; k += gridDim.x * kCorrBlock) {
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
    if (A.kind[k] == 0 && !A.recip_on) continue;
    if (A.kind[k] == 1) {
      if (r2 <= A.P.lj_rc2) {
        const float2 c = A.lj_table[A.ltype[i] * A.P.n_types + A.ltype[j]];
        float sl, el;
        dyn_lj(rinv * rinv, c.x, c.y, &sl, &el);
        s_sr += DYN_SCALE_LJ_14 * sl;
        acc[0] += (double)(DYN_SCALE_LJ_14 * el);
      }
      if (r2 < A.P.rc2) {
        float sc, ec;
        dyn_coul_erfc(r2, rinv, qq, A.P.beta, A.P.beta2, A.P.two_b_sqrtpi, &sc, &ec);
        s_sr += 0.833333333f * sc;
        acc[1] += (double)(0.833333333f * ec);
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
    const float s = s_sr + s_rc;
    const int64_t q0 = dyn_fx(s * d0), q1 = dyn_fx(s * d1), q2 = dyn_fx(s * d2);
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + i), (unsigned long long)q0);
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + A.n_sites + i), (unsigned long long)q1);
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + 2 * (size_t)A.n_sites + i), (unsigned long long)q2);
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + j), (unsigned long long)(-q0));
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + A.n_sites + j), (unsigned long long)(-q1));
    atomicAdd(reinterpret_cast<unsigned long long*>(A.fx + 2 * (size_t)A.n_sites + j), (unsigned long long)(-q2));
  }
  for (int v = 0; v < 5; ++v) {
    red[threadIdx.x] = acc[v];
    __syncthreads();
    for (int s = kCorrBlock / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) A.partials[6 * blockIdx.x + v] = red[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) A.partials[6 * blockIdx.x + 5] = 0.0;
}

__global__ void k_eng_pme_input(const float4* __restrict__ wpos, const uint32_t* __restrict__ qsite,
                                uint32_t nq, float4* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nq) out[k] = wpos[qsite[k]];
}

// Site-order forces: site accumulators plus the slot accumulators of the
// pair kernel.
__global__ void k_eng_finalize(const int64_t* __restrict__ fx_site,
                               const int64_t* __restrict__ fx_slot,
                               const uint32_t* __restrict__ slot_of, uint32_t n_sites,
                               uint32_t n_slots, int64_t* __restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sites) return;
  const uint32_t k = slot_of[s];
  for (int c = 0; c < 3; ++c)
    out[(size_t)c * n_sites + s] = fx_site[(size_t)c * n_sites + s] + fx_slot[(size_t)c * n_slots + k];
}

// Fixed-order sums of the per-row and per-block partials.
__global__ void k_eng_energy_reduce(const double* __restrict__ pp, uint32_t n_rows,
                                    const double* __restrict__ cp, uint32_t n_cblk,
                                    double* __restrict__ out) {
  __shared__ double red[256];
  for (int v = 0; v < 14; ++v) {
    double a = 0.0;
    if (v < kNumEnergy) {
      for (uint32_t r = threadIdx.x; r < n_rows; r += 256) a += pp[(size_t)r * kNumEnergy + v];
    } else {
      const int w = v - kNumEnergy;
      for (uint32_t r = threadIdx.x; r < n_cblk; r += 256) a += cp[(size_t)r * 6 + w];
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[v] = red[0];
    __syncthreads();
  }
}

__global__ void k_eng_max_disp(const float4* __restrict__ wpos, const float4* __restrict__ ref,
                               uint32_t n, EngGeom g, unsigned int* __restrict__ out_bits) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float4 a = wpos[s], b = ref[s];
  float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  dx -= g.L[0] * rintf(dx * g.invL[0]);
  dy -= g.L[1] * rintf(dy * g.invL[1]);
  dz -= g.L[2] * rintf(dz * g.invL[2]);
  atomicMax(out_bits, __float_as_uint(dx * dx + dy * dy + dz * dz));
}

// scan helpers (single workgroup), as in compat.hip
constexpr int kScanBlock = 1024;
__global__ void __launch_bounds__(kScanBlock)
k_eng_scan(const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ out,
           uint32_t* __restrict__ maxv) {
  __shared__ uint32_t sh[kScanBlock];
  __shared__ uint32_t shm[kScanBlock];
  const int t = threadIdx.x;
  const uint32_t per = (n + kScanBlock - 1) / kScanBlock;
  const uint32_t b = t * per;
  const uint32_t e = min(n, b + per);
  uint32_t s = 0, m = 0;
  for (uint32_t i = b; i < e; ++i) {
    s += in[i];
    m = max(m, in[i]);
  }
  sh[t] = s;
  shm[t] = m;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const uint32_t v = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += v;
    __syncthreads();
  }
  for (int off = kScanBlock / 2; off > 0; off >>= 1) {
    if (t < off) shm[t] = max(shm[t], shm[t + off]);
    __syncthreads();
  }
  uint32_t run = sh[t] - s;
  for (uint32_t i = b; i < e; ++i) {
    out[i] = run;
    run += in[i];
  }
  if (t == kScanBlock - 1) out[n] = sh[t];
  if (t == 0 && maxv) *maxv = shm[0];
}

}  // namespace

// ---------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------
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
  double e_self = 0.0, q_total = 0.0;
  // topology
  DevBuf<float> q;
  DevBuf<int32_t> ltype;
  DevBuf<uint8_t> cls;
  DevBuf<int32_t> wmol;
  DevBuf<float2> lj_table;
  DevBuf<uint32_t> ex_ptr, ex_idx;
  DevBuf<uint32_t> corr_i, corr_j;
  DevBuf<uint8_t> corr_kind;
  uint32_t n_corr = 0;
  DevBuf<uint32_t> qsite;
  uint32_t n_q = 0;
  float c6_oo = 0.f, c12_oo = 0.f;
  // positions
  DevBuf<float4> pos_own;
  const float4* pos_dev = nullptr;
  PinnedBuf<float4> h_pos;
  DevBuf<float4> wpos, wpos_ref;
  // clustering
  EngGeom geom{};
  std::vector<int> h_col_class;
  DevBuf<int> col_class;
  DevBuf<uint32_t> col_of, col_count, col_start, col_ncl, col_cl_start, cursor, items, scan_max;
  uint32_t n_cols = 0, n_cl = 0, n_slots = 0;
  DevBuf<int32_t> slot_site, slot_type;
  DevBuf<uint8_t> cl_class;
  DevBuf<uint32_t> cl_col, slot_of;
  DevBuf<float4> xyzq, bb_c, bb_h;
  DevBuf<uint32_t> row_cnt, row_ptr, cj;
  DevBuf<unsigned long long> cmask;
  uint64_t n_entries = 0, n_builds = 0;
  // accumulators
  DevBuf<int64_t> fx_slot, fx_site, fx_out;
  DevBuf<double> pair_part, corr_part, eng_out;
  unsigned corr_blocks = 0;
  DevBuf<unsigned int> disp_bits;
  // PME
  bool pme_on = false;
  PmeCore pme;
  DevBuf<float4> pme_xyzq;
  // host copies
  PinnedBuf<int64_t> h_fx;
  PinnedBuf<double> h_eng;
  bool fx_fresh = false;

  ~dyn_engine() {
    if (stream) (void)hipStreamDestroy(stream);
  }

  void set_geometry();
  void build_list();
  void compute(uint32_t flags);
};

void dyn_engine::set_geometry() {
  const float V = box[0] * box[1] * box[2];
  uint32_t ncls[3] = {n_dyn, n_water, 3 * n_water};
  int off = 0;
  h_col_class.clear();
  for (int c = 0; c < 3; ++c) {
    ClassGrid& g = geom.cg[c];
    if (ncls[c] == 0) {
      g.ncx = g.ncy = 0;
      g.wx = g.wy = 1.f;
      g.col_off = off;
      continue;
    }
    const double s = std::cbrt(8.0 * (double)V / (double)ncls[c]);
    g.ncx = std::max(1, std::min(256, (int)std::floor(box[0] / s)));
    g.ncy = std::max(1, std::min(256, (int)std::floor(box[1] / s)));
    g.wx = box[0] / g.ncx;
    g.wy = box[1] / g.ncy;
    g.col_off = off;
    off += g.ncx * g.ncy;
    for (int k = 0; k < g.ncx * g.ncy; ++k) h_col_class.push_back(c);
  }
  geom.n_cols = off;
  n_cols = (uint32_t)off;
  for (int a = 0; a < 3; ++a) {
    geom.L[a] = box[a];
    geom.invL[a] = 1.0f / box[a];
  }
  col_class.reserve(std::max<uint32_t>(n_cols, 1));
  col_of.reserve(n_sites);
  col_count.reserve(n_cols + 1);
  col_start.reserve(n_cols + 1);
  col_ncl.reserve(n_cols + 1);
  col_cl_start.reserve(n_cols + 1);
  cursor.reserve(n_cols + 1);
  items.reserve(n_sites);
  scan_max.reserve(2);
  DYN_HIP(hipMemcpyAsync(col_class.p, h_col_class.data(), sizeof(int) * n_cols,
                         hipMemcpyHostToDevice, stream));
  DYN_HIP(hipStreamSynchronize(stream));
  list_valid = false;
}

void dyn_engine::build_list() {
  // bin
  DYN_HIP(hipMemsetAsync(col_count.p, 0, 4 * (size_t)n_cols, stream));
  DYN_HIP(hipMemsetAsync(cursor.p, 0, 4 * (size_t)n_cols, stream));
  {
    Profiler::Scope ps(prof, "nbl_prepare_bin");
    hipLaunchKernelGGL(k_eng_prepare, dim3(grid_for(n_sites, 256)), dim3(256), 0, stream,
                       pos_dev, n_sites, q.p, cls.p, geom, 1, wpos.p, col_of.p, col_count.p);
    DYN_CHECK_LAUNCH("k_eng_prepare");
  }
  {
    Profiler::Scope ps(prof, "nbl_scan");
    hipLaunchKernelGGL(k_eng_scan, dim3(1), dim3(kScanBlock), 0, stream, col_count.p, n_cols,
                       col_start.p, scan_max.p);
    hipLaunchKernelGGL(k_eng_col_clusters, dim3(grid_for(n_cols, 256)), dim3(256), 0, stream,
                       col_count.p, n_cols, col_ncl.p);
    hipLaunchKernelGGL(k_eng_scan, dim3(1), dim3(kScanBlock), 0, stream, col_ncl.p, n_cols,
                       col_cl_start.p, (uint32_t*)nullptr);
    DYN_CHECK_LAUNCH("k_eng_scan");
  }
  uint32_t hv[2] = {0, 0};
  DYN_HIP(hipMemcpyAsync(&hv[0], scan_max.p, 4, hipMemcpyDeviceToHost, stream));
  DYN_HIP(hipMemcpyAsync(&hv[1], col_cl_start.p + n_cols, 4, hipMemcpyDeviceToHost, stream));
  DYN_HIP(hipStreamSynchronize(stream));
  if (hv[0] > (uint32_t)kColCap)
    throw Error(DYN_ERR_CAPACITY, fmt("a cluster column holds %u sites; at most %d are supported",
                                      hv[0], kColCap));
  n_cl = hv[1];
  n_slots = n_cl * DYN_CL_SIZE;
  slot_site.reserve(n_slots);
  slot_type.reserve(n_slots);
  cl_class.reserve(n_cl);
  cl_col.reserve(n_cl);
  xyzq.reserve(n_slots);
  bb_c.reserve(n_cl);
  bb_h.reserve(n_cl);
  row_cnt.reserve(n_cl + 1);
  row_ptr.reserve(n_cl + 1);
  fx_slot.reserve(3 * (size_t)n_slots);
  pair_part.reserve((size_t)kNumEnergy * n_cl);
  {
    Profiler::Scope ps(prof, "nbl_scatter");
    hipLaunchKernelGGL(k_eng_scatter, dim3(grid_for(n_sites, 256)), dim3(256), 0, stream,
                       n_sites, col_of.p, col_start.p, cursor.p, items.p);
    DYN_CHECK_LAUNCH("k_eng_scatter");
  }
  {
    Profiler::Scope ps(prof, "nbl_colsort");
    hipLaunchKernelGGL(k_eng_colsort, dim3(n_cols), dim3(kSortBlock), 0, stream, col_start.p,
                       items.p, col_cl_start.p, wpos.p, col_class.p, slot_site.p, cl_class.p,
                       cl_col.p);
    DYN_CHECK_LAUNCH("k_eng_colsort");
  }
  {
    Profiler::Scope ps(prof, "nbl_slots_bbox");
    hipLaunchKernelGGL(k_eng_slots, dim3(grid_for(n_slots, 256)), dim3(256), 0, stream,
                       slot_site.p, n_slots, wpos.p, ltype.p, type_none, xyzq.p, slot_type.p,
                       slot_of.p, 1);
    hipLaunchKernelGGL(k_eng_bbox, dim3(grid_for(n_cl, 256)), dim3(256), 0, stream, xyzq.p,
                       slot_site.p, n_cl, bb_c.p, bb_h.p);
    DYN_CHECK_LAUNCH("k_eng_bbox");
  }
  ListArgs A;
  A.g = geom;
  const float rl = prm.cutoff + std::max(0.f, prm.list_buffer);
  A.rlist2 = rl * rl;
  A.bb_c = bb_c.p;
  A.bb_h = bb_h.p;
  A.col_cl_start = col_cl_start.p;
  A.cl_class = cl_class.p;
  A.slot_site = slot_site.p;
  A.wmol = wmol.p;
  A.ex_ptr = ex_ptr.p;
  A.ex_idx = ex_idx.p;
  A.n_dyn = n_dyn;
  {
    Profiler::Scope ps(prof, "nbl_list_count");
    hipLaunchKernelGGL(k_eng_list<0>, dim3(n_cl), dim3(kWave), 0, stream, A, n_cl, row_cnt.p,
                       (const uint32_t*)nullptr, (uint32_t*)nullptr,
                       (unsigned long long*)nullptr);
    DYN_CHECK_LAUNCH("k_eng_list<0>");
    hipLaunchKernelGGL(k_eng_scan, dim3(1), dim3(kScanBlock), 0, stream, row_cnt.p, n_cl,
                       row_ptr.p, scan_max.p);
    DYN_CHECK_LAUNCH("k_eng_scan");
  }
  DYN_HIP(hipMemcpyAsync(&hv[0], scan_max.p, 4, hipMemcpyDeviceToHost, stream));
  DYN_HIP(hipMemcpyAsync(&hv[1], row_ptr.p + n_cl, 4, hipMemcpyDeviceToHost, stream));
  DYN_HIP(hipStreamSynchronize(stream));
  if (hv[0] > (uint32_t)kListCap)
    throw Error(DYN_ERR_CAPACITY, fmt("a cluster has %u cluster-pair entries; at most %d are "
                                      "supported (reduce list_buffer)", hv[0], kListCap));
  n_entries = hv[1];
  cj.reserve(std::max<uint64_t>(n_entries, 1));
  cmask.reserve(std::max<uint64_t>(n_entries, 1));
  {
    Profiler::Scope ps(prof, "nbl_list_fill");
    hipLaunchKernelGGL(k_eng_list<1>, dim3(n_cl), dim3(kWave), 0, stream, A, n_cl,
                       (uint32_t*)nullptr, row_ptr.p, cj.p, cmask.p);
    DYN_CHECK_LAUNCH("k_eng_list<1>");
  }
  DYN_HIP(hipMemcpyAsync(wpos_ref.p, wpos.p, sizeof(float4) * n_sites, hipMemcpyDeviceToDevice,
                         stream));
  list_valid = true;
  ++n_builds;
}

void dyn_engine::compute(uint32_t flags) {
  DYN_REQUIRE(box_set, "box not set");
  if (!pos_set) throw Error(DYN_ERR_STATE, "positions not set");
  bool rebuild = (flags & DYN_COMPUTE_REBUILD) || !list_valid;
  if (!rebuild) {
    if (prm.list_buffer <= 0.f) {
      rebuild = true;
    } else {
      hipLaunchKernelGGL(k_eng_prepare, dim3(grid_for(n_sites, 256)), dim3(256), 0, stream,
                         pos_dev, n_sites, q.p, cls.p, geom, 0, wpos.p, (uint32_t*)nullptr,
                         (uint32_t*)nullptr);
      DYN_CHECK_LAUNCH("k_eng_prepare");
      DYN_HIP(hipMemsetAsync(disp_bits.p, 0, 4, stream));
      hipLaunchKernelGGL(k_eng_max_disp, dim3(grid_for(n_sites, 256)), dim3(256), 0, stream,
                         wpos.p, wpos_ref.p, n_sites, geom, disp_bits.p);
      unsigned int bits = 0;
      DYN_HIP(hipMemcpyAsync(&bits, disp_bits.p, 4, hipMemcpyDeviceToHost, stream));
      DYN_HIP(hipStreamSynchronize(stream));
      float d2;
      memcpy(&d2, &bits, 4);
      const float half = 0.5f * prm.list_buffer;
      if (!(d2 < half * half)) rebuild = true;
    }
  }
  if (rebuild) {
    build_list();
  } else {
    Profiler::Scope ps(prof, "slots");
    hipLaunchKernelGGL(k_eng_slots, dim3(grid_for(n_slots, 256)), dim3(256), 0, stream,
                       slot_site.p, n_slots, wpos.p, ltype.p, type_none, xyzq.p, slot_type.p,
                       slot_of.p, 0);
    DYN_CHECK_LAUNCH("k_eng_slots");
  }
  DynPairParams P;
  for (int a = 0; a < 3; ++a) {
    P.ext[a] = box[a];
    P.inv_ext[a] = 1.0f / box[a];
  }
  P.rc2 = prm.cutoff * prm.cutoff;
  P.lj_rc2 = prm.lj_cutoff * prm.lj_cutoff;
  P.beta = prm.beta;
  P.beta2 = prm.beta * prm.beta;
  P.two_b_sqrtpi = 2.0f * prm.beta * DYN_INV_SQRT_PI;
  P.c6_oo = c6_oo;
  P.c12_oo = c12_oo;
  P.n_types = n_types;
  DYN_HIP(hipMemsetAsync(fx_slot.p, 0, sizeof(int64_t) * 3 * n_slots, stream));
  DYN_HIP(hipMemsetAsync(fx_site.p, 0, sizeof(int64_t) * 3 * n_sites, stream));
  DYN_HIP(hipMemsetAsync(pair_part.p, 0, sizeof(double) * kNumEnergy * n_cl, stream));
  DYN_HIP(hipMemsetAsync(corr_part.p, 0, sizeof(double) * 6 * corr_blocks, stream));
  if (flags & DYN_COMPUTE_SHORT) {
    Profiler::Scope ps(prof, "pairs");
    PairArgs A;
    A.P = P;
    A.xyzq = xyzq.p;
    A.slot_type = slot_type.p;
    A.cl_class = cl_class.p;
    A.row_ptr = row_ptr.p;
    A.cj = cj.p;
    A.cmask = cmask.p;
    A.lj_table = lj_table.p;
    A.fx = fx_slot.p;
    A.n_slots = n_slots;
    A.partials = pair_part.p;
    A.n_rows = n_cl;
    const unsigned blocks = (n_cl + (kPairBlock / kWave) - 1) / (kPairBlock / kWave);
    hipLaunchKernelGGL(k_eng_pairs, dim3(blocks), dim3(kPairBlock), 0, stream, A);
    DYN_CHECK_LAUNCH("k_eng_pairs");
  }
  if ((flags & DYN_COMPUTE_CORR) && n_corr) {
    Profiler::Scope ps(prof, "corr");
    CorrArgs C;
    C.P = P;
    C.wpos = wpos.p;
    C.ltype = ltype.p;
    C.lj_table = lj_table.p;
    C.ci = corr_i.p;
    C.cj = corr_j.p;
    C.kind = corr_kind.p;
    C.n = n_corr;
    C.mode = (int)prm.mode;
    C.recip_on = (pme_on && (flags & DYN_COMPUTE_RECIP)) ? 1 : 0;
    C.fx = fx_site.p;
    C.n_sites = n_sites;
    C.partials = corr_part.p;
    hipLaunchKernelGGL(k_eng_corr, dim3(corr_blocks), dim3(kCorrBlock), 0, stream, C);
    DYN_CHECK_LAUNCH("k_eng_corr");
  }
  const bool recip_res_needed = pme_on && (flags & DYN_COMPUTE_RECIP);
  if (recip_res_needed) {
    {
      Profiler::Scope ps(prof, "pme_input");
      hipLaunchKernelGGL(k_eng_pme_input, dim3(grid_for(n_q, 256)), dim3(256), 0, stream, wpos.p,
                         qsite.p, n_q, pme_xyzq.p);
      DYN_CHECK_LAUNCH("k_eng_pme_input");
    }
    pme.run(n_q, pme_xyzq.p, qsite.p, fx_site.p, 1, n_sites, nullptr);
  }
  {
    Profiler::Scope ps(prof, "finalize");
    hipLaunchKernelGGL(k_eng_finalize, dim3(grid_for(n_sites, 256)), dim3(256), 0, stream,
                       fx_site.p, fx_slot.p, slot_of.p, n_sites, n_slots, fx_out.p);
    hipLaunchKernelGGL(k_eng_energy_reduce, dim3(1), dim3(256), 0, stream, pair_part.p, n_cl,
                       corr_part.p, corr_blocks, eng_out.p);
    DYN_CHECK_LAUNCH("k_eng_energy_reduce");
  }
  if (recip_res_needed)
    DYN_HIP(hipMemcpyAsync(eng_out.p + 14, pme.res_dev.p, 2 * sizeof(double),
                           hipMemcpyDeviceToDevice, stream));
  else
    DYN_HIP(hipMemsetAsync(eng_out.p + 14, 0, 2 * sizeof(double), stream));
  computed = true;
  fx_fresh = false;
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
    DYN_REQUIRE(p.list_buffer >= 0.f && p.list_buffer < 10.f, "list_buffer must be in [0, 10)");
    DYN_REQUIRE(n_dyn + 4ull * n_water > 0, "empty system");
    DYN_REQUIRE(n_dyn + 4ull * n_water < (1ull << 31), "system too large");
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
    E->prof.stream = E->stream;
    E->prm = p;
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
    std::vector<float2> tab((size_t)nt * nt);
    for (int a = 0; a < nt; ++a)
      for (int b = 0; b < nt; ++b) {
        // Lorentz–Berthelot in f32 (src/non_bonded.rs:1081-1086), then
        // c6 = 4εσ⁶, c12 = 4εσ¹² in f64.
        const float s_ab = 0.5f * (sig[a] + sig[b]);
        const float e_ab = std::sqrt(eps[a] * eps[b]);
        const double s6 = std::pow((double)s_ab, 6.0);
        tab[(size_t)a * nt + b] = make_float2((float)(4.0 * e_ab * s6), (float)(4.0 * e_ab * s6 * s6));
      }
    E->c6_oo = tab[(size_t)E->type_o * nt + E->type_o].x;
    E->c12_oo = tab[(size_t)E->type_o * nt + E->type_o].y;
    // per-site topology
    const float qh = 0.6791f * 18.2223f, qm = -2.0f * qh;
    std::vector<float> hq(ns);
    std::vector<int32_t> ht(ns), hw(ns);
    std::vector<uint8_t> hc(ns);
    for (uint32_t i = 0; i < n_dyn; ++i) {
      DYN_REQUIRE(type_dyn[i] >= 0 && (uint32_t)type_dyn[i] < n_types, "atom %u: bad type", i);
      DYN_REQUIRE(std::isfinite(q_dyn[i]), "atom %u: non-finite charge", i);
      hq[i] = q_dyn[i];
      ht[i] = type_dyn[i];
      hw[i] = -1;
      hc[i] = DYN_CLS_SOLUTE;
    }
    const float wq[4] = {0.f, qm, qh, qh};
    for (uint32_t m = 0; m < n_water; ++m)
      for (int k = 0; k < 4; ++k) {
        const uint32_t s = n_dyn + 4 * m + k;
        hq[s] = wq[k];
        ht[s] = k == 0 ? E->type_o : E->type_none;
        hw[s] = (int32_t)m;
        hc[s] = k == 0 ? DYN_CLS_WATER_O : DYN_CLS_WATER_Q;
      }
    // exclusions (1-2, 1-3 and 1-4 all leave the main kernel), CSR both ways
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
    up(E->cls, hc);
    up(E->lj_table, tab);
    up(E->ex_ptr, hptr);
    up(E->ex_idx, hidx);
    up(E->corr_i, ci);
    up(E->corr_j, cjv);
    up(E->corr_kind, ck);
    up(E->qsite, hqs);
    E->n_corr = (uint32_t)ci.size();
    E->corr_blocks = grid_for(std::max<uint32_t>(E->n_corr, 1), kCorrBlock, 256);
    E->corr_part.reserve(6 * (size_t)E->corr_blocks);
    E->pos_own.reserve(ns);
    E->h_pos.reserve(ns);
    E->wpos.reserve(ns);
    E->wpos_ref.reserve(ns);
    E->slot_of.reserve(ns);
    E->fx_site.reserve(3 * (size_t)ns);
    E->fx_out.reserve(3 * (size_t)ns);
    E->h_fx.reserve(3 * (size_t)ns);
    E->eng_out.reserve(16);
    E->h_eng.reserve(16);
    E->disp_bits.reserve(1);
    E->pme_xyzq.reserve(std::max<uint32_t>(E->n_q, 1));
    E->pme_on = p.grid[0] > 0 && p.grid[1] > 0 && p.grid[2] > 0;
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
    DeviceGuard g(eng->device);
    DYN_REQUIRE(box, "box is null");
    for (int a = 0; a < 3; ++a) {
      DYN_REQUIRE(box[a] > 0.f && std::isfinite(box[a]), "box lengths must be > 0");
      DYN_REQUIRE(eng->prm.cutoff + eng->prm.list_buffer < 0.5f * box[a],
                  "cutoff + list_buffer must be below half the box (minimum image)");
    }
    for (int a = 0; a < 3; ++a) eng->box[a] = box[a];
    eng->box_set = true;
    eng->set_geometry();
    if (eng->pme_on) {
      if (eng->pme.ngrid == 0)
        eng->pme.init(eng->device, eng->stream, &eng->prof, eng->prm.grid, eng->box,
                      eng->prm.beta, (int)eng->prm.order);
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
    DYN_REQUIRE(xyzw_device, "null device pointer");
    DYN_REQUIRE(((uintptr_t)xyzw_device & 15) == 0, "device positions must be 16-byte aligned");
    eng->pos_dev = reinterpret_cast<const float4*>(xyzw_device);
    eng->pos_set = true;
  });
}

extern "C" int dyn_engine_compute(dyn_engine* eng, uint32_t flags) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    eng->compute(flags);
    if (!(flags & DYN_COMPUTE_ASYNC)) DYN_HIP(hipStreamSynchronize(eng->stream));
  });
}

extern "C" int dyn_engine_sync(dyn_engine* eng) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    DYN_HIP(hipStreamSynchronize(eng->stream));
  });
}

static void fetch_fx(dyn_engine* eng) {
  if (!eng->computed) throw Error(DYN_ERR_STATE, "no compute yet");
  if (eng->fx_fresh) return;
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
    DYN_HIP(hipMemcpyAsync(eng->h_eng.p, eng->eng_out.p, 16 * sizeof(double),
                           hipMemcpyDeviceToHost, eng->stream));
    DYN_HIP(hipStreamSynchronize(eng->stream));
    const double* v = eng->h_eng.p;
    // v[0..7]: e_lj_ww, e_lj_s, e_c_ww, e_c_s, virial_pairs, pair counts by kind
    // v[8..12]: e_lj14, e_c14, e_corr, vir_sr14, vir_recip_corr; v[14..15]: pme energy, virial
    dyn_energies E{};
    E.lj = v[0] + v[1] + v[8];
    E.coul_sr = v[2] + v[3] + v[9];
    E.recip = eng->pme_on ? v[14] : 0.0;
    double self = eng->pme_on ? eng->e_self : 0.0;
    if (eng->pme_on && eng->prm.mode == DYN_MODE_STANDARD) {
      const double V = (double)eng->box[0] * eng->box[1] * eng->box[2];
      self += -M_PI * eng->q_total * eng->q_total / (2.0 * V * (double)eng->prm.beta * eng->prm.beta);
    }
    E.self = self;
    E.excl_corr = v[10];
    E.total = E.lj + E.coul_sr + E.recip + E.self + E.excl_corr;
    E.reference = v[1] + v[3] + v[8] + v[9] + E.recip + E.self;
    E.virial_sr = v[4] + v[11];
    E.virial_recip = (eng->pme_on ? v[15] : 0.0) + v[12];
    E.dh_dl = 0.0;
    E.pair_count_kind[0] = (uint64_t)std::llround(v[5]);
    E.pair_count_kind[1] = (uint64_t)std::llround(v[6]);
    E.pair_count_kind[2] = (uint64_t)std::llround(v[7]);
    E.pair_count = E.pair_count_kind[0] + E.pair_count_kind[1] + E.pair_count_kind[2];
    E.list_entries = eng->n_entries;
    E.list_builds = eng->n_builds;
    *out = E;
  });
}

extern "C" int dyn_engine_profile(dyn_engine* eng, int enable) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    eng->prof.reset();
    eng->prof.enabled = enable != 0;
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

extern "C" int dyn_engine_debug_export(dyn_engine* eng, uint32_t* n_slots, uint32_t* n_clusters,
                                       uint64_t* n_entries, int32_t* slot_site, float* xyzq,
                                       int32_t* slot_type, uint8_t* cl_class, uint32_t* row_ptr,
                                       uint32_t* cj, uint64_t* cmask, int64_t* fx_slot) {
  ENG_CHECK(eng);
  DYN_API_BODY(eng, {
    DeviceGuard g(eng->device);
    if (!eng->list_valid) throw Error(DYN_ERR_STATE, "no list built yet");
    if (n_slots) *n_slots = eng->n_slots;
    if (n_clusters) *n_clusters = eng->n_cl;
    if (n_entries) *n_entries = eng->n_entries;
    hipStream_t s = eng->stream;
    const size_t ns = eng->n_slots, nc = eng->n_cl, ne = eng->n_entries;
    if (slot_site) DYN_HIP(hipMemcpyAsync(slot_site, eng->slot_site.p, 4 * ns, hipMemcpyDeviceToHost, s));
    if (xyzq) DYN_HIP(hipMemcpyAsync(xyzq, eng->xyzq.p, 16 * ns, hipMemcpyDeviceToHost, s));
    if (slot_type) DYN_HIP(hipMemcpyAsync(slot_type, eng->slot_type.p, 4 * ns, hipMemcpyDeviceToHost, s));
    if (cl_class) DYN_HIP(hipMemcpyAsync(cl_class, eng->cl_class.p, nc, hipMemcpyDeviceToHost, s));
    if (row_ptr) DYN_HIP(hipMemcpyAsync(row_ptr, eng->row_ptr.p, 4 * (nc + 1), hipMemcpyDeviceToHost, s));
    if (cj && ne) DYN_HIP(hipMemcpyAsync(cj, eng->cj.p, 4 * ne, hipMemcpyDeviceToHost, s));
    if (cmask && ne) DYN_HIP(hipMemcpyAsync(cmask, eng->cmask.p, 8 * ne, hipMemcpyDeviceToHost, s));
    if (fx_slot) DYN_HIP(hipMemcpyAsync(fx_slot, eng->fx_slot.p, 24 * ns, hipMemcpyDeviceToHost, s));
    DYN_HIP(hipStreamSynchronize(s));
  });
}
