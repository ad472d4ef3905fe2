// replay.cpp — host replays of the GPU kernels' fixed-point sums
// (test infrastructure only).
//
// Compiles the product's arithmetic headers (dyn_math.h, pair_ref.h,
// engine.h) with g++ and contraction off, evaluates every pair exactly as the
// kernels do, and sums the int64 fixed-point contributions; since integer
// addition is order-independent, the GPU's atomically accumulated sums must
// match these bit for bit.  For the engine, the 8×8 wave tile and its xor-tree
// reductions are emulated lane by lane.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "dyn_math.h"
#include "engine.h"
#include "pair_ref.h"

extern "C" {

// ---- arithmetic probes ----------------------------------------------------
float orc_rsqrtf(float x) { return dyn_rsqrtf(x); }
float orc_exp_neg(float y) { return dyn_exp_neg(y); }
float orc_erfcx(float x) { return dyn_erfcx(x); }
float orc_erfcf(float x) { return dyn_erfcf(x); }
float orc_inv_cbrt(float x) { return dyn_inv_cbrt(x); }
int64_t orc_fx(float f) { return dyn_fx(f); }
void orc_coul_erfc(float r2, float qq, float beta, float* s, float* e) {
  const float rinv = dyn_rsqrtf(r2);
  dyn_coul_erfc(r2, rinv, qq, beta, beta * beta, 2.0f * beta * DYN_INV_SQRT_PI, s, e);
}
// The pair tiles' tabulated erfc Coulomb (dyn_coul_erfc_tab) and its table.
static const float4* coul_table() {
  static std::vector<float> tab;
  if (tab.empty()) {
    tab.resize(8 * DYN_COUL_TAB_N);
    dyn_coul_tab_build(tab.data());
  }
  return reinterpret_cast<const float4*>(tab.data());
}
void orc_coul_erfc_tab(float r2, float qq, float beta, float* s, float* e) {
  const float rinv = dyn_rsqrtf(r2);
  dyn_coul_erfc_tab(r2, rinv, qq, beta, coul_table(), s, e);
}
// The two tabulated functions at x: φ(x) and erfc(x).
void orc_coul_tab_eval(float x, float* phi, float* erfc_x) {
  *phi = dyn_tab_cubic(coul_table(), 0, x);
  *erfc_x = dyn_tab_cubic(coul_table(), 1, x);
}
void orc_coul_erf(float r2, float qq, float beta, float* s, float* e) {
  const float rinv = dyn_rsqrtf(r2);
  dyn_coul_erf(r2, rinv, qq, beta, beta * beta, 2.0f * beta * DYN_INV_SQRT_PI, s, e);
}
void orc_lj(float r2, float c6, float c12, float* s, float* e) {
  const float rinv = dyn_rsqrtf(r2);
  dyn_lj(rinv * rinv, c6, c12, s, e);
}
void orc_soft_core(float r2, float sigma, float eps, float lambda, float* s, float* e,
                   float* dhdl) {
  dyn_lj_soft_core(r2, sigma, eps, lambda, s, e, dhdl);
}
void orc_bspline4(float w, float* th, float* dth) { dyn_bspline<4>(w, th, dth); }
void orc_bspline5(float w, float* th, float* dth) { dyn_bspline<5>(w, th, dth); }
float orc_min_image_ref(float dv, float ext) { return dyn_min_image_ref(dv, ext); }
float orc_norm2_ref(float x, float y, float z) { return dyn_norm2_ref(x, y, z); }

// ---- reference pair-list kernel ----------------------------------------------
// pos: all sites, class-major (solute, O, M, H0, H1), 3 floats each.
void orc_ref_pairs(const float* pos, uint32_t n_dyn, uint32_t n_water, uint32_t n,
                   const uint32_t* tgt, const uint32_t* src, const float* sigma,
                   const float* eps, const float* q_tgt, const float* q_src,
                   const uint8_t* at_t, const uint8_t* ws_t, const uint8_t* at_s,
                   const uint8_t* ws_s, const uint8_t* scale14, const uint8_t* calc_lj,
                   const uint8_t* calc_coul, const uint8_t* symmetric, const uint8_t* alch,
                   const float* ext, float cutoff, float alpha, float lj_cutoff,
                   int coulomb_disabled, int lj_disabled, double lambda, int energy_all,
                   int64_t* fx, double* energy, double* virial, double* dhdl) {
  DynRefParams P;
  for (int k = 0; k < 3; ++k) {
    P.ext[k] = ext[k];
    P.inv_ext[k] = 1.0f / ext[k];
  }
  P.cutoff = cutoff;
  P.beta = alpha;
  P.beta2 = alpha * alpha;
  P.two_b_sqrtpi = 2.0f * alpha * DYN_INV_SQRT_PI;
  P.lj_cutoff = lj_cutoff;
  P.coulomb_disabled = coulomb_disabled;
  P.lj_disabled = lj_disabled;
  P.sch = dyn_schedule((float)lambda);
  const uint32_t ns = n_dyn + 4 * n_water;
  memset(fx, 0, sizeof(int64_t) * 3 * ns);
  double e = 0, w = 0, d = 0;
  for (uint32_t k = 0; k < n; ++k) {
    DynPairRec r{};
    r.tgt = at_t[k] == 0 ? tgt[k] : n_dyn + (ws_t[k] - 1u) * n_water + tgt[k];
    r.src = at_s[k] == 0 ? src[k] : n_dyn + (ws_s[k] - 1u) * n_water + src[k];
    r.sigma = sigma[k];
    r.eps = eps[k];
    r.q_tgt = q_tgt[k];
    r.q_src = q_src[k];
    uint32_t f = 0;
    if (scale14[k]) f |= DYN_PF_SCALE14;
    if (calc_lj[k]) f |= DYN_PF_LJ;
    if (calc_coul[k]) f |= DYN_PF_COUL;
    if (symmetric[k]) f |= DYN_PF_SYMMETRIC;
    if (alch && alch[k]) f |= DYN_PF_ALCH;
    if (at_t[k] == 0) f |= DYN_PF_TGT_SOLUTE;
    r.flags = f;
    float ff[3], ee, dh;
    double vir;
    if (!dyn_ref_pair(pos, r, P, ff, &ee, &dh, &vir)) continue;
    w += vir;
    if (energy_all || (f & DYN_PF_TGT_SOLUTE)) e += ee;
    if (f & DYN_PF_ALCH) d += dh;
    for (int c = 0; c < 3; ++c) {
      const int64_t v = dyn_fx(ff[c]);
      fx[3 * r.tgt + c] += v;
      if (f & DYN_PF_SYMMETRIC) fx[3 * r.src + c] -= v;
    }
  }
  *energy = e;
  *virial = w;
  *dhdl = d;
}

}  // extern "C"

// ---- native engine pair tiles -----------------------------------------------
// Rows of j entries in groups of 8, cut into work items of ≤ DYN_CHUNK
// groups; each lane (ia, jb) accumulates the force on i slot 8I+ia over the
// item's groups in list order, then the 8 j lanes of each i slot are summed
// by the xor-8/16/32 tree and converted to fixed point once per item.  A j
// entry at or beyond n_slots is the null slot (padding).
template <int KIND>
static void tile(const float* xyzq, const int32_t* slot_type, const int32_t* slot_gid,
                 uint32_t n_slots, int type_none, uint32_t I, const uint32_t* jent, uint64_t mask,
                 const DynPairParams& P, const float2* lj, const float2* lj_se, float fi[64][3],
                 double* e) {
  for (int l = 0; l < 64; ++l) {
    const int ia = l & 7, jb = l >> 3;
    const uint32_t js = jent[jb] & DYN_JSLOT_MASK;
    const bool null_j = js >= n_slots;
    const float* a = xyzq + 4 * (size_t)(I * DYN_CL_SIZE + ia);
    const float pi[4] = {a[0], a[1], a[2], a[3]};
    float pj[4] = {0.f, 0.f, 0.f, 0.f};
    if (!null_j)
      for (int k = 0; k < 4; ++k) pj[k] = xyzq[4 * (size_t)js + k];
    const int ti = slot_type[I * DYN_CL_SIZE + ia];
    const int tj = null_j ? type_none : slot_type[js];
    const int gj = null_j ? DYN_GID_PAD : slot_gid[js];
    bool active = dyn_slot_pair_valid(slot_gid[I * DYN_CL_SIZE + ia], gj);
    if (KIND == DYN_KIND_SS) active = active && ((mask >> l) & 1ull);
    float d[3], s, elj, ec, r2, dh = 0.f;
    int hit;
    // the alchemical variant equals the plain one on every lane that does
    // not cross the decoupled molecule's boundary
    if (P.alch_on && KIND <= DYN_KIND_SQ)
      dyn_tile_pair_alch<KIND>(pi, ti, pj, tj, active, P, lj, lj_se, d, &s, &elj, &ec, &r2, &hit, &dh);
    else
      dyn_tile_pair<KIND>(pi, ti, pj, tj, active, P, lj, d, &s, &elj, &ec, &r2, &hit);
    fi[l][0] = dyn_fmaf(s, d[0], fi[l][0]);
    fi[l][1] = dyn_fmaf(s, d[1], fi[l][1]);
    fi[l][2] = dyn_fmaf(s, d[2], fi[l][2]);
    if (e) {
      // e: e_lj water, e_lj solute, e_coul water, e_coul solute, virial, dH/dλ
      // (every pair twice: once from each of its two rows)
      const bool solute = KIND == DYN_KIND_SS || KIND == DYN_KIND_SO || KIND == DYN_KIND_SQ;
      if (KIND != DYN_KIND_SQ && KIND != DYN_KIND_QQ) e[solute ? 1 : 0] += elj;
      if (KIND != DYN_KIND_SO && KIND != DYN_KIND_OO) e[solute ? 3 : 2] += ec;
      e[4] += s * r2;
      if (P.alch_on && KIND <= DYN_KIND_SQ) e[5] += dh;
    }
  }
}

extern "C" {

// row_ptr and cmask per group, cj per entry (8 per group), items (I, first
// group).  n_types includes water O and "none" (the last type).
void orc_engine_tiles(const float* xyzq, const int32_t* slot_type, const int32_t* slot_gid,
                      const uint8_t* cl_class, const uint32_t* row_ptr, const uint32_t* cj,
                      const uint64_t* cmask, const uint32_t* items, uint64_t n_items,
                      uint32_t n_slots, const float* lj_table, const float* lj_se_table,
                      int n_types, const float* box, float cutoff, float lj_cutoff, float beta,
                      float c6_oo, float c12_oo, int alch_on, float lambda, int64_t* fx,
                      double* energies, double* item_energies) {
  DynPairParams P;
  for (int k = 0; k < 3; ++k) {
    P.ext[k] = box[k];
    P.inv_ext[k] = 1.0f / box[k];
  }
  P.rc2 = cutoff * cutoff;
  P.lj_rc2 = lj_cutoff * lj_cutoff;
  P.beta = beta;
  P.beta2 = beta * beta;
  P.two_b_sqrtpi = 2.0f * beta * DYN_INV_SQRT_PI;
  P.c6_oo = c6_oo;
  P.c12_oo = c12_oo;
  P.n_types = n_types;
  P.alch_on = alch_on;
  P.sch = dyn_schedule(lambda);
  static std::vector<float> coul_tab;
  if (coul_tab.empty()) {
    coul_tab.resize(8 * DYN_COUL_TAB_N);
    dyn_coul_tab_build(coul_tab.data());
  }
  P.coul_tab = reinterpret_cast<const float4*>(coul_tab.data());
  const float2* lj = reinterpret_cast<const float2*>(lj_table);
  const float2* lse = reinterpret_cast<const float2*>(lj_se_table);
  memset(fx, 0, sizeof(int64_t) * 3 * n_slots);
  if (energies)
    for (int k = 0; k < 6; ++k) energies[k] = 0.0;
  static float fi[64][3];
  for (uint64_t it = 0; it < n_items; ++it) {
    const uint32_t I = items[2 * it], g0 = items[2 * it + 1];
    // a work item ends where the next one of its row begins (the items of a
    // row are consecutive), else at the row's end: any work-item length
    const bool next_same = it + 1 < n_items && items[2 * (it + 1)] == I;
    const uint32_t g1 = next_same ? items[2 * (it + 1) + 1] : row_ptr[I + 1];
    memset(fi, 0, sizeof fi);
    const int ci = cl_class[I];
    double ie[6] = {0, 0, 0, 0, 0, 0};
    double* e_acc = (energies || item_energies) ? ie : nullptr;
    for (uint32_t g = g0; g < g1; ++g) {
      const uint32_t* je = cj + (size_t)g * DYN_CL_SIZE;
      const uint64_t mask = cmask[g];
      const int tn = n_types - 1;
      switch (dyn_kind(ci, (int)(je[0] >> DYN_JCLS_SHIFT))) {
        case DYN_KIND_QQ: tile<DYN_KIND_QQ>(xyzq, slot_type, slot_gid, n_slots, tn, I, je, mask, P, lj, lse, fi, e_acc); break;
        case DYN_KIND_OO: tile<DYN_KIND_OO>(xyzq, slot_type, slot_gid, n_slots, tn, I, je, mask, P, lj, lse, fi, e_acc); break;
        case DYN_KIND_SS: tile<DYN_KIND_SS>(xyzq, slot_type, slot_gid, n_slots, tn, I, je, mask, P, lj, lse, fi, e_acc); break;
        case DYN_KIND_SO: tile<DYN_KIND_SO>(xyzq, slot_type, slot_gid, n_slots, tn, I, je, mask, P, lj, lse, fi, e_acc); break;
        case DYN_KIND_SQ: tile<DYN_KIND_SQ>(xyzq, slot_type, slot_gid, n_slots, tn, I, je, mask, P, lj, lse, fi, e_acc); break;
        default: break;
      }
    }
    for (int k = 0; k < 6; ++k) {
      if (energies) energies[k] += ie[k];
      if (item_energies) item_energies[6 * it + k] = ie[k];
    }
    float v[64][3];
    memcpy(v, fi, sizeof v);
    for (int m : {8, 16, 32}) {
      float h[64][3];
      for (int l = 0; l < 64; ++l)
        for (int c = 0; c < 3; ++c) h[l][c] = v[l][c] + v[l ^ m][c];
      memcpy(v, h, sizeof v);
    }
    for (int ia = 0; ia < 8; ++ia) {
      const uint32_t slot = I * DYN_CL_SIZE + ia;
      for (int c = 0; c < 3; ++c) fx[(size_t)c * n_slots + slot] += dyn_fx(v[ia][c]);
    }
  }
}

// LJ table as the engine builds it: types + water O + "none" (zero).
void orc_engine_lj_table(const float* sig_eps, int n_types_in, float* table_out) {
  const int nt = n_types_in + 2;
  std::vector<float> sig(nt, 0.f), eps(nt, 0.f);
  for (int t = 0; t < n_types_in; ++t) {
    sig[t] = sig_eps[2 * t];
    eps[t] = sig_eps[2 * t + 1];
  }
  sig[n_types_in] = 1.777167268f * (2.0f / 1.122462048309373f);
  eps[n_types_in] = 0.2128008130f;
  for (int a = 0; a < nt; ++a)
    for (int b = 0; b < nt; ++b) {
      const float s_ab = 0.5f * (sig[a] + sig[b]);
      const float e_ab = std::sqrt(eps[a] * eps[b]);
      const double s6 = std::pow((double)s_ab, 6.0);
      table_out[2 * (a * nt + b)] = (float)(4.0 * e_ab * s6);
      table_out[2 * (a * nt + b) + 1] = (float)(4.0 * e_ab * s6 * s6);
    }
}

// ---- SPME charge spreading -----------------------------------------------------
void orc_pme_spread(const float* xyzq, uint32_t n, const int* K, const float* box, int order,
                    int64_t* grid) {
  float sc[3];
  for (int a = 0; a < 3; ++a) sc[a] = (float)((double)K[a] / (double)box[a]);
  memset(grid, 0, sizeof(int64_t) * (size_t)K[0] * K[1] * K[2]);
  for (uint32_t i = 0; i < n; ++i) {
    const float* p = xyzq + 4 * (size_t)i;
    if (p[3] == 0.0f) continue;
    int k0[3];
    float th[3][5], dth[5];
    for (int a = 0; a < 3; ++a) {
      const float u = p[a] * sc[a];
      const float fl = floorf(u);
      const float w = u - fl;
      int k = (int)fl % K[a];
      if (k < 0) k += K[a];
      k0[a] = k;
      if (order == 4) dyn_bspline<4>(w, th[a], dth);
      else dyn_bspline<5>(w, th[a], dth);
    }
    for (int jx = 0; jx < order; ++jx) {
      const int gx = (k0[0] + K[0] - order + 1 + jx) % K[0];
      const float qx = p[3] * th[0][jx];
      for (int jy = 0; jy < order; ++jy) {
        const int gy = (k0[1] + K[1] - order + 1 + jy) % K[1];
        const float qxy = qx * th[1][jy];
        for (int jz = 0; jz < order; ++jz) {
          const int gz = (k0[2] + K[2] - order + 1 + jz) % K[2];
          const float v = qxy * th[2][jz];
          grid[((size_t)gx * K[1] + gy) * K[2] + gz] += dyn_fx_d((double)v, DYN_GRID_FX_SCALE);
        }
      }
    }
  }
}

}  // extern "C"

// ---- engine cluster-pair list ---------------------------------------------
// Host restatement of the engine's list build (dynamics_amd/csrc/engine.hip:
// binning, column sort, clusters, bounding boxes, candidate walk, exclusion
// masks) with the same f32 operations, so the GPU list must equal it entry
// for entry.  Geometry follows the engine's rule: per class, columns of
// width ≈ the edge of a cube holding 8 sites of the class, at most 64 per
// axis.
namespace {

struct ListGeom {
  int ncx[3], ncy[3], col_off[3], n[3];
  float wx[3], wy[3];
  int n_cols;
};

float wrap_f(float x, float L, float invL) {
  float w = x - L * std::floor(x * invL);
  if (!(w < L)) w = 0.0f;
  if (w < 0.0f) w = 0.0f;
  return w;
}

}  // namespace

extern "C" {

// pos4: n_sites × (x, y, z, ·) in the engine's site order.  ex_ptr/ex_idx:
// per solute site, the sorted partners of its 1-2, 1-3 and 1-4 pairs.
// Outputs (caller-sized): slot_site[8·cap_cl], cl_class[cap_cl],
// row_ptr[cap_cl + 1] and cmask[cap_entries] per group of 8 j entries
// (cmask written for solute–solute groups, all ones elsewhere),
// cj[8·cap_entries] (slot | class << 30; padding names slot 2^30 − 1).
// Returns 0, or −1 when a capacity is exceeded.
int orc_engine_list(uint32_t n_dyn, uint32_t n_water, const float* pos4, const float* box,
                    float cutoff, float list_buffer, const uint32_t* ex_ptr,
                    const uint32_t* ex_idx, uint32_t cap_cl, uint64_t cap_entries,
                    uint32_t* n_cl_out, int32_t* slot_site, uint8_t* cl_class, uint32_t* row_ptr,
                    uint32_t* cj, uint64_t* cmask, uint64_t* n_entries_out) {
  const uint32_t ns = n_dyn + 4 * n_water;
  const float L[3] = {box[0], box[1], box[2]};
  const float invL[3] = {1.0f / box[0], 1.0f / box[1], 1.0f / box[2]};
  ListGeom g{};
  const double V = (double)box[0] * box[1] * box[2];
  const uint32_t ncls[3] = {n_dyn, n_water, 3 * n_water};
  int off = 0;
  for (int c = 0; c < 3; ++c) {
    g.n[c] = (int)ncls[c];
    g.col_off[c] = off;
    if (ncls[c] == 0) {
      g.ncx[c] = g.ncy[c] = 0;
      g.wx[c] = g.wy[c] = 1.f;
      continue;
    }
    const double s = std::cbrt(8.0 * V / (double)ncls[c]);
    g.ncx[c] = std::max(1, std::min(64, (int)std::floor(box[0] / s)));
    g.ncy[c] = std::max(1, std::min(64, (int)std::floor(box[1] / s)));
    g.wx[c] = box[0] / g.ncx[c];
    g.wy[c] = box[1] / g.ncy[c];
    off += g.ncx[c] * g.ncy[c];
  }
  g.n_cols = off;
  // wrap and bin
  std::vector<float> w(4 * (size_t)ns);
  std::vector<std::vector<uint32_t>> col(g.n_cols);
  std::vector<int> col_cls(g.n_cols);
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < g.ncx[c] * g.ncy[c]; ++k) col_cls[g.col_off[c] + k] = c;
  for (uint32_t s = 0; s < ns; ++s) {
    for (int a = 0; a < 3; ++a) w[4 * s + a] = wrap_f(pos4[4 * s + a], L[a], invL[a]);
    const int c = s < n_dyn ? 0 : (((s - n_dyn) & 3u) == 0 ? 1 : 2);
    int cx = (int)(w[4 * s] / g.wx[c]), cy = (int)(w[4 * s + 1] / g.wy[c]);
    cx = std::min(std::max(cx, 0), g.ncx[c] - 1);
    cy = std::min(std::max(cy, 0), g.ncy[c] - 1);
    col[g.col_off[c] + cx * g.ncy[c] + cy].push_back(s);
  }
  // sort columns by (z, site), cut into clusters
  std::vector<uint32_t> col_cl(g.n_cols + 1, 0);
  std::vector<float> bbc, bbh;
  std::vector<int32_t> site;
  std::vector<uint8_t> cls;
  for (int k = 0; k < g.n_cols; ++k) {
    auto& v = col[k];
    std::vector<uint64_t> key(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
      uint32_t zb;
      memcpy(&zb, &w[4 * v[i] + 2], 4);
      key[i] = ((uint64_t)zb << 32) | v[i];
    }
    std::sort(key.begin(), key.end());
    const size_t ncl = (v.size() + DYN_CL_SIZE - 1) / DYN_CL_SIZE;
    col_cl[k + 1] = col_cl[k] + (uint32_t)ncl;
    for (size_t q = 0; q < ncl; ++q) {
      float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
      for (int m = 0; m < DYN_CL_SIZE; ++m) {
        const size_t r = q * DYN_CL_SIZE + m;
        if (r < v.size()) {
          const uint32_t s = (uint32_t)(key[r] & 0xffffffffu);
          site.push_back((int32_t)s);
          for (int a = 0; a < 3; ++a) {
            lo[a] = std::fmin(lo[a], w[4 * s + a]);
            hi[a] = std::fmax(hi[a], w[4 * s + a]);
          }
        } else {
          site.push_back(-1);
        }
      }
      for (int a = 0; a < 3; ++a) {
        bbc.push_back(0.5f * (lo[a] + hi[a]));
        bbh.push_back(0.5f * (hi[a] - lo[a]));
      }
      cls.push_back((uint8_t)col_cls[k]);
    }
  }
  const uint32_t n_cl = col_cl[g.n_cols];
  if (n_cl > cap_cl) return -1;
  *n_cl_out = n_cl;
  memcpy(slot_site, site.data(), sizeof(int32_t) * site.size());
  memcpy(cl_class, cls.data(), cls.size());
  std::vector<uint32_t> slot_of(ns, 0);
  for (size_t k = 0; k < site.size(); ++k)
    if (site[k] >= 0) slot_of[site[k]] = (uint32_t)k;
  // exclusion groups and slot positions as the engine stores them
  std::vector<int32_t> gid(site.size());
  for (size_t k = 0; k < site.size(); ++k) {
    const int32_t st = site[k];
    gid[k] = st < 0 ? DYN_GID_PAD
                    : ((uint32_t)st < n_dyn ? -2 - st : (int32_t)(((uint32_t)st - n_dyn) / 4));
  }
  auto slot_pos = [&](size_t k, float p[3]) {
    const int32_t st = site[k] >= 0 ? site[k] : site[k & ~(size_t)(DYN_CL_SIZE - 1)];
    for (int a = 0; a < 3; ++a) p[a] = w[4 * (size_t)st + a];
  };
  // row walk: candidate clusters in the engine's order, then the j slots
  // within rlist of an i slot they interact with, class segments padded to 8
  const float rlist = cutoff + std::max(0.f, list_buffer);
  const float rlist2 = rlist * rlist;
  const float col_r2 = rlist2 * 1.0001f + 1e-3f;
  const uint32_t null_entry = DYN_JSLOT_MASK;
  uint64_t ne = 0;  // groups
  std::vector<uint32_t> row;
  row_ptr[0] = 0;
  for (uint32_t I = 0; I < n_cl; ++I) {
    const float c[3] = {bbc[3 * I], bbc[3 * I + 1], bbc[3 * I + 2]};
    const float h[3] = {bbh[3 * I], bbh[3 * I + 1], bbh[3 * I + 2]};
    const int ci = cls[I];
    float pi[8][3];
    for (int ia = 0; ia < 8; ++ia) slot_pos(I * DYN_CL_SIZE + ia, pi[ia]);
    row.clear();
    for (int k = 0; k < 3; ++k) {
      if (dyn_kind(ci, k) == DYN_KIND_NONE || g.n[k] == 0) continue;
      const uint32_t tag = (uint32_t)k << DYN_JCLS_SHIFT;
      int x0 = (int)std::floor((c[0] - h[0] - rlist) / g.wx[k]);
      int x1 = (int)std::floor((c[0] + h[0] + rlist) / g.wx[k]);
      int y0 = (int)std::floor((c[1] - h[1] - rlist) / g.wy[k]);
      int y1 = (int)std::floor((c[1] + h[1] + rlist) / g.wy[k]);
      bool full_x = false, full_y = false;
      if (x1 - x0 + 1 >= g.ncx[k]) {
        x0 = 0;
        x1 = g.ncx[k] - 1;
        full_x = true;
      }
      if (y1 - y0 + 1 >= g.ncy[k]) {
        y0 = 0;
        y1 = g.ncy[k] - 1;
        full_y = true;
      }
      for (int xx = x0; xx <= x1; ++xx)
        for (int yy = y0; yy <= y1; ++yy) {
          float dx = 0.f, dy = 0.f;
          if (!full_x)
            dx = std::fmax(0.f, std::fmax((float)xx * g.wx[k] - (c[0] + h[0]),
                                          (c[0] - h[0]) - (float)(xx + 1) * g.wx[k]));
          if (!full_y)
            dy = std::fmax(0.f, std::fmax((float)yy * g.wy[k] - (c[1] + h[1]),
                                          (c[1] - h[1]) - (float)(yy + 1) * g.wy[k]));
          if (!(dx * dx + dy * dy < col_r2)) continue;
          const int cx = ((xx % g.ncx[k]) + g.ncx[k]) % g.ncx[k];
          const int cy = ((yy % g.ncy[k]) + g.ncy[k]) % g.ncy[k];
          const int cc = g.col_off[k] + cx * g.ncy[k] + cy;
          // the column's periodic image as the walk reached it
          const float sx = (float)((xx - cx) / g.ncx[k]) * L[0];
          const float sy = (float)((yy - cy) / g.ncy[k]) * L[1];
          for (uint32_t J = col_cl[cc]; J < col_cl[cc + 1]; ++J) {
            float d[3], kz = 0.f;
            for (int a = 0; a < 3; ++a) {
              float v = c[a] - bbc[3 * J + a];
              const float k = std::rint(v * invL[a]);
              if (a == 2) kz = k;
              v = std::fabs(v - L[a] * k);
              d[a] = std::fmax(0.f, v - h[a] - bbh[3 * J + a]);
            }
            if (!(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < rlist2)) continue;
            // z: through the image of J's box nearest I's when no slot pair
            // can be within rlist in another image (rlist and both
            // half-heights short of half the box), else per slot pair
            const bool zc = rlist + h[2] + bbh[3 * J + 2] < 0.45f * L[2];
            const float sz = kz * L[2];
            for (int jb = 0; jb < 8; ++jb) {
              const size_t js = (size_t)J * DYN_CL_SIZE + jb;
              float pj[3];
              slot_pos(js, pj);
              // within rlist of a non-padding i slot (x and y through the
              // column's image unless the column range spans the box)
              bool keep = false;
              const float px = pj[0] + sx, py = pj[1] + sy;
              for (int ia = 0; ia < 8 && !keep; ++ia) {
                if (gid[I * DYN_CL_SIZE + ia] == DYN_GID_PAD) continue;
                float d0, d1;
                if (!full_x && !full_y) {
                  d0 = pi[ia][0] - px;
                  d1 = pi[ia][1] - py;
                } else {
                  d0 = dyn_min_image_rint(pi[ia][0] - pj[0], L[0], invL[0]);
                  d1 = dyn_min_image_rint(pi[ia][1] - pj[1], L[1], invL[1]);
                }
                const float d2 = zc ? pi[ia][2] - (pj[2] + sz)
                                    : dyn_min_image_rint(pi[ia][2] - pj[2], L[2], invL[2]);
                const float r2 = dyn_fmaf(d2, d2, dyn_fmaf(d1, d1, d0 * d0));
                keep = r2 < rlist2;
              }
              if (keep && gid[js] != DYN_GID_PAD) row.push_back((uint32_t)js | tag);
            }
          }
        }
      while (row.size() % DYN_CL_SIZE) row.push_back(null_entry | tag);
    }
    const uint64_t ng = row.size() / DYN_CL_SIZE;
    if (ne + ng > cap_entries) return -1;
    for (uint64_t gq = 0; gq < ng; ++gq) {
      const uint32_t* je = row.data() + gq * DYN_CL_SIZE;
      uint64_t mk = ~0ull;
      if (ci == DYN_CLS_SOLUTE && (je[0] >> DYN_JCLS_SHIFT) == DYN_CLS_SOLUTE)
        for (int ia = 0; ia < DYN_CL_SIZE; ++ia) {
          const int32_t sa = site[I * DYN_CL_SIZE + ia];
          if (sa < 0) continue;
          for (uint32_t e = ex_ptr[sa]; e < ex_ptr[sa + 1]; ++e) {
            const uint32_t sl = slot_of[ex_idx[e]];
            for (int jb = 0; jb < DYN_CL_SIZE; ++jb)
              if ((je[jb] & DYN_JSLOT_MASK) == sl) mk &= ~(1ull << (ia + DYN_CL_SIZE * jb));
          }
        }
      cmask[ne + gq] = mk;
      memcpy(cj + (ne + gq) * DYN_CL_SIZE, je, sizeof(uint32_t) * DYN_CL_SIZE);
    }
    ne += ng;
    row_ptr[I + 1] = (uint32_t)ne;
  }
  *n_entries_out = ne;
  return 0;
}

}  // extern "C"
