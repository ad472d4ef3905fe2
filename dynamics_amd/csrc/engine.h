// engine.h — device-side types of the native nonbonded engine, shared by the
// kernels and the host fixed-point replay.
#pragma once

#include "dyn_math.h"

// Site classes.  Clusters hold sites of one class only, so a cluster pair
// knows from its two classes which interactions exist: solute–solute LJ +
// Coulomb, solute–O LJ, solute–charged Coulomb, O–O LJ, charged–charged
// Coulomb; O–charged pairs have neither and are never listed.
#define DYN_CLS_SOLUTE 0
#define DYN_CLS_WATER_O 1
#define DYN_CLS_WATER_Q 2

#define DYN_CL_SIZE 8  // sites per cluster; a wave holds 8 i slots × 8 j slots
#define DYN_CHUNK 16   // default j groups (of 8 j slots) per work item of the pair kernel

// A row of the pair list (one i cluster) holds j slots: every slot of a
// nearby cluster that lies within the list radius of at least one i slot it
// interacts with.  The slots are grouped by class, each class segment padded
// to a multiple of 8 with the null slot, and a list entry stores
// slot | class << DYN_JCLS_SHIFT.  A group is 8 consecutive entries.
#define DYN_JCLS_SHIFT 30
#define DYN_JSLOT_MASK ((1u << DYN_JCLS_SHIFT) - 1u)

// Exclusion group of a cluster slot: the water molecule (≥ 0) for water
// sites, −2 − site for solute atoms (unique), −1 for padding.  Two slots
// interact only if their groups differ and neither is padding; solute–solute
// cluster pairs additionally carry an explicit 64-bit mask for the 1-2, 1-3
// and 1-4 exclusions.
#define DYN_GID_PAD (-1)

DYN_HD bool dyn_slot_pair_valid(int gi, int gj) {
  return gi != gj && gi != DYN_GID_PAD && gj != DYN_GID_PAD;
}

// Interaction kinds of a cluster pair (from the two classes).
#define DYN_KIND_SS 0  // LJ (type table) + Coulomb
#define DYN_KIND_SO 1  // LJ (type table)
#define DYN_KIND_SQ 2  // Coulomb
#define DYN_KIND_OO 3  // LJ (O–O constants)
#define DYN_KIND_QQ 4  // Coulomb
#define DYN_KIND_NONE 5

DYN_HD int dyn_kind(int ci, int cj) {
  if (ci > cj) {
    const int t = ci;
    ci = cj;
    cj = t;
  }
  if (ci == DYN_CLS_SOLUTE) return cj == DYN_CLS_SOLUTE ? DYN_KIND_SS : (cj == DYN_CLS_WATER_O ? DYN_KIND_SO : DYN_KIND_SQ);
  if (ci == DYN_CLS_WATER_O) return cj == DYN_CLS_WATER_O ? DYN_KIND_OO : DYN_KIND_NONE;
  return DYN_KIND_QQ;
}

// Type word of a slot: the LJ type in the low bits, DYN_TYPE_ALCH when the
// site belongs to the decoupled (alchemical) molecule.
#define DYN_TYPE_ALCH 0x10000
#define DYN_TYPE_MASK 0xffff

struct DynPairParams {
  float ext[3], inv_ext[3];
  float rc2;        // Coulomb (and list) cutoff squared
  float lj_rc2;     // LJ cutoff squared
  float beta, beta2, two_b_sqrtpi;
  float c6_oo, c12_oo;
  int n_types;      // LJ type tables are (n_types × n_types) float2
  int alch_on;      // a decoupled molecule exists
  DynSchedule sch;  // staged decoupling schedule at the current λ
  const float4* coul_tab;  // the erfc Coulomb table (dyn_coul_tab_build)
};

// One lane of an 8×8 tile: the pair (i, j) of the given kind.  Returns the
// scalar s (F_i = s·d) and energies; s = e = 0 for an inactive lane.
// `active` already includes the list mask; the cutoff test happens here.
template <int KIND>
DYN_HD void dyn_tile_pair(const float pi[4], int ti, const float pj[4], int tj, bool active,
                          const DynPairParams& P, const float2* lj_table, float d[3],
                          float* s_out, float* e_lj, float* e_c, float* r2_out, int* hit) {
  float d0 = pi[0] - pj[0], d1 = pi[1] - pj[1], d2 = pi[2] - pj[2];
  d0 = dyn_min_image_rint(d0, P.ext[0], P.inv_ext[0]);
  d1 = dyn_min_image_rint(d1, P.ext[1], P.inv_ext[1]);
  d2 = dyn_min_image_rint(d2, P.ext[2], P.inv_ext[2]);
  d[0] = d0;
  d[1] = d1;
  d[2] = d2;
  const float r2 = dyn_fmaf(d2, d2, dyn_fmaf(d1, d1, d0 * d0));
  const bool in = active && r2 < P.rc2 && r2 > 1e-16f;
  const float r2s = in ? r2 : 1.0f;
  const float rinv = dyn_rsqrtf(r2s);
  float s = 0.f, elj = 0.f, ec = 0.f;
  if (KIND == DYN_KIND_SS || KIND == DYN_KIND_SO || KIND == DYN_KIND_OO) {
    float c6, c12;
    if (KIND == DYN_KIND_OO) {
      c6 = P.c6_oo;
      c12 = P.c12_oo;
    } else {
      const float2 c = lj_table[(ti & DYN_TYPE_MASK) * P.n_types + (tj & DYN_TYPE_MASK)];
      c6 = c.x;
      c12 = c.y;
    }
    float sl, el;
    dyn_lj(rinv * rinv, c6, c12, &sl, &el);
    const bool lj_in = in && r2 <= P.lj_rc2;
    s = lj_in ? sl : 0.f;
    elj = lj_in ? el : 0.f;
  }
  if (KIND == DYN_KIND_SS || KIND == DYN_KIND_SQ || KIND == DYN_KIND_QQ) {
    float sc, ee;
    dyn_coul_erfc_tab(r2s, rinv, pi[3] * pj[3], P.beta, P.coul_tab, &sc, &ee);
    s = KIND == DYN_KIND_SS ? s + (in ? sc : 0.f) : (in ? sc : 0.f);
    ec = in ? ee : 0.f;
  }
  *s_out = s;
  *e_lj = elj;
  *e_c = ec;
  *r2_out = in ? r2 : 0.f;
  *hit = in ? 1 : 0;
}

// FP64 variant for the precision sweep: the same interactions from the same
// f32 inputs (positions, charges, combined σ and ε), every operation in
// double with the math library's erfc, exp and sqrt; c6/c12 from a double
// table.  Not bit-replayed on the host: it is
// checked against the FP64 oracle instead.
template <int KIND>
DYN_HD void dyn_tile_pair_f64(const float pi[4], int ti, const float pj[4], int tj, bool active,
                              const DynPairParams& P, const double2* lj64, double d[3],
                              double* s_out, double* e_lj, double* e_c, double* r2_out, int* hit) {
  double r2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double L = (double)P.ext[a];
    double v = (double)pi[a] - (double)pj[a];
    v -= L * rint(v / L);
    d[a] = v;
    r2 += v * v;
  }
  const bool in = active && r2 < (double)P.rc2 && r2 > 1e-16;
  const double r2s = in ? r2 : 1.0;
  const double rinv = 1.0 / sqrt(r2s);
  double s = 0.0, elj = 0.0, ec = 0.0;
  if (KIND == DYN_KIND_SS || KIND == DYN_KIND_SO || KIND == DYN_KIND_OO) {
    // the O–O pair too comes from the table (its f32 constants are rounded)
    const double2 c = lj64[(ti & DYN_TYPE_MASK) * P.n_types + (tj & DYN_TYPE_MASK)];
    const double c6 = c.x, c12 = c.y;
    const double rinv2 = rinv * rinv, rinv6 = rinv2 * rinv2 * rinv2;
    const bool lj_in = in && r2 <= (double)P.lj_rc2;
    s = lj_in ? (12.0 * c12 * rinv6 - 6.0 * c6) * rinv6 * rinv2 : 0.0;
    elj = lj_in ? (c12 * rinv6 - c6) * rinv6 : 0.0;
  }
  if (KIND == DYN_KIND_SS || KIND == DYN_KIND_SQ || KIND == DYN_KIND_QQ) {
    const double qq = (double)pi[3] * (double)pj[3];
    const double beta = (double)P.beta, r = r2s * rinv;
    const double g = erfc(beta * r) * rinv;
    const double sc = qq * (g + 1.1283791670955126 * beta * exp(-beta * beta * r2s)) * rinv * rinv;
    s += in ? sc : 0.0;
    ec = in ? qq * g : 0.0;
  }
  *s_out = s;
  *e_lj = elj;
  *e_c = ec;
  *r2_out = in ? r2 : 0.0;
  *hit = in ? 1 : 0;
}

// Alchemical variant (src/cuda/cuda.cu:367-566): a lane whose two sites lie
// on different sides of the decoupled molecule (exactly one carries
// DYN_TYPE_ALCH) takes Beutler soft-core LJ at the schedule's LJ λ and
// Coulomb scaled by the schedule's factor, and returns dH/dλ; every other
// lane returns dyn_tile_pair's result bit for bit.  lj_se holds the
// combined (σ, ε) of each type pair.
template <int KIND>
DYN_HD void dyn_tile_pair_alch(const float pi[4], int ti, const float pj[4], int tj, bool active,
                               const DynPairParams& P, const float2* lj_table,
                               const float2* lj_se, float d[3], float* s_out, float* e_lj,
                               float* e_c, float* r2_out, int* hit, float* dh) {
  dyn_tile_pair<KIND>(pi, ti, pj, tj, active, P, lj_table, d, s_out, e_lj, e_c, r2_out, hit);
  *dh = 0.f;
  if (((ti ^ tj) & DYN_TYPE_ALCH) == 0 || !*hit) return;
  const float r2 = *r2_out;
  const float rinv = dyn_rsqrtf(r2);
  float s = 0.f, elj = 0.f, ec = 0.f, dhh = 0.f;
  if (KIND == DYN_KIND_SS || KIND == DYN_KIND_SO) {
    if (r2 <= P.lj_rc2) {
      const float2 se = lj_se[(ti & DYN_TYPE_MASK) * P.n_types + (tj & DYN_TYPE_MASK)];
      float ssc, esc, dd;
      dyn_lj_soft_core(r2, se.x, se.y, P.sch.lj_lambda, &ssc, &esc, &dd);
      s = ssc;
      elj = esc;
      dhh = dd * P.sch.lj_dlambda_dlambda;
    }
  }
  if (KIND == DYN_KIND_SS || KIND == DYN_KIND_SQ) {
    float sc, ee;
    dyn_coul_erfc(r2, rinv, pi[3] * pj[3], P.beta, P.beta2, P.two_b_sqrtpi, &sc, &ee);
    s = s + sc * P.sch.coulomb_scale;
    ec = ee * P.sch.coulomb_scale;
    dhh = dhh + ee * P.sch.coulomb_dscale_dlambda;
  }
  *s_out = s;
  *e_lj = elj;
  *e_c = ec;
  *dh = dhh;
}
