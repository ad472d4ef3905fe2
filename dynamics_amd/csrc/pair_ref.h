// pair_ref.h — one pair of the reference's flattened pair list, evaluated as
// the reference's CUDA kernel does (src/cuda/cuda.cu:185-362 and the
// alchemical variant :367-566), in the IEEE-only arithmetic of dyn_math.h.
// Shared by the gfx950 kernel and the host fixed-point replay.
#pragma once

#include "dyn_math.h"

// Packed pair record (32 bytes, two 16-byte loads).  tgt/src are global site
// indices: solute atoms first, then water O, M, H0, H1 blocks (class-major),
// which folds the reference's (atom_type, water_site, index) triple into one
// index (src/cuda/util.cu:69-84).
struct DynPairRec {
  uint32_t tgt, src;
  float sigma, eps;
  float q_tgt, q_src;
  uint32_t flags;
  uint32_t pad;
};

#define DYN_PF_SCALE14 1u
#define DYN_PF_LJ 2u
#define DYN_PF_COUL 4u
#define DYN_PF_SYMMETRIC 8u
#define DYN_PF_ALCH 16u
#define DYN_PF_TGT_SOLUTE 32u

struct DynRefParams {
  float ext[3], inv_ext[3];
  float cutoff, beta, beta2, two_b_sqrtpi;
  float lj_cutoff;  // +inf: no LJ cutoff (the CUDA kernel's behaviour)
  int coulomb_disabled, lj_disabled;
  DynSchedule sch;
};

// Returns false when the pair is skipped (r² < 1e-16, as the kernel does).
// f: force on tgt (src receives −f when symmetric).  e: pair energy,
// dhdl: dH/dλ of the pair, vir: d·f accumulated by the caller in f64.
DYN_HD bool dyn_ref_pair(const float* pos, const DynPairRec& p, const DynRefParams& P,
                         float f[3], float* e, float* dhdl, double* vir) {
  const float* a = pos + 3u * p.tgt;
  const float* b = pos + 3u * p.src;
  float d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  d0 = dyn_min_image_rint(d0, P.ext[0], P.inv_ext[0]);
  d1 = dyn_min_image_rint(d1, P.ext[1], P.inv_ext[1]);
  d2 = dyn_min_image_rint(d2, P.ext[2], P.inv_ext[2]);
  const float r2 = dyn_fmaf(d2, d2, dyn_fmaf(d1, d1, d0 * d0));
  if (r2 < 1e-16f) return false;
  const float inv_r = dyn_rsqrtf(r2);
  const float r = r2 * inv_r;
  const float dir0 = d0 * inv_r, dir1 = d1 * inv_r, dir2 = d2 * inv_r;
  const bool alch = (p.flags & DYN_PF_ALCH) != 0;

  float mag_lj = 0.f, e_lj = 0.f, dh = 0.f;
  if ((p.flags & DYN_PF_LJ) && !P.lj_disabled && r <= P.lj_cutoff) {
    if (alch) {
      float s, ee, dd;
      dyn_lj_soft_core(r2, p.sigma, p.eps, P.sch.lj_lambda, &s, &ee, &dd);
      mag_lj = s * r;  // |F| along dir
      e_lj = ee;
      dh += dd * P.sch.lj_dlambda_dlambda;
    } else {
      const float sr = p.sigma * inv_r;
      const float sr2 = sr * sr;
      const float sr4 = sr2 * sr2;
      const float sr6 = sr4 * sr2;
      const float sr12 = sr6 * sr6;
      mag_lj = 24.0f * p.eps * dyn_fmaf(2.f, sr12, -sr6) * inv_r;
      e_lj = 4.f * p.eps * (sr12 - sr6);
    }
  }
  float mag_c = 0.f, e_c = 0.f;
  if ((p.flags & DYN_PF_COUL) && !P.coulomb_disabled && r < P.cutoff) {
    float s, ee;
    dyn_coul_erfc(r2, inv_r, p.q_tgt * p.q_src, P.beta, P.beta2, P.two_b_sqrtpi, &s, &ee);
    mag_c = s * r;
    e_c = ee;
  }
  if (p.flags & DYN_PF_SCALE14) {
    mag_lj *= 0.5f;
    e_lj *= 0.5f;
    dh *= 0.5f;
    mag_c *= 0.833333333f;
    e_c *= 0.833333333f;
  }
  float mag;
  if (alch) {
    mag = mag_lj + mag_c * P.sch.coulomb_scale;
    *e = e_lj + e_c * P.sch.coulomb_scale;
    dh += e_c * P.sch.coulomb_dscale_dlambda;
  } else {
    mag = mag_lj + mag_c;
    *e = e_lj + e_c;
  }
  f[0] = dir0 * mag;
  f[1] = dir1 * mag;
  f[2] = dir2 * mag;
  *dhdl = dh;
  *vir = (double)d0 * (double)f[0] + (double)d1 * (double)f[1] + (double)d2 * (double)f[2];
  return true;
}
