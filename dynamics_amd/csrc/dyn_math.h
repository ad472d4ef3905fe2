// dyn_math.h — pair arithmetic shared by the gfx950 kernels and the host-side
// fixed-point replay.
//
// Every function here uses only operations whose results IEEE-754 fixes
// exactly (add, sub, mul, fma, rint, ldexp, integer bit moves, exact
// conversions), so a host build of this header produces bit-identical
// per-pair forces and therefore bit-identical fixed-point sums.  Nothing here
// calls a vendor transcendental: rsqrt is a bit-level seed plus three Newton
// steps, exp is Cody–Waite reduction plus a degree-6 polynomial, and erfc is
// exp(-x²)·erfcx(x) with a degree-16 Chebyshev-derived polynomial for erfcx on
// [0, 4] (coefficients fitted here with numpy against scipy.special.erfcx; the
// fit's max relative error evaluated in f32 is 1.4e-7).
//
// Physics follows the reference:
//   LJ 12-6:           src/forces.rs:10-21, src/cuda/cuda.cu:40-63
//   erfc Coulomb:      src/cuda/cuda.cu:147-180 (F = qq (erfc(βr)/r² + 2β/√π e^{-β²r²}/r) r̂,
//                      E = qq erfc(βr)/r, zero at r ≥ cutoff)
//   soft-core LJ:      src/cuda/cuda.cu:73-144, src/non_bonded.rs:883-941
//   staged schedule:   src/alchemical.rs:82-100
//   min image (GPU):   src/cuda/util.cu:53-59 (rint, round half to even)
//   min image (CPU):   src/barostat.rs:162-170 (division then round half away)
//
// Compile with contraction off (-ffp-contract=off on both hipcc and g++):
// every fused multiply-add below is written out as dyn_fmaf().
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DYN_HD __host__ __device__ inline
#else
#include <math.h>
#include <string.h>
#define DYN_HD static inline
struct float2 {
  float x, y;
};
struct float4 {
  float x, y, z, w;
};
struct double2 {
  double x, y;
};
#endif

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

// ---------------------------------------------------------------------------
// Constants shared with the reference (values and meaning).
// ---------------------------------------------------------------------------
// src/non_bonded.rs:55-56 (1-4 scale factors, Amber convention).
#define DYN_SCALE_LJ_14 0.5f
#define DYN_SCALE_COUL_14 (1.0f / 1.2f)
// src/non_bonded.rs:63 (sqrt of 332.0522 kcal·Å/(mol·e²)).
#define DYN_CHARGE_UNIT_SCALER 18.2223f
// src/cuda/util.cu:18
#define DYN_INV_SQRT_PI 0.5641895835477563f
// src/alchemical.rs:55-65
#define DYN_SOFT_CORE_ALPHA 0.5f
#define DYN_SOFT_CORE_SIGMA_MIN 3.0f

// Fixed-point force scale: 2^32 units per kcal/(mol·Å).  Range ±2^31
// kcal/(mol·Å) in an int64 accumulator, resolution 2.3e-10.
#define DYN_FX_SCALE 4294967296.0
#define DYN_FX_INV_SCALE (1.0 / 4294967296.0)
// A single contribution is clamped to ±2^18 before conversion so the
// magic-number conversion below stays exact (needs |f·2^32| < 2^51).
#define DYN_FX_CLAMP 262144.0f
// Fixed-point charge-grid scale for SPME spreading: 2^32 units per charge
// unit (charges are in sqrt(kcal·Å/mol)).
#define DYN_GRID_FX_SCALE 4294967296.0

// ---------------------------------------------------------------------------
// Primitive operations
// ---------------------------------------------------------------------------
DYN_HD float dyn_fmaf(float a, float b, float c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fmaf(a, b, c);
#else
  return fmaf(a, b, c);
#endif
}

DYN_HD double dyn_fma(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fma(a, b, c);
#else
  return fma(a, b, c);
#endif
}

DYN_HD float dyn_floorf(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_floorf(x);
#else
  return floorf(x);
#endif
}

DYN_HD float dyn_rintf(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_rintf(x);
#else
  return rintf(x);
#endif
}

// Round half away from zero (Rust f32::round).
DYN_HD float dyn_roundf(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_roundf(x);
#else
  return roundf(x);
#endif
}

DYN_HD float dyn_ldexpf(float x, int e) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_ldexpf(x, e);
#else
  return ldexpf(x, e);
#endif
}

DYN_HD uint32_t dyn_f2u(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(f);
#else
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
#endif
}

DYN_HD float dyn_u2f(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __uint_as_float(u);
#else
  float f;
  memcpy(&f, &u, 4);
  return f;
#endif
}

DYN_HD int64_t dyn_d2bits(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (int64_t)__double_as_longlong(d);
#else
  int64_t u;
  memcpy(&u, &d, 8);
  return u;
#endif
}

DYN_HD float dyn_minf(float a, float b) { return a < b ? a : b; }
DYN_HD float dyn_maxf(float a, float b) { return a > b ? a : b; }

// 1/sqrt(x) for normal positive x: bit-level seed, three Newton steps.
// Max relative error 1.2e-7 over [1e-12, 1e12] (checked in tests).
DYN_HD float dyn_rsqrtf(float x) {
  float y = dyn_u2f(0x5f375a86u - (dyn_f2u(x) >> 1));
  const float hx = 0.5f * x;
  y = y * dyn_fmaf(-hx, y * y, 1.5f);
  y = y * dyn_fmaf(-hx, y * y, 1.5f);
  y = y * dyn_fmaf(-hx, y * y, 1.5f);
  return y;
}

// exp(y) for y in [-80, 0] (callers clamp).  Cody–Waite: y = n ln2 + r,
// |r| <= ln2/2, exp(r) by a degree-6 polynomial (max rel error 7.4e-8),
// scale by 2^n with an exact ldexp.
DYN_HD float dyn_exp_neg(float y) {
  y = dyn_maxf(y, -80.0f);
  const float n = dyn_rintf(y * 1.44269502f);
  float r = dyn_fmaf(n, -0.693145752f, y);  // ln2 high part (exact product)
  r = dyn_fmaf(n, -1.42860677e-06f, r);     // ln2 low part
  float p = 0x1.6d940cp-10f;
  p = dyn_fmaf(p, r, 0x1.125b1cp-7f);
  p = dyn_fmaf(p, r, 0x1.55547ep-5f);
  p = dyn_fmaf(p, r, 0x1.555430p-3f);
  p = dyn_fmaf(p, r, 0x1.000000p-1f);
  p = dyn_fmaf(p, r, 1.0f);
  p = dyn_fmaf(p, r, 1.0f);
  return dyn_ldexpf(p, (int)n);
}

// erfcx(x) = exp(x²) erfc(x) on [0, 4]; x is clamped into that range (the
// pair kernels only evaluate x = β r with r < cutoff).
DYN_HD float dyn_erfcx(float x) {
  x = dyn_minf(dyn_maxf(x, 0.0f), 4.0f);
  const float t = dyn_fmaf(x, 0.5f, -1.0f);
  float p = 0x1.1d0172p-12f;
  p = dyn_fmaf(p, t, -0x1.1b8896p-11f);
  p = dyn_fmaf(p, t, 0x1.226feep-21f);
  p = dyn_fmaf(p, t, -0x1.72ec3cp-15f);
  p = dyn_fmaf(p, t, 0x1.dce904p-10f);
  p = dyn_fmaf(p, t, -0x1.b0d348p-9f);
  p = dyn_fmaf(p, t, 0x1.26a02ep-8f);
  p = dyn_fmaf(p, t, -0x1.0564bap-7f);
  p = dyn_fmaf(p, t, 0x1.d63ce0p-7f);
  p = dyn_fmaf(p, t, -0x1.854ac8p-6f);
  p = dyn_fmaf(p, t, 0x1.367e3cp-5f);
  p = dyn_fmaf(p, t, -0x1.e0054ap-5f);
  p = dyn_fmaf(p, t, 0x1.64920cp-4f);
  p = dyn_fmaf(p, t, -0x1.fa9e76p-4f);
  p = dyn_fmaf(p, t, 0x1.5672b2p-3f);
  p = dyn_fmaf(p, t, -0x1.b57032p-3f);
  p = dyn_fmaf(p, t, 0x1.058672p-2f);
  return p;
}

// erfc(x) for x >= 0 from the two pieces above.
DYN_HD float dyn_erfcf(float x) {
  return dyn_exp_neg(-(x * x)) * dyn_erfcx(x);
}

// ---------------------------------------------------------------------------
// Fixed-point conversion.  round-to-nearest-even of f·2^32 as int64, computed
// with one f64 fma against 1.5·2^52 so the integer lands in the low mantissa
// bits (exact for |f·2^32| < 2^51; contributions are clamped to ±2^18).
// ---------------------------------------------------------------------------
DYN_HD int64_t dyn_fx(float f) {
  f = dyn_minf(dyn_maxf(f, -DYN_FX_CLAMP), DYN_FX_CLAMP);
  const double d = dyn_fma((double)f, DYN_FX_SCALE, 6755399441055744.0);
  return dyn_d2bits(d) - (int64_t)0x4338000000000000LL;
}

DYN_HD int64_t dyn_fx_d(double f, double scale) {
  const double d = dyn_fma(f, scale, 6755399441055744.0);
  return dyn_d2bits(d) - (int64_t)0x4338000000000000LL;
}

// dyn_fx for a double value (FP64 pair option): same clamp, same units.
DYN_HD int64_t dyn_fx64(double f) {
  f = f < -(double)DYN_FX_CLAMP ? -(double)DYN_FX_CLAMP : (f > (double)DYN_FX_CLAMP ? (double)DYN_FX_CLAMP : f);
  return dyn_fx_d(f, DYN_FX_SCALE);
}

// ---------------------------------------------------------------------------
// Minimum image.
// ---------------------------------------------------------------------------
// Kernel convention: dv - rint(dv·inv_ext)·ext with one rounding (fma).
DYN_HD float dyn_min_image_rint(float dv, float ext, float inv_ext) {
  return dyn_fmaf(-dyn_rintf(dv * inv_ext), ext, dv);
}

// The reference's CPU convention, reproduced bit for bit: true division,
// round half away from zero, then a separate multiply and subtract
// (src/barostat.rs:162-170).  Used by the neighbor-list build so its pair
// set matches the reference's brute-force build exactly.
DYN_HD float dyn_min_image_ref(float dv, float ext) {
  const float k = dyn_roundf(dv / ext);
  const float m = k * ext;
  return dv - m;
}

// |d|² as the reference's Vec3::magnitude_squared: (x² + y²) + z².
DYN_HD float dyn_norm2_ref(float x, float y, float z) {
  const float a = x * x;
  const float b = y * y;
  const float c = z * z;
  return (a + b) + c;
}

// ---------------------------------------------------------------------------
// Pair interactions.  All return the scalar s with F_on_i = s · d, d = r_i - r_j,
// so the virial contribution d·F is s·r².
// ---------------------------------------------------------------------------
// LJ with combined c6 = 4εσ⁶, c12 = 4εσ¹².
DYN_HD void dyn_lj(float rinv2, float c6, float c12, float* s, float* e) {
  const float rinv6 = rinv2 * rinv2 * rinv2;
  const float t12 = c12 * rinv6;
  *e = (t12 - c6) * rinv6;
  *s = dyn_fmaf(12.0f, t12, -6.0f * c6) * rinv6 * rinv2;
}

// erfc-screened Coulomb.  r2 = r², rinv = 1/r, qq = q_i q_j (Amber-scaled
// charges), beta = Ewald splitting parameter, beta2 = β², two_b_sqrtpi =
// 2β/√π.
DYN_HD void dyn_coul_erfc(float r2, float rinv, float qq, float beta,
                          float beta2, float two_b_sqrtpi, float* s, float* e) {
  const float r = r2 * rinv;
  const float ex = dyn_exp_neg(-(r2 * beta2));
  const float g = dyn_erfcx(r * beta);
  const float qe = qq * ex;
  const float gr = g * rinv;
  *e = qe * gr;
  *s = qe * (rinv * rinv) * (gr + two_b_sqrtpi);
}

// Tabulated erfc Coulomb of the pair kernel.  With x = βr the force scalar
// is s = qq/r³ · φ(x), φ(x) = erfc(x) + (2/√π) x e^{−x²}, and the energy
// E = qq/r · erfc(x).  Both functions of x are tabulated on [0, 4] as
// cubic Hermite pieces (values and derivatives from the closed forms, in
// double; φ'(x) = −(4/√π) x² e^{−x²}, erfc'(x) = −(2/√π) e^{−x²}) over
// DYN_COUL_TAB_N intervals: entry i holds φ's cubic on interval i, entry
// DYN_COUL_TAB_N + i erfc's (so a forces-only evaluation needs the first
// half only), each as (c0, c1, c2, c3) in the local coordinate t ∈ [0, 1).
// Interpolation error ≤ h⁴/384·max|f⁗| ≈ 1e-10 (h = 1/128), below f32
// rounding, so the table reproduces the closed forms to f32 accuracy with
// one load and three fmas in place of exp(−x²)·erfcx(x).  Evaluation uses
// only IEEE-exact operations (mul, floor, conversion, fma), so the host
// replay, given the same table, reproduces it bit for bit.
#define DYN_COUL_TAB_N 512
#define DYN_COUL_TAB_XMAX 4.0f
#define DYN_COUL_TAB_SCALE 128.0f  // intervals per unit of x

// The table's cubic at u = x · DYN_COUL_TAB_SCALE (u given).
DYN_HD float dyn_tab_cubic_u(const float4* tab, int entry_off, float u) {
  const float fl = dyn_floorf(u);
  int i = (int)fl;
  i = i < 0 ? 0 : (i > DYN_COUL_TAB_N - 1 ? DYN_COUL_TAB_N - 1 : i);
  const float t = u - fl;
  const float4 c = tab[entry_off * DYN_COUL_TAB_N + i];
  return dyn_fmaf(dyn_fmaf(dyn_fmaf(c.w, t, c.z), t, c.y), t, c.x);
}

DYN_HD float dyn_tab_cubic(const float4* tab, int entry_off, float x) {
  return dyn_tab_cubic_u(tab, entry_off, x * DYN_COUL_TAB_SCALE);
}

// u = βr·DYN_COUL_TAB_SCALE in one product: the scale is a power of two, so
// r·(β·scale) equals (r·β)·scale bit for bit.
DYN_HD void dyn_coul_erfc_tab(float r2, float rinv, float qq, float beta, const float4* tab,
                              float* s, float* e) {
  const float u = (r2 * rinv) * (beta * DYN_COUL_TAB_SCALE);
  const float phi = dyn_tab_cubic_u(tab, 0, u);
  *s = qq * ((rinv * rinv) * rinv) * phi;
  *e = qq * rinv * dyn_tab_cubic_u(tab, 1, u);
}

// The table (2·DYN_COUL_TAB_N float4 as 8·DYN_COUL_TAB_N floats), built on
// the host in double from the closed forms (a host function).
static inline void dyn_coul_tab_build(float* out) {
  const double h = 1.0 / (double)DYN_COUL_TAB_SCALE;
  const double k2 = 1.1283791670955126;  // 2/√π
  auto phi = [&](double x) { return erfc(x) + k2 * x * exp(-x * x); };
  auto dphi = [&](double x) { return -2.0 * k2 * x * x * exp(-x * x); };
  auto ef = [&](double x) { return erfc(x); };
  auto def = [&](double x) { return -k2 * exp(-x * x); };
  for (int i = 0; i < DYN_COUL_TAB_N; ++i) {
    const double x0 = i * h, x1 = (i + 1) * h;
    for (int which = 0; which < 2; ++which) {
      const double f0 = which ? ef(x0) : phi(x0), f1 = which ? ef(x1) : phi(x1);
      const double d0 = h * (which ? def(x0) : dphi(x0)), d1 = h * (which ? def(x1) : dphi(x1));
      float* c = out + 4 * (which * DYN_COUL_TAB_N + i);
      c[0] = (float)f0;
      c[1] = (float)d0;
      c[2] = (float)(3.0 * (f1 - f0) - 2.0 * d0 - d1);
      c[3] = (float)(2.0 * (f0 - f1) + d0 + d1);
    }
  }
}

// erf-screened Coulomb qq·erf(βr)/r, i.e. the part of a pair's Coulomb
// interaction carried by the reciprocal sum.  Used for exclusion
// corrections.  For small βr the direct series avoids cancellation.
DYN_HD void dyn_coul_erf(float r2, float rinv, float qq, float beta,
                         float beta2, float two_b_sqrtpi, float* s, float* e) {
  const float r = r2 * rinv;
  const float x = r * beta;
  const float ex = dyn_exp_neg(-(r2 * beta2));
  float erf_x;
  if (x < 0.5f) {
    // erf(x) = 2/√π Σ (-1)^n x^(2n+1)/(n!(2n+1)), 7 terms: rel err < 1e-9 for x < 0.5.
    const float x2 = x * x;
    float p = -1.0f / 75600.0f;
    p = dyn_fmaf(p, x2, 1.0f / 9360.0f);
    p = dyn_fmaf(p, x2, -1.0f / 1320.0f);
    p = dyn_fmaf(p, x2, 1.0f / 216.0f);
    p = dyn_fmaf(p, x2, -1.0f / 42.0f);
    p = dyn_fmaf(p, x2, 1.0f / 10.0f);
    p = dyn_fmaf(p, x2, -1.0f / 3.0f);
    p = dyn_fmaf(p, x2, 1.0f);
    erf_x = 1.1283791670955126f * x * p;
  } else {
    erf_x = 1.0f - ex * dyn_erfcx(x);
  }
  const float er = erf_x * rinv;
  *e = qq * er;
  // d/dr [erf(βr)/r] = -erf/r² + (2β/√π) e^{-β²r²}/r ; s = -(dE/dr)/r
  *s = qq * (rinv * rinv) * (er - two_b_sqrtpi * ex);
}

// Staged decoupling schedule (src/alchemical.rs:82-100).
struct DynSchedule {
  float coulomb_scale;
  float coulomb_dscale_dlambda;
  float lj_lambda;
  float lj_dlambda_dlambda;
};

DYN_HD DynSchedule dyn_schedule(float lambda) {
  lambda = dyn_minf(1.0f, dyn_maxf(0.0f, lambda));
  DynSchedule s;
  if (lambda < 0.5f) {
    s.coulomb_scale = 1.0f - 2.0f * lambda;
    s.coulomb_dscale_dlambda = -2.0f;
    s.lj_lambda = 0.0f;
    s.lj_dlambda_dlambda = 0.0f;
  } else {
    s.coulomb_scale = 0.0f;
    s.coulomb_dscale_dlambda = 0.0f;
    s.lj_lambda = 2.0f * lambda - 1.0f;
    s.lj_dlambda_dlambda = 2.0f;
  }
  return s;
}

// Beutler soft-core LJ decoupling with p = 1 (src/cuda/cuda.cu:73-144):
// V = (1-λ) V_LJ(r_sc), r_sc⁶ = r⁶ + α σ_sc⁶ λ, σ_sc = max(σ, σ_min).
// Written with sigma/eps (not c6/c12) because σ_sc depends on σ.
// Returns s (F = s·d), energy and dV/dλ.  r_sc⁻² = (r_sc⁶)^(-1/3) comes from
// a Newton iteration for the inverse cube root (bit-level seed), and
// r_sc⁻¹ = r_sc⁻² · rsqrt(r_sc⁻²); no division or pow is needed.
DYN_HD float dyn_inv_cbrt(float a) {
  // y ≈ a^(-1/3); Newton on f(y) = y^-3 - a: y ← y (4 - a y³)/3.
  float y = dyn_u2f(0x54aaaaaau - dyn_f2u(a) / 3u);
  for (int k = 0; k < 5; ++k) {
    const float y3 = y * y * y;
    y = y * dyn_fmaf(-a, y3, 4.0f) * (1.0f / 3.0f);
  }
  return y;
}

DYN_HD void dyn_lj_soft_core(float r2, float sigma, float eps, float lambda,
                             float* s, float* e, float* dhdl) {
  *s = 0.0f;
  *e = 0.0f;
  *dhdl = 0.0f;
  if (eps == 0.0f || r2 < 0.0f) return;
  lambda = dyn_minf(1.0f, dyn_maxf(0.0f, lambda));
  const float scale = 1.0f - lambda;
  const float ss = dyn_maxf(sigma, DYN_SOFT_CORE_SIGMA_MIN);
  const float ss2 = ss * ss;
  const float ss6 = ss2 * ss2 * ss2;
  const float r6 = r2 * r2 * r2;
  const float rsc6 = r6 + DYN_SOFT_CORE_ALPHA * ss6 * lambda;
  if (!(rsc6 > 0.0f)) return;
  const float inv_rsc2 = dyn_inv_cbrt(rsc6);
  const float inv_rsc = inv_rsc2 * dyn_rsqrtf(inv_rsc2);
  const float sg2 = sigma * sigma;
  const float sr2 = sg2 * inv_rsc2;
  const float sr6 = sr2 * sr2 * sr2;
  const float sr12 = sr6 * sr6;
  const float hard_fmag = 24.0f * eps * dyn_fmaf(2.0f, sr12, -sr6) * inv_rsc;  // |F_LJ(r_sc)|
  const float hard_e = 4.0f * eps * (sr12 - sr6);
  // F(r) = scale · F_LJ(r_sc) · (r/r_sc)^5 r̂, so s = F/r = scale·F_LJ·r⁴/r_sc⁵
  const float inv_rsc5 = inv_rsc2 * inv_rsc2 * inv_rsc;
  *s = hard_fmag * scale * (r2 * r2) * inv_rsc5;
  *e = hard_e * scale;
  // dr_sc/dλ = α σ_sc⁶ / (6 r_sc⁵)
  const float drsc = DYN_SOFT_CORE_ALPHA * ss6 * inv_rsc5 * (1.0f / 6.0f);
  *dhdl = -hard_e - scale * hard_fmag * drsc;
}

// ---------------------------------------------------------------------------
// Cardinal B-spline weights for SPME (Essmann et al. 1995, eq. 4.1 recursion).
// u is the fractional grid coordinate within its cell (w in [0,1)); the
// weights are for grid points k0-(order-1) .. k0 where k0 = floor(position).
// th[j] is the weight for offset j (grid index k0 - order + 1 + j), dth its
// derivative with respect to the fractional coordinate.
// ---------------------------------------------------------------------------
// R = float for the f32 paths (bit-identical on device and host), double for
// the FP64 SPME option.
template <int ORDER, class R = float>
DYN_HD void dyn_bspline(R w, R* th, R* dth) {
  // order 2
  R m[ORDER];
  m[0] = R(1) - w;
  m[1] = w;
  for (int k = 2; k < ORDER - 1; ++k) {
    // raise from order k to k+1
    const R div = R(1) / (R)k;
    m[k] = div * w * m[k - 1];
    for (int j = 1; j < k; ++j) {
      m[k - j] = div * ((w + (R)j) * m[k - j - 1] + ((R)(k - j + 1) - w) * m[k - j]);
    }
    m[0] = div * (R(1) - w) * m[0];
  }
  // derivative of the order-ORDER spline from the order-(ORDER-1) values
  dth[0] = -m[0];
  for (int j = 1; j < ORDER - 1; ++j) dth[j] = m[j - 1] - m[j];
  dth[ORDER - 1] = m[ORDER - 2];
  // final raise to ORDER
  {
    const int k = ORDER - 1;
    const R div = R(1) / (R)k;
    m[k] = div * w * m[k - 1];
    for (int j = 1; j < k; ++j) {
      m[k - j] = div * ((w + (R)j) * m[k - j - 1] + ((R)(k - j + 1) - w) * m[k - j]);
    }
    m[0] = div * (R(1) - w) * m[0];
  }
  for (int j = 0; j < ORDER; ++j) th[j] = m[j];
}
