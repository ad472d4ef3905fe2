// md.hip — device-resident molecular-dynamics step around the nonbonded
// engine: positions, velocities and forces stay in HBM; one step enqueues
// integration kernels, the engine's evaluation and the bonded forces on the
// engine's stream, with no host round trip.
//
// Reference semantics:
//   Langevin-middle step        src/integrate.rs:100-190
//   velocity Verlet (γ = 0)     src/integrate.rs:191-260
//   kick / drift                src/integrate.rs:478-590
//   rigid-water drift           src/solvent/opc_settle.rs:107-198
//   M-site force projection     src/solvent/mod.rs:397-407 (C_H, C_O :89-90)
//   Langevin velocity update    src/thermostat.rs:308-362
//   SHAKE / RATTLE              src/bonded.rs:141-304
//   bonded terms                src/bonded_forces.rs:11-200
//   pressure, C-rescale         src/barostat.rs:324-413, :419-431
// Units as the reference: Å, ps, amu, kcal/mol; accelerations use
// KCAL_TO_NATIVE = 418.4 (src/lib.rs:196).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <numeric>
#include <vector>

#include "common.h"
#include "dyn_math.h"
#include "engine_internal.h"

using namespace dyn;

namespace {

constexpr float kKcalToNative = 418.4f;
constexpr float kKb = 0.83144626f;  // k_B in Å²/ps²/K per amu (src/thermostat.rs:19)
constexpr float kMassO = 16.0f, kMassH = 1.008f;  // src/solvent/mod.rs:58-59
constexpr float kMassW = kMassO + 2.0f * kMassH;
constexpr float kOEpR = 0.15939833f;               // O–M distance (src/solvent/mod.rs:66)
constexpr float kRa = 0.5395199719801114f;         // O to H–H midpoint (opc_settle.rs:41)
constexpr float kCH = (kOEpR / kRa) / 2.0f;
constexpr float kCO = 1.0f - 2.0f * kCH;
constexpr float kMaxAccel = 1e5f;                  // Å/ps² (src/integrate.rs:26)
constexpr float kEpsConstr = 1.0e-8f;              // src/bonded.rs:10
constexpr int kShakeMaxIt = 100;                   // src/bonded.rs:16
constexpr int kMaxClusterCons = 16;                // constraints of one cluster
constexpr int kBlock = 256;

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 ld(const float4& p) { return {p.x, p.y, p.z}; }

struct Box {
  float L[3];
};
// src/barostat.rs:141-170: dv − round(dv/ext)·ext, and rem_euclid wrap into [0, L).
// Reciprocal, square root and reciprocal square root by the hardware
// instructions (about 1 ulp): the integrator's constraint, rigid-body and
// bonded arithmetic needs no correctly rounded division.
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

__device__ __forceinline__ V3 min_image(V3 d, const Box& b) {
  return {dyn_min_image_ref(d.x, b.L[0]), dyn_min_image_ref(d.y, b.L[1]),
          dyn_min_image_ref(d.z, b.L[2])};
}
__device__ __forceinline__ float wrap_coord(float x, float L) {
  // a coordinate within one box length of the box: one exact add or
  // subtract (the same result as fmod); anything farther takes fmod
  float r = x;
  if (r >= L) r -= L;
  else if (r < 0.0f) r += L;
  if (!(r >= 0.0f && r < L)) {
    r = fmodf(x, L);
    if (r < 0.0f) r += L;
  }
  if (!(r < L)) r = 0.0f;
  return r;
}

// Counter-based normal deviates: SplitMix64 finalizer over (seed, step,
// index), Box–Muller on two 24-bit uniforms, with the hardware log2, square
// root and cosine (v_cos_f32 takes its argument in turns): a thermostat's
// noise needs the distribution, not the last bits.
__device__ __host__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float normal_dev(uint64_t seed, uint64_t step, uint64_t idx) {
  const uint64_t h = mix64(mix64(mix64(seed) + step) + idx);
  const float u1 = ((float)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
  const float u2 = ((float)((h >> 16) & 0xFFFFFFull) + 0.5f) * (1.0f / 16777216.0f);
  // −2 ln u1 = −2 ln 2 · log2 u1
  return __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1)) *
         __builtin_amdgcn_cosf(u2);
}

// The same deviate in double on the host (the barostat's noise).
inline double normal_host(uint64_t seed, uint64_t step, uint64_t idx) {
  const uint64_t h = mix64(mix64(mix64(seed) + step) + idx);
  const double u1 = ((double)(h >> 40) + 0.5) * (1.0 / 16777216.0);
  const double u2 = ((double)((h >> 16) & 0xFFFFFFull) + 0.5) * (1.0 / 16777216.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
// Stream index of the barostat's deviate (beyond every site's 3 indices).
constexpr uint64_t kBarostatIndex = 0xB0B0000000000000ull;

struct Fx {
  const int64_t* nb;  // engine forces, site order, component-major
  const int64_t* bd;  // bonded forces over solute atoms, component-major
  uint32_t n_sites, n_dyn;
  // parts != 0: the engine left its forces as parts (no finalize launch);
  // they are summed here, as the finalize kernel would
  ForceParts P;
  int parts;
  int unused;  // no implicit padding: argument packs are compared byte for byte
};
__device__ __forceinline__ void engine_force(const Fx& F, uint32_t s, int64_t f[3]) {
  if (F.parts) {
    site_force_sum(F.P, s, f);
  } else {
    for (int c = 0; c < 3; ++c) f[c] = F.nb[(size_t)c * F.n_sites + s];
  }
}
__device__ __forceinline__ V3 force_of(const Fx& F, uint32_t s) {
  int64_t f[3];
  engine_force(F, s, f);
  for (int c = 0; c < 3; ++c)
    if (s < F.n_dyn) f[c] += F.bd[(size_t)c * F.n_dyn + s];
  return {(float)((double)f[0] * DYN_FX_INV_SCALE), (float)((double)f[1] * DYN_FX_INV_SCALE),
          (float)((double)f[2] * DYN_FX_INV_SCALE)};
}

// Rotate r by angle |ω|·dt about ω (Rodrigues; series for tiny angles).
__device__ __forceinline__ V3 rodrigues(V3 r, V3 w, float dt) {
  const V3 wd = w * dt;
  const float th = hw_sqrt(dot(wd, wd));
  if (th < 1e-12f) {
    const V3 wxr = cross(wd, r);
    return r + wxr + cross(wd, wxr) * 0.5f;
  }
  const V3 n = wd * hw_rcp(th);
  float c, s;
  if (th < 0.5f) {
    // Taylor series to th⁹ / th⁸: below 3e-10 absolute for th < 0.5
    const float t2 = th * th;
    s = th * (1.0f - t2 * (1.0f / 6.0f) * (1.0f - t2 * (1.0f / 20.0f) *
                                                   (1.0f - t2 * (1.0f / 42.0f) * (1.0f - t2 * (1.0f / 72.0f)))));
    c = 1.0f - t2 * 0.5f * (1.0f - t2 * (1.0f / 12.0f) *
                                      (1.0f - t2 * (1.0f / 30.0f) * (1.0f - t2 * (1.0f / 56.0f))));
  } else {
    c = cosf(th);
    s = sinf(th);
  }
  return r * c + cross(n, r) * s + n * (dot(n, r) * (1.0f - c));
}

// Rigid drift of one water over dt: centre-of-mass translation plus the
// rotation whose angular velocity reproduces the molecule's angular
// momentum; velocities become the rigid-body ones; O is wrapped, the H
// sites follow it, and M is placed on the H–O–H bisector.
// One water's sites in registers: O, M, H0, H1 positions and velocities.
struct WaterRegs {
  float4 x[4], v[4];
};

__device__ __forceinline__ void rigid_drift(WaterRegs& W, float dt, const Box& B) {
  const V3 o = ld(W.x[0]);
  const V3 h0 = o + min_image(ld(W.x[2]) - o, B);
  const V3 h1 = o + min_image(ld(W.x[3]) - o, B);
  const V3 vo = ld(W.v[0]), vh0 = ld(W.v[2]), vh1 = ld(W.v[3]);
  const V3 rc = (o * kMassO + h0 * kMassH + h1 * kMassH) * (1.0f / kMassW);
  const V3 vc = (vo * kMassO + vh0 * kMassH + vh1 * kMassH) * (1.0f / kMassW);
  const V3 rO = o - rc, rH0 = h0 - rc, rH1 = h1 - rc;
  const V3 uO = vo - vc, uH0 = vh0 - vc, uH1 = vh1 - vc;
  const V3 Lm = cross(rO, uO) * kMassO + cross(rH0, uH0) * kMassH + cross(rH1, uH1) * kMassH;
  float I[6] = {0, 0, 0, 0, 0, 0};  // xx, yy, zz, xy, xz, yz
  const V3 rs[3] = {rO, rH0, rH1};
  const float ms[3] = {kMassO, kMassH, kMassH};
  for (int k = 0; k < 3; ++k) {
    const V3 r = rs[k];
    const float m = ms[k], r2 = dot(r, r);
    I[0] += m * (r2 - r.x * r.x);
    I[1] += m * (r2 - r.y * r.y);
    I[2] += m * (r2 - r.z * r.z);
    I[3] += -m * r.x * r.y;
    I[4] += -m * r.x * r.z;
    I[5] += -m * r.y * r.z;
  }
  const float ixx = I[0], iyy = I[1], izz = I[2], ixy = I[3], ixz = I[4], iyz = I[5];
  const float det = ixx * (iyy * izz - iyz * iyz) - ixy * (ixy * izz - iyz * ixz) +
                    ixz * (ixy * iyz - iyy * ixz);
  V3 w = {0.f, 0.f, 0.f};
  if (fabsf(det) >= 1.0e-12f) {
    const float id = hw_rcp(det);
    const float a00 = (iyy * izz - iyz * iyz) * id, a01 = (ixz * iyz - ixy * izz) * id,
                a02 = (ixy * iyz - ixz * iyy) * id, a11 = (ixx * izz - ixz * ixz) * id,
                a12 = (ixz * ixy - ixx * iyz) * id, a22 = (ixx * iyy - ixy * ixy) * id;
    w = {a00 * Lm.x + a01 * Lm.y + a02 * Lm.z, a01 * Lm.x + a11 * Lm.y + a12 * Lm.z,
         a02 * Lm.x + a12 * Lm.y + a22 * Lm.z};
  }
  const V3 dlt = vc * dt;
  const V3 rO2 = rodrigues(rO, w, dt), rH02 = rodrigues(rH0, w, dt), rH12 = rodrigues(rH1, w, dt);
  const V3 no = rc + dlt + rO2, nh0 = rc + dlt + rH02, nh1 = rc + dlt + rH12;
  const V3 wo = {wrap_coord(no.x, B.L[0]), wrap_coord(no.y, B.L[1]), wrap_coord(no.z, B.L[2])};
  const V3 wh0 = wo + min_image(nh0 - wo, B), wh1 = wo + min_image(nh1 - wo, B);
  const V3 bis = (wh0 - wo) + (wh1 - wo);
  const float bl2 = dot(bis, bis);
  const V3 wm = bl2 > 0.0f ? wo + bis * (kOEpR * hw_rsqrt(bl2)) : wo;
  W.x[0] = make_float4(wo.x, wo.y, wo.z, 0.f);
  W.x[1] = make_float4(wm.x, wm.y, wm.z, 0.f);
  W.x[2] = make_float4(wh0.x, wh0.y, wh0.z, 0.f);
  W.x[3] = make_float4(wh1.x, wh1.y, wh1.z, 0.f);
  const V3 vO2 = vc + cross(w, rO2), vH02 = vc + cross(w, rH02), vH12 = vc + cross(w, rH12);
  W.v[0] = make_float4(vO2.x, vO2.y, vO2.z, 0.f);
  W.v[2] = make_float4(vH02.x, vH02.y, vH02.z, 0.f);
  W.v[3] = make_float4(vH12.x, vH12.y, vH12.z, 0.f);
  const V3 vm = (vH02 + vH12) * 0.5f;
  W.v[1] = make_float4(vm.x, vm.y, vm.z, 0.f);
}

// Acceleration of solute atom t from the current forces, clamped to
// MAX_ACCEL.
__device__ __forceinline__ V3 accel_atom(const Fx& F, const float* __restrict__ acc_conv,
                                         uint32_t t) {
  V3 a = force_of(F, t) * acc_conv[t];
  const float a2 = dot(a, a);
  if (a2 > kMaxAccel * kMaxAccel) a = a * (kMaxAccel / sqrtf(a2));
  return a;
}

__device__ __forceinline__ float4 kick_by(float4 vi, V3 a, float h) {
  vi.x += a.x * h;
  vi.y += a.y * h;
  vi.z += a.z * h;
  return vi;
}

// Kick of solute atom t: v += a·h from the current forces.
__device__ __forceinline__ void kick_atom(const Fx& F, const float* __restrict__ acc_conv,
                                          uint32_t t, float h, float4* __restrict__ v) {
  v[t] = kick_by(v[t], accel_atom(F, acc_conv, t), h);
}

__device__ __forceinline__ void drift_atom(uint32_t t, float h, float4* __restrict__ x,
                                           const float4* __restrict__ v) {
  const float4 vi = v[t];
  float4 xi = x[t];
  xi.x += vi.x * h;
  xi.y += vi.y * h;
  xi.z += vi.z * h;
  x[t] = xi;
}

// Accelerations of the O and H sites of the water whose O is site base,
// with the M-site force projected onto them.
struct WaterAcc {
  V3 a[3];  // O, H0, H1
};
__device__ __forceinline__ WaterAcc water_accel(const Fx& F, uint32_t base) {
  const V3 fm = force_of(F, base + 1);
  const V3 fo = force_of(F, base) + fm * kCO;
  const V3 f0 = force_of(F, base + 2) + fm * kCH;
  const V3 f1 = force_of(F, base + 3) + fm * kCH;
  const float co = kKcalToNative / kMassO, ch = kKcalToNative / kMassH;
  return {{fo * co, f0 * ch, f1 * ch}};
}

// Kick of the water whose O is site base (M carries no mass).
__device__ __forceinline__ void kick_water(const Fx& F, uint32_t base, float h,
                                           float4* __restrict__ v) {
  const WaterAcc wa = water_accel(F, base);
  const uint32_t ss[3] = {base, base + 2, base + 3};
#pragma unroll
  for (int k = 0; k < 3; ++k) v[ss[k]] = kick_by(v[ss[k]], wa.a[k], h);
}

// The same kick on a water held in registers.
__device__ __forceinline__ void kick_water(const WaterAcc& wa, float h, WaterRegs& W) {
  const int ks[3] = {0, 2, 3};
#pragma unroll
  for (int k = 0; k < 3; ++k) W.v[ks[k]] = kick_by(W.v[ks[k]], wa.a[k], h);
}

// Langevin velocity update (Ornstein–Uhlenbeck) of site s: v ← c·v + σ·ξ per
// component, σ² = k_B T (1 − c²)/m.
__device__ __forceinline__ void langevin_site(float4* __restrict__ v, uint32_t s, float c,
                                              float sig, uint64_t seed, uint64_t step) {
  float4 vv = v[s];
  vv.x = c * vv.x + sig * normal_dev(seed, step, 3ull * s + 0);
  vv.y = c * vv.y + sig * normal_dev(seed, step, 3ull * s + 1);
  vv.z = c * vv.z + sig * normal_dev(seed, step, 3ull * s + 2);
  v[s] = vv;
}

struct ConsArgs {
  const uint32_t* cl_ptr;     // per cluster: range of constraint indices
  const uint2* ij;            // per constraint
  const float* r0sq;
  const float* inv_m;         // 1/m_i + 1/m_j
  const float* mass;          // per solute atom
  uint32_t n_clusters;
  float tol;
};

// SHAKE over cluster c of coupled constraints (one thread), in the
// constraints' input order, until the largest correction is below tol.
__device__ __forceinline__ void shake_cluster(const ConsArgs& C, uint32_t c,
                                              float4* __restrict__ x) {
  const uint32_t k0 = C.cl_ptr[c], k1 = C.cl_ptr[c + 1];
  for (int it = 0; it < kShakeMaxIt; ++it) {
    float max_corr = 0.0f;
    for (uint32_t k = k0; k < k1; ++k) {
      const uint2 p = C.ij[k];
      const float4 a = x[p.x], b = x[p.y];
      const V3 d = v3(b.x - a.x, b.y - a.y, b.z - a.z);
      const float d2 = dot(d, d);
      const float lam = (d2 - C.r0sq[k]) * hw_rcp(2.0f * C.inv_m[k] * fmaxf(d2, kEpsConstr));
      const V3 corr = d * lam;
      const float mi = hw_rcp(C.mass[p.x]), mj = hw_rcp(C.mass[p.y]);
      x[p.x] = make_float4(a.x + corr.x * mi, a.y + corr.y * mi, a.z + corr.z * mi, a.w);
      x[p.y] = make_float4(b.x - corr.x * mj, b.y - corr.y * mj, b.z - corr.z * mj, b.w);
      max_corr = fmaxf(max_corr, dot(corr, corr));
    }
    if (max_corr < C.tol * C.tol) break;
  }
}

// RATTLE: one pass over cluster c's constraints removing the relative
// velocity along each bond.
__device__ __forceinline__ void rattle_cluster(const ConsArgs& C, uint32_t c,
                                               const float4* __restrict__ x,
                                               float4* __restrict__ v) {
  for (uint32_t k = C.cl_ptr[c]; k < C.cl_ptr[c + 1]; ++k) {
    const uint2 p = C.ij[k];
    const float4 a = x[p.x], b = x[p.y];
    const float4 va = v[p.x], vb = v[p.y];
    const V3 r = v3(b.x - a.x, b.y - a.y, b.z - a.z);
    const V3 dv = v3(vb.x - va.x, vb.y - va.y, vb.z - va.z);
    const float lam = dot(dv, r) * hw_rcp(C.inv_m[k] * fmaxf(dot(r, r), kEpsConstr));
    const V3 corr = r * lam;
    const float mi = hw_rcp(C.mass[p.x]), mj = hw_rcp(C.mass[p.y]);
    v[p.x] = make_float4(va.x + corr.x * mi, va.y + corr.y * mi, va.z + corr.z * mi, va.w);
    v[p.y] = make_float4(vb.x - corr.x * mj, vb.y - corr.y * mj, vb.z - corr.z * mj, vb.w);
  }
}

// A solute pressure group the step kernels keep in registers: a constraint
// cluster that is a star of at most 4 constraints around one atom (bonds to
// hydrogen) — atom[0] the centre, atom[k+1] the other atom of constraint k
// (cluster order), bit k of flip set when constraint k lists the spoke first
// — or a lone atom (n = 0).  natoms = 0: another cluster (the general path).
constexpr int kStarMax = 4;
struct StarDesc {
  uint32_t atom[kStarMax + 1];
  uint32_t n, flip, natoms;
};

// SHAKE and RATTLE of a star cluster held in registers: the same operations,
// in the same order, as shake_cluster / rattle_cluster.
__device__ __forceinline__ void shake_star(const ConsArgs& C, uint32_t c, const StarDesc& D,
                                           const float (&im)[kStarMax + 1],
                                           float4 (&X)[kStarMax + 1]) {
  const uint32_t k0 = C.cl_ptr[c];
  for (int it = 0; it < kShakeMaxIt; ++it) {
    float max_corr = 0.0f;
#pragma unroll
    for (int k = 0; k < kStarMax; ++k) {
      if (k < (int)D.n) {
        const bool f = (D.flip >> k) & 1u;
        const float4 a = f ? X[k + 1] : X[0], b = f ? X[0] : X[k + 1];
        const float mi = f ? im[k + 1] : im[0], mj = f ? im[0] : im[k + 1];
        const V3 d = v3(b.x - a.x, b.y - a.y, b.z - a.z);
        const float d2 = dot(d, d);
        const float lam = (d2 - C.r0sq[k0 + k]) * hw_rcp(2.0f * C.inv_m[k0 + k] * fmaxf(d2, kEpsConstr));
        const V3 corr = d * lam;
        const float4 na = make_float4(a.x + corr.x * mi, a.y + corr.y * mi, a.z + corr.z * mi, a.w);
        const float4 nb = make_float4(b.x - corr.x * mj, b.y - corr.y * mj, b.z - corr.z * mj, b.w);
        X[0] = f ? nb : na;
        X[k + 1] = f ? na : nb;
        max_corr = fmaxf(max_corr, dot(corr, corr));
      }
    }
    if (max_corr < C.tol * C.tol) break;
  }
}

__device__ __forceinline__ void rattle_star(const ConsArgs& C, uint32_t c, const StarDesc& D,
                                            const float (&im)[kStarMax + 1],
                                            const float4 (&X)[kStarMax + 1],
                                            float4 (&V)[kStarMax + 1]) {
  const uint32_t k0 = C.cl_ptr[c];
#pragma unroll
  for (int k = 0; k < kStarMax; ++k) {
    if (k < (int)D.n) {
      const bool f = (D.flip >> k) & 1u;
      const float4 a = f ? X[k + 1] : X[0], b = f ? X[0] : X[k + 1];
      const float4 va = f ? V[k + 1] : V[0], vb = f ? V[0] : V[k + 1];
      const float mi = f ? im[k + 1] : im[0], mj = f ? im[0] : im[k + 1];
      const V3 r = v3(b.x - a.x, b.y - a.y, b.z - a.z);
      const V3 dv = v3(vb.x - va.x, vb.y - va.y, vb.z - va.z);
      const float lam = dot(dv, r) * hw_rcp(C.inv_m[k0 + k] * fmaxf(dot(r, r), kEpsConstr));
      const V3 corr = r * lam;
      const float4 na = make_float4(va.x + corr.x * mi, va.y + corr.y * mi, va.z + corr.z * mi, va.w);
      const float4 nb = make_float4(vb.x - corr.x * mj, vb.y - corr.y * mj, vb.z - corr.z * mj, vb.w);
      V[0] = f ? nb : na;
      V[k + 1] = f ? na : nb;
    }
  }
}

__device__ __forceinline__ float4 drift_value(float4 xi, float4 vi, float h) {
  xi.x += vi.x * h;
  xi.y += vi.y * h;
  xi.z += vi.z * h;
  return xi;
}

__device__ __forceinline__ float4 langevin_value(float4 vv, uint32_t s, float c, float sig,
                                                 uint64_t seed, uint64_t step) {
  vv.x = c * vv.x + sig * normal_dev(seed, step, 3ull * s + 0);
  vv.y = c * vv.y + sig * normal_dev(seed, step, 3ull * s + 1);
  vv.z = c * vv.z + sig * normal_dev(seed, step, 3ull * s + 2);
  return vv;
}

struct BondedArgs {
  const float4* x;
  Box B;
  uint32_t n_dyn;
  uint32_t n_bond, n_angle, n_dihe;
  const uint32_t* bond;   // 2 per term
  const float2* bond_p;   // k, r0
  const uint32_t* angle;  // 3 per term (vertex in the middle)
  const float2* angle_p;  // k, θ0
  const uint32_t* dihe;   // 4 per term
  const float4* dihe_p;   // k, periodicity (an integer 0..6), cos(phase), sin(phase)
  int64_t* f;             // fixed point, component-major over solute atoms
  double* partials;       // 4 per block: bond, angle, dihedral energy, virial
  uint32_t nblk;          // workgroups of the system (one partial each)
};

__device__ __forceinline__ void add_f(int64_t* f, uint32_t nd, uint32_t i, V3 v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(f + i), (unsigned long long)dyn_fx(v.x));
  atomicAdd(reinterpret_cast<unsigned long long*>(f + nd + i), (unsigned long long)dyn_fx(v.y));
  atomicAdd(reinterpret_cast<unsigned long long*>(f + 2 * (size_t)nd + i), (unsigned long long)dyn_fx(v.z));
}

// Harmonic bonds and angles (U = k(r − r0)², k(θ − θ0)²) and periodic
// dihedrals (U = k(1 + cos(nφ − γ)), φ the IUPAC torsion angle); one
// thread per term, fixed-point force sums.  The virial Σ r_a·F_a of each
// term is taken relative to one of its atoms (src/bonded.rs:41-46, :79-83,
// :124-130; the same sign convention for every term).
__device__ __forceinline__ void k_md_bonded(BondedArgs A) {
  __shared__ double red[4][kBlock / 64];
  if (blockIdx.x >= A.nblk) return;  // partials exist for the system's own blocks
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  double e[4] = {0.0, 0.0, 0.0, 0.0};
  const uint32_t nd = A.n_dyn;
  if (t < A.n_bond) {
    const uint32_t i = A.bond[2 * t], j = A.bond[2 * t + 1];
    const V3 d = min_image(ld(A.x[j]) - ld(A.x[i]), A.B);
    const float r = hw_sqrt(dot(d, d));
    const float2 p = A.bond_p[t];
    const float dr = r - p.y;
    const V3 fi = d * (2.0f * p.x * dr * hw_rcp(fmaxf(r, 1e-8f)));
    add_f(A.f, nd, i, fi);
    add_f(A.f, nd, j, fi * -1.0f);
    e[0] = (double)(p.x * dr * dr);
    e[3] = -(double)dot(d, fi);  // (r_i − r_j)·F_i
  } else if (t - A.n_bond < A.n_angle) {
    const uint32_t q = t - A.n_bond;
    const uint32_t i = A.angle[3 * q], j = A.angle[3 * q + 1], k = A.angle[3 * q + 2];
    const V3 b0 = min_image(ld(A.x[i]) - ld(A.x[j]), A.B);
    const V3 b2 = min_image(ld(A.x[k]) - ld(A.x[j]), A.B);
    const float l0s = dot(b0, b0), l2s = dot(b2, b2);
    if (l0s >= 1e-8f && l2s >= 1e-8f) {
      const float il0 = hw_rsqrt(l0s), il2 = hw_rsqrt(l2s);
      const V3 u0 = b0 * il0, u2 = b2 * il2;
      const float ct = fminf(1.0f, fmaxf(-1.0f, dot(u0, u2)));
      const float th = acosf(ct);
      const float ist = hw_rcp(fmaxf(hw_sqrt(1.0f - ct * ct), 1e-8f));
      const float2 p = A.angle_p[q];
      const float dth = th - p.y;
      const float dedt = 2.0f * p.x * dth;
      const V3 g0 = (u0 * ct - u2) * (il0 * ist);
      const V3 g2 = (u2 * ct - u0) * (il2 * ist);
      const V3 f0 = g0 * -dedt, f2 = g2 * -dedt;
      add_f(A.f, nd, i, f0);
      add_f(A.f, nd, k, f2);
      add_f(A.f, nd, j, (f0 + f2) * -1.0f);
      e[1] = (double)(p.x * dth * dth);
      e[3] = (double)dot(b0, f0) + (double)dot(b2, f2);
    }
  } else if (t - A.n_bond - A.n_angle < A.n_dihe) {
    const uint32_t q = t - A.n_bond - A.n_angle;
    const uint32_t i = A.dihe[4 * q], j = A.dihe[4 * q + 1], k = A.dihe[4 * q + 2],
                   l = A.dihe[4 * q + 3];
    const V3 b1 = min_image(ld(A.x[j]) - ld(A.x[i]), A.B);
    const V3 b2 = min_image(ld(A.x[k]) - ld(A.x[j]), A.B);
    const V3 b3 = min_image(ld(A.x[l]) - ld(A.x[k]), A.B);
    const V3 n1 = cross(b1, b2), n2 = cross(b2, b3);
    const float n1s = dot(n1, n1), n2s = dot(n2, n2), b2s = dot(b2, b2);
    const float b2l = hw_sqrt(b2s);
    if (n1s >= 1e-6f && n2s >= 1e-6f && b2l >= 1e-6f) {
      // cos φ, sin φ of the IUPAC torsion φ = atan2(|b2| b1·n2, n1·n2); then
      // cos nφ, sin nφ by the angle-addition recurrence and the phase by one
      // more rotation: cos(nφ − γ), sin(nφ − γ) without trigonometric calls
      const float4 p = A.dihe_p[q];
      const float y = b2l * dot(b1, n2), x = dot(n1, n2);
      const float rinv = hw_rsqrt(x * x + y * y);
      const float c1 = x * rinv, s1 = y * rinv;
      float cn = 1.0f, sn = 0.0f;
      const int nper = (int)p.y;
      for (int k = 0; k < nper; ++k) {
        const float c2 = cn * c1 - sn * s1;
        sn = sn * c1 + cn * s1;
        cn = c2;
      }
      const float cos_arg = cn * p.z + sn * p.w;   // cos(nφ − γ)
      const float sin_arg = sn * p.z - cn * p.w;   // sin(nφ − γ)
      const float dedphi = -p.x * p.y * sin_arg;
      // ∂φ/∂x_i = −|b2|/|n1|² n1, ∂φ/∂x_l = |b2|/|n2|² n2; the inner atoms
      // take the lever-arm split of the two
      const V3 gi = n1 * (-b2l * hw_rcp(n1s)), gl = n2 * (b2l * hw_rcp(n2s));
      const float ib2s = hw_rcp(b2s);
      const float ta = dot(b1, b2) * ib2s, tb = dot(b3, b2) * ib2s;
      const V3 gj = gl * tb - gi * (1.0f + ta);
      const V3 gk = gi * ta - gl * (1.0f + tb);
      add_f(A.f, nd, i, gi * -dedphi);
      add_f(A.f, nd, j, gj * -dedphi);
      add_f(A.f, nd, k, gk * -dedphi);
      add_f(A.f, nd, l, gl * -dedphi);
      e[2] = (double)(p.x * (1.0f + cos_arg));
      // relative to atom j: r_i − r_j = −b1, r_k − r_j = b2, r_l − r_j = b2 + b3
      e[3] = -(double)dot(b1, gi * -dedphi) + (double)dot(b2, gk * -dedphi) +
             (double)dot(b2 + b3, gl * -dedphi);
    }
  }
  if (!A.partials) return;  // forces only
  // wave sums by shuffles, then the block's waves in a fixed order
  for (int c = 0; c < 4; ++c)
    for (int m = 32; m >= 1; m >>= 1) e[c] += __shfl_xor(e[c], m, 64);
  if ((threadIdx.x & 63) == 0)
    for (int c = 0; c < 4; ++c) red[c][threadIdx.x / 64] = e[c];
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += red[threadIdx.x][w];
    A.partials[4 * blockIdx.x + threadIdx.x] = t;
  }
}

__device__ __forceinline__ void k_md_zero(int64_t* __restrict__ f, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = 0;
}

// Kinetic energies (amu Å²/ps²) and the bonded-energy partials, reduced by
// one workgroup.  out: [0] kinetic over solute atoms and water O/H site
// velocities (the reference's measure, src/thermostat.rs:46-63), [1..3]
// bond, angle, dihedral energy, [4] kinetic of the rigid-body motion (solute
// atoms as they are; per water the centre-of-mass and rotational part,
// ½ M v_com² + ½ ω·L with I ω = L about the centre of mass).
__device__ __forceinline__ void
k_md_energies(const float4* __restrict__ x, const float4* __restrict__ v,
              const float* __restrict__ mass, uint32_t n_dyn, uint32_t n_water, Box B,
              const double* __restrict__ bpart, uint32_t n_bblk, double* __restrict__ out) {
  __shared__ double red[5][1024];
  const uint32_t t = threadIdx.x;
  double a[5] = {0, 0, 0, 0, 0};
  for (uint32_t i = t; i < n_dyn; i += 1024) {
    const float4 u = v[i];
    const double k = 0.5 * (double)mass[i] * ((double)u.x * u.x + (double)u.y * u.y + (double)u.z * u.z);
    a[0] += k;
    a[4] += k;
  }
  for (uint32_t m = t; m < n_water; m += 1024) {
    const uint32_t b = n_dyn + 4 * m;
    const uint32_t ss[3] = {b, b + 2, b + 3};
    const double ms[3] = {kMassO, kMassH, kMassH};
    double r[3][3], w[3][3], vc[3] = {0, 0, 0}, rc[3] = {0, 0, 0};
    const float4 o = x[b];
    for (int k = 0; k < 3; ++k) {
      const float4 p = x[ss[k]], u = v[ss[k]];
      const V3 d = min_image(v3(p.x - o.x, p.y - o.y, p.z - o.z), B);
      r[k][0] = d.x;
      r[k][1] = d.y;
      r[k][2] = d.z;
      w[k][0] = u.x;
      w[k][1] = u.y;
      w[k][2] = u.z;
      a[0] += 0.5 * ms[k] * (w[k][0] * w[k][0] + w[k][1] * w[k][1] + w[k][2] * w[k][2]);
      for (int c = 0; c < 3; ++c) {
        vc[c] += ms[k] * w[k][c] / (double)kMassW;
        rc[c] += ms[k] * r[k][c] / (double)kMassW;
      }
    }
    double L[3] = {0, 0, 0}, I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 3; ++k) {
      double q[3], u[3];
      for (int c = 0; c < 3; ++c) {
        q[c] = r[k][c] - rc[c];
        u[c] = w[k][c] - vc[c];
      }
      L[0] += ms[k] * (q[1] * u[2] - q[2] * u[1]);
      L[1] += ms[k] * (q[2] * u[0] - q[0] * u[2]);
      L[2] += ms[k] * (q[0] * u[1] - q[1] * u[0]);
      const double q2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
      for (int c = 0; c < 3; ++c)
        for (int e = 0; e < 3; ++e) I[c][e] += ms[k] * ((c == e ? q2 : 0.0) - q[c] * q[e]);
    }
    const double det = I[0][0] * (I[1][1] * I[2][2] - I[1][2] * I[2][1]) -
                       I[0][1] * (I[1][0] * I[2][2] - I[1][2] * I[2][0]) +
                       I[0][2] * (I[1][0] * I[2][1] - I[1][1] * I[2][0]);
    double krot = 0.0;
    if (fabs(det) > 1e-18) {
      const double inv[3][3] = {
          {(I[1][1] * I[2][2] - I[1][2] * I[2][1]) / det, (I[0][2] * I[2][1] - I[0][1] * I[2][2]) / det,
           (I[0][1] * I[1][2] - I[0][2] * I[1][1]) / det},
          {(I[1][2] * I[2][0] - I[1][0] * I[2][2]) / det, (I[0][0] * I[2][2] - I[0][2] * I[2][0]) / det,
           (I[0][2] * I[1][0] - I[0][0] * I[1][2]) / det},
          {(I[1][0] * I[2][1] - I[1][1] * I[2][0]) / det, (I[0][1] * I[2][0] - I[0][0] * I[2][1]) / det,
           (I[0][0] * I[1][1] - I[0][1] * I[1][0]) / det}};
      for (int c = 0; c < 3; ++c) {
        const double om = inv[c][0] * L[0] + inv[c][1] * L[1] + inv[c][2] * L[2];
        krot += 0.5 * om * L[c];
      }
    }
    a[4] += 0.5 * (double)kMassW * (vc[0] * vc[0] + vc[1] * vc[1] + vc[2] * vc[2]) + krot;
  }
  for (uint32_t r = t; r < n_bblk; r += 1024)
    for (int c = 0; c < 3; ++c) a[1 + c] += bpart[4 * r + c];
  for (int c = 0; c < 5; ++c) red[c][t] = a[c];
  __syncthreads();
  for (int s2 = 512; s2 > 0; s2 >>= 1) {
    if (t < (uint32_t)s2)
      for (int c = 0; c < 5; ++c) red[c][t] += red[c][t + s2];
    __syncthreads();
  }
  if (t < 5) out[t] = red[t][0];
}

// Pressure groups: every water (O, H0, H1; M is a virtual site), every
// cluster of solute atoms joined by constraints, every other solute atom
// alone.  Forces inside a group (the rigid water's, the constraints') cancel
// in the group virial, so no constraint virial is needed.
struct GroupArgs {
  const uint32_t* grp_ptr;    // solute groups: range of grp_atoms
  const uint32_t* grp_atoms;
  const float* mass;
  uint32_t n_sgroups, n_water, n_dyn;
  Box B;
};

__device__ __forceinline__ void force_of_d(const Fx& F, uint32_t s, double f[3]) {
  int64_t v[3];
  engine_force(F, s, v);
  for (int c = 0; c < 3; ++c) {
    if (s < F.n_dyn) v[c] += F.bd[(size_t)c * F.n_dyn + s];
    f[c] = (double)v[c] * DYN_FX_INV_SCALE;
  }
}

// Members of group g: site indices and masses; returns the member count.
__device__ __forceinline__ int group_sites(const GroupArgs& G, uint32_t g,
                                           uint32_t site[kMaxClusterCons + 1],
                                           float mass[kMaxClusterCons + 1]) {
  if (g < G.n_sgroups) {
    const uint32_t a0 = G.grp_ptr[g], a1 = G.grp_ptr[g + 1];
    int n = 0;
    for (uint32_t a = a0; a < a1 && n <= kMaxClusterCons; ++a, ++n) {
      site[n] = G.grp_atoms[a];
      mass[n] = G.mass[site[n]];
    }
    return n;
  }
  const uint32_t base = G.n_dyn + 4u * (g - G.n_sgroups);
  site[0] = base;
  site[1] = base + 2;
  site[2] = base + 3;
  mass[0] = kMassO;
  mass[1] = mass[2] = kMassH;
  return 3;
}

// Mass-weighted mean of the members' offsets from the first member
// (minimum image): the centre of mass is x[site[0]] + dr.
__device__ __forceinline__ double group_com(const GroupArgs& G, const float4* __restrict__ x,
                                            const uint32_t* site, const float* mass, int n,
                                            double dr[3]) {
  const float4 ref = x[site[0]];
  double M = 0.0;
  dr[0] = dr[1] = dr[2] = 0.0;
  for (int k = 0; k < n; ++k) {
    const float4 p = x[site[k]];
    const V3 d = min_image(v3(p.x - ref.x, p.y - ref.y, p.z - ref.z), G.B);
    M += (double)mass[k];
    dr[0] += (double)mass[k] * d.x;
    dr[1] += (double)mass[k] * d.y;
    dr[2] += (double)mass[k] * d.z;
  }
  for (int c = 0; c < 3; ++c) dr[c] /= M;
  return M;
}

// Per group: kinetic energy of the centre of mass ½M|V|² (amu Å²/ps²) and
// the intra-group virial Σ (r_a − R)·F_a (kcal/mol; a water's M-site force
// projected onto O and the H sites as the kick does); one partial pair per
// workgroup.
__device__ __forceinline__ void k_md_press(GroupArgs G, const float4* __restrict__ x,
                                           const float4* __restrict__ v, Fx F,
                                           double* __restrict__ partials, uint32_t nblk) {
  __shared__ double red[2][kBlock];
  if (blockIdx.x >= nblk) return;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  double kin = 0.0, cvir = 0.0;
  if (g < G.n_sgroups + G.n_water) {
    uint32_t site[kMaxClusterCons + 1];
    float m[kMaxClusterCons + 1];
    const int n = group_sites(G, g, site, m);
    double dr[3];
    const double M = group_com(G, x, site, m, n, dr);
    double pv[3] = {0, 0, 0};
    for (int k = 0; k < n; ++k) {
      const float4 u = v[site[k]];
      pv[0] += (double)m[k] * u.x;
      pv[1] += (double)m[k] * u.y;
      pv[2] += (double)m[k] * u.z;
    }
    kin = 0.5 * (pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2]) / M;
    if (n > 1) {
      const bool water = g >= G.n_sgroups;
      double fm[3] = {0, 0, 0};
      if (water) force_of_d(F, site[0] + 1, fm);
      const float4 ref = x[site[0]];
      for (int k = 0; k < n; ++k) {
        const float4 p = x[site[k]];
        const V3 d = min_image(v3(p.x - ref.x, p.y - ref.y, p.z - ref.z), G.B);
        double f[3];
        force_of_d(F, site[k], f);
        if (water) {
          const double c = k == 0 ? (double)kCO : (double)kCH;
          for (int e = 0; e < 3; ++e) f[e] += c * fm[e];
        }
        cvir += ((double)d.x - dr[0]) * f[0] + ((double)d.y - dr[1]) * f[1] +
                ((double)d.z - dr[2]) * f[2];
      }
    }
  }
  red[0][threadIdx.x] = kin;
  red[1][threadIdx.x] = cvir;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = red[0][0];
    partials[2 * blockIdx.x + 1] = red[1][0];
  }
}

// Isotropic rescaling by μ about the origin: every group's centre of mass
// moves to μR (its members keep their relative positions; a water's M site
// moves with it) and every velocity is multiplied by 1/μ.  dm = μ − 1.
__device__ __forceinline__ void k_md_rescale(GroupArgs G, float4* __restrict__ x,
                                             float4* __restrict__ v, float dm, float inv_mu) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G.n_sgroups + G.n_water) return;
  uint32_t site[kMaxClusterCons + 1];
  float m[kMaxClusterCons + 1];
  const int n = group_sites(G, g, site, m);
  double dr[3];
  (void)group_com(G, x, site, m, n, dr);
  const float4 ref = x[site[0]];
  const float sx = dm * (float)((double)ref.x + dr[0]);
  const float sy = dm * (float)((double)ref.y + dr[1]);
  const float sz = dm * (float)((double)ref.z + dr[2]);
  const bool water = g >= G.n_sgroups;
  const int nm = water ? 4 : n;
  for (int k = 0; k < nm; ++k) {
    // water: O, M, H0, H1 (M moves with the molecule)
    const uint32_t s = water ? G.n_dyn + 4u * (g - G.n_sgroups) + (uint32_t)k : site[k];
    float4 p = x[s], u = v[s];
    p.x += sx;
    p.y += sy;
    p.z += sz;
    u.x *= inv_mu;
    u.y *= inv_mu;
    u.z *= inv_mu;
    x[s] = p;
    v[s] = u;
  }
}

// The integration around one force evaluation, one thread per pressure
// group (a cluster of constrained solute atoms, a lone atom or a water), so
// each group's kick, drift, constraints and thermostat run back to back in
// one launch.  Solute groups [0, n_clusters) are the constraint clusters.
struct StepArgs {
  Fx F;
  const float* acc_conv;  // per solute atom: KCAL_TO_NATIVE / mass
  GroupArgs G;
  ConsArgs C;
  float h;                // dt/2
  float dt;
  int langevin;           // Langevin-middle (else velocity Verlet)
  float lang_c;           // e^{−γ dt}
  const float* sigma_dyn; // per solute atom: Langevin σ
  float sigma_o, sigma_h;
  uint64_t seed;
  int64_t* fb;            // bonded force sums to clear for the evaluation
  const StarDesc* stars;  // per solute pressure group
  uint32_t water0;        // thread of the first water (n_sgroups rounded up to a wave)
  // the engine's status word: after a failed evaluation (a list or SPME
  // buffer overflowed) the remaining steps of the call leave the state as it
  // is, and the call is repeated from its start
  const unsigned int* status;
};

// Before the force evaluation.  Langevin-middle (src/integrate.rs:100-160):
// kick ½ + drift ½, SHAKE, RATTLE, Langevin, RATTLE, drift ½, SHAKE.
// Velocity Verlet (:213-230): kick ½ + drift 1, SHAKE, RATTLE.  Water: the
// same steps with the rigid drift and no constraints.  Clears the group's
// bonded force sums.  PREV: the group first finishes the previous step
// (k_md_step_b's kick ½ and RATTLE, with the same forces) in registers, so
// two steps meet in one launch; the operations and their order are those of
// the two separate launches.  The index of the step that begins here (it
// seeds the Langevin noise) is *step_dev + PREV.  Thread g < n_sgroups
// holds a solute group, thread water0 + w water w (water0 rounds n_sgroups
// up to a whole wave, so no wave runs both paths).
template <bool PREV>
__device__ __forceinline__ void step_group(const StepArgs& A, float4* __restrict__ x,
                                           float4* __restrict__ v, uint32_t g,
                                           const uint64_t* __restrict__ step_dev) {
  const float hd = A.langevin ? A.h : A.dt;
  if (g < A.G.n_sgroups && A.stars[g].natoms) {
    // a star cluster or a lone atom: the whole sequence in registers
    const StarDesc D = A.stars[g];
    const bool cons = D.n > 0;
    float4 X[kStarMax + 1], V[kStarMax + 1];
    V3 acc[kStarMax + 1];
    float im[kStarMax + 1], sg[kStarMax + 1];
#pragma unroll
    for (int k = 0; k <= kStarMax; ++k) {
      const uint32_t t = D.atom[k <= (int)D.n ? k : 0];
      X[k] = x[t];
      V[k] = v[t];
      acc[k] = accel_atom(A.F, A.acc_conv, t);
      im[k] = hw_rcp(A.C.mass[t]);
      sg[k] = A.langevin ? A.sigma_dyn[t] : 0.f;
    }
    if (PREV) {
#pragma unroll
      for (int k = 0; k <= kStarMax; ++k) V[k] = kick_by(V[k], acc[k], A.h);
      if (cons) rattle_star(A.C, g, D, im, X, V);
    }
#pragma unroll
    for (int k = 0; k <= kStarMax; ++k) {
      V[k] = kick_by(V[k], acc[k], A.h);
      X[k] = drift_value(X[k], V[k], hd);
    }
    if (cons) {
      shake_star(A.C, g, D, im, X);
      rattle_star(A.C, g, D, im, X, V);
    }
    if (A.langevin) {
      const uint64_t step = *step_dev + PREV;
#pragma unroll
      for (int k = 0; k <= kStarMax; ++k)
        if (k <= (int)D.n) V[k] = langevin_value(V[k], D.atom[k], A.lang_c, sg[k], A.seed, step);
      if (cons) rattle_star(A.C, g, D, im, X, V);
#pragma unroll
      for (int k = 0; k <= kStarMax; ++k) X[k] = drift_value(X[k], V[k], A.h);
      if (cons) shake_star(A.C, g, D, im, X);
    }
    const uint32_t nd = A.G.n_dyn;
#pragma unroll
    for (int k = 0; k <= kStarMax; ++k)
      if (k <= (int)D.n) {
        const uint32_t t = D.atom[k];
        x[t] = X[k];
        v[t] = V[k];
        A.fb[t] = 0;
        A.fb[nd + t] = 0;
        A.fb[2 * (size_t)nd + t] = 0;
      }
    return;
  }
  if (g < A.G.n_sgroups) {
    const uint32_t a0 = A.G.grp_ptr[g], a1 = A.G.grp_ptr[g + 1];
    const bool cons = g < A.C.n_clusters;
    if (PREV) {
      for (uint32_t a = a0; a < a1; ++a) kick_atom(A.F, A.acc_conv, A.G.grp_atoms[a], A.h, v);
      if (cons) rattle_cluster(A.C, g, x, v);
    }
    for (uint32_t a = a0; a < a1; ++a) {
      const uint32_t t = A.G.grp_atoms[a];
      kick_atom(A.F, A.acc_conv, t, A.h, v);
      drift_atom(t, hd, x, v);
    }
    if (cons) {
      shake_cluster(A.C, g, x);
      rattle_cluster(A.C, g, x, v);
    }
    if (A.langevin) {
      const uint64_t step = *step_dev + PREV;
      for (uint32_t a = a0; a < a1; ++a) {
        const uint32_t t = A.G.grp_atoms[a];
        langevin_site(v, t, A.lang_c, A.sigma_dyn[t], A.seed, step);
      }
      if (cons) rattle_cluster(A.C, g, x, v);
      for (uint32_t a = a0; a < a1; ++a) drift_atom(A.G.grp_atoms[a], A.h, x, v);
      if (cons) shake_cluster(A.C, g, x);
    }
    const uint32_t nd = A.G.n_dyn;
    for (uint32_t a = a0; a < a1; ++a) {
      const uint32_t t = A.G.grp_atoms[a];
      A.fb[t] = 0;
      A.fb[nd + t] = 0;
      A.fb[2 * (size_t)nd + t] = 0;
    }
    return;
  }
  if (g < A.water0) return;
  const uint32_t base = A.G.n_dyn + 4u * (g - A.water0);
  WaterRegs W;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    W.x[k] = x[base + k];
    W.v[k] = v[base + k];
  }
  const WaterAcc wa = water_accel(A.F, base);
  if (PREV) kick_water(wa, A.h, W);
  kick_water(wa, A.h, W);
  rigid_drift(W, hd, A.G.B);
  if (A.langevin) {
    const uint64_t step = *step_dev + PREV;
    W.v[0] = langevin_value(W.v[0], base, A.lang_c, A.sigma_o, A.seed, step);
    W.v[2] = langevin_value(W.v[2], base + 2, A.lang_c, A.sigma_h, A.seed, step);
    W.v[3] = langevin_value(W.v[3], base + 3, A.lang_c, A.sigma_h, A.seed, step);
    rigid_drift(W, A.h, A.G.B);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    x[base + k] = W.x[k];
    v[base + k] = W.v[k];
  }
}

__device__ __forceinline__ void k_md_step_a(StepArgs A, float4* __restrict__ x,
                                            float4* __restrict__ v,
                                            const uint64_t* __restrict__ step_dev) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (*A.status || g >= A.water0 + A.G.n_water) return;
  step_group<false>(A, x, v, g, step_dev);
}

// The end of one step and the beginning of the next in one launch (no
// energies or pressure are observed between them).  step_dev[0] is the
// index of the step that ends here; the launch advances it once every
// workgroup has used it: each workgroup counts itself in step_dev[1] when
// its threads are done, and the last one stores the new index and clears
// the count.  No fence precedes the count: a workgroup's reads of the
// index have completed (their values are used before the barrier), and the
// positions and velocities it wrote are read by later launches only, which
// the end of this launch makes visible (a device-scope fence per workgroup
// would write back the L2 each time).
__device__ __forceinline__ void k_md_step_ba(StepArgs A, float4* __restrict__ x,
                                             float4* __restrict__ v,
                                             uint64_t* __restrict__ step_dev) {
  // in place of the engine's finalize launch: the evaluation's bookkeeping
  if (A.F.parts && blockIdx.x == 0 && threadIdx.x == 0) evaluation_flags_end(A.F.P.fl, A.F.P.forced);
  if (*A.status) return;  // every workgroup reads the same word: none counts itself
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < A.water0 + A.G.n_water) step_group<true>(A, x, v, g, step_dev);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned* done = reinterpret_cast<unsigned*>(step_dev + 1);
    if (atomicAdd(done, 1u) == gridDim.x - 1) {
      step_dev[0] = step_dev[0] + 1;
      *done = 0u;
    }
  }
}

// After the force evaluation: kick ½ with the new forces, RATTLE; the step's
// last launch advances the device step index.
__device__ __forceinline__ void k_md_step_b(StepArgs A, float4* __restrict__ x,
                                            float4* __restrict__ v,
                                            uint64_t* __restrict__ step_inc) {
  if (*A.status) return;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) ++*step_inc;
  if (g >= A.water0 + A.G.n_water) return;
  if (g < A.G.n_sgroups && A.stars[g].natoms) {
    const StarDesc D = A.stars[g];
    float4 X[kStarMax + 1], V[kStarMax + 1];
    float im[kStarMax + 1];
#pragma unroll
    for (int k = 0; k <= kStarMax; ++k) {
      const uint32_t t = D.atom[k <= (int)D.n ? k : 0];
      X[k] = x[t];
      V[k] = kick_by(v[t], accel_atom(A.F, A.acc_conv, t), A.h);
      im[k] = hw_rcp(A.C.mass[t]);
    }
    if (D.n) rattle_star(A.C, g, D, im, X, V);
#pragma unroll
    for (int k = 0; k <= kStarMax; ++k)
      if (k <= (int)D.n) v[D.atom[k]] = V[k];
    return;
  }
  if (g < A.G.n_sgroups) {
    for (uint32_t a = A.G.grp_ptr[g]; a < A.G.grp_ptr[g + 1]; ++a)
      kick_atom(A.F, A.acc_conv, A.G.grp_atoms[a], A.h, v);
    if (g < A.C.n_clusters) rattle_cluster(A.C, g, x, v);
    return;
  }
  if (g >= A.water0) kick_water(A.F, A.G.n_dyn + 4u * (g - A.water0), A.h, v);
}

}  // namespace

// ---------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------
// An engine call's failure, with the engine's code and message.
static void engine_call(int rc, dyn_engine* eng) {
  if (rc != DYN_OK) throw Error(rc, fmt("engine: %s", dyn_engine_last_error(eng)));
}

namespace dyn {
// Graphs of runs of steady MD steps (begun by the previous step's last
// launch, no energies, joined to the next step), one per SPME cadence phase
// of the members.  A graph holds the kernel arguments and buffer addresses
// of its capture, so it is used only while the device layout generation is
// the one it was captured at; a run is captured the second time it is met
// at one generation (the first, direct run has uploaded every argument pack
// and sized every buffer).
struct StepGraphs {
  struct Entry {
    hipGraphExec_t exec = nullptr;
    uint64_t gen = 0;       // layout generation of exec
    uint64_t seen_gen = 0;  // generation of the last direct run
    std::vector<Profiler::Rec> recs;  // timed scopes recorded into the graph
  };
  std::map<std::vector<uint32_t>, Entry> cache;
  bool disabled = false;  // a capture was refused: direct launches from then on
  uint64_t launches = 0;  // graph launches so far
  void drop(Entry& e) {
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    e.exec = nullptr;
    e.recs.clear();
  }
  StepGraphs() = default;
  StepGraphs(const StepGraphs&) = delete;
  StepGraphs& operator=(const StepGraphs&) = delete;
  ~StepGraphs() {
    for (auto& kv : cache) drop(kv.second);
  }
};
}  // namespace dyn

struct dyn_md {
  std::string err;
  dyn_engine* eng = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;  // the engine's stream (or the batch's)
  dyn_md_params prm{};
  uint32_t n_dyn = 0, n_water = 0, n_sites = 0;
  float box[3] = {0, 0, 0};
  DevBuf<float4> x, v;
  DevBuf<float4> x_ck, v_ck;  // state at the start of the current dyn_md_step call
  DevBuf<int64_t> fb_ck;      // bonded forces and partials at that start
  DevBuf<double> bpart_ck;
  DevBuf<float> mass, acc_conv, lang_sigma;
  float sigma_o = 0.f, sigma_h = 0.f, lang_c = 1.f;
  uint32_t n_bond = 0, n_angle = 0, n_dihe = 0;
  DevBuf<uint32_t> bond, angle, dihe;
  DevBuf<float2> bond_p, angle_p;
  DevBuf<float4> dihe_p;
  DevBuf<int64_t> fb;
  DevBuf<double> bpart;
  unsigned bblocks = 1;
  uint32_t n_clusters = 0, n_cons = 0;
  DevBuf<StarDesc> stars;     // per cluster (n = 0: not a star)
  DevBuf<uint32_t> cl_ptr;
  DevBuf<uint2> cons_ij;
  DevBuf<float> cons_r0sq, cons_invm;
  DevBuf<double> eout;
  PinnedBuf<double> h_out;
  uint64_t step = 0;
  DevBuf<uint64_t> step_dev;  // the step index (the Langevin noise's), a workgroup count
  uint64_t n_eval = 0;  // force evaluations since set_state (SPME cadence)
  bool state_set = false, forces_valid = false;
  // pressure groups (solute constraint clusters and lone atoms; waters
  // follow) and the pressure of the last barostat step
  uint32_t n_sgroups = 0;
  DevBuf<uint32_t> grp_ptr, grp_atoms;
  unsigned pblocks = 1;
  DevBuf<double> ppart;
  PinnedBuf<double> h_ppart, h_bpart;
  double pressure = std::nan(""), virial = std::nan(""), kin_com = std::nan("");
  uint64_t n_pcouple = 0;  // barostat rescalings so far
  uint64_t n_repeats = 0;  // step calls run again from their start (a buffer outgrew its capacity)
  uint64_t graph_steps = 0;  // steps run inside graph launches
  double mass_total = 0.0;  // amu: solute atoms and the waters' O and H
  bool eng_mode_reference = true;
  ArgCache own_cache, side_cache;  // a pass over this MD alone (main, second stream)
  StepGraphs graphs;
  bool in_batch = false;
};

namespace dyn {

// The MD handles of one pass (sharing a stream), with their engines.
struct MdSet {
  std::vector<dyn_md*> m;
  EngineSet es;
  BatchRun R;     // the MD kernels' launches (own argument packs)
  BatchRun side;  // MD work on the engines' second stream (bonded forces)
  Profiler* prof = nullptr;
  StepGraphs* graphs = nullptr;
  // steady evaluations stage their list builds beside the pair kernel (off
  // for the repeat of a call that found a stale list)
  bool staged_builds = true;
};

}  // namespace dyn

namespace {

// One evaluation that waits: a list that outgrew its buffers is rebuilt once
// with enlarged ones.
void engine_set_compute_sync(EngineSet& S, uint32_t flags) {
  engine_set_enqueue(S, flags);
  try {
    engine_set_check(S);
  } catch (const Error& e) {
    if (e.code != DYN_ERR_CAPACITY) throw;
    engine_set_enqueue(S, flags | DYN_COMPUTE_REBUILD);
    engine_set_check(S);
  }
}

// Nonbonded forces (all engines of the set) and bonded forces.  The first
// evaluation after set_state is synchronous, so that a list that outgrows
// its buffers is rebuilt at once; later ones are enqueued only (a failure
// is found at the wait that ends the step call).  Nonbonded energies are
// evaluated only when asked (the last step of a step call), the forces
// alone otherwise.
void md_forces(MdSet& S, bool energies, bool fb_clear = true, bool fuse = false) {
  bool first = false, cadence_skip = true;
  for (dyn_md* m : S.m) {
    first = first || !m->forces_valid;
    // the first evaluation after set_state (or a replay) always runs SPME
    if (!(m->forces_valid && m->prm.spme_ratio > 1 && m->n_eval % m->prm.spme_ratio != 0))
      cadence_skip = false;
  }
  uint32_t flags = DYN_COMPUTE_ALL | (energies ? 0u : DYN_COMPUTE_NO_ENERGY);
  if (cadence_skip) flags = (flags & ~DYN_COMPUTE_RECIP) | DYN_COMPUTE_RECIP_REUSE;
  for (dyn_md* m : S.m) ++m->n_eval;
  if (fb_clear) {  // else the step's first launch has cleared the sums
    const unsigned gz = [&] {
      unsigned g = 1;
      for (dyn_md* m : S.m) g = std::max(g, grid_for(3 * (size_t)std::max(m->n_dyn, 1u), kBlock));
      return g;
    }();
    launch<k_md_zero>(S.R, "k_md_zero", 0, gz, kBlock, 0, [&](uint32_t i) {
      dyn_md& m = *S.m[i];
      return args<k_md_zero>(m.fb.p, 3 * (size_t)m.n_dyn);
    });
  }
  bool any = false;
  unsigned gb = 1;
  for (dyn_md* m : S.m) {
    any = any || (m->n_bond + m->n_angle + m->n_dihe) > 0;
    gb = std::max(gb, m->bblocks);
  }
  // bonded forces on the engines' second stream, beside the pair kernel
  auto bonded = [&](const BatchRun& RS) {
    if (!any) return;
    const BatchRun& RB = RS.stream == S.side.stream ? S.side : S.R;
    Profiler::Scope ps(*S.prof, "md_bonded");
    launch<k_md_bonded, kBlock>(RB, "k_md_bonded", 0, gb, kBlock, 0, [&](uint32_t i) {
      dyn_md& m = *S.m[i];
      BondedArgs A;
      std::memset(&A, 0, sizeof A);
      A.x = m.x.p;
      for (int a = 0; a < 3; ++a) A.B.L[a] = m.box[a];
      A.n_dyn = m.n_dyn;
      A.n_bond = m.n_bond;
      A.n_angle = m.n_angle;
      A.n_dihe = m.n_dihe;
      A.bond = m.bond.p;
      A.bond_p = m.bond_p.p;
      A.angle = m.angle.p;
      A.angle_p = m.angle_p.p;
      A.dihe = m.dihe.p;
      A.dihe_p = m.dihe_p.p;
      A.f = m.fb.p;
      A.partials = energies ? m.bpart.p : nullptr;
      A.nblk = m.bblocks;
      return args<k_md_bonded>(A);
    });
  };
  if (first) {
    engine_set_compute_sync(S.es, flags);
    bonded(S.R);
  } else {
    engine_set_enqueue(S.es,
                       flags | DYN_COMPUTE_ASYNC | (S.staged_builds ? kComputeStagedBuild : 0u) |
                           (fuse ? kComputeFusedFinalize : 0u),
                       bonded);
  }
  for (dyn_md* m : S.m) m->forces_valid = true;
}

// The forces of the engine's last evaluation: its force array, or its parts
// when the evaluation left the sum to the step kernel.
Fx force_view(const dyn_md& m) {
  Fx F;
  std::memset(&F, 0, sizeof F);
  F.nb = engine_forces_device(m.eng);
  F.bd = m.fb.p;
  F.n_sites = m.n_sites;
  F.n_dyn = m.n_dyn;
  if (engine_forces_stale(m.eng)) {
    F.P = engine_force_parts(m.eng);
    F.parts = 1;
  }
  return F;
}

ConsArgs cons_args(const dyn_md& M) {
  ConsArgs C;
  std::memset(&C, 0, sizeof C);
  C.cl_ptr = M.cl_ptr.p;
  C.ij = M.cons_ij.p;
  C.r0sq = M.cons_r0sq.p;
  C.inv_m = M.cons_invm.p;
  C.mass = M.mass.p;
  C.n_clusters = M.n_clusters;
  C.tol = M.prm.shake_tol;
  return C;
}

GroupArgs group_args(const dyn_md& M) {
  GroupArgs G;
  std::memset(&G, 0, sizeof G);
  G.grp_ptr = M.grp_ptr.p;
  G.grp_atoms = M.grp_atoms.p;
  G.mass = M.mass.p;
  G.n_sgroups = M.n_sgroups;
  G.n_water = M.n_water;
  G.n_dyn = M.n_dyn;
  for (int a = 0; a < 3; ++a) G.B.L[a] = M.box[a];
  return G;
}

// Enqueue the pressure parts of every MD of the set (the forces of the last
// evaluation, the current velocities) and their copies to the host.
void md_pressure_enqueue(MdSet& S) {
  Profiler::Scope ps(*S.prof, "md_pressure");
  unsigned g = 1;
  for (dyn_md* m : S.m) g = std::max(g, m->pblocks);
  launch<k_md_press, kBlock>(S.R, "k_md_press", 0, g, kBlock, 0, [&](uint32_t i) {
    dyn_md& m = *S.m[i];
    const Fx F = force_view(m);
    return args<k_md_press>(group_args(m), (const float4*)m.x.p, (const float4*)m.v.p, F,
                            m.ppart.p, m.pblocks);
  });
  for (dyn_md* m : S.m) {
    DYN_HIP(hipMemcpyAsync(m->h_ppart.p, m->ppart.p, sizeof(double) * 2 * m->pblocks,
                           hipMemcpyDeviceToHost, S.R.stream));
    DYN_HIP(hipMemcpyAsync(m->h_bpart.p, m->bpart.p, sizeof(double) * 4 * m->bblocks,
                           hipMemcpyDeviceToHost, S.R.stream));
  }
}

constexpr double kBarPerKcalMolA3 = 69476.95457055373;  // src/barostat.rs:22
constexpr double kKbBarA3 = 138.0649;                   // k_B in bar·Å³/K, src/barostat.rs:26

// Pressure (bar) from the copies md_pressure_enqueue made, after the wait:
// P = (2 K_com + W)/(3V) over the pressure groups (src/barostat.rs:419-431
// with the group virial W = W_nonbonded + W_bonded − Σ_g Σ_a (r_a − R_g)·F_a).
double md_pressure_finish(dyn_md& m) {
  double kin = 0.0, cvir = 0.0, wb = 0.0;
  for (unsigned b = 0; b < m.pblocks; ++b) {
    kin += m.h_ppart.p[2 * b];
    cvir += m.h_ppart.p[2 * b + 1];
  }
  if (m.n_bond + m.n_angle + m.n_dihe)
    for (unsigned b = 0; b < m.bblocks; ++b) wb += m.h_bpart.p[4 * b + 3];
  dyn_energies E{};
  engine_call(dyn_engine_get_energies(m.eng, &E), m.eng);
  const double V = (double)m.box[0] * m.box[1] * m.box[2];
  m.kin_com = kin / (double)kKcalToNative;
  m.virial = E.virial_sr + E.virial_recip + wb - cvir;
  m.pressure = (2.0 * m.kin_com + m.virial) / (3.0 * V) * kBarPerKcalMolA3;
  return m.pressure;
}

// C-rescale length factor μ (Barostat::scale_factor, src/barostat.rs:324-351):
// ΔlnV = (κ/τ)(P − P₀)Δt + √(2κ k_B T Δt/(τV))·ξ, clamped to ±0.1, μ = e^{ΔlnV/3};
// Δt = nstpcouple·dt, ξ the step's deviate of the counter-based stream.
double c_rescale_mu(const dyn_md& m, double p_bar, uint64_t step) {
  const dyn_md_params& P = m.prm;
  const double dt = (double)P.dt * P.nstpcouple;
  const double V = (double)m.box[0] * m.box[1] * m.box[2];
  const double det = (P.compressibility / (double)P.tau_p) * (p_bar - (double)P.pressure_target) * dt;
  const double sig = std::sqrt(2.0 * P.compressibility * kKbBarA3 * (double)P.temperature * dt /
                               ((double)P.tau_p * V));
  const double xi = normal_host(P.seed, step, kBarostatIndex);
  const double dlnv = std::min(0.1, std::max(-0.1, det + sig * xi));
  return std::exp(dlnv / 3.0);
}

// Thread of the first water in the step kernels: the solute groups rounded
// up to a whole wave.
uint32_t step_water0(const dyn_md& m) { return (m.n_sgroups + 63u) / 64u * 64u; }

unsigned step_grid(const MdSet& S) {
  unsigned g = 1;
  for (dyn_md* m : S.m) g = std::max(g, grid_for(step_water0(*m) + m->n_water, kBlock));
  return g;
}

StepArgs step_args(const dyn_md& m) {
  StepArgs A;
  std::memset(&A, 0, sizeof A);
  A.F = force_view(m);
  A.acc_conv = m.acc_conv.p;
  A.G = group_args(m);
  A.C = cons_args(m);
  A.h = 0.5f * m.prm.dt;
  A.dt = m.prm.dt;
  A.langevin = m.prm.gamma > 0.f ? 1 : 0;
  A.lang_c = m.lang_c;
  A.sigma_dyn = m.lang_sigma.p;
  A.sigma_o = m.sigma_o;
  A.sigma_h = m.sigma_h;
  A.seed = m.prm.seed;
  A.fb = m.fb.p;
  A.stars = m.stars.p;
  A.water0 = step_water0(m);
  A.status = engine_status_device(m.eng);
  return A;
}

unsigned group_grid(const MdSet& S) {
  unsigned g = 1;
  for (dyn_md* m : S.m) g = std::max(g, grid_for(m->n_sgroups + m->n_water, kBlock));
  return g;
}

// Pressure coupling of every MD of the set at the end of a step's force
// evaluation (energies on): measure, wait, rescale positions, velocities and
// box (the engine takes the new box: a list rebuild and, with a mesh
// spacing, a grid re-plan at the next evaluation — regen_pme).
void md_barostat(MdSet& S) {
  md_pressure_enqueue(S);
  engine_set_check(S.es);  // waits; a list that overflowed throws (the call replays)
  std::vector<double> mu(S.m.size(), 1.0);
  for (size_t i = 0; i < S.m.size(); ++i) {
    dyn_md& m = *S.m[i];
    const double p = md_pressure_finish(m);
    if (std::isfinite(p)) mu[i] = c_rescale_mu(m, p, m.step);
    if (!(std::isfinite(mu[i]) && mu[i] > 0.0) || std::fabs(mu[i] - 1.0) < 1e-12) mu[i] = 1.0;
  }
  Profiler::Scope ps(*S.prof, "md_rescale");
  launch<k_md_rescale>(S.R, "k_md_rescale", 0, group_grid(S), kBlock, 0, [&](uint32_t i) {
    dyn_md& m = *S.m[i];
    return args<k_md_rescale>(group_args(m), m.x.p, m.v.p, (float)(mu[i] - 1.0),
                              (float)(1.0 / mu[i]));
  });
  for (size_t i = 0; i < S.m.size(); ++i) {
    dyn_md& m = *S.m[i];
    if (mu[i] == 1.0) continue;
    float nb[3];
    for (int a = 0; a < 3; ++a) nb[a] = (float)((double)m.box[a] * mu[i]);
    engine_call(dyn_engine_set_box(m.eng, nb), m.eng);
    for (int a = 0; a < 3; ++a) m.box[a] = nb[a];
    ++m.n_pcouple;
  }
}

// One step of every MD of the set, as src/integrate.rs:100-190 (γ > 0,
// Langevin-middle) or :191-260 (γ = 0, velocity Verlet without thermostat):
// the integration before the force evaluation in one launch, the last
// half-kick after it in another.  Step fractions are in units of each
// system's dt; the scheme (Langevin or not) must be the same for the whole
// set.  begun: the previous call's last launch has already done this step's
// first part.  may_join: nothing observes the state between this step and
// the next, so this step's last launch may also do the next step's first
// part (k_md_step_ba); returns whether it did.
bool md_one_step(MdSet& S, bool energies, bool begun, bool may_join) {
  bool any_invalid = false;
  for (dyn_md* m : S.m) any_invalid = any_invalid || !m->forces_valid;
  if (any_invalid) md_forces(S, false);
  // pressure coupling after this step's force evaluation (the set shares
  // nstpcouple and the step index)
  const uint32_t nst = S.m[0]->prm.nstpcouple;
  const bool couple = nst > 0 && (S.m[0]->step + 1) % nst == 0;
  const bool langevin = S.m[0]->prm.gamma > 0.f;
  energies = energies || couple;
  if (!begun) {
    Profiler::Scope ps(*S.prof, "md_step_a");
    launch<k_md_step_a>(S.R, "k_md_step_a", 0, step_grid(S), kBlock, 0, [&](uint32_t i) {
      dyn_md& m = *S.m[i];
      return args<k_md_step_a>(step_args(m), m.x.p, m.v.p, (const uint64_t*)m.step_dev.p);
    });
  }
  // a joined step without pressure coupling reads the forces in k_md_step_ba
  // only: that launch sums them from their parts (no finalize launch)
  md_forces(S, energies, /*fb_clear=*/false, /*fuse=*/may_join && !couple);
  if (couple && langevin) md_barostat(S);  // before the last half-kick (src/integrate.rs:172-192)
  const bool join = may_join && !(couple && !langevin);
  if (join) {
    Profiler::Scope ps(*S.prof, "md_step_ba");
    launch<k_md_step_ba>(S.R, "k_md_step_ba", 0, step_grid(S), kBlock, 0, [&](uint32_t i) {
      dyn_md& m = *S.m[i];
      return args<k_md_step_ba>(step_args(m), m.x.p, m.v.p, m.step_dev.p);
    });
  } else {
    Profiler::Scope ps(*S.prof, "md_step_b");
    launch<k_md_step_b>(S.R, "k_md_step_b", 0, step_grid(S), kBlock, 0, [&](uint32_t i) {
      dyn_md& m = *S.m[i];
      return args<k_md_step_b>(step_args(m), m.x.p, m.v.p, m.step_dev.p);
    });
  }
  if (couple && !langevin) md_barostat(S);  // at the end of the step (src/integrate.rs:241-310)
  for (dyn_md* m : S.m) ++m->step;
  return join;
}

// n steps of every MD of the set, then a wait for them.  Later force
// evaluations are only enqueued, so a cluster-pair list that outgrows its
// buffers (or a stale list) is found at the wait; the engine has enlarged
// the buffers by then, and the steps run again from the state they started
// from: positions, velocities, step counters, box, and the start's forces
// and energies (so a repeat keeps the SPME cadence of the original call and
// the same random stream, which the step index seeds).  A call that gives
// up restores that start state too before it reports the error.
namespace {
struct MdCk {
  uint64_t step, eval, pcouple;
  float box[3];
  bool forces_valid;
};

void md_checkpoint(MdSet& S, std::vector<MdCk>& ck) {
  ck.resize(S.m.size());
  for (size_t i = 0; i < S.m.size(); ++i) {
    dyn_md& m = *S.m[i];
    const size_t bytes = 16 * (size_t)m.n_sites;
    DYN_HIP(hipMemcpyAsync(m.x_ck.p, m.x.p, bytes, hipMemcpyDeviceToDevice, S.R.stream));
    DYN_HIP(hipMemcpyAsync(m.v_ck.p, m.v.p, bytes, hipMemcpyDeviceToDevice, S.R.stream));
    ck[i] = {m.step, m.n_eval, m.n_pcouple, {m.box[0], m.box[1], m.box[2]}, m.forces_valid};
    if (m.forces_valid) {
      DYN_HIP(hipMemcpyAsync(m.fb_ck.p, m.fb.p, 24 * (size_t)std::max(m.n_dyn, 1u),
                             hipMemcpyDeviceToDevice, S.R.stream));
      DYN_HIP(hipMemcpyAsync(m.bpart_ck.p, m.bpart.p, 32 * (size_t)m.bblocks,
                             hipMemcpyDeviceToDevice, S.R.stream));
      engine_checkpoint(m.eng);  // on the same stream as the set's work
    }
  }
}

// Waits for whatever the failed attempt left in flight, then puts every MD
// of the set back to its checkpoint.
void md_restore(MdSet& S, const std::vector<MdCk>& ck) {
  (void)hipStreamSynchronize(S.R.stream);
  for (size_t i = 0; i < S.m.size(); ++i) {
    dyn_md& m = *S.m[i];
    const size_t bytes = 16 * (size_t)m.n_sites;
    DYN_HIP(hipMemcpyAsync(m.x.p, m.x_ck.p, bytes, hipMemcpyDeviceToDevice, S.R.stream));
    DYN_HIP(hipMemcpyAsync(m.v.p, m.v_ck.p, bytes, hipMemcpyDeviceToDevice, S.R.stream));
    m.step = ck[i].step;
    m.n_eval = ck[i].eval;
    if (m.n_pcouple != ck[i].pcouple) {  // the box of the start
      engine_call(dyn_engine_set_box(m.eng, ck[i].box), m.eng);
      for (int a = 0; a < 3; ++a) m.box[a] = ck[i].box[a];
      m.n_pcouple = ck[i].pcouple;
    }
    m.forces_valid = ck[i].forces_valid;
    if (m.forces_valid) {
      DYN_HIP(hipMemcpyAsync(m.fb.p, m.fb_ck.p, 24 * (size_t)std::max(m.n_dyn, 1u),
                             hipMemcpyDeviceToDevice, S.R.stream));
      DYN_HIP(hipMemcpyAsync(m.bpart.p, m.bpart_ck.p, 32 * (size_t)m.bblocks,
                             hipMemcpyDeviceToDevice, S.R.stream));
      engine_restore(m.eng);
    }
    DYN_HIP(hipMemcpyAsync(m.step_dev.p, &ck[i].step, 8, hipMemcpyHostToDevice, S.R.stream));
    DYN_HIP(hipStreamSynchronize(S.R.stream));
  }
}
}  // namespace

// Steady steps of a call from step k on: steps the previous step's last
// launch has begun, that are not the call's last and couple no pressure.
uint32_t steady_steps(const MdSet& S, uint32_t k, uint32_t n, bool begun) {
  if (!begun) return 0;
  const uint32_t nst = S.m[0]->prm.nstpcouple;
  uint32_t j = k;
  for (; j + 1 < n; ++j) {
    const uint64_t step = S.m[0]->step + (j - k);
    if (nst > 0 && (step + 1) % nst == 0) break;
  }
  return j - k;
}

constexpr uint32_t kGraphSteps = 16;  // steps per graph (a multiple of every SPME ratio in use)

bool graphs_usable(const MdSet& S) {
  // a profile of every kernel runs them one by one; timing one kernel (the
  // pair kernel) records its events into the graphs
  // a repeat (it builds before the pair kernel) runs without the graphs of
  // staged steps
  if (!S.graphs || S.graphs->disabled || (S.prof->enabled && !S.prof->only) ||
      S.staged_builds != (S.m[0]->prm.staged_builds != 0))
    return false;
  for (dyn_md* m : S.m)
    if (!engine_graph_safe(m->eng) || !m->forces_valid || m->prm.direct_launch) return false;
  return true;
}

// kGraphSteps steady steps of every MD of the set: as one graph launch when
// a graph of this run is current, else directly (and captured the second
// time the run is met at the same layout).
void md_graph_block(MdSet& S) {
  StepGraphs& C = *S.graphs;
  std::vector<uint32_t> key;
  for (dyn_md* m : S.m)
    key.push_back(m->prm.spme_ratio > 1 ? (uint32_t)(m->n_eval % m->prm.spme_ratio) : 0u);
  StepGraphs::Entry& E = C.cache[key];
  const uint64_t gen = g_layout_gen.load();
  auto steps = [&]() {
    for (uint32_t k = 0; k < kGraphSteps; ++k) md_one_step(S, false, true, true);
  };
  if (E.exec && E.gen == gen) {
    // the graph enqueues the work; the host keeps its bookkeeping
    g_launch_mode = kLaunchDry;
    try {
      steps();
    } catch (...) {
      g_launch_mode = kLaunchDirect;
      throw;
    }
    g_launch_mode = kLaunchDirect;
    DYN_HIP(hipGraphLaunch(E.exec, S.R.stream));
    ++C.launches;
    for (dyn_md* m : S.m) m->graph_steps += kGraphSteps;
    for (const auto& r : E.recs) S.prof->pending.push_back(r);
    return;
  }
  C.drop(E);
  if (E.seen_gen != gen) {
    // the direct run may itself upload argument packs (new device copies):
    // the layout it leaves is the one a capture would find
    steps();
    E.seen_gen = g_layout_gen.load();
    return;
  }
  DYN_HIP(hipStreamBeginCapture(S.R.stream, hipStreamCaptureModeThreadLocal));
  g_launch_mode = kLaunchCapture;
  S.prof->capture_sink = &E.recs;
  hipGraph_t graph = nullptr;
  try {
    steps();
  } catch (const std::exception& ex) {
    // the refusal may come between a fork and its join: join the second
    // stream into the capture so that it ends cleanly, then drop the graph
    // (nothing of the recorded steps has run)
    if (S.es.side.stream && S.es.side.stream != S.R.stream)
      stream_order(S.es.side.stream, S.R.stream, S.es.ev_join);
    S.prof->capture_sink = nullptr;
    g_launch_mode = kLaunchDirect;
    (void)hipStreamEndCapture(S.R.stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    C.disabled = true;
    g_capture_refusal = ex.what();
    throw CaptureRefused(std::string("the steps could not be recorded into a graph: ") + ex.what());
  }
  S.prof->capture_sink = nullptr;
  g_launch_mode = kLaunchDirect;
  DYN_HIP(hipStreamEndCapture(S.R.stream, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess || g_layout_gen.load() != gen) {
    // nothing of the recorded steps has run: the call repeats without graphs
    if (exec) (void)hipGraphExecDestroy(exec);
    C.disabled = true;
    throw CaptureRefused("a graph of the steps could not be made");
  }
  E.exec = exec;
  E.gen = gen;
  DYN_HIP(hipGraphLaunch(E.exec, S.R.stream));
  ++C.launches;
  for (dyn_md* m : S.m) m->graph_steps += kGraphSteps;
  for (const auto& r : E.recs) S.prof->pending.push_back(r);
}

void md_run(MdSet& S, uint32_t n) {
  if (!n) return;
  std::vector<MdCk> ck;
  md_checkpoint(S, ck);
  for (int attempt = 0;; ++attempt) {
    // a repeat builds its lists before the pair kernel (it may follow a
    // stale staged list)
    S.staged_builds = attempt == 0 && S.m[0]->prm.staged_builds != 0;
    try {
      const auto t0 = std::chrono::steady_clock::now();
      // a barostat step waits mid-call, and may find an overflow there
      bool begun = false;
      uint32_t k = 0;
      while (k < n) {
        if (steady_steps(S, k, n, begun) >= kGraphSteps && graphs_usable(S)) {
          md_graph_block(S);
          k += kGraphSteps;
          continue;
        }
        begun = md_one_step(S, k + 1 == n, begun, k + 1 < n);
        ++k;
      }
      const auto t1 = std::chrono::steady_clock::now();
      engine_set_check(S.es);
      const auto t2 = std::chrono::steady_clock::now();
      S.prof->add_host("md_issue_steps", std::chrono::duration<double, std::milli>(t1 - t0).count());
      S.prof->add_host("md_wait", std::chrono::duration<double, std::milli>(t2 - t1).count());
      for (dyn_md* m : S.m) m->n_repeats += (uint64_t)attempt;
      return;
    } catch (const CaptureRefused&) {
      // graphs are off for this set from now on; the call runs again
    } catch (const Error& e) {
      // a buffer that overflowed has been enlarged: the call runs again (twice
      // at most); any other failure ends it
      if (e.code != DYN_ERR_CAPACITY || attempt >= 2) {
        try {
          md_restore(S, ck);
        } catch (const Error&) {
          // the original error is the one to report
        }
        throw;
      }
    }
    md_restore(S, ck);
  }
}

// A set of one: this MD and its engine.
MdSet self_set(dyn_md* m) {
  MdSet S;
  S.m = {m};
  S.es = engine_self_set(m->eng);
  S.prof = engine_profiler(m->eng);
  S.R.stream = m->stream;
  S.R.cache = &m->own_cache;
  S.R.n = 1;
  S.side.stream = S.es.side.stream;
  S.side.cache = &m->side_cache;
  S.side.n = 1;
  S.graphs = &m->graphs;
  return S;
}

}  // namespace

// ---------------------------------------------------------------------------
// C entry points
// ---------------------------------------------------------------------------
#define MD_CHECK(m)                        \
  if (!(m)) {                              \
    set_global_error("null MD handle");    \
    return DYN_ERR_ARG;                    \
  }

extern "C" int dyn_md_create(dyn_md** out, dyn_engine* eng, const dyn_md_params* p,
                             const float* mass_dyn, uint32_t n_bonds, const uint32_t* bond_ij,
                             const float* bond_k_r0, uint32_t n_angles, const uint32_t* angle_ijk,
                             const float* angle_k_theta0, uint32_t n_dihedrals,
                             const uint32_t* dihedral_ijkl, const float* dihedral_k_n_phase,
                             uint32_t n_constraints, const uint32_t* constraint_ij,
                             const float* constraint_r0) {
  try {
    DYN_REQUIRE(out && eng && p, "null argument");
    *out = nullptr;
    DYN_REQUIRE(p->dt > 0.f && p->dt < 0.01f, "dt must be in (0, 0.01) ps");
    DYN_REQUIRE(p->temperature >= 0.f && std::isfinite(p->temperature), "bad temperature");
    DYN_REQUIRE(p->gamma >= 0.f && std::isfinite(p->gamma), "gamma must be ≥ 0");
    DYN_REQUIRE(p->shake_tol > 0.f, "shake_tol must be > 0");
    if (p->nstpcouple) {
      DYN_REQUIRE(p->tau_p > 0.f && std::isfinite(p->tau_p), "tau_p must be > 0");
      DYN_REQUIRE(p->compressibility > 0.0 && std::isfinite(p->compressibility),
                  "compressibility must be > 0");
      DYN_REQUIRE(std::isfinite(p->pressure_target), "bad pressure_target");
    }
    DYN_REQUIRE(p->spme_ratio <= 64, "spme_ratio must be at most 64");
    std::unique_ptr<dyn_md> M(new dyn_md());
    M->eng = eng;
    M->device = engine_device(eng);
    M->stream = engine_stream(eng);
    DeviceGuard g(M->device);
    M->prm = *p;
    engine_sizes(eng, &M->n_dyn, &M->n_water);
    engine_box(eng, M->box);
    DYN_REQUIRE(M->box[0] > 0.f, "set the engine's box first");
    const uint32_t nd = M->n_dyn, nw = M->n_water, ns = nd + 4 * nw;
    M->n_sites = ns;
    DYN_REQUIRE(nd == 0 || mass_dyn, "mass_dyn is null");
    std::vector<float> m(nd), conv(nd), sig(nd);
    const float c = std::exp(-p->gamma * p->dt);
    const float s2 = std::max(0.0f, 1.0f - c * c);
    const float num = kKb * p->temperature * s2;
    for (uint32_t i = 0; i < nd; ++i) {
      DYN_REQUIRE(mass_dyn[i] > 0.f && std::isfinite(mass_dyn[i]), "atom %u: mass must be > 0", i);
      m[i] = mass_dyn[i];
      conv[i] = kKcalToNative / m[i];
      sig[i] = std::sqrt(num / m[i]);
    }
    M->mass_total = (double)kMassW * nw;
    for (uint32_t i = 0; i < nd; ++i) M->mass_total += (double)m[i];
    M->eng_mode_reference = engine_mode(eng) == DYN_MODE_REFERENCE;
    M->lang_c = c;
    M->sigma_o = std::sqrt(num / kMassO);
    M->sigma_h = std::sqrt(num / kMassH);
    auto up = [&](auto& buf, const void* src, size_t count, size_t elem) {
      buf.reserve(std::max<size_t>(count, 1));
      if (count)
        DYN_HIP(hipMemcpyAsync(buf.p, src, count * elem, hipMemcpyHostToDevice, M->stream));
    };
    up(M->mass, m.data(), nd, 4);
    up(M->acc_conv, conv.data(), nd, 4);
    up(M->lang_sigma, sig.data(), nd, 4);
    auto check_idx = [&](const uint32_t* idx, uint32_t n, int per, const char* what) {
      DYN_REQUIRE(n == 0 || idx, "%s indices are null", what);
      for (uint64_t k = 0; k < (uint64_t)n * per; ++k)
        DYN_REQUIRE(idx[k] < nd, "%s term %llu: atom %u out of range", what,
                    (unsigned long long)(k / per), idx[k]);
    };
    check_idx(bond_ij, n_bonds, 2, "bond");
    check_idx(angle_ijk, n_angles, 3, "angle");
    check_idx(dihedral_ijkl, n_dihedrals, 4, "dihedral");
    check_idx(constraint_ij, n_constraints, 2, "constraint");
    DYN_REQUIRE(n_bonds == 0 || bond_k_r0, "bond parameters are null");
    DYN_REQUIRE(n_angles == 0 || angle_k_theta0, "angle parameters are null");
    DYN_REQUIRE(n_dihedrals == 0 || dihedral_k_n_phase, "dihedral parameters are null");
    M->n_bond = n_bonds;
    M->n_angle = n_angles;
    M->n_dihe = n_dihedrals;
    up(M->bond, bond_ij, 2 * (size_t)n_bonds, 4);
    up(M->bond_p, bond_k_r0, n_bonds, 8);
    up(M->angle, angle_ijk, 3 * (size_t)n_angles, 4);
    up(M->angle_p, angle_k_theta0, n_angles, 8);
    up(M->dihe, dihedral_ijkl, 4 * (size_t)n_dihedrals, 4);
    std::vector<float4> dp(n_dihedrals);
    for (uint32_t k = 0; k < n_dihedrals; ++k) {
      const float per = dihedral_k_n_phase[3 * k + 1];
      DYN_REQUIRE(per >= 0.f && per <= 6.f && per == std::floor(per),
                  "dihedral %u: periodicity %g is not an integer in [0, 6]", k, (double)per);
      const double ph = (double)dihedral_k_n_phase[3 * k + 2];
      dp[k] = make_float4(dihedral_k_n_phase[3 * k], per, (float)std::cos(ph), (float)std::sin(ph));
    }
    up(M->dihe_p, dp.data(), n_dihedrals, 16);
    M->bblocks = grid_for(std::max<uint32_t>(n_bonds + n_angles + n_dihedrals, 1), kBlock);
    M->bpart.reserve(4 * (size_t)M->bblocks);
    DYN_HIP(hipMemsetAsync(M->bpart.p, 0, sizeof(double) * 4 * M->bblocks, M->stream));
    M->fb.reserve(3 * (size_t)std::max(nd, 1u));
    // constraint clusters: connected components of the constraint graph
    std::vector<uint32_t> parent(nd);
    std::iota(parent.begin(), parent.end(), 0u);
    std::function<uint32_t(uint32_t)> find = [&](uint32_t a) {
      while (parent[a] != a) a = parent[a] = parent[parent[a]];
      return a;
    };
    for (uint32_t k = 0; k < n_constraints; ++k) {
      DYN_REQUIRE(constraint_r0 && constraint_r0[k] > 0.f, "constraint %u: r0 must be > 0", k);
      DYN_REQUIRE(constraint_ij[2 * k] != constraint_ij[2 * k + 1], "constraint %u joins an atom to itself", k);
      parent[find(constraint_ij[2 * k])] = find(constraint_ij[2 * k + 1]);
    }
    std::vector<std::vector<uint32_t>> groups;
    std::vector<int64_t> gid(nd, -1);
    for (uint32_t k = 0; k < n_constraints; ++k) {
      const uint32_t r = find(constraint_ij[2 * k]);
      if (gid[r] < 0) {
        gid[r] = (int64_t)groups.size();
        groups.emplace_back();
      }
      groups[gid[r]].push_back(k);
    }
    std::vector<uint32_t> ptr(1, 0);
    std::vector<uint2> ij;
    std::vector<float> r0sq, invm;
    for (auto& gq : groups) {
      DYN_REQUIRE(gq.size() <= (size_t)kMaxClusterCons, "a constraint cluster has %zu constraints "
                  "(at most %d: constrain bonds to hydrogen only)", gq.size(), kMaxClusterCons);
      for (uint32_t k : gq) {
        const uint32_t i = constraint_ij[2 * k], j = constraint_ij[2 * k + 1];
        ij.push_back(make_uint2(i, j));
        r0sq.push_back(constraint_r0[k] * constraint_r0[k]);
        invm.push_back(1.0f / m[i] + 1.0f / m[j]);
      }
      ptr.push_back((uint32_t)ij.size());
    }
    M->n_clusters = (uint32_t)groups.size();
    M->n_cons = n_constraints;
    // star clusters (≤ 4 constraints sharing one atom) and lone atoms run in
    // registers; the descriptors of the lone atoms are added with the
    // pressure groups below
    std::vector<StarDesc> st(groups.size());
    {
      for (size_t c = 0; c < groups.size(); ++c) {
        StarDesc& D = st[c];
        std::memset(&D, 0, sizeof D);
        const auto& gq = groups[c];
        if (gq.size() > (size_t)kStarMax) continue;
        const uint32_t a = constraint_ij[2 * gq[0]], b = constraint_ij[2 * gq[0] + 1];
        for (uint32_t center : {a, b}) {
          bool ok = true;
          for (uint32_t k : gq)
            ok = ok && (constraint_ij[2 * k] == center || constraint_ij[2 * k + 1] == center);
          // the spokes must be distinct atoms (no constraint listed twice)
          for (size_t u = 0; ok && u < gq.size(); ++u)
            for (size_t w = 0; w < u; ++w) {
              const uint32_t su = constraint_ij[2 * gq[u]] == center ? constraint_ij[2 * gq[u] + 1]
                                                                     : constraint_ij[2 * gq[u]];
              const uint32_t sw = constraint_ij[2 * gq[w]] == center ? constraint_ij[2 * gq[w] + 1]
                                                                     : constraint_ij[2 * gq[w]];
              ok = ok && su != sw;
            }
          if (!ok) continue;
          D.atom[0] = center;
          for (size_t k = 0; k < gq.size(); ++k) {
            const uint32_t i = constraint_ij[2 * gq[k]], j = constraint_ij[2 * gq[k] + 1];
            D.atom[k + 1] = i == center ? j : i;
            if (i != center) D.flip |= 1u << k;
          }
          D.n = (uint32_t)gq.size();
          D.natoms = D.n + 1;
          break;
        }
      }
    }
    // pressure groups: the atoms of each constraint cluster, then every atom
    // in no constraint alone
    {
      std::vector<uint32_t> gptr(1, 0), gat;
      std::vector<std::vector<uint32_t>> members(groups.size());
      std::vector<char> in_cons(nd, 0);
      for (uint32_t k = 0; k < n_constraints; ++k)
        in_cons[constraint_ij[2 * k]] = in_cons[constraint_ij[2 * k + 1]] = 1;
      for (uint32_t a = 0; a < nd; ++a)
        if (in_cons[a]) members[gid[find(a)]].push_back(a);
      for (auto& mem : members) {
        gat.insert(gat.end(), mem.begin(), mem.end());
        gptr.push_back((uint32_t)gat.size());
      }
      for (uint32_t a = 0; a < nd; ++a)
        if (!in_cons[a]) {
          gat.push_back(a);
          gptr.push_back((uint32_t)gat.size());
          StarDesc D;
          std::memset(&D, 0, sizeof D);
          D.atom[0] = a;
          D.natoms = 1;
          st.push_back(D);
        }
      M->n_sgroups = (uint32_t)gptr.size() - 1;
      if (st.empty()) st.resize(1);
      up(M->stars, st.data(), st.size(), sizeof(StarDesc));
      up(M->grp_ptr, gptr.data(), gptr.size(), 4);
      up(M->grp_atoms, gat.data(), gat.size(), 4);
      M->pblocks = grid_for(std::max<uint32_t>(M->n_sgroups + nw, 1), kBlock);
      M->ppart.reserve(2 * (size_t)M->pblocks);
      M->h_ppart.reserve(2 * (size_t)M->pblocks);
      M->h_bpart.reserve(4 * (size_t)M->bblocks);
    }
    up(M->cl_ptr, ptr.data(), ptr.size(), 4);
    up(M->cons_ij, ij.data(), ij.size(), 8);
    up(M->cons_r0sq, r0sq.data(), r0sq.size(), 4);
    up(M->cons_invm, invm.data(), invm.size(), 4);
    M->x.reserve(ns);
    M->v.reserve(ns);
    M->x_ck.reserve(ns);
    M->fb_ck.reserve(3 * (size_t)std::max(nd, 1u));
    M->bpart_ck.reserve(4 * (size_t)M->bblocks);
    M->v_ck.reserve(ns);
    M->step_dev.reserve(2);  // the step index, the workgroup count of k_md_step_ba
    DYN_HIP(hipMemsetAsync(M->step_dev.p, 0, 16, M->stream));
    M->eout.reserve(8);
    M->h_out.reserve(8);
    DYN_HIP(hipStreamSynchronize(M->stream));
    *out = M.release();
    return DYN_OK;
  } catch (const Error& e) {
    set_global_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_global_error(e.what());
    return DYN_ERR_HIP;
  }
}

extern "C" void dyn_md_destroy(dyn_md* md) {
  if (!md) return;
  try {
    DeviceGuard g(md->device);
    (void)hipStreamSynchronize(md->stream);
    delete md;
  } catch (...) {
  }
}

extern "C" const char* dyn_md_last_error(const dyn_md* md) {
  return md ? md->err.c_str() : dyn_global_last_error();
}

extern "C" int dyn_md_set_state(dyn_md* md, const float* pos_dyn, const float* vel_dyn,
                                const float* pos_water, const float* vel_water) {
  MD_CHECK(md);
  DYN_API_BODY(md, {
    DeviceGuard g(md->device);
    const uint32_t nd = md->n_dyn, nw = md->n_water, ns = md->n_sites;
    DYN_REQUIRE(nd == 0 || (pos_dyn && vel_dyn), "solute state is null");
    DYN_REQUIRE(nw == 0 || (pos_water && vel_water), "water state is null");
    std::vector<float4> hx(ns), hv(ns);
    for (uint32_t i = 0; i < nd; ++i) {
      hx[i] = make_float4(pos_dyn[3 * i], pos_dyn[3 * i + 1], pos_dyn[3 * i + 2], 0.f);
      hv[i] = make_float4(vel_dyn[3 * i], vel_dyn[3 * i + 1], vel_dyn[3 * i + 2], 0.f);
    }
    for (uint32_t m = 0; m < nw; ++m) {
      for (int k = 0; k < 4; ++k) {
        const float* p = pos_water + 12 * (size_t)m + 3 * k;
        hx[nd + 4 * m + k] = make_float4(p[0], p[1], p[2], 0.f);
      }
      // velocities of O, H0, H1; M takes the mean of the H velocities
      const float* q = vel_water + 9 * (size_t)m;
      hv[nd + 4 * m] = make_float4(q[0], q[1], q[2], 0.f);
      hv[nd + 4 * m + 2] = make_float4(q[3], q[4], q[5], 0.f);
      hv[nd + 4 * m + 3] = make_float4(q[6], q[7], q[8], 0.f);
      hv[nd + 4 * m + 1] = make_float4(0.5f * (q[3] + q[6]), 0.5f * (q[4] + q[7]), 0.5f * (q[5] + q[8]), 0.f);
    }
    for (uint32_t s = 0; s < ns; ++s)
      DYN_REQUIRE(std::isfinite(hx[s].x + hx[s].y + hx[s].z + hv[s].x + hv[s].y + hv[s].z),
                  "non-finite state at site %u", s);
    DYN_HIP(hipMemcpyAsync(md->x.p, hx.data(), 16 * (size_t)ns, hipMemcpyHostToDevice, md->stream));
    DYN_HIP(hipMemcpyAsync(md->v.p, hv.data(), 16 * (size_t)ns, hipMemcpyHostToDevice, md->stream));
    DYN_HIP(hipStreamSynchronize(md->stream));
    engine_call(dyn_engine_set_positions_device(md->eng, reinterpret_cast<const float*>(md->x.p)),
                md->eng);
    md->state_set = true;
    md->forces_valid = false;
    md->n_eval = 0;
  });
}

extern "C" int dyn_md_step(dyn_md* md, uint32_t n_steps) {
  MD_CHECK(md);
  DYN_API_BODY(md, {
    DeviceGuard g(md->device);
    if (!md->state_set) throw Error(DYN_ERR_STATE, "state not set");
    if (md->in_batch) throw Error(DYN_ERR_STATE, "this MD is stepped by a batch (dyn_batch_step)");
    MdSet S = self_set(md);
    md_run(S, n_steps);
  });
}

extern "C" int dyn_md_get_state(dyn_md* md, float* pos_dyn, float* vel_dyn, float* pos_water,
                                float* vel_water) {
  MD_CHECK(md);
  DYN_API_BODY(md, {
    DeviceGuard g(md->device);
    if (!md->state_set) throw Error(DYN_ERR_STATE, "state not set");
    engine_call(dyn_engine_sync(md->eng), md->eng);
    const uint32_t nd = md->n_dyn, nw = md->n_water, ns = md->n_sites;
    std::vector<float4> hx(ns), hv(ns);
    DYN_HIP(hipMemcpyAsync(hx.data(), md->x.p, 16 * (size_t)ns, hipMemcpyDeviceToHost, md->stream));
    DYN_HIP(hipMemcpyAsync(hv.data(), md->v.p, 16 * (size_t)ns, hipMemcpyDeviceToHost, md->stream));
    DYN_HIP(hipStreamSynchronize(md->stream));
    for (uint32_t i = 0; i < nd; ++i) {
      if (pos_dyn) memcpy(pos_dyn + 3 * i, &hx[i], 12);
      if (vel_dyn) memcpy(vel_dyn + 3 * i, &hv[i], 12);
    }
    for (uint32_t m = 0; m < nw; ++m) {
      if (pos_water)
        for (int k = 0; k < 4; ++k) memcpy(pos_water + 12 * (size_t)m + 3 * k, &hx[nd + 4 * m + k], 12);
      if (vel_water) {
        memcpy(vel_water + 9 * (size_t)m, &hv[nd + 4 * m], 12);
        memcpy(vel_water + 9 * (size_t)m + 3, &hv[nd + 4 * m + 2], 12);
        memcpy(vel_water + 9 * (size_t)m + 6, &hv[nd + 4 * m + 3], 12);
      }
    }
  });
}

extern "C" int dyn_md_get_energies(dyn_md* md, dyn_md_energies* out) {
  MD_CHECK(md);
  DYN_API_BODY(md, {
    DeviceGuard g(md->device);
    DYN_REQUIRE(out, "out is null");
    if (!md->state_set) throw Error(DYN_ERR_STATE, "state not set");
    MdSet S = self_set(md);
    if (!md->forces_valid) md_forces(S, true);
    Box B;
    for (int a = 0; a < 3; ++a) B.L[a] = md->box[a];
    launch<k_md_energies, 1024>(S.R, "k_md_energies", 0, 1, 1024, 0, [&](uint32_t) {
      return args<k_md_energies>(md->x.p, md->v.p, md->mass.p, md->n_dyn, md->n_water, B,
                                 md->bpart.p,
                                 (md->n_bond + md->n_angle + md->n_dihe) ? md->bblocks : 0u,
                                 md->eout.p);
    });
    DYN_HIP(hipMemcpyAsync(md->h_out.p, md->eout.p, 5 * sizeof(double), hipMemcpyDeviceToHost,
                           md->stream));
    engine_call(dyn_engine_sync(md->eng), md->eng);
    DYN_HIP(hipStreamSynchronize(md->stream));
    dyn_md_energies E{};
    E.kinetic = md->h_out.p[0] / (double)kKcalToNative;
    E.kinetic_rigid = md->h_out.p[4] / (double)kKcalToNative;
    E.bond = md->h_out.p[1];
    E.angle = md->h_out.p[2];
    E.dihedral = md->h_out.p[3];
    // degrees of freedom: 3 per solute atom less one per constraint, 6 per
    // rigid water
    const double dof = 3.0 * md->n_dyn - (double)md->n_cons + 6.0 * md->n_water;
    E.dof = dof;
    E.temperature = dof > 0 ? 2.0 * md->h_out.p[4] / (dof * (double)kKb) : 0.0;
    E.steps = md->step;
    // pressure of the current velocities and the last evaluation's forces
    md_pressure_enqueue(S);
    engine_call(dyn_engine_sync(md->eng), md->eng);
    DYN_HIP(hipStreamSynchronize(md->stream));
    try {
      E.pressure = md_pressure_finish(*md);
    } catch (const Error& e) {
      if (e.code != DYN_ERR_STATE) throw;  // the last evaluation skipped energies
      md->pressure = md->virial = md->kin_com = std::nan("");
      E.pressure = md->pressure;
    }
    E.virial = md->virial;
    E.kinetic_com = md->kin_com;
    for (int a = 0; a < 3; ++a) E.box[a] = md->box[a];
    E.volume = (double)md->box[0] * md->box[1] * md->box[2];
    E.pcouple_steps = md->n_pcouple;
    E.potential_bonded = E.bond + E.angle + E.dihedral;
    {
      dyn_energies EE{};
      const int rc = dyn_engine_get_energies(md->eng, &EE);
      if (rc == DYN_OK) {
        E.potential_nonbonded = md->eng_mode_reference ? EE.reference : EE.total;
        E.dh_dl = EE.dh_dl;
      } else if (rc == DYN_ERR_STATE) {  // the last evaluation skipped energies
        E.potential_nonbonded = E.dh_dl = std::nan("");
      } else {
        engine_call(rc, md->eng);
      }
    }
    E.potential = E.potential_nonbonded + E.potential_bonded;
    E.density = md->mass_total / E.volume;
    E.repeats = md->n_repeats;
    E.graph_steps = md->graph_steps;
    *out = E;
  });
}

// ---------------------------------------------------------------------------
// Batch of independent systems: one launch per kernel for all of them
// ---------------------------------------------------------------------------
struct dyn_batch {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<dyn_md*> m;
  hipStream_t side_stream = nullptr;  // second stream (SPME, bonded), same priority
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  ArgCache eng_cache, md_cache, eng_side_cache, md_side_cache;
  Profiler prof;
  StepGraphs graphs;

  MdSet set() {
    MdSet S;
    const uint32_t n = (uint32_t)m.size();
    S.m = m;
    for (dyn_md* x : m) S.es.e.push_back(x->eng);
    S.es.R = BatchRun{stream, &eng_cache, n};
    S.es.prof = &prof;
    S.R = BatchRun{stream, &md_cache, n};
    // fully profiled steps stay on one stream (their per-kernel times), and
    // so do the steps of members that ask for one stream
    if ((!prof.enabled || prof.only) && !engine_single_stream(m[0]->eng)) {
      S.es.side = BatchRun{side_stream, &eng_side_cache, n};
      S.es.ev_fork = ev_fork;
      S.es.ev_join = ev_join;
      S.side = BatchRun{side_stream, &md_side_cache, n};
    }
    S.prof = &prof;
    S.graphs = &graphs;
    return S;
  }
  // hand the members back to their own streams
  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (dyn_md* x : m) {
      engine_set_stream(x->eng, nullptr);
      x->stream = engine_stream(x->eng);
      x->in_batch = false;
    }
    m.clear();
    if (side_stream) (void)hipStreamSynchronize(side_stream);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side_stream) (void)hipStreamDestroy(side_stream);
    if (stream) (void)hipStreamDestroy(stream);
    stream = side_stream = nullptr;
    ev_fork = ev_join = nullptr;
  }
};

extern "C" int dyn_batch_create(dyn_batch** out, uint32_t n, dyn_md* const* mds) {
  try {
    DYN_REQUIRE(out && mds && n > 0, "null argument or empty batch");
    *out = nullptr;
    std::unique_ptr<dyn_batch> B(new dyn_batch());
    B->device = mds[0] ? mds[0]->device : 0;
    for (uint32_t i = 0; i < n; ++i) {
      dyn_md* x = mds[i];
      DYN_REQUIRE(x, "MD handle %u is null", i);
      DYN_REQUIRE(!x->in_batch, "MD handle %u is in another batch", i);
      DYN_REQUIRE(x->device == B->device, "MD handle %u is on device %d, not %d", i, x->device,
                  B->device);
      DYN_REQUIRE((x->prm.gamma > 0.f) == (mds[0]->prm.gamma > 0.f),
                  "MD handle %u: Langevin and plain velocity Verlet cannot share a batch", i);
      DYN_REQUIRE(x->prm.nstpcouple == mds[0]->prm.nstpcouple,
                  "MD handle %u: members of a batch share nstpcouple", i);
      DYN_REQUIRE((x->prm.staged_builds != 0) == (mds[0]->prm.staged_builds != 0),
                  "MD handle %u: members of a batch share staged_builds", i);
      DYN_REQUIRE(!x->prm.nstpcouple || x->step == mds[0]->step, "MD handle %u: members of a "
                  "batch with pressure coupling share the step count (%llu vs %llu)", i,
                  (unsigned long long)x->step, (unsigned long long)mds[0]->step);
      for (uint32_t j = 0; j < i; ++j)
        DYN_REQUIRE(mds[j] != x && mds[j]->eng != x->eng, "MD handles %u and %u share an engine",
                    j, i);
    }
    DeviceGuard g(B->device);
    for (uint32_t i = 0; i < n; ++i) DYN_HIP(hipStreamSynchronize(mds[i]->stream));
    DYN_HIP(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    DYN_HIP(hipStreamCreateWithFlags(&B->side_stream, hipStreamNonBlocking));
    DYN_HIP(hipEventCreateWithFlags(&B->ev_fork, hipEventDisableTiming));
    DYN_HIP(hipEventCreateWithFlags(&B->ev_join, hipEventDisableTiming));
    B->prof.stream = B->stream;
    for (uint32_t i = 0; i < n; ++i) {
      dyn_md* x = mds[i];
      engine_set_stream(x->eng, B->stream);
      x->stream = B->stream;
      x->own_cache.invalidate();
      x->in_batch = true;
      B->m.push_back(x);
    }
    *out = B.release();
    return DYN_OK;
  } catch (const Error& e) {
    set_global_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_global_error(e.what());
    return DYN_ERR_HIP;
  }
}

extern "C" void dyn_batch_destroy(dyn_batch* b) {
  if (!b) return;
  try {
    DeviceGuard g(b->device);
    b->release();
    delete b;
  } catch (...) {
  }
}

extern "C" const char* dyn_batch_last_error(const dyn_batch* b) {
  return b ? b->err.c_str() : "null batch handle";
}

extern "C" int dyn_batch_step(dyn_batch* b, uint32_t n_steps) {
  if (!b) {
    set_global_error("null batch handle");
    return DYN_ERR_ARG;
  }
  DYN_API_BODY(b, {
    DeviceGuard g(b->device);
    for (size_t i = 0; i < b->m.size(); ++i)
      if (!b->m[i]->state_set) throw Error(DYN_ERR_STATE, fmt("MD %zu: state not set", i));
    MdSet S = b->set();
    md_run(S, n_steps);
  });
}

extern "C" int dyn_batch_profile(dyn_batch* b, int enable) {
  if (!b) return DYN_ERR_ARG;
  DYN_API_BODY(b, {
    DeviceGuard g(b->device);
    b->prof.reset();
    const bool on = enable != 0;
    const char* only = enable == 2 ? "pairs" : nullptr;
    if (on != b->prof.enabled || only != b->prof.only) g_layout_gen.fetch_add(1);
    b->prof.enabled = on;
    b->prof.only = only;
  });
}

extern "C" int dyn_batch_profile_read(dyn_batch* b, uint32_t max, const char** names, double* ms,
                                      uint64_t* counts, uint32_t* n) {
  if (!b) return DYN_ERR_ARG;
  DYN_API_BODY(b, {
    DeviceGuard g(b->device);
    b->prof.drain();
    const uint32_t k = (uint32_t)b->prof.order.size();
    if (n) *n = k;
    for (uint32_t i = 0; i < k && i < max; ++i) {
      const auto& nm = b->prof.order[i];
      auto it = b->prof.totals.find(nm);
      if (names) names[i] = it->first.c_str();
      if (ms) ms[i] = it->second.first;
      if (counts) counts[i] = it->second.second;
    }
  });
}
