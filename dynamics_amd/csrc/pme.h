// pme.h — SPME reciprocal-space core shared by the dyn_pme_* entry points
// and the native engine.  Smooth PME (Essmann et al., J. Chem. Phys. 103,
// 8577 (1995)) with cardinal B-splines of order 4 or 5:
//   Q(k)   = Σ_i q_i Π_a M_n(u_ia − k_a),       u_ia = K_a x_ia / L_a
//   E_rec  = ½ Σ_{m≠0} G(m) |F[Q](m)|²,
//   G(m)   = exp(−π² |m|²/β²) / (π V |m|²) · Π_a |b_a(m_a)|²
//   φ      = F⁻¹[G · F[Q]]  (unnormalised transforms), F_i = −Σ_k ∂Q(k)/∂r_i · φ(k)
//   W_rec  = Σ_m ½ G(m)|F[Q](m)|² (1 − 2π²|m|²/β²)   (trace of the k-space virial)
// Charge spreading accumulates in int64 fixed point (2^-32 units), so the
// grid is independent of the order of the atomic adds.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "batch.h"
#include "common.h"
#include "wave.h"

namespace dyn {

constexpr int kBrick = 8;  // grid points per brick edge

// Fractional coordinate → base index (mod K) and in-cell weight.
template <class R>
__device__ __forceinline__ void frac_coord(R x, R sc, int K, int& k0, R& w) {
  const R u = x * sc;
  const R fl = floor(u);
  w = u - fl;
  int k = (int)fl % K;
  if (k < 0) k += K;
  k0 = k;
}

__device__ __forceinline__ uint32_t brick_of(const int k0[3], const uint32_t nbk[3]) {
  return ((uint32_t)(k0[0] / kBrick) * nbk[1] + (uint32_t)(k0[1] / kBrick)) * nbk[2] +
         (uint32_t)(k0[2] / kBrick);
}

// What listing charges under their home bricks needs: grid size and K/L
// per axis, bricks per axis, the per-brick counts and fixed-capacity lists,
// and where an overflow is flagged.
struct BinArgs {
  uint32_t* count;
  uint32_t* items;
  unsigned int* status;
  int K[3];
  float sc[3];
  uint32_t nb[3];
  uint32_t cap;
  unsigned int status_bit;
  uint32_t unused;  // no implicit padding: argument packs are compared byte for byte
};
static_assert(sizeof(BinArgs) == 72, "BinArgs has no padding");

// Lists charge i (position p) under its home brick.  Every lane of the wave
// must call it (valid marks the lanes with a charge): one atomic per
// distinct brick of the wave.
__device__ __forceinline__ void pme_bin_charge(const BinArgs& B, float4 p, bool valid,
                                               uint32_t i) {
  uint32_t b = 0xFFFFFFFFu;  // no brick
  if (valid) {
    const float xs[3] = {p.x, p.y, p.z};
    int k0[3];
    for (int a = 0; a < 3; ++a) {
      float w;
      frac_coord<float>(xs[a], B.sc[a], B.K[a], k0[a], w);
    }
    b = brick_of(k0, B.nb);
  }
  const uint32_t rank = wave_rank(b, valid, B.count);
  if (valid) {
    if (rank < B.cap) B.items[(size_t)b * B.cap + rank] = i;
    else atomicOr(B.status, B.status_bit);
  }
}

struct FftAxis {
  int N;
  int nstages;
  int radix[16];
};

FftAxis make_fft_axis(int N);  // throws if N has a prime factor > 5
bool is_235_smooth(uint32_t n);

// Position/charge input for spreading and gathering: float4 (x, y, z, q).
// `site` maps each charge to the fixed-point force row it adds into (engine
// use); null when forces are written to a float array instead.
struct PmeCore {
  int device = 0;
  hipStream_t stream = nullptr;
  Profiler* prof = nullptr;
  uint32_t K[3] = {0, 0, 0};
  float L[3] = {0, 0, 0};
  float beta = 0.f;
  int order = 4;
  size_t ngrid = 0;
  FftAxis ax[3];
  DevBuf<FftAxis> ax_dev;       // device copy of ax
  bool f64 = false;             // FP64 option: the *64 buffers replace their f32 twins
  DevBuf<int64_t> grid_fx;      // K0·K1·K2 fixed-point charge grid
  DevBuf<float2> cgrid;         // complex working grid
  DevBuf<float> G;              // influence function
  DevBuf<double> bsq;           // 1/|b_a(m)|² per axis (K0 + K1 + K2 values)
  DevBuf<float2> tw[3];         // e^{-2πi k/N} tables per axis
  DevBuf<double2> cgrid64;
  DevBuf<double> G64;
  DevBuf<double2> tw64[3];
  DevBuf<int64_t> partials;     // energy/virial partials of the convolution pass (2^-32 fixed point)
  PinnedBuf<double> h_part;
  DevBuf<double> res_dev;       // [energy, virial] reduced on the device
  // Spreading by grid brick (8³ points): each charge is listed under its
  // home brick (the brick of its base grid point) in a fixed-capacity list,
  // and each brick gathers the charges of the home bricks whose stencils can
  // reach it.
  uint32_t nbk[3] = {0, 0, 0};  // bricks per axis
  DevBuf<uint32_t> brick_count;  // per brick: charges listed (zeroed after use)
  DevBuf<uint32_t> brick_items;  // cap_brick slots per brick
  uint32_t cap_brick = 0;
  DevBuf<unsigned int> status;   // standalone runs: 1 = a brick list overflowed
  void ensure_cap(uint32_t n);
  void grow();
  // the binning arguments of the current grid and lists (ensure_cap first)
  BinArgs bin_args(unsigned int* status, unsigned int status_bit) const;

  void init(int device, hipStream_t s, Profiler* p, const uint32_t k[3], const float l[3],
            float beta, int order, bool fp64 = false);
  void set_box(const float l[3]);  // recomputes G (on the device, asynchronously)
  void set_moduli();                // bsq for the current K and order
  // Enqueue spread → FFT → convolve → inverse FFT → gather for n charges.
  // xyzq: device float4 array (positions wrapped or not; wrapped internally).
  // If site != null, forces are atomically added (fixed point, 2^32) into
  // fx_sites[3·site[i] + c] (component-interleaved); else written as floats
  // into f_out[3i + c].  Energy and virial (without self energy) end up in
  // res_dev[0], res_dev[1] after the final reduction kernel (or in res[0],
  // res[1] when res is given).  Forces are multiplied by fscale (exactly 1
  // leaves them unchanged).
  void run(uint32_t n, const float4* xyzq, const uint32_t* site, int64_t* fx_sites,
           int fx_component_major, uint32_t fx_nsites, float* f_out, float fscale = 1.0f,
           double* res = nullptr);
  ArgCache cache;  // device argument packs of run()
};

// One system's part of a batched SPME pass (the arguments of PmeCore::run;
// core == nullptr: the system takes no part).
struct PmeRun {
  PmeCore* core = nullptr;
  uint32_t n = 0;
  const float4* xyzq = nullptr;
  const uint32_t* site = nullptr;
  int64_t* fx = nullptr;
  int comp_major = 0;  // bit 0: component-major fx; bit 1: each site has one charge
                       // (plain adds instead of atomics)
  uint32_t nsites = 0;
  float* f_out = nullptr;
  float fscale = 1.0f;
  double* res = nullptr;  // [energy, virial]
  unsigned int* status = nullptr;  // a full brick list sets status_bit here (null: the core's)
  unsigned int status_bit = 0;
  bool binned = false;  // the charges are already listed (bin_args, after ensure_cap(n))
};

// runs[0 .. R.n): one per system of the pass; tag separates passes that run
// more than once per evaluation (their launch sites keep separate packs).
void pme_run(const BatchRun& R, Profiler& prof, const PmeRun* runs, int tag);

}  // namespace dyn
