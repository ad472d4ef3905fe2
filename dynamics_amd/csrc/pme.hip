// pme.hip — SPME reciprocal space on gfx950: fixed-point B-spline charge
// spreading, a mixed-radix (2/3/4/5) Stockham FFT staged through LDS, the
// influence-function convolution fused into the x-axis pass with FP64 energy
// and virial reductions, and the force gather.  Replaces the `ewald` crate's
// PmeRecip (used at src/non_bonded.rs:681-795, :850-875).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "common.h"
#include "wave.h"
#include "dyn_math.h"
#include "pme.h"

#pragma clang fp contract(off)

namespace dyn {

bool is_235_smooth(uint32_t n) {
  if (n == 0) return false;
  for (uint32_t p : {2u, 3u, 5u})
    while (n % p == 0) n /= p;
  return n == 1;
}

FftAxis make_fft_axis(int N) {
  FftAxis a;
  a.N = N;
  a.nstages = 0;
  int n = N;
  // radix-4 first, then 2, 3, 5
  while (n % 4 == 0) {
    a.radix[a.nstages++] = 4;
    n /= 4;
  }
  while (n % 2 == 0) {
    a.radix[a.nstages++] = 2;
    n /= 2;
  }
  while (n % 3 == 0) {
    a.radix[a.nstages++] = 3;
    n /= 3;
  }
  while (n % 5 == 0) {
    a.radix[a.nstages++] = 5;
    n /= 5;
  }
  if (n != 1) throw Error(DYN_ERR_ARG, fmt("grid size %d has a prime factor above 5", N));
  return a;
}

}  // namespace dyn

using namespace dyn;

namespace {

constexpr int kFftBlock = 256;
constexpr int kFftCap = 2048;  // complex elements per workgroup (two LDS buffers: 32 KB)

// Complex arithmetic on float2 (the default SPME) or double2 (the FP64
// option).
template <class C>
struct CxT;
template <>
struct CxT<float2> {
  using R = float;
  __device__ static float2 mk(float x, float y) { return make_float2(x, y); }
  static constexpr float c3 = -0.5f, s3 = 0.866025403784438647f;
  static constexpr float c51 = 0.309016994374947424f, c52 = -0.809016994374947424f;
  static constexpr float s51 = 0.951056516295153572f, s52 = 0.587785252292473129f;
};
template <>
struct CxT<double2> {
  using R = double;
  __device__ static double2 mk(double x, double y) { return make_double2(x, y); }
  static constexpr double c3 = -0.5, s3 = 0.866025403784438647;
  static constexpr double c51 = 0.309016994374947424, c52 = -0.809016994374947424;
  static constexpr double s51 = 0.951056516295153572, s52 = 0.587785252292473129;
};

template <class C>
__device__ __forceinline__ C cadd(C a, C b) { return CxT<C>::mk(a.x + b.x, a.y + b.y); }
template <class C>
__device__ __forceinline__ C csub(C a, C b) { return CxT<C>::mk(a.x - b.x, a.y - b.y); }
template <class C>
__device__ __forceinline__ C cscale(C a, typename CxT<C>::R s) { return CxT<C>::mk(a.x * s, a.y * s); }
// multiply by SIGN·i
template <int SIGN, class C>
__device__ __forceinline__ C cmul_si(C a) {
  return SIGN < 0 ? CxT<C>::mk(a.y, -a.x) : CxT<C>::mk(-a.y, a.x);
}
// v · w where w = tw (SIGN = −1) or conj(tw) (SIGN = +1)
template <int SIGN, class C>
__device__ __forceinline__ C ctw(C v, C t) {
  const typename CxT<C>::R ti = SIGN < 0 ? t.y : -t.y;
  return CxT<C>::mk(v.x * t.x - v.y * ti, v.x * ti + v.y * t.x);
}

template <int SIGN, class C>
__device__ __forceinline__ void dft2(C* v) {
  const C a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <int SIGN, class C>
__device__ __forceinline__ void dft4(C* v) {
  const C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
  const C s13 = cadd(v[1], v[3]), d13 = cmul_si<SIGN>(csub(v[1], v[3]));
  v[0] = cadd(s02, s13);
  v[2] = csub(s02, s13);
  v[1] = cadd(d02, d13);
  v[3] = csub(d02, d13);
}
template <int SIGN, class C>
__device__ __forceinline__ void dft3(C* v) {
  const C t = cadd(v[1], v[2]);
  const C d = cscale(cmul_si<SIGN>(csub(v[1], v[2])), CxT<C>::s3);
  const C a = cadd(v[0], cscale(t, CxT<C>::c3));
  v[0] = cadd(v[0], t);
  v[1] = cadd(a, d);
  v[2] = csub(a, d);
}
template <int SIGN, class C>
__device__ __forceinline__ void dft5(C* v) {
  const auto c1 = CxT<C>::c51, c2 = CxT<C>::c52, s1 = CxT<C>::s51, s2 = CxT<C>::s52;
  const C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const C d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
  const C a1 = cadd(v[0], cadd(cscale(t1, c1), cscale(t2, c2)));
  const C a2 = cadd(v[0], cadd(cscale(t1, c2), cscale(t2, c1)));
  const C b1 = cmul_si<SIGN>(cadd(cscale(d1, s1), cscale(d2, s2)));
  const C b2 = cmul_si<SIGN>(csub(cscale(d1, s2), cscale(d2, s1)));
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

// One radix-R Stockham stage with compile-time sizes (line length N,
// sub-transform length NS): the index arithmetic folds to shifts and masks.
template <int SIGN, int N, int R, int NS, class C>
__device__ __forceinline__ void fft_stage(const C* __restrict__ a, C* __restrict__ b, int nl,
                                          const C* __restrict__ tw) {
  constexpr int NR = N / R, TWS = N / (NS * R);
  const int total = nl * NR;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const int l = i / NR;
    const int j = i - l * NR;
    const int k = j % NS;
    const C* src = a + l * N;
    C* dst = b + l * N;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      C x = src[j + r * NR];
      if (r) x = ctw<SIGN>(x, tw[(r * k * TWS) % N]);
      v[r] = x;
    }
    if constexpr (R == 4) dft4<SIGN>(v);
    else dft2<SIGN>(v);
    const int d = (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[d + r * NS] = v[r];
  }
  __syncthreads();
}

// x / d for 0 ≤ x < 2²² by a float reciprocal and one correction step.
__device__ __forceinline__ int fast_div(int x, int d, float rd) {
  int q = (int)((float)x * rd);
  const int rem = x - q * d;
  if (rem < 0) --q;
  else if (rem >= d) ++q;
  return q;
}

// One radix-R Stockham stage with the line length and sub-transform length
// known at run time (the general grids): the radix is a template parameter so
// the butterfly stays in registers, and the divisions use fast_div.  The
// twiddle index r·k·(N/(Ns·R)) stays below N.
template <int SIGN, int R, class C>
__device__ __forceinline__ void fft_stage_rt(const C* __restrict__ a, C* __restrict__ b, int nl,
                                             int N, int Ns, const C* __restrict__ tw) {
  const int NR = N / R, tws = N / (Ns * R);
  const float rNR = 1.0f / (float)NR, rNs = 1.0f / (float)Ns;
  const int total = nl * NR;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const int l = fast_div(i, NR, rNR);
    const int j = i - l * NR;
    const int q = fast_div(j, Ns, rNs);
    const int k = j - q * Ns;
    const C* src = a + l * N;
    C* dst = b + l * N;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      C x = src[j + r * NR];
      if (r) x = ctw<SIGN>(x, tw[r * k * tws]);
      v[r] = x;
    }
    if constexpr (R == 4) dft4<SIGN>(v);
    else if constexpr (R == 2) dft2<SIGN>(v);
    else if constexpr (R == 3) dft3<SIGN>(v);
    else dft5<SIGN>(v);
    const int d = q * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[d + r * Ns] = v[r];
  }
  __syncthreads();
}

// The same transform as lds_fft for the power-of-two lengths of common grids
// (radix 4 first, then 2, the plan of make_fft_axis).
template <int SIGN, int N, class C>
__device__ __forceinline__ C* lds_fft_pow2(C* a, C* b, int nl, const C* __restrict__ tw) {
  static_assert(N == 32 || N == 64 || N == 128, "fixed plans for 32, 64 and 128 points");
  fft_stage<SIGN, N, 4, 1>(a, b, nl, tw);
  fft_stage<SIGN, N, 4, 4>(b, a, nl, tw);
  if constexpr (N == 32) {
    fft_stage<SIGN, N, 2, 16>(a, b, nl, tw);
    return b;
  } else {
    fft_stage<SIGN, N, 4, 16>(a, b, nl, tw);
    if constexpr (N == 64) return b;
    fft_stage<SIGN, N, 2, 64>(b, a, nl, tw);
    return a;
  }
}

// Stockham autosort FFT of nl lines of length N held in LDS as a[l·N + n].
// Returns the buffer holding the (natural-order) result.  NF: a line length
// every system of the launch shares (32, 64 or 128: the fixed plan), or 0.
template <int SIGN, int NF, class C>
__device__ C* lds_fft(C* a, C* b, int nl, const FftAxis& ax, const C* __restrict__ tw) {
  if constexpr (NF != 0) return lds_fft_pow2<SIGN, NF>(a, b, nl, tw);
  const int N = ax.N;
  int Ns = 1;
  for (int s = 0; s < ax.nstages; ++s) {
    const int R = ax.radix[s];
    if (R == 4) fft_stage_rt<SIGN, 4>(a, b, nl, N, Ns, tw);
    else if (R == 2) fft_stage_rt<SIGN, 2>(a, b, nl, N, Ns, tw);
    else if (R == 3) fft_stage_rt<SIGN, 3>(a, b, nl, N, Ns, tw);
    else fft_stage_rt<SIGN, 5>(a, b, nl, N, Ns, tw);
    C* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// The stage of fft_stage on nl lines stored interleaved (element n of line l
// at a[n·NL + l]): consecutive threads take consecutive lines, so a wave's
// LDS accesses are contiguous.  Every line sees the arithmetic of fft_stage.
template <int SIGN, int N, int R, int NS, int NL, class C>
__device__ __forceinline__ void fft_stage_t(const C* __restrict__ a, C* __restrict__ b,
                                            const C* __restrict__ tw) {
  constexpr int NR = N / R, TWS = N / (NS * R);
  for (int i = threadIdx.x; i < NL * NR; i += blockDim.x) {
    const int j = i / NL;
    const int l = i - j * NL;
    const int k = j % NS;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      C x = a[(j + r * NR) * NL + l];
      if (r) x = ctw<SIGN>(x, tw[(r * k * TWS) % N]);
      v[r] = x;
    }
    if constexpr (R == 4) dft4<SIGN>(v);
    else dft2<SIGN>(v);
    const int d = (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) b[(d + r * NS) * NL + l] = v[r];
  }
  __syncthreads();
}

// lds_fft_pow2 on NL interleaved lines.
template <int SIGN, int N, int NL, class C>
__device__ __forceinline__ C* lds_fft_pow2_t(C* a, C* b, const C* __restrict__ tw) {
  static_assert(N == 32 || N == 64, "plane passes for 32 and 64 points");
  fft_stage_t<SIGN, N, 4, 1, NL>(a, b, tw);
  fft_stage_t<SIGN, N, 4, 4, NL>(b, a, tw);
  if constexpr (N == 32) {
    fft_stage_t<SIGN, N, 2, 16, NL>(a, b, tw);
  } else {
    fft_stage_t<SIGN, N, 4, 16, NL>(a, b, tw);
  }
  return b;
}

struct PassGeom {
  int K[3];
  int lines;      // total lines in this pass
  int L;          // lines per workgroup
};

// Address of element n of line `line` for the axis pass AXIS (0 = x, 1 = y, 2 = z).
template <int AXIS>
__device__ __forceinline__ size_t pass_addr(const PassGeom& g, int line, int n) {
  if (AXIS == 2) return (size_t)line * g.K[2] + n;
  if (AXIS == 1) {
    const int x = line / g.K[2], z = line - x * g.K[2];
    return ((size_t)x * g.K[1] + n) * g.K[2] + z;
  }
  const int y = line / g.K[2], z = line - y * g.K[2];
  return ((size_t)n * g.K[1] + y) * g.K[2] + z;
}

// LDS element e of a group ↔ (line l, point n).  z lines are contiguous in
// memory so n runs fastest; for y and x lines, consecutive lines are
// consecutive in z, so l runs fastest and loads stay coalesced.
template <int AXIS>
__device__ __forceinline__ void split_e(int e, int L, int N, int& l, int& n) {
  if (AXIS == 2) {
    l = e / N;
    n = e - l * N;
  } else {
    n = e / L;
    l = e - n * L;
  }
}

// One FFT pass along AXIS over complex data in place.  FROM_FX: the input is
// the int64 fixed-point charge grid (z pass only).
template <int AXIS, int SIGN, bool FROM_FX, class C, int NF>
__device__ __forceinline__ void
k_fft_pass(C* __restrict__ data, const int64_t* __restrict__ grid_fx, PassGeom g,
           const FftAxis* __restrict__ axp, const C* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* smem = reinterpret_cast<C*>(smem_raw);
  using R = typename CxT<C>::R;
  const FftAxis& ax = *axp;  // device copy: uniform loads, no private copy
  const int N = ax.N;
  const int line0 = blockIdx.x * g.L;
  if (line0 >= g.lines) return;  // beyond this system's lines
  const int nl = min(g.L, g.lines - line0);
  C* A = smem;
  C* B = smem + g.L * N;
  const int tot = nl * N;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int l, n;
    split_e<AXIS>(e, nl, N, l, n);
    const size_t addr = pass_addr<AXIS>(g, line0 + l, n);
    C v;
    if (FROM_FX) v = CxT<C>::mk((R)((double)grid_fx[addr] * (1.0 / DYN_GRID_FX_SCALE)), R(0));
    else v = data[addr];
    A[l * N + n] = v;
  }
  __syncthreads();
  C* Rs = lds_fft<SIGN, NF>(A, B, nl, ax, tw);
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int l, n;
    split_e<AXIS>(e, nl, N, l, n);
    data[pass_addr<AXIS>(g, line0 + l, n)] = Rs[l * N + n];
  }
}

// The z and y passes of one direction in one launch (single precision,
// fixed plans, NY·NZ ≤ kPlaneCap): a workgroup holds the x plane of its
// block index in LDS and transforms its z lines then its y lines (forward,
// from the fixed-point charge grid) or its y lines then its z lines
// (inverse).  Every line sees the stages, twiddles and order of the
// separate passes, so the result is the same to the bit.
constexpr int kPlaneCap = 4096;
constexpr int kPlaneBlock = 512;
template <int SIGN, bool FROM_FX, int NY, int NZ>
__device__ __forceinline__ void k_fft_plane(float2* __restrict__ data,
                                            const int64_t* __restrict__ grid_fx, int nx,
                                            const float2* __restrict__ twy,
                                            const float2* __restrict__ twz) {
  static_assert(NY * NZ <= kPlaneCap, "plane exceeds the LDS buffers");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float2* A = reinterpret_cast<float2*>(smem_raw);
  float2* B = A + NY * NZ;
  const int x = blockIdx.x;
  if (x >= nx) return;  // beyond this system's planes
  const size_t base = (size_t)x * NY * NZ;
  for (int e = threadIdx.x; e < NY * NZ; e += blockDim.x) {
    if (FROM_FX) A[e] = make_float2((float)((double)grid_fx[base + e] * (1.0 / DYN_GRID_FX_SCALE)), 0.0f);
    else A[e] = data[base + e];
  }
  __syncthreads();
  float2* out;
  if (SIGN < 0) {
    float2* rz = lds_fft_pow2<SIGN, NZ>(A, B, NY, twz);  // NY contiguous z lines
    out = lds_fft_pow2_t<SIGN, NY, NZ>(rz, rz == A ? B : A, twy);
  } else {
    float2* ry = lds_fft_pow2_t<SIGN, NY, NZ>(A, B, twy);
    out = lds_fft_pow2<SIGN, NZ>(ry, ry == A ? B : A, NY, twz);
  }
  for (int e = threadIdx.x; e < NY * NZ; e += blockDim.x) data[base + e] = out[e];
}

struct ConvArgs {
  float invL[3];
  double pi2_over_b2;  // π²/β²
};

// x pass fused with the convolution: forward FFT along x, multiply by G(m)
// (accumulating ½G|F[Q]|² and its virial weight in f64), inverse FFT along x.
template <class CX, int NF>
__device__ __forceinline__ void
k_fft_conv_x(CX* __restrict__ data, PassGeom g, const FftAxis* __restrict__ axp,
             const CX* __restrict__ tw, const typename CxT<CX>::R* __restrict__ G, ConvArgs C,
             int64_t* __restrict__ partials) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  CX* smem = reinterpret_cast<CX*>(smem_raw);
  const FftAxis& ax = *axp;
  const int N = ax.N;
  const int line0 = blockIdx.x * g.L;
  if (line0 >= g.lines) return;  // beyond this system's lines
  const int nl = min(g.L, g.lines - line0);
  CX* A = smem;
  CX* B = smem + g.L * N;
  const int tot = nl * N;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int l, n;
    split_e<0>(e, nl, N, l, n);
    A[l * N + n] = data[pass_addr<0>(g, line0 + l, n)];
  }
  __syncthreads();
  CX* R = lds_fft<-1, NF>(A, B, nl, ax, tw);
  CX* O = (R == A) ? B : A;
  // the influence function of the block's points into the free LDS buffer
  // (coalesced reads), for the energy sums and the product
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int l, n;
    split_e<0>(e, nl, N, l, n);
    O[l * N + n].x = G[pass_addr<0>(g, line0 + l, n)];
  }
  __syncthreads();
  // energy and virial weight per line (a wave per line, lanes over its
  // points, shuffle sums), each line's pair rounded to 2^-32 fixed point and
  // summed as integers: the workgroup's partial does not depend on how many
  // lines it holds, so the total is the same for any grouping of the lines
  {
    __shared__ long long red[2][kFftBlock / 64];
    long long qe = 0, qw = 0;
    const int wv = threadIdx.x / 64, lane = threadIdx.x & 63;
    for (int l = wv; l < nl; l += kFftBlock / 64) {
      const int line = line0 + l;
      const int my = line / g.K[2], mz = line - my * g.K[2];
      const int myp = my <= g.K[1] / 2 ? my : my - g.K[1];
      const int mzp = mz <= g.K[2] / 2 ? mz : mz - g.K[2];
      const double fy = myp * (double)C.invL[1], fz = mzp * (double)C.invL[2];
      double e_acc = 0.0, w_acc = 0.0;
      for (int n = lane; n < N; n += 64) {
        const auto gv = O[l * N + n].x;
        const CX v = R[l * N + n];
        const double en = 0.5 * (double)gv * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
        const int mxp = n <= g.K[0] / 2 ? n : n - g.K[0];
        const double fx = mxp * (double)C.invL[0];
        const double msq = fx * fx + fy * fy + fz * fz;
        e_acc += en;
        w_acc += en * (1.0 - 2.0 * C.pi2_over_b2 * msq);
      }
      for (int m = 32; m >= 1; m >>= 1) {
        e_acc += __shfl_xor(e_acc, m, 64);
        w_acc += __shfl_xor(w_acc, m, 64);
      }
      qe += (long long)dyn_fx64(e_acc);
      qw += (long long)dyn_fx64(w_acc);
    }
    if (lane == 0) {
      red[0][wv] = qe;
      red[1][wv] = qw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      long long te = 0, tw2 = 0;
      for (int k = 0; k < kFftBlock / 64; ++k) {
        te += red[0][k];
        tw2 += red[1][k];
      }
      partials[2 * blockIdx.x] = te;
      partials[2 * blockIdx.x + 1] = tw2;
    }
  }
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int l, n;
    split_e<0>(e, nl, N, l, n);
    const auto gv = O[l * N + n].x;
    CX v = R[l * N + n];
    v.x *= gv;
    v.y *= gv;
    R[l * N + n] = v;
  }
  __syncthreads();
  CX* R2 = lds_fft<+1, NF>(R, O, nl, ax, tw);
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int l, n;
    split_e<0>(e, nl, N, l, n);
    data[pass_addr<0>(g, line0 + l, n)] = R2[l * N + n];
  }
}


struct SplineArgs {
  int K[3];
  float sc[3];    // K/L
  double scd[3];  // K/L in double (FP64 option)
};

constexpr int kBrickBlock = 256;

// Per charge: its home brick (the brick of its base grid point) and a rank
// in it; the charge goes to slot rank of the brick's fixed-capacity list
// (cap slots per brick).  A full brick sets *status |= status_bit (the caller
// enlarges the lists and evaluates again).  The rank order depends on atomic
// timing; the brick sums are integers, so the spread grid does not.
__device__ __forceinline__ void k_pme_bin(const float4* __restrict__ xyzq, uint32_t n, BinArgs B) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  pme_bin_charge(B, i < n ? xyzq[i] : make_float4(0.f, 0.f, 0.f, 0.f), i < n, i);
}

// Home bricks along one axis whose charges can reach the points [lo, hi]
// (lo a brick start): a charge with base point k0 covers k0 − ORDER + 1 … k0
// (mod K), so k0 runs over lo … hi + ORDER − 1 (mod K): the bricks of
// lo … e1 (e1 < K), then, past the wrap, those of 0 … e2 that come before
// lo's brick.  At most 3 bricks (ORDER ≤ 5, 8-point bricks).
struct HomeRange {
  int b0, n1, n;  // first brick, bricks before the wrap, bricks in all
  __device__ int brick(int j) const { return j < n1 ? b0 + j : j - n1; }
};
template <int ORDER>
__device__ __forceinline__ HomeRange home_bricks(int lo, int hi, int K) {
  const int len = min(hi - lo + ORDER, K);
  const int last = lo + len - 1;
  HomeRange r;
  r.b0 = lo / kBrick;
  r.n1 = min(last, K - 1) / kBrick - r.b0 + 1;
  const int e2 = last - K;  // last point past the wrap (< 0: none)
  r.n = r.n1 + (e2 >= 0 && r.b0 > 0 ? min(e2 / kBrick, r.b0 - 1) + 1 : 0);
  return r;
}

// One workgroup per brick: every charge of the reaching home bricks adds its
// stencil points inside the brick to an LDS brick (int64 LDS adds of the
// same fixed-point values the replay sums), which is then stored once.  The
// home bricks' charge lists are walked as one flattened index space.
template <int ORDER, class R>
__device__ __forceinline__ void
k_pme_spread(const float4* __restrict__ xyzq, uint32_t n, SplineArgs S, uint32_t nbx,
             uint32_t nby, uint32_t nbz, const uint32_t* __restrict__ count,
             const uint32_t* __restrict__ items, uint32_t cap, int64_t* __restrict__ grid) {
  __shared__ unsigned long long acc[kBrick * kBrick * kBrick];
  __shared__ uint32_t s_c0[27], s_beg[28];
  const uint32_t b = blockIdx.x;
  if (b >= nbx * nby * nbz) return;  // beyond this system's bricks
  const int bz = (int)(b % nbz), by = (int)((b / nbz) % nby), bx = (int)(b / (nbz * nby));
  const int lo[3] = {bx * kBrick, by * kBrick, bz * kBrick};
  const int hi[3] = {min(lo[0] + kBrick, S.K[0]) - 1, min(lo[1] + kBrick, S.K[1]) - 1,
                     min(lo[2] + kBrick, S.K[2]) - 1};
  for (int k = threadIdx.x; k < kBrick * kBrick * kBrick; k += kBrickBlock) acc[k] = 0ull;
  if (threadIdx.x == 0) {
    const HomeRange hx = home_bricks<ORDER>(lo[0], hi[0], S.K[0]);
    const HomeRange hy = home_bricks<ORDER>(lo[1], hi[1], S.K[1]);
    const HomeRange hz = home_bricks<ORDER>(lo[2], hi[2], S.K[2]);
    uint32_t tot = 0;
    int q = 0;
    for (int ix = 0; ix < hx.n; ++ix)
      for (int iy = 0; iy < hy.n; ++iy)
        for (int iz = 0; iz < hz.n; ++iz) {
          const uint32_t h = ((uint32_t)hx.brick(ix) * nby + (uint32_t)hy.brick(iy)) * nbz +
                             (uint32_t)hz.brick(iz);
          s_c0[q] = h * cap;
          s_beg[q] = tot;
          tot += min(count[h], cap);
          ++q;
        }
    for (; q < 28; ++q) s_beg[q] = tot;
  }
  __syncthreads();
  const uint32_t total = s_beg[27];
  for (uint32_t t = threadIdx.x; t < total; t += kBrickBlock) {
    int q = 0;
    while (q < 26 && s_beg[q + 1] <= t) ++q;
    const uint32_t ci = items[s_c0[q] + (t - s_beg[q])];
    if (ci >= n) continue;
    const float4 p = xyzq[ci];
    if (p.w == 0.0f) continue;
    int k0[3];
    R th[3][ORDER], dth[ORDER];
    const R xs[3] = {(R)p.x, (R)p.y, (R)p.z};
    for (int a = 0; a < 3; ++a) {
      R w;
      frac_coord<R>(xs[a], std::is_same<R, float>::value ? (R)S.sc[a] : (R)S.scd[a], S.K[a], k0[a], w);
      dyn_bspline<ORDER, R>(w, th[a], dth);
    }
    for (int jx = 0; jx < ORDER; ++jx) {
      const int gx = (k0[0] + S.K[0] - ORDER + 1 + jx) % S.K[0];
      if (gx < lo[0] || gx > hi[0]) continue;
      const R qx = (R)p.w * th[0][jx];
      for (int jy = 0; jy < ORDER; ++jy) {
        const int gy = (k0[1] + S.K[1] - ORDER + 1 + jy) % S.K[1];
        if (gy < lo[1] || gy > hi[1]) continue;
        const R qxy = qx * th[1][jy];
        for (int jz = 0; jz < ORDER; ++jz) {
          const int gz = (k0[2] + S.K[2] - ORDER + 1 + jz) % S.K[2];
          if (gz < lo[2] || gz > hi[2]) continue;
          const R v = qxy * th[2][jz];
          const int li = ((gx - lo[0]) * kBrick + (gy - lo[1])) * kBrick + (gz - lo[2]);
          atomicAdd(&acc[li], (unsigned long long)dyn_fx_d((double)v, DYN_GRID_FX_SCALE));
        }
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kBrick * kBrick * kBrick; k += kBrickBlock) {
    const int x = lo[0] + k / (kBrick * kBrick), y = lo[1] + (k / kBrick) % kBrick,
              z = lo[2] + k % kBrick;
    if (x <= hi[0] && y <= hi[1] && z <= hi[2])
      grid[((size_t)x * S.K[1] + y) * S.K[2] + z] = (int64_t)acc[k];
  }
}

template <class R>
__device__ __forceinline__ int64_t fx_of(R f) {
  if constexpr (std::is_same<R, float>::value) return dyn_fx(f);
  else return dyn_fx64(f);
}

constexpr int kGatherBlock = 128;
template <int ORDER, class C>
__device__ __forceinline__ void k_pme_gather(const float4* __restrict__ xyzq, uint32_t n, SplineArgs S,
                             const C* __restrict__ phi, const uint32_t* __restrict__ site,
                             int64_t* __restrict__ fx, int comp_major, uint32_t nsites,
                             float* __restrict__ fout, float fscale,
                             const int64_t* __restrict__ partials, int nbp,
                             double* __restrict__ res, uint32_t* __restrict__ brick_count,
                             uint32_t nbk) {
  // partials: the convolution's per-workgroup energy and virial sums in
  // 2^-32 fixed point (nbp workgroups)
  using R = typename CxT<C>::R;
  constexpr bool F32 = std::is_same<R, float>::value;
  // the spread has read the brick counts: clear them for the next binning
  if (brick_count)
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nbk; k += gridDim.x * blockDim.x)
      brick_count[k] = 0u;
  if (res && blockIdx.x == 0) {
    // the first workgroup also sums the convolution's energy and virial
    // partials (fixed order: strided per thread, then waves, then lanes)
    __shared__ long long red[2][kGatherBlock / 64];
    long long e = 0, w = 0;
    for (int b = threadIdx.x; b < nbp; b += kGatherBlock) {
      e += (long long)partials[2 * b];
      w += (long long)partials[2 * b + 1];
    }
    for (int m = 32; m >= 1; m >>= 1) {
      e += __shfl_xor(e, m, 64);
      w += __shfl_xor(w, m, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x / 64] = e;
      red[1][threadIdx.x / 64] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      long long te = 0, tw = 0;
      for (int k = 0; k < kGatherBlock / 64; ++k) {
        te += red[0][k];
        tw += red[1][k];
      }
      res[0] = (double)te * DYN_FX_INV_SCALE;
      res[1] = (double)tw * DYN_FX_INV_SCALE;
    }
  }
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = xyzq[i];
  R f0 = R(0), f1 = R(0), f2 = R(0);
  if (p.w != 0.0f) {
    int k0[3];
    R th[3][ORDER], dth[3][ORDER];
    const R xs[3] = {(R)p.x, (R)p.y, (R)p.z};
    R sc[3];
    for (int a = 0; a < 3; ++a) {
      sc[a] = F32 ? (R)S.sc[a] : (R)S.scd[a];
      R w;
      frac_coord<R>(xs[a], sc[a], S.K[a], k0[a], w);
      dyn_bspline<ORDER, R>(w, th[a], dth[a]);
    }
    R sx = R(0), sy = R(0), sz = R(0);
    for (int jx = 0; jx < ORDER; ++jx) {
      const int gx = (k0[0] + S.K[0] - ORDER + 1 + jx) % S.K[0];
      for (int jy = 0; jy < ORDER; ++jy) {
        const int gy = (k0[1] + S.K[1] - ORDER + 1 + jy) % S.K[1];
        const C* row = phi + ((size_t)gx * S.K[1] + gy) * S.K[2];
        const R txy = th[0][jx] * th[1][jy];
        const R dxy = dth[0][jx] * th[1][jy];
        const R tdy = th[0][jx] * dth[1][jy];
        for (int jz = 0; jz < ORDER; ++jz) {
          const int gz = (k0[2] + S.K[2] - ORDER + 1 + jz) % S.K[2];
          const R ph = row[gz].x;
          if constexpr (F32) {
            sx = dyn_fmaf(ph, dxy * th[2][jz], sx);
            sy = dyn_fmaf(ph, tdy * th[2][jz], sy);
            sz = dyn_fmaf(ph, txy * dth[2][jz], sz);
          } else {
            sx = fma(ph, dxy * th[2][jz], sx);
            sy = fma(ph, tdy * th[2][jz], sy);
            sz = fma(ph, txy * dth[2][jz], sz);
          }
        }
      }
    }
    f0 = -(R)p.w * sc[0] * sx * (R)fscale;
    f1 = -(R)p.w * sc[1] * sy * (R)fscale;
    f2 = -(R)p.w * sc[2] * sz * (R)fscale;
  }
  if (site) {
    const uint32_t s = site[i];
    if (comp_major & 2) {
      // every site has one charge: no other thread of this pass touches it
      // (a second pass over the same sites is a later launch)
      fx[s] += fx_of(f0);
      fx[nsites + s] += fx_of(f1);
      fx[2 * (size_t)nsites + s] += fx_of(f2);
    } else if (comp_major) {
      atomicAdd(reinterpret_cast<unsigned long long*>(fx + s), (unsigned long long)fx_of(f0));
      atomicAdd(reinterpret_cast<unsigned long long*>(fx + nsites + s), (unsigned long long)fx_of(f1));
      atomicAdd(reinterpret_cast<unsigned long long*>(fx + 2 * (size_t)nsites + s), (unsigned long long)fx_of(f2));
    } else {
      atomicAdd(reinterpret_cast<unsigned long long*>(fx + 3 * (size_t)s), (unsigned long long)fx_of(f0));
      atomicAdd(reinterpret_cast<unsigned long long*>(fx + 3 * (size_t)s + 1), (unsigned long long)fx_of(f1));
      atomicAdd(reinterpret_cast<unsigned long long*>(fx + 3 * (size_t)s + 2), (unsigned long long)fx_of(f2));
    }
  } else {
    fout[3 * i] = (float)f0;
    fout[3 * i + 1] = (float)f1;
    fout[3 * i + 2] = (float)f2;
  }
}

// Lines per workgroup of an FFT pass: about 512 points per workgroup (two
// per thread), at most kFftCap/N lines, and few enough lines that the pass
// over all nsys systems of a launch still has ≥ 512 workgroups.
PassGeom pass_geom(const uint32_t K[3], int axis, uint32_t nsys) {
  PassGeom g;
  for (int a = 0; a < 3; ++a) g.K[a] = (int)K[a];
  const int N = (int)K[axis];
  g.lines = (int)(K[0] * K[1] * K[2] / K[axis]);
  const int l_work = std::max(1, 512 / N);
  const int l_par = (int)std::max<int64_t>(1, (int64_t)g.lines * std::max<uint32_t>(nsys, 1) / 512);
  int L = std::max(1, std::min(kFftCap / N, std::min(l_work, l_par)));
  int p2 = 1;
  while (p2 * 2 <= L) p2 *= 2;
  g.L = p2;
  return g;
}

}  // namespace

namespace dyn {

void PmeCore::init(int dev, hipStream_t s, Profiler* p, const uint32_t k[3], const float l[3],
                   float b, int ord, bool fp64) {
  f64 = fp64;
  device = dev;
  stream = s;
  prof = p;
  for (int a = 0; a < 3; ++a) {
    DYN_REQUIRE(k[a] >= (uint32_t)ord && k[a] <= 1024, "grid size %u out of range [%d, 1024]",
                k[a], ord);
    DYN_REQUIRE(is_235_smooth(k[a]), "grid size %u must factor into 2, 3 and 5", k[a]);
    K[a] = k[a];
    ax[a] = make_fft_axis((int)k[a]);
  }
  ax_dev.reserve(3);
  DYN_HIP(hipMemcpyAsync(ax_dev.p, ax, sizeof ax, hipMemcpyHostToDevice, s));
  DYN_REQUIRE(ord == 4 || ord == 5, "B-spline order must be 4 or 5 (got %d)", ord);
  DYN_REQUIRE(b > 0.f && std::isfinite(b), "beta must be > 0");
  order = ord;
  beta = b;
  ngrid = (size_t)K[0] * K[1] * K[2];
  grid_fx.reserve(ngrid);
  for (int a = 0; a < 3; ++a) nbk[a] = (K[a] + kBrick - 1) / kBrick;
  const size_t nb = (size_t)nbk[0] * nbk[1] * nbk[2];
  brick_count.reserve(nb + 1);
  cap_brick = 0;
  status.reserve(1);
  // the binning pass counts into zeroed counts; the gather clears them again
  DYN_HIP(hipMemsetAsync(brick_count.p, 0, sizeof(uint32_t) * (nb + 1), stream));
  DYN_HIP(hipMemsetAsync(status.p, 0, sizeof(unsigned int), stream));
  if (f64) {
    cgrid64.reserve(ngrid);
    G64.reserve(ngrid);
  } else {
    cgrid.reserve(ngrid);
    G.reserve(ngrid);
  }
  for (int a = 0; a < 3; ++a) {
    std::vector<float2> t(K[a]);
    std::vector<double2> t64(K[a]);
    for (uint32_t m = 0; m < K[a]; ++m) {
      const double ang = -2.0 * M_PI * (double)m / (double)K[a];
      t[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
      t64[m] = make_double2(std::cos(ang), std::sin(ang));
    }
    if (f64) {
      tw64[a].reserve(K[a]);
      DYN_HIP(hipMemcpyAsync(tw64[a].p, t64.data(), sizeof(double2) * K[a], hipMemcpyHostToDevice,
                             stream));
    } else {
      tw[a].reserve(K[a]);
      DYN_HIP(hipMemcpyAsync(tw[a].p, t.data(), sizeof(float2) * K[a], hipMemcpyHostToDevice,
                             stream));
    }
    DYN_HIP(hipStreamSynchronize(stream));
  }
  // energy and virial partials: one fixed-point pair per x-pass workgroup
  // (at most one per line)
  partials.reserve(2 * (size_t)K[1] * K[2]);
  res_dev.reserve(2);
  set_moduli();
  set_box(l);
}

// Influence function at every grid point (set_box), in double:
// G(m) = exp(−π²|m|²/β²)/(π V |m|²) · Π_a |b_a(m_a)|², m_a = m'_a/L_a with
// m'_a the signed index; G(0) = 0.  bsq: 1/|b_a|² per axis, axes concatenated.
template <class R>
__global__ void __launch_bounds__(256)
k_pme_influence(R* __restrict__ G, const double* __restrict__ bsq, uint32_t K0, uint32_t K1,
                uint32_t K2, double il0, double il1, double il2, double inv_piv, double pi2_b2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)K0 * K1 * K2) return;
  const uint32_t mz = (uint32_t)(i % K2);
  const size_t r = i / K2;
  const uint32_t my = (uint32_t)(r % K1), mx = (uint32_t)(r / K1);
  const double fx = (mx <= K0 / 2 ? (double)mx : (double)mx - K0) * il0;
  const double fy = (my <= K1 / 2 ? (double)my : (double)my - K1) * il1;
  const double fz = (mz <= K2 / 2 ? (double)mz : (double)mz - K2) * il2;
  const double msq = fx * fx + fy * fy + fz * fz;
  double val = 0.0;
  if (msq > 0)
    val = exp(-pi2_b2 * msq) * inv_piv / msq * bsq[mx] * bsq[K0 + my] * bsq[K0 + K1 + mz];
  G[i] = (R)val;
}

// |b(m)|⁻² per axis (Essmann eq. 4.4) from the order-n cardinal B-spline at
// the integers; a vanishing modulus (odd order at m = K/2) takes the mean of
// its neighbours.
void PmeCore::set_moduli() {
  std::vector<double> Mn(order + 1, 0.0);
  if (order == 4) {
    Mn[1] = 1.0 / 6;
    Mn[2] = 4.0 / 6;
    Mn[3] = 1.0 / 6;
  } else {
    Mn[1] = 1.0 / 24;
    Mn[2] = 11.0 / 24;
    Mn[3] = 11.0 / 24;
    Mn[4] = 1.0 / 24;
  }
  std::vector<double> all;
  for (int a = 0; a < 3; ++a) {
    const int Ka = (int)K[a];
    std::vector<double> mod(Ka);
    for (int m = 0; m < Ka; ++m) {
      double re = 0, im = 0;
      for (int k = 0; k <= order - 2; ++k) {
        const double ang = 2.0 * M_PI * (double)m * k / Ka;
        re += Mn[k + 1] * std::cos(ang);
        im += Mn[k + 1] * std::sin(ang);
      }
      mod[m] = re * re + im * im;
    }
    for (int m = 0; m < Ka; ++m)
      if (mod[m] < 1e-7) mod[m] = 0.5 * (mod[(m - 1 + Ka) % Ka] + mod[(m + 1) % Ka]);
    for (int m = 0; m < Ka; ++m) all.push_back(1.0 / mod[m]);
  }
  bsq.reserve(all.size());
  DYN_HIP(hipMemcpyAsync(bsq.p, all.data(), sizeof(double) * all.size(), hipMemcpyHostToDevice,
                         stream));
  DYN_HIP(hipStreamSynchronize(stream));
}

// A new box: the influence function is recomputed on the device (the grid
// and its FFT plan stay; the reference rebuilds its PmeRecip instead,
// src/non_bonded.rs:850-875).
void PmeCore::set_box(const float l[3]) {
  for (int a = 0; a < 3; ++a) {
    DYN_REQUIRE(l[a] > 0.f && std::isfinite(l[a]), "box length must be > 0");
    L[a] = l[a];
  }
  const double V = (double)L[0] * L[1] * L[2];
  const double b2 = (double)beta * beta;
  const unsigned grid = (unsigned)((ngrid + 255) / 256);
  const double il0 = 1.0 / (double)L[0], il1 = 1.0 / (double)L[1], il2 = 1.0 / (double)L[2];
  if (f64)
    hipLaunchKernelGGL((k_pme_influence<double>), dim3(grid), dim3(256), 0, stream, G64.p, bsq.p,
                       K[0], K[1], K[2], il0, il1, il2, 1.0 / (M_PI * V), M_PI * M_PI / b2);
  else
    hipLaunchKernelGGL((k_pme_influence<float>), dim3(grid), dim3(256), 0, stream, G.p, bsq.p,
                       K[0], K[1], K[2], il0, il1, il2, 1.0 / (M_PI * V), M_PI * M_PI / b2);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(DYN_ERR_HIP, fmt("launch of k_pme_influence: %s", hipGetErrorString(e)));
}

// Brick lists for n charges: three times the mean per brick plus 64, in
// multiples of 32 (grown further when a brick overflows).
void PmeCore::ensure_cap(uint32_t n) {
  const size_t nb = (size_t)nbk[0] * nbk[1] * nbk[2];
  const double mean = (double)n * (double)(kBrick * kBrick * kBrick) / (double)ngrid;
  const uint32_t want = ((uint32_t)(3.0 * mean) + 64 + 31) & ~31u;
  if (want > cap_brick) cap_brick = want;
  brick_items.reserve(nb * cap_brick);
}

BinArgs PmeCore::bin_args(unsigned int* st, unsigned int bit) const {
  BinArgs B;
  std::memset(&B, 0, sizeof B);
  for (int a = 0; a < 3; ++a) {
    B.K[a] = (int)K[a];
    B.sc[a] = (float)((double)K[a] / (double)L[a]);
    B.nb[a] = nbk[a];
  }
  B.count = brick_count.p;
  B.items = brick_items.p;
  B.cap = cap_brick;
  B.status = st ? st : status.p;
  B.status_bit = st ? bit : 1u;
  return B;
}
void PmeCore::grow() {
  cap_brick = std::max<uint32_t>(64, 2 * cap_brick);
  brick_items.reserve((size_t)nbk[0] * nbk[1] * nbk[2] * cap_brick);
}

void PmeCore::run(uint32_t n, const float4* xyzq, const uint32_t* site, int64_t* fx_sites,
                  int comp_major, uint32_t nsites, float* f_out, float fscale, double* res) {
  PmeRun r;
  r.core = this;
  r.n = n;
  r.xyzq = xyzq;
  r.site = site;
  r.fx = fx_sites;
  r.comp_major = comp_major;
  r.nsites = nsites;
  r.f_out = f_out;
  r.fscale = fscale;
  r.res = res ? res : res_dev.p;
  BatchRun R;
  R.stream = stream;
  R.cache = &cache;
  R.n = 1;
  // a brick list that overflowed: enlarge and evaluate again
  for (int attempt = 0;; ++attempt) {
    pme_run(R, *prof, &r, 0);
    unsigned int st = 0;
    DYN_HIP(hipMemcpyAsync(&st, status.p, sizeof st, hipMemcpyDeviceToHost, stream));
    DYN_HIP(hipStreamSynchronize(stream));
    if (!st) break;
    DYN_HIP(hipMemsetAsync(status.p, 0, sizeof st, stream));
    if (attempt >= 4) throw Error(DYN_ERR_CAPACITY, "SPME brick lists keep overflowing");
    grow();
  }
}

// One SPME evaluation per system of the pass, each stage one launch for all
// of them (per precision and B-spline order present).  A system with a null
// core takes part in no stage (its packs describe empty work: no charges, no
// bricks, no lines, no result); neither does it in the launches of the other
// precision or order.
void pme_run(const BatchRun& R, Profiler& prof, const PmeRun* runs, int tag) {
  const uint32_t ns = R.n;
  std::vector<SplineArgs> S(ns);
  std::vector<PassGeom> gz(ns), gy(ns), gx(ns);
  unsigned g_n = 1, g_b = 1, bz = 1, by = 1, bx = 1;
  size_t shz = 0, shy = 0, shx = 0;
  bool any = false, ord[2][2] = {{false, false}, {false, false}}, prec[2] = {false, false};
  auto shm = [](const PassGeom& g, int N, size_t elem) { return (size_t)2 * g.L * N * elem; };
  for (uint32_t i = 0; i < ns; ++i) {
    const PmeRun& r = runs[i];
    std::memset(&S[i], 0, sizeof S[i]);
    std::memset(&gz[i], 0, sizeof gz[i]);
    gy[i] = gx[i] = gz[i];
    if (!r.core) continue;
    const PmeCore& c = *r.core;
    any = true;
    ord[c.f64][c.order == 5] = true;
    prec[c.f64] = true;
    for (int a = 0; a < 3; ++a) {
      S[i].K[a] = (int)c.K[a];
      S[i].sc[a] = (float)((double)c.K[a] / (double)c.L[a]);
      S[i].scd[a] = (double)c.K[a] / (double)c.L[a];
    }
    r.core->ensure_cap(r.n);
    gz[i] = pass_geom(c.K, 2, ns);
    gy[i] = pass_geom(c.K, 1, ns);
    gx[i] = pass_geom(c.K, 0, ns);
    const size_t el = c.f64 ? sizeof(double2) : sizeof(float2);
    g_n = std::max(g_n, grid_for(r.n, 256));
    g_b = std::max(g_b, c.nbk[0] * c.nbk[1] * c.nbk[2]);
    bz = std::max(bz, (unsigned)((gz[i].lines + gz[i].L - 1) / gz[i].L));
    by = std::max(by, (unsigned)((gy[i].lines + gy[i].L - 1) / gy[i].L));
    bx = std::max(bx, (unsigned)((gx[i].lines + gx[i].L - 1) / gx[i].L));
    shz = std::max(shz, shm(gz[i], (int)c.K[2], el));
    shy = std::max(shy, shm(gy[i], (int)c.K[1], el));
    shx = std::max(shx, shm(gx[i], (int)c.K[0], el));
  }
  if (!any) return;
  auto on = [&](uint32_t i) { return runs[i].core != nullptr; };
  auto nq = [&](uint32_t i) { return on(i) ? runs[i].n : 0u; };
  auto nb = [&](uint32_t i, int a) { return on(i) ? runs[i].core->nbk[a] : 0u; };
  auto cp = [&](uint32_t i) { return runs[i].core; };
  // this system in the launch of precision F64 (and order ORDER when given)
  auto mine = [&](uint32_t i, bool f64, int order) {
    return on(i) && cp(i)->f64 == f64 && (order == 0 || cp(i)->order == order);
  };
  // every grid point belongs to exactly one brick and is stored by it, so
  // the grid needs no clearing
  {
    Profiler::Scope ps(prof, "pme_spread");
    bool bin_any = false;
    for (uint32_t i = 0; i < ns; ++i) bin_any = bin_any || (on(i) && !runs[i].binned);
    if (bin_any)
      launch<k_pme_bin>(R, "k_pme_bin", tag, g_n, 256, 0, [&](uint32_t i) {
        const bool o = on(i) && !runs[i].binned;
        BinArgs B;
        std::memset(&B, 0, sizeof B);
        if (o) B = cp(i)->bin_args(runs[i].status, runs[i].status_bit);
        return args<k_pme_bin>(runs[i].xyzq, o ? nq(i) : 0u, B);
      });
    auto spread = [&](auto F, const char* name, bool f64, int order) {
      launch<decltype(F)::value, kBrickBlock>(R, name, tag, g_b, kBrickBlock, 0, [&](uint32_t i) {
        const bool m = mine(i, f64, order);
        return args<decltype(F)::value>(runs[i].xyzq, m ? nq(i) : 0u, S[i], m ? nb(i, 0) : 0u,
                                        m ? nb(i, 1) : 0u, m ? nb(i, 2) : 0u,
                                        m ? cp(i)->brick_count.p : nullptr,
                                        m ? cp(i)->brick_items.p : nullptr,
                                        m ? cp(i)->cap_brick : 0u,
                                        m ? cp(i)->grid_fx.p : nullptr);
      });
    };
#define DYN_SPREAD(O, T, F64)                                                                    \
  if (ord[F64][O == 5])                                                                          \
    spread(std::integral_constant<decltype(&k_pme_spread<O, T>), &k_pme_spread<O, T>>{},         \
           "k_pme_spread<" #O ", " #T ">", F64, O);
    DYN_SPREAD(4, float, 0)
    DYN_SPREAD(5, float, 0)
    DYN_SPREAD(4, double, 1)
    DYN_SPREAD(5, double, 1)
#undef DYN_SPREAD
  }
  // FFT stage over the systems of one precision: C = float2 or double2
  auto fft = [&](auto F, auto Ctag, const char* name, const std::vector<PassGeom>& g, int axis,
                 unsigned gx_, size_t shm_, bool from_fx) {
    using C = typename decltype(Ctag)::type;
    constexpr bool F64 = std::is_same<C, double2>::value;
    launch<decltype(F)::value, kFftBlock>(R, name, tag, gx_, kFftBlock, shm_, [&](uint32_t i) {
      const bool m = mine(i, F64, 0);
      C* grid = nullptr;
      const C* tw = nullptr;
      if (m) {
        if constexpr (F64) {
          grid = cp(i)->cgrid64.p;
          tw = cp(i)->tw64[axis].p;
        } else {
          grid = cp(i)->cgrid.p;
          tw = cp(i)->tw[axis].p;
        }
      }
      PassGeom ge = g[i];
      if (!m) ge.lines = 0;
      return args<decltype(F)::value>(grid, from_fx && m ? cp(i)->grid_fx.p : nullptr, ge,
                                      m ? cp(i)->ax_dev.p + axis : nullptr, tw);
    });
  };
  auto conv = [&](auto F, auto Ctag, const char* name) {
    using C = typename decltype(Ctag)::type;
    constexpr bool F64 = std::is_same<C, double2>::value;
    launch<decltype(F)::value, kFftBlock>(R, name, tag, bx, kFftBlock, shx, [&](uint32_t i) {
      const bool m = mine(i, F64, 0);
      ConvArgs Cv;
      std::memset(&Cv, 0, sizeof Cv);
      C* grid = nullptr;
      const C* tw = nullptr;
      const typename CxT<C>::R* G = nullptr;
      PassGeom ge = gx[i];
      if (m) {
        for (int a = 0; a < 3; ++a) Cv.invL[a] = 1.0f / cp(i)->L[a];
        Cv.pi2_over_b2 = M_PI * M_PI / ((double)cp(i)->beta * cp(i)->beta);
        if constexpr (F64) {
          grid = cp(i)->cgrid64.p;
          tw = cp(i)->tw64[0].p;
          G = cp(i)->G64.p;
        } else {
          grid = cp(i)->cgrid.p;
          tw = cp(i)->tw[0].p;
          G = cp(i)->G.p;
        }
      } else {
        ge.lines = 0;
      }
      return args<decltype(F)::value>(grid, ge, m ? cp(i)->ax_dev.p : nullptr, tw, G, Cv,
                                      m ? cp(i)->partials.p : nullptr);
    });
  };
#define DYN_IC(F) std::integral_constant<decltype(&F), &F>{}
  struct TF {
    using type = float2;
  };
  struct TD {
    using type = double2;
  };
  // a line length shared by every single-precision system of the pass
  // along an axis (32, 64, 128) selects the fixed-plan FFT kernels
  auto shared_n = [&](int axis) {
    int n = -1;
    for (uint32_t i = 0; i < ns; ++i) {
      if (!mine(i, false, 0)) continue;
      const int k = (int)cp(i)->K[axis];
      if (n < 0) n = k;
      else if (n != k) return 0;
    }
    return (n == 32 || n == 64 || n == 128) ? n : 0;
  };
#define DYN_FFT_F32(AXIS, SIGN, FX, NAME, G, B, SH)                                               \
  switch (shared_n(AXIS)) {                                                                       \
    case 32: fft(DYN_IC((k_fft_pass<AXIS, SIGN, FX, float2, 32>)), TF{}, NAME, G, AXIS, B, SH, FX); break; \
    case 64: fft(DYN_IC((k_fft_pass<AXIS, SIGN, FX, float2, 64>)), TF{}, NAME, G, AXIS, B, SH, FX); break; \
    case 128: fft(DYN_IC((k_fft_pass<AXIS, SIGN, FX, float2, 128>)), TF{}, NAME, G, AXIS, B, SH, FX); break; \
    default: fft(DYN_IC((k_fft_pass<AXIS, SIGN, FX, float2, 0>)), TF{}, NAME, G, AXIS, B, SH, FX); \
  }
  // the z and y passes of each direction as one plane launch when every
  // single-precision system shares fixed-plan y and z lengths
  auto plane = [&](auto F, const char* name, bool from_fx, int ny, int nz) {
    unsigned gp = 1;
    for (uint32_t i = 0; i < ns; ++i)
      if (mine(i, false, 0)) gp = std::max(gp, cp(i)->K[0]);
    launch<decltype(F)::value, kPlaneBlock>(R, name, tag, gp, kPlaneBlock,
                                            (size_t)2 * ny * nz * sizeof(float2), [&](uint32_t i) {
      const bool m = mine(i, false, 0);
      return args<decltype(F)::value>(m ? cp(i)->cgrid.p : nullptr,
                                      from_fx && m ? cp(i)->grid_fx.p : nullptr,
                                      m ? (int)cp(i)->K[0] : 0, m ? cp(i)->tw[1].p : nullptr,
                                      m ? cp(i)->tw[2].p : nullptr);
    });
  };
#define DYN_PLANE(SIGN, FX, NAME)                                                                  \
  switch (ny * 1000 + nz) {                                                                       \
    case 32032: plane(DYN_IC((k_fft_plane<SIGN, FX, 32, 32>)), NAME, FX, 32, 32); break;           \
    case 32064: plane(DYN_IC((k_fft_plane<SIGN, FX, 32, 64>)), NAME, FX, 32, 64); break;           \
    case 64032: plane(DYN_IC((k_fft_plane<SIGN, FX, 64, 32>)), NAME, FX, 64, 32); break;           \
    default: plane(DYN_IC((k_fft_plane<SIGN, FX, 64, 64>)), NAME, FX, 64, 64);                     \
  }
  for (int pr = 0; pr < 2; ++pr) {
    if (!prec[pr]) continue;
    if (pr == 0) {
      const int ny = shared_n(1), nz = shared_n(2);
      const bool planes = ny > 0 && nz > 0 && ny <= 64 && nz <= 64;
      if (planes) {
        Profiler::Scope ps(prof, "pme_fft_zy_fwd");
        DYN_PLANE(-1, true, "k_fft_plane fwd")
      } else {
        { Profiler::Scope ps(prof, "pme_fft_z_fwd");
          DYN_FFT_F32(2, -1, true, "k_fft_pass z fwd", gz, bz, shz) }
        { Profiler::Scope ps(prof, "pme_fft_y_fwd");
          DYN_FFT_F32(1, -1, false, "k_fft_pass y fwd", gy, by, shy) }
      }
      { Profiler::Scope ps(prof, "pme_fft_x_conv");
        switch (shared_n(0)) {
          case 32: conv(DYN_IC((k_fft_conv_x<float2, 32>)), TF{}, "k_fft_conv_x"); break;
          case 64: conv(DYN_IC((k_fft_conv_x<float2, 64>)), TF{}, "k_fft_conv_x"); break;
          case 128: conv(DYN_IC((k_fft_conv_x<float2, 128>)), TF{}, "k_fft_conv_x"); break;
          default: conv(DYN_IC((k_fft_conv_x<float2, 0>)), TF{}, "k_fft_conv_x");
        } }
      if (planes) {
        Profiler::Scope ps(prof, "pme_fft_yz_inv");
        DYN_PLANE(1, false, "k_fft_plane inv")
      } else {
        { Profiler::Scope ps(prof, "pme_fft_y_inv");
          DYN_FFT_F32(1, 1, false, "k_fft_pass y inv", gy, by, shy) }
        { Profiler::Scope ps(prof, "pme_fft_z_inv");
          DYN_FFT_F32(2, 1, false, "k_fft_pass z inv", gz, bz, shz) }
      }
    } else {
      { Profiler::Scope ps(prof, "pme_fft_z_fwd");
        fft(DYN_IC((k_fft_pass<2, -1, true, double2, 0>)), TD{}, "k_fft_pass z fwd f64", gz, 2, bz, shz, true); }
      { Profiler::Scope ps(prof, "pme_fft_y_fwd");
        fft(DYN_IC((k_fft_pass<1, -1, false, double2, 0>)), TD{}, "k_fft_pass y fwd f64", gy, 1, by, shy, false); }
      { Profiler::Scope ps(prof, "pme_fft_x_conv");
        conv(DYN_IC((k_fft_conv_x<double2, 0>)), TD{}, "k_fft_conv_x f64"); }
      { Profiler::Scope ps(prof, "pme_fft_y_inv");
        fft(DYN_IC((k_fft_pass<1, 1, false, double2, 0>)), TD{}, "k_fft_pass y inv f64", gy, 1, by, shy, false); }
      { Profiler::Scope ps(prof, "pme_fft_z_inv");
        fft(DYN_IC((k_fft_pass<2, 1, false, double2, 0>)), TD{}, "k_fft_pass z inv f64", gz, 2, bz, shz, false); }
    }
  }
#undef DYN_FFT_F32
#undef DYN_PLANE
  {
    Profiler::Scope ps(prof, "pme_gather");
    const unsigned gg = std::max(1u, (g_n * 256 + kGatherBlock - 1) / kGatherBlock);
    auto gather = [&](auto F, auto Ctag, int order) {
      using C = typename decltype(Ctag)::type;
      constexpr bool F64 = std::is_same<C, double2>::value;
      launch<decltype(F)::value, kGatherBlock>(R, "k_pme_gather",
                                               tag * 4 + (F64 ? 2 : 0) + (order == 5), gg,
                                               kGatherBlock, 0, [&](uint32_t i) {
        const bool m = mine(i, F64, order);
        const C* phi = nullptr;
        if (m) {
          if constexpr (F64) phi = cp(i)->cgrid64.p;
          else phi = cp(i)->cgrid.p;
        }
        return args<decltype(F)::value>(runs[i].xyzq, m ? nq(i) : 0u, S[i], phi, runs[i].site,
                                        runs[i].fx, runs[i].comp_major, runs[i].nsites,
                                        runs[i].f_out, runs[i].fscale,
                                        m ? (const int64_t*)cp(i)->partials.p : nullptr,
                                        m ? (gx[i].lines + gx[i].L - 1) / gx[i].L : 0,
                                        m ? runs[i].res : nullptr,
                                        m ? cp(i)->brick_count.p : nullptr,
                                        m ? nb(i, 0) * nb(i, 1) * nb(i, 2) : 0u);
      });
    };
    if (ord[0][0]) gather(DYN_IC((k_pme_gather<4, float2>)), TF{}, 4);
    if (ord[0][1]) gather(DYN_IC((k_pme_gather<5, float2>)), TF{}, 5);
    if (ord[1][0]) gather(DYN_IC((k_pme_gather<4, double2>)), TD{}, 4);
    if (ord[1][1]) gather(DYN_IC((k_pme_gather<5, double2>)), TD{}, 5);
  }
#undef DYN_IC
}

}  // namespace dyn

// ---------------------------------------------------------------------------
// Standalone PmeRecip replacement
// ---------------------------------------------------------------------------
struct dyn_pme {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  Profiler prof;
  PmeCore core;
  DevBuf<float4> xyzq;
  DevBuf<float> fout;
  PinnedBuf<float4> h_xyzq;
  bool computed = false;
  ~dyn_pme() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

extern "C" int dyn_pme_create(dyn_pme** out, int device, uint32_t nx, uint32_t ny, uint32_t nz,
                              float lx, float ly, float lz, float alpha, uint32_t order) {
  try {
    DYN_REQUIRE(out != nullptr, "out is null");
    *out = nullptr;
    int ndev = 0;
    DYN_HIP(hipGetDeviceCount(&ndev));
    DYN_REQUIRE(device >= 0 && device < ndev, "device %d not present (%d devices)", device, ndev);
    std::unique_ptr<dyn_pme> p(new dyn_pme());
    p->device = device;
    DeviceGuard g(device);
    DYN_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    p->prof.stream = p->stream;
    const uint32_t K[3] = {nx, ny, nz};
    const float L[3] = {lx, ly, lz};
    p->core.init(device, p->stream, &p->prof, K, L, alpha, (int)order, false);
    *out = p.release();
    return DYN_OK;
  } catch (const Error& e) {
    set_global_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_global_error(e.what());
    return DYN_ERR_HIP;
  }
}

extern "C" void dyn_pme_destroy(dyn_pme* pme) {
  if (!pme) return;
  try {
    DeviceGuard g(pme->device);
    delete pme;
  } catch (...) {
  }
}

extern "C" const char* dyn_pme_last_error(const dyn_pme* pme) {
  return pme ? pme->err.c_str() : dyn_global_last_error();
}

extern "C" int dyn_pme_forces_and_virial(dyn_pme* pme, uint32_t n, const float* pos,
                                         const float* q, float* f_out, double* energy,
                                         double* virial) {
  if (!pme) {
    set_global_error("null handle");
    return DYN_ERR_ARG;
  }
  DYN_API_BODY(pme, {
    DeviceGuard g(pme->device);
    DYN_REQUIRE(n == 0 || (pos && q), "null positions or charges");
    const size_t nn = std::max<uint32_t>(n, 1);
    pme->xyzq.reserve(nn);
    pme->fout.reserve(3 * nn);
    pme->h_xyzq.reserve(nn);
    double q2 = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      DYN_REQUIRE(std::isfinite(pos[3 * i]) && std::isfinite(pos[3 * i + 1]) &&
                      std::isfinite(pos[3 * i + 2]) && std::isfinite(q[i]),
                  "non-finite input at charge %u", i);
      pme->h_xyzq.p[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], q[i]);
      q2 += (double)q[i] * (double)q[i];
    }
    if (n)
      DYN_HIP(hipMemcpyAsync(pme->xyzq.p, pme->h_xyzq.p, sizeof(float4) * n,
                             hipMemcpyHostToDevice, pme->stream));
    pme->core.run(n, pme->xyzq.p, nullptr, nullptr, 0, 0, pme->fout.p);
    double res[2] = {0, 0};
    DYN_HIP(hipMemcpyAsync(res, pme->core.res_dev.p, sizeof(res), hipMemcpyDeviceToHost,
                           pme->stream));
    if (n && f_out)
      DYN_HIP(hipMemcpyAsync(f_out, pme->fout.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost,
                             pme->stream));
    DYN_HIP(hipStreamSynchronize(pme->stream));
    const double e_self = (double)pme->core.beta / std::sqrt(M_PI) * q2;
    if (energy) *energy = res[0] - e_self;
    if (virial) *virial = res[1];
    pme->computed = true;
  });
}

extern "C" int dyn_pme_get_fixed_grid(dyn_pme* pme, int64_t* out, uint64_t n_values) {
  if (!pme) {
    set_global_error("null handle");
    return DYN_ERR_ARG;
  }
  DYN_API_BODY(pme, {
    DeviceGuard g(pme->device);
    if (!pme->computed) throw Error(DYN_ERR_STATE, "no compute yet");
    DYN_REQUIRE(out && n_values >= pme->core.ngrid, "need %zu values", pme->core.ngrid);
    DYN_HIP(hipMemcpyAsync(out, pme->core.grid_fx.p, sizeof(int64_t) * pme->core.ngrid,
                           hipMemcpyDeviceToHost, pme->stream));
    DYN_HIP(hipStreamSynchronize(pme->stream));
  });
}

extern "C" int dyn_pme_grid_n(float lx, float ly, float lz, float spacing, uint32_t out[3]) {
  try {
    DYN_REQUIRE(out, "out is null");
    DYN_REQUIRE(spacing > 0.f && std::isfinite(spacing), "spacing must be > 0");
    const float l[3] = {lx, ly, lz};
    for (int a = 0; a < 3; ++a) {
      DYN_REQUIRE(l[a] > 0.f && std::isfinite(l[a]), "box length must be > 0");
      uint32_t n = (uint32_t)std::ceil((double)l[a] / (double)spacing);
      if (n < 5) n = 5;
      while (!is_235_smooth(n)) ++n;
      out[a] = n;
    }
    return DYN_OK;
  } catch (const Error& e) {
    set_global_error(e.what());
    return e.code;
  }
}
