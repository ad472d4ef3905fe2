// compat.hip — the reference's device boundary (src/gpu_interface.rs) on
// gfx950: the flattened-pair-list force kernel with 64-bit fixed-point
// accumulation, the exact neighbor-list build of src/neighbors.rs:180-281,
// and the 1-4 reciprocal correction of src/non_bonded.rs:763-790.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <vector>

#include "common.h"
#include "dyn_math.h"
#include "pair_ref.h"

#pragma clang fp contract(off)

namespace dyn {
static thread_local std::string g_err;
void set_global_error(const std::string& m) { g_err = m; }
}  // namespace dyn

using namespace dyn;

extern "C" int dyn_version(void) { return 100; }
extern "C" int dyn_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
extern "C" const char* dyn_global_last_error(void) { return dyn::g_err.c_str(); }

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
namespace {

constexpr int kRefBlock = 256;

__device__ __forceinline__ void atomic_add_fx(int64_t* addr, int64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(addr), static_cast<unsigned long long>(v));
}

// Block-wide sum of a double in a fixed tree order (deterministic).
template <int BLOCK>
__device__ double block_sum_f64(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// ---- scratch of the allocation-free 1-4 correction ------------------------
// dyn_nb_correct_14_noalloc stages its inputs, fixed-point sums and partials
// here instead of in buffers the context allocates.  Each device holds its
// own copy for the life of the process; host code uses the address from
// hipGetSymbolAddress (c14_scratch_ptr), never the symbol itself.
constexpr uint32_t kC14MaxAtoms = 1u << 16;  // solute atoms of one call
constexpr uint32_t kC14MaxPairs = 1u << 18;  // 1-4 pairs of one call
constexpr unsigned kC14MaxBlocks = 1024u;    // grid of k_corr14_vacuum (grid_for(n, 256, 1024))
constexpr size_t c14_align(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kC14OffFx = 0;                                                      // int64 [3·atoms]
constexpr size_t kC14OffF = kC14OffFx + c14_align(8ull * 3 * kC14MaxAtoms);          // double [3·atoms]
constexpr size_t kC14OffPart = kC14OffF + c14_align(8ull * 3 * kC14MaxAtoms);        // double [blocks + 1]
constexpr size_t kC14OffPos = kC14OffPart + c14_align(8ull * (kC14MaxBlocks + 1));   // float [3·atoms]
constexpr size_t kC14OffQ = kC14OffPos + c14_align(4ull * 3 * kC14MaxAtoms);         // float [atoms]
constexpr size_t kC14OffI = kC14OffQ + c14_align(4ull * kC14MaxAtoms);               // uint32 [pairs]
constexpr size_t kC14OffJ = kC14OffI + c14_align(4ull * kC14MaxPairs);               // uint32 [pairs]
constexpr size_t kC14ScratchBytes = kC14OffJ + c14_align(4ull * kC14MaxPairs);       // 6.3 MB

// Zero n int64 values (a fill kernel, not a memset: callers may capture it).
__global__ void k_fill_zero_i64(int64_t* __restrict__ p, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    p[k] = 0;
}

// f[k] += fx[k]·2⁻³² (the caller's forces, staged in f) and w_out = the sum
// of the per-block partials in block order: the arithmetic and order of the
// host-side finish of dyn_nb_correct_14.
__global__ void k_c14_finish(const int64_t* __restrict__ fx, double* __restrict__ f, size_t n,
                             const double* __restrict__ part, unsigned nblk,
                             double* __restrict__ w_out) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    f[k] += (double)fx[k] * DYN_FX_INV_SCALE;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double w = 0;
    for (unsigned b = 0; b < nblk; ++b) w += part[b];
    *w_out = w;
  }
}

// One thread per pair, grid-stride (src/cuda/cuda.cu:232-239).  Forces go
// into int64 fixed-point sums (order-independent); energy, virial and dH/dλ
// into per-block f64 partials reduced in a fixed order.
__global__ void __launch_bounds__(kRefBlock)
k_ref_pairs(const DynPairRec* __restrict__ pairs, uint32_t n, const float* __restrict__ pos,
            DynRefParams P, int energy_all, int64_t* __restrict__ fx,
            double* __restrict__ partials) {
  __shared__ double sh[kRefBlock];
  double e_acc = 0.0, w_acc = 0.0, dh_acc = 0.0;
  for (uint32_t i = blockIdx.x * kRefBlock + threadIdx.x; i < n; i += gridDim.x * kRefBlock) {
    const DynPairRec p = pairs[i];
    float f[3], e, dh;
    double vir;
    if (!dyn_ref_pair(pos, p, P, f, &e, &dh, &vir)) continue;
    w_acc += vir;
    if (energy_all || (p.flags & DYN_PF_TGT_SOLUTE)) e_acc += (double)e;
    if (p.flags & DYN_PF_ALCH) dh_acc += (double)dh;
    int64_t* ft = fx + 3u * p.tgt;
    const int64_t q0 = dyn_fx(f[0]), q1 = dyn_fx(f[1]), q2 = dyn_fx(f[2]);
    atomic_add_fx(ft + 0, q0);
    atomic_add_fx(ft + 1, q1);
    atomic_add_fx(ft + 2, q2);
    if (p.flags & DYN_PF_SYMMETRIC) {
      int64_t* fs = fx + 3u * p.src;
      atomic_add_fx(fs + 0, -q0);
      atomic_add_fx(fs + 1, -q1);
      atomic_add_fx(fs + 2, -q2);
    }
  }
  const double e_b = block_sum_f64<kRefBlock>(e_acc, sh);
  const double w_b = block_sum_f64<kRefBlock>(w_acc, sh);
  const double d_b = block_sum_f64<kRefBlock>(dh_acc, sh);
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x + 0] = e_b;
    partials[3 * blockIdx.x + 1] = w_b;
    partials[3 * blockIdx.x + 2] = d_b;
  }
}

// 1-4 vacuum correction (src/non_bonded.rs:763-790), force only.
__global__ void k_corr14_vacuum(const uint32_t* __restrict__ ii, const uint32_t* __restrict__ jj,
                                uint32_t n, const float* __restrict__ pos,
                                const float* __restrict__ q, DynRefParams P,
                                int64_t* __restrict__ fx, double* __restrict__ partials) {
  __shared__ double sh[kRefBlock];
  double w_acc = 0.0;
  for (uint32_t k = blockIdx.x * kRefBlock + threadIdx.x; k < n; k += gridDim.x * kRefBlock) {
    const uint32_t i = ii[k], j = jj[k];
    float d0 = pos[3 * i] - pos[3 * j];
    float d1 = pos[3 * i + 1] - pos[3 * j + 1];
    float d2 = pos[3 * i + 2] - pos[3 * j + 2];
    d0 = dyn_min_image_rint(d0, P.ext[0], P.inv_ext[0]);
    d1 = dyn_min_image_rint(d1, P.ext[1], P.inv_ext[1]);
    d2 = dyn_min_image_rint(d2, P.ext[2], P.inv_ext[2]);
    const float r2 = dyn_fmaf(d2, d2, dyn_fmaf(d1, d1, d0 * d0));
    if (r2 < 1e-12f) continue;
    const float inv_r = dyn_rsqrtf(r2);
    // df = (1/1.2 − 1)·qq/r² along r̂ = (…)·qq/r³ · d
    const float s = (DYN_SCALE_COUL_14 - 1.0f) * (q[i] * q[j]) * (inv_r * inv_r * inv_r);
    const float f0 = d0 * s, f1 = d1 * s, f2 = d2 * s;
    w_acc += (double)d0 * (double)f0 + (double)d1 * (double)f1 + (double)d2 * (double)f2;
    const int64_t q0 = dyn_fx(f0), q1 = dyn_fx(f1), q2 = dyn_fx(f2);
    atomic_add_fx(fx + 3 * i + 0, q0);
    atomic_add_fx(fx + 3 * i + 1, q1);
    atomic_add_fx(fx + 3 * i + 2, q2);
    atomic_add_fx(fx + 3 * j + 0, -q0);
    atomic_add_fx(fx + 3 * j + 1, -q1);
    atomic_add_fx(fx + 3 * j + 2, -q2);
  }
  const double w_b = block_sum_f64<kRefBlock>(w_acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = w_b;
}

// ---- neighbor-list build ---------------------------------------------------
struct NblGrid {
  int nc[3];        // cells per axis
  double lo[3];     // binning origin (0) and cell width per axis
  double cw[3];
  double ext[3];
};

__device__ __forceinline__ int nbl_cell_1d(float x, double ext, double cw, int nc) {
  double w = fmod((double)x, ext);
  if (w < 0) w += ext;
  int c = (int)(w / cw);
  if (c >= nc) c = nc - 1;
  if (c < 0) c = 0;
  return c;
}

__global__ void k_nbl_cell_count(const float* __restrict__ p, uint32_t n, NblGrid G,
                                 uint32_t* __restrict__ cell_of, uint32_t* __restrict__ counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int cx = nbl_cell_1d(p[3 * i], G.ext[0], G.cw[0], G.nc[0]);
  const int cy = nbl_cell_1d(p[3 * i + 1], G.ext[1], G.cw[1], G.nc[1]);
  const int cz = nbl_cell_1d(p[3 * i + 2], G.ext[2], G.cw[2], G.nc[2]);
  const uint32_t c = (uint32_t)((cx * G.nc[1] + cy) * G.nc[2] + cz);
  cell_of[i] = c;
  atomicAdd(&counts[c], 1u);
}

__global__ void k_nbl_cell_fill(uint32_t n, const uint32_t* __restrict__ cell_of,
                                const uint32_t* __restrict__ cell_start,
                                uint32_t* __restrict__ cursor, uint32_t* __restrict__ items) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = cell_of[i];
  const uint32_t slot = atomicAdd(&cursor[c], 1u);
  items[cell_start[c] + slot] = i;
}

// Single-workgroup exclusive scan of n uint32 values into out[0..n]
// (out[n] = total), also writing the maximum element to *maxv.
constexpr int kScanBlock = 1024;
__global__ void __launch_bounds__(kScanBlock)
k_scan_u32(const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ out,
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
  // Hillis–Steele inclusive scan over per-thread sums
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
  uint32_t run = sh[t] - s;  // exclusive prefix of this thread's range
  for (uint32_t i = b; i < e; ++i) {
    out[i] = run;
    run += in[i];
  }
  if (t == kScanBlock - 1) out[n] = sh[t];
  if (t == 0) *maxv = shm[0];
}

constexpr int kNblBlock = 256;
constexpr int kNblCap = 4096;  // max neighbors per row handled in LDS

struct NblRowArgs {
  const float* outer;
  const float* inner;
  const uint8_t* is_static;
  const uint32_t* cell_start;
  const uint32_t* items;
  NblGrid G;
  float ext[3];
  float thr;
  int symmetric;
  int offs;  // 1 if ±1 neighbour cells are used on every axis, else per-axis flags below
  int use_nb[3];
};

// The reference's exact pair test (src/neighbors.rs:221-222).
__device__ __forceinline__ bool nbl_accept(const float* po, const float* pi, const float ext[3],
                                           float thr) {
  const float dx = dyn_min_image_ref(po[0] - pi[0], ext[0]);
  const float dy = dyn_min_image_ref(po[1] - pi[1], ext[1]);
  const float dz = dyn_min_image_ref(po[2] - pi[2], ext[2]);
  return dyn_norm2_ref(dx, dy, dz) < thr;
}

// One workgroup per outer point.  PASS 0 counts accepted inner points; PASS 1
// collects them in LDS, bitonic-sorts ascending and writes the row.
template <int PASS>
__global__ void __launch_bounds__(kNblBlock)
k_nbl_rows(NblRowArgs A, uint32_t n_outer, uint32_t* __restrict__ row_count,
           const uint32_t* __restrict__ row_ptr, uint32_t* __restrict__ col) {
  __shared__ uint32_t buf[PASS == 1 ? kNblCap : 1];
  __shared__ uint32_t cnt;
  const uint32_t i = blockIdx.x;
  if (i >= n_outer) return;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const float po[3] = {A.outer[3 * i], A.outer[3 * i + 1], A.outer[3 * i + 2]};
  const bool st_i = A.is_static ? A.is_static[i] != 0 : false;
  const int cx = nbl_cell_1d(po[0], A.G.ext[0], A.G.cw[0], A.G.nc[0]);
  const int cy = nbl_cell_1d(po[1], A.G.ext[1], A.G.cw[1], A.G.nc[1]);
  const int cz = nbl_cell_1d(po[2], A.G.ext[2], A.G.cw[2], A.G.nc[2]);
  uint32_t local = 0;
  for (int ox = -A.use_nb[0]; ox <= A.use_nb[0]; ++ox)
    for (int oy = -A.use_nb[1]; oy <= A.use_nb[1]; ++oy)
      for (int oz = -A.use_nb[2]; oz <= A.use_nb[2]; ++oz) {
        const int x = (cx + ox + A.G.nc[0]) % A.G.nc[0];
        const int y = (cy + oy + A.G.nc[1]) % A.G.nc[1];
        const int z = (cz + oz + A.G.nc[2]) % A.G.nc[2];
        const uint32_t c = (uint32_t)((x * A.G.nc[1] + y) * A.G.nc[2] + z);
        const uint32_t b = A.cell_start[c], e = A.cell_start[c + 1];
        for (uint32_t k = b + threadIdx.x; k < e; k += kNblBlock) {
          const uint32_t j = A.items[k];
          if (A.symmetric && j == i) continue;
          if (A.is_static && st_i && A.is_static[j]) continue;
          if (!nbl_accept(po, A.inner + 3 * j, A.ext, A.thr)) continue;
          if (PASS == 0) {
            ++local;
          } else {
            const uint32_t slot = atomicAdd(&cnt, 1u);
            if (slot < kNblCap) buf[slot] = j;
          }
        }
      }
  if (PASS == 0) {
    atomicAdd(&cnt, local);
    __syncthreads();
    if (threadIdx.x == 0) row_count[i] = cnt;
    return;
  } else {
    __syncthreads();
    const uint32_t m = min(cnt, (uint32_t)kNblCap);
    uint32_t np2 = 1;
    while (np2 < m) np2 <<= 1;
    for (uint32_t k = m + threadIdx.x; k < np2; k += kNblBlock) buf[k] = 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t size = 2; size <= np2; size <<= 1) {
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t t = threadIdx.x; t < np2; t += kNblBlock) {
          const uint32_t partner = t ^ stride;
          if (partner > t) {
            const bool up = (t & size) == 0;
            const uint32_t a = buf[t], b = buf[partner];
            if ((a > b) == up) {
              buf[t] = b;
              buf[partner] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    const uint32_t base = row_ptr[i];
    for (uint32_t k = threadIdx.x; k < m; k += kNblBlock) col[base + k] = buf[k];
  }
}

}  // namespace

// The allocation-free 1-4 correction's device memory (kC14ScratchBytes per
// device, held until the process ends).  In host code this name is only a
// stand-in for the symbol: kernels get the address c14_scratch_ptr returns.
static __device__ __align__(256) unsigned char scratch[kC14ScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (that device must be current):
// looked up once per device, at context creation or at the first
// allocation-free call there.
unsigned char* c14_scratch_ptr(int device) {
  DYN_REQUIRE(device >= 0 && device < kMaxDevices, "device %d beyond the scratch table", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(DYN_ERR_HIP, fmt("hipGetSymbolAddress(scratch) on device %d: %s", device,
                                   hipGetErrorString(e)));
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[device];
}
}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct dyn_nb_ctx {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t n_dyn = 0, n_water = 0, n_sites = 0;
  float cutoff = 10.f, alpha = 0.3f;
  dyn_nb_opts opts{INFINITY, 0};
  // device state
  DevBuf<float> pos;          // n_sites × 3, class-major (solute, O, M, H0, H1)
  DevBuf<int64_t> fx;         // n_sites × 3 fixed point
  DevBuf<DynPairRec> pairs;
  uint32_t n_pairs = 0;
  bool pairs_set = false;
  bool any_alch = false;
  DevBuf<double> partials;
  PinnedBuf<float> h_pos;
  PinnedBuf<int64_t> h_fx;
  PinnedBuf<double> h_part;
  // neighbor build scratch
  DevBuf<float> nb_outer, nb_inner;
  DevBuf<uint8_t> nb_static;
  DevBuf<uint32_t> nb_cell_of, nb_counts, nb_cell_start, nb_cursor, nb_items, nb_row_count,
      nb_row_ptr, nb_col, nb_max;
  DevBuf<uint32_t> c14_i, c14_j;
  DevBuf<float> c14_pos, c14_q;
  DevBuf<int64_t> c14_fx;
  bool computed = false;

  ~dyn_nb_ctx() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {
unsigned ref_grid(uint32_t n) {
  // Enough blocks to fill 256 CUs several times; fixed for a given n so the
  // per-block partial sums (and hence energies) are deterministic.
  return grid_for(n, kRefBlock, 4096);
}

DynRefParams make_ref_params(const dyn_nb_ctx* c, const float ext[3], bool coul_dis,
                             bool lj_dis, double lambda) {
  DynRefParams P;
  for (int k = 0; k < 3; ++k) {
    P.ext[k] = ext[k];
    P.inv_ext[k] = 1.0f / ext[k];
  }
  P.cutoff = c->cutoff;
  P.beta = c->alpha;
  P.beta2 = c->alpha * c->alpha;
  P.two_b_sqrtpi = 2.0f * c->alpha * DYN_INV_SQRT_PI;
  P.lj_cutoff = c->opts.lj_cutoff;
  P.coulomb_disabled = coul_dis;
  P.lj_disabled = lj_dis;
  P.sch = dyn_schedule((float)lambda);
  return P;
}
}  // namespace

extern "C" int dyn_nb_create(dyn_nb_ctx** out, int device, uint32_t n_dyn, uint32_t n_water,
                             float cutoff_ewald, float alpha_ewald, const dyn_nb_opts* opts) {
  try {
    DYN_REQUIRE(out != nullptr, "out is null");
    *out = nullptr;
    DYN_REQUIRE(cutoff_ewald > 0.f && std::isfinite(cutoff_ewald), "cutoff must be > 0");
    DYN_REQUIRE(alpha_ewald > 0.f && std::isfinite(alpha_ewald), "alpha must be > 0");
    DYN_REQUIRE(alpha_ewald * cutoff_ewald <= 4.0f,
                "alpha·cutoff = %g exceeds 4, the range of the erfc approximation",
                alpha_ewald * cutoff_ewald);
    int ndev = 0;
    DYN_HIP(hipGetDeviceCount(&ndev));
    DYN_REQUIRE(device >= 0 && device < ndev, "device %d not present (%d devices)", device, ndev);
    std::unique_ptr<dyn_nb_ctx> c(new dyn_nb_ctx());
    c->device = device;
    DeviceGuard g(device);
    DYN_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->n_dyn = n_dyn;
    c->n_water = n_water;
    c->n_sites = n_dyn + 4 * n_water;
    c->cutoff = cutoff_ewald;
    c->alpha = alpha_ewald;
    if (opts) c->opts = *opts;
    // Zero-size systems keep one element, as the reference pads allocations
    // to avoid zero-length device buffers (src/gpu_interface.rs:130-134).
    const size_t ns = std::max<uint32_t>(c->n_sites, 1);
    c->pos.reserve(ns * 3);
    c->fx.reserve(ns * 3);
    c->h_pos.reserve(ns * 3);
    c->h_fx.reserve(ns * 3);
    DYN_HIP(hipMemsetAsync(c->pos.p, 0, c->pos.bytes(), c->stream));
    DYN_HIP(hipMemsetAsync(c->fx.p, 0, c->fx.bytes(), c->stream));
    DYN_HIP(hipStreamSynchronize(c->stream));
    (void)c14_scratch_ptr(device);  // this device's scratch address, once
    *out = c.release();
    return DYN_OK;
  } catch (const Error& e) {
    set_global_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_global_error(e.what());
    return DYN_ERR_HIP;
  }
}

extern "C" void dyn_nb_destroy(dyn_nb_ctx* ctx) {
  if (!ctx) return;
  try {
    DeviceGuard g(ctx->device);
    delete ctx;
  } catch (...) {
  }
}

extern "C" const char* dyn_nb_last_error(const dyn_nb_ctx* ctx) {
  return ctx ? ctx->err.c_str() : dyn::g_err.c_str();
}

extern "C" int dyn_nb_set_pairs(dyn_nb_ctx* ctx, uint32_t n, const uint32_t* tgt,
                                const uint32_t* src, const float* sigma, const float* eps,
                                const float* q_tgt, const float* q_src,
                                const uint8_t* atom_type_tgt, const uint8_t* water_site_tgt,
                                const uint8_t* atom_type_src, const uint8_t* water_site_src,
                                const uint8_t* scale14, const uint8_t* calc_lj,
                                const uint8_t* calc_coul, const uint8_t* symmetric,
                                const uint8_t* alch) {
  if (!ctx) {
    set_global_error("null context");
    return DYN_ERR_ARG;
  }
  DYN_API_BODY(ctx, {
    DeviceGuard g(ctx->device);
    if (n > 0) {
      DYN_REQUIRE(tgt && src && sigma && eps && q_tgt && q_src && atom_type_tgt &&
                      water_site_tgt && atom_type_src && water_site_src && scale14 &&
                      calc_lj && calc_coul && symmetric,
                  "null pair array");
    }
    const uint32_t nd = ctx->n_dyn, nw = ctx->n_water;
    auto site = [&](uint8_t at, uint8_t ws, uint32_t idx, uint32_t k) -> uint32_t {
      if (at == 0) {
        DYN_REQUIRE(idx < nd, "pair %u: solute index %u out of range (n_dyn %u)", k, idx, nd);
        return idx;
      }
      DYN_REQUIRE(at == 1, "pair %u: atom_type %u is neither 0 nor 1", k, (unsigned)at);
      DYN_REQUIRE(ws >= 1 && ws <= 4, "pair %u: water_site %u not in 1..4", k, (unsigned)ws);
      DYN_REQUIRE(idx < nw, "pair %u: water index %u out of range (n_water %u)", k, idx, nw);
      return nd + (uint32_t)(ws - 1) * nw + idx;
    };
    std::vector<DynPairRec> recs(n);
    bool any_alch = false;
    for (uint32_t k = 0; k < n; ++k) {
      DynPairRec r{};
      r.tgt = site(atom_type_tgt[k], water_site_tgt[k], tgt[k], k);
      r.src = site(atom_type_src[k], water_site_src[k], src[k], k);
      r.sigma = sigma[k];
      r.eps = eps[k];
      r.q_tgt = q_tgt[k];
      r.q_src = q_src[k];
      uint32_t f = 0;
      if (scale14[k]) f |= DYN_PF_SCALE14;
      if (calc_lj[k]) f |= DYN_PF_LJ;
      if (calc_coul[k]) f |= DYN_PF_COUL;
      if (symmetric[k]) f |= DYN_PF_SYMMETRIC;
      if (alch && alch[k]) {
        f |= DYN_PF_ALCH;
        any_alch = true;
      }
      if (atom_type_tgt[k] == 0) f |= DYN_PF_TGT_SOLUTE;
      r.flags = f;
      recs[k] = r;
    }
    ctx->pairs.reserve(std::max<uint32_t>(n, 1));
    if (n) {
      DYN_HIP(hipMemcpyAsync(ctx->pairs.p, recs.data(), n * sizeof(DynPairRec),
                             hipMemcpyHostToDevice, ctx->stream));
    }
    ctx->n_pairs = n;
    ctx->any_alch = any_alch;
    ctx->pairs_set = true;
    const unsigned grid = ref_grid(n);
    ctx->partials.reserve(3 * (size_t)grid);
    ctx->h_part.reserve(3 * (size_t)grid);
    DYN_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int dyn_nb_compute(dyn_nb_ctx* ctx, const float* pos_dyn, const float* pos_w_o,
                              const float* pos_w_m, const float* pos_w_h0,
                              const float* pos_w_h1, const float cell_extent[3],
                              uint8_t coulomb_disabled, uint8_t lj_disabled,
                              double lambda_alch, double* f_dyn, double* f_water,
                              double* virial, double* energy, double* dh_dl) {
  if (!ctx) {
    set_global_error("null context");
    return DYN_ERR_ARG;
  }
  DYN_API_BODY(ctx, {
    DeviceGuard g(ctx->device);
    if (!ctx->pairs_set) throw Error(DYN_ERR_STATE, "dyn_nb_compute before dyn_nb_set_pairs");
    DYN_REQUIRE(cell_extent != nullptr, "cell_extent is null");
    for (int k = 0; k < 3; ++k)
      DYN_REQUIRE(cell_extent[k] > 0.f && std::isfinite(cell_extent[k]),
                  "cell extent must be positive and finite");
    const uint32_t nd = ctx->n_dyn, nw = ctx->n_water;
    DYN_REQUIRE(nd == 0 || pos_dyn, "pos_dyn is null");
    DYN_REQUIRE(nw == 0 || (pos_w_o && pos_w_m && pos_w_h0 && pos_w_h1), "water positions null");
    // upload_positions (src/gpu_interface.rs:414-455): one staging buffer,
    // one copy.
    float* hp = ctx->h_pos.p;
    if (nd) memcpy(hp, pos_dyn, sizeof(float) * 3 * nd);
    const float* wsrc[4] = {pos_w_o, pos_w_m, pos_w_h0, pos_w_h1};
    for (int s = 0; s < 4; ++s)
      if (nw) memcpy(hp + 3 * (nd + s * nw), wsrc[s], sizeof(float) * 3 * nw);
    for (size_t k = 0; k < 3 * (size_t)ctx->n_sites; ++k)
      DYN_REQUIRE(std::isfinite(hp[k]), "non-finite position (site %zu)", k / 3);
    if (ctx->n_sites)
      DYN_HIP(hipMemcpyAsync(ctx->pos.p, hp, sizeof(float) * 3 * ctx->n_sites,
                             hipMemcpyHostToDevice, ctx->stream));
    DYN_HIP(hipMemsetAsync(ctx->fx.p, 0, ctx->fx.bytes(), ctx->stream));
    const DynRefParams P =
        make_ref_params(ctx, cell_extent, coulomb_disabled != 0, lj_disabled != 0, lambda_alch);
    const uint32_t n = ctx->n_pairs;
    const unsigned grid = ref_grid(n);
    if (n) {
      hipLaunchKernelGGL(k_ref_pairs, dim3(grid), dim3(kRefBlock), 0, ctx->stream,
                         ctx->pairs.p, n, ctx->pos.p, P, (int)ctx->opts.energy_all_pairs,
                         ctx->fx.p, ctx->partials.p);
      DYN_CHECK_LAUNCH("k_ref_pairs");
      DYN_HIP(hipMemcpyAsync(ctx->h_part.p, ctx->partials.p, sizeof(double) * 3 * grid,
                             hipMemcpyDeviceToHost, ctx->stream));
    }
    if (ctx->n_sites)
      DYN_HIP(hipMemcpyAsync(ctx->h_fx.p, ctx->fx.p, sizeof(int64_t) * 3 * ctx->n_sites,
                             hipMemcpyDeviceToHost, ctx->stream));
    DYN_HIP(hipStreamSynchronize(ctx->stream));
    double e = 0, w = 0, d = 0;
    if (n)
      for (unsigned b = 0; b < grid; ++b) {
        e += ctx->h_part.p[3 * b];
        w += ctx->h_part.p[3 * b + 1];
        d += ctx->h_part.p[3 * b + 2];
      }
    const int64_t* F = ctx->h_fx.p;
    if (f_dyn)
      for (size_t k = 0; k < 3 * (size_t)nd; ++k) f_dyn[k] = (double)F[k] * DYN_FX_INV_SCALE;
    if (f_water) {
      // ForcesOnWaterMol order: f_o, f_h0, f_h1, f_m; device blocks O, M, H0, H1.
      const int block_of_field[4] = {0, 2, 3, 1};
      for (uint32_t m = 0; m < nw; ++m)
        for (int fld = 0; fld < 4; ++fld) {
          const int64_t* src = F + 3 * ((size_t)nd + (size_t)block_of_field[fld] * nw + m);
          for (int k = 0; k < 3; ++k) f_water[12 * (size_t)m + 3 * fld + k] = (double)src[k] * DYN_FX_INV_SCALE;
        }
    }
    if (virial) *virial = w;
    if (energy) *energy = e;
    if (dh_dl) *dh_dl = ctx->any_alch ? d : 0.0;
    ctx->computed = true;
  });
}

extern "C" int dyn_nb_get_fixed_forces(dyn_nb_ctx* ctx, int64_t* out, uint64_t n_values) {
  if (!ctx) {
    set_global_error("null context");
    return DYN_ERR_ARG;
  }
  DYN_API_BODY(ctx, {
    if (!ctx->computed) throw Error(DYN_ERR_STATE, "no compute yet");
    DYN_REQUIRE(out, "out is null");
    DYN_REQUIRE(n_values >= 3ull * ctx->n_sites, "need %llu values",
                (unsigned long long)(3ull * ctx->n_sites));
    memcpy(out, ctx->h_fx.p, sizeof(int64_t) * 3 * ctx->n_sites);
  });
}

extern "C" int dyn_nb_correct_14(dyn_nb_ctx* ctx, uint32_t n14, const uint32_t* i14,
                                 const uint32_t* j14, const float* pos_dyn, const float* q_dyn,
                                 const float cell_extent[3], double* f_dyn, double* virial) {
  if (!ctx) {
    set_global_error("null context");
    return DYN_ERR_ARG;
  }
  DYN_API_BODY(ctx, {
    DeviceGuard g(ctx->device);
    const uint32_t nd = ctx->n_dyn;
    if (n14 == 0) return DYN_OK;
    DYN_REQUIRE(i14 && j14 && pos_dyn && q_dyn && cell_extent && f_dyn, "null argument");
    for (uint32_t k = 0; k < n14; ++k)
      DYN_REQUIRE(i14[k] < nd && j14[k] < nd, "1-4 pair %u index out of range", k);
    for (int k = 0; k < 3; ++k) DYN_REQUIRE(cell_extent[k] > 0.f, "cell extent must be > 0");
    ctx->c14_i.reserve(n14);
    ctx->c14_j.reserve(n14);
    ctx->c14_pos.reserve(3 * (size_t)nd);
    ctx->c14_q.reserve(nd);
    ctx->c14_fx.reserve(3 * (size_t)nd);
    DYN_HIP(hipMemcpyAsync(ctx->c14_i.p, i14, 4 * (size_t)n14, hipMemcpyHostToDevice, ctx->stream));
    DYN_HIP(hipMemcpyAsync(ctx->c14_j.p, j14, 4 * (size_t)n14, hipMemcpyHostToDevice, ctx->stream));
    DYN_HIP(hipMemcpyAsync(ctx->c14_pos.p, pos_dyn, 12 * (size_t)nd, hipMemcpyHostToDevice, ctx->stream));
    DYN_HIP(hipMemcpyAsync(ctx->c14_q.p, q_dyn, 4 * (size_t)nd, hipMemcpyHostToDevice, ctx->stream));
    DYN_HIP(hipMemsetAsync(ctx->c14_fx.p, 0, ctx->c14_fx.bytes(), ctx->stream));
    const DynRefParams P = make_ref_params(ctx, cell_extent, false, false, 0.0);
    const unsigned grid = grid_for(n14, kRefBlock, 1024);
    DevBuf<double> part;
    part.reserve(grid);
    hipLaunchKernelGGL(k_corr14_vacuum, dim3(grid), dim3(kRefBlock), 0, ctx->stream,
                       ctx->c14_i.p, ctx->c14_j.p, n14, ctx->c14_pos.p, ctx->c14_q.p, P,
                       ctx->c14_fx.p, part.p);
    DYN_CHECK_LAUNCH("k_corr14_vacuum");
    std::vector<int64_t> hf(3 * (size_t)nd);
    std::vector<double> hp(grid);
    DYN_HIP(hipMemcpyAsync(hf.data(), ctx->c14_fx.p, 8 * hf.size(), hipMemcpyDeviceToHost, ctx->stream));
    DYN_HIP(hipMemcpyAsync(hp.data(), part.p, 8 * (size_t)grid, hipMemcpyDeviceToHost, ctx->stream));
    DYN_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t k = 0; k < hf.size(); ++k) f_dyn[k] += (double)hf[k] * DYN_FX_INV_SCALE;
    double w = 0;
    for (unsigned b = 0; b < grid; ++b) w += hp[b];
    if (virial) *virial += w;
  });
}

// dyn_nb_correct_14 without any allocation: the same kernel and the same
// finishing arithmetic, with inputs, sums and partials staged in this
// device's `scratch` (through its looked-up address).  Calls on one device
// share that memory; overlapping calls there race.
extern "C" int dyn_nb_correct_14_noalloc(dyn_nb_ctx* ctx, uint32_t n14, const uint32_t* i14,
                                         const uint32_t* j14, const float* pos_dyn,
                                         const float* q_dyn, const float cell_extent[3],
                                         double* f_dyn, double* virial) {
  if (!ctx) {
    set_global_error("null context");
    return DYN_ERR_ARG;
  }
  DYN_API_BODY(ctx, {
    DeviceGuard g(ctx->device);
    const uint32_t nd = ctx->n_dyn;
    if (n14 == 0) return DYN_OK;
    DYN_REQUIRE(i14 && j14 && pos_dyn && q_dyn && cell_extent && f_dyn, "null argument");
    if (nd > kC14MaxAtoms || n14 > kC14MaxPairs)
      throw Error(DYN_ERR_LIMIT,
                  fmt("the allocation-free 1-4 correction takes at most %u solute atoms and %u "
                      "pairs (got %u atoms, %u pairs); use dyn_nb_correct_14",
                      kC14MaxAtoms, kC14MaxPairs, nd, n14));
    for (uint32_t k = 0; k < n14; ++k)
      DYN_REQUIRE(i14[k] < nd && j14[k] < nd, "1-4 pair %u index out of range", k);
    for (int k = 0; k < 3; ++k) DYN_REQUIRE(cell_extent[k] > 0.f, "cell extent must be > 0");
    unsigned char* base = c14_scratch_ptr(ctx->device);
    int64_t* fx = reinterpret_cast<int64_t*>(base + kC14OffFx);
    double* f = reinterpret_cast<double*>(base + kC14OffF);
    double* part = reinterpret_cast<double*>(base + kC14OffPart);
    float* pos = reinterpret_cast<float*>(base + kC14OffPos);
    float* q = reinterpret_cast<float*>(base + kC14OffQ);
    uint32_t* ii = reinterpret_cast<uint32_t*>(base + kC14OffI);
    uint32_t* jj = reinterpret_cast<uint32_t*>(base + kC14OffJ);
    hipStream_t s = ctx->stream;
    DYN_HIP(hipMemcpyAsync(ii, i14, 4 * (size_t)n14, hipMemcpyHostToDevice, s));
    DYN_HIP(hipMemcpyAsync(jj, j14, 4 * (size_t)n14, hipMemcpyHostToDevice, s));
    DYN_HIP(hipMemcpyAsync(pos, pos_dyn, 12 * (size_t)nd, hipMemcpyHostToDevice, s));
    DYN_HIP(hipMemcpyAsync(q, q_dyn, 4 * (size_t)nd, hipMemcpyHostToDevice, s));
    DYN_HIP(hipMemcpyAsync(f, f_dyn, 24 * (size_t)nd, hipMemcpyHostToDevice, s));
    const size_t n3 = 3 * (size_t)nd;
    const unsigned gz = grid_for(n3, 256, 1024);
    hipLaunchKernelGGL(k_fill_zero_i64, dim3(gz), dim3(256), 0, s, fx, n3);
    DYN_CHECK_LAUNCH("k_fill_zero_i64");
    const DynRefParams P = make_ref_params(ctx, cell_extent, false, false, 0.0);
    const unsigned grid = grid_for(n14, kRefBlock, kC14MaxBlocks);
    hipLaunchKernelGGL(k_corr14_vacuum, dim3(grid), dim3(kRefBlock), 0, s, ii, jj, n14, pos, q, P,
                       fx, part);
    DYN_CHECK_LAUNCH("k_corr14_vacuum");
    hipLaunchKernelGGL(k_c14_finish, dim3(gz), dim3(256), 0, s, fx, f, n3, part, grid,
                       part + kC14MaxBlocks);
    DYN_CHECK_LAUNCH("k_c14_finish");
    double w = 0.0;
    DYN_HIP(hipMemcpyAsync(f_dyn, f, 24 * (size_t)nd, hipMemcpyDeviceToHost, s));
    DYN_HIP(hipMemcpyAsync(&w, part + kC14MaxBlocks, 8, hipMemcpyDeviceToHost, s));
    DYN_HIP(hipStreamSynchronize(s));
    if (virial) *virial += w;
  });
}

extern "C" int dyn_nb_build_neighbors(dyn_nb_ctx* ctx, const float* posits_outer,
                                      uint32_t n_outer, const float* posits_inner,
                                      uint32_t n_inner, const uint8_t* is_static,
                                      const float cell_extent[3], uint8_t symmetric,
                                      float skin_sq_w_cutoff, uint32_t* row_ptr,
                                      uint32_t* col_idx, uint64_t col_capacity,
                                      uint64_t* n_total) {
  if (!ctx) {
    set_global_error("null context");
    return DYN_ERR_ARG;
  }
  DYN_API_BODY(ctx, {
    DeviceGuard g(ctx->device);
    DYN_REQUIRE(row_ptr && n_total && cell_extent, "null argument");
    DYN_REQUIRE(!(is_static && !symmetric),
                "Invalid neighbor build config; can't pass static indices if non-symmetric.");
    if (symmetric) DYN_REQUIRE(n_inner == n_outer, "symmetric=true requires identical sets");
    DYN_REQUIRE(skin_sq_w_cutoff > 0.f && std::isfinite(skin_sq_w_cutoff),
                "skin_sq_w_cutoff must be positive");
    for (int k = 0; k < 3; ++k)
      DYN_REQUIRE(cell_extent[k] > 0.f && std::isfinite(cell_extent[k]), "bad cell extent");
    *n_total = 0;
    if (n_outer == 0) {
      row_ptr[0] = 0;
      return DYN_OK;
    }
    if (n_inner == 0) {
      memset(row_ptr, 0, sizeof(uint32_t) * (n_outer + 1));
      return DYN_OK;
    }
    DYN_REQUIRE(posits_outer && posits_inner, "null positions");
    // Cell grid over the box with cell width ≥ r_list (+ margin for the f32
    // rounding of the exact test); candidates come from the 27 neighbour
    // cells, or every cell along an axis with fewer than 3 cells.
    const double r_list = std::sqrt((double)skin_sq_w_cutoff) * (1.0 + 1e-5) + 1e-3;
    NblGrid G;
    int use_nb[3];
    for (int k = 0; k < 3; ++k) {
      G.ext[k] = cell_extent[k];
      int nc = (int)std::floor(cell_extent[k] / r_list);
      if (nc < 3) {
        nc = 1;
        use_nb[k] = 0;
      } else {
        use_nb[k] = 1;
      }
      nc = std::min(nc, 128);
      G.nc[k] = nc;
      G.cw[k] = cell_extent[k] / nc;
      G.lo[k] = 0.0;
    }
    const uint32_t ncell = (uint32_t)(G.nc[0] * G.nc[1] * G.nc[2]);
    ctx->nb_outer.reserve(3 * (size_t)n_outer);
    ctx->nb_inner.reserve(3 * (size_t)n_inner);
    DYN_HIP(hipMemcpyAsync(ctx->nb_outer.p, posits_outer, 12 * (size_t)n_outer,
                           hipMemcpyHostToDevice, ctx->stream));
    DYN_HIP(hipMemcpyAsync(ctx->nb_inner.p, posits_inner, 12 * (size_t)n_inner,
                           hipMemcpyHostToDevice, ctx->stream));
    if (is_static) {
      ctx->nb_static.reserve(n_outer);
      DYN_HIP(hipMemcpyAsync(ctx->nb_static.p, is_static, n_outer, hipMemcpyHostToDevice,
                             ctx->stream));
    }
    ctx->nb_cell_of.reserve(n_inner);
    ctx->nb_counts.reserve(ncell);
    ctx->nb_cell_start.reserve(ncell + 1);
    ctx->nb_cursor.reserve(ncell);
    ctx->nb_items.reserve(n_inner);
    ctx->nb_row_count.reserve(n_outer);
    ctx->nb_row_ptr.reserve((size_t)n_outer + 1);
    ctx->nb_max.reserve(2);
    DYN_HIP(hipMemsetAsync(ctx->nb_counts.p, 0, 4 * (size_t)ncell, ctx->stream));
    DYN_HIP(hipMemsetAsync(ctx->nb_cursor.p, 0, 4 * (size_t)ncell, ctx->stream));
    hipLaunchKernelGGL(k_nbl_cell_count, dim3(grid_for(n_inner, 256)), dim3(256), 0,
                       ctx->stream, ctx->nb_inner.p, n_inner, G, ctx->nb_cell_of.p,
                       ctx->nb_counts.p);
    DYN_CHECK_LAUNCH("k_nbl_cell_count");
    hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(kScanBlock), 0, ctx->stream, ctx->nb_counts.p,
                       ncell, ctx->nb_cell_start.p, ctx->nb_max.p);
    DYN_CHECK_LAUNCH("k_scan_u32");
    hipLaunchKernelGGL(k_nbl_cell_fill, dim3(grid_for(n_inner, 256)), dim3(256), 0,
                       ctx->stream, n_inner, ctx->nb_cell_of.p, ctx->nb_cell_start.p,
                       ctx->nb_cursor.p, ctx->nb_items.p);
    DYN_CHECK_LAUNCH("k_nbl_cell_fill");
    NblRowArgs A;
    A.outer = ctx->nb_outer.p;
    A.inner = ctx->nb_inner.p;
    A.is_static = is_static ? ctx->nb_static.p : nullptr;
    A.cell_start = ctx->nb_cell_start.p;
    A.items = ctx->nb_items.p;
    A.G = G;
    for (int k = 0; k < 3; ++k) {
      A.ext[k] = cell_extent[k];
      A.use_nb[k] = use_nb[k];
    }
    A.thr = skin_sq_w_cutoff;
    A.symmetric = symmetric ? 1 : 0;
    A.offs = 1;
    hipLaunchKernelGGL(k_nbl_rows<0>, dim3(n_outer), dim3(kNblBlock), 0, ctx->stream, A,
                       n_outer, ctx->nb_row_count.p, (const uint32_t*)nullptr,
                       (uint32_t*)nullptr);
    DYN_CHECK_LAUNCH("k_nbl_rows<0>");
    hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(kScanBlock), 0, ctx->stream,
                       ctx->nb_row_count.p, n_outer, ctx->nb_row_ptr.p, ctx->nb_max.p);
    DYN_CHECK_LAUNCH("k_scan_u32");
    uint32_t maxrow = 0;
    DYN_HIP(hipMemcpyAsync(row_ptr, ctx->nb_row_ptr.p, 4 * ((size_t)n_outer + 1),
                           hipMemcpyDeviceToHost, ctx->stream));
    DYN_HIP(hipMemcpyAsync(&maxrow, ctx->nb_max.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    DYN_HIP(hipStreamSynchronize(ctx->stream));
    const uint64_t total = row_ptr[n_outer];
    *n_total = total;
    if (maxrow > (uint32_t)kNblCap)
      throw Error(DYN_ERR_CAPACITY,
                  fmt("a row has %u neighbors; the build handles at most %d", maxrow, kNblCap));
    if (total > col_capacity)
      throw Error(DYN_ERR_CAPACITY, fmt("col_idx needs %llu entries, capacity %llu",
                                        (unsigned long long)total,
                                        (unsigned long long)col_capacity));
    if (total == 0) return DYN_OK;
    DYN_REQUIRE(col_idx, "col_idx is null");
    ctx->nb_col.reserve(total);
    hipLaunchKernelGGL(k_nbl_rows<1>, dim3(n_outer), dim3(kNblBlock), 0, ctx->stream, A,
                       n_outer, (uint32_t*)nullptr, ctx->nb_row_ptr.p, ctx->nb_col.p);
    DYN_CHECK_LAUNCH("k_nbl_rows<1>");
    DYN_HIP(hipMemcpyAsync(col_idx, ctx->nb_col.p, 4 * total, hipMemcpyDeviceToHost,
                           ctx->stream));
    DYN_HIP(hipStreamSynchronize(ctx->stream));
  });
}
