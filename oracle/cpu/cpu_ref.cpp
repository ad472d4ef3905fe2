// cpu_ref.cpp — C++ restatement of the reference's CPU nonbonded path, used
// as the timed CPU baseline (test infrastructure only; never linked into the
// product).  One evaluation = what the reference does for one step at a
// fresh neighbor list:
//   build_all_neighbors   brute force over solute atoms and water O proxies
//                         (src/neighbors.rs:102-156, :180-281), f32 test
//   setup_pairs           flattened pair list (src/non_bonded.rs:541-677)
//   calc_force_cpu        per-thread f64 force buffers, f32 pair math of
//                         f_nonbonded_cpu + force_e_lj + the erfc short-range
//                         Coulomb, reduce (src/non_bonded.rs:258-412,
//                         :950-1057, src/forces.rs:10-21)
//   SPME                  B-spline order-n spreading, 3D FFT, influence
//                         function, gather, in f64 (the reference calls the
//                         `ewald` crate here, which is not available; this is
//                         the substitute, src/non_bonded.rs:681-795)
// Threads: std::thread, static chunks (the reference uses rayon).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

struct Input {  // one system, as the caller lays it out
  uint32_t n_std, n_wat;
  const float* pos_std;  // 3·n_std
  const float* q_std;
  const float* sig_std;
  const float* eps_std;
  const float* pos_w;    // n_wat × 4 sites (O, M, H0, H1) × 3
  uint32_t n_excl, n_14;
  const uint32_t* excl;  // pairs (i < j)
  const uint32_t* p14;
  float ext[3];
  float cutoff, lj_cutoff, skin, alpha;
  uint32_t K[3];
  uint32_t order;
};

namespace {

using cd = std::complex<double>;

struct V3 {
  float x, y, z;
};

inline float round_away(float v) { return std::round(v); }
inline V3 min_image(V3 d, const float ext[3]) {
  // src/barostat.rs:162-170
  d.x = d.x - round_away(d.x / ext[0]) * ext[0];
  d.y = d.y - round_away(d.y / ext[1]) * ext[1];
  d.z = d.z - round_away(d.z / ext[2]) * ext[2];
  return d;
}
inline float norm2(V3 d) { return d.x * d.x + d.y * d.y + d.z * d.z; }

void parallel_for(int threads, size_t n, const std::function<void(int, size_t, size_t)>& fn) {
  if (threads <= 1 || n < 2) {
    fn(0, 0, n);
    return;
  }
  std::vector<std::thread> ts;
  const size_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    ts.emplace_back(fn, t, b, e);
  }
  for (auto& t : ts) t.join();
}

// build_neighbors: symmetric lists come back full and ascending.
std::vector<std::vector<uint32_t>> build_neighbors(const std::vector<V3>& outer,
                                                   const std::vector<V3>& inner, bool symmetric,
                                                   const float ext[3], float thr, int threads) {
  const size_t no = outer.size(), ni = inner.size();
  std::vector<std::vector<uint32_t>> half(no);
  parallel_for(threads, no, [&](int, size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      const V3 p = outer[i];
      for (size_t j = symmetric ? i + 1 : 0; j < ni; ++j) {
        V3 d{p.x - inner[j].x, p.y - inner[j].y, p.z - inner[j].z};
        if (norm2(min_image(d, ext)) < thr) half[i].push_back((uint32_t)j);
      }
    }
  });
  if (!symmetric) return half;
  std::vector<std::vector<uint32_t>> full(no);
  for (size_t i = 0; i < no; ++i)
    for (uint32_t j : half[i]) {
      full[i].push_back(j);
      full[j].push_back((uint32_t)i);
    }
  return full;
}

struct Pair {
  // body: index < n_std → solute atom, else water site (mol·4 + site) + n_std
  uint32_t tgt, src;
  float sigma, eps;
  uint8_t scale14, calc_lj, calc_coul, involves_std;
};



constexpr float O_SIGMA = 1.777167268f * (2.0f / 1.122462048309373f);
constexpr float O_EPS = 0.2128008130f;
constexpr float Q_H = 0.6791f * 18.2223f;
constexpr float Q_M = -2.0f * Q_H;

inline uint64_t key(uint32_t a, uint32_t b) { return ((uint64_t)a << 32) | b; }

std::vector<Pair> setup_pairs(const Input& in, const std::vector<std::vector<uint32_t>>& ss,
                              const std::vector<std::vector<uint32_t>>& sw,
                              const std::vector<std::vector<uint32_t>>& ww,
                              const std::vector<uint64_t>& excl_sorted,
                              const std::vector<uint64_t>& s14_sorted) {
  std::vector<Pair> out;
  const uint32_t ns = in.n_std;
  for (uint32_t i = 0; i < ns; ++i)
    for (uint32_t j : ss[i]) {
      if (j <= i) continue;
      const uint64_t k = key(i, j);
      if (std::binary_search(excl_sorted.begin(), excl_sorted.end(), k)) continue;
      Pair p{};
      p.tgt = i;
      p.src = j;
      p.sigma = 0.5f * (in.sig_std[i] + in.sig_std[j]);
      p.eps = std::sqrt(in.eps_std[i] * in.eps_std[j]);
      p.scale14 = std::binary_search(s14_sorted.begin(), s14_sorted.end(), k);
      p.calc_lj = p.calc_coul = 1;
      p.involves_std = 1;
      out.push_back(p);
    }
  for (uint32_t i = 0; i < ns; ++i)
    for (uint32_t w : sw[i])
      for (uint32_t site = 0; site < 4; ++site) {
        Pair p{};
        p.tgt = i;
        p.src = ns + 4 * w + site;
        p.sigma = 0.5f * (in.sig_std[i] + O_SIGMA);
        p.eps = std::sqrt(in.eps_std[i] * O_EPS);
        p.calc_lj = site == 0;
        p.calc_coul = site != 0;
        p.involves_std = 1;
        out.push_back(p);
      }
  for (uint32_t a = 0; a < in.n_wat; ++a)
    for (uint32_t b : ww[a]) {
      if (b <= a) continue;
      for (uint32_t s0 = 0; s0 < 4; ++s0)
        for (uint32_t s1 = 0; s1 < 4; ++s1) {
          const bool lj = s0 == 0 && s1 == 0, c = s0 != 0 && s1 != 0;
          if (!(lj || c)) continue;
          Pair p{};
          p.tgt = ns + 4 * a + s0;
          p.src = ns + 4 * b + s1;
          p.sigma = O_SIGMA;
          p.eps = O_EPS;
          p.calc_lj = lj;
          p.calc_coul = c;
          out.push_back(p);
        }
    }
  return out;
}

// ---- small complex FFT (Stockham, radix 2/3/4/5), double ------------------------
void fft_line(cd* a, cd* b, int N, int sign) {
  int n = N, Ns = 1;
  std::vector<int> radices;
  while (n % 4 == 0) { radices.push_back(4); n /= 4; }
  while (n % 2 == 0) { radices.push_back(2); n /= 2; }
  while (n % 3 == 0) { radices.push_back(3); n /= 3; }
  while (n % 5 == 0) { radices.push_back(5); n /= 5; }
  cd* src = a;
  cd* dst = b;
  for (int R : radices) {
    const int NR = N / R;
    for (int j = 0; j < NR; ++j) {
      const int k = j % Ns;
      cd v[5];
      for (int r = 0; r < R; ++r) {
        const double ang = sign * 2.0 * M_PI * r * k / (double)(Ns * R);
        v[r] = src[j + r * NR] * cd(std::cos(ang), std::sin(ang));
      }
      cd y[5];
      for (int q = 0; q < R; ++q) {
        y[q] = 0;
        for (int r = 0; r < R; ++r) {
          const double ang = sign * 2.0 * M_PI * r * q / R;
          y[q] += v[r] * cd(std::cos(ang), std::sin(ang));
        }
      }
      const int d = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; ++r) dst[d + r * Ns] = y[r];
    }
    std::swap(src, dst);
    Ns *= R;
  }
  if (src != a) std::copy(src, src + N, a);
}

void fft3d(std::vector<cd>& g, const uint32_t K[3], int sign, int threads) {
  const int n0 = K[0], n1 = K[1], n2 = K[2];
  for (int axis = 2; axis >= 0; --axis) {
    const int N = (int)K[axis];
    const size_t lines = (size_t)n0 * n1 * n2 / N;
    parallel_for(threads, lines, [&](int, size_t b, size_t e) {
      std::vector<cd> A(N), B(N);
      for (size_t l = b; l < e; ++l) {
        size_t base, stride;
        if (axis == 2) { base = l * n2; stride = 1; }
        else if (axis == 1) { base = (l / n2) * (size_t)n1 * n2 + l % n2; stride = n2; }
        else { base = l; stride = (size_t)n1 * n2; }
        for (int i = 0; i < N; ++i) A[i] = g[base + i * stride];
        fft_line(A.data(), B.data(), N, sign);
        for (int i = 0; i < N; ++i) g[base + i * stride] = A[i];
      }
    });
  }
}

double Mn(double x, int n) {
  if (n == 2) return (x >= 0 && x <= 2) ? 1 - std::fabs(x - 1) : 0.0;
  return (x * Mn(x, n - 1) + (n - x) * Mn(x - 1, n - 1)) / (n - 1);
}

struct Spme {
  uint32_t K[3];
  float L[3];
  double beta;
  int order;
  std::vector<double> G, vf;
  void init(const Input& in) {
    for (int a = 0; a < 3; ++a) {
      K[a] = in.K[a];
      L[a] = in.ext[a];
    }
    beta = in.alpha;
    order = (int)in.order;
    std::vector<double> bsq[3];
    for (int a = 0; a < 3; ++a) {
      bsq[a].resize(K[a]);
      for (uint32_t m = 0; m < K[a]; ++m) {
        cd s = 0;
        for (int k = 0; k <= order - 2; ++k)
          s += Mn(k + 1.0, order) * std::exp(cd(0, 2 * M_PI * m * k / (double)K[a]));
        double md = std::norm(s);
        bsq[a][m] = md;
      }
      for (uint32_t m = 0; m < K[a]; ++m)
        if (bsq[a][m] < 1e-7) bsq[a][m] = 0.5 * (bsq[a][(m + K[a] - 1) % K[a]] + bsq[a][(m + 1) % K[a]]);
      for (auto& v : bsq[a]) v = 1.0 / v;
    }
    const double V = (double)L[0] * L[1] * L[2];
    G.assign((size_t)K[0] * K[1] * K[2], 0.0);
    vf.assign(G.size(), 0.0);
    for (uint32_t x = 0; x < K[0]; ++x)
      for (uint32_t y = 0; y < K[1]; ++y)
        for (uint32_t z = 0; z < K[2]; ++z) {
          const double fx = (x <= K[0] / 2 ? (int)x : (int)x - (int)K[0]) / (double)L[0];
          const double fy = (y <= K[1] / 2 ? (int)y : (int)y - (int)K[1]) / (double)L[1];
          const double fz = (z <= K[2] / 2 ? (int)z : (int)z - (int)K[2]) / (double)L[2];
          const double msq = fx * fx + fy * fy + fz * fz;
          const size_t i = ((size_t)x * K[1] + y) * K[2] + z;
          if (msq > 0) {
            G[i] = std::exp(-M_PI * M_PI * msq / (beta * beta)) / (M_PI * V * msq) * bsq[0][x] * bsq[1][y] * bsq[2][z];
            vf[i] = 1.0 - 2.0 * M_PI * M_PI * msq / (beta * beta);
          }
        }
  }
  // forces into f (3 per charge); returns energy (minus self) and virial
  void run(const std::vector<V3>& pos, const std::vector<float>& q, std::vector<double>& f,
           double* energy, double* virial, int threads) {
    const size_t n = pos.size();
    const int o = order;
    std::vector<cd> grid((size_t)K[0] * K[1] * K[2], 0.0);
    std::vector<int> k0(3 * n);
    std::vector<double> th(3 * n * o), dth(3 * n * o);
    for (size_t i = 0; i < n; ++i) {
      const float p[3] = {pos[i].x, pos[i].y, pos[i].z};
      for (int a = 0; a < 3; ++a) {
        double u = (double)p[a] / L[a] * K[a];
        u -= K[a] * std::floor(u / K[a]);
        const double fl = std::floor(u);
        const double w = u - fl;
        k0[3 * i + a] = (int)fl;
        for (int j = 0; j < o; ++j) {
          th[(3 * i + a) * o + j] = Mn(w + o - 1 - j, o);
          dth[(3 * i + a) * o + j] = Mn(w + o - 1 - j, o - 1) - Mn(w + o - 2 - j, o - 1);
        }
      }
      for (int jx = 0; jx < o; ++jx) {
        const int gx = (k0[3 * i] - o + 1 + jx + 2 * K[0]) % K[0];
        for (int jy = 0; jy < o; ++jy) {
          const int gy = (k0[3 * i + 1] - o + 1 + jy + 2 * K[1]) % K[1];
          for (int jz = 0; jz < o; ++jz) {
            const int gz = (k0[3 * i + 2] - o + 1 + jz + 2 * K[2]) % K[2];
            grid[((size_t)gx * K[1] + gy) * K[2] + gz] += (double)q[i] * th[(3 * i) * o + jx] *
                                                            th[(3 * i + 1) * o + jy] *
                                                            th[(3 * i + 2) * o + jz];
          }
        }
      }
    }
    fft3d(grid, K, -1, threads);
    double e = 0, w = 0;
    for (size_t i = 0; i < grid.size(); ++i) {
      const double en = 0.5 * G[i] * std::norm(grid[i]);
      e += en;
      w += en * vf[i];
      grid[i] *= G[i];
    }
    fft3d(grid, K, +1, threads);
    double q2 = 0;
    for (float v : q) q2 += (double)v * v;
    *energy = e - beta / std::sqrt(M_PI) * q2;
    *virial = w;
    f.assign(3 * n, 0.0);
    parallel_for(threads, n, [&](int, size_t b, size_t en_) {
      for (size_t i = b; i < en_; ++i) {
        double s[3] = {0, 0, 0};
        for (int jx = 0; jx < o; ++jx) {
          const int gx = (k0[3 * i] - o + 1 + jx + 2 * K[0]) % K[0];
          for (int jy = 0; jy < o; ++jy) {
            const int gy = (k0[3 * i + 1] - o + 1 + jy + 2 * K[1]) % K[1];
            for (int jz = 0; jz < o; ++jz) {
              const int gz = (k0[3 * i + 2] - o + 1 + jz + 2 * K[2]) % K[2];
              const double ph = grid[((size_t)gx * K[1] + gy) * K[2] + gz].real();
              const double tx = th[(3 * i) * o + jx], ty = th[(3 * i + 1) * o + jy], tz = th[(3 * i + 2) * o + jz];
              s[0] += ph * dth[(3 * i) * o + jx] * ty * tz;
              s[1] += ph * tx * dth[(3 * i + 1) * o + jy] * tz;
              s[2] += ph * tx * ty * dth[(3 * i + 2) * o + jz];
            }
          }
        }
        for (int a = 0; a < 3; ++a) f[3 * i + a] = -(double)q[i] * K[a] / L[a] * s[a];
      }
    });
  }
};

}  // namespace

extern "C" {

struct CpuRefStats {
  double us_neighbors, us_setup_pairs, us_short, us_spme;
  double energy_short, virial_short, energy_recip, virial_recip;
  uint64_t n_list_pairs, n_interacting;
  int steps;
};

// Evaluations over a trajectory at the reference's cadence: frame k (sites
// ordered solute, then per water O, M, H0, H1; 3 floats each) replaces the
// positions of step k.  The neighbor lists and the flattened pair list are
// rebuilt only when the largest displacement since the last build, of a
// solute atom (plain difference) or a water O (minimum image of the wrapped
// position), reaches half the skin (src/neighbors.rs:56-83, :159-169); SPME
// runs on steps divisible by spme_ratio and its forces, energy and virial
// are reused in between (src/lib.rs:1282-1341; with spme_ratio 1 every
// step).  With frames == NULL every step evaluates the input positions and
// rebuilds (the cost of one full evaluation).  rebuilds_out counts builds.
int cpuref_run_traj(const Input* in, int steps, const float* frames, int spme_ratio, int threads,
                    CpuRefStats* st, double* forces_out, int* rebuilds_out);

// Runs `steps` evaluations (neighbors + setup_pairs + short range + SPME each)
// and fills per-component wall times summed over the steps.  forces_out
// (optional) receives the last step's forces, sites ordered solute, then per
// water O, M, H0, H1.
int cpuref_run(const Input* in, int steps, int threads, CpuRefStats* st, double* forces_out) {
  return cpuref_run_traj(in, steps, nullptr, 1, threads, st, forces_out, nullptr);
}

int cpuref_run_traj(const Input* in, int steps, const float* frames, int spme_ratio, int threads,
                    CpuRefStats* st, double* forces_out, int* rebuilds_out) {
  using clk = std::chrono::steady_clock;
  auto us = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::micro>(b - a).count();
  };
  memset(st, 0, sizeof *st);
  const uint32_t ns = in->n_std, nw = in->n_wat;
  const size_t nsites = ns + 4 * (size_t)nw;
  std::vector<V3> pos(nsites);
  for (uint32_t i = 0; i < ns; ++i) pos[i] = {in->pos_std[3 * i], in->pos_std[3 * i + 1], in->pos_std[3 * i + 2]};
  for (size_t k = 0; k < 4 * (size_t)nw; ++k) pos[ns + k] = {in->pos_w[3 * k], in->pos_w[3 * k + 1], in->pos_w[3 * k + 2]};
  std::vector<float> qsite(nsites);
  for (uint32_t i = 0; i < ns; ++i) qsite[i] = in->q_std[i];
  for (uint32_t m = 0; m < nw; ++m) {
    qsite[ns + 4 * m] = 0.f;
    qsite[ns + 4 * m + 1] = Q_M;
    qsite[ns + 4 * m + 2] = Q_H;
    qsite[ns + 4 * m + 3] = Q_H;
  }
  std::vector<uint64_t> ex, s14;
  for (uint32_t k = 0; k < in->n_excl; ++k) ex.push_back(key(in->excl[2 * k], in->excl[2 * k + 1]));
  for (uint32_t k = 0; k < in->n_14; ++k) s14.push_back(key(in->p14[2 * k], in->p14[2 * k + 1]));
  std::sort(ex.begin(), ex.end());
  std::sort(s14.begin(), s14.end());
  for (auto k : s14) ex.erase(std::remove(ex.begin(), ex.end(), k), ex.end());
  Spme spme;
  spme.init(*in);
  const float rl = in->cutoff + in->skin;
  const float thr = rl * rl;
  const float two_b_sqrtpi = 2.0f * in->alpha * 0.5641895835477563f;
  std::vector<double> F(3 * nsites);
  std::vector<Pair> pairs;
  std::vector<V3> apos_ref, wpos_ref;  // positions at the last build
  const float half_skin_sq = (0.5f * in->skin) * (0.5f * in->skin);
  int rebuilds = 0;
  std::vector<double> f_rec_kept(3 * nsites, 0.0);
  double e_rec_kept = 0.0, w_rec_kept = 0.0;
  double e_rec = 0.0, w_rec = 0.0;
  const int ratio = std::max(1, spme_ratio);
  for (int step = 0; step < steps; ++step) {
    auto t0 = clk::now();
    if (frames) {
      const float* fr = frames + (size_t)step * 3 * nsites;
      for (size_t k = 0; k < nsites; ++k) pos[k] = {fr[3 * k], fr[3 * k + 1], fr[3 * k + 2]};
    }
    std::vector<V3> apos(pos.begin(), pos.begin() + ns);
    std::vector<V3> wpos(nw);
    for (uint32_t m = 0; m < nw; ++m) {
      V3 o = pos[ns + 4 * m];
      float* c[3] = {&o.x, &o.y, &o.z};
      for (int a = 0; a < 3; ++a) {
        float r = std::fmod(*c[a], in->ext[a]);
        if (r < 0) r += std::fabs(in->ext[a]);
        *c[a] = r;
      }
      wpos[m] = o;
    }
    bool rebuild = !frames || step == 0;
    if (!rebuild) {
      float mx = 0.f;
      for (uint32_t i = 0; i < ns; ++i) {
        const V3 d{apos[i].x - apos_ref[i].x, apos[i].y - apos_ref[i].y, apos[i].z - apos_ref[i].z};
        mx = std::max(mx, norm2(d));
      }
      for (uint32_t m = 0; m < nw; ++m) {
        const V3 d = min_image(V3{wpos[m].x - wpos_ref[m].x, wpos[m].y - wpos_ref[m].y,
                                  wpos[m].z - wpos_ref[m].z}, in->ext);
        mx = std::max(mx, norm2(d));
      }
      rebuild = mx >= half_skin_sq;
    }
    auto t1 = t0, t2 = t0;
    if (rebuild) {
      auto nss = build_neighbors(apos, apos, true, in->ext, thr, threads);
      auto nsw = build_neighbors(apos, wpos, false, in->ext, thr, threads);
      auto nww = build_neighbors(wpos, wpos, true, in->ext, thr, threads);
      t1 = clk::now();
      pairs = setup_pairs(*in, nss, nsw, nww, ex, s14);
      t2 = clk::now();
      apos_ref = apos;
      wpos_ref = wpos;
      ++rebuilds;
    }
    // calc_force_cpu
    const int T = std::max(1, threads);
    std::vector<std::vector<double>> bufs(T, std::vector<double>(3 * nsites, 0.0));
    std::vector<double> vir(T, 0.0), en(T, 0.0);
    std::vector<uint64_t> hits(T, 0);
    parallel_for(T, pairs.size(), [&](int t, size_t b, size_t e) {
      double* f = bufs[t].data();
      double w = 0, ee = 0;
      uint64_t h = 0;
      for (size_t k = b; k < e; ++k) {
        const Pair& p = pairs[k];
        const V3 a = pos[p.tgt], c = pos[p.src];
        const V3 diff = min_image(V3{a.x - c.x, a.y - c.y, a.z - c.z}, in->ext);
        const float d2 = norm2(diff);
        if (d2 < 1e-12f) continue;
        const float dist = std::sqrt(d2);
        const float inv = 1.0f / dist;
        const V3 dir{diff.x * inv, diff.y * inv, diff.z * inv};
        float mag = 0.f, e_p = 0.f;
        bool hit = false;
        if (p.calc_lj && dist <= in->lj_cutoff) {
          const float sr = p.sigma * inv, sr2 = sr * sr, sr4 = sr2 * sr2, sr6 = sr4 * sr2,
                      sr12 = sr6 * sr6;
          float m = 24.f * p.eps * std::fma(2.0f, sr12, -sr6) * inv;
          float el = 4.f * p.eps * (sr12 - sr6);
          if (p.scale14) {
            m *= 0.5f;
            el *= 0.5f;
          }
          mag += m;
          e_p += el;
          hit = true;
        }
        if (p.calc_coul && dist < in->cutoff) {
          const float qq = qsite[p.tgt] * qsite[p.src];
          const float ar = in->alpha * dist;
          const float ec = std::erfc(ar);
          const float ex = std::exp(-(ar * ar));
          float m = qq * std::fma(ec, inv * inv, two_b_sqrtpi * ex * inv);
          float el = qq * inv * ec;
          if (p.scale14) {
            m *= 1.0f / 1.2f;
            el *= 1.0f / 1.2f;
          }
          mag += m;
          e_p += el;
          hit = true;
        }
        const V3 fv{dir.x * mag, dir.y * mag, dir.z * mag};
        f[3 * p.tgt] += fv.x;
        f[3 * p.tgt + 1] += fv.y;
        f[3 * p.tgt + 2] += fv.z;
        f[3 * p.src] -= fv.x;
        f[3 * p.src + 1] -= fv.y;
        f[3 * p.src + 2] -= fv.z;
        w += (double)(diff.x * fv.x + diff.y * fv.y + diff.z * fv.z);
        if (p.involves_std) ee += e_p;
        h += hit;
      }
      vir[t] = w;
      en[t] = ee;
      hits[t] = h;
    });
    std::fill(F.begin(), F.end(), 0.0);
    for (int t = 0; t < T; ++t)
      for (size_t k = 0; k < F.size(); ++k) F[k] += bufs[t][k];
    double w_sr = 0, e_sr = 0;
    uint64_t nh = 0;
    for (int t = 0; t < T; ++t) {
      w_sr += vir[t];
      e_sr += en[t];
      nh += hits[t];
    }
    auto t3 = clk::now();
    if (step % ratio != 0) {
      // the reference's cached SPME forces, energy and virial
      for (size_t k = 0; k < F.size(); ++k) F[k] += f_rec_kept[k];
      e_rec = e_rec_kept;
      w_rec = w_rec_kept;
      auto t4 = clk::now();
      st->us_neighbors += us(t0, t1);
      st->us_setup_pairs += us(t1, t2);
      st->us_short += us(t2, t3);
      st->us_spme += us(t3, t4);
      st->energy_short = e_sr;
      st->virial_short = w_sr;
      st->energy_recip = e_rec;
      st->virial_recip = w_rec;
      st->n_list_pairs = pairs.size();
      st->n_interacting = nh;
      st->steps = step + 1;
      continue;
    }
    // SPME over charged sites, wrapped
    std::vector<V3> qpos;
    std::vector<float> qv;
    std::vector<size_t> qidx;
    for (size_t s = 0; s < nsites; ++s)
      if (qsite[s] != 0.f) {
        V3 p = pos[s];
        float* c[3] = {&p.x, &p.y, &p.z};
        for (int a = 0; a < 3; ++a) {
          float r = std::fmod(*c[a], in->ext[a]);
          if (r < 0) r += in->ext[a];
          *c[a] = r;
        }
        qpos.push_back(p);
        qv.push_back(qsite[s]);
        qidx.push_back(s);
      }
    std::vector<double> fr;
    spme.run(qpos, qv, fr, &e_rec, &w_rec, threads);
    std::vector<double> before;
    if (ratio > 1) before = F;
    for (size_t k = 0; k < qidx.size(); ++k)
      for (int a = 0; a < 3; ++a) F[3 * qidx[k] + a] += fr[3 * k + a];
    // 1-4 vacuum correction of handle_spme_recip (src/non_bonded.rs:763-790)
    for (uint32_t k = 0; k < in->n_14; ++k) {
      const uint32_t i = in->p14[2 * k], j = in->p14[2 * k + 1];
      const V3 d = min_image(V3{pos[i].x - pos[j].x, pos[i].y - pos[j].y, pos[i].z - pos[j].z}, in->ext);
      const float r = std::sqrt(norm2(d));
      if (std::fabs(r) < 1e-6f) continue;
      const float inv = 1.0f / r;
      const float m = qsite[i] * qsite[j] * inv * inv * (1.0f / 1.2f - 1.0f);
      const V3 df{d.x * inv * m, d.y * inv * m, d.z * inv * m};
      F[3 * i] += df.x;
      F[3 * i + 1] += df.y;
      F[3 * i + 2] += df.z;
      F[3 * j] -= df.x;
      F[3 * j + 1] -= df.y;
      F[3 * j + 2] -= df.z;
      w_rec += (double)(d.x * df.x + d.y * df.y + d.z * df.z);
    }
    if (ratio > 1) {
      // what the reference caches: SPME with its 1-4 correction
      for (size_t k = 0; k < F.size(); ++k) f_rec_kept[k] = F[k] - before[k];
      e_rec_kept = e_rec;
      w_rec_kept = w_rec;
    }
    auto t4 = clk::now();
    st->us_neighbors += us(t0, t1);
    st->us_setup_pairs += us(t1, t2);
    st->us_short += us(t2, t3);
    st->us_spme += us(t3, t4);
    st->energy_short = e_sr;
    st->virial_short = w_sr;
    st->energy_recip = e_rec;
    st->virial_recip = w_rec;
    st->n_list_pairs = pairs.size();
    st->n_interacting = nh;
    st->steps = step + 1;
  }
  if (forces_out) memcpy(forces_out, F.data(), sizeof(double) * F.size());
  if (rebuilds_out) *rebuilds_out = rebuilds;
  return 0;
}

}  // extern "C"
