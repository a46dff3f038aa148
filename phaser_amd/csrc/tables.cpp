// tables.cpp -- FP64 construction of the kernels' constant tables.
//
// All tables are computed in double with three-term recurrences and stored
// as float.  References: spherical-sampler.cc:29-64 (query grid),
// makeweights.c:68-100 (quadrature weights), s2_primitive.c:48-75,185-208
// (L2-normed Legendre), makeWigner.c:180-210,324-387 (L2-normed Wigner d),
// laplace-pyramid.cc:41-104 and pyramid-struct.cc:17-29 (fusion structure),
// phase-aligner.cc:27-30 (voxel edges, Eigen LinSpaced >= 3.3.4).
#include "tables.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <thread>

namespace phz {

int seanindex(int m, int l, int bw) {
  const int L = bw - 1;
  if (m >= 0) return m * (L + 1) - ((m * (m - 1)) / 2) + (l - m);
  return ((L * (L + 3)) / 2) + 1 + ((L + m) * (L + m + 1) / 2) + (l - std::abs(m));
}

std::vector<float> twiddles(int L) {
  std::vector<float> t(2 * (size_t)L);
  for (int e = 0; e < L; ++e) {
    const double a = -2.0 * M_PI * (double)e / (double)L;
    t[2 * e] = (float)std::cos(a);
    t[2 * e + 1] = (float)std::sin(a);
  }
  return t;
}

static void weights(int bw, std::vector<double>* w) {
  w->assign(2 * (size_t)bw, 0.0);
  const double fudge = M_PI / (double)(4 * bw);
  for (int j = 0; j < 2 * bw; ++j) {
    double s = 0.0;
    for (int k = 0; k < bw; ++k) s += std::sin((double)((2 * j + 1) * (2 * k + 1)) * fudge) / (double)(2 * k + 1);
    s *= std::sin((double)(2 * j + 1) * fudge) * 2.0 / (double)bw;
    (*w)[j] = s;
  }
}

// L2-normed P~_l^m(cos theta) for l = m..bw-1 by the three-term recurrence
static void legendre_column(int bw, int m, double theta, double* out) {
  double pmm;
  if (m == 0) {
    pmm = std::sqrt(0.5);
  } else {
    const double md = m;
    double c = std::sqrt(md + 0.5);
    for (int i = 0; i < m; ++i) c *= std::sqrt((md - i / 2.0) / (md - i));
    c *= std::pow(2.0, -md / 2.0);
    if (m & 1) c = -c;
    pmm = c * std::pow(std::sin(theta), md);
  }
  const double x = std::cos(theta);
  double pp = 0.0, p = pmm;
  out[0] = p;
  for (int l = m; l < bw - 1; ++l) {
    const double an = std::sqrt((2.0 * l + 3) / (2.0 * l + 1) * (l - m + 1.0) / (l + m + 1.0)) *
                      (2.0 * l + 1) / (l - m + 1.0);
    const double cn = (l == 0) ? 0.0
                               : -std::sqrt((2.0 * l + 3) / (2.0 * l - 1) * (l - m + 1.0) / (l + m + 1.0) *
                                            (double)(l - m) / (double)(l + m)) *
                                     (double)(l + m) / (l - m + 1.0);
    const double nx = an * x * p + cn * pp;
    pp = p;
    p = nx;
    out[l - m + 1] = p;
  }
}

// d~^l_{m1,m2}(beta) for l = max(|m1|,|m2|)..bw-1 (any signs)
static void wigner_column(int bw, int m1, int m2, double beta, double* out) {
  const int a1 = std::abs(m1), a2 = std::abs(m2);
  const int l = a1 > a2 ? a1 : a2;
  const int delta = l - (a1 < a2 ? a1 : a2);
  double norm = std::sqrt((2.0 * l + 1.0) / 2.0);
  for (int i = 0; i < delta; ++i) norm *= std::sqrt((2.0 * l - i) / (i + 1.0));
  int cp, sp;
  double sign = 1.0;
  if (l == a1) {
    if (m1 >= 0) {
      cp = l + m2;
      sp = l - m2;
      if ((l - m2) & 1) sign = -1.0;
    } else {
      cp = l - m2;
      sp = l + m2;
    }
  } else if (m2 >= 0) {
    cp = l + m1;
    sp = l - m1;
  } else {
    cp = l - m1;
    sp = l + m1;
    if ((l + m1) & 1) sign = -1.0;
  }
  const double s2 = std::sin(0.5 * beta), c2 = std::cos(0.5 * beta), cb = std::cos(beta);
  double pp = 0.0, p = norm * sign * std::pow(s2, (double)sp) * std::pow(c2, (double)cp);
  out[0] = p;
  for (int j = l; j < bw - 1; ++j) {
    const double dj = j, d1 = m1, d2 = m2;
    const double t5 = 1.0 / std::sqrt(((dj + 1) * (dj + 1) - d1 * d1) * ((dj + 1) * (dj + 1) - d2 * d2));
    const double bn = std::sqrt((2 * dj + 3) / (2 * dj + 1)) * (dj + 1) * (2 * dj + 1) * t5;
    double an = 0.0, cn = 0.0;
    if (j != 0) {
      an = -std::sqrt((2 * dj + 3) / (2 * dj - 1)) * (dj + 1) / dj *
           std::sqrt((dj * dj - d1 * d1) * (dj * dj - d2 * d2)) * t5;
      cn = -d1 * d2 / (dj * (dj + 1));
    }
    const double nx = an * pp + bn * (cb + cn) * p;
    pp = p;
    p = nx;
    out[j - l + 1] = p;
  }
}

void pyramid_bounds(uint32_t n, int levels, float div, uint32_t* coeffs, uint32_t* lower,
                    uint32_t* upper) {
  for (int lvl = 0; lvl < levels; ++lvl) {
    const uint32_t cur = lvl > 0 ? coeffs[lvl - 1] : n;
    const float lb = (float)cur / div;
    upper[lvl] = (uint32_t)std::round((float)cur - lb);
    lower[lvl] = (uint32_t)std::round(lb);
    coeffs[lvl] = upper[lvl] - lower[lvl];
  }
}

namespace {
struct Run {
  uint64_t a, b;  // [a, b)
  int mode;
};

// modes over [0, n_lvl) of level lvl, offset by `base` in level-0 indices
void level_runs(int lvl, int levels, uint64_t base, const uint32_t* cf, const uint32_t* lo,
                uint64_t n_lvl, float div, std::vector<Run>* out) {
  if (lvl == levels) {  // the averaged low pass
    out->push_back({base, base + n_lvl, 1});
    return;
  }
  const uint64_t lb = lo[lvl], ub = lo[lvl] + cf[lvl];
  // expand (laplace-pyramid.cc:41-58) writes recon into [lbe, n - lbe)
  const uint64_t lbe = (uint64_t)std::round((float)n_lvl / div);
  const uint64_t width = n_lvl >= 2 * lbe ? n_lvl - 2 * lbe : 0;
  const uint64_t w = width < cf[lvl] ? width : cf[lvl];
  if (lb > 0) out->push_back({base, base + lb, 0});
  std::vector<Run> inner;
  level_runs(lvl + 1, levels, 0, cf, lo, cf[lvl], div, &inner);
  for (const Run& r : inner) {
    const uint64_t a = r.a, b = r.b < w ? r.b : w;
    if (a < b) out->push_back({base + lbe + a, base + lbe + b, r.mode});
  }
  if (lbe + w < ub) out->push_back({base + lbe + w, base + ub, 2});
  if (ub < n_lvl) out->push_back({base + ub, base + n_lvl, 0});
}
}  // namespace

bool fusion_runs(uint64_t n, uint64_t limit, FusionRuns* res) {
  const int levels = 5;
  const float div = 4.0f;
  uint32_t cf[levels], lo[levels], hi[levels];
  pyramid_bounds((uint32_t)n, levels, div, cf, lo, hi);
  std::vector<Run> shifted;
  level_runs(0, levels, 0, cf, lo, n, div, &shifted);
  // unshifted index q reads the mode at shifted position (q + floor(n/2)) mod n
  const uint64_t fl = n / 2;
  std::vector<Run> uns;
  for (const Run& r : shifted) {
    // shifted p in [a,b) -> q = (p - fl) mod n
    uint64_t a = (r.a + n - fl) % n, b = a + (r.b - r.a);
    if (b <= n) {
      uns.push_back({a, b, r.mode});
    } else {
      uns.push_back({a, n, r.mode});
      uns.push_back({0, b - n, r.mode});
    }
  }
  // sort by start and merge
  for (size_t i = 0; i < uns.size(); ++i)
    for (size_t j = i + 1; j < uns.size(); ++j)
      if (uns[j].a < uns[i].a) std::swap(uns[i], uns[j]);
  res->count = 0;
  for (const Run& r : uns) {
    if (r.a >= limit) break;
    if (res->count > 0 && (int)res->mode[res->count - 1] == r.mode) continue;
    if (res->count >= FusionRuns::kMax) return false;
    res->start[res->count] = (uint32_t)r.a;
    res->mode[res->count] = (uint32_t)r.mode;
    ++res->count;
  }
  return true;
}

void build_s2_tables(int bw, S2Tables* t) {
  const int n = 2 * bw;
  t->bw = bw;
  t->n = n;
  // query grid exactly as spherical-sampler.cc (float angles, float sin/cos)
  t->grid.assign(4 * (size_t)n * n, 0.0f);
  t->theta.resize(n);
  for (int i = 0; i < n; ++i) {
    const float th = (float)((M_PI * (double)(2 * i + 1)) / (double)(4 * bw));
    t->theta[i] = th;
    for (int j = 0; j < n; ++j) {
      const float ph = (float)((2 * M_PI * (double)j) / (double)(2 * bw));
      volatile float st = sinf(th), ct = cosf(th), sp = sinf(ph), cp = cosf(ph);
      const float one = 1.0f;
      volatile float x = one * st;
      x = x * cp;
      volatile float y = one * st;
      y = y * sp;
      volatile float z = one * ct;
      float* g = &t->grid[4 * ((size_t)i * n + j)];
      g[0] = x;
      g[1] = y;
      g[2] = z;
    }
  }
  // weighted Legendre table [m][i][l-m]
  std::vector<double> w;
  weights(bw, &w);
  t->leg_off.assign(bw + 1, 0);
  for (int m = 0; m < bw; ++m) t->leg_off[m + 1] = t->leg_off[m] + n * (bw - m);
  t->legendre.assign((size_t)t->leg_off[bw], 0.0f);
  std::vector<double> col(bw);
  for (int m = 0; m < bw; ++m)
    for (int i = 0; i < n; ++i) {
      const double th = M_PI * (2.0 * i + 1.0) / (4.0 * bw);
      legendre_column(bw, m, th, col.data());
      for (int l = m; l < bw; ++l)
        t->legendre[(size_t)t->leg_off[m] + (size_t)i * (bw - m) + (l - m)] = (float)(w[i] * col[l - m]);
    }
  // Wigner blocks a <= b, [l-b][k]
  const int nblk = bw * (bw + 1) / 2;
  t->wig_off.assign(nblk + 1, 0);
  for (int b = 0; b < bw; ++b)
    for (int a = 0; a <= b; ++a) {
      const int id = b * (b + 1) / 2 + a;
      t->wig_off[id + 1] = t->wig_off[id] + (int64_t)n * (bw - b);
    }
  t->wigner.assign((size_t)t->wig_off[nblk], 0.0f);
  {  // blocks are independent: spread the orders b over host threads
    std::atomic<int> next_b{0};
    auto work = [&]() {
      std::vector<double> cb(bw);
      for (int b; (b = next_b.fetch_add(1)) < bw;)
        for (int a = 0; a <= b; ++a) {
          const int id = b * (b + 1) / 2 + a;
          float* dst = &t->wigner[(size_t)t->wig_off[id]];
          for (int k = 0; k < n; ++k) {
            const double beta = M_PI * (2.0 * k + 1.0) / (4.0 * bw);
            wigner_column(bw, a, b, beta, cb.data());
            for (int l = b; l < bw; ++l) dst[(size_t)(l - b) * n + k] = (float)cb[l - b];
          }
        }
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nth = (int)std::min<unsigned>(std::min(hw, 16u), (unsigned)std::max(1, bw / 8));
    std::vector<std::thread> pool;
    for (int i = 1; i < nth; ++i) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
  }
  t->wig_norm.resize(bw);
  for (int l = 0; l < bw; ++l) t->wig_norm[l] = (float)(2. * M_PI * std::sqrt(2. / (2. * l + 1.)));
  t->tw_n = twiddles(n);
  fusion_runs((uint64_t)bw * bw, (uint64_t)bw * bw, &t->runs);
}

void build_voxel_tables(int N, float lo, float hi, VoxelTables* t) {
  t->N = N;
  t->lo = lo;
  t->hi = hi;
  t->edges.resize(N);
  if (N == 1) {
    t->edges[0] = hi;
  } else {
    const float step = (hi - lo) / (float)(N - 1);
    const bool flip = std::fabs(hi) < std::fabs(lo);
    for (int i = 0; i < N; ++i) {
      volatile float prod = flip ? (float)(N - 1 - i) * step : (float)i * step;
      if (flip)
        t->edges[i] = (i == 0) ? lo : hi - prod;
      else
        t->edges[i] = (i == N - 1) ? hi : lo + prod;
    }
  }
  t->tw_N = twiddles(N);
  const uint64_t T = (uint64_t)N * N * N;
  const uint64_t H = (uint64_t)N * N * (N / 2 + 1);
  fusion_runs(T, H, &t->runs);
}

}  // namespace phz
