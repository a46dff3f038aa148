/*
 * phaser_oracle.c -- FP64 CPU restatement of PHASER's correlation core.
 *
 * TEST INFRASTRUCTURE ONLY (see phaser_oracle.h).  Built with
 * -ffp-contract=off so that float expressions round exactly as the
 * reference's scalar C++ does.
 */
#include "phaser_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static void* xcalloc(size_t n, size_t sz) {
  void* p = calloc(n ? n : 1, sz);
  if (!p) abort();
  return p;
}

/* ======================================================================== */
/* FFT: recursive mixed-radix decimation in time with a twiddle table.       */
/* ======================================================================== */
typedef struct {
  double re, im;
} cplx;

static int smallest_factor(int n) {
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (int p = 3; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

/* out[0..n) = DFT of in[j*istride], j < n.  tw[e] = exp(sign 2 pi i e / N0),
 * and this level's n divides N0 with twiddle step ts = N0 / n. */
static void fft_rec(const cplx* in, int64_t istride, cplx* out, int n,
                    const cplx* tw, int ts, cplx* tmp) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  const int p = smallest_factor(n);
  const int m = n / p;
  for (int r = 0; r < p; ++r)
    fft_rec(in + r * istride, istride * p, out + (int64_t)r * m, m, tw,
            ts * p, tmp);
  const int N0 = n * ts;
  for (int k = 0; k < m; ++k) {
    for (int r = 0; r < p; ++r) {
      const cplx a = out[(int64_t)r * m + k];
      const cplx w = tw[((int64_t)r * k * ts) % N0];
      tmp[r].re = a.re * w.re - a.im * w.im;
      tmp[r].im = a.re * w.im + a.im * w.re;
    }
    for (int q = 0; q < p; ++q) {
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < p; ++r) {
        const cplx w = tw[((int64_t)r * q * m * ts) % N0];
        sr += tmp[r].re * w.re - tmp[r].im * w.im;
        si += tmp[r].re * w.im + tmp[r].im * w.re;
      }
      out[k + (int64_t)q * m].re = sr;
      out[k + (int64_t)q * m].im = si;
    }
  }
}

void po_fft(double* data, int n, int64_t stride, int sign) {
  if (n <= 1) return;
  cplx* in = (cplx*)xcalloc(n, sizeof(cplx));
  cplx* out = (cplx*)xcalloc(n, sizeof(cplx));
  cplx* tw = (cplx*)xcalloc(n, sizeof(cplx));
  cplx* tmp = (cplx*)xcalloc(n, sizeof(cplx));
  for (int e = 0; e < n; ++e) {
    const double a = 2.0 * M_PI * (double)e / (double)n;
    tw[e].re = cos(a);
    tw[e].im = sign * sin(a);
  }
  for (int j = 0; j < n; ++j) {
    in[j].re = data[2 * j * stride];
    in[j].im = data[2 * j * stride + 1];
  }
  fft_rec(in, 1, out, n, tw, 1, tmp);
  for (int j = 0; j < n; ++j) {
    data[2 * j * stride] = out[j].re;
    data[2 * j * stride + 1] = out[j].im;
  }
  free(in);
  free(out);
  free(tw);
  free(tmp);
}

/* ======================================================================== */
/* S^2 sampling                                                              */
/* ======================================================================== */
/* spherical-sampler.cc:29-64: theta and phi are computed in double, stored
 * as float (Point_t), and converted with float sin/cos. */
void po_sampler_grid(int bw, float* xyz) {
  const size_t n = 2 * (size_t)bw;
  size_t q = 0;
  for (size_t i = 0; i < n; ++i) {
    const float th = (float)((M_PI * (double)(2 * i + 1)) / (double)(4 * bw));
    for (size_t j = 0; j < n; ++j) {
      const float ph = (float)((2 * M_PI * (double)j) / (double)(2 * bw));
      const float dist = 1.0f;
      xyz[3 * q + 0] = dist * sinf(th) * cosf(ph);
      xyz[3 * q + 1] = dist * sinf(th) * sinf(ph);
      xyz[3 * q + 2] = dist * cosf(th);
      ++q;
    }
  }
}

/* spherical-projection.cc:19-41 (mixed float/double exactly as written). */
void po_project(const float* pts, int64_t P, float* proj, double* range) {
  for (int64_t i = 0; i < P; ++i) {
    const float x = pts[4 * i], y = pts[4 * i + 1], z = pts[4 * i + 2];
    const double dist_xy = (double)sqrtf(x * x + y * y);
    const double dist = sqrt(dist_xy * dist_xy + (double)(z * z));
    const float pz = (float)((double)z / dist);
    range[i] = dist;
    const double tmp_x = (double)acosf(pz);
    const double tmp_y = fmod((double)atan2f(y, x) + 2 * M_PI, 2 * M_PI);
    proj[3 * i + 0] = (float)(sin(tmp_x) * cos(tmp_y));
    proj[3 * i + 1] = (float)(sin(tmp_x) * sin(tmp_y));
    proj[3 * i + 2] = pz;
  }
}

/* FLANN L2_Simple<float>: ((0 + d0^2) + d1^2) + d2^2 in float. */
static inline float l2f(const float* a, const float* b) {
  const float d0 = a[0] - b[0];
  const float d1 = a[1] - b[1];
  const float d2 = a[2] - b[2];
  float r = 0.0f;
  r += d0 * d0;
  r += d1 * d1;
  r += d2 * d2;
  return r;
}

void po_nn_brute(const float* proj, int64_t P, const float* queries, int nq,
                 float max_sq_dist, int32_t* nn) {
  for (int q = 0; q < nq; ++q) {
    float best = INFINITY;
    int32_t bi = -1;
    for (int64_t i = 0; i < P; ++i) {
      const float d = l2f(queries + 3 * q, proj + 3 * i);
      if (d < best) {
        best = d;
        bi = (int32_t)i;
      }
    }
    nn[q] = (bi >= 0 && best <= max_sq_dist) ? bi : -1;
  }
}

/* ---- kd-tree with exact search (same answer as po_nn_brute) ---- */
typedef struct {
  const float* pts;
  int32_t* idx;
  int32_t* node_lo;  /* per node: range in idx */
  int32_t* node_hi;
  int32_t* node_left;
  int32_t* node_right;
  int8_t* node_dim;
  float* node_split;
  int n_nodes, cap;
} kdtree;

/* qsort comparator context: thread-local, so that registrations may run
 * concurrently on several host threads */
static _Thread_local int kd_cmp_dim;
static _Thread_local const float* kd_cmp_pts;
static int kd_cmp(const void* a, const void* b) {
  const float va = kd_cmp_pts[3 * *(const int32_t*)a + kd_cmp_dim];
  const float vb = kd_cmp_pts[3 * *(const int32_t*)b + kd_cmp_dim];
  return (va < vb) ? -1 : (va > vb) ? 1 : 0;
}

static int kd_build(kdtree* t, int lo, int hi) {
  const int id = t->n_nodes++;
  t->node_lo[id] = lo;
  t->node_hi[id] = hi;
  t->node_left[id] = t->node_right[id] = -1;
  if (hi - lo <= 8) return id;
  float mn[3] = {INFINITY, INFINITY, INFINITY};
  float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = lo; i < hi; ++i)
    for (int d = 0; d < 3; ++d) {
      const float v = t->pts[3 * t->idx[i] + d];
      if (v < mn[d]) mn[d] = v;
      if (v > mx[d]) mx[d] = v;
    }
  int dim = 0;
  for (int d = 1; d < 3; ++d)
    if (mx[d] - mn[d] > mx[dim] - mn[dim]) dim = d;
  kd_cmp_dim = dim;
  kd_cmp_pts = t->pts;
  qsort(t->idx + lo, hi - lo, sizeof(int32_t), kd_cmp);
  const int mid = (lo + hi) / 2;
  t->node_dim[id] = (int8_t)dim;
  t->node_split[id] = t->pts[3 * t->idx[mid] + dim];
  const int l = kd_build(t, lo, mid);
  const int r = kd_build(t, mid, hi);
  t->node_left[id] = l;
  t->node_right[id] = r;
  return id;
}

static void kd_search(const kdtree* t, int node, const float* q, float* best,
                      int32_t* bi) {
  if (t->node_left[node] < 0) {
    for (int i = t->node_lo[node]; i < t->node_hi[node]; ++i) {
      const int32_t p = t->idx[i];
      const float d = l2f(q, t->pts + 3 * p);
      if (d < *best || (d == *best && p < *bi)) {
        *best = d;
        *bi = p;
      }
    }
    return;
  }
  const int dim = t->node_dim[node];
  const double diff = (double)q[dim] - (double)t->node_split[node];
  const int first = diff < 0 ? t->node_left[node] : t->node_right[node];
  const int second = diff < 0 ? t->node_right[node] : t->node_left[node];
  kd_search(t, first, q, best, bi);
  /* conservative prune: float distances differ from exact ones by a few
   * ulps, so only skip when the plane is clearly farther than the best. */
  if (diff * diff <= (double)*best * (1.0 + 1e-5) + 1e-12)
    kd_search(t, second, q, best, bi);
}

void po_nn_kdtree(const float* proj, int64_t P, const float* queries, int nq,
                  float max_sq_dist, int32_t* nn) {
  kdtree t;
  memset(&t, 0, sizeof(t));
  t.pts = proj;
  t.cap = (int)(2 * P / 4 + 64);
  t.idx = (int32_t*)xcalloc(P, sizeof(int32_t));
  for (int64_t i = 0; i < P; ++i) t.idx[i] = (int32_t)i;
  t.node_lo = (int32_t*)xcalloc(t.cap, sizeof(int32_t));
  t.node_hi = (int32_t*)xcalloc(t.cap, sizeof(int32_t));
  t.node_left = (int32_t*)xcalloc(t.cap, sizeof(int32_t));
  t.node_right = (int32_t*)xcalloc(t.cap, sizeof(int32_t));
  t.node_dim = (int8_t*)xcalloc(t.cap, sizeof(int8_t));
  t.node_split = (float*)xcalloc(t.cap, sizeof(float));
  if (P > 0) kd_build(&t, 0, (int)P);
  for (int q = 0; q < nq; ++q) {
    float best = INFINITY;
    int32_t bi = -1;
    if (P > 0) kd_search(&t, 0, queries + 3 * q, &best, &bi);
    nn[q] = (bi >= 0 && best <= max_sq_dist) ? bi : -1;
  }
  free(t.idx);
  free(t.node_lo);
  free(t.node_hi);
  free(t.node_left);
  free(t.node_right);
  free(t.node_dim);
  free(t.node_split);
}

/* point-cloud.cc:66-72 */
float po_squared_voxel_size(int low_pass_lower, int low_pass_upper,
                            int n_voxels) {
  const float v =
      (float)(abs(low_pass_lower) + abs(low_pass_upper)) / (float)n_voxels;
  return v * v;
}

/* spherical-sampler.cc:20-27 + point-cloud.cc:127-157 +
 * function-value.cc:15-45: one neighbour, so the average is the value. */
void po_sample_cloud(int bw, const float* pts, int64_t P, float max_sq_dist,
                     int use_kdtree, double* grids) {
  po_sample_cloud_ex(bw, pts, NULL, P, max_sq_dist, use_kdtree, grids);
}

void po_sample_cloud_ex(int bw, const float* pts, const float* extra, int64_t P,
                        float max_sq_dist, int use_kdtree, double* grids) {
  const int n = 2 * bw, nq = n * n;
  float* q = (float*)xcalloc(3 * (size_t)nq, sizeof(float));
  float* proj = (float*)xcalloc(3 * (size_t)P, sizeof(float));
  double* range = (double*)xcalloc(P, sizeof(double));
  int32_t* nn = (int32_t*)xcalloc(nq, sizeof(int32_t));
  po_sampler_grid(bw, q);
  po_project(pts, P, proj, range);
  if (use_kdtree)
    po_nn_kdtree(proj, P, q, nq, max_sq_dist, nn);
  else
    po_nn_brute(proj, P, q, nq, max_sq_dist, nn);
  memset(grids, 0, sizeof(double) * 4 * (size_t)nq);
  for (int i = 0; i < nq; ++i) {
    if (nn[i] < 0) continue;
    grids[i] = range[nn[i]];
    grids[(size_t)nq + i] = (double)pts[4 * (size_t)nn[i] + 3];
    if (extra) { /* point-cloud.cc:152-155: the nearest point's own values */
      grids[2 * (size_t)nq + i] = (double)extra[2 * (size_t)nn[i]];
      grids[3 * (size_t)nq + i] = (double)extra[2 * (size_t)nn[i] + 1];
    }
  }
  free(q);
  free(proj);
  free(range);
  free(nn);
}

/* ======================================================================== */
/* S^2 transform                                                             */
/* ======================================================================== */
int po_seanindex(int m, int l, int bw) {
  const int bigL = bw - 1;
  if (m >= 0) return m * (bigL + 1) - ((m * (m - 1)) / 2) + (l - m);
  return ((bigL * (bigL + 3)) / 2) + 1 + ((bigL + m) * (bigL + m + 1) / 2) +
         (l - abs(m));
}

void po_makeweights(int bw, double* w) {
  const double fudge = M_PI / ((double)(4 * bw));
  for (int j = 0; j < 2 * bw; ++j) {
    double s = 0.0;
    for (int k = 0; k < bw; ++k)
      s += 1. / ((double)(2 * k + 1)) * sin((double)((2 * j + 1) * (2 * k + 1)) * fudge);
    s *= sin((double)(2 * j + 1) * fudge);
    s *= 2. / ((double)bw);
    w[j] = s;
    w[j + 2 * bw] = s * sin((double)(2 * j + 1) * fudge);
  }
}

static double l2_an(int m, int l) {
  return (sqrt((((double)(2 * l + 3)) / ((double)(2 * l + 1))) *
               (((double)(l - m + 1)) / ((double)(l + m + 1)))) *
          (((double)(2 * l + 1)) / ((double)(l - m + 1))));
}

static double l2_cn(int m, int l) {
  if (l == 0) return 0.0;
  return (-1.0 *
          sqrt((((double)(2 * l + 3)) / ((double)(2 * l - 1))) *
               (((double)(l - m + 1)) / ((double)(l + m + 1))) *
               (((double)(l - m)) / ((double)(l + m)))) *
          (((double)(l + m)) / ((double)(l - m + 1))));
}

/* P~_m^m, s2_primitive.c:185-208 (m=0 uses the literal sqrt(1/2) of
 * s2_cospmls.c:346-347). */
static double pmm_l2(int m, double theta) {
  if (m == 0) return 0.707106781186547;
  const double md = (double)m;
  double mcons = sqrt(md + 0.5), id = 0.0;
  for (int i = 0; i < m; ++i) {
    mcons *= sqrt((md - (id / 2.0)) / (md - id));
    id += 1.0;
  }
  mcons *= pow(2.0, -md / 2.0);
  if (m % 2) mcons *= -1.0;
  return mcons * pow(sin(theta), md);
}

void po_legendre(int bw, int m, double* out) {
  const int n = 2 * bw;
  for (int i = 0; i < n; ++i) {
    const double th = M_PI * (2.0 * i + 1.0) / (4.0 * bw);
    const double x = cos(th);
    double prevprev = 0.0, prev = pmm_l2(m, th);
    out[i] = prev;
    for (int k = 0; k < bw - m - 1; ++k) {
      const double t1 = l2_cn(m, m + k) * prevprev;
      const double t3 = l2_an(m, m + k) * (prev * x);
      const double cur = t3 + t1;
      out[(size_t)(k + 1) * n + i] = cur;
      prevprev = prev;
      prev = cur;
    }
  }
}

/* FST_semi_memo with dataformat=1 (real samples).  The reference reaches
 * sum_i w_i P~_l^m(cos th_i) F_m(th_i) through a DCT of the weighted data
 * and cosine-series tables (s2_legendreTransforms.c:437-522); the direct
 * quadrature below is the same sum. */
void po_fst(int bw, const double* data, double* re, double* im) {
  const int n = 2 * bw;
  double* row = (double*)xcalloc(2 * (size_t)n, sizeof(double));
  double* Fr = (double*)xcalloc((size_t)bw * n, sizeof(double));
  double* Fi = (double*)xcalloc((size_t)bw * n, sizeof(double));
  double* w = (double*)xcalloc(4 * (size_t)bw, sizeof(double));
  double* P = (double*)xcalloc((size_t)bw * n, sizeof(double));
  const double scale = (1. / ((double)n)) * sqrt(2. * M_PI);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      row[2 * j] = data[(size_t)i * n + j];
      row[2 * j + 1] = 0.0;
    }
    po_fft(row, n, 1, -1);
    for (int m = 0; m < bw; ++m) {
      Fr[(size_t)m * n + i] = row[2 * m] * scale;
      Fi[(size_t)m * n + i] = row[2 * m + 1] * scale;
    }
  }
  po_makeweights(bw, w);
  for (int m = 0; m < bw; ++m) {
    po_legendre(bw, m, P);
    for (int l = m; l < bw; ++l) {
      double sr = 0.0, si = 0.0;
      for (int i = 0; i < n; ++i) {
        const double wp = w[i] * P[(size_t)(l - m) * n + i];
        sr += wp * Fr[(size_t)m * n + i];
        si += wp * Fi[(size_t)m * n + i];
      }
      re[po_seanindex(m, l, bw)] = sr;
      im[po_seanindex(m, l, bw)] = si;
    }
  }
  /* s2_semi_memo.c:349-373 */
  double pow_one = 1.0;
  for (int i = 1; i < bw; ++i) {
    pow_one *= -1.0;
    for (int j = i; j < bw; ++j) {
      const int d0 = po_seanindex(i, j, bw), d1 = po_seanindex(-i, j, bw);
      re[d1] = pow_one * re[d0];
      im[d1] = -pow_one * im[d0];
    }
  }
  free(row);
  free(Fr);
  free(Fi);
  free(w);
  free(P);
}

/* ======================================================================== */
/* Laplace pyramid fusion                                                    */
/* ======================================================================== */
void po_pyramid_struct(uint32_t n_coeffs, int n_levels, float div,
                       uint32_t* coeffs, uint32_t* lower, uint32_t* upper) {
  for (int lvl = 0; lvl < n_levels; ++lvl) {
    const uint32_t n = lvl > 0 ? coeffs[lvl - 1] : n_coeffs;
    const float lb = (float)n / div;
    const uint32_t ub = (uint32_t)roundf((float)n - lb);
    lower[lvl] = (uint32_t)roundf(lb);
    upper[lvl] = ub;
    coeffs[lvl] = ub - lower[lvl];
  }
}

static double energy(const double* c) {
  const double e = sqrt(c[0] * c[0] + c[1] * c[1]);
  return e * e;
}

void po_fuse_channels(const double* channels, int C, uint32_t n_coeffs,
                      int n_levels, float div, double* out) {
  uint32_t* cf = (uint32_t*)xcalloc(n_levels, sizeof(uint32_t));
  uint32_t* lo = (uint32_t*)xcalloc(n_levels, sizeof(uint32_t));
  uint32_t* hi = (uint32_t*)xcalloc(n_levels, sizeof(uint32_t));
  po_pyramid_struct(n_coeffs, n_levels, div, cf, lo, hi);
  /* coefficients[c] starts as the channel data, then becomes the low pass */
  double** coef = (double**)xcalloc(C, sizeof(double*));
  double** fused = (double**)xcalloc(n_levels, sizeof(double*));
  uint32_t* fused_n = (uint32_t*)xcalloc(n_levels, sizeof(uint32_t));
  for (int c = 0; c < C; ++c) {
    coef[c] = (double*)xcalloc(2 * (size_t)n_coeffs, sizeof(double));
    memcpy(coef[c], channels + 2 * (size_t)c * n_coeffs,
           sizeof(double) * 2 * n_coeffs);
  }
  double* lap = NULL;
  for (int lvl = 0; lvl < n_levels; ++lvl) {
    const uint32_t n_lap = lvl > 0 ? cf[lvl - 1] : n_coeffs;
    const uint32_t n_low = cf[lvl];
    /* reduce (laplace-pyramid.cc:13-39) for every channel */
    lap = (double*)xcalloc(2 * (size_t)C * n_lap, sizeof(double));
    double** low = (double**)xcalloc(C, sizeof(double*));
    for (int c = 0; c < C; ++c) {
      low[c] = (double*)xcalloc(2 * (size_t)n_low + 2, sizeof(double));
      double* L = lap + 2 * (size_t)c * n_lap;
      memcpy(L, coef[c], sizeof(double) * 2 * n_lap);
      for (uint32_t i = lo[lvl]; i < hi[lvl]; ++i) {
        const uint32_t k = i - lo[lvl];
        low[c][2 * k] = coef[c][2 * i];
        low[c][2 * k + 1] = coef[c][2 * i + 1];
        L[2 * i] = L[2 * i] - low[c][2 * k];
        L[2 * i + 1] = L[2 * i + 1] - low[c][2 * k + 1];
      }
    }
    /* fuseLevelByMaxCoeff (laplace-pyramid.cc:106-122) */
    fused[lvl] = (double*)xcalloc(2 * (size_t)n_lap, sizeof(double));
    fused_n[lvl] = n_lap;
    for (uint32_t i = 0; i < n_lap; ++i) {
      int best = 0;
      double be = energy(lap + 2 * (size_t)i);
      for (int c = 1; c < C; ++c) {
        const double e = energy(lap + 2 * ((size_t)c * n_lap + i));
        if (e > be) {
          be = e;
          best = c;
        }
      }
      fused[lvl][2 * i] = lap[2 * ((size_t)best * n_lap + i)];
      fused[lvl][2 * i + 1] = lap[2 * ((size_t)best * n_lap + i) + 1];
    }
    for (int c = 0; c < C; ++c) {
      free(coef[c]);
      coef[c] = low[c];
    }
    free(low);
    free(lap);
  }
  /* fuseLastLowPassLayer (laplace-pyramid.cc:124-136, :179-184) */
  const uint32_t n_last = cf[n_levels - 1];
  double* recon = (double*)xcalloc(2 * (size_t)n_last + 2, sizeof(double));
  for (uint32_t i = 0; i < n_last; ++i) {
    double ar = 0, ai = 0;
    for (int c = 0; c < C; ++c) {
      ar += coef[c][2 * i];
      ai += coef[c][2 * i + 1];
    }
    recon[2 * i] = ar / (double)C;
    recon[2 * i + 1] = ai / (double)C;
  }
  uint32_t recon_n = n_last;
  /* expand (laplace-pyramid.cc:41-58) from the coarsest level up */
  for (int lvl = n_levels - 1; lvl >= 0; --lvl) {
    const uint32_t nc = fused_n[lvl];
    const uint32_t lb = (uint32_t)roundf((float)nc / div);
    const uint32_t ub = nc - lb;
    for (uint32_t i = lb; i < ub; ++i) {
      const uint32_t k = i - lb;
      if (k >= recon_n) break; /* the reference would read past the end */
      fused[lvl][2 * i] = recon[2 * k];
      fused[lvl][2 * i + 1] = recon[2 * k + 1];
    }
    free(recon);
    recon = fused[lvl];
    recon_n = nc;
  }
  memcpy(out, recon, sizeof(double) * 2 * n_coeffs);
  free(recon);
  for (int c = 0; c < C; ++c) free(coef[c]);
  free(coef);
  free(fused);
  free(fused_n);
  free(cf);
  free(lo);
  free(hi);
}

/* Two tag runs with channel values (1, 3) and (3, 1): an index fused by the
 * max-energy pick reads 3 in both runs, a low-pass mean reads 2, and an
 * index whose expand was dropped (the .5-rounding case) keeps 0. */
void po_fusion_modes(uint32_t n_coeffs, int n_levels, float div,
                     uint8_t* mode) {
  double* ch = (double*)xcalloc(4 * (size_t)n_coeffs, sizeof(double));
  double* o1 = (double*)xcalloc(2 * (size_t)n_coeffs, sizeof(double));
  double* o2 = (double*)xcalloc(2 * (size_t)n_coeffs, sizeof(double));
  for (uint32_t i = 0; i < n_coeffs; ++i) {
    ch[2 * i] = 1.0;
    ch[2 * ((size_t)n_coeffs + i)] = 3.0;
  }
  po_fuse_channels(ch, 2, n_coeffs, n_levels, div, o1);
  for (uint32_t i = 0; i < n_coeffs; ++i) {
    ch[2 * i] = 3.0;
    ch[2 * ((size_t)n_coeffs + i)] = 1.0;
  }
  po_fuse_channels(ch, 2, n_coeffs, n_levels, div, o2);
  for (uint32_t i = 0; i < n_coeffs; ++i) {
    /* pick -> 3 in both runs; mean -> 2 in both; zero -> 0 */
    const double a = o1[2 * i], b = o2[2 * i];
    if (a == 3.0 && b == 3.0)
      mode[i] = 0;
    else if (a == 2.0 && b == 2.0)
      mode[i] = 1;
    else
      mode[i] = 2;
  }
  free(ch);
  free(o1);
  free(o2);
}

static void rotate_left(void* base, size_t n, size_t k, size_t esz) {
  if (n == 0) return;
  k %= n;
  if (k == 0) return;
  unsigned char* tmp = (unsigned char*)xcalloc(n, esz);
  unsigned char* b = (unsigned char*)base;
  memcpy(tmp, b + k * esz, (n - k) * esz);
  memcpy(tmp + (n - k) * esz, b, k * esz);
  memcpy(b, tmp, n * esz);
  free(tmp);
}

void po_fftshift(double* c, size_t n) {
  rotate_left(c, n, (n % 2 == 0) ? (n >> 1) : (n >> 1) + 1, 2 * sizeof(double));
}
void po_ifftshift(double* c, size_t n) {
  rotate_left(c, n, n >> 1, 2 * sizeof(double));
}
void po_fftshift_real(double* v, size_t n) {
  rotate_left(v, n, (n % 2 == 0) ? (n >> 1) : (n >> 1) + 1, sizeof(double));
}
void po_ifftshift_real(double* v, size_t n) {
  rotate_left(v, n, n >> 1, sizeof(double));
}

/* ======================================================================== */
/* SO(3)                                                                     */
/* ======================================================================== */
/* makeWigner.c:180-210 */
static void l2_3term_so3(int j, int m1, int m2, double* an, double* bn,
                         double* cn) {
  const double dj = (double)j, dm1 = (double)m1, dm2 = (double)m2;
  const double t1 = sqrt((2 * dj + 3) / (2 * dj + 1));
  const double t3 = (dj + 1.) * (2 * dj + 1);
  double t5 = sqrt(((dj + 1) * (dj + 1) - dm1 * dm1) *
                   ((dj + 1) * (dj + 1) - dm2 * dm2));
  t5 = 1 / t5;
  if (j == 0) {
    *an = 0;
    *cn = 0;
  } else {
    const double t2 = sqrt((2 * dj + 3) / (2 * dj - 1)) * (dj + 1.) / dj;
    const double t4 = sqrt((dj * dj - dm1 * dm1) * (dj * dj - dm2 * dm2));
    *an = -t2 * t4 * t5;
    *cn = -dm1 * dm2 / (dj * (dj + 1));
  }
  *bn = t1 * t3 * t5;
}

/* makeWigner.c:324-387: d~ at degree max(|m1|,|m2|) from half-angle powers */
static double wig_spec(int m1, int m2, double s2, double c2) {
  const int a1 = abs(m1), a2 = abs(m2);
  const int l = a1 > a2 ? a1 : a2;
  const int delta = l - (a1 < a2 ? a1 : a2);
  const double dl = (double)l;
  double norm = 1.0, sign = 1.0;
  int cp, sp;
  for (int i = 0; i < delta; ++i)
    norm *= sqrt((2. * dl - ((double)i)) / (((double)i) + 1.));
  norm *= sqrt((2. * dl + 1.) / 2.);
  if (l == a1) {
    if (m1 >= 0) {
      cp = l + m2;
      sp = l - m2;
      if ((l - m2) % 2) sign = -1.;
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
    if ((l + m1) % 2) sign = -1.;
  }
  return norm * sign * pow(s2, (double)sp) * pow(c2, (double)cp);
}

void po_wigner_d(int bw, int m1, int m2, double* out) {
  const int n = 2 * bw;
  const int a1 = abs(m1), a2 = abs(m2);
  const int l0 = a1 > a2 ? a1 : a2;
  for (int k = 0; k < n; ++k) {
    const double twoN = (double)(2 * n);
    const double beta = ((2.0 * ((double)k) + 1.0) * M_PI) / twoN;
    const double cb = cos(beta);
    const double s2 = sin(0.5 * ((2.0 * ((double)k) + 1.0) * M_PI) / twoN);
    const double c2 = cos(0.5 * ((2.0 * ((double)k) + 1.0) * M_PI) / twoN);
    double prevprev = 0.0, prev = wig_spec(m1, m2, s2, c2);
    out[k] = prev;
    for (int i = 0; i < bw - l0 - 1; ++i) {
      double an, bn, cn;
      l2_3term_so3(l0 + i, m1, m2, &an, &bn, &cn);
      const double t0 = an * prevprev;
      const double t1 = bn * (cn + cb);
      const double cur = t0 + prev * t1;
      out[(size_t)(i + 1) * n + k] = cur;
      prevprev = prev;
      prev = cur;
    }
  }
}

/* Which (m1,m2) the reference synthesises with flag=1; the others are
 * conjugate copies of (-m1,-m2) (soft_fftw.c:1240-1780). */
static int so3_direct(int m1, int m2) {
  if (m1 > 0) return m2 != -m1;
  if (m1 == 0) return m2 >= 0;
  return m2 == -m1;
}

/* so3_correlate_fftw.c:102-161 for one (l, m1, m2) */
static void so3_combine(int bw, int l, int m1, int m2, const double* fr,
                        const double* fi, const double* hr, const double* hi,
                        double* re, double* im) {
  const double wigNorm = 2. * M_PI * sqrt(2. / (2. * l + 1.));
  const int ds = po_seanindex(-m1, l, bw), dp = po_seanindex(-m2, l, bw);
  const double fudge = ((m1 + m2) % 2) ? -1.0 : 1.0;
  const double sR = fr[ds], sI = fi[ds];
  const double pR = fudge * hr[dp], pI = fudge * (-hi[dp]);
  *re = wigNorm * (sR * pR - sI * pI);
  *im = wigNorm * (sR * pI + sI * pR);
}

void po_so3_synthesis(int bw, const double* f_re, const double* f_im,
                      const double* h_re, const double* h_im, double* S) {
  const int n = 2 * bw;
  memset(S, 0, sizeof(double) * 2 * (size_t)n * n * n);
  double* d = (double*)xcalloc((size_t)bw * n, sizeof(double));
  double* cr = (double*)xcalloc(bw, sizeof(double));
  double* ci = (double*)xcalloc(bw, sizeof(double));
  for (int m1 = -(bw - 1); m1 < bw; ++m1) {
    for (int m2 = -(bw - 1); m2 < bw; ++m2) {
      if (!so3_direct(m1, m2)) continue;
      const int l0 = abs(m1) > abs(m2) ? abs(m1) : abs(m2);
      po_wigner_d(bw, m1, m2, d);
      for (int l = l0; l < bw; ++l)
        so3_combine(bw, l, m1, m2, f_re, f_im, h_re, h_im, &cr[l - l0],
                    &ci[l - l0]);
      const int M1 = (m1 + n) % n, M2 = (m2 + n) % n;
      double* dst = S + 2 * (((size_t)M1 * n + M2) * n);
      for (int k = 0; k < n; ++k) {
        double a = 0.0, b = 0.0;
        for (int j = 0; j < bw - l0; ++j) {
          a += d[(size_t)j * n + k] * cr[j];
          b += d[(size_t)j * n + k] * ci[j];
        }
        dst[2 * k] = a;
        dst[2 * k + 1] = b;
      }
    }
  }
  for (int m1 = -(bw - 1); m1 < bw; ++m1) {
    for (int m2 = -(bw - 1); m2 < bw; ++m2) {
      if (so3_direct(m1, m2)) continue;
      const int M1 = (m1 + n) % n, M2 = (m2 + n) % n;
      const int S1 = (-m1 + n) % n, S2 = (-m2 + n) % n;
      double* dst = S + 2 * (((size_t)M1 * n + M2) * n);
      const double* src = S + 2 * (((size_t)S1 * n + S2) * n);
      for (int k = 0; k < n; ++k) {
        dst[2 * k] = src[2 * k];
        dst[2 * k + 1] = -src[2 * k + 1];
      }
    }
  }
  free(d);
  free(cr);
  free(ci);
}

void po_so3_correlate(int bw, const double* f_re, const double* f_im,
                      const double* h_re, const double* h_im,
                      double* out_cplx, double* mag) {
  const int n = 2 * bw;
  const size_t n3 = (size_t)n * n * n;
  double* S = (double*)xcalloc(2 * n3, sizeof(double));
  po_so3_synthesis(bw, f_re, f_im, h_re, h_im, S);
  /* transpose + FFT over m1 (alpha), transpose + FFT over m2 (gamma):
   * soft_fftw.c:1852-1901; output [k][alpha][gamma] */
  double* A = (double*)xcalloc(2 * (size_t)n * n, sizeof(double));
  const double dn = (1. / ((double)n)) * (((double)bw) / M_PI);
  for (int k = 0; k < n; ++k) {
    for (int M1 = 0; M1 < n; ++M1)
      for (int M2 = 0; M2 < n; ++M2) {
        const double* s = S + 2 * (((size_t)M1 * n + M2) * n + k);
        A[2 * ((size_t)M1 * n + M2)] = s[0];
        A[2 * ((size_t)M1 * n + M2) + 1] = s[1];
      }
    for (int M2 = 0; M2 < n; ++M2) po_fft(A + 2 * M2, n, n, -1);
    for (int a = 0; a < n; ++a) po_fft(A + 2 * (size_t)a * n, n, 1, -1);
    for (size_t e = 0; e < (size_t)n * n; ++e) {
      const double re = A[2 * e] * dn, im = A[2 * e + 1] * dn;
      const size_t o = (size_t)k * n * n + e;
      if (out_cplx) {
        out_cplx[2 * o] = re;
        out_cplx[2 * o + 1] = im;
      }
      if (mag) mag[o] = re * re + im * im;
    }
  }
  free(A);
  free(S);
}

void po_spherical_correlate_laplace(int bw, int C, const double* f,
                                    const double* h, double* mag,
                                    double* f_fused, double* h_fused) {
  po_spherical_correlate_laplace_pad(bw, bw, C, f, h, mag, f_fused, h_fused);
}

void po_spherical_correlate_laplace_pad(int bw, int bw_out, int C, const double* f,
                                        const double* h, double* mag,
                                        double* f_fused, double* h_fused) {
  const int n = 2 * bw;
  const size_t nc = (size_t)bw * bw;
  double* fch = (double*)xcalloc(2 * nc * C, sizeof(double));
  double* hch = (double*)xcalloc(2 * nc * C, sizeof(double));
  double* re = (double*)xcalloc(nc, sizeof(double));
  double* im = (double*)xcalloc(nc, sizeof(double));
  for (int c = 0; c < C; ++c) {
    for (int which = 0; which < 2; ++which) {
      const double* src = (which == 0 ? f : h) + (size_t)c * n * n;
      double* dst = (which == 0 ? fch : hch) + 2 * nc * c;
      po_fst(bw, src, re, im);
      po_fftshift_real(re, nc); /* spherical-correlation-laplace.cc:101-106 */
      po_fftshift_real(im, nc);
      for (size_t i = 0; i < nc; ++i) {
        dst[2 * i] = re[i];
        dst[2 * i + 1] = im[i];
      }
    }
  }
  double* F = (double*)xcalloc(2 * nc, sizeof(double));
  double* H = (double*)xcalloc(2 * nc, sizeof(double));
  po_fuse_channels(fch, C, (uint32_t)nc, 5, 4.0f, F);
  po_fuse_channels(hch, C, (uint32_t)nc, 5, 4.0f, H);
  double* fr = (double*)xcalloc(nc, sizeof(double));
  double* fi = (double*)xcalloc(nc, sizeof(double));
  double* hr = (double*)xcalloc(nc, sizeof(double));
  double* hi = (double*)xcalloc(nc, sizeof(double));
  for (size_t i = 0; i < nc; ++i) {
    fr[i] = F[2 * i];
    fi[i] = F[2 * i + 1];
    hr[i] = H[2 * i];
    hi[i] = H[2 * i + 1];
  }
  po_ifftshift_real(fr, nc);
  po_ifftshift_real(fi, nc);
  po_ifftshift_real(hr, nc);
  po_ifftshift_real(hi, nc);
  if (f_fused)
    for (size_t i = 0; i < nc; ++i) {
      f_fused[2 * i] = fr[i];
      f_fused[2 * i + 1] = fi[i];
    }
  if (h_fused)
    for (size_t i = 0; i < nc; ++i) {
      h_fused[2 * i] = hr[i];
      h_fused[2 * i + 1] = hi[i];
    }
  if (bw_out == bw) {
    po_so3_correlate(bw, fr, fi, hr, hi, NULL, mag);
  } else {
    /* so3CombineCoef_fftw(bw, bw_out, bw - 1, ...) then the inverse SO(3)
     * transform at bw_out (spherical-correlation.cc:248-278): the same as
     * the correlation at bw_out of coefficients that are zero for l >= bw */
    const size_t no = (size_t)bw_out * bw_out;
    double* p[4];
    const double* src[4] = {fr, fi, hr, hi};
    for (int k = 0; k < 4; ++k) {
      p[k] = (double*)xcalloc(no, sizeof(double));
      for (int l = 0; l < bw; ++l)
        for (int m = -l; m <= l; ++m)
          p[k][po_seanindex(m, l, bw_out)] = src[k][po_seanindex(m, l, bw)];
    }
    po_so3_correlate(bw_out, p[0], p[1], p[2], p[3], NULL, mag);
    for (int k = 0; k < 4; ++k) free(p[k]);
  }
  free(fch);
  free(hch);
  free(re);
  free(im);
  free(F);
  free(H);
  free(fr);
  free(fi);
  free(hr);
  free(hi);
}

void po_spherical_correlate_blend(int bw, const double* f_range,
                                  const double* f_int, const double* h_range,
                                  const double* h_int, double* mag) {
  const int n = 2 * bw;
  const size_t nn = (size_t)n * n, nc = (size_t)bw * bw;
  double* f = (double*)xcalloc(nn, sizeof(double));
  double* h = (double*)xcalloc(nn, sizeof(double));
  for (size_t i = 0; i < nn; ++i) {
    f[i] = 0.6 * f_int[i] + 0.4 * f_range[i];
    h[i] = 0.6 * h_int[i] + 0.4 * h_range[i];
  }
  double* fr = (double*)xcalloc(nc, sizeof(double));
  double* fi = (double*)xcalloc(nc, sizeof(double));
  double* hr = (double*)xcalloc(nc, sizeof(double));
  double* hi = (double*)xcalloc(nc, sizeof(double));
  po_fst(bw, f, fr, fi);
  po_fst(bw, h, hr, hi);
  po_so3_correlate(bw, fr, fi, hr, hi, NULL, mag);
  free(f);
  free(h);
  free(fr);
  free(fi);
  free(hr);
  free(hi);
}

/* ======================================================================== */
/* voxels and spatial correlation                                            */
/* ======================================================================== */
void po_linspaced(int N, float lo, float hi, float* E) {
  if (N == 1) {
    E[0] = hi;
    return;
  }
  const float step = (hi - lo) / (float)(N - 1);
  const int flip = fabsf(hi) < fabsf(lo);
  for (int i = 0; i < N; ++i) {
    if (flip)
      E[i] = (i == 0) ? lo : (hi - (float)(N - 1 - i) * step);
    else
      E[i] = (i == N - 1) ? hi : (lo + (float)i * step);
  }
}

int po_histc(double x, const float* E, int N) {
  if (x < (double)E[0] || x > (double)E[N - 1]) return -1;
  size_t l = 0, h = (size_t)N - 1, k = l;
  while ((h - l) > 1) {
    k = (h + l) / 2;
    if (x < (double)E[k])
      h = k;
    else
      l = k;
  }
  return (x == (double)E[h]) ? (int)h : (int)l;
}

void po_voxelize(const float* pts, int64_t P, int N, float lo, float hi,
                 double* ch, double* hist) {
  po_voxelize_ex(pts, NULL, P, N, lo, hi, ch, hist);
}

void po_voxelize_ex(const float* pts, const float* extra, int64_t P, int N, float lo,
                    float hi, double* ch, double* hist) {
  const uint32_t nf = (uint32_t)N * N * N;
  float* E = (float*)xcalloc(N, sizeof(float));
  double* hs = hist ? hist : (double*)xcalloc(nf, sizeof(double));
  po_linspaced(N, lo, hi, E);
  memset(ch, 0, sizeof(double) * 4 * (size_t)nf);
  memset(hs, 0, sizeof(double) * nf);
  for (int64_t i = 0; i < P; ++i) {
    const float x = pts[4 * i], y = pts[4 * i + 1], z = pts[4 * i + 2];
    const uint32_t bx = (uint32_t)(int64_t)po_histc((double)x, E, N);
    const uint32_t by = (uint32_t)(int64_t)po_histc((double)y, E, N);
    const uint32_t bz = (uint32_t)(int64_t)po_histc((double)z, E, N);
    /* SignalUtils::Sub2Ind in uint32 (signal-utils.cc:5-9) */
    const uint32_t lin = (bx * (uint32_t)N + by) + ((uint32_t)N * (uint32_t)N * bz);
    if (lin > nf) continue;  /* phase-aligner.cc:105-109 */
    if (lin == nf) continue; /* the reference would write one past the end */
    ch[lin] = ch[lin] + (double)pts[4 * i + 3];
    ch[(size_t)nf + lin] = ch[(size_t)nf + lin] + (double)sqrtf(x * x + y * y + z * z);
    if (extra) { /* phase-aligner.cc:113-118 */
      ch[2 * (size_t)nf + lin] = ch[2 * (size_t)nf + lin] + (double)extra[2 * i];
      ch[3 * (size_t)nf + lin] = ch[3 * (size_t)nf + lin] + (double)extra[2 * i + 1];
    }
    hs[lin] = hs[lin] + 1;
  }
  for (int c = 0; c < 4; ++c)
    for (uint32_t v = 0; v < nf; ++v) {
      const double r = ch[(size_t)c * nf + v] / hs[v];
      ch[(size_t)c * nf + v] = isfinite(r) ? r : 0.0;
    }
  free(E);
  if (!hist) free(hs);
}

void po_r2c3(int N, const double* in, double* out) {
  const size_t n3 = (size_t)N * N * N;
  const int h = N / 2 + 1;
  double* A = (double*)xcalloc(2 * n3, sizeof(double));
  for (size_t i = 0; i < n3; ++i) A[2 * i] = in[i];
  for (size_t r = 0; r < (size_t)N * N; ++r) po_fft(A + 2 * r * N, N, 1, -1);
  for (int d0 = 0; d0 < N; ++d0)
    for (int d2 = 0; d2 < N; ++d2)
      po_fft(A + 2 * ((size_t)d0 * N * N + d2), N, N, -1);
  for (int d1 = 0; d1 < N; ++d1)
    for (int d2 = 0; d2 < N; ++d2)
      po_fft(A + 2 * ((size_t)d1 * N + d2), N, (int64_t)N * N, -1);
  memset(out, 0, sizeof(double) * 2 * n3);
  for (int d0 = 0; d0 < N; ++d0)
    for (int d1 = 0; d1 < N; ++d1)
      for (int k = 0; k < h; ++k) {
        const size_t s = ((size_t)d0 * N + d1) * N + k;
        const size_t d = ((size_t)d0 * N + d1) * h + k;
        out[2 * d] = A[2 * s];
        out[2 * d + 1] = A[2 * s + 1];
      }
  free(A);
}

void po_c2r3(int N, const double* in, double* out) {
  const int h = N / 2 + 1;
  const size_t n3 = (size_t)N * N * N;
  double* Y = (double*)xcalloc(2 * (size_t)N * N * h, sizeof(double));
  memcpy(Y, in, sizeof(double) * 2 * (size_t)N * N * h);
  /* complex backward transforms over d0 and d1 for every k2 */
  for (int d1 = 0; d1 < N; ++d1)
    for (int k = 0; k < h; ++k)
      po_fft(Y + 2 * ((size_t)d1 * h + k), N, (int64_t)N * h, +1);
  for (int d0 = 0; d0 < N; ++d0)
    for (int k = 0; k < h; ++k)
      po_fft(Y + 2 * ((size_t)d0 * N * h + k), N, h, +1);
  /* 1-D c2r along d2: Hermitian completion, real parts of k=0 and N/2 */
  double* Z = (double*)xcalloc(2 * (size_t)N, sizeof(double));
  for (size_t r = 0; r < (size_t)N * N; ++r) {
    const double* y = Y + 2 * r * h;
    memset(Z, 0, sizeof(double) * 2 * N);
    Z[0] = y[0];
    for (int k = 1; k < h; ++k) {
      if (2 * k == N) {
        Z[2 * k] = y[2 * k];
        continue;
      }
      Z[2 * k] = y[2 * k];
      Z[2 * k + 1] = y[2 * k + 1];
      Z[2 * (N - k)] = y[2 * k];
      Z[2 * (N - k) + 1] = -y[2 * k + 1];
    }
    po_fft(Z, N, 1, +1);
    for (int j = 0; j < N; ++j) out[r * N + j] = Z[2 * j];
  }
  (void)n3;
  free(Y);
  free(Z);
}

void po_ind2sub(uint32_t idx, uint32_t rows, uint32_t cols, uint32_t* xyz) {
  xyz[1] = idx % cols;
  const int u = (int)(idx / cols);
  xyz[0] = (uint32_t)u % rows;
  xyz[2] = (uint32_t)u / rows;
}

uint32_t po_sub2ind(uint32_t i, uint32_t j, uint32_t k, uint32_t rows,
                    uint32_t cols) {
  return (i * cols + j) + (rows * cols * k);
}
#define sub2ind po_sub2ind

static int all_zero(const double* v, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (v[i] != 0.0) return 0;
  return 1;
}

/* spatial-correlation.cc:154-160 */
static void cmul_conj(const double* F, const double* G, double* C) {
  C[0] = F[0] * G[0] - F[1] * (-G[1]);
  C[1] = F[0] * (-G[1]) + F[1] * G[0];
}

/* SpatialCorrelation's padded correlation size, spatial-correlation.cc:
 * 17-24: N^3 times ((N + 2 pad) / N)^3 computed in double and truncated to
 * uint32, which is (N + 2 pad)^3 - 1 for many (N, pad) (e.g. N = 3, pad = 4:
 * 1330, not 1331).  The correlation vector has this many entries
 * (getCorrelationSize, :219-221), so the argmax never sees the last c2r
 * output then. */
uint32_t po_padded_size(int N, int pad) {
  const uint32_t T = (uint32_t)N * N * N;
  const uint32_t padding_per_dim = 2u * (uint32_t)pad;
  const double ratio = (N + padding_per_dim) / (double)N;
  const double factor = ratio * ratio * ratio;
  return (uint32_t)(T * factor);
}

/* The padded inverse of spatial-correlation.cc:142-161,221-230 and
 * spatial-correlation-laplace.cc:29-37: product entry i (i < N^3, in the
 * product's own order) goes to Sub2Ind(x + pad, y + pad, z + pad, Np, Np) of
 * a zeroed buffer of (N + 2 pad)^3 entries, with (x, y, z) = Ind2Sub(i, N, N)
 * (computeZeroPaddedIndex; the identity when pad = 0); `shifted` (the Laplace
 * class): IFFTShift of the buffer's first N^3 entries; then the c2r of size
 * Np = N + 2 pad over the buffer's first Np^2 (Np/2 + 1) entries.
 * c: Np^3. */
void po_spatial_pad_inverse(int N, int pad, int shifted, const double* prod, double* c) {
  const uint32_t T = (uint32_t)N * N * N;
  const int Np = N + 2 * pad;
  const size_t Tp = (size_t)Np * Np * Np;
  double* C = (double*)xcalloc(2 * Tp, sizeof(double));
  for (uint32_t idx = 0; idx < T; ++idx) {
    uint32_t p = idx;
    if (pad != 0) {
      uint32_t ijk[3];
      po_ind2sub(idx, N, N, ijk);
      p = sub2ind(ijk[0] + pad, ijk[1] + pad, ijk[2] + pad, Np, Np);
    }
    C[2 * (size_t)p] = prod[2 * (size_t)idx];
    C[2 * (size_t)p + 1] = prod[2 * (size_t)idx + 1];
  }
  if (shifted) po_ifftshift(C, T);
  po_c2r3(Np, C, c);
  free(C);
}

void po_spatial_correlate(int N, int pad, const double* f, const double* g,
                          double* c) {
  const uint32_t T = (uint32_t)N * N * N;
  double* F = (double*)xcalloc(2 * (size_t)T, sizeof(double));
  double* G = (double*)xcalloc(2 * (size_t)T, sizeof(double));
  double* P = (double*)xcalloc(2 * (size_t)T, sizeof(double));
  po_r2c3(N, f, F);
  po_r2c3(N, g, G);
  for (uint32_t idx = 0; idx < T; ++idx)
    cmul_conj(F + 2 * (size_t)idx, G + 2 * (size_t)idx, P + 2 * (size_t)idx);
  po_spatial_pad_inverse(N, pad, 0, P, c);
  free(F);
  free(G);
  free(P);
}

/* spatial-correlation-laplace.cc:17-89.  The reference reuses one r2c
 * output buffer for every channel and shifts it in place, so the tail
 * beyond N^2(N/2+1) carries an earlier channel's data; fusion is per index
 * and the c2r reads only the first N^2(N/2+1) unshifted entries, so the
 * tail never reaches the output and a zero tail is used here. */
void po_spatial_correlate_laplace(int N, int C, const double* f,
                                  const double* g, double* c) {
  po_spatial_correlate_laplace_pad(N, 0, C, f, g, c);
}

/* The same with phaser_core_spatial_zero_padding = pad: the fused product
 * (shifted order) through po_spatial_pad_inverse; c: (N + 2 pad)^3. */
void po_spatial_correlate_laplace_pad(int N, int pad, int C, const double* f,
                                      const double* g, double* c) {
  const uint32_t T = (uint32_t)N * N * N;
  double* Fc = (double*)xcalloc(2 * (size_t)T * C, sizeof(double));
  double* Gc = (double*)xcalloc(2 * (size_t)T * C, sizeof(double));
  for (int ch = 0; ch < C; ++ch) {
    /* an all-zero channel (reflectivity / ambient of a PointXYZI cloud)
     * transforms and shifts to exact zeros: its calloc'ed block stays */
    if (!all_zero(f + (size_t)ch * T, T)) {
      po_r2c3(N, f + (size_t)ch * T, Fc + 2 * (size_t)ch * T);
      po_fftshift(Fc + 2 * (size_t)ch * T, T);
    }
    if (!all_zero(g + (size_t)ch * T, T)) {
      po_r2c3(N, g + (size_t)ch * T, Gc + 2 * (size_t)ch * T);
      po_fftshift(Gc + 2 * (size_t)ch * T, T);
    }
  }
  double* F = (double*)xcalloc(2 * (size_t)T, sizeof(double));
  double* G = (double*)xcalloc(2 * (size_t)T, sizeof(double));
  po_fuse_channels(Fc, C, T, 5, 4.0f, F);
  free(Fc);
  po_fuse_channels(Gc, C, T, 5, 4.0f, G);
  free(Gc);
  double* Cc = (double*)xcalloc(2 * (size_t)T, sizeof(double));
  for (uint32_t i = 0; i < T; ++i)
    cmul_conj(F + 2 * (size_t)i, G + 2 * (size_t)i, Cc + 2 * (size_t)i);
  po_spatial_pad_inverse(N, pad, 1, Cc, c);
  free(F);
  free(G);
  free(Cc);
}

/* ======================================================================== */
/* evaluation and pose                                                       */
/* ======================================================================== */
uint32_t po_argmax(const double* v, size_t n) {
  size_t best = 0;
  for (size_t i = 1; i < n; ++i)
    if (v[i] > v[best]) best = i;
  return (uint32_t)best;
}

void po_zyz_from_index(uint32_t loc, int bw, double* zyz) {
  const int32_t ii = (int32_t)floor(loc / (4. * bw * bw));
  const int32_t tmp = (int32_t)(loc - (ii * 4. * bw * bw));
  const int32_t jj = (int32_t)floor(tmp / (2. * bw));
  const int32_t kk = (int32_t)(loc - (uint32_t)(ii * 4 * bw * bw) - (uint32_t)(jj * (2 * bw)));
  const double b = (double)bw;
  zyz[0] = M_PI * jj / b;
  zyz[1] = M_PI * (2 * ii + 1) / (4. * bw);
  zyz[2] = M_PI * kk / b;
}

static void quat_mul(const double* a, const double* b, double* r) {
  const double w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  const double x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
  const double y = a[0] * b[2] + a[2] * b[0] + a[3] * b[1] - a[1] * b[3];
  const double z = a[0] * b[3] + a[3] * b[0] + a[1] * b[2] - a[2] * b[1];
  r[0] = w;
  r[1] = x;
  r[2] = y;
  r[3] = z;
}

/* Eigen AngleAxisd(z0,Z) * AngleAxisd(z1,Y) * AngleAxisd(z2,Z) */
void po_quat_from_zyz(const double* zyz, double* q) {
  const double qa[4] = {cos(0.5 * zyz[0]), 0, 0, sin(0.5 * zyz[0])};
  const double qb[4] = {cos(0.5 * zyz[1]), 0, sin(0.5 * zyz[1]), 0};
  const double qc[4] = {cos(0.5 * zyz[2]), 0, 0, sin(0.5 * zyz[2])};
  double t[4];
  quat_mul(qa, qb, t);
  quat_mul(t, qc, q);
}

void po_xyz_from_quat(const double* q, double* xyz) {
  const double qw = q[0], qx = q[1], qy = q[2], qz = q[3];
  const double r11 = -2 * (qy * qz - qw * qx);
  const double r12 = qw * qw - qx * qx - qy * qy + qz * qz;
  const double r21 = 2 * (qx * qz + qw * qy);
  const double r31 = -2 * (qx * qy - qw * qz);
  const double r32 = qw * qw + qx * qx - qy * qy - qz * qz;
  xyz[2] = -atan2(r31, r32);
  xyz[1] = asin(r21);
  xyz[0] = atan2(r11, r12);
}

static void mat4_mul(const float* A, const float* B, float* R) {
  float t[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      float s = A[4 * i + 0] * B[0 * 4 + j];
      s = s + A[4 * i + 1] * B[1 * 4 + j];
      s = s + A[4 * i + 2] * B[2 * 4 + j];
      s = s + A[4 * i + 3] * B[3 * 4 + j];
      t[4 * i + j] = s;
    }
  memcpy(R, t, sizeof(t));
}

void po_transform_xyz(float a, float b, float c, float* T) {
  float X[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  float Y[16], Z[16];
  memcpy(Y, X, sizeof(X));
  memcpy(Z, X, sizeof(X));
  X[5] = cosf(a);
  X[6] = -sinf(a);
  X[9] = sinf(a);
  X[10] = cosf(a);
  Y[0] = cosf(b);
  Y[2] = sinf(b);
  Y[8] = -sinf(b);
  Y[10] = cosf(b);
  Z[0] = cosf(c);
  Z[1] = sinf(c);
  Z[4] = -sinf(c);
  Z[5] = cosf(c);
  float XY[16];
  mat4_mul(X, Y, XY);
  mat4_mul(XY, Z, T);
}

void po_apply_transform(const float* T, const float* pts, int64_t P,
                        float* out) {
  for (int64_t i = 0; i < P; ++i) {
    const float x = pts[4 * i], y = pts[4 * i + 1], z = pts[4 * i + 2];
    const float it = pts[4 * i + 3];
    out[4 * i + 0] = T[0] * x + T[1] * y + T[2] * z + T[3];
    out[4 * i + 1] = T[4] * x + T[5] * y + T[6] * z + T[7];
    out[4 * i + 2] = T[8] * x + T[9] * y + T[10] * z + T[11];
    out[4 * i + 3] = it;
  }
}

double po_translation_from_index(double index, uint32_t N, int lo, int hi) {
  if (index == 0) return 0.0;
  const double half = N / 2.0;
  const double width = abs(lo) + abs(hi) + 1;
  if (index <= half) return index * width / N;
  return (index - N) * width / N;
}

void po_register(const po_config* cfg, const float* target, int64_t Pt,
                 const float* source, int64_t Ps, po_result* res,
                 float* reg_out, double* rot_corr, double* trans_corr) {
  po_register_ex(cfg, target, NULL, Pt, source, NULL, Ps, res, reg_out, rot_corr,
                 trans_corr);
}

void po_register_ex(const po_config* cfg, const float* target, const float* target_extra,
                    int64_t Pt, const float* source, const float* source_extra, int64_t Ps,
                    po_result* res, float* reg_out, double* rot_corr, double* trans_corr) {
  const int bw = cfg->bw, n = 2 * bw, N = cfg->n_voxels;
  const size_t nn = (size_t)n * n;
  const size_t T3 = (size_t)N * N * N;
  memset(res, 0, sizeof(*res));
  const float max_sq =
      po_squared_voxel_size(cfg->lp_lower, cfg->lp_upper, cfg->n_voxels);
  double t0 = now_ms();
  double* fg = (double*)xcalloc(4 * nn, sizeof(double));
  double* hg = (double*)xcalloc(4 * nn, sizeof(double));
  po_sample_cloud_ex(bw, target, target_extra, Pt, max_sq, cfg->use_kdtree, fg);
  po_sample_cloud_ex(bw, source, source_extra, Ps, max_sq, cfg->use_kdtree, hg);
  double t1 = now_ms();
  /* the SO(3) grid at bw_out = bw + spherical zero padding */
  const int bw_out = bw + (cfg->method == 1 ? 0 : cfg->sph_pad);
  const size_t no3 = (size_t)8 * bw_out * bw_out * bw_out;
  double* mag = rot_corr ? rot_corr : (double*)xcalloc(no3, sizeof(double));
  if (cfg->method == 1)
    po_spherical_correlate_blend(bw, fg, fg + nn, hg, hg + nn, mag);
  else
    po_spherical_correlate_laplace_pad(bw, bw_out, 4, fg, hg, mag, NULL, NULL);
  res->rot_index = po_argmax(mag, no3);
  res->rot_value = mag[res->rot_index];
  po_zyz_from_index(res->rot_index, bw_out, res->zyz);
  po_quat_from_zyz(res->zyz, res->quat);
  po_xyz_from_quat(res->quat, res->euler_xyz);
  const int skip = cfg->method == 1 ? cfg->skip_rot_axes : 0;
  po_transform_xyz((skip & 1) ? 0.f : (float)res->euler_xyz[0],
                   (skip & 2) ? 0.f : (float)res->euler_xyz[1],
                   (skip & 4) ? 0.f : (float)res->euler_xyz[2], res->T_rot);
  float* reg = (float*)xcalloc(4 * (size_t)Ps, sizeof(float));
  po_apply_transform(res->T_rot, source, Ps, reg);
  double t2 = now_ms();
  double* fv = (double*)xcalloc(4 * T3, sizeof(double));
  double* gv = (double*)xcalloc(4 * T3, sizeof(double));
  /* the rotated source keeps each point's reflectivity and ambient */
  po_voxelize_ex(target, target_extra, Pt, N, (float)cfg->lower, (float)cfg->upper, fv, NULL);
  po_voxelize_ex(reg, source_extra, Ps, N, (float)cfg->lower, (float)cfg->upper, gv, NULL);
  /* spatial zero padding: the grid of Np = N + 2 pad per axis, of which the
   * first po_padded_size entries form the correlation vector (phase-aligner.
   * cc:66-68), indexed and converted to metres with Np voxels
   * (getNumberOfVoxels, :137-139; phase-correlation-eval.cc:22-27) */
  const int Np = N + 2 * cfg->spat_pad;
  const size_t Tg = (size_t)Np * Np * Np;
  const size_t Tc = cfg->spat_pad ? po_padded_size(N, cfg->spat_pad) : T3;
  double* c = trans_corr ? trans_corr : (double*)xcalloc(Tg, sizeof(double));
  po_spatial_correlate_laplace_pad(N, cfg->spat_pad, 4, fv, gv, c);
  res->trans_index = po_argmax(c, Tc);
  res->trans_value = c[res->trans_index];
  uint32_t xyz[3];
  po_ind2sub(res->trans_index, Np, Np, xyz);
  for (int d = 0; d < 3; ++d)
    res->translation[d] = po_translation_from_index((double)xyz[d], Np,
                                                    cfg->lower, cfg->upper);
  if (reg_out)
    for (int64_t i = 0; i < Ps; ++i) {
      reg_out[4 * i + 0] = reg[4 * i + 0] + (float)res->translation[0];
      reg_out[4 * i + 1] = reg[4 * i + 1] + (float)res->translation[1];
      reg_out[4 * i + 2] = reg[4 * i + 2] + (float)res->translation[2];
      reg_out[4 * i + 3] = reg[4 * i + 3];
    }
  double t3 = now_ms();
  res->t_sample_ms = t1 - t0;
  res->t_rot_ms = t2 - t1;
  res->t_trans_ms = t3 - t2;
  free(fg);
  free(hg);
  if (!rot_corr) free(mag);
  free(reg);
  free(fv);
  free(gv);
  if (!trans_corr) free(c);
}

/* ZScorePeakExtraction::extractPeaks + calculateSmoothedZScore
 * (z-score-peak-extraction.cc:42-112), the smoothed z-score detector over a
 * correlation vector.  lag = min(lag_max, (uint32)(n lag_percentile)) (:44-45);
 * nothing when n <= lag + 2 (:57-58); the running mean and standard
 * deviation (divisor lag - 1) of the window of the lag values before the
 * last processed index, starting with [0, lag) (:68-72); from i = lag + 1,
 * inputs <= filter_threshold are skipped without updating the window
 * (:77-79); an input farther than threshold deviations from the mean is a
 * signal when it is above the mean (:81-83) and is replaced by
 * influence x + (1 - influence) * (the previous, possibly replaced, input)
 * (:86); the window statistics then move to [i - lag, i) (:107-110).  The
 * signals come out in increasing order (a std::set); at most max_peaks are
 * stored, the return value counts all.  Window sums run in index order (the
 * reference's Eigen reductions may add in another order: last-bit
 * differences decide only exact threshold ties). */
size_t po_zscore_peaks(const double* corr, size_t n, double lag_percentile, uint32_t lag_max,
                       double threshold, double influence, double filter_threshold,
                       uint32_t* peaks, size_t max_peaks) {
  const uint32_t by_pct = (uint32_t)((double)n * lag_percentile);
  const double lag = (double)(lag_max < by_pct ? lag_max : by_pct);
  const uint32_t n_input = (uint32_t)n;
  if (n_input <= lag + 2) return 0;
  const uint32_t L = (uint32_t)lag;
  double* in = (double*)xcalloc(n, sizeof(double));
  memcpy(in, corr, n * sizeof(double));
  double sum = 0.0;
  for (uint32_t j = 0; j < L; ++j) sum += in[j];
  double avg = sum / (double)L;
  const double lagm1 = lag - 1;
  double sq = 0.0;
  for (uint32_t j = 0; j < L; ++j) sq += (in[j] - avg) * (in[j] - avg);
  double dev = sqrt(sq / lagm1);
  const double inv_influence = 1.0 - influence;
  size_t count = 0;
  for (uint32_t i = (uint32_t)(lag + 1u); i < n_input; ++i) {
    const double x = in[i];
    if (x <= filter_threshold) continue;
    if (fabs(x - avg) > threshold * dev) {
      if (x >= avg) {
        if (count < max_peaks) peaks[count] = i;
        ++count;
      }
      in[i] = influence * x + inv_influence * in[i - 1];
    }
    const uint32_t start = (uint32_t)(i - lag);
    sum = 0.0;
    for (uint32_t j = start; j < start + L; ++j) sum += in[j];
    avg = sum / (double)L;
    sq = 0.0;
    for (uint32_t j = start; j < start + L; ++j) sq += (in[j] - avg) * (in[j] - avg);
    dev = sqrt(sq / lagm1);
  }
  free(in);
  return count;
}
