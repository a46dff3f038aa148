/*
 * synth.c -- seeded synthetic lidar scan pairs for registration workloads.
 *
 * Scene (world frame): ground plane z = -ground_depth, n_boxes yawed boxes
 * with sides U[0.5,5] m and n_cylinders vertical cylinders, every object
 * inside scene_radius and at least 14 m from the origin, each surface with
 * its own intensity U[0,1].  A spinning lidar with `rings` elevation rings
 * (linearly spaced in [elev_min, elev_max]) x `azimuths` columns casts rays
 * and keeps the first hit closer than max_range.
 *
 * Source: the scan from the origin.  Target: the same scene in the frame
 * x_t = R X + t (R Haar-uniform, t in [-tmax,tmax]^3): by default the source
 * points moved rigidly (target_mode 0), or a new scan cast from the moved
 * sensor (target_mode 1).  Registering source onto target recovers (R, t).  The target gets
 * Gaussian noise, an azimuth crop that keeps `overlap` of the ring, and
 * `outlier_frac` uniform outliers inside a 45 m ball.  Every point stays
 * inside the +-50 m voxel box in its own frame and after any rotation.
 *
 * Deterministic across machines given the seed (xoshiro256** PRNG, double
 * arithmetic, libm).  Host only; not part of the GPU path.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/phaser_synth.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
  uint64_t s[4];
} rng_t;

static uint64_t splitmix(uint64_t* x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static void rng_seed(rng_t* r, uint64_t seed) {
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix(&seed);
}
static inline uint64_t rotl(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}
static uint64_t rng_next(rng_t* r) {
  uint64_t* s = r->s;
  const uint64_t res = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return res;
}
static double U01(rng_t* r) { return (rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static double U(rng_t* r, double a, double b) { return a + (b - a) * U01(r); }
static double gauss(rng_t* r) {
  double u1 = U01(r);
  if (u1 < 1e-300) u1 = 1e-300;
  const double u2 = U01(r);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

#define MAX_OBJ 256
typedef struct {
  int n_box, n_cyl;
  double bc[MAX_OBJ][3], bh[MAX_OBJ][3], byaw[MAX_OBJ], bI[MAX_OBJ][6];
  double cc[MAX_OBJ][2], cr[MAX_OBJ], cz0[MAX_OBJ], cz1[MAX_OBJ],
      cI[MAX_OBJ][2];
  double ground_z, ground_I;
} scene_t;

static void make_scene(const phz_synth_config* cfg, rng_t* r, scene_t* sc) {
  sc->n_box = cfg->n_boxes < MAX_OBJ ? cfg->n_boxes : MAX_OBJ;
  sc->n_cyl = cfg->n_cylinders < MAX_OBJ ? cfg->n_cylinders : MAX_OBJ;
  sc->ground_z = -cfg->ground_depth;
  sc->ground_I = U01(r);
  for (int i = 0; i < sc->n_box; ++i) {
    const double rad = U(r, 14.0, cfg->scene_radius - 4.0);
    const double ang = U(r, 0, 2 * M_PI);
    const double sx = U(r, 0.5, 5.0), sy = U(r, 0.5, 5.0), sz = U(r, 0.5, 5.0);
    sc->bc[i][0] = rad * cos(ang);
    sc->bc[i][1] = rad * sin(ang);
    sc->bc[i][2] = sc->ground_z + 0.5 * sz;
    sc->bh[i][0] = 0.5 * sx;
    sc->bh[i][1] = 0.5 * sy;
    sc->bh[i][2] = 0.5 * sz;
    sc->byaw[i] = U(r, 0, 2 * M_PI);
    for (int f = 0; f < 6; ++f) sc->bI[i][f] = U01(r);
  }
  for (int i = 0; i < sc->n_cyl; ++i) {
    const double rad = U(r, 14.0, cfg->scene_radius - 2.0);
    const double ang = U(r, 0, 2 * M_PI);
    sc->cc[i][0] = rad * cos(ang);
    sc->cc[i][1] = rad * sin(ang);
    sc->cr[i] = U(r, 0.2, 1.5);
    sc->cz0[i] = sc->ground_z;
    sc->cz1[i] = sc->ground_z + U(r, 2.0, 8.0);
    sc->cI[i][0] = U01(r);
    sc->cI[i][1] = U01(r);
  }
}

/* first hit along o + s d, s in (smin, smax); returns intensity or -1 */
static double cast(const scene_t* sc, const double* o, const double* d,
                   double smin, double smax, double* s_out) {
  double best = smax, I = -1.0;
  if (d[2] != 0.0) {
    const double s = (sc->ground_z - o[2]) / d[2];
    if (s > smin && s < best) {
      best = s;
      I = sc->ground_I;
    }
  }
  for (int i = 0; i < sc->n_box; ++i) {
    const double c = cos(sc->byaw[i]), sn = sin(sc->byaw[i]);
    const double px = o[0] - sc->bc[i][0], py = o[1] - sc->bc[i][1],
                 pz = o[2] - sc->bc[i][2];
    const double lo[3] = {c * px + sn * py, -sn * px + c * py, pz};
    const double ld[3] = {c * d[0] + sn * d[1], -sn * d[0] + c * d[1], d[2]};
    double t0 = -INFINITY, t1 = INFINITY;
    int face0 = -1;
    int ok = 1;
    for (int a = 0; a < 3 && ok; ++a) {
      if (fabs(ld[a]) < 1e-15) {
        if (fabs(lo[a]) > sc->bh[i][a]) ok = 0;
        continue;
      }
      double ta = (-sc->bh[i][a] - lo[a]) / ld[a];
      double tb = (sc->bh[i][a] - lo[a]) / ld[a];
      int fa = 2 * a, fb = 2 * a + 1;
      if (ta > tb) {
        const double tt = ta;
        ta = tb;
        tb = tt;
        const int ff = fa;
        fa = fb;
        fb = ff;
      }
      if (ta > t0) {
        t0 = ta;
        face0 = fa;
      }
      if (tb < t1) t1 = tb;
      if (t0 > t1) ok = 0;
    }
    if (ok && t0 > smin && t0 < best && face0 >= 0) {
      best = t0;
      I = sc->bI[i][face0];
    }
  }
  for (int i = 0; i < sc->n_cyl; ++i) {
    const double px = o[0] - sc->cc[i][0], py = o[1] - sc->cc[i][1];
    const double a = d[0] * d[0] + d[1] * d[1];
    const double b = 2 * (px * d[0] + py * d[1]);
    const double cq = px * px + py * py - sc->cr[i] * sc->cr[i];
    if (a > 1e-15) {
      const double disc = b * b - 4 * a * cq;
      if (disc >= 0) {
        const double s = (-b - sqrt(disc)) / (2 * a);
        const double z = o[2] + s * d[2];
        if (s > smin && s < best && z >= sc->cz0[i] && z <= sc->cz1[i]) {
          best = s;
          I = sc->cI[i][0];
        }
      }
    }
    if (d[2] != 0.0) {
      const double s = (sc->cz1[i] - o[2]) / d[2];
      const double x = px + s * d[0], y = py + s * d[1];
      if (s > smin && s < best && x * x + y * y <= sc->cr[i] * sc->cr[i]) {
        best = s;
        I = sc->cI[i][1];
      }
    }
  }
  *s_out = best;
  return I;
}

static void random_rotation(rng_t* r, double* R) {
  /* uniform unit quaternion (Shoemake) */
  const double u1 = U01(r), u2 = U01(r), u3 = U01(r);
  const double a = sqrt(1 - u1), b = sqrt(u1);
  const double w = a * sin(2 * M_PI * u2), x = a * cos(2 * M_PI * u2);
  const double y = b * sin(2 * M_PI * u3), z = b * cos(2 * M_PI * u3);
  R[0] = 1 - 2 * (y * y + z * z);
  R[1] = 2 * (x * y - z * w);
  R[2] = 2 * (x * z + y * w);
  R[3] = 2 * (x * y + z * w);
  R[4] = 1 - 2 * (x * x + z * z);
  R[5] = 2 * (y * z - x * w);
  R[6] = 2 * (x * z - y * w);
  R[7] = 2 * (y * z + x * w);
  R[8] = 1 - 2 * (x * x + y * y);
}

/* scan from world origin `o`, sensor-to-world rotation Rsw; points are
 * written in the sensor frame. */
static int64_t scan(const phz_synth_config* cfg, const scene_t* sc,
                    const double* o, const double* Rsw, float* out,
                    int64_t cap) {
  int64_t k = 0;
  for (int ri = 0; ri < cfg->rings; ++ri) {
    const double el =
        (cfg->rings == 1)
            ? 0.0
            : (cfg->elev_min_deg +
               (cfg->elev_max_deg - cfg->elev_min_deg) * ri / (cfg->rings - 1)) *
                  M_PI / 180.0;
    for (int ai = 0; ai < cfg->azimuths; ++ai) {
      const double az = 2 * M_PI * ai / cfg->azimuths;
      const double ds[3] = {cos(el) * cos(az), cos(el) * sin(az), sin(el)};
      const double dw[3] = {Rsw[0] * ds[0] + Rsw[1] * ds[1] + Rsw[2] * ds[2],
                            Rsw[3] * ds[0] + Rsw[4] * ds[1] + Rsw[5] * ds[2],
                            Rsw[6] * ds[0] + Rsw[7] * ds[1] + Rsw[8] * ds[2]};
      double s;
      const double I = cast(sc, o, dw, cfg->min_range, cfg->max_range, &s);
      if (I < 0) continue;
      if (k >= cap) return k;
      out[4 * k + 0] = (float)(s * ds[0]);
      out[4 * k + 1] = (float)(s * ds[1]);
      out[4 * k + 2] = (float)(s * ds[2]);
      out[4 * k + 3] = (float)I;
      ++k;
    }
  }
  return k;
}

void phz_synth_default_config(phz_synth_config* c) {
  memset(c, 0, sizeof(*c));
  c->seed = 0;
  c->rings = 64;
  c->azimuths = 1024;
  c->elev_min_deg = -25.0f;
  c->elev_max_deg = 15.0f;
  c->min_range = 0.5f;
  c->max_range = 40.0f;
  c->ground_depth = 2.0f;
  c->n_boxes = 20;
  c->n_cylinders = 5;
  c->scene_radius = 40.0f;
  c->max_translation = 5.0f;
  c->noise_sigma = 0.02f;
  c->overlap = 1.0f;
  c->outlier_frac = 0.0f;
  c->random_rotation = 1;
  c->target_mode = 0;
}

int64_t phz_synth_capacity(const phz_synth_config* c) {
  return (int64_t)c->rings * c->azimuths +
         (int64_t)ceil(c->outlier_frac * c->rings * c->azimuths) + 16;
}

int phz_synth_pair(const phz_synth_config* cfg, float* src, int64_t cap_src,
                   int64_t* n_src, float* tgt, int64_t cap_tgt,
                   int64_t* n_tgt, double* R_out, double* t_out) {
  if (!cfg || !src || !tgt || !n_src || !n_tgt) return -1;
  rng_t r;
  rng_seed(&r, cfg->seed * 0x100000001B3ull + 17);
  scene_t* sc = (scene_t*)calloc(1, sizeof(scene_t));
  if (!sc) return -2;
  make_scene(cfg, &r, sc);
  double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, t[3] = {0, 0, 0};
  if (cfg->random_rotation) random_rotation(&r, R);
  /* sensor origin o = -R^T t must stay above the ground */
  for (int it = 0; it < 1000; ++it) {
    for (int d = 0; d < 3; ++d)
      t[d] = U(&r, -cfg->max_translation, cfg->max_translation);
    const double oz = -(R[2] * t[0] + R[5] * t[1] + R[8] * t[2]);
    if (oz > -cfg->ground_depth + 0.5) break;
  }
  const double o0[3] = {0, 0, 0}, I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  *n_src = scan(cfg, sc, o0, I3, src, cap_src);
  const double o1[3] = {-(R[0] * t[0] + R[3] * t[1] + R[6] * t[2]),
                        -(R[1] * t[0] + R[4] * t[1] + R[7] * t[2]),
                        -(R[2] * t[0] + R[5] * t[1] + R[8] * t[2])};
  const double Rt[9] = {R[0], R[3], R[6], R[1], R[4], R[7], R[2], R[5], R[8]};
  int64_t nt;
  if (cfg->target_mode == 1) {
    nt = scan(cfg, sc, o1, Rt, tgt, cap_tgt);
  } else {
    nt = *n_src < cap_tgt ? *n_src : cap_tgt;
    for (int64_t i = 0; i < nt; ++i) {
      const double x = src[4 * i], y = src[4 * i + 1], z = src[4 * i + 2];
      for (int d = 0; d < 3; ++d)
        tgt[4 * i + d] = (float)(R[3 * d] * x + R[3 * d + 1] * y + R[3 * d + 2] * z + t[d]);
      tgt[4 * i + 3] = src[4 * i + 3];
    }
  }
  /* crop to `overlap` of the azimuth ring, add noise */
  int64_t k = 0;
  const double keep = cfg->overlap >= 1.0f ? 2 * M_PI + 1 : cfg->overlap * 2 * M_PI;
  for (int64_t i = 0; i < nt; ++i) {
    double az = atan2(tgt[4 * i + 1], tgt[4 * i]);
    if (az < 0) az += 2 * M_PI;
    if (az >= keep) continue;
    float p[4];
    for (int d = 0; d < 3; ++d)
      p[d] = (float)(tgt[4 * i + d] + cfg->noise_sigma * gauss(&r));
    p[3] = tgt[4 * i + 3];
    memcpy(tgt + 4 * k, p, sizeof(p));
    ++k;
  }
  const int64_t n_out = (int64_t)floor(cfg->outlier_frac * k);
  for (int64_t i = 0; i < n_out && k < cap_tgt; ++i) {
    double v[3], n2;
    do {
      for (int d = 0; d < 3; ++d) v[d] = U(&r, -1, 1);
      n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    } while (n2 > 1.0 || n2 < 1e-6);
    const double rad = 45.0 * cbrt(U01(&r));
    const double s = rad / sqrt(n2);
    tgt[4 * k + 0] = (float)(v[0] * s);
    tgt[4 * k + 1] = (float)(v[1] * s);
    tgt[4 * k + 2] = (float)(v[2] * s);
    tgt[4 * k + 3] = (float)U01(&r);
    ++k;
  }
  *n_tgt = k;
  if (R_out) memcpy(R_out, R, sizeof(R));
  if (t_out) memcpy(t_out, t, sizeof(t));
  free(sc);
  return 0;
}
