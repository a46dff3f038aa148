/* oracle_sanitize.c -- drives the oracle's registration (every option of
 * po_config) and the z-score detector on small synthetic clouds, for a build
 * of oracle/phaser_oracle.c with AddressSanitizer and UndefinedBehavior-
 * Sanitizer (tests/test_oracle_sanitize.py).  Prints "all ok". */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "phaser_oracle.h"

static uint64_t rng = 88172645463325252ull;
static double uni(void) {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return (double)(rng >> 11) * (1.0 / 9007199254740992.0);
}

/* a room-like cloud: points on the faces of a box, plus a few inside */
static void cloud(float* p, int64_t P, float shift) {
  for (int64_t i = 0; i < P; ++i) {
    float x = (float)(uni() * 30 - 15), y = (float)(uni() * 20 - 10), z = (float)(uni() * 6 - 3);
    const int face = (int)(uni() * 7);
    if (face == 0) x = -15; else if (face == 1) x = 15;
    else if (face == 2) y = -10; else if (face == 3) y = 10;
    else if (face == 4) z = -3; else if (face == 5) z = 3;
    p[4 * i + 0] = x + shift;
    p[4 * i + 1] = y;
    p[4 * i + 2] = z;
    p[4 * i + 3] = (float)(uni() * 100);
  }
}

int main(void) {
  const int64_t P = 3000;
  float* t = malloc(sizeof(float) * 4 * P);
  float* s = malloc(sizeof(float) * 4 * P);
  float* te = malloc(sizeof(float) * 2 * P);
  float* reg = malloc(sizeof(float) * 4 * P);
  cloud(t, P, 0.f);
  cloud(s, P, 1.5f);
  for (int64_t i = 0; i < 2 * P; ++i) te[i] = (float)(uni() * 50);
  struct { int bw, N, kd, sph_pad, method, skip, spat_pad, lp_lo, lp_hi; } cases[] = {
      {8, 16, 1, 0, 0, 0, 0, 0, 100000}, {8, 16, 0, 0, 0, 0, 0, 0, 100000},
      {8, 12, 1, 4, 0, 0, 0, 0, 100000}, {8, 16, 1, 0, 1, 2, 0, 0, 100000},
      {6, 10, 1, 0, 0, 0, 2, 0, 100000}, {8, 15, 1, 0, 0, 0, 1, 2, 9}};
  for (size_t c = 0; c < sizeof(cases) / sizeof(cases[0]); ++c) {
    po_config cfg = {0};
    cfg.bw = cases[c].bw;
    cfg.n_voxels = cases[c].N;
    cfg.lower = -50;
    cfg.upper = 50;
    cfg.lp_lower = cases[c].lp_lo;
    cfg.lp_upper = cases[c].lp_hi;
    cfg.use_kdtree = cases[c].kd;
    cfg.sph_pad = cases[c].sph_pad;
    cfg.method = cases[c].method;
    cfg.skip_rot_axes = cases[c].skip;
    cfg.spat_pad = cases[c].spat_pad;
    const int Bo = cfg.bw + cfg.sph_pad, Np = cfg.n_voxels + 2 * cfg.spat_pad;
    double* rc = malloc(sizeof(double) * 8 * (size_t)Bo * Bo * Bo);
    double* tc = malloc(sizeof(double) * (size_t)Np * Np * Np);
    po_result r;
    if (c == 1)
      po_register_ex(&cfg, t, te, P, s, NULL, P, &r, reg, rc, tc);
    else
      po_register(&cfg, t, P, s, P, &r, reg, rc, tc);
    printf("case %zu: rot %u trans %u\n", c, r.rot_index, r.trans_index);
    uint32_t sig[4096];
    const size_t n = po_zscore_peaks(rc, 8 * (size_t)Bo * Bo * Bo, 0.05, 20, 4.33, 0.05, 0.405,
                                     sig, 4096);
    printf("  z-score signals %zu\n", n);
    free(rc);
    free(tc);
  }
  free(t);
  free(s);
  free(te);
  free(reg);
  printf("all ok\n");
  return 0;
}
