/*
 * phaser_synth.h -- seeded synthetic scan pairs (host-side workload
 * generator shared by tests, the benchmark and the CPU baseline).
 * No counterpart in the reference, which ships no synthetic-pair generator
 * beyond its mock registrars (sph-registration-mock-*.cc).
 */
#ifndef PHASER_SYNTH_H_
#define PHASER_SYNTH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint64_t seed;
  int32_t rings;           /* elevation rings */
  int32_t azimuths;        /* columns per ring */
  float elev_min_deg;      /* lowest ring elevation */
  float elev_max_deg;      /* highest ring elevation */
  float min_range;         /* m */
  float max_range;         /* m, keeps every point inside the +-50 m box */
  float ground_depth;      /* ground plane at z = -ground_depth */
  int32_t n_boxes;
  int32_t n_cylinders;
  float scene_radius;      /* objects within this radius */
  float max_translation;   /* t uniform in [-max, max]^3 */
  float noise_sigma;       /* target Gaussian noise, m */
  float overlap;           /* fraction of the target azimuth ring kept */
  float outlier_frac;      /* uniform outliers, fraction of target points */
  int32_t random_rotation; /* 0: R = I */
  int32_t target_mode;     /* 0: target = R src + t (same points, rigidly
                              moved); 1: re-cast from the moved sensor */
} phz_synth_config;

void phz_synth_default_config(phz_synth_config* cfg);
/* upper bound on the points of either cloud */
int64_t phz_synth_capacity(const phz_synth_config* cfg);
/* src/tgt: x,y,z,intensity float quadruples.  R_out: row-major 3x3, t_out: 3.
 * The target satisfies x_t = R x_s + t (+ noise) where scans overlap. */
int phz_synth_pair(const phz_synth_config* cfg, float* src, int64_t cap_src,
                   int64_t* n_src, float* tgt, int64_t cap_tgt,
                   int64_t* n_tgt, double* R_out, double* t_out);

#ifdef __cplusplus
}
#endif
#endif /* PHASER_SYNTH_H_ */
