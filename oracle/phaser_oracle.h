/*
 * phaser_oracle.h -- FP64 CPU restatement of PHASER's correlation core.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP implementation in phaser_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product path never calls
 * it and must fail loudly if its own HIP library is missing.
 *
 * Every function restates the behaviour of one reference function (cited as
 * file:line, paths relative to the reference repository root).  Arithmetic is
 * FP64 except where the reference itself computes in float (projection, query
 * grid, voxel edges, point transforms); those places use float on purpose.
 *
 * Each function is pinned by tests/test_oracle.py, tests/test_golden.py,
 * tests/test_channels_oracle.py, tests/test_sph_padding_oracle.py and
 * tests/test_sph_method_oracle.py (po_spherical_correlate_blend and
 * po_config.method) (the reference tests' known answers, closed forms, brute
 * force, library routines).
 * Parity unpinned (no reference output fixes them): the last ulp of SOFT's
 * semi-naive DCT route and of FFTW (here direct quadrature and an own FFT, equal
 * in exact arithmetic); FLANN's choice among exactly equidistant neighbours
 * (here the lowest index); PCL's transform summation order (PCL 1.8 scalar
 * order assumed); Eigen LinSpaced (>= 3.3.4 semantics assumed).
 */
#ifndef PHASER_ORACLE_H_
#define PHASER_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FFT (stand-in for FFTW's arithmetic, own mixed-radix algorithm) ---- */
/* In-place complex DFT of n interleaved (re,im) doubles with element stride
 * `stride` (in complex elements).  sign=-1: forward e^{-2 pi i jk/n};
 * sign=+1: backward.  Unnormalised, like FFTW. */
void po_fft(double* data, int n, int64_t stride, int sign);

/* ---- S^2 sampling ------------------------------------------------------ */
/* a1: DH query grid, spherical-sampler.cc:29-64.  xyz: (2bw)^2 x 3 floats,
 * index i*2bw + j (theta-major). */
void po_sampler_grid(int bw, float* xyz);
/* a2: spherical-projection.cc:19-41.  pts: P x 4 floats (x,y,z,intensity).
 * proj: P x 3 floats (unit sphere), range: P doubles. */
void po_project(const float* pts, int64_t P, float* proj, double* range);
/* a3: exact 1-NN of each query among proj (float L2, FLANN L2_Simple order),
 * point-cloud.cc:93-157.  Ties go to the lowest point index.  nn[q] = -1 when
 * the squared distance exceeds max_sq_dist (point-cloud.cc:140-142). */
void po_nn_brute(const float* proj, int64_t P, const float* queries, int nq,
                 float max_sq_dist, int32_t* nn);
/* Same result through a kd-tree (used for the CPU baseline timing). */
void po_nn_kdtree(const float* proj, int64_t P, const float* queries, int nq,
                  float max_sq_dist, int32_t* nn);
/* point-cloud.cc:66-72 */
float po_squared_voxel_size(int low_pass_lower, int low_pass_upper,
                            int n_voxels);
/* a3+a4: sample a cloud on the DH grid.  grids: 4 x (2bw)^2 doubles in the
 * spherical channel order {range, intensity, reflectivity, ambient}
 * (spherical-combined-worker.cc:23-57).  use_kdtree selects po_nn_kdtree. */
void po_sample_cloud(int bw, const float* pts, int64_t P, float max_sq_dist,
                     int use_kdtree, double* grids);
/* The same with per-point (reflectivity, ambient) pairs (extra: 2 x P floats,
 * NULL = none, as for a PointXYZI cloud): grids 2 and 3 then hold the
 * nearest point's values (point-cloud.cc:152-155). */
void po_sample_cloud_ex(int bw, const float* pts, const float* extra, int64_t P,
                        float max_sq_dist, int use_kdtree, double* grids);

/* ---- S^2 transform ------------------------------------------------------ */
int po_seanindex(int m, int l, int bw); /* s2_primitive.c:300-313 */
void po_makeweights(int bw, double* w); /* makeweights.c:68-100, 4bw */
/* L2-normed associated Legendre P~_l^m(cos theta) for l = m..bw-1 at the 2bw
 * DH nodes.  out[(l-m)*2bw + i].  s2_primitive.c:48-75,185-208 */
void po_legendre(int bw, int m, double* out);
/* a7: FST_semi_memo(dataformat=1), s2_semi_memo.c:123-375.  data: (2bw)^2
 * real samples (theta-major).  re/im: bw^2 coefficients in seanindex order. */
void po_fst(int bw, const double* data, double* re, double* im);

/* ---- Laplace fusion (laplace-pyramid.cc, pyramid-struct.cc) -------------- */
/* pyramid-struct.cc:17-29.  Arrays of length n_levels. */
void po_pyramid_struct(uint32_t n_coeffs, int n_levels, float div,
                       uint32_t* coeffs, uint32_t* lower, uint32_t* upper);
/* laplace-pyramid.cc:60-104 restated literally.  channels: C arrays of
 * n_coeffs complex (interleaved), contiguous.  out: n_coeffs complex. */
void po_fuse_channels(const double* channels, int C, uint32_t n_coeffs,
                      int n_levels, float div, double* out);
/* Per-index fusion mode derived by running po_fuse_channels' structure on
 * tags: mode[i] = 0 pick max-energy channel, 1 channel mean, 2 forced zero. */
void po_fusion_modes(uint32_t n_coeffs, int n_levels, float div,
                     uint8_t* mode);
/* signal-utils.h:29-47, std::rotate semantics on complex arrays. */
void po_fftshift(double* cplx, size_t n);
void po_ifftshift(double* cplx, size_t n);
void po_fftshift_real(double* v, size_t n);
void po_ifftshift_real(double* v, size_t n);

/* ---- SO(3) ------------------------------------------------------------- */
/* L2-normed Wigner d~^l_{m1,m2}(beta_k), l = max(|m1|,|m2|)..bw-1, at the 2bw
 * points beta_k = pi(2k+1)/(4bw).  out[(l-l0)*2bw + k].
 * makeWigner.c:180-210,324-387,634-698. */
void po_wigner_d(int bw, int m1, int m2, double* out);
/* a9-a11: so3CombineCoef_fftw + Inverse_SO3_Naive_fftw(flag=1) + |.|^2.
 * f (signal) and h (pattern): bw^2 coefficients each (seanindex order).
 * out_cplx (optional, (2bw)^3 complex) and mag ((2bw)^3 doubles), layout
 * [beta][alpha][gamma].  so3_correlate_fftw.c:102-161,
 * soft_fftw.c:1036-1912, spherical-correlation.cc:273-278. */
void po_so3_correlate(int bw, const double* f_re, const double* f_im,
                      const double* h_re, const double* h_im,
                      double* out_cplx, double* mag);
/* Intermediate for GPU parity: S(m1,m2,beta_k) for every (m1,m2), as the
 * reference leaves it before the 2-D FFT (direct or conjugate-copied,
 * Nyquist rows zero).  S: [M1][M2][k] complex, M = m mod 2bw. */
void po_so3_synthesis(int bw, const double* f_re, const double* f_im,
                      const double* h_re, const double* h_im, double* S);
/* a7+a8+a9-a11: SphericalCorrelationLaplace::correlateSampledSignals,
 * spherical-correlation-laplace.cc:14-40.  f,h: C x (2bw)^2 samples.
 * Optional fused coefficient outputs (bw^2 complex each, unshifted). */
void po_spherical_correlate_laplace(int bw, int C, const double* f,
                                    const double* h, double* mag,
                                    double* f_fused, double* h_fused);
/* SphericalCorrelation::correlateSignals (single blended channel,
 * spherical-correlation.cc:54-80,103-112): signal = 0.6 I + 0.4 R. */
void po_spherical_correlate_blend(int bw, const double* f_range,
                                  const double* f_int, const double* h_range,
                                  const double* h_int, double* mag);

/* ---- voxels and spatial correlation ------------------------------------ */
/* Eigen::VectorXf::LinSpaced (Eigen >= 3.3.4 semantics), phase-aligner.cc:27-30 */
void po_linspaced(int N, float lo, float hi, float* E);
/* include/igl/histc.cpp:34-73 */
int po_histc(double x, const float* E, int N);
/* a15: PhaseAligner::discretizePointcloud, phase-aligner.cc:71-131.
 * ch: 4 x N^3 doubles in the spatial order {intensity, range, refl, ambient};
 * hist (optional): N^3 counts. */
void po_voxelize(const float* pts, int64_t P, int N, float lo, float hi,
                 double* ch, double* hist);
/* The same with per-point (reflectivity, ambient) pairs (NULL = none):
 * channels 2 and 3 then hold their voxel means (phase-aligner.cc:113-131). */
void po_voxelize_ex(const float* pts, const float* extra, int64_t P, int N, float lo,
                    float hi, double* ch, double* hist);
/* 3-D r2c into a flat N^3 complex buffer with a zero tail
 * (spatial-correlation.cc:30-57): out[d0][d1][k2], k2 <= N/2. */
void po_r2c3(int N, const double* in, double* out_cplx);
/* 3-D c2r of the first N^2(N/2+1) complex entries; imaginary parts of the
 * k2 = 0 and k2 = N/2 planes are ignored (FFTW hc2r behaviour). */
void po_c2r3(int N, const double* in_cplx, double* out);
/* SpatialCorrelation::correlateSignals, spatial-correlation.cc:163-206
 * (single channel, optional zero padding).  c: (N+2pad)^3. */
void po_spatial_correlate(int N, int pad, const double* f, const double* g,
                          double* c);
/* spatial-correlation.cc:17-24: entries of the padded correlation vector,
 * (uint32)(N^3 ((N + 2 pad) / N)^3) in double (may be (N + 2 pad)^3 - 1). */
uint32_t po_padded_size(int N, int pad);
/* spatial-correlation.cc:142-161,221-230 (+ spatial-correlation-laplace.cc:
 * 36-37 when shifted): N^3 complex products (interleaved) -> the padded
 * buffer, IFFTShift of its first N^3 entries if shifted, c2r of size
 * N + 2 pad.  c: (N + 2 pad)^3. */
void po_spatial_pad_inverse(int N, int pad, int shifted, const double* prod, double* c);
/* SpatialCorrelationLaplace with zero padding pad.  c: (N + 2 pad)^3. */
void po_spatial_correlate_laplace_pad(int N, int pad, int C, const double* f,
                                      const double* g, double* c);
/* a16-a19: SpatialCorrelationLaplace::correlateSignals,
 * spatial-correlation-laplace.cc:17-51.  f,g: C x N^3.  c: N^3. */
void po_spatial_correlate_laplace(int N, int C, const double* f,
                                  const double* g, double* c);

/* ---- evaluation and pose ---------------------------------------------- */
uint32_t po_argmax(const double* v, size_t n); /* first max, z-score-eval.cc:46 */
void po_zyz_from_index(uint32_t idx, int bw, double* zyz); /* rotation-utils.cc:190-207 */
void po_quat_from_zyz(const double* zyz, double* q);       /* :101-107, q=(w,x,y,z) */
void po_xyz_from_quat(const double* q, double* xyz);       /* :86-99,166-173 */
/* rotation-utils.cc:17-23,109-143: T = Tx(a) Ty(b) Tz(c), float, row-major */
void po_transform_xyz(float a, float b, float c, float* T);
/* PCL transformPointCloud (1.8 scalar order): out may alias pts. */
void po_apply_transform(const float* T, const float* pts, int64_t P,
                        float* out);
double po_translation_from_index(double index, uint32_t N, int lo, int hi);
void po_ind2sub(uint32_t idx, uint32_t rows, uint32_t cols, uint32_t* xyz);
/* signal-utils.cc:5-9 */
uint32_t po_sub2ind(uint32_t i, uint32_t j, uint32_t k, uint32_t rows,
                    uint32_t cols);

/* ZScorePeakExtraction (z-score-peak-extraction.cc:42-112): the signals of
 * a correlation vector in increasing order; stores at most max_peaks,
 * returns the count.  Flags z_score_lag_percentile (0.05), z_score_lag_max
 * (20), z_score_threshold (4.33), z_score_influence (0.05),
 * z_score_filter_threshold (0.405) (:12-25). */
size_t po_zscore_peaks(const double* corr, size_t n, double lag_percentile, uint32_t lag_max,
                       double threshold, double influence, double filter_threshold,
                       uint32_t* peaks, size_t max_peaks);

/* ---- a21: SphOptRegistration::registerPointCloud ------------------------ */
typedef struct {
  int bw;            /* phaser_core_spherical_bandwidth */
  int n_voxels;      /* phaser_core_spatial_n_voxels */
  int lower;         /* phaser_core_spatial_discretize_lower */
  int upper;         /* phaser_core_spatial_discretize_upper */
  int lp_lower;      /* phaser_core_spatial_low_pass_lower_bound (NN radius) */
  int lp_upper;      /* phaser_core_spatial_low_pass_upper_bound */
  int use_kdtree;    /* NN route (results identical) */
  int sph_pad;       /* phaser_core_spherical_zero_padding: SO(3) grid at bw + sph_pad */
  int method;        /* 0: SphOptRegistration (Laplace fusion of the four channels);
                        1: SphRegistration (sph-registration.cc:82-158): one sampled
                        channel 0.6 intensity + 0.4 range (spherical-correlation.cc:54-80,
                        103-112), transformed and correlated unshifted, no padding */
  int skip_rot_axes; /* method 1: bits 0/1/2 = refine_rot_x/y/z false, that Euler angle
                        zeroed before RotateAroundXYZ (sph-registration.cc:114-119) */
  int spat_pad;      /* phaser_core_spatial_zero_padding: translation grid of
                        (n_voxels + 2 spat_pad)^3, argmax over its first
                        po_padded_size entries */
} po_config;

typedef struct {
  uint32_t rot_index;    /* argmax of the SO(3) grid */
  uint32_t trans_index;  /* argmax of the spatial grid */
  double rot_value;
  double trans_value;
  double zyz[3];
  double quat[4];        /* w,x,y,z */
  double euler_xyz[3];
  float T_rot[16];       /* row-major 4x4 applied to the source */
  double translation[3];
  double t_sample_ms, t_rot_ms, t_trans_ms;
} po_result;

/* SphericalCorrelationLaplace with zero padding (spherical-correlation.cc:
 * 21-27, 248-278): fused coefficients at bw, correlation on the SO(3) grid
 * of bw_out >= bw (mag: (2 bw_out)^3). */
void po_spherical_correlate_laplace_pad(int bw, int bw_out, int C, const double* f,
                                        const double* h, double* mag,
                                        double* f_fused, double* h_fused);

/* target = cloud_prev (signal f), source = cloud_cur (pattern h).
 * reg_out (optional): registered source, P_s x 4.  rot_corr: (2 (bw +
 * sph_pad))^3, trans_corr: (N + 2 spat_pad)^3 (both optional). */
void po_register(const po_config* cfg, const float* target, int64_t Pt,
                 const float* source, int64_t Ps, po_result* res,
                 float* reg_out, double* rot_corr, double* trans_corr);
/* The same for clouds with (reflectivity, ambient) per point (2 x P floats
 * each, NULL = none): all four channels of the fusions are then live. */
void po_register_ex(const po_config* cfg, const float* target, const float* target_extra,
                    int64_t Pt, const float* source, const float* source_extra, int64_t Ps,
                    po_result* res, float* reg_out, double* rot_corr, double* trans_corr);

#ifdef __cplusplus
}
#endif
#endif /* PHASER_ORACLE_H_ */
