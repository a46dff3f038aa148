/*
 * phaser_amd.h -- C ABI of the MI355X correlation core of PHASER.
 *
 * Plain C types only.  Every call returns 0 on success, a negative status
 * on error (-1 invalid argument, -2 HIP runtime error, -3 out of capacity,
 * -4 a registered pair failed one of the reference's CHECKs on its peak
 * windows: every pair's result is filled and phz_result.status names it);
 * phz_last_error() returns the message of the calling thread's last error.
 * A context owns one HIP device, one stream, the resident tables for its
 * bandwidth and voxel count, and all device workspaces.  A context is not
 * thread-safe: use one per host thread (the reference's objects are not
 * thread-safe either, spherical-correlation.h:19-77).
 *
 * Point clouds are arrays of float quadruples (x, y, z, intensity), the
 * fields of the reference's PCL PointXYZI (point-types.h:10-14).
 * Arguments named d_* are device pointers; all others are host pointers.
 *
 * References below are paths in the PHASER repository.
 */
#ifndef PHASER_AMD_H_
#define PHASER_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phz_ctx phz_ctx;

/* Flags of core-gflags.cc:6-39 plus capacities. */
typedef struct {
  int32_t bandwidth;              /* phaser_core_spherical_bandwidth */
  int32_t spherical_zero_padding; /* phaser_core_spherical_zero_padding: the SO(3)
                                     correlation on the grid of Bo = bandwidth + padding
                                     (<= 160) from coefficients of degree < bandwidth
                                     (spherical-correlation.cc:21-27, 248-278); rotation
                                     grids and indices are then (2 Bo)^3 */
  int32_t n_voxels;               /* phaser_core_spatial_n_voxels */
  int32_t discretize_lower;       /* phaser_core_spatial_discretize_lower */
  int32_t discretize_upper;       /* phaser_core_spatial_discretize_upper */
  int32_t spatial_zero_padding;   /* phaser_core_spatial_zero_padding: the translation grid
                                     of Np = n_voxels + 2 padding per axis (Np <= 256;
                                     spatial-correlation.cc:11-61,221-230), trans_corr
                                     outputs then hold Np^3 floats, trans_index decodes
                                     with Np (phase-aligner.cc:137-139) and the argmax
                                     skips entries past the reference's truncated padded
                                     size (spatial-correlation.cc:17-24) */
  int32_t low_pass_lower;         /* phaser_core_spatial_low_pass_lower_bound */
  int32_t low_pass_upper;         /* phaser_core_spatial_low_pass_upper_bound */
  int64_t max_points;             /* capacity per cloud */
  int32_t max_batch;              /* pairs per batched call */
  int32_t device;                 /* HIP device ordinal */
  int32_t precision;              /* Legendre and Wigner-d contractions: 0 = FP32,
                                     1 = split FP16 (hi+lo) on MFMA, FP32 accumulation */
  int32_t bingham_peak_neighbors;  /* bingham_peak_neighbors (bingham-peak-based-eval.cc:
                                      8-10), 0 .. 2^20 */
  int32_t gaussian_peak_neighbors; /* gaussian_peak_neighbors (gaussian-peak-based-eval.cc:
                                      9-11), 0 .. 2^20 */
  int32_t channels;  /* per-point signal channels: 2 (range and intensity; reflectivity and
                        ambient are zero, as for a PointXYZI cloud; default) or 4 (clouds
                        may also carry reflectivity and ambient, phz_register_channels) */
  int32_t method;    /* PHZ_METHOD_SPH_OPT (0, SphOptRegistration: Laplace fusion of the
                        channels) or PHZ_METHOD_SPH (1, SphRegistration, sph-registration.cc:
                        82-158: one channel 0.6 intensity + 0.4 range, spherical-correlation.cc:
                        54-112; no spherical zero padding, which its SphericalCorrelation(bw)
                        does not take; reflectivity and ambient, with channels = 4, reach the
                        translation's fusion only).  Applies to the registration entry points;
                        the SO(3) stage entry points are SphericalCorrelationLaplace's */
  int32_t skip_rot_axes; /* bits 0/1/2: refine_rot_x/y/z false (sph-registration.cc:114-119),
                            that Euler angle left out of T_rot and the registered cloud;
                            PHZ_METHOD_SPH only */
  int32_t rot_eval;  /* rot_evaluation_algorithm (sph-registration.cc:21-23,58-65):
                        PHZ_EVAL_BINGHAM (0, BinghamPeakBasedEval, the window of
                        bingham_peak_neighbors) or PHZ_EVAL_MIXTURE (1, BmmPeakBasedEval:
                        bmm_peak_neighbors, weights normalised also for one sample, their
                        sum checked > 0; with the single argmax signal of ZScoreEval,
                        z-score-eval.cc:56-70, the mixture's one component is this
                        Bingham fit, whose mode is returned: the reference refits
                        Metropolis-Hastings samples of it drawn with an unseeded random
                        generator, bingham-mixture.cc:55-78, bingham.cc:126-190) */
  int32_t pos_eval;  /* pos_evaluation_algorithm: PHZ_EVAL_GAUSSIAN (0,
                        GaussianPeakBasedEval, gaussian_peak_neighbors) or PHZ_EVAL_MIXTURE
                        (1, GmmPeakBasedEval, gmm-peak-based-eval.cc:27-105: the window of
                        gmm_peak_neighbors ends at the peak when peak + k > N^3, and an end
                        of N^3 fails its CHECK_LT (status PHZ_PAIR_TRANS_WINDOW); one
                        signal's mixture is that Gaussian, gaussian-mixture.cc:57-78) */
  int32_t bmm_peak_neighbors;  /* bmm-peak-based-eval.cc:7-9 (default 2), 0 .. 2^20 */
  int32_t gmm_peak_neighbors;  /* gmm-peak-based-eval.cc:12-14 (default 2), 0 .. 2^20 */
} phz_config;

#define PHZ_EVAL_BINGHAM 0
#define PHZ_EVAL_GAUSSIAN 0
#define PHZ_EVAL_MIXTURE 1

#define PHZ_METHOD_SPH_OPT 0
#define PHZ_METHOD_SPH 1

/* The pose and uncertainty parts of model::RegistrationResult
 * (registration-result.cc:13-99, sph-opt-registration.cc:49-95): the
 * rotation argmax and the Bingham fit to its +-bingham_peak_neighbors
 * linear-index window (BinghamPeakBasedEval, bingham-peak-based-eval.cc:
 * 22-90; Bingham::fit, bingham.cc:239-296), the translation argmax and the
 * Gaussian fit to its +-gaussian_peak_neighbors window
 * (GaussianPeakBasedEval, gaussian-peak-based-eval.cc:26-102; gaussian.cc:
 * 14-56).  The window weights are the grid values (the z-score normalisation
 * is disabled in the reference, z-score-eval.cc:44-68).  With 0 neighbours
 * (the default) both fits have one sample: quat is the argmax's quaternion
 * and translation the argmax's translation.  A window the reference rejects
 * with a failed CHECK (a Bingham window reaching index (2B)^3, or weights that
 * do not sum to a positive value) makes the call return -4 (phz_result.status). */
typedef struct {
  uint32_t rot_index;     /* argmax of the (2B)^3 SO(3) grid [beta][alpha][gamma] */
  uint32_t trans_index;   /* argmax of the Np^3 grid (Np = N + 2 spatial padding),
                             x*Np + y + Np^2*z */
  float rot_value;        /* |C|^2 at rot_index */
  float trans_value;      /* correlation at trans_index */
  double zyz[3];          /* of the argmax, rotation-utils.cc:197-207 */
  double quat[4];         /* w,x,y,z: the Bingham mode (bingham.cc:110-113), signed
                             to agree with the argmax's quaternion
                             (rotation-utils.cc:101-107) */
  double euler_xyz[3];    /* of quat, rotation-utils.cc:86-99 */
  float T_rot[16];        /* row-major, applied to the source (RotateAroundXYZ) */
  double translation[3];  /* the Gaussian mean (gaussian.cc:48-50), of the argmax's
                             translation-utils.cc:40-52 translations */
  double trans_cov[9];    /* row-major Gaussian covariance (gaussian.cc:51-55); NaN
                             where the reference takes the root of a negative weight */
  double rot_omega[4];    /* eigenvalues of the Bingham moment matrix over their sum,
                             ascending (bingham.cc:241-246): the statistic the
                             concentrations are fitted to */
  double rot_axes[16];    /* row-major 4x4 whose columns are the matching unit
                             eigenvectors (the Bingham M, last column = quat) */
  int32_t status;         /* 0, or the failed checks (the call then returns -4):
                             PHZ_PAIR_ROT_WINDOW, PHZ_PAIR_ROT_WEIGHTS,
                             PHZ_PAIR_TRANS_WEIGHTS */
  int32_t reserved;
} phz_result;

#define PHZ_PAIR_ROT_WINDOW 1    /* Bingham window end >= (2Bo)^3: CHECK_LT, bingham-peak-based-eval.cc:69 */
#define PHZ_PAIR_ROT_WEIGHTS 2   /* Bingham window weights sum <= 0: CHECK_GT, :88 */
#define PHZ_PAIR_TRANS_WEIGHTS 4 /* Gaussian window weights sum <= 0: CHECK_GT, gaussian-peak-based-eval.cc:100 */
#define PHZ_PAIR_TRANS_WINDOW 8  /* GMM window end >= N^3: CHECK_LT, gmm-peak-based-eval.cc:81 */

/* Defaults of core-gflags.cc (bandwidth 75 there; 64 here so that the
 * default is the benchmarked configuration), max_points 262144. */
void phz_default_config(phz_config* cfg);
/* Replaces the constructors of SphOptRegistration (sph-opt-registration.cc:
 * 21-31), SphericalCorrelation (spherical-correlation.cc:21-28) and
 * PhaseAligner (phase-aligner.cc:19-45): builds every table once. */
int phz_create(const phz_config* cfg, phz_ctx** out);
void phz_destroy(phz_ctx* ctx);
const char* phz_last_error(void);
int phz_get_config(const phz_ctx* ctx, phz_config* out);

/* ---- end to end -------------------------------------------------------- */
/* SphOptRegistration::registerPointCloud(cloud_prev = target,
 * cloud_cur = source) (sph-opt-registration.cc:34-47).  rot_corr ((2Bo)^3
 * floats, |C|^2), trans_corr ((N + 2 spatial padding)^3 floats) and
 * registered (Ps x 4 floats,
 * the registered source) are optional outputs (NULL to skip). */
int phz_register(phz_ctx* ctx, const float* target, int64_t Pt, const float* source,
                 int64_t Ps, phz_result* result, float* rot_corr, float* trans_corr,
                 float* registered);
/* The same on device-resident clouds and outputs.  stream: a hipStream_t
 * (NULL = the context's stream).  Returns after the result is on the host. */
int phz_register_device(phz_ctx* ctx, const float* d_target, int64_t Pt,
                        const float* d_source, int64_t Ps, phz_result* result,
                        float* d_rot_corr, float* d_trans_corr, float* d_registered,
                        void* stream);
/* n independent pairs in one pass (n <= max_batch).  If shared_target is
 * nonzero only d_targets[0]/Pt[0] is read and every source registers against
 * it; the target's spectra are then computed once.  d_rot_corr/d_trans_corr
 * (optional) hold n consecutive grids.  Point buffers aligned to 16 bytes
 * (8 for the reflectivity / ambient pairs) are gathered into the context in
 * one launch; otherwise each cloud is copied on its own. */
int phz_register_pairs_device(phz_ctx* ctx, int n, int shared_target,
                              const float* const* d_targets, const int64_t* Pt,
                              const float* const* d_sources, const int64_t* Ps,
                              phz_result* results, float* d_rot_corr, float* d_trans_corr,
                              void* stream);

/* The same from host clouds (copied to the device on the context stream). */
int phz_register_pairs(phz_ctx* ctx, int n, int shared_target, const float* const* targets,
                       const int64_t* Pt, const float* const* sources, const int64_t* Ps,
                       phz_result* results);
/* phz_register for clouds that carry reflectivity and ambient as well
 * (model::PointCloud's reflectivities_ and ambient_points_, filled from a
 * PLY by point-cloud.cc:346-348): target_ra / source_ra hold (reflectivity,
 * ambient) per point (Pt x 2 / Ps x 2 floats; NULL: none, a zero channel).
 * They enter the S^2 samples (point-cloud.cc:152-155) and the voxel means
 * (phase-aligner.cc:113-118), so all four channels of both Laplace fusions
 * are live.  Needs a context created with channels = 4. */
int phz_register_channels(phz_ctx* ctx, const float* target, const float* target_ra, int64_t Pt,
                          const float* source, const float* source_ra, int64_t Ps,
                          phz_result* result, float* rot_corr, float* trans_corr,
                          float* registered);
/* phz_register_pairs_device for clouds with reflectivity and ambient:
 * d_targets_ra / d_sources_ra hold one device pointer per cloud to its
 * (reflectivity, ambient) pairs (either array, or any entry, may be NULL: a
 * zero channel).  Needs a context created with channels = 4. */
int phz_register_pairs_channels_device(phz_ctx* ctx, int n, int shared_target,
                                       const float* const* d_targets,
                                       const float* const* d_targets_ra, const int64_t* Pt,
                                       const float* const* d_sources,
                                       const float* const* d_sources_ra, const int64_t* Ps,
                                       phz_result* results, float* d_rot_corr,
                                       float* d_trans_corr, void* stream);
/* Device bytes owned by the context (tables and workspaces). */
int phz_memory_usage(const phz_ctx* ctx, size_t* bytes);
/* The device bytes phz_create would allocate for cfg (validates cfg; no
 * device call), for admission control by a memory budget. */
int phz_estimate_memory(const phz_config* cfg, size_t* bytes);

/* ---- stages (host buffers) ---------------------------------------------- */
/* SphericalSampler::sampleUniformly + the range and intensity channels of
 * SphericalCombinedWorker::run (spherical-sampler.cc:20-27,
 * spherical-combined-worker.cc:19-57).  grids: 2 x (2B)^2 floats
 * {range, intensity}, theta-major. */
int phz_sample_s2(phz_ctx* ctx, const float* pts, int64_t P, float* grids);
/* FST_semi_memo over nsig real grids (s2_semi_memo.c:123-375).
 * coef: nsig x B^2 complex (interleaved re,im) in seanindex order. */
int phz_s2_transform(phz_ctx* ctx, int nsig, const float* grids, float* coef);
/* SphericalCorrelation::correlate + inverseTransform on given coefficients
 * (spherical-correlation.cc:248-278): so3CombineCoef_fftw +
 * Inverse_SO3_Naive_fftw + |.|^2.  corr ((2B)^3) and synthesis (the
 * (m1,m2>=0,beta) half-plane, [m2][m1+B-1][k] complex) are optional. */
int phz_so3_correlate_coef(phz_ctx* ctx, const float* f_coef, const float* h_coef,
                           float* corr, float* synthesis, uint32_t* argmax);
/* SphericalCorrelationLaplace::correlateSampledSignals
 * (spherical-correlation-laplace.cc:14-40): C channel grids per signal
 * (2 <= C <= 8).  corr optional. */
int phz_spherical_correlate(phz_ctx* ctx, int C, const float* f_grids, const float* h_grids,
                            float* corr, uint32_t* argmax);
/* PhaseAligner::discretizePointcloud (phase-aligner.cc:71-131): means of
 * intensity and range per voxel, out: 2 x N^3 {intensity, range}; counts
 * (N^3 floats) optional. */
int phz_voxelize(phz_ctx* ctx, const float* pts, int64_t P, float* out, float* counts);
/* SpatialCorrelationLaplace::correlateSignals (spatial-correlation-laplace.
 * cc:17-51): C voxel channels per signal (2 <= C <= 8), corr optional:
 * (N + 2 spatial_zero_padding)^3 floats, the padded correlation of the
 * context's spatial zero padding (argmax over the reference's truncated
 * padded size). */
int phz_spatial_correlate(phz_ctx* ctx, int C, const float* f, const float* g, float* corr,
                          uint32_t* argmax);
/* SpatialCorrelation::correlateSignals (spatial-correlation.cc:163-206),
 * single channel, no padding.  n_voxels may differ from the context's
 * (any size <= 256). */
int phz_spatial_correlate_single(phz_ctx* ctx, int n_voxels, const float* f, const float* g,
                                 float* corr, uint32_t* argmax);
/* SpatialCorrelation(n_voxels, zero_padding)::correlateSignals
 * (spatial-correlation.cc:11-61,142-161,163-206,221-230), single channel:
 * each product of the two r2c spectra goes to computeZeroPaddedIndex in a
 * zeroed buffer of (N + 2 pad)^3 entries, whose c2r of size N + 2 pad is the
 * correlation.  corr: (N + 2 pad)^3 floats (optional); argmax over the
 * reference's correlation vector, whose length is the truncated double
 * product N^3 ((N + 2 pad) / N)^3 (getCorrelationSize, :17-24,219-221:
 * (N + 2 pad)^3 - 1 for some sizes, leaving the last entry out).
 * n_voxels + 2 zero_padding <= 256. */
int phz_spatial_correlate_single_padded(phz_ctx* ctx, int n_voxels, int zero_padding,
                                        const float* f, const float* g, float* corr,
                                        uint32_t* argmax);
/* SpatialCorrelationLowPass(n_voxels, zero_padding, lower, upper)::
 * correlateSignals (spatial-correlation-low-pass.cc:21-94): the r2c spectra
 * of one channel (the reference takes f[1], the range channel of
 * PhaseAligner's order), FFTShifted over N^3, multiplied only at the shifted
 * indices Sub2Ind(i, j, k, N, N), i, j, k in [lower, min(upper, N)]
 * (padded as phz_spatial_correlate_single_padded), IFFTShift, c2r.  The
 * reference leaves the rest of its product buffer and the spectra's tails as
 * the previous call left them; this is its first call on a fresh object
 * (zeros).  A cube reaching coordinate N, or an empty one, is rejected (-1)
 * where the reference's CHECKs abort.  corr: (N + 2 pad)^3 floats or NULL. */
int phz_spatial_correlate_low_pass(phz_ctx* ctx, int n_voxels, int zero_padding, int lower,
                                   int upper, const float* f, const float* g, float* corr,
                                   uint32_t* argmax);
/* SphericalCorrelationLowPass::correlateSampledSignals (spherical-
 * correlation-low-pass.cc:14-59): the S^2 transforms of one sampled grid
 * per signal (f1[0], f2[0]: (2B)^2 floats each), every coefficient zeroed
 * outside [min(B^2, lower), min(B^2, upper)) of the FFTShifted seanindex
 * order (phaser_core_spherical_low_pass_lower/upper_bound), then the SO(3)
 * correlation |C|^2 (corr: (2B)^3 floats or NULL).  A context without
 * spherical zero padding. */
int phz_spherical_correlate_low_pass(phz_ctx* ctx, const float* f_grid, const float* h_grid,
                                     int lower, int upper, float* corr, uint32_t* argmax);
/* No counterpart in the reference.  phz_spatial_correlate_single for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): it calls no hipMalloc,
 * hipFree or allocating routine and runs on a file-scope device array of
 * 25.6 MB per device (twiddle tables, spectra, workspace and argmax keys),
 * set up once per device by phz_create and held until the process ends.
 * n_voxels in [2, 128] (other sizes return -1).
 * d_f, d_g: N^3 device floats; d_corr: N^3 device floats or NULL; argmax:
 * host pointer or NULL (then the call does not wait for the stream).
 * stream: a hipStream_t of the context's device (NULL = the context's).
 * Every call on one device shares that device's array: calls that overlap
 * on one device race and must be serialised by the caller.  Results equal
 * phz_spatial_correlate_single's on the same inputs. */
int phz_spatial_correlate_single_noalloc(phz_ctx* ctx, int n_voxels, const float* d_f,
                                         const float* d_g, float* d_corr, uint32_t* argmax,
                                         void* stream);

/* ---- preprocessing (CloudPreProcessor, cloud-pre-processor.cc:30-44) --- */
/* VoxelGridCmd (voxel-grid-cmd.cc:11-24): PCL VoxelGrid (1.8) with leaf
 * size `leaf` on all three axes (phaser_voxel_grid_leaf_size, 0.25) and
 * every field averaged.  pts: P x 4 floats (x, y, z, intensity), replaced in
 * place by the voxel centroids ordered by voxel index; *P_out = their count.
 * Non-finite points are dropped; a grid of more than INT32_MAX voxels leaves
 * the cloud unchanged (as PCL).  Each voxel's points are summed in their
 * input order (PCL's unstable sort may add them in another order: last-bit
 * differences of voxels with three or more points). */
int phz_voxel_grid_filter(phz_ctx* ctx, float* pts, int64_t P, float leaf, int64_t* P_out);
/* PassThroughGndFilterCmd (pass-through-gnd-filter-cmd.cc:14-26): PCL
 * PassThrough on field 0 (x), 1 (y), 2 (z) or 3 (intensity): with negative = 0 the points
 * whose field lies in [lower, upper] are kept, with negative != 0 those
 * outside it (the command: z, -5, 3, negative, phaser_pass_through_gnd_
 * filter_limit_min/max); non-finite points are dropped.  pts is compacted
 * in place in input order; *P_out = the kept count. */
int phz_pass_through_filter(phz_ctx* ctx, float* pts, int64_t P, int field, float lower,
                            float upper, int negative, int64_t* P_out);

/* ZScorePeakExtraction (z-score-peak-extraction.cc:12-112): the smoothed
 * z-score detector over a correlation vector (the peak extraction that
 * ZScoreEval can run, z-score-eval.cc:70, before the Bingham / Gaussian /
 * mixture fits).  Defaults of the flags in phz_zscore_default_params. */
typedef struct {
  double lag_percentile;   /* z_score_lag_percentile (0.05), in (0, 1) */
  uint32_t lag_max;        /* z_score_lag_max (20); the window lag =
                              min(lag_max, (uint32)(n lag_percentile)) <= 64 */
  uint32_t reserved;
  double threshold;        /* z_score_threshold (4.33) */
  double influence;        /* z_score_influence (0.05), in [0, 1] */
  double filter_threshold; /* z_score_filter_threshold (0.405) */
} phz_zscore_params;
void phz_zscore_default_params(phz_zscore_params* p);
/* corr: n host floats; the signal indices in increasing order into peaks (at
 * most max_peaks stored), *count = all signals.  Runs on the device: one
 * wavefront, n / 64 speculative steps plus one per replaced input. */
int phz_zscore_peaks(phz_ctx* ctx, const float* corr, int64_t n, const phz_zscore_params* params,
                     uint32_t* peaks, int64_t max_peaks, int64_t* count);

/* ---- tuning -------------------------------------------------------------
 * No counterpart in the reference (FFTW plans are its only tuning,
 * spatial-correlation.cc:52-60).  Kernel shape of the three column-FFT
 * passes of the spatial correlation (forward along x, forward along z,
 * inverse along x): 8, 16, 32 or 64 columns per workgroup of the LDS
 * Stockham kernel (limited by LDS; the tile never changes a line's
 * arithmetic, so these give bit-identical results), or 1 for the register
 * four-step kernel (n = 64, 128, 256; same transform, other rounding). */
typedef struct {
  int32_t fwd_x, fwd_z, inv_x;   /* 0 = default */
  int32_t plane_fwd;   /* first forward pass on whole z planes in LDS (N <= 128):
                          1 = on, 0 or -1 = off (default) */
  int32_t nn_lanes;    /* lanes per query of the S^2 nearest-neighbour search:
                          1, 2, 4, 8 or 16, or 0 = automatic (2 for launches of 64
                          clouds or more, else 4); exact for every value */
  int32_t fused_reg;   /* the r2c, cross-spectrum and c2r passes on the register
                          four-step kernels (N = 64, 128, 256): 1 = on, 0/-1 = off */
  int32_t so3_reg;     /* the SO(3) alpha and gamma passes on the register kernels
                          (B = 32, 64, 128): 1 = on, 0/-1 = off */
  int32_t occ_fwd;     /* with fused_reg and the register engine for the forward
                          z pass: that pass runs inside the cross-spectrum
                          kernel (its output is never stored), and with the
                          register engine for the x pass too, the r2c, x and z
                          passes read and write only the voxel rows and z
                          planes that hold points (the rest of the transform
                          is zero): 0 or 1 = on (default), -1 = off */
  int32_t inv_plane;   /* with fused_reg and the register engine for the inverse x
                          pass (N = 64, 128): inverse x, c2r and argmax in one pass
                          over whole z planes in LDS: 0 or 1 = on (default), -1 = off */
  int32_t nn_xcd;      /* S^2 nearest-neighbour workgroups ordered so that each of the
                          8 XCDs works through its own clouds (launches of a multiple
                          of 8 clouds): 0 = default, 1 = on, -1 = off */
  int32_t nn_seed;     /* how the exact S^2 nearest-neighbour search starts and prunes
                          (the result is the same exact nearest point either way):
                          bit 0 = seeds (each query starts from a point near its
                          coarse cell, found by jump flooding over the cell grid: an
                          upper bound that prunes the walk from its first band),
                          bit 1 = bands skipped on their sector occupancy bits in LDS;
                          0 = default (neither: at 64 pairs per launch neither
                          shortened the search), -1 = neither, 1, 2 or 3 */
  int32_t xz_slab;     /* with occ_fwd and inv_plane on (two channels, N = 64, 128):
                          the x pass, the forward z pass, fusion, the cross
                          spectrum and the inverse z pass in one kernel per
                          (pair, y frequency) with the (z, x) plane in LDS, so
                          the x-transformed spectra are never stored (exact,
                          other rounding): 1 = on, 0 or -1 = off (default: at
                          N = 128 it took 1.08 ms per 64 pairs against 0.85 ms
                          for the separate x pass and cross-spectrum kernel) */
  float ms_fwd_x, ms_fwd_z, ms_inv_x;  /* per launch at max_batch (phz_autotune) */
  float ms_plane;      /* plane pass per launch, or the r2c + x passes it replaces */
  float ms_fused;      /* the three fused passes at max_batch, chosen form */
  float ms_so3;        /* SO(3) alpha + gamma at max_batch, chosen form */
} phz_tuning;
/* Times every tile of each pass, and the plane pass against the r2c + x
 * passes it replaces, on the context's buffers at max_batch (repeats
 * launches each, after one warm-up round); keeps the fastest. */
int phz_autotune(phz_ctx* ctx, int repeats, phz_tuning* out);
int phz_get_tuning(const phz_ctx* ctx, phz_tuning* out);
int phz_set_tuning(phz_ctx* ctx, const phz_tuning* tuning);
/* Nonzero: the register entry points replay the pipeline from captured HIP
 * graphs (one per batch size, shared-target flag and output grids; at most
 * 32 kept), which removes per-kernel launch overhead; inputs are still copied
 * and results fetched outside the graph.  Off while stage timing is on. */
int phz_set_graphs(phz_ctx* ctx, int enable);

/* ---- statistics --------------------------------------------------------- */
/* Per-stage device time measured with HIP events on the context's stream,
 * the counterpart of the reference's StatisticsManager timings
 * (sph-registration.cc:142-146,168-189; statistic-utils.h:10-18).
 * phz_set_profiling resets the accumulators. */
typedef struct {
  char name[32];
  double total_ms;
  int64_t calls;
} phz_stage_time;
int phz_set_profiling(phz_ctx* ctx, int enable);
int phz_get_stage_times(phz_ctx* ctx, phz_stage_time* out, int max, int* count);

/* ---- device memory -------------------------------------------------------
 * No counterpart in the reference (a CPU library).  For callers without a
 * HIP allocator of their own, so that one HIP runtime serves the process. */
int phz_device_malloc(phz_ctx* ctx, size_t bytes, void** ptr);
int phz_device_free(phz_ctx* ctx, void* ptr);
/* kind: 0 host->host, 1 host->device, 2 device->host, 3 device->device;
 * waits for the context's stream first. */
int phz_memcpy(phz_ctx* ctx, void* dst, const void* src, size_t bytes, int kind);
int phz_synchronize(phz_ctx* ctx);
/* A plain streaming copy kernel (bytes a multiple of 16) run `repeats`
 * times on the context's stream; *ms = mean duration per launch.  The HBM
 * reference against which kernel bandwidth and traffic counters are read. */
int phz_copy_kernel(phz_ctx* ctx, void* d_dst, const void* d_src, size_t bytes, int repeats,
                    float* ms);
/* The dense matrix-core rate measured on this device: 8192 waves each run
 * `iters` rounds of four independent accumulator chains of one MFMA
 * instruction on register data (dtype 0: v_mfma_f32_32x32x16_f16, 1: the
 * bf16 form, 2: v_mfma_f32_32x32x2_f32); *tflops = 2 M N K flops per
 * instruction over the launch's duration.  The MFMA reference beside
 * phz_copy_kernel's HBM one. */
int phz_mfma_peak(phz_ctx* ctx, int dtype, int iters, double* tflops);

#ifdef __cplusplus
}
#endif
#endif /* PHASER_AMD_H_ */
