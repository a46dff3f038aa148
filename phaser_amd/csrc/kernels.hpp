// kernels.hpp -- launchers for the phaser_amd HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "tables.hpp"

namespace phz {

// ------------------------------------------------------------- S^2 sampling
// Clouds are float4 (x,y,z,intensity).  `nclouds` clouds of up to `cap`
// points each, cloud c at pts + c*cap with npts[c] points.  Cells: n bands
// (theta) x n sectors (phi centred on the query columns).
struct CellBuffers {
  int nb, ns;        // colatitude bands x azimuth sectors of the search grid
  int* cell_count;   // nclouds * n*n, then the band colatitude ranges below
  // nclouds * nb * 2 past the counts (cleared with them): per band the
  // complement of the smallest and the largest colatitude bits of its points
  // (non-negative floats order as their bits; 0 = none yet)
  unsigned* band_th;
  // nclouds * nb past the band ranges: per band its longest cyclic run of
  // empty sectors, (length << 16) | first sector (rewritten by every call)
  unsigned* band_gap;
  int* cell_start;   // nclouds * (n*n + 1)
  int* band_start;   // nclouds * (n + 1)
  int4* next_right;  // nclouds * n*n: nearest non-empty sector to the right and
                     // its point range {sector, start, end, 0} (-1: empty band)
  int4* next_left;   // nclouds * n*n: the same to the left
  float4* sorted;    // nclouds * cap + kNnScanPad: projected x,y,z + original index bits
  float4* proj;      // nclouds * cap
  const float4* pts;  // the clouds the cells were built from (cloud c at pts + c cap): the
                      // search reads its samples' range and intensity there
  int* cell;         // nclouds * cap
  int* seed_pos;     // nclouds * kSeedCells: per coarse cell the sorted position
                     // of a point near its centre (-1: none)
  unsigned long long* band_mask;  // nclouds * nb * ceil(ns / 64): sector occupancy
};

// points of padding past the last cloud's sorted points: the search's
// branch-free batches read up to 4 x 16 points past a cell's range
constexpr int64_t kNnScanPad = 64;

// The S^2 search's coarse seed grid: F x F cells per coarse cell, F the
// smallest with ceil(n / F)^2 <= kSeedCells.
constexpr int kSeedCells = 4096;
inline int seed_factor(int n) {
  int F = 1;
  while (((n + F - 1) / F) * ((n + F - 1) / F) > kSeedCells) ++F;
  return F;
}

// How the exact nearest-neighbour search starts and prunes (phz_tuning.
// nn_seed): seeds = every query starts from its coarse cell's seed point (a
// real point: an upper bound on its nearest distance); mask = bands whose
// nearest occupied sectors are already out of reach are skipped on the
// band's occupancy bits in LDS, without a memory round trip.
// phz_tuning.nn_seed's default: neither (measured at 64 pairs per launch on
// gfx950, re-cast targets, with bands bounded by their points' colatitude
// range: the NN stage 0.454 ms plain, 0.540 ms with seeds, 0.474 ms with the
// mask; the search now also skips a band from its longest run of empty
// sectors, the mask's main case, without the mask's bit scans)
constexpr int kNnSeedDefault = 0;
struct NnSeeds {
  const int* pos = nullptr;
  const unsigned long long* mask = nullptr;
  int F = 0, cns = 0;
  int use_seeds = 0, use_mask = 0;
};

void launch_sample_s2(const float4* pts, const int* npts, int64_t cap, int nclouds, int n,
                      const float4* grid, const float* theta, float max_sq, CellBuffers cb,
                      float* grids /* nclouds x 2 x n*n : range, intensity */, hipStream_t s);
// the two halves of launch_sample_s2: projection + cell sort, then the NN search
void launch_s2_cells(const float4* pts, const int* npts, int64_t cap, int nclouds, int n,
                     CellBuffers cb, hipStream_t s);
// lanes per query: 1, 2, 4, 8 or 16 (others: 4); 0 = nn_auto_lanes(nclouds),
// 2 for launches of 64 clouds or more (many queries: the search's control
// flow is shared by more queries per wave) and 4 below (fewer queries: more
// lanes split each one's scans)
int nn_auto_lanes(int nclouds);
// extra (nullable): per-point (reflectivity, ambient) of each cloud slot
// (cloud * cap + original index); xgrids then receives nclouds x 2 x n*n
// grids of the nearest point's values (zero where none is accepted).
// xcd: order the workgroups so that each XCD works through its own clouds
// (needs a multiple of 8 clouds; otherwise the plain 2-D grid).
// blend: one grid per cloud, 0.6 intensity + 0.4 range (SphRegistration's
// channel, spherical-correlation.cc:103-112), at grids + cloud * n*n.
// seed_stride > 1 (with cb.seed_pos): two passes, the coarse queries at that
// stride first, then every query seeded by them (the same exact result)
void launch_s2_nn(int64_t cap, int nclouds, int n, const float4* grid, const float* theta,
                  float max_sq, CellBuffers cb, float* grids, int lanes, hipStream_t s,
                  const float2* extra = nullptr, float* xgrids = nullptr, bool xcd = false,
                  bool blend = false, int seed_mode = 0);

// ------------------------------------------------------------- S^2 transform
// sig: S real grids of n*n.  F: S x B x n complex scratch.  coef: S x B*B
// complex (seanindex order).
// precision 0: FP32 Legendre contraction; 1: split-FP16 MFMA (kernels_mfma.hip)
void launch_fst(const float* sig, int S, int bw, const float* legendre, const int* leg_off_dev,
                const float2* tw_n, float2* F, float2* coef, int precision, hipStream_t s);
void launch_legendre_mfma(const float2* F, int S, int bw, const float* legendre,
                          const int* leg_off, float2* coef, hipStream_t s);

// Laplace fusion of C channels (null = all-zero channel) of `count` complex
// coefficients, mode runs in unshifted order.  out: count complex.
struct ChannelSet {
  const float2* ch[8];
  int C;
};
// nsets channel sets (each channel pointer advanced by set_stride per set)
// are fused in one launch into out[set][count].
void launch_fuse(ChannelSet cs, uint32_t count, const FusionRuns& runs, float2* out,
                 hipStream_t s, int nsets = 1, int64_t set_stride = 0);
// keys[pair] = max over partials[pair*count .. +count)
void launch_reduce_keys(const unsigned long long* partials, int count, int npairs,
                        unsigned long long* keys, hipStream_t s);

// ------------------------------------------------------------- SO(3)
// S buffer: [m2 in 0..B-1][m1+B-1 in 0..2B-2][k in 0..n-1] complex per pair.
// precision 0: FP32 Wigner-d GEMV per pair; 1: split-FP16 MFMA over the batch
// (wig_hl, hl_off: the table split into FP16 halves by launch_wig_split,
// needed with precision 1)
void launch_so3_synth(int bw, const float2* f_coef, int64_t f_stride, const float2* h_coef,
                      int64_t h_stride, int npairs, const float* wigner,
                      const int64_t* wig_off_dev, const float* wig_norm, float2* S,
                      int precision, hipStream_t s, const uint16_t* wig_hl = nullptr,
                      const int64_t* hl_off = nullptr);
void launch_so3_synth_mfma(int bw, const float2* f, int64_t f_stride, const float2* h,
                           int64_t h_stride, int npairs, const uint16_t* hl, const int64_t* hl_off,
                           const float* wig_norm, float2* S, hipStream_t s);
// the Wigner table of bandwidth bw split once into FP16 hi / lo halves in the
// synthesis GEMM's layout: offsets per block (in halves) and the total
int64_t wig_split_offsets(int bw, std::vector<int64_t>* off);
void launch_wig_split(int bw, const float* wig, const int64_t* wig_off, const int64_t* hl_off,
                      uint16_t* hl, hipStream_t s);
// FFT over m1 -> alpha: U[k][m2][alpha] per pair
// reg: register four-step kernels (bw = 32, 64, 128; see so3_reg_ok)
bool so3_reg_ok(int bw);
void launch_so3_fft_alpha(int bw, int npairs, const float2* S, const FftPlan& plan,
                          const float2* tw, float2* U, hipStream_t s, bool reg = false);
// FFT over m2 -> gamma (Hermitian), |C|^2, optional grid write, argmax key
// via per-workgroup partials (so3_gamma_blocks(bw) per pair)
int so3_gamma_blocks(int bw);
void launch_so3_fft_gamma(int bw, int npairs, const float2* U, const FftPlan& plan,
                          const float2* tw, float* corr /*nullable*/, int64_t corr_stride,
                          unsigned long long* partials, unsigned long long* keys,
                          hipStream_t s, bool reg = false);

// One input cloud of a batched call on device pointers: its points (x, y, z,
// intensity), its (reflectivity, ambient) pairs or null, and its count.
struct GatherEntry {
  const float4* pts;
  const float2* extra;
  int64_t n;
};
// Copies cloud i < ncl of the device table tab to pts + i cap (and, when
// extra is given, its pairs, or zeros for a null entry, to extra + i cap)
// and its count to npts[i], all in one launch; max_n bounds every n.
void launch_gather_clouds(const GatherEntry* tab, int ncl, int64_t max_n, float4* pts,
                          float2* extra, int64_t cap, int* npts, hipStream_t s);

// CloudPreProcessor's filters (kernels_pre.hip).  launch_bbox: min / max of
// the finite points' x, y, z into d_out[6] (d_part: 6 x the returned block
// count floats, at most 256 blocks).
int launch_bbox(const float4* pts, int64_t P, float* d_part, float* d_out, hipStream_t s);
void launch_voxel_grid_keys(const float4* pts, int64_t P, float3 inv, int3 min_b, int3 mul,
                            uint32_t* keys, uint32_t* vals, hipStream_t s);
// stable sort of (key, point) pairs, runs of equal keys (unique, counts,
// nruns) and their offsets; temp == null: *temp_bytes = the size needed
hipError_t voxel_grid_sort_runs(void* temp, size_t* temp_bytes, const uint32_t* keys_in,
                                uint32_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                                int n, uint32_t* unique, uint32_t* counts, uint32_t* offsets,
                                int* nruns, hipStream_t s);
void launch_voxel_grid_centroids(const float4* pts, const uint32_t* order, const uint32_t* counts,
                                 const uint32_t* offsets, const int* nruns, int max_runs,
                                 float4* out, hipStream_t s);
// kept points in order into out, their count into nsel; temp == null: size query
hipError_t pass_through_select(void* temp, size_t* temp_bytes, const float4* pts, int64_t P,
                               int field, float lo, float hi, int negative, unsigned char* keep,
                               float4* out, int* nsel, hipStream_t s);

// ZScorePeakExtraction (kernels_peaks.hip): nvec vectors of n floats,
// stride apart; lag <= 64; signals of vector v in increasing order at
// peaks + v * max_peaks (at most max_peaks stored), their count in counts[v]
void launch_zscore_peaks(const float* corr, int64_t stride, int nvec, uint32_t n, int lag,
                         double threshold, double influence, double filter, uint32_t* peaks,
                         int64_t max_peaks, unsigned long long* counts, hipStream_t s);

// Device-side pose of each pair: the argmaxes, the Bingham and Gaussian
// fits to their peak windows and the float transform T (row-major 4x4)
// applied to the source.  status: kPose* bits of the checks the reference
// enforces with CHECK (a failed one aborts the reference's process).
struct DevPose {
  uint32_t rot_index, trans_index;
  float rot_value, trans_value;
  double zyz[3];
  double quat[4];
  double euler[3];
  double translation[3];
  double trans_cov[9];
  double rot_omega[4];
  double rot_axes[16];
  float T[16];
  int32_t status;
  int32_t pad;
};
constexpr int32_t kPoseRotWindow = 1;   // Bingham window end >= (2B)^3
constexpr int32_t kPoseRotWeights = 2;  // Bingham window weights sum <= 0
constexpr int32_t kPoseTransWeights = 4;  // Gaussian window weights sum <= 0
constexpr int32_t kPoseTransWindow = 8;   // GMM window end >= N^3

// One wavefront per pair.  corr: the pairs' (2B)^3 grids at corr_stride
// floats apart (read only when k > 0); k: bingham_peak_neighbors.
// nsets coefficient sets from the layout of stride in_stride into the
// layout of `count` entries: out[set][j] = map[j] >= 0 ? in[set][map[j]] : 0
// (spherical zero padding: degree-limited coefficients at a larger bandwidth)
void launch_pad_coef(const float2* in, int64_t in_stride, const int* map, int64_t count,
                     float2* out, int nsets, hipStream_t s);
// skip_axes: bits 0/1/2 leave the x/y/z Euler angle out of the transform;
// mixture: BmmPeakBasedEval's rules (weights normalised and checked also
// for a single sample)
void launch_rot_pose(int bw, int npairs, const unsigned long long* keys, const float* corr,
                     int64_t corr_stride, int k, DevPose* pose, hipStream_t s,
                     int skip_axes = 0, bool mixture = false);

// ------------------------------------------------------------- spatial
// voxel accumulators per cloud: sumI, sumR (float) and count (uint) of N^3
void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
void launch_copy_f4(float4* dst, const float4* src, size_t n, hipStream_t s);
// dense MFMA issue rate (dtype 0 f16, 1 bf16, 2 f32): returns the flops the
// launch performs; out: blocks * 256 floats
double launch_mfma_peak(int dtype, int iters, int blocks, float* out, hipStream_t s);
// occupancy of the voxel grids, kept beside the accumulators by
// launch_voxel_scatter (sets) and launch_voxel_clear (resets):
// row[cloud][z*N + x] and plane[cloud][z] are 1 where some voxel of that
// row / z plane holds a point, else 0 (null: not kept)
struct VoxelOcc {
  uint8_t* row = nullptr;
  uint8_t* plane = nullptr;
};
// Clouds that also carry (reflectivity, ambient) per point: extra[cloud *
// cap + i], summed into sumF / sumA (same layout as sumI); null: none.
struct VoxelExtra {
  const float2* extra = nullptr;
  float* sumF = nullptr;
  float* sumA = nullptr;
};
void launch_voxel_scatter(const float4* pts, const int* npts, int64_t cap, int nclouds,
                          const DevPose* pose /*nullable: no transform*/, int pose_stride,
                          int N, const float* edges, float* sumI, float* sumR, uint32_t* cnt,
                          VoxelOcc occ, hipStream_t s, VoxelExtra ex = {});
// forward 3-D r2c of (mean intensity, mean range) per cloud:
// columns per workgroup of the three column-FFT passes (0 = default);
// chosen by phz_autotune.  The tile width never changes the arithmetic of a
// line, so every choice gives bit-identical spectra.
struct SpatialTiles {
  int fwd_x = 0, fwd_z = 0, inv_x = 0;
  bool fused_reg = false;  // r2c, cross_z and c2r passes on the register engine
};
// tile value kRegisterEngine selects the register four-step kernel
// (fft_reg.hpp) for L = 64, 128, 256
constexpr int kRegisterEngine = 1;
int fft_cols_default_tile(int L);
bool fft_cols_tile_ok(int L, int it);
// one column pass over nbatch half spectra: which = 0 forward along x,
// 1 forward along z, 2 inverse along x, 3 inverse along z; it = columns per
// workgroup
void launch_spatial_fft_pass(int N, int nbatch, float2* buf, int which, int it,
                             const FftPlan& plan, const float2* tw, hipStream_t s);

// spec[cloud][ch in {I,R}][z][x][k2], k2 < N/2+1.  launch_spatial_forward =
// launch_spatial_r2c (voxel means -> r2c along y) + launch_spatial_fft_fwd
// (FFT along x, then z, of nspec = 2*nclouds half spectra).
void launch_spatial_r2c(int N, int nclouds, const float* sumI, const float* sumR,
                        const uint32_t* cnt, const FftPlan& plan, const float2* tw, float2* spec,
                        hipStream_t s);
void launch_spatial_fft_fwd(int N, int nspec, float2* spec, const FftPlan& plan,
                            const float2* tw, const SpatialTiles& tiles, hipStream_t s);
// register-engine forms of the r2c, cross_z and c2r passes (fft_reg.hpp),
// for N where spatial_reg_ok(N); same arguments as the LDS forms
bool spatial_reg_ok(int N);
// row_occ (nullable: every row read) skips the empty rows' loads; with
// sparse_out their (zero) spectrum rows are not written either, and only
// launch_spatial_fft_x_occ may read the spectra afterwards
void launch_spatial_r2c_reg(int N, int nclouds, const float* sumI, const float* sumR,
                            const uint32_t* cnt, const uint8_t* row_occ, bool sparse_out,
                            const float2* tw, float2* spec, hipStream_t s);
// the sparse r2c over listed occupied rows, one workgroup per z plane (the
// same output as launch_spatial_r2c_reg with sparse_out)
void launch_spatial_r2c_rows(int N, int nclouds, const float* sumI, const float* sumR,
                             const uint32_t* cnt, VoxelOcc occ, const float2* tw, float2* spec,
                             hipStream_t s);
// the forward x pass (launch_spatial_fft_pass which = 0) on the register
// engine over the spectra of nbatch / 2 clouds left by a sparse_out r2c:
// reads only the rows that hold points (the others are zero and were never
// written) and skips the lines of planes without points, which stay
// unwritten (read them as zero).  spatial_reg_ok(N) only.
void launch_spatial_fft_x_occ(int N, int nbatch, float2* spec, VoxelOcc occ, const float2* tw,
                              hipStream_t s);
void launch_spatial_cross_z_reg(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                                int64_t g_stride, const FusionRuns& runs, bool fuse,
                                const float2* tw, float2* work, hipStream_t s);
// the forward z pass fused into the cross spectrum: F and G hold the
// x-transformed spectra (launch_spatial_fft_x_occ, or launch_spatial_fft_pass
// which = 0), each pair transforms its channels
// along z in registers, fuses, multiplies and transforms back along z into
// work, as launch_spatial_cross_z_reg does after a separate z pass.  f_occ /
// g_occ (nullable: every plane read): the z-plane occupancy of pair 0's
// clouds, advanced per pair by the given strides.  Only for channel sets
// whose channels 0 and 1 are present and the rest absent
// (spatial_cross_zz_ok).
// The same for four live channels (F.ch[0..3], G.ch[0..3], C = 4).
bool spatial_cross_zz4_ok(int N, const ChannelSet& F, const ChannelSet& G);
void launch_spatial_cross_zz4_reg(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                                  int64_t g_stride, const uint8_t* f_occ, int64_t f_occ_stride,
                                  const uint8_t* g_occ, int64_t g_occ_stride,
                                  const FusionRuns& runs, const float2* tw, float2* work,
                                  hipStream_t s);
bool spatial_cross_zz_ok(int N, const ChannelSet& F, const ChannelSet& G);
void launch_spatial_cross_zz_reg(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                                 int64_t g_stride, const uint8_t* f_occ, int64_t f_occ_stride,
                                 const uint8_t* g_occ, int64_t g_occ_stride,
                                 const FusionRuns& runs, const float2* tw, float2* work,
                                 hipStream_t s);
int spatial_c2r_reg_blocks(int N);
// inverse FFT along x + c2r along y + argmax in one pass over whole z planes
// held in LDS (N = 64, 128): replaces launch_spatial_fft_inv followed by
// launch_spatial_c2r_reg; partials: npairs * N entries
// slab: work in the [pair][k2][z][x] layout of launch_spatial_cross_slab
bool spatial_inv_plane_ok(int N);
void launch_spatial_inv_plane(int N, int npairs, const float2* work, const float2* tw,
                              float* corr, unsigned long long* partials,
                              unsigned long long* keys, hipStream_t s, bool slab = false);
// the x pass, the z pass, fusion, cross spectrum and inverse z pass in one
// launch over (pair, k2) slabs with the (z, x) plane in LDS (N = 64, 128;
// two live channels as spatial_cross_zz_ok): F / G are the half spectra of
// launch_spatial_r2c_rows (only occupied rows written), f_rows / f_planes
// and g_rows / g_planes the occupancy bytes of the first target and source
// clouds (one cloud per pair; f_shared: target 0 for all).  Output:
// work[pair][k2][z][x], for launch_spatial_inv_plane(.., slab = true).
bool spatial_cross_slab_ok(int N, const ChannelSet& F, const ChannelSet& G);
void launch_spatial_cross_slab(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                               int64_t g_stride, const uint8_t* f_rows, const uint8_t* f_planes,
                               bool f_shared, const uint8_t* g_rows, const uint8_t* g_planes,
                               const FusionRuns& runs, const float2* tw, float2* work,
                               hipStream_t s);
void launch_spatial_c2r_reg(int N, int npairs, const float2* work, const float2* tw, float* corr,
                            unsigned long long* partials, unsigned long long* keys,
                            hipStream_t s);
// plane form of launch_spatial_r2c + the x pass (whole z planes in LDS), for
// N where spatial_plane_ok(N); follow with launch_spatial_fft_pass(which=1)
bool spatial_plane_ok(int N);
void launch_spatial_plane_fwd(int N, int nclouds, const float* sumI, const float* sumR,
                              const uint32_t* cnt, const FftPlan& plan, const float2* tw,
                              float2* spec, hipStream_t s);
void launch_spatial_forward(int N, int nclouds, const float* sumI, const float* sumR,
                            const uint32_t* cnt, const FftPlan& plan, const float2* tw,
                            float2* spec, hipStream_t s);
// forward from dense real channels (stage API): chans[c][N^3] in order,
// nch channels (<= 8) -> spec[c][z][x][k2]
void launch_spatial_forward_dense(int N, int nch, const float* chans, const FftPlan& plan,
                                  const float2* tw, float2* spec, const SpatialTiles& tiles,
                                  hipStream_t s);
// the same for two channels at separate addresses a and b (spec[0], spec[1]),
// with the default column tiles
void launch_spatial_forward_two(int N, const float* a, const float* b, const FftPlan& plan,
                                const float2* tw, float2* spec, hipStream_t s);
// phz_spatial_correlate_single_noalloc (noalloc.hip): the single-channel
// spatial correlation on a file-scope device array.  noalloc_setup runs once
// per device with that device current (phz_create): it takes the array's
// device address and writes the twiddle tables; spatial_single_noalloc then
// runs on that address only.
size_t noalloc_scratch_bytes();
int noalloc_max_n();
int noalloc_setup(int device, hipStream_t s);
int spatial_single_noalloc(int device, int N, const float* d_f, const float* d_g, float* d_corr,
                           uint32_t* argmax, hipStream_t s);
// fused cross spectrum + inverse 3-D c2r + argmax for npairs pairs.  F
// channels and G channels (null = zero channel) in the reference's spatial
// order; pair p reads its channels at ch[c] + p*stride (stride 0: shared).
// work: npairs * N^2(N/2+1); corr (nullable): npairs * N^3; partials:
// npairs * spatial_c2r_blocks(N) entries of scratch; keys: npairs.
int spatial_c2r_blocks(int N);
void launch_spatial_inverse(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                            int64_t g_stride, const FusionRuns& runs, bool fuse,
                            const FftPlan& plan, const float2* tw, float2* work,
                            float* corr /*nullable*/, unsigned long long* partials,
                            unsigned long long* keys, const SpatialTiles& tiles,
                            hipStream_t s);
// its three passes: cross spectrum + inverse FFT along z; inverse FFT along
// x; c2r along y + argmax
void launch_spatial_cross_z(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                            int64_t g_stride, const FusionRuns& runs, bool fuse,
                            const FftPlan& plan, const float2* tw, float2* work, hipStream_t s);
void launch_spatial_fft_inv(int N, int npairs, float2* work, const FftPlan& plan,
                            const float2* tw, const SpatialTiles& tiles, hipStream_t s);
// limit: only linear indices below it enter the argmax (the reference's
// padded correlation vector is shorter than the c2r output, po_padded_size)
void launch_spatial_c2r(int N, int npairs, const float2* work, const FftPlan& plan,
                        const float2* tw, float* corr, unsigned long long* partials,
                        unsigned long long* keys, hipStream_t s, uint32_t limit = 0xffffffffu);
// spatial zero padding (spatial-correlation.cc:142-161,221-230): the fused
// (or single-channel) cross spectrum over the H unshifted half-spectrum
// entries, without the inverse transform -> work[pair][H] ...
void launch_spatial_cross_product(int N, int npairs, ChannelSet F, int64_t f_stride,
                                  ChannelSet G, int64_t g_stride, const FusionRuns& runs,
                                  bool fuse, float2* work, hipStream_t s);
// ... then the c2r input of size Np = N + 2 pad: out[pair][Np^2 (Np/2+1)],
// with the Laplace class's IFFTShift when shifted; follow with
// launch_spatial_fft_pass which = 3 (inverse z) and 2 (inverse x) and
// launch_spatial_c2r at Np.  cube_lo <= cube_hi keeps only the products at
// shifted indices Sub2Ind(i, j, k) with i, j, k in [cube_lo, cube_hi]
// (SpatialCorrelationLowPass); cube_lo > cube_hi keeps all.
void launch_spatial_pad_gather(int N, int pad, bool shifted, int npairs, const float2* work,
                               float2* out, hipStream_t s, int cube_lo = 1, int cube_hi = 0);
// SphericalCorrelationLowPass (spherical-correlation-low-pass.cc:25-59):
// nsets coefficient sets of count = B^2 entries (stride apart), each zeroed
// outside [lower, upper) of its FFTShift order
void launch_coef_band(float2* coef, int64_t stride, int nsets, uint32_t count, uint32_t lower,
                      uint32_t upper, hipStream_t s);
// zero the accumulators of the voxels the points touched (same transform and
// binning as launch_voxel_scatter), so they need no full clear before reuse
void launch_voxel_clear(const float4* pts, const int* npts, int64_t cap, int nclouds,
                        const DevPose* pose, int pose_stride, int N, const float* edges,
                        float* sumI, float* sumR, uint32_t* cnt, VoxelOcc occ, hipStream_t s,
                        VoxelExtra ex = {});
// One wavefront per pair.  corr: the pairs' N^3 grids at corr_stride floats
// apart (read only when k > 0); k: gaussian_peak_neighbors.
// n_corr: entries of the correlation vector (0 = N^3; with spatial zero
// padding the reference's truncated padded size), the Gaussian window's end
// mixture: GmmPeakBasedEval's window rule (end = peak when peak + k > n_corr,
// an end of n_corr fails kPoseTransWindow)
void launch_trans_pose(int N, int lower, int upper, int npairs,
                       const unsigned long long* keys, const float* corr, int64_t corr_stride,
                       int k, DevPose* pose, hipStream_t s, uint32_t n_corr = 0,
                       bool mixture = false);
// out = T_rot * p + translation (PCL scalar order), intensity kept
void launch_apply_pose(const float4* pts, int64_t P, const DevPose* pose, float4* out,
                       hipStream_t s);

}  // namespace phz
