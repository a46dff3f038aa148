// capi.hip -- context, resident tables, workspaces and the C ABI entry
// points of phaser_amd (see include/phaser_amd.h for the reference
// interfaces each one replaces).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/phaser_amd.h"
#include "kernels.hpp"

namespace phz {

thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

FftPlan make_plan(int L) {
  FftPlan p;
  p.L = L;
  p.nst = 0;
  if (L < 1) {
    p.nst = -1;
    return p;
  }
  int r = L;
  while (r % 4 == 0 && p.nst < 12) {
    p.radix[p.nst++] = 4;
    r /= 4;
  }
  const int small[3] = {2, 3, 5};
  for (int f : small)
    while (r % f == 0 && p.nst < 12) {
      p.radix[p.nst++] = f;
      r /= f;
    }
  // remaining primes, smallest first, each a direct-DFT stage
  for (int f = 7; r > 1 && p.nst < 12; f += 2)
    while (r % f == 0 && p.nst < 12) {
      p.radix[p.nst++] = f;
      r /= f;
    }
  if (r != 1) p.nst = -1;
  return p;
}

// SpatialCorrelation's total_n_voxels_padded_ (spatial-correlation.cc:17-24):
// N^3 ((N + 2 pad) / N)^3 in double, truncated
uint32_t padded_size(int N, int pad) {
  const uint32_t T = (uint32_t)N * N * N;
  const uint32_t per_dim = 2u * (uint32_t)pad;
  const double ratio = (N + per_dim) / (double)N;
  const double factor = ratio * ratio * ratio;
  return (uint32_t)(T * factor);
}

bool smooth235(int L) {
  if (L < 1) return false;
  for (int f : {2, 3, 5})
    while (L % f == 0) L /= f;
  return L == 1;
}

}  // namespace phz

using namespace phz;

// largest bandwidth (and bandwidth + spherical zero padding): the S^2
// search grid of the nearest-neighbour kernels holds up to 2 * 160 bands
// and sectors (the reference's tuned driver run uses 150 padded to 160,
// run_phaser_core_driver:11-15)
constexpr int kMaxBandwidth = 160;

// XCD-ordered S^2 nearest-neighbour workgroups by default (measured on
// gfx950 at B = 64, 64 clouds per launch: k_nn 1.34 -> 1.18 ms)
constexpr bool kNnXcdDefault = true;

struct phz_ctx {
  phz_config cfg{};
  int B = 0, n = 0, N = 0, h = 0;
  // SO(3) correlation bandwidth and grid: Bo = B + spherical_zero_padding
  // (spherical-correlation.cc:21-27), no = 2 Bo
  int Bo = 0, no = 0;
  int64_t H = 0, T3 = 0, cap = 0;
  // spatial zero padding sp (spatial-correlation.cc:11-61): the translation
  // grid of Np = N + 2 sp per axis (T3p = Np^3 floats, Hp = Np^2 (Np/2 + 1)
  // half-spectrum entries), of which the first Tc entries are the
  // correlation vector (the reference's truncated padded size); without
  // padding Np = N, T3p = Tc = N^3
  int sp = 0, Np = 0;
  int64_t Hp = 0, T3p = 0;
  uint32_t Tc = 0;
  int max_clouds = 0;
  hipStream_t stream = nullptr;
  S2Tables s2;
  VoxelTables vox;
  FftPlan plan_n{}, plan_N{}, plan_no{}, plan_Np{};
  SpatialTiles tiles{};
  bool plane_fwd = false;  // whole-plane first forward pass
  bool so3_reg = false;    // SO(3) alpha/gamma passes on the register engine
  bool occ_fwd = true;     // forward passes skip rows / planes without points
  bool inv_plane = true;   // inverse x + c2r as one pass over z planes in LDS
  bool nn_xcd = kNnXcdDefault;  // XCD-ordered NN workgroups
  int nn_seed = kNnSeedDefault;  // NN seeds (bit 0) and LDS band skipping (bit 1)
  bool xz_slab = false;    // x and z passes + cross spectrum per (pair, k2) slab in LDS
  phz_tuning tuning{};
  float max_sq = 0.f;
  // resident tables
  float4* d_grid = nullptr;
  float* d_theta = nullptr;
  float* d_leg = nullptr;
  int* d_leg_off = nullptr;
  float* d_wig = nullptr;
  int64_t* d_wig_off = nullptr;
  float* d_wig_norm = nullptr;
  uint16_t* d_wig_hl = nullptr;      // precision 1: the table in FP16 halves
  int64_t* d_wig_hl_off = nullptr;
  float2* d_tw_n = nullptr;
  float2* d_tw_no = nullptr;     // exp(-2 pi i e / no) (d_tw_n when unpadded)
  int* d_pad_map = nullptr;      // Bo-layout coefficient -> B-layout index or -1
  float2* d_fused_o = nullptr;   // fused coefficients in the Bo layout (d_fused when unpadded)
  float* d_edges = nullptr;
  float2* d_tw_N = nullptr;
  float2* d_tw_Np = nullptr;     // exp(-2 pi i e / Np) (d_tw_N when unpadded)
  float2* d_work_pad = nullptr;  // padded c2r inputs, Hp per pair (spatial padding only)
  // workspaces
  float4* d_pts = nullptr;
  int* d_npts = nullptr;
  CellBuffers cells{};
  float* d_grids = nullptr;
  float2* d_F = nullptr;
  float2* d_coef = nullptr;
  float2* d_fused = nullptr;
  float2* d_S = nullptr;
  float2* d_U = nullptr;
  float* d_corr_rot = nullptr;
  unsigned long long* d_keys = nullptr;
  unsigned long long* d_partials = nullptr;  // per-workgroup argmax keys
  int64_t partials_per_pair = 0;
  DevPose* d_pose = nullptr;
  float* d_sumI = nullptr;
  float* d_sumR = nullptr;
  uint32_t* d_cnt = nullptr;
  uint8_t* d_row_occ = nullptr;    // [cloud][z][x], 1 where the row holds points
  uint8_t* d_plane_occ = nullptr;  // [cloud][z]
  float2* d_spec = nullptr;
  float2* d_work = nullptr;
  float* d_corr_trans = nullptr;
  // clouds with reflectivity and ambient (channels = 4)
  float2* d_extra = nullptr;   // [cloud slot][point] (reflectivity, ambient)
  float* d_sumF = nullptr;     // voxel sums of reflectivity / ambient, as d_sumR
  float* d_sumA = nullptr;
  float2* d_spec2 = nullptr;   // their spectra, laid out as d_spec
  float* d_peak_rot = nullptr;    // per-pair grids for the peak-window fits when the
  float* d_peak_trans = nullptr;  // caller asks for no grids (neighbours > 0 only)
  float4* d_reg = nullptr;
  int* h_npts = nullptr;  // pinned staging for per-cloud point counts
  // the input table of a batched call on device pointers: pinned staging,
  // its device copy, and an event after the upload (the staging is not
  // rewritten before the previous upload has run)
  GatherEntry* h_gather = nullptr;
  GatherEntry* d_gather = nullptr;
  hipEvent_t gather_done = nullptr;
  bool gather_pending = false;
  std::vector<void*> allocs;
  size_t alloc_bytes = 0;
  // per-stage HIP event timing (phz_set_profiling)
  struct Timing {
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    double total_ms = 0.0;
    int64_t calls = 0;
    bool pending = false;
  };
  bool profiling = false;
  std::vector<Timing> timings;
  // captured pipelines (phz_set_graphs), keyed by batch shape and outputs
  struct GraphEntry {
    int np;
    bool shared;
    float* rot;
    float* trans;
    hipGraphExec_t exec;
  };
  bool use_graphs = false;
  std::vector<GraphEntry> graphs;  // most recently built last
};

namespace {
// one stage per kernel where a kernel dominates it (spatial passes), so
// that stage time / launches is that kernel's mean launch duration
enum Stage {
  kS2Cells, kS2NN, kS2Transform, kFuse, kSo3Synth, kSo3Alpha, kSo3Gamma, kRotPose,
  kVoxelScatter, kSpatialR2c, kSpatialFftFwd, kVoxelClear, kSpatialCrossZ, kSpatialFftInv,
  kSpatialC2r, kTransPose, kNumStages
};
const char* kStageNames[kNumStages] = {
    "s2_cells", "s2_nn", "s2_transform", "fuse", "so3_synth", "so3_fft_alpha", "so3_fft_gamma",
    "rot_pose", "voxel_scatter", "spatial_r2c", "spatial_fft_fwd", "voxel_clear",
    "spatial_cross_z", "spatial_fft_inv", "spatial_c2r", "trans_pose"};

struct StageScope {
  phz_ctx* c;
  int st;
  hipStream_t s;
  StageScope(phz_ctx* c_, int st_, hipStream_t s_) : c(c_), st(st_), s(s_) {
    if (c->profiling) (void)hipEventRecord(c->timings[st].a, s);
  }
  ~StageScope() {
    if (c->profiling) {
      (void)hipEventRecord(c->timings[st].b, s);
      c->timings[st].pending = true;
    }
  }
};

void collect_timings(phz_ctx* c) {
  if (!c->profiling) return;
  for (auto& t : c->timings)
    if (t.pending) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
        t.total_ms += ms;
        ++t.calls;
      }
      t.pending = false;
    }
}
}  // namespace

namespace {

template <class T>
int dalloc(phz_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  const size_t bytes = sizeof(T) * (count ? count : 1);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    return -2;
  }
  c->allocs.push_back(q);
  c->alloc_bytes += bytes;
  *p = (T*)q;
  return 0;
}

// A synchronous copy on the context's own (non-blocking) stream.  Never the
// legacy stream: while another thread captures a graph, HIP rejects legacy
// stream work ("operation would make the legacy stream depend on a
// capturing blocking stream"), and contexts are created and used by
// concurrent threads (the request server, multi-stream callers).
int copy_sync(phz_ctx* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  PHZ_HIP(hipMemcpyAsync(dst, src, bytes, kind, c->stream));
  PHZ_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

template <class T>
int upload(phz_ctx* c, T** p, const std::vector<T>& v) {
  if (int rc = dalloc(c, p, v.size())) return rc;
  return copy_sync(c, *p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}

#define TRY(x)                \
  do {                        \
    int _rc = (x);            \
    if (_rc) return _rc;      \
  } while (0)

// peak-window neighbours of the chosen evaluations: BinghamPeakBasedEval /
// BmmPeakBasedEval and GaussianPeakBasedEval / GmmPeakBasedEval
int rot_neighbors(const phz_config& cf) {
  return cf.rot_eval == PHZ_EVAL_MIXTURE ? cf.bmm_peak_neighbors : cf.bingham_peak_neighbors;
}
int pos_neighbors(const phz_config& cf) {
  return cf.pos_eval == PHZ_EVAL_MIXTURE ? cf.gmm_peak_neighbors : cf.gaussian_peak_neighbors;
}

// Device bytes init_ctx allocates for a configuration, in closed form (the
// same list of buffers; init_ctx checks that the two agree).
size_t context_bytes(const phz_config& cf) {
  const int64_t B = cf.bandwidth, n = 2 * B, N = cf.n_voxels;
  const int64_t pad = cf.spherical_zero_padding, Bo = B + pad, no = 2 * Bo;
  const int64_t nq = n * n, T3 = N * N * N, H = N * N * (N / 2 + 1);
  const int64_t sp = cf.spatial_zero_padding, Np = N + 2 * sp, T3p = Np * Np * Np;
  const int64_t Hp = Np * Np * (Np / 2 + 1);
  const int64_t P = cf.max_batch, M = 2 * P, cap = cf.max_points;
  const int64_t G = std::max<int64_t>((cf.channels == 4 ? 4 : 2) * M, 16);
  const int64_t MV = std::max<int64_t>(M, 16);
  const int64_t ppp = std::max(so3_gamma_blocks((int)Bo), spatial_c2r_blocks((int)Np));
  size_t s = 0;
  auto add = [&](int64_t count, size_t elem) { s += elem * (size_t)(count ? count : 1); };
  // tables
  add(nq, sizeof(float4));
  add(n, sizeof(float));
  add(n * B * (B + 1) / 2, sizeof(float));
  add(B + 1, sizeof(int));
  add(no * Bo * (Bo + 1) * (Bo + 2) / 6, sizeof(float));
  add(Bo * (Bo + 1) / 2 + 1, sizeof(int64_t));
  add(Bo, sizeof(float));
  if (cf.precision == 1) {
    std::vector<int64_t> hl_off;
    add(wig_split_offsets((int)Bo, &hl_off), sizeof(uint16_t));
    add(Bo * (Bo + 1) / 2, sizeof(int64_t));
  }
  add(n, sizeof(float2));
  add(N, sizeof(float2));
  add(N, sizeof(float));
  if (pad > 0) {
    add(no, sizeof(float2));
    add(Bo * Bo, sizeof(int));
    add(std::max<int64_t>(M, 2) * Bo * Bo, sizeof(float2));
  }
  // clouds and the S^2 search grid
  add(M * cap, sizeof(float4));
  add(M, sizeof(int));
  add(M, sizeof(GatherEntry));
  add(M * (nq + 3 * n), sizeof(int));
  add(M * (nq + 1), sizeof(int));
  add(M * (n + 1), sizeof(int));
  add(M * nq, sizeof(int4));
  add(M * nq, sizeof(int4));
  add(M * cap + kNnScanPad, sizeof(float4));
  add(M * cap, sizeof(float4));
  add(M * cap, sizeof(int));
  add(M * kSeedCells, sizeof(int));
  add(M * n * ((n + 63) / 64), sizeof(unsigned long long));
  // spherical spectra and SO(3) grids
  add(G * nq, sizeof(float));
  add(G * B * n, sizeof(float2));
  add(G * B * B, sizeof(float2));
  add(std::max<int64_t>(M, 2) * B * B, sizeof(float2));
  add(P * Bo * (2 * Bo - 1) * no, sizeof(float2));
  add(P * no * Bo * no, sizeof(float2));
  add(no * no * no, sizeof(float));
  add(2 * P, sizeof(unsigned long long));
  add(P * ppp, sizeof(unsigned long long));
  add(P, sizeof(DevPose));
  // voxel accumulators, spectra and the translation grid
  add(MV * T3, sizeof(float));
  add(M * T3, sizeof(float));
  add(M * T3, sizeof(uint32_t));
  add((M * N * N + 3) / 4 * 4, 1);
  add((M * N + 3) / 4 * 4, 1);
  add(std::max<int64_t>(2 * M, 16) * H, sizeof(float2));
  add(P * H, sizeof(float2));
  add(T3p, sizeof(float));
  add(cap, sizeof(float4));
  if (sp > 0) {
    add(Np, sizeof(float2));
    add(P * Hp, sizeof(float2));
  }
  if (rot_neighbors(cf) > 0) add(P * no * no * no, sizeof(float));
  if (pos_neighbors(cf) > 0) add(P * T3p, sizeof(float));
  if (cf.channels == 4) {
    add(M * cap, sizeof(float2));
    add(M * T3, sizeof(float));
    add(M * T3, sizeof(float));
    add(std::max<int64_t>(2 * M, 16) * H, sizeof(float2));
  }
  return s;
}

int check_config(const phz_config& cf) {
  PHZ_CHECK(cf.bandwidth >= 2 && cf.bandwidth <= kMaxBandwidth, "bandwidth must be in [2, 160]");
  PHZ_CHECK(cf.spherical_zero_padding >= 0 &&
                cf.bandwidth + cf.spherical_zero_padding <= kMaxBandwidth,
            "spherical zero padding must be >= 0 with bandwidth + padding <= 160");
  PHZ_CHECK(cf.n_voxels >= 2 && cf.n_voxels <= 256, "n_voxels must be in [2, 256]");
  PHZ_CHECK(cf.spatial_zero_padding >= 0 &&
                cf.n_voxels + 2 * (int64_t)cf.spatial_zero_padding <= 256,
            "spatial zero padding must be >= 0 with n_voxels + 2 * padding <= 256");
  PHZ_CHECK(cf.discretize_lower < cf.discretize_upper, "discretize bounds must be increasing");
  PHZ_CHECK(cf.max_points >= 1 && cf.max_points <= (1ll << 26), "max_points out of range");
  PHZ_CHECK(cf.max_batch >= 1 && cf.max_batch <= 1024, "max_batch out of range");
  PHZ_CHECK(cf.precision == 0 || cf.precision == 1, "precision must be 0 (fp32) or 1 (split fp16)");
  PHZ_CHECK(cf.bingham_peak_neighbors >= 0 && cf.bingham_peak_neighbors <= (1 << 20),
            "bingham_peak_neighbors must be in [0, 2^20]");
  PHZ_CHECK(cf.gaussian_peak_neighbors >= 0 && cf.gaussian_peak_neighbors <= (1 << 20),
            "gaussian_peak_neighbors must be in [0, 2^20]");
  PHZ_CHECK(cf.bmm_peak_neighbors >= 0 && cf.bmm_peak_neighbors <= (1 << 20),
            "bmm_peak_neighbors must be in [0, 2^20]");
  PHZ_CHECK(cf.gmm_peak_neighbors >= 0 && cf.gmm_peak_neighbors <= (1 << 20),
            "gmm_peak_neighbors must be in [0, 2^20]");
  PHZ_CHECK(cf.rot_eval == PHZ_EVAL_BINGHAM || cf.rot_eval == PHZ_EVAL_MIXTURE,
            "rot_eval must be PHZ_EVAL_BINGHAM (0) or PHZ_EVAL_MIXTURE (1)");
  PHZ_CHECK(cf.pos_eval == PHZ_EVAL_GAUSSIAN || cf.pos_eval == PHZ_EVAL_MIXTURE,
            "pos_eval must be PHZ_EVAL_GAUSSIAN (0) or PHZ_EVAL_MIXTURE (1)");
  PHZ_CHECK(cf.channels == 0 || cf.channels == 2 || cf.channels == 4,
            "channels must be 2 or 4 (0 means 2)");
  PHZ_CHECK(cf.method == PHZ_METHOD_SPH_OPT || cf.method == PHZ_METHOD_SPH,
            "method must be PHZ_METHOD_SPH_OPT (0) or PHZ_METHOD_SPH (1)");
  PHZ_CHECK(cf.skip_rot_axes >= 0 && cf.skip_rot_axes <= 7, "skip_rot_axes must be in [0, 7]");
  PHZ_CHECK(cf.method == PHZ_METHOD_SPH || cf.skip_rot_axes == 0,
            "skip_rot_axes (refine_rot_x/y/z) belongs to PHZ_METHOD_SPH");
  PHZ_CHECK(cf.method != PHZ_METHOD_SPH || cf.spherical_zero_padding == 0,
            "PHZ_METHOD_SPH correlates without spherical zero padding");
  return 0;
}

int init_ctx(phz_ctx* c) {
  const phz_config& cf = c->cfg;
  TRY(check_config(cf));
  c->B = cf.bandwidth;
  c->n = 2 * c->B;
  c->Bo = c->B + cf.spherical_zero_padding;
  c->no = 2 * c->Bo;
  c->N = cf.n_voxels;
  c->h = c->N / 2 + 1;
  c->H = (int64_t)c->N * c->N * c->h;
  c->T3 = (int64_t)c->N * c->N * c->N;
  c->sp = cf.spatial_zero_padding;
  c->Np = c->N + 2 * c->sp;
  c->T3p = (int64_t)c->Np * c->Np * c->Np;
  c->Hp = (int64_t)c->Np * c->Np * (c->Np / 2 + 1);
  c->Tc = c->sp ? padded_size(c->N, c->sp) : (uint32_t)c->T3;
  c->cap = cf.max_points;
  c->max_clouds = 2 * cf.max_batch;
  c->plan_n = make_plan(c->n);
  c->plan_no = make_plan(c->no);
  c->plan_N = make_plan(c->N);
  c->plan_Np = make_plan(c->Np);
  {
    const float v = (float)(std::abs(cf.low_pass_lower) + std::abs(cf.low_pass_upper)) /
                    (float)cf.n_voxels;  // point-cloud.cc:66-72
    c->max_sq = v * v;
  }
  PHZ_HIP(hipSetDevice(cf.device));
  PHZ_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  // the no-allocation entry point's device array: its address and twiddle
  // tables, once per device (this device is current)
  TRY(noalloc_setup(cf.device, c->stream));
  build_s2_tables(c->B, &c->s2);
  build_voxel_tables(c->N, (float)cf.discretize_lower, (float)cf.discretize_upper, &c->vox);
  for (int i = 1; i < c->N; ++i)  // the voxel binning relies on it
    PHZ_CHECK(c->vox.edges[i] > c->vox.edges[i - 1], "voxel edges must be strictly increasing");
  PHZ_CHECK(c->s2.runs.count > 0 && c->vox.runs.count > 0, "fusion structure has too many runs");
  {
    std::vector<float4> g(c->s2.grid.size() / 4);
    std::memcpy(g.data(), c->s2.grid.data(), sizeof(float) * c->s2.grid.size());
    TRY(upload(c, &c->d_grid, g));
  }
  TRY(upload(c, &c->d_theta, c->s2.theta));
  TRY(upload(c, &c->d_leg, c->s2.legendre));
  TRY(upload(c, &c->d_leg_off, c->s2.leg_off));
  // the SO(3) tables at Bo (with padding: the Wigner-d blocks and beta grid
  // of the larger bandwidth; the S^2 tables stay at B)
  S2Tables so3_pad;
  if (c->Bo != c->B) build_s2_tables(c->Bo, &so3_pad);
  const S2Tables& so3t = c->Bo != c->B ? so3_pad : c->s2;
  TRY(upload(c, &c->d_wig, so3t.wigner));
  TRY(upload(c, &c->d_wig_off, so3t.wig_off));
  TRY(upload(c, &c->d_wig_norm, so3t.wig_norm));
  if (cf.precision == 1) {
    std::vector<int64_t> hl_off;
    const int64_t hl_halves = wig_split_offsets(c->Bo, &hl_off);
    TRY(upload(c, &c->d_wig_hl_off, hl_off));
    TRY(dalloc(c, &c->d_wig_hl, (size_t)hl_halves));
    launch_wig_split(c->Bo, c->d_wig, c->d_wig_off, c->d_wig_hl_off, c->d_wig_hl, c->stream);
    PHZ_HIP(hipGetLastError());
  }
  {
    std::vector<float2> t(c->n), u(c->N);
    std::memcpy(t.data(), c->s2.tw_n.data(), sizeof(float2) * c->n);
    std::memcpy(u.data(), c->vox.tw_N.data(), sizeof(float2) * c->N);
    TRY(upload(c, &c->d_tw_n, t));
    if (c->Bo != c->B) {
      std::vector<float2> to(c->no);
      std::memcpy(to.data(), so3t.tw_n.data(), sizeof(float2) * c->no);
      TRY(upload(c, &c->d_tw_no, to));
      // Bo-layout index j of (m, l) -> the B-layout index, or -1 for l >= B
      // (so3CombineCoef_fftw with degLim = B - 1, so3_correlate_fftw.c:102-118)
      std::vector<int> map((size_t)c->Bo * c->Bo, -1);
      for (int l = 0; l < c->Bo; ++l)
        for (int m = -l; m <= l; ++m)
          map[seanindex(m, l, c->Bo)] = l < c->B ? seanindex(m, l, c->B) : -1;
      TRY(upload(c, &c->d_pad_map, map));
    } else {
      c->d_tw_no = c->d_tw_n;
    }
    TRY(upload(c, &c->d_tw_N, u));
    if (c->sp) {
      const std::vector<float> tp = twiddles(c->Np);
      std::vector<float2> v(c->Np);
      std::memcpy(v.data(), tp.data(), sizeof(float2) * c->Np);
      TRY(upload(c, &c->d_tw_Np, v));
    } else {
      c->d_tw_Np = c->d_tw_N;
    }
  }
  TRY(upload(c, &c->d_edges, c->vox.edges));
  const int M = c->max_clouds, P = cf.max_batch;
  const int64_t cap = c->cap, nq = (int64_t)c->n * c->n;
  // NN search grid at the query resolution: measured on gfx950 at B = 64,
  // an (n/2)^2 grid (fewer cells, twice the points per cell) made the
  // search 2.5x slower
  c->cells.nb = c->n;
  c->cells.ns = c->n;
  TRY(dalloc(c, &c->d_pts, (size_t)M * cap));
  c->cells.pts = c->d_pts;
  TRY(dalloc(c, &c->d_npts, (size_t)M));
  TRY(dalloc(c, &c->d_gather, (size_t)M));
  TRY(dalloc(c, &c->cells.cell_count, (size_t)M * (nq + 3 * c->n)));
  c->cells.band_th = (unsigned*)(c->cells.cell_count + (size_t)M * nq);
  c->cells.band_gap = c->cells.band_th + (size_t)M * 2 * c->n;
  TRY(dalloc(c, &c->cells.cell_start, (size_t)M * (nq + 1)));
  TRY(dalloc(c, &c->cells.band_start, (size_t)M * (c->n + 1)));
  TRY(dalloc(c, &c->cells.next_right, (size_t)M * nq));
  TRY(dalloc(c, &c->cells.next_left, (size_t)M * nq));
  TRY(dalloc(c, &c->cells.sorted, (size_t)M * cap + kNnScanPad));
  TRY(dalloc(c, &c->cells.proj, (size_t)M * cap));
  TRY(dalloc(c, &c->cells.cell, (size_t)M * cap));
  TRY(dalloc(c, &c->cells.seed_pos, (size_t)M * kSeedCells));
  TRY(dalloc(c, &c->cells.band_mask, (size_t)M * c->n * ((c->n + 63) / 64)));
  // stage entry points use up to 16 grids / signals; with four channels the
  // reflectivity and ambient grids of all clouds follow the range and
  // intensity grids
  const int nch = cf.channels == 4 ? 4 : 2;
  const int G = M * nch > 16 ? M * nch : 16;
  TRY(dalloc(c, &c->d_grids, (size_t)G * nq));
  TRY(dalloc(c, &c->d_F, (size_t)G * c->B * c->n));
  TRY(dalloc(c, &c->d_coef, (size_t)G * c->B * c->B));
  TRY(dalloc(c, &c->d_fused, (size_t)(M > 2 ? M : 2) * c->B * c->B));
  const int64_t Bo = c->Bo, no = c->no;
  if (c->Bo != c->B)
    TRY(dalloc(c, &c->d_fused_o, (size_t)(M > 2 ? M : 2) * Bo * Bo));
  else
    c->d_fused_o = c->d_fused;
  TRY(dalloc(c, &c->d_S, (size_t)P * Bo * (2 * Bo - 1) * no));
  TRY(dalloc(c, &c->d_U, (size_t)P * no * Bo * no));
  TRY(dalloc(c, &c->d_corr_rot, (size_t)no * no * no));
  TRY(dalloc(c, &c->d_keys, (size_t)2 * P));
  c->partials_per_pair = std::max(so3_gamma_blocks(c->Bo), spatial_c2r_blocks(c->Np));
  TRY(dalloc(c, &c->d_partials, (size_t)P * c->partials_per_pair));
  TRY(dalloc(c, &c->d_pose, (size_t)P));
  const int MV = M > 16 ? M : 16;
  TRY(dalloc(c, &c->d_sumI, (size_t)MV * c->T3));
  TRY(dalloc(c, &c->d_sumR, (size_t)M * c->T3));
  TRY(dalloc(c, &c->d_cnt, (size_t)M * c->T3));
  const size_t occ_row_bytes = ((size_t)M * c->N * c->N + 3) / 4 * 4;
  const size_t occ_plane_bytes = ((size_t)M * c->N + 3) / 4 * 4;
  TRY(dalloc(c, &c->d_row_occ, occ_row_bytes));
  TRY(dalloc(c, &c->d_plane_occ, occ_plane_bytes));
  TRY(dalloc(c, &c->d_spec, (size_t)(2 * M > 16 ? 2 * M : 16) * c->H));
  TRY(dalloc(c, &c->d_work, (size_t)P * c->H));
  TRY(dalloc(c, &c->d_corr_trans, (size_t)c->T3p));
  TRY(dalloc(c, &c->d_reg, (size_t)cap));
  if (c->sp) TRY(dalloc(c, &c->d_work_pad, (size_t)P * c->Hp));
  if (rot_neighbors(cf) > 0)
    TRY(dalloc(c, &c->d_peak_rot, (size_t)P * no * no * no));
  if (pos_neighbors(cf) > 0)
    TRY(dalloc(c, &c->d_peak_trans, (size_t)P * c->T3p));
  if (cf.channels == 4) {
    TRY(dalloc(c, &c->d_extra, (size_t)M * cap));
    TRY(dalloc(c, &c->d_sumF, (size_t)M * c->T3));
    TRY(dalloc(c, &c->d_sumA, (size_t)M * c->T3));
    TRY(dalloc(c, &c->d_spec2, (size_t)(2 * M > 16 ? 2 * M : 16) * c->H));
  }
  PHZ_HIP(hipHostMalloc((void**)&c->h_npts, sizeof(int) * M, hipHostMallocDefault));
  PHZ_HIP(hipHostMalloc((void**)&c->h_gather, sizeof(GatherEntry) * M, hipHostMallocDefault));
  PHZ_HIP(hipEventCreateWithFlags(&c->gather_done, hipEventDisableTiming));
  // voxel accumulators stay zero between calls: every pipeline run clears
  // exactly the voxels it touched
  launch_fill_u32((uint32_t*)c->d_sumI, (size_t)MV * c->T3, 0u, c->stream);
  launch_fill_u32((uint32_t*)c->d_sumR, (size_t)M * c->T3, 0u, c->stream);
  launch_fill_u32(c->d_cnt, (size_t)M * c->T3, 0u, c->stream);
  launch_fill_u32((uint32_t*)c->d_row_occ, occ_row_bytes / 4, 0u, c->stream);
  launch_fill_u32((uint32_t*)c->d_plane_occ, occ_plane_bytes / 4, 0u, c->stream);
  if (cf.channels == 4) {
    launch_fill_u32((uint32_t*)c->d_sumF, (size_t)M * c->T3, 0u, c->stream);
    launch_fill_u32((uint32_t*)c->d_sumA, (size_t)M * c->T3, 0u, c->stream);
  }
  PHZ_HIP(hipStreamSynchronize(c->stream));
  PHZ_CHECK(c->alloc_bytes == context_bytes(cf), "internal error: device memory estimate is stale");
  c->tuning.fwd_x = c->tuning.fwd_z = c->tuning.inv_x = fft_cols_default_tile(c->N);
  c->plane_fwd = false;  // measured slower at N = 128 (one workgroup per CU)
  c->tuning.plane_fwd = -1;
  c->tuning.nn_lanes = 0;  // automatic, by clouds per launch
  c->tiles.fused_reg = false;
  c->tuning.fused_reg = -1;
  c->so3_reg = false;
  c->tuning.so3_reg = -1;
  c->occ_fwd = true;
  c->tuning.occ_fwd = 1;
  c->inv_plane = true;
  c->tuning.inv_plane = 1;
  c->nn_xcd = kNnXcdDefault;
  c->tuning.nn_xcd = kNnXcdDefault ? 1 : -1;
  c->nn_seed = kNnSeedDefault;
  c->tuning.nn_seed = kNnSeedDefault ? kNnSeedDefault : -1;
  c->xz_slab = false;
  c->tuning.xz_slab = -1;
  return 0;
}

// occupancy bytes of cloud slots [first, ...)
VoxelOcc occ_from(phz_ctx* c, int first) {
  VoxelOcc o;
  o.row = c->d_row_occ + (int64_t)first * c->N * c->N;
  o.plane = c->d_plane_occ + (int64_t)first * c->N;
  return o;
}

// reflectivity / ambient of cloud slots [first, ...) (none with two channels)
VoxelExtra extra_from(phz_ctx* c, int first) {
  VoxelExtra e;
  if (c->cfg.channels == 4) {
    e.extra = c->d_extra + (int64_t)first * c->cap;
    e.sumF = c->d_sumF + (int64_t)first * c->T3;
    e.sumA = c->d_sumA + (int64_t)first * c->T3;
  }
  return e;
}

ChannelSet channels(const float2* a, const float2* b, int C) {
  ChannelSet cs{};
  cs.C = C;
  cs.ch[0] = a;
  cs.ch[1] = b;
  return cs;
}

// The registration pipeline on cloud slots [0, nt) targets and
// [nt, nt+np) sources already resident in ctx->d_pts / d_npts.
int run_pairs(phz_ctx* c, int np, bool shared, hipStream_t s, float* d_rot_corr,
              float* d_trans_corr) {
  // the peak-window fits read the grids: written to the context's own
  // buffers when the caller wants none
  const int kb = rot_neighbors(c->cfg), kg = pos_neighbors(c->cfg);
  const bool bmm = c->cfg.rot_eval == PHZ_EVAL_MIXTURE, gmm = c->cfg.pos_eval == PHZ_EVAL_MIXTURE;
  if (!d_rot_corr && kb > 0) d_rot_corr = c->d_peak_rot;
  if (!d_trans_corr && kg > 0) d_trans_corr = c->d_peak_trans;
  const int nt = shared ? 1 : np;
  const int ncl = nt + np;
  const int B = c->B, n = c->n, N = c->N;
  const int64_t nq = (int64_t)n * n, BB = (int64_t)B * B;
  // four channels: each cloud's reflectivity and ambient grids follow all
  // clouds' range and intensity grids (and so do their coefficients)
  const bool four = c->cfg.channels == 4;
  // rotation
  {
    StageScope t(c, kS2Cells, s);
    launch_s2_cells(c->d_pts, c->d_npts, c->cap, ncl, n, c->cells, s);
  }
  // SphRegistration: one sampled channel 0.6 intensity + 0.4 range per cloud
  // (spherical-correlation.cc:103-112), transformed straight into the
  // correlation's coefficients; no fusion, and reflectivity / ambient take
  // no part in the rotation
  const bool sph = c->cfg.method == PHZ_METHOD_SPH;
  {
    StageScope t(c, kS2NN, s);
    launch_s2_nn(c->cap, ncl, n, c->d_grid, c->d_theta, c->max_sq, c->cells, c->d_grids,
                 c->tuning.nn_lanes, s, four && !sph ? c->d_extra : nullptr,
                 four && !sph ? c->d_grids + 2 * ncl * nq : nullptr, c->nn_xcd, sph,
                 c->nn_seed);
  }
  {
    StageScope t(c, kS2Transform, s);
    if (sph)
      launch_fst(c->d_grids, ncl, B, c->d_leg, c->d_leg_off, c->d_tw_n, c->d_F, c->d_fused,
                 c->cfg.precision, s);
    else
      launch_fst(c->d_grids, (four ? 4 : 2) * ncl, B, c->d_leg, c->d_leg_off, c->d_tw_n, c->d_F,
                 c->d_coef, c->cfg.precision, s);
  }
  if (!sph) {  // spherical order {range, intensity, refl, amb} (refl = amb = 0 with
            // two channels), one launch for all clouds
    StageScope t(c, kFuse, s);
    ChannelSet cs = channels(c->d_coef, c->d_coef + BB, 4);
    if (four) {
      cs.ch[2] = c->d_coef + 2 * ncl * BB;
      cs.ch[3] = c->d_coef + 2 * ncl * BB + BB;
    }
    launch_fuse(cs, (uint32_t)BB, c->s2.runs, c->d_fused, s, ncl, 2 * BB);
    // spherical zero padding: the coefficients of degree < B in the layout of
    // the correlation bandwidth Bo, zero above
    if (c->Bo != B)
      launch_pad_coef(c->d_fused, BB, c->d_pad_map, (int64_t)c->Bo * c->Bo, c->d_fused_o, ncl, s);
  }
  // the SO(3) correlation at Bo = B + spherical_zero_padding
  const int Bo = c->Bo;
  const int64_t BBo = (int64_t)Bo * Bo, no3 = (int64_t)c->no * c->no * c->no;
  {
    StageScope t(c, kSo3Synth, s);
    launch_so3_synth(Bo, c->d_fused_o, shared ? 0 : BBo, c->d_fused_o + nt * BBo, BBo, np,
                     c->d_wig, c->d_wig_off, c->d_wig_norm, c->d_S, c->cfg.precision, s,
                     c->d_wig_hl, c->d_wig_hl_off);
  }
  {
    StageScope t(c, kSo3Alpha, s);
    launch_so3_fft_alpha(Bo, np, c->d_S, c->plan_no, c->d_tw_no, c->d_U, s, c->so3_reg);
  }
  {
    StageScope t(c, kSo3Gamma, s);
    launch_so3_fft_gamma(Bo, np, c->d_U, c->plan_no, c->d_tw_no, d_rot_corr, no3, c->d_partials,
                         c->d_keys, s, c->so3_reg);
  }
  {
    StageScope t(c, kRotPose, s);
    launch_rot_pose(Bo, np, c->d_keys, d_rot_corr, no3, kb, c->d_pose, s, c->cfg.skip_rot_axes,
                    bmm);
  }
  // translation (accumulators are zero on entry and cleared again below)
  {
    StageScope t(c, kVoxelScatter, s);
    launch_voxel_scatter(c->d_pts, c->d_npts, c->cap, nt, nullptr, 0, N, c->d_edges, c->d_sumI,
                         c->d_sumR, c->d_cnt, occ_from(c, 0), s, extra_from(c, 0));
    launch_voxel_scatter(c->d_pts + nt * c->cap, c->d_npts + nt, c->cap, np, c->d_pose, 1, N,
                         c->d_edges, c->d_sumI + nt * c->T3, c->d_sumR + nt * c->T3,
                         c->d_cnt + nt * c->T3, occ_from(c, nt), s, extra_from(c, nt));
  }
  // spatial order {intensity, range, refl=0, amb=0}; pair p: target p (or
  // the shared target 0) against source nt+p
  const float2* F = c->d_spec;
  const float2* G = c->d_spec + (int64_t)(2 * nt) * c->H;
  const ChannelSet Fc = channels(F, F + c->H, 4), Gc = channels(G, G + c->H, 4);
  const int64_t fs = shared ? 0 : 2 * c->H, gs = 2 * c->H;
  // zz: the forward z pass runs inside the cross-spectrum kernel; occ: the
  // r2c and x passes before it also skip rows / planes without points
  // four channels: the reflectivity / ambient spectra in d_spec2 (layout of
  // d_spec) take the same passes; channels 2 and 3 of both sides come from it
  ChannelSet F4 = Fc, G4 = Gc;
  if (four) {
    F4.ch[2] = c->d_spec2;
    F4.ch[3] = c->d_spec2 + c->H;
    G4.ch[2] = c->d_spec2 + (int64_t)(2 * nt) * c->H;
    G4.ch[3] = G4.ch[2] + c->H;
  }
  // spatial zero padding: dense forward passes, the cross spectrum alone,
  // then the padded inverse at Np (below)
  const bool padded = c->sp > 0;
  const bool zz = !padded && c->occ_fwd && c->tiles.fused_reg &&
                  c->tuning.fwd_z == kRegisterEngine &&
                  (four ? spatial_cross_zz4_ok(N, F4, G4) : spatial_cross_zz_ok(N, Fc, Gc));
  const bool occ = zz && !c->plane_fwd && c->tuning.fwd_x == kRegisterEngine;
  const bool plane = c->plane_fwd && !four;
  const bool inv_plane = !padded && c->inv_plane && c->tiles.fused_reg &&
                         c->tuning.inv_x == kRegisterEngine && spatial_inv_plane_ok(N);
  // slab: the x pass, the forward z pass, fusion, cross spectrum and
  // inverse z pass in one kernel per (pair, k2), written in the slab layout
  // that only the inverse plane pass reads
  const bool slab = occ && !four && inv_plane && c->xz_slab && spatial_cross_slab_ok(N, Fc, Gc);
  {  // y (and with the plane pass also x) first, then the remaining columns
    StageScope t(c, kSpatialR2c, s);
    if (plane)
      launch_spatial_plane_fwd(N, ncl, c->d_sumI, c->d_sumR, c->d_cnt, c->plan_N, c->d_tw_N,
                               c->d_spec, s);
    else if (occ) {
      launch_spatial_r2c_rows(N, ncl, c->d_sumI, c->d_sumR, c->d_cnt, occ_from(c, 0), c->d_tw_N,
                              c->d_spec, s);
      if (four)
        launch_spatial_r2c_rows(N, ncl, c->d_sumF, c->d_sumA, c->d_cnt, occ_from(c, 0),
                                c->d_tw_N, c->d_spec2, s);
    } else if (c->tiles.fused_reg) {
      launch_spatial_r2c_reg(N, ncl, c->d_sumI, c->d_sumR, c->d_cnt, c->d_row_occ, false,
                             c->d_tw_N, c->d_spec, s);
      if (four)
        launch_spatial_r2c_reg(N, ncl, c->d_sumF, c->d_sumA, c->d_cnt, c->d_row_occ, false,
                               c->d_tw_N, c->d_spec2, s);
    } else {
      launch_spatial_r2c(N, ncl, c->d_sumI, c->d_sumR, c->d_cnt, c->plan_N, c->d_tw_N,
                         c->d_spec, s);
      if (four)
        launch_spatial_r2c(N, ncl, c->d_sumF, c->d_sumA, c->d_cnt, c->plan_N, c->d_tw_N,
                           c->d_spec2, s);
    }
  }
  {
    StageScope t(c, kSpatialFftFwd, s);
    for (int set = 0; set < (four ? 2 : 1); ++set) {
      float2* spec = set ? c->d_spec2 : c->d_spec;
      if (occ) {
        if (!slab) launch_spatial_fft_x_occ(N, 2 * ncl, spec, occ_from(c, 0), c->d_tw_N, s);
      } else if (!plane)
        launch_spatial_fft_pass(N, 2 * ncl, spec, 0, c->tiles.fwd_x, c->plan_N, c->d_tw_N, s);
      if (!zz)
        launch_spatial_fft_pass(N, 2 * ncl, spec, 1, c->tiles.fwd_z, c->plan_N, c->d_tw_N, s);
    }
  }
  {
    {
      StageScope t(c, kSpatialCrossZ, s);
      if (slab)
        launch_spatial_cross_slab(N, np, Fc, fs, Gc, gs, occ_from(c, 0).row, occ_from(c, 0).plane,
                                  shared, occ_from(c, nt).row, occ_from(c, nt).plane, c->vox.runs,
                                  c->d_tw_N, c->d_work, s);
      else if (padded)
        launch_spatial_cross_product(N, np, four ? F4 : Fc, fs, four ? G4 : Gc, gs, c->vox.runs,
                                     true, c->d_work, s);
      else if (four && zz)
        launch_spatial_cross_zz4_reg(N, np, F4, fs, G4, gs, occ ? c->d_plane_occ : nullptr,
                                     shared ? 0 : N,
                                     occ ? c->d_plane_occ + (int64_t)nt * N : nullptr, N,
                                     c->vox.runs, c->d_tw_N, c->d_work, s);
      else if (four && c->tiles.fused_reg)
        launch_spatial_cross_z_reg(N, np, F4, fs, G4, gs, c->vox.runs, true, c->d_tw_N,
                                   c->d_work, s);
      else if (four)
        launch_spatial_cross_z(N, np, F4, fs, G4, gs, c->vox.runs, true, c->plan_N, c->d_tw_N,
                               c->d_work, s);
      else if (zz)
        launch_spatial_cross_zz_reg(N, np, Fc, fs, Gc, gs, occ ? c->d_plane_occ : nullptr,
                                    shared ? 0 : N, occ ? c->d_plane_occ + (int64_t)nt * N : nullptr,
                                    N, c->vox.runs, c->d_tw_N, c->d_work, s);
      else if (c->tiles.fused_reg)
        launch_spatial_cross_z_reg(N, np, Fc, fs, Gc, gs, c->vox.runs, true, c->d_tw_N, c->d_work,
                                   s);
      else
        launch_spatial_cross_z(N, np, Fc, fs, Gc, gs, c->vox.runs, true, c->plan_N, c->d_tw_N,
                               c->d_work, s);
    }
    // after the last reader of the occupancy bytes (the cross spectrum
    // reads the plane bytes)
    {
      StageScope t(c, kVoxelClear, s);
      launch_voxel_clear(c->d_pts, c->d_npts, c->cap, nt, nullptr, 0, N, c->d_edges, c->d_sumI,
                         c->d_sumR, c->d_cnt, occ_from(c, 0), s, extra_from(c, 0));
      launch_voxel_clear(c->d_pts + nt * c->cap, c->d_npts + nt, c->cap, np, c->d_pose, 1, N,
                         c->d_edges, c->d_sumI + nt * c->T3, c->d_sumR + nt * c->T3,
                         c->d_cnt + nt * c->T3, occ_from(c, nt), s, extra_from(c, nt));
    }
    if (padded) {  // spatial-correlation.cc:142-161,221-230; spatial-correlation-laplace.cc:36-37
      const int Np = c->Np, tile = fft_cols_default_tile(Np);
      {
        StageScope t(c, kSpatialFftInv, s);
        launch_spatial_pad_gather(N, c->sp, true, np, c->d_work, c->d_work_pad, s);
        launch_spatial_fft_pass(Np, np, c->d_work_pad, 3, tile, c->plan_Np, c->d_tw_Np, s);
        launch_spatial_fft_pass(Np, np, c->d_work_pad, 2, tile, c->plan_Np, c->d_tw_Np, s);
      }
      StageScope t(c, kSpatialC2r, s);
      launch_spatial_c2r(Np, np, c->d_work_pad, c->plan_Np, c->d_tw_Np, d_trans_corr,
                         c->d_partials, c->d_keys + np, s, c->Tc);
    } else if (inv_plane) {  // inverse x and c2r in one pass over whole z planes
      StageScope t(c, kSpatialC2r, s);
      launch_spatial_inv_plane(N, np, c->d_work, c->d_tw_N, d_trans_corr, c->d_partials,
                               c->d_keys + np, s, slab);
    } else {
      {
        StageScope t(c, kSpatialFftInv, s);
        launch_spatial_fft_inv(N, np, c->d_work, c->plan_N, c->d_tw_N, c->tiles, s);
      }
      StageScope t(c, kSpatialC2r, s);
      if (c->tiles.fused_reg)
        launch_spatial_c2r_reg(N, np, c->d_work, c->d_tw_N, d_trans_corr, c->d_partials,
                               c->d_keys + np, s);
      else
        launch_spatial_c2r(N, np, c->d_work, c->plan_N, c->d_tw_N, d_trans_corr, c->d_partials,
                           c->d_keys + np, s);
    }
  }
  {
    StageScope t(c, kTransPose, s);
    // with padding the window and the metres use Np voxels and the
    // correlation vector's Tc entries (phase-correlation-eval.cc:19-27)
    launch_trans_pose(c->Np, c->cfg.discretize_lower, c->cfg.discretize_upper, np,
                      c->d_keys + np, d_trans_corr, c->T3p, kg, c->d_pose, s, c->Tc, gmm);
  }
  PHZ_HIP(hipGetLastError());
  return 0;
}

void clear_graphs(phz_ctx* c) {
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
}

// run_pairs through a captured graph when graphs are on (and stage timing is
// off): one graph per (pairs, shared target, output grids); capture records
// kernel launches only (fills are kernels, no memset nodes)
int exec_pairs(phz_ctx* c, int np, bool shared, hipStream_t s, float* rot, float* trans) {
  if (!c->use_graphs || c->profiling) return run_pairs(c, np, shared, s, rot, trans);
  for (const auto& g : c->graphs)
    if (g.np == np && g.shared == shared && g.rot == rot && g.trans == trans) {
      PHZ_HIP(hipGraphLaunch(g.exec, s));
      return 0;
    }
  hipGraph_t graph = nullptr;
  PHZ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = run_pairs(c, np, shared, s, rot, trans);
  const hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) {
    set_error(std::string("graph capture failed: ") + hipGetErrorString(e));
    return -2;
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    set_error(std::string("graph instantiation failed: ") + hipGetErrorString(ei));
    return -2;
  }
  if (c->graphs.size() >= 32) {
    (void)hipGraphExecDestroy(c->graphs.front().exec);
    c->graphs.erase(c->graphs.begin());
  }
  c->graphs.push_back({np, shared, rot, trans, exec});
  PHZ_HIP(hipGraphLaunch(exec, s));
  return 0;
}

std::string pair_status_text(int32_t st) {
  if (st & kPoseTransWindow)
    return "the Gaussian-mixture peak window reaches past the translation grid "
           "(gmm-peak-based-eval.cc:81 CHECK_LT(end, n))";
  if (st & kPoseRotWindow)
    return "the Bingham peak window reaches past the rotation grid "
           "(bingham-peak-based-eval.cc:69 CHECK_LT(end, n))";
  if (st & kPoseRotWeights)
    return "the Bingham peak window's weights do not sum to a positive value "
           "(bingham-peak-based-eval.cc:88 CHECK_GT)";
  return "the Gaussian peak window's weights do not sum to a positive value "
         "(gaussian-peak-based-eval.cc:100 CHECK_GT)";
}

void to_result(const DevPose& d, phz_result* r) {
  r->rot_index = d.rot_index;
  r->trans_index = d.trans_index;
  r->rot_value = d.rot_value;
  r->trans_value = d.trans_value;
  for (int i = 0; i < 3; ++i) {
    r->zyz[i] = d.zyz[i];
    r->euler_xyz[i] = d.euler[i];
    r->translation[i] = d.translation[i];
  }
  for (int i = 0; i < 4; ++i) {
    r->quat[i] = d.quat[i];
    r->rot_omega[i] = d.rot_omega[i];
  }
  for (int i = 0; i < 16; ++i) {
    r->T_rot[i] = d.T[i];
    r->rot_axes[i] = d.rot_axes[i];
  }
  for (int i = 0; i < 9; ++i) r->trans_cov[i] = d.trans_cov[i];
  static_assert(kPoseRotWindow == PHZ_PAIR_ROT_WINDOW && kPoseRotWeights == PHZ_PAIR_ROT_WEIGHTS &&
                    kPoseTransWeights == PHZ_PAIR_TRANS_WEIGHTS &&
                    kPoseTransWindow == PHZ_PAIR_TRANS_WINDOW,
                "pose status bits are the API's");
  r->status = d.status;
  r->reserved = 0;
}

int fetch_results(phz_ctx* c, int np, phz_result* out, hipStream_t s) {
  std::vector<DevPose> hp(np);
  PHZ_HIP(hipMemcpyAsync(hp.data(), c->d_pose, sizeof(DevPose) * np, hipMemcpyDeviceToHost, s));
  PHZ_HIP(hipStreamSynchronize(s));
  collect_timings(c);
  for (int p = 0; p < np; ++p) to_result(hp[p], &out[p]);
  // the reference's CHECKs on the peak windows (its process aborts there):
  // every result is filled, the first failing pair is named
  for (int p = 0; p < np; ++p) {
    const int32_t st = hp[p].status;
    if (!st) continue;
    set_error("pair " + std::to_string(p) + ": " + pair_status_text(st));
    return -4;
  }
  return 0;
}

// extra: (reflectivity, ambient) per point, or null (zero channels); only a
// context with four channels keeps them
int load_cloud(phz_ctx* c, int slot, const float* pts, int64_t P, hipMemcpyKind kind,
               hipStream_t s, const float* extra = nullptr) {
  PHZ_CHECK(P >= 0 && P <= c->cap, "point count exceeds max_points");
  PHZ_CHECK(P == 0 || pts != nullptr, "null point buffer");
  PHZ_CHECK(!extra || c->cfg.channels == 4,
            "reflectivity and ambient need a context created with channels = 4");
  if (P > 0)
    PHZ_HIP(hipMemcpyAsync(c->d_pts + slot * c->cap, pts, sizeof(float4) * P, kind, s));
  if (c->cfg.channels == 4 && P > 0) {
    if (extra)
      PHZ_HIP(hipMemcpyAsync(c->d_extra + slot * c->cap, extra, sizeof(float2) * P, kind, s));
    else
      launch_fill_u32((uint32_t*)(c->d_extra + slot * c->cap), (size_t)2 * P, 0u, s);
  }
  c->h_npts[slot] = (int)P;
  return 0;
}

// Stages slots [0, ncl) from device buffers: the table of pointers and
// counts goes up in one copy and one kernel gathers every cloud (instead
// of a copy per cloud, each a launch of its own).  Returns 1 without doing
// anything when a buffer is not aligned to its element (16 bytes for the
// points, 8 for the pairs): the caller then copies cloud by cloud.
int gather_clouds(phz_ctx* c, int ncl, const float* const* pts, const float* const* extra,
                  const int64_t* counts, hipStream_t s) {
  PHZ_CHECK(ncl >= 0 && ncl <= c->max_clouds, "cloud count exceeds the context's slots");
  int64_t max_n = 0;
  for (int i = 0; i < ncl; ++i) {
    const int64_t P = counts[i];
    PHZ_CHECK(P >= 0 && P <= c->cap, "point count exceeds max_points");
    PHZ_CHECK(P == 0 || pts[i] != nullptr, "null point buffer");
    PHZ_CHECK(!(extra && extra[i]) || c->cfg.channels == 4,
              "reflectivity and ambient need a context created with channels = 4");
    if (((uintptr_t)pts[i] & 15u) || (extra && ((uintptr_t)extra[i] & 7u))) return 1;
    max_n = P > max_n ? P : max_n;
  }
  if (c->gather_pending) PHZ_HIP(hipEventSynchronize(c->gather_done));
  for (int i = 0; i < ncl; ++i) {
    c->h_gather[i] = {(const float4*)pts[i], extra ? (const float2*)extra[i] : nullptr, counts[i]};
    c->h_npts[i] = (int)counts[i];
  }
  PHZ_HIP(hipMemcpyAsync(c->d_gather, c->h_gather, sizeof(GatherEntry) * ncl,
                         hipMemcpyHostToDevice, s));
  PHZ_HIP(hipEventRecord(c->gather_done, s));
  c->gather_pending = true;
  launch_gather_clouds(c->d_gather, ncl, max_n, c->d_pts, c->cfg.channels == 4 ? c->d_extra : nullptr,
                       c->cap, c->d_npts, s);
  PHZ_HIP(hipGetLastError());
  return 0;
}

// copies the staged counts of slots [0, ncl) to the device
int commit_counts(phz_ctx* c, int ncl, hipStream_t s) {
  PHZ_HIP(hipMemcpyAsync(c->d_npts, c->h_npts, sizeof(int) * ncl, hipMemcpyHostToDevice, s));
  return 0;
}

}  // namespace

namespace {
// device buffers of one stage call, freed on every return
struct CallBuffers {
  std::vector<void*> ptrs;
  ~CallBuffers() {
    for (void* p : ptrs) hipFree(p);
  }
  template <class T>
  hipError_t get(T** p, size_t count) {
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, sizeof(T) * (count ? count : 1));
    if (e == hipSuccess) ptrs.push_back(q);
    *p = (T*)q;
    return e;
  }
};
}  // namespace

extern "C" {

void phz_default_config(phz_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->bandwidth = 64;
  cfg->spherical_zero_padding = 0;
  cfg->n_voxels = 128;
  cfg->discretize_lower = -50;
  cfg->discretize_upper = 50;
  cfg->spatial_zero_padding = 0;
  cfg->low_pass_lower = 0;
  cfg->low_pass_upper = 100000;
  cfg->max_points = 262144;
  cfg->max_batch = 1;
  cfg->device = 0;
  cfg->channels = 2;
  cfg->bmm_peak_neighbors = 2;  // bmm-peak-based-eval.cc:7-9
  cfg->gmm_peak_neighbors = 2;  // gmm-peak-based-eval.cc:12-14
}

const char* phz_last_error(void) { return g_last_error.c_str(); }

void phz_zscore_default_params(phz_zscore_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->lag_percentile = 0.05;  // z-score-peak-extraction.cc:12-25
  p->lag_max = 20;
  p->threshold = 4.33;
  p->influence = 0.05;
  p->filter_threshold = 0.405;
}

int phz_create(const phz_config* cfg, phz_ctx** out) {
  PHZ_CHECK(cfg != nullptr && out != nullptr, "null argument");
  *out = nullptr;
  phz_ctx* c = new phz_ctx();
  c->cfg = *cfg;
  const int rc = init_ctx(c);
  if (rc) {
    const std::string msg = g_last_error;
    phz_destroy(c);
    set_error(msg);
    return rc;
  }
  *out = c;
  return 0;
}

void phz_destroy(phz_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  clear_graphs(c);
  for (auto& t : c->timings) {
    if (t.a) hipEventDestroy(t.a);
    if (t.b) hipEventDestroy(t.b);
  }
  for (void* p : c->allocs) hipFree(p);
  if (c->h_npts) hipHostFree(c->h_npts);
  if (c->h_gather) hipHostFree(c->h_gather);
  if (c->gather_done) hipEventDestroy(c->gather_done);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int phz_get_config(const phz_ctx* c, phz_config* out) {
  PHZ_CHECK(c && out, "null argument");
  *out = c->cfg;
  return 0;
}

int phz_register(phz_ctx* c, const float* target, int64_t Pt, const float* source, int64_t Ps,
                 phz_result* result, float* rot_corr, float* trans_corr, float* registered) {
  return phz_register_channels(c, target, nullptr, Pt, source, nullptr, Ps, result, rot_corr,
                               trans_corr, registered);
}

int phz_register_channels(phz_ctx* c, const float* target, const float* target_ra, int64_t Pt,
                          const float* source, const float* source_ra, int64_t Ps,
                          phz_result* result, float* rot_corr, float* trans_corr,
                          float* registered) {
  PHZ_CHECK(c && result, "null argument");
  hipStream_t s = c->stream;
  TRY(load_cloud(c, 0, target, Pt, hipMemcpyHostToDevice, s, target_ra));
  TRY(load_cloud(c, 1, source, Ps, hipMemcpyHostToDevice, s, source_ra));
  TRY(commit_counts(c, 2, s));
  TRY(exec_pairs(c, 1, false, s, rot_corr ? c->d_corr_rot : nullptr,
                 trans_corr ? c->d_corr_trans : nullptr));
  if (registered && Ps > 0) launch_apply_pose(c->d_pts + c->cap, Ps, c->d_pose, c->d_reg, s);
  // -4 (a failed peak-window check) still fills every output: the grids and
  // the registered cloud are copied before the status is returned
  const int rc = fetch_results(c, 1, result, s);
  if (rc && rc != -4) return rc;
  const std::string pair_msg = g_last_error;
  const int64_t n3 = (int64_t)c->no * c->no * c->no;
  if (rot_corr)
    TRY(copy_sync(c, rot_corr, c->d_corr_rot, sizeof(float) * n3, hipMemcpyDeviceToHost));
  if (trans_corr)
    TRY(copy_sync(c, trans_corr, c->d_corr_trans, sizeof(float) * c->T3p, hipMemcpyDeviceToHost));
  if (registered && Ps > 0)
    TRY(copy_sync(c, registered, c->d_reg, sizeof(float4) * Ps, hipMemcpyDeviceToHost));
  if (rc) set_error(pair_msg);
  return rc;
}

int phz_register_pairs_device(phz_ctx* c, int n, int shared_target, const float* const* d_targets,
                              const int64_t* Pt, const float* const* d_sources, const int64_t* Ps,
                              phz_result* results, float* d_rot_corr, float* d_trans_corr,
                              void* stream) {
  return phz_register_pairs_channels_device(c, n, shared_target, d_targets, nullptr, Pt,
                                            d_sources, nullptr, Ps, results, d_rot_corr,
                                            d_trans_corr, stream);
}

int phz_register_pairs_channels_device(phz_ctx* c, int n, int shared_target,
                                       const float* const* d_targets,
                                       const float* const* d_targets_ra, const int64_t* Pt,
                                       const float* const* d_sources,
                                       const float* const* d_sources_ra, const int64_t* Ps,
                                       phz_result* results, float* d_rot_corr,
                                       float* d_trans_corr, void* stream) {
  PHZ_CHECK(c && d_targets && Pt && d_sources && Ps && results, "null argument");
  PHZ_CHECK(n >= 1 && n <= c->cfg.max_batch, "pair count exceeds max_batch");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const int nt = shared_target ? 1 : n;
  // slots: targets [0, nt), sources [nt, nt + n)
  std::vector<const float*> pts(nt + n), extra(nt + n);
  std::vector<int64_t> counts(nt + n);
  for (int t = 0; t < nt; ++t) {
    pts[t] = d_targets[t];
    extra[t] = d_targets_ra ? d_targets_ra[t] : nullptr;
    counts[t] = Pt[t];
  }
  for (int p = 0; p < n; ++p) {
    pts[nt + p] = d_sources[p];
    extra[nt + p] = d_sources_ra ? d_sources_ra[p] : nullptr;
    counts[nt + p] = Ps[p];
  }
  const int g = gather_clouds(c, nt + n, pts.data(), extra.data(), counts.data(), s);
  if (g < 0) return g;
  if (g > 0) {
    for (int i = 0; i < nt + n; ++i)
      TRY(load_cloud(c, i, pts[i], counts[i], hipMemcpyDeviceToDevice, s, extra[i]));
    TRY(commit_counts(c, nt + n, s));
  }
  TRY(exec_pairs(c, n, shared_target != 0, s, d_rot_corr, d_trans_corr));
  return fetch_results(c, n, results, s);
}

int phz_register_pairs(phz_ctx* c, int n, int shared_target, const float* const* targets,
                       const int64_t* Pt, const float* const* sources, const int64_t* Ps,
                       phz_result* results) {
  PHZ_CHECK(c && targets && Pt && sources && Ps && results, "null argument");
  PHZ_CHECK(n >= 1 && n <= c->cfg.max_batch, "pair count exceeds max_batch");
  hipStream_t s = c->stream;
  const int nt = shared_target ? 1 : n;
  for (int t = 0; t < nt; ++t) TRY(load_cloud(c, t, targets[t], Pt[t], hipMemcpyHostToDevice, s));
  for (int p = 0; p < n; ++p)
    TRY(load_cloud(c, nt + p, sources[p], Ps[p], hipMemcpyHostToDevice, s));
  TRY(commit_counts(c, nt + n, s));
  TRY(exec_pairs(c, n, shared_target != 0, s, nullptr, nullptr));
  return fetch_results(c, n, results, s);
}

int phz_memory_usage(const phz_ctx* c, size_t* bytes) {
  PHZ_CHECK(c && bytes, "null argument");
  *bytes = c->alloc_bytes;
  return 0;
}

int phz_estimate_memory(const phz_config* cfg, size_t* bytes) {
  PHZ_CHECK(cfg && bytes, "null argument");
  TRY(check_config(*cfg));
  *bytes = context_bytes(*cfg);
  return 0;
}

int phz_register_device(phz_ctx* c, const float* d_target, int64_t Pt, const float* d_source,
                        int64_t Ps, phz_result* result, float* d_rot_corr, float* d_trans_corr,
                        float* d_registered, void* stream) {
  PHZ_CHECK(c && result, "null argument");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const float* pts[2] = {d_target, d_source};
  const int64_t counts[2] = {Pt, Ps};
  const int g = gather_clouds(c, 2, pts, nullptr, counts, s);
  if (g < 0) return g;
  if (g > 0) {
    TRY(load_cloud(c, 0, d_target, Pt, hipMemcpyDeviceToDevice, s));
    TRY(load_cloud(c, 1, d_source, Ps, hipMemcpyDeviceToDevice, s));
    TRY(commit_counts(c, 2, s));
  }
  TRY(exec_pairs(c, 1, false, s, d_rot_corr, d_trans_corr));
  if (d_registered && Ps > 0)
    launch_apply_pose(c->d_pts + c->cap, Ps, c->d_pose, (float4*)d_registered, s);
  return fetch_results(c, 1, result, s);
}

// ---------------------------------------------------------------- stages
int phz_sample_s2(phz_ctx* c, const float* pts, int64_t P, float* grids) {
  PHZ_CHECK(c && grids, "null argument");
  hipStream_t s = c->stream;
  TRY(load_cloud(c, 0, pts, P, hipMemcpyHostToDevice, s));
  TRY(commit_counts(c, 1, s));
  launch_sample_s2(c->d_pts, c->d_npts, c->cap, 1, c->n, c->d_grid, c->d_theta, c->max_sq,
                   c->cells, c->d_grids, s);
  PHZ_HIP(hipGetLastError());
  PHZ_HIP(hipMemcpyAsync(grids, c->d_grids, sizeof(float) * 2 * c->n * c->n,
                         hipMemcpyDeviceToHost, s));
  PHZ_HIP(hipStreamSynchronize(s));
  return 0;
}

int phz_s2_transform(phz_ctx* c, int nsig, const float* grids, float* coef) {
  PHZ_CHECK(c && grids && coef, "null argument");
  PHZ_CHECK(nsig >= 1 && nsig <= 16, "nsig must be in [1, 16]");
  hipStream_t s = c->stream;
  const int64_t nq = (int64_t)c->n * c->n, BB = (int64_t)c->B * c->B;
  PHZ_HIP(hipMemcpyAsync(c->d_grids, grids, sizeof(float) * nsig * nq, hipMemcpyHostToDevice, s));
  launch_fst(c->d_grids, nsig, c->B, c->d_leg, c->d_leg_off, c->d_tw_n, c->d_F, c->d_coef,
             c->cfg.precision, s);
  PHZ_HIP(hipGetLastError());
  PHZ_HIP(hipMemcpyAsync(coef, c->d_coef, sizeof(float2) * nsig * BB, hipMemcpyDeviceToHost, s));
  PHZ_HIP(hipStreamSynchronize(s));
  return 0;
}

static int so3_from_fused(phz_ctx* c, const float2* d_f, const float2* d_h, float* corr,
                          float* synthesis, uint32_t* argmax, hipStream_t s) {
  const int B = c->B, n = c->n;
  launch_so3_synth(B, d_f, 0, d_h, 0, 1, c->d_wig, c->d_wig_off, c->d_wig_norm, c->d_S,
                   c->cfg.precision, s, c->d_wig_hl, c->d_wig_hl_off);
  launch_so3_fft_alpha(B, 1, c->d_S, c->plan_n, c->d_tw_n, c->d_U, s, c->so3_reg);
  launch_so3_fft_gamma(B, 1, c->d_U, c->plan_n, c->d_tw_n, corr ? c->d_corr_rot : nullptr,
                       (int64_t)n * n * n, c->d_partials, c->d_keys, s, c->so3_reg);
  PHZ_HIP(hipGetLastError());
  unsigned long long key = 0;
  PHZ_HIP(hipMemcpyAsync(&key, c->d_keys, sizeof(key), hipMemcpyDeviceToHost, s));
  PHZ_HIP(hipStreamSynchronize(s));
  if (argmax) *argmax = key_index(key);
  if (corr)
    TRY(copy_sync(c, corr, c->d_corr_rot, sizeof(float) * n * n * n, hipMemcpyDeviceToHost));
  if (synthesis)
    TRY(copy_sync(c, synthesis, c->d_S, sizeof(float2) * (size_t)B * (2 * B - 1) * n,
                      hipMemcpyDeviceToHost));
  return 0;
}

int phz_so3_correlate_coef(phz_ctx* c, const float* f_coef, const float* h_coef, float* corr,
                           float* synthesis, uint32_t* argmax) {
  PHZ_CHECK(c && f_coef && h_coef, "null argument");
  PHZ_CHECK(c->Bo == c->B, "the SO(3) stage entry points run without spherical zero padding");
  hipStream_t s = c->stream;
  const int64_t BB = (int64_t)c->B * c->B;
  PHZ_HIP(hipMemcpyAsync(c->d_fused, f_coef, sizeof(float2) * BB, hipMemcpyHostToDevice, s));
  PHZ_HIP(hipMemcpyAsync(c->d_fused + BB, h_coef, sizeof(float2) * BB, hipMemcpyHostToDevice, s));
  return so3_from_fused(c, c->d_fused, c->d_fused + BB, corr, synthesis, argmax, s);
}

int phz_spherical_correlate(phz_ctx* c, int C, const float* f_grids, const float* h_grids,
                            float* corr, uint32_t* argmax) {
  PHZ_CHECK(c && f_grids && h_grids, "null argument");
  PHZ_CHECK(C >= 2 && C <= 8, "channel count must be in [2, 8]");
  PHZ_CHECK(c->Bo == c->B, "the SO(3) stage entry points run without spherical zero padding");
  hipStream_t s = c->stream;
  const int64_t nq = (int64_t)c->n * c->n, BB = (int64_t)c->B * c->B;
  PHZ_HIP(hipMemcpyAsync(c->d_grids, f_grids, sizeof(float) * C * nq, hipMemcpyHostToDevice, s));
  PHZ_HIP(hipMemcpyAsync(c->d_grids + C * nq, h_grids, sizeof(float) * C * nq,
                         hipMemcpyHostToDevice, s));
  launch_fst(c->d_grids, 2 * C, c->B, c->d_leg, c->d_leg_off, c->d_tw_n, c->d_F, c->d_coef,
             c->cfg.precision, s);
  for (int w = 0; w < 2; ++w) {
    ChannelSet cs{};
    cs.C = C;
    for (int k = 0; k < C; ++k) cs.ch[k] = c->d_coef + (int64_t)(w * C + k) * BB;
    launch_fuse(cs, (uint32_t)BB, c->s2.runs, c->d_fused + w * BB, s);
  }
  return so3_from_fused(c, c->d_fused, c->d_fused + BB, corr, nullptr, argmax, s);
}

int phz_voxelize(phz_ctx* c, const float* pts, int64_t P, float* out, float* counts) {
  PHZ_CHECK(c && out, "null argument");
  hipStream_t s = c->stream;
  TRY(load_cloud(c, 0, pts, P, hipMemcpyHostToDevice, s));
  TRY(commit_counts(c, 1, s));
  launch_voxel_scatter(c->d_pts, c->d_npts, c->cap, 1, nullptr, 0, c->N, c->d_edges, c->d_sumI,
                       c->d_sumR, c->d_cnt, occ_from(c, 0), s);
  PHZ_HIP(hipGetLastError());
  std::vector<float> sI(c->T3), sR(c->T3);
  std::vector<uint32_t> k(c->T3);
  PHZ_HIP(hipMemcpyAsync(sI.data(), c->d_sumI, sizeof(float) * c->T3, hipMemcpyDeviceToHost, s));
  PHZ_HIP(hipMemcpyAsync(sR.data(), c->d_sumR, sizeof(float) * c->T3, hipMemcpyDeviceToHost, s));
  PHZ_HIP(hipMemcpyAsync(k.data(), c->d_cnt, sizeof(uint32_t) * c->T3, hipMemcpyDeviceToHost, s));
  launch_voxel_clear(c->d_pts, c->d_npts, c->cap, 1, nullptr, 0, c->N, c->d_edges, c->d_sumI,
                     c->d_sumR, c->d_cnt, occ_from(c, 0), s);
  PHZ_HIP(hipStreamSynchronize(s));
  for (int64_t v = 0; v < c->T3; ++v) {
    out[v] = k[v] ? sI[v] / (float)k[v] : 0.f;
    out[c->T3 + v] = k[v] ? sR[v] / (float)k[v] : 0.f;
    if (counts) counts[v] = (float)k[v];
  }
  return 0;
}

int phz_spatial_correlate(phz_ctx* c, int C, const float* f, const float* g, float* corr,
                          uint32_t* argmax) {
  PHZ_CHECK(c && f && g, "null argument");
  PHZ_CHECK(C >= 2 && C <= 8, "channel count must be in [2, 8]");
  hipStream_t s = c->stream;
  const int64_t T = c->T3, H = c->H;
  PHZ_HIP(hipMemcpyAsync(c->d_sumI, f, sizeof(float) * C * T, hipMemcpyHostToDevice, s));
  PHZ_HIP(hipMemcpyAsync(c->d_sumI + C * T, g, sizeof(float) * C * T, hipMemcpyHostToDevice, s));
  launch_spatial_forward_dense(c->N, 2 * C, c->d_sumI, c->plan_N, c->d_tw_N, c->d_spec, c->tiles,
                               s);
  ChannelSet F{}, G{};
  F.C = G.C = C;
  for (int k = 0; k < C; ++k) {
    F.ch[k] = c->d_spec + k * H;
    G.ch[k] = c->d_spec + (C + k) * H;
  }
  if (c->sp) {  // the context's spatial zero padding: grid (N + 2 pad)^3
    launch_spatial_cross_product(c->N, 1, F, 0, G, 0, c->vox.runs, true, c->d_work, s);
    launch_spatial_pad_gather(c->N, c->sp, true, 1, c->d_work, c->d_work_pad, s);
    const int tile = fft_cols_default_tile(c->Np);
    launch_spatial_fft_pass(c->Np, 1, c->d_work_pad, 3, tile, c->plan_Np, c->d_tw_Np, s);
    launch_spatial_fft_pass(c->Np, 1, c->d_work_pad, 2, tile, c->plan_Np, c->d_tw_Np, s);
    launch_spatial_c2r(c->Np, 1, c->d_work_pad, c->plan_Np, c->d_tw_Np,
                       corr ? c->d_corr_trans : nullptr, c->d_partials, c->d_keys, s, c->Tc);
  } else {
    launch_spatial_inverse(c->N, 1, F, 0, G, 0, c->vox.runs, true, c->plan_N, c->d_tw_N,
                           c->d_work, corr ? c->d_corr_trans : nullptr, c->d_partials, c->d_keys,
                           c->tiles, s);
  }
  // the dense channels were staged in the voxel accumulators: zero them again
  launch_fill_u32((uint32_t*)c->d_sumI, (size_t)2 * C * T, 0u, s);
  PHZ_HIP(hipGetLastError());
  unsigned long long key = 0;
  PHZ_HIP(hipMemcpyAsync(&key, c->d_keys, sizeof(key), hipMemcpyDeviceToHost, s));
  PHZ_HIP(hipStreamSynchronize(s));
  if (argmax) *argmax = key_index(key);
  if (corr) TRY(copy_sync(c, corr, c->d_corr_trans, sizeof(float) * c->T3p, hipMemcpyDeviceToHost));
  return 0;
}

// SpatialCorrelation(N, pad)::correlateSignals on one channel pair, with
// buffers allocated for the call (spatial-correlation.cc:11-61,163-206);
// cube_lo <= cube_hi: SpatialCorrelationLowPass (spatial-correlation-low-
// pass.cc:39-94), the product of the FFTShifted spectra at the shifted
// indices of the [cube_lo, cube_hi]^3 cube only, IFFTShift, c2r
static int spatial_single(phz_ctx* c, int N, int pad, const float* f, const float* g,
                          float* corr, uint32_t* argmax, int cube_lo = 1, int cube_hi = 0) {
  PHZ_CHECK(c && f && g, "null argument");
  const bool low_pass = cube_lo <= cube_hi;
  const bool gather = pad > 0 || low_pass;
  PHZ_CHECK(N >= 1 && N <= 256, "n_voxels must be in [1, 256]");
  PHZ_CHECK(pad >= 0 && N + 2 * pad <= 256, "zero padding must be >= 0 with n_voxels + 2 * padding <= 256");
  PHZ_CHECK(N > 1 || pad == 0, "zero padding needs n_voxels >= 2");
  FftPlan plan = make_plan(N);
  const int Np = N + 2 * pad;
  const FftPlan plan_p = make_plan(Np);
  hipStream_t s = c->stream;
  const int64_t T = (int64_t)N * N * N, H = (int64_t)N * N * (N / 2 + 1);
  const int64_t Tp = (int64_t)Np * Np * Np, Hp = (int64_t)Np * Np * (Np / 2 + 1);
  const uint32_t limit = pad ? padded_size(N, pad) : (uint32_t)T;
  std::vector<float> tw = twiddles(N), twp = twiddles(Np);
  float *d_in = nullptr, *d_corr = nullptr;
  float2 *d_spec = nullptr, *d_work = nullptr, *d_work_pad = nullptr, *d_tw = nullptr,
         *d_twp = nullptr;
  unsigned long long* d_key = nullptr;
  int rc = 0;
  auto fail = [&](const char* what, hipError_t e) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    rc = -2;
  };
  hipError_t e;
  if ((e = hipMalloc(&d_in, sizeof(float) * 2 * T)) != hipSuccess) fail("hipMalloc", e);
  if (!rc && (e = hipMalloc(&d_spec, sizeof(float2) * 2 * H)) != hipSuccess) fail("hipMalloc", e);
  if (!rc && (e = hipMalloc(&d_work, sizeof(float2) * H)) != hipSuccess) fail("hipMalloc", e);
  if (!rc && (e = hipMalloc(&d_tw, sizeof(float2) * N)) != hipSuccess) fail("hipMalloc", e);
  if (!rc && (e = hipMalloc(&d_corr, sizeof(float) * Tp)) != hipSuccess) fail("hipMalloc", e);
  if (!rc && gather && (e = hipMalloc(&d_work_pad, sizeof(float2) * Hp)) != hipSuccess)
    fail("hipMalloc", e);
  if (!rc && gather && (e = hipMalloc(&d_twp, sizeof(float2) * Np)) != hipSuccess)
    fail("hipMalloc", e);
  const int nblk = spatial_c2r_blocks(Np);
  if (!rc && (e = hipMalloc(&d_key, sizeof(unsigned long long) * (1 + nblk))) != hipSuccess)
    fail("hipMalloc", e);
  if (!rc) {
    hipMemcpyAsync(d_in, f, sizeof(float) * T, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(d_in + T, g, sizeof(float) * T, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(d_tw, tw.data(), sizeof(float2) * N, hipMemcpyHostToDevice, s);
    if (gather) hipMemcpyAsync(d_twp, twp.data(), sizeof(float2) * Np, hipMemcpyHostToDevice, s);
    if (N == 1) {
      plan.L = 1;
      plan.nst = 0;
    }
    launch_spatial_forward_dense(N, 2, d_in, plan, d_tw, d_spec, SpatialTiles{}, s);
    ChannelSet F{}, G{};
    F.C = G.C = 1;
    F.ch[0] = d_spec;
    G.ch[0] = d_spec + H;
    FusionRuns none{};
    if (gather) {  // spatial-correlation.cc:142-161,221-230; shifted in the low-pass class
      launch_spatial_cross_product(N, 1, F, 0, G, 0, none, false, d_work, s);
      launch_spatial_pad_gather(N, pad, low_pass, 1, d_work, d_work_pad, s, cube_lo, cube_hi);
      const int tile = fft_cols_default_tile(Np);
      launch_spatial_fft_pass(Np, 1, d_work_pad, 3, tile, plan_p, d_twp, s);
      launch_spatial_fft_pass(Np, 1, d_work_pad, 2, tile, plan_p, d_twp, s);
      launch_spatial_c2r(Np, 1, d_work_pad, plan_p, d_twp, d_corr, d_key + 1, d_key, s, limit);
    } else {
      launch_spatial_inverse(N, 1, F, 0, G, 0, none, false, plan, d_tw, d_work, d_corr, d_key + 1,
                             d_key, SpatialTiles{}, s);
    }
    if ((e = hipGetLastError()) != hipSuccess) fail("launch", e);
  }
  unsigned long long key = 0;
  if (!rc && (e = hipMemcpyAsync(&key, d_key, sizeof(key), hipMemcpyDeviceToHost, s)) != hipSuccess)
    fail("copy", e);
  if (!rc && (e = hipStreamSynchronize(s)) != hipSuccess) fail("sync", e);
  if (!rc && corr &&
      ((e = hipMemcpyAsync(corr, d_corr, sizeof(float) * Tp, hipMemcpyDeviceToHost, s)) !=
           hipSuccess ||
       (e = hipStreamSynchronize(s)) != hipSuccess))
    fail("copy", e);
  if (!rc && argmax) *argmax = key_index(key);
  hipFree(d_in);
  hipFree(d_spec);
  hipFree(d_work);
  hipFree(d_work_pad);
  hipFree(d_tw);
  hipFree(d_twp);
  hipFree(d_corr);
  hipFree(d_key);
  return rc;
}

int phz_zscore_peaks(phz_ctx* c, const float* corr, int64_t n, const phz_zscore_params* zp,
                     uint32_t* peaks, int64_t max_peaks, int64_t* count) {
  PHZ_CHECK(c && corr && zp && count && (peaks || max_peaks == 0), "null argument");
  PHZ_CHECK(n >= 0 && n <= 0xffffffffll, "vector length must fit uint32 (the reference's n_input)");
  PHZ_CHECK(max_peaks >= 0, "max_peaks must be >= 0");
  PHZ_CHECK(zp->lag_percentile > 0 && zp->lag_percentile < 1,
            "z_score_lag_percentile must be in (0, 1)");  // z-score-peak-extraction.cc:35-36
  PHZ_CHECK(zp->threshold > 0, "z_score_threshold must be > 0");
  PHZ_CHECK(zp->influence >= 0 && zp->influence <= 1, "z_score_influence must be in [0, 1]");
  // :44-45: lag = min(lag_max, (uint32)(n * lag_percentile))
  const uint32_t by_pct = (uint32_t)((double)n * zp->lag_percentile);
  const uint32_t lag = zp->lag_max < by_pct ? zp->lag_max : by_pct;
  PHZ_CHECK(lag <= 64, "the device detector keeps windows of at most 64 values (lag)");
  hipStream_t s = c->stream;
  float* d_corr = nullptr;
  uint32_t* d_peaks = nullptr;
  unsigned long long* d_count = nullptr;
  int rc = 0;
  hipError_t e;
  auto fail = [&](const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    rc = -2;
  };
  if ((e = hipMalloc(&d_corr, sizeof(float) * (n ? n : 1))) != hipSuccess) fail("hipMalloc");
  if (!rc && (e = hipMalloc(&d_peaks, sizeof(uint32_t) * (max_peaks ? max_peaks : 1))) != hipSuccess)
    fail("hipMalloc");
  if (!rc && (e = hipMalloc(&d_count, sizeof(unsigned long long))) != hipSuccess) fail("hipMalloc");
  unsigned long long k = 0;
  if (!rc) {
    if (n) hipMemcpyAsync(d_corr, corr, sizeof(float) * n, hipMemcpyHostToDevice, s);
    launch_zscore_peaks(d_corr, n, 1, (uint32_t)n, (int)lag, zp->threshold, zp->influence,
                        zp->filter_threshold, d_peaks, max_peaks, d_count, s);
    if ((e = hipGetLastError()) != hipSuccess) fail("launch");
  }
  if (!rc && (e = hipMemcpyAsync(&k, d_count, sizeof(k), hipMemcpyDeviceToHost, s)) != hipSuccess)
    fail("copy");
  if (!rc && (e = hipStreamSynchronize(s)) != hipSuccess) fail("sync");
  const int64_t stored = (int64_t)k < max_peaks ? (int64_t)k : max_peaks;
  if (!rc && stored > 0 &&
      ((e = hipMemcpyAsync(peaks, d_peaks, sizeof(uint32_t) * stored, hipMemcpyDeviceToHost, s)) !=
           hipSuccess ||
       (e = hipStreamSynchronize(s)) != hipSuccess))
    fail("copy");
  if (!rc) *count = (int64_t)k;
  hipFree(d_corr);
  hipFree(d_peaks);
  hipFree(d_count);
  return rc;
}

int phz_voxel_grid_filter(phz_ctx* c, float* pts, int64_t P, float leaf, int64_t* P_out) {
  PHZ_CHECK(c && P_out && (pts || P == 0), "null argument");
  PHZ_CHECK(P >= 0 && P < (1ll << 31), "point count must be in [0, 2^31)");
  PHZ_CHECK(leaf > 0.f && std::isfinite(leaf), "leaf size must be positive");
  *P_out = P;
  if (P == 0) return 0;
  hipStream_t s = c->stream;
  CallBuffers buf;
  float4 *d_pts, *d_out;
  float *d_part, *d_box;
  uint32_t *kin, *kout, *vin, *vout, *uniq, *cnt, *off;
  int* d_nruns;
  PHZ_HIP(buf.get(&d_pts, P));
  PHZ_HIP(buf.get(&d_out, P));
  PHZ_HIP(buf.get(&d_part, 256 * 6));
  PHZ_HIP(buf.get(&d_box, 6));
  for (uint32_t** q : {&kin, &kout, &vin, &vout, &uniq, &cnt, &off}) PHZ_HIP(buf.get(q, P));
  PHZ_HIP(buf.get(&d_nruns, 1));
  PHZ_HIP(hipMemcpyAsync(d_pts, pts, sizeof(float4) * P, hipMemcpyHostToDevice, s));
  launch_bbox(d_pts, P, d_part, d_box, s);
  float box[6];
  PHZ_HIP(hipMemcpyAsync(box, d_box, sizeof(box), hipMemcpyDeviceToHost, s));
  PHZ_HIP(hipStreamSynchronize(s));
  if (!(box[0] <= box[3])) {  // no finite point: nothing survives the filter
    *P_out = 0;
    return 0;
  }
  // voxel_grid.hpp (PCL 1.8): inverse leaf in float, the overflow check
  // (the input is returned unchanged when the grid exceeds INT32_MAX voxels)
  const float inv = 1.0f / leaf;
  int64_t d[3];
  for (int k = 0; k < 3; ++k) d[k] = static_cast<int64_t>((box[3 + k] - box[k]) * inv) + 1;
  if (d[0] * d[1] * d[2] > static_cast<int64_t>(std::numeric_limits<int32_t>::max())) return 0;
  int min_b[3], max_b[3];
  for (int k = 0; k < 3; ++k) {
    min_b[k] = static_cast<int>(std::floor(box[k] * inv));
    max_b[k] = static_cast<int>(std::floor(box[3 + k] * inv));
  }
  const int div0 = max_b[0] - min_b[0] + 1, div1 = max_b[1] - min_b[1] + 1;
  launch_voxel_grid_keys(d_pts, P, make_float3(inv, inv, inv), make_int3(min_b[0], min_b[1], min_b[2]),
                         make_int3(1, div0, div0 * div1), kin, vin, s);
  size_t tb = 0;
  PHZ_HIP(voxel_grid_sort_runs(nullptr, &tb, kin, kout, vin, vout, (int)P, uniq, cnt, off, d_nruns, s));
  void* temp = nullptr;
  PHZ_HIP(buf.get((unsigned char**)&temp, tb));
  PHZ_HIP(voxel_grid_sort_runs(temp, &tb, kin, kout, vin, vout, (int)P, uniq, cnt, off, d_nruns, s));
  launch_voxel_grid_centroids(d_pts, vout, cnt, off, d_nruns, (int)P, d_out, s);
  PHZ_HIP(hipGetLastError());
  int nruns = 0;
  PHZ_HIP(hipMemcpyAsync(&nruns, d_nruns, sizeof(int), hipMemcpyDeviceToHost, s));
  PHZ_HIP(hipStreamSynchronize(s));
  uint32_t last = 0;
  if (nruns > 0) TRY(copy_sync(c, &last, uniq + nruns - 1, sizeof(last), hipMemcpyDeviceToHost));
  if (nruns > 0 && last == 0xffffffffu) --nruns;  // the non-finite points' run
  if (nruns > 0)
    TRY(copy_sync(c, pts, d_out, sizeof(float4) * nruns, hipMemcpyDeviceToHost));
  *P_out = nruns;
  return 0;
}

int phz_pass_through_filter(phz_ctx* c, float* pts, int64_t P, int field, float lower, float upper,
                            int negative, int64_t* P_out) {
  PHZ_CHECK(c && P_out && (pts || P == 0), "null argument");
  PHZ_CHECK(P >= 0 && P < (1ll << 31), "point count must be in [0, 2^31)");
  PHZ_CHECK(field >= 0 && field <= 3, "field must be 0 (x), 1 (y), 2 (z) or 3 (intensity)");
  *P_out = 0;
  if (P == 0) return 0;
  hipStream_t s = c->stream;
  CallBuffers buf;
  float4 *d_pts, *d_out;
  unsigned char* keep;
  int* d_n;
  PHZ_HIP(buf.get(&d_pts, P));
  PHZ_HIP(buf.get(&d_out, P));
  PHZ_HIP(buf.get(&keep, P));
  PHZ_HIP(buf.get(&d_n, 1));
  PHZ_HIP(hipMemcpyAsync(d_pts, pts, sizeof(float4) * P, hipMemcpyHostToDevice, s));
  size_t tb = 0;
  PHZ_HIP(pass_through_select(nullptr, &tb, d_pts, P, field, lower, upper, negative, keep, d_out,
                              d_n, s));
  void* temp = nullptr;
  PHZ_HIP(buf.get((unsigned char**)&temp, tb));
  PHZ_HIP(pass_through_select(temp, &tb, d_pts, P, field, lower, upper, negative, keep, d_out, d_n,
                              s));
  int n = 0;
  PHZ_HIP(hipMemcpyAsync(&n, d_n, sizeof(int), hipMemcpyDeviceToHost, s));
  PHZ_HIP(hipStreamSynchronize(s));
  if (n > 0) TRY(copy_sync(c, pts, d_out, sizeof(float4) * n, hipMemcpyDeviceToHost));
  *P_out = n;
  return 0;
}

int phz_spatial_correlate_single(phz_ctx* c, int N, const float* f, const float* g, float* corr,
                                 uint32_t* argmax) {
  return spatial_single(c, N, 0, f, g, corr, argmax);
}

int phz_spatial_correlate_single_padded(phz_ctx* c, int N, int zero_padding, const float* f,
                                        const float* g, float* corr, uint32_t* argmax) {
  return spatial_single(c, N, zero_padding, f, g, corr, argmax);
}

int phz_spatial_correlate_low_pass(phz_ctx* c, int N, int zero_padding, int lower, int upper,
                                   const float* f, const float* g, float* corr, uint32_t* argmax) {
  PHZ_CHECK(N >= 2, "n_voxels must be >= 2");
  // spatial-correlation-low-pass.cc:11-19: the upper bound is capped at
  // n_voxels; an index of the cube with a coordinate N reaches past N^3
  // (Sub2Ind, :86-87) and fails the CHECK_LT of complexMulSeqUsingIndices
  // (spatial-correlation.cc:153), an empty cube its CHECK(!indices.empty())
  const int hi = upper < N ? upper : N;
  PHZ_CHECK(lower >= 0, "the low-pass lower bound must be >= 0");
  PHZ_CHECK(lower <= hi, "empty low-pass cube (spatial-correlation.cc:145 CHECK)");
  PHZ_CHECK(hi < N, "low-pass cube reaches index N^3 (spatial-correlation.cc:153 CHECK_LT)");
  return spatial_single(c, N, zero_padding, f, g, corr, argmax, lower, hi);
}

int phz_spherical_correlate_low_pass(phz_ctx* c, const float* f_grid, const float* h_grid,
                                     int lower, int upper, float* corr, uint32_t* argmax) {
  PHZ_CHECK(c && f_grid && h_grid, "null argument");
  PHZ_CHECK(c->Bo == c->B, "the SO(3) stage entry points run without spherical zero padding");
  // spherical-correlation-low-pass.cc:26-27
  PHZ_CHECK(lower >= 0, "the spherical low-pass lower bound must be >= 0");
  PHZ_CHECK(upper > 0, "the spherical low-pass upper bound must be > 0");
  hipStream_t s = c->stream;
  const int64_t nq = (int64_t)c->n * c->n, BB = (int64_t)c->B * c->B;
  const uint32_t lo = (uint32_t)std::min<int64_t>(BB, lower);
  const uint32_t hi = (uint32_t)std::min<int64_t>(BB, upper);
  PHZ_HIP(hipMemcpyAsync(c->d_grids, f_grid, sizeof(float) * nq, hipMemcpyHostToDevice, s));
  PHZ_HIP(hipMemcpyAsync(c->d_grids + nq, h_grid, sizeof(float) * nq, hipMemcpyHostToDevice, s));
  launch_fst(c->d_grids, 2, c->B, c->d_leg, c->d_leg_off, c->d_tw_n, c->d_F, c->d_coef,
             c->cfg.precision, s);
  launch_coef_band(c->d_coef, BB, 2, (uint32_t)BB, lo, hi, s);
  return so3_from_fused(c, c->d_coef, c->d_coef + BB, corr, nullptr, argmax, s);
}

int phz_spatial_correlate_single_noalloc(phz_ctx* c, int n_voxels, const float* d_f,
                                         const float* d_g, float* d_corr, uint32_t* argmax,
                                         void* stream) {
  PHZ_CHECK(c, "null argument");
  return spatial_single_noalloc(c->cfg.device, n_voxels, d_f, d_g, d_corr, argmax,
                                stream ? (hipStream_t)stream : c->stream);
}

// ---------------------------------------------------------------- statistics
int phz_set_profiling(phz_ctx* c, int enable) {
  PHZ_CHECK(c, "null argument");
  PHZ_HIP(hipStreamSynchronize(c->stream));
  if (c->timings.empty()) {
    c->timings.resize(kNumStages);
    for (int i = 0; i < kNumStages; ++i) {
      c->timings[i].name = kStageNames[i];
      PHZ_HIP(hipEventCreate(&c->timings[i].a));
      PHZ_HIP(hipEventCreate(&c->timings[i].b));
    }
  }
  for (auto& t : c->timings) {
    t.total_ms = 0.0;
    t.calls = 0;
    t.pending = false;
  }
  c->profiling = enable != 0;
  return 0;
}

int phz_get_stage_times(phz_ctx* c, phz_stage_time* out, int max, int* count) {
  PHZ_CHECK(c && count, "null argument");
  int k = 0;
  for (const auto& t : c->timings) {
    if (out && k < max) {
      std::memset(&out[k], 0, sizeof(phz_stage_time));
      std::strncpy(out[k].name, t.name, sizeof(out[k].name) - 1);
      out[k].total_ms = t.total_ms;
      out[k].calls = t.calls;
    }
    ++k;
  }
  *count = k;
  return 0;
}

// ---------------------------------------------------------------- device memory
int phz_device_malloc(phz_ctx* c, size_t bytes, void** ptr) {
  PHZ_CHECK(c && ptr, "null argument");
  PHZ_HIP(hipSetDevice(c->cfg.device));
  PHZ_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return 0;
}

int phz_device_free(phz_ctx* c, void* ptr) {
  PHZ_CHECK(c, "null argument");
  if (ptr) PHZ_HIP(hipFree(ptr));
  return 0;
}

int phz_memcpy(phz_ctx* c, void* dst, const void* src, size_t bytes, int kind) {
  PHZ_CHECK(c && dst && src, "null argument");
  PHZ_CHECK(kind >= 0 && kind <= 3, "kind must be 0 h2h, 1 h2d, 2 d2h or 3 d2d");
  const hipMemcpyKind k[4] = {hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                              hipMemcpyDeviceToDevice};
  PHZ_HIP(hipStreamSynchronize(c->stream));
  TRY(copy_sync(c, dst, src, bytes, k[kind]));
  return 0;
}

static void apply_tuning(phz_ctx* c, const phz_tuning& t) {
  c->tuning = t;
  c->tiles.fwd_x = fft_cols_tile_ok(c->N, t.fwd_x) ? t.fwd_x : 0;
  c->tiles.fwd_z = fft_cols_tile_ok(c->N, t.fwd_z) ? t.fwd_z : 0;
  c->tiles.inv_x = fft_cols_tile_ok(c->N, t.inv_x) ? t.inv_x : 0;
  c->tuning.fwd_x = c->tiles.fwd_x ? c->tiles.fwd_x : fft_cols_default_tile(c->N);
  c->tuning.fwd_z = c->tiles.fwd_z ? c->tiles.fwd_z : fft_cols_default_tile(c->N);
  c->tuning.inv_x = c->tiles.inv_x ? c->tiles.inv_x : fft_cols_default_tile(c->N);
  c->plane_fwd = t.plane_fwd > 0 && spatial_plane_ok(c->N);
  c->tuning.plane_fwd = c->plane_fwd ? 1 : -1;
  c->tuning.nn_lanes = (t.nn_lanes == 1 || t.nn_lanes == 2 || t.nn_lanes == 4 ||
                        t.nn_lanes == 8 || t.nn_lanes == 16)
                           ? t.nn_lanes
                           : 0;
  c->tiles.fused_reg = t.fused_reg > 0 && spatial_reg_ok(c->N);
  c->tuning.fused_reg = c->tiles.fused_reg ? 1 : -1;
  c->so3_reg = t.so3_reg > 0 && so3_reg_ok(c->Bo);
  c->tuning.so3_reg = c->so3_reg ? 1 : -1;
  c->occ_fwd = t.occ_fwd >= 0;
  c->tuning.occ_fwd = c->occ_fwd ? 1 : -1;
  c->inv_plane = t.inv_plane >= 0;
  c->tuning.inv_plane = c->inv_plane ? 1 : -1;
  c->nn_xcd = t.nn_xcd == 0 ? kNnXcdDefault : t.nn_xcd > 0;
  c->tuning.nn_xcd = c->nn_xcd ? 1 : -1;
  c->nn_seed = t.nn_seed == 0 ? kNnSeedDefault : (t.nn_seed > 0 ? (t.nn_seed & 3) : 0);
  c->tuning.nn_seed = c->nn_seed ? c->nn_seed : -1;
  c->xz_slab = t.xz_slab > 0;
  c->tuning.xz_slab = c->xz_slab ? 1 : -1;
}

int phz_set_tuning(phz_ctx* c, const phz_tuning* t) {
  PHZ_CHECK(c && t, "null argument");
  apply_tuning(c, *t);
  clear_graphs(c);  // captured kernels follow the old choices
  return 0;
}

int phz_set_graphs(phz_ctx* c, int enable) {
  PHZ_CHECK(c, "null argument");
  if (!enable) clear_graphs(c);
  c->use_graphs = enable != 0;
  return 0;
}

int phz_get_tuning(const phz_ctx* c, phz_tuning* t) {
  PHZ_CHECK(c && t, "null argument");
  *t = c->tuning;
  return 0;
}

int phz_autotune(phz_ctx* c, int repeats, phz_tuning* out) {
  PHZ_CHECK(c, "null argument");
  PHZ_CHECK(repeats >= 1 && repeats <= 1000, "repeats out of range");
  hipStream_t s = c->stream;
  const int N = c->N, P = c->cfg.max_batch, nspec = 2 * c->max_clouds;
  // timing runs on the context's own spectra buffers at full batch size
  launch_fill_u32((uint32_t*)c->d_spec, (size_t)2 * nspec * c->H, 0u, s);
  launch_fill_u32((uint32_t*)c->d_work, (size_t)2 * P * c->H, 0u, s);
  hipEvent_t a, b;
  PHZ_HIP(hipEventCreate(&a));
  PHZ_HIP(hipEventCreate(&b));
  const int cands[5] = {kRegisterEngine, 8, 16, 32, 64};
  phz_tuning best{};
  float best_ms[3] = {1e30f, 1e30f, 1e30f};
  int rc = 0;
  for (int site = 0; site < 3 && !rc; ++site)
    for (int it : cands) {
      if (!fft_cols_tile_ok(N, it)) continue;
      float ms = 0.f;
      for (int pass = 0; pass < 2; ++pass) {  // pass 0 warms up
        if (hipEventRecord(a, s) != hipSuccess) rc = -2;
        for (int r = 0; r < repeats; ++r)
          launch_spatial_fft_pass(N, site == 2 ? P : nspec, site == 2 ? c->d_work : c->d_spec,
                                  site, it, c->plan_N, c->d_tw_N, s);
        if (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
            hipEventElapsedTime(&ms, a, b) != hipSuccess)
          rc = -2;
      }
      ms /= (float)repeats;
      if (ms < best_ms[site]) {
        best_ms[site] = ms;
        (site == 0 ? best.fwd_x : site == 1 ? best.fwd_z : best.inv_x) = it;
      }
    }
  // the plane pass against the r2c + x passes it replaces (best x tile)
  float plane_ms = 0.f, classic_ms = 0.f;
  if (!rc && spatial_plane_ok(N)) {
    const int ncl = c->max_clouds;
    for (int variant = 0; variant < 2 && !rc; ++variant) {
      float ms = 0.f;
      for (int pass = 0; pass < 2; ++pass) {
        if (hipEventRecord(a, s) != hipSuccess) rc = -2;
        for (int r = 0; r < repeats; ++r) {
          if (variant == 0) {
            launch_spatial_plane_fwd(N, ncl, c->d_sumI, c->d_sumR, c->d_cnt, c->plan_N,
                                     c->d_tw_N, c->d_spec, s);
          } else {
            launch_spatial_r2c(N, ncl, c->d_sumI, c->d_sumR, c->d_cnt, c->plan_N, c->d_tw_N,
                               c->d_spec, s);
            launch_spatial_fft_pass(N, nspec, c->d_spec, 0, best.fwd_x, c->plan_N, c->d_tw_N, s);
          }
        }
        if (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
            hipEventElapsedTime(&ms, a, b) != hipSuccess)
          rc = -2;
      }
      (variant == 0 ? plane_ms : classic_ms) = ms / (float)repeats;
    }
    best.plane_fwd = plane_ms <= classic_ms ? 1 : -1;
  } else {
    best.plane_fwd = -1;
  }
  // the fused passes (r2c, cross_z, c2r): register engine against LDS kernels
  float fused_ms[2] = {0.f, 0.f};
  if (!rc && spatial_reg_ok(N)) {
    const int ncl = c->max_clouds;
    const float2* F = c->d_spec;
    const float2* G = c->d_spec + (int64_t)2 * c->H;
    for (int variant = 0; variant < 2 && !rc; ++variant) {
      float ms = 0.f;
      for (int pass = 0; pass < 2; ++pass) {
        if (hipEventRecord(a, s) != hipSuccess) rc = -2;
        for (int r = 0; r < repeats; ++r) {
          if (variant == 1) {
            launch_spatial_r2c_reg(N, ncl, c->d_sumI, c->d_sumR, c->d_cnt, nullptr, false,
                                   c->d_tw_N, c->d_spec, s);
            launch_spatial_cross_z_reg(N, P, channels(F, F + c->H, 4), 0, channels(G, G + c->H, 4),
                                       0, c->vox.runs, true, c->d_tw_N, c->d_work, s);
            launch_spatial_c2r_reg(N, P, c->d_work, c->d_tw_N, nullptr, c->d_partials,
                                   c->d_keys, s);
          } else {
            launch_spatial_r2c(N, ncl, c->d_sumI, c->d_sumR, c->d_cnt, c->plan_N, c->d_tw_N,
                               c->d_spec, s);
            launch_spatial_cross_z(N, P, channels(F, F + c->H, 4), 0, channels(G, G + c->H, 4), 0,
                                   c->vox.runs, true, c->plan_N, c->d_tw_N, c->d_work, s);
            launch_spatial_c2r(N, P, c->d_work, c->plan_N, c->d_tw_N, nullptr, c->d_partials,
                               c->d_keys, s);
          }
        }
        if (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
            hipEventElapsedTime(&ms, a, b) != hipSuccess)
          rc = -2;
      }
      fused_ms[variant] = ms / (float)repeats;
    }
    best.fused_reg = fused_ms[1] <= fused_ms[0] ? 1 : -1;
  } else {
    best.fused_reg = -1;
  }
  // SO(3) alpha + gamma passes: register engine against LDS kernels
  float so3_ms[2] = {0.f, 0.f};
  if (!rc && so3_reg_ok(c->Bo)) {
    const int B = c->Bo, n = c->no;
    launch_fill_u32((uint32_t*)c->d_S, (size_t)2 * P * B * (2 * B - 1) * n, 0u, s);
    for (int variant = 0; variant < 2 && !rc; ++variant) {
      float ms = 0.f;
      for (int pass = 0; pass < 2; ++pass) {
        if (hipEventRecord(a, s) != hipSuccess) rc = -2;
        for (int r = 0; r < repeats; ++r) {
          launch_so3_fft_alpha(B, P, c->d_S, c->plan_no, c->d_tw_no, c->d_U, s, variant == 1);
          launch_so3_fft_gamma(B, P, c->d_U, c->plan_no, c->d_tw_no, nullptr, (int64_t)n * n * n,
                               c->d_partials, c->d_keys, s, variant == 1);
        }
        if (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
            hipEventElapsedTime(&ms, a, b) != hipSuccess)
          rc = -2;
      }
      so3_ms[variant] = ms / (float)repeats;
    }
    best.so3_reg = so3_ms[1] <= so3_ms[0] ? 1 : -1;
  } else {
    best.so3_reg = -1;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) {
    set_error("autotune timing failed");
    return rc;
  }
  best.nn_lanes = c->tuning.nn_lanes;  // data-dependent: set with phz_set_tuning
  best.nn_xcd = c->tuning.nn_xcd;
  best.nn_seed = c->tuning.nn_seed;
  best.xz_slab = c->tuning.xz_slab;
  best.ms_fwd_x = best_ms[0];
  best.ms_fwd_z = best_ms[1];
  best.ms_inv_x = best_ms[2];
  apply_tuning(c, best);
  clear_graphs(c);
  c->tuning.ms_fwd_x = best_ms[0];
  c->tuning.ms_fwd_z = best_ms[1];
  c->tuning.ms_inv_x = best_ms[2];
  c->tuning.ms_plane = best.plane_fwd > 0 ? plane_ms : classic_ms;
  c->tuning.ms_fused = best.fused_reg > 0 ? fused_ms[1] : fused_ms[0];
  c->tuning.ms_so3 = best.so3_reg > 0 ? so3_ms[1] : so3_ms[0];
  if (out) *out = c->tuning;
  return 0;
}

int phz_mfma_peak(phz_ctx* c, int dtype, int iters, double* tflops) {
  PHZ_CHECK(c && tflops, "null argument");
  PHZ_CHECK(dtype >= 0 && dtype <= 2, "dtype must be 0 (f16), 1 (bf16) or 2 (f32)");
  PHZ_CHECK(iters >= 1 && iters <= (1 << 20), "iters must be in [1, 2^20]");
  constexpr int kBlocks = 2048;  // 8192 waves: 8 per SIMD of the 256 CUs
  CallBuffers buf;
  float* out;
  PHZ_HIP(buf.get(&out, (size_t)kBlocks * 256));
  hipEvent_t a, b;
  PHZ_HIP(hipEventCreate(&a));
  PHZ_HIP(hipEventCreate(&b));
  launch_mfma_peak(dtype, iters, kBlocks, out, c->stream);  // warm-up
  PHZ_HIP(hipEventRecord(a, c->stream));
  const double flops = launch_mfma_peak(dtype, iters, kBlocks, out, c->stream);
  PHZ_HIP(hipEventRecord(b, c->stream));
  PHZ_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  PHZ_HIP(hipEventElapsedTime(&ms, a, b));
  hipEventDestroy(a);
  hipEventDestroy(b);
  PHZ_HIP(hipGetLastError());
  *tflops = flops / (ms * 1e-3) / 1e12;
  return 0;
}

int phz_copy_kernel(phz_ctx* c, void* dst, const void* src, size_t bytes, int repeats,
                    float* ms) {
  PHZ_CHECK(c && dst && src, "null argument");
  PHZ_CHECK(bytes % 16 == 0 && repeats >= 1, "bytes must be a multiple of 16, repeats >= 1");
  hipEvent_t a, b;
  PHZ_HIP(hipEventCreate(&a));
  PHZ_HIP(hipEventCreate(&b));
  PHZ_HIP(hipEventRecord(a, c->stream));
  for (int r = 0; r < repeats; ++r)
    launch_copy_f4((float4*)dst, (const float4*)src, bytes / 16, c->stream);
  PHZ_HIP(hipEventRecord(b, c->stream));
  PHZ_HIP(hipEventSynchronize(b));
  float t = 0.f;
  PHZ_HIP(hipEventElapsedTime(&t, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (ms) *ms = t / (float)repeats;
  return 0;
}

int phz_synchronize(phz_ctx* c) {
  PHZ_CHECK(c, "null argument");
  PHZ_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"
