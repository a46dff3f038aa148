// phaser_host.cc -- the C++ classes of include/phaser_core.hpp over the C
// ABI of include/phaser_amd.h.  Host-side orchestration only: every numeric
// stage runs in libphaser_amd.so on the GPU.
#include "phaser_core.hpp"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "phaser_amd.h"
#include "phaser_ply.h"

namespace phaser_core {
int32_t FLAGS_phaser_core_spherical_bandwidth = 75;
int32_t FLAGS_phaser_core_spherical_zero_padding = 0;
int32_t FLAGS_phaser_core_spherical_low_pass_lower_bound = 0;
int32_t FLAGS_phaser_core_spherical_low_pass_upper_bound = 100000;
int32_t FLAGS_phaser_core_spatial_n_voxels = 100;
int32_t FLAGS_phaser_core_spatial_discretize_lower = -50;
int32_t FLAGS_phaser_core_spatial_discretize_upper = 50;
int32_t FLAGS_phaser_core_spatial_zero_padding = 0;
int32_t FLAGS_phaser_core_spatial_low_pass_lower_bound = 0;
int32_t FLAGS_phaser_core_spatial_low_pass_upper_bound = 100000;
}  // namespace phaser_core

int32_t FLAGS_spherical_bandwith = 75;
std::string FLAGS_alignment_algorithm = "phase";
std::string FLAGS_rot_evaluation_algorithm = "bingham";
std::string FLAGS_pos_evaluation_algorithm = "gaussian";
bool FLAGS_refine_rot_x = true, FLAGS_refine_rot_y = true, FLAGS_refine_rot_z = true;
bool FLAGS_estimate_rotation = true, FLAGS_estimate_translation = true;
int32_t FLAGS_bingham_peak_neighbors = 0;
int32_t FLAGS_gaussian_peak_neighbors = 0;
int32_t FLAGS_bmm_peak_neighbors = 2;  // bmm-peak-based-eval.cc:7-9
int32_t FLAGS_gmm_peak_neighbors = 2;  // gmm-peak-based-eval.cc:12-14
std::string FLAGS_phaser_ply_intensity_str = "intensity";
std::string FLAGS_phaser_ply_reflectivity_str = "reflectivity";
std::string FLAGS_phaser_ply_ambient_str = "noise";
std::string FLAGS_phaser_ply_range_str = "range";

namespace {

// glog CHECK equivalent: the reference aborts on every failed precondition
[[noreturn]] void fatal(const std::string& what) {
  std::fprintf(stderr, "F phaser_core: %s\n", what.c_str());
  std::fflush(stderr);
  std::abort();
}

void check(int rc, const char* call) {
  if (rc != 0) fatal(std::string(call) + " failed: " + phz_last_error());
}

#define CHECK_TRUE(cond, msg) \
  do {                        \
    if (!(cond)) fatal(msg);  \
  } while (0)

constexpr int64_t kDefaultCapacity = 262144;

}  // namespace

namespace phaser_core {

std::shared_ptr<phz_ctx> acquireContext(int bw, int n_voxels, int64_t max_points, int channels,
                                        int sph_pad, int method, int skip_rot_axes,
                                        int eval_mask, int spat_pad) {
  if (spat_pad < 0) spat_pad = FLAGS_phaser_core_spatial_zero_padding;
  static std::mutex mu;
  static std::map<std::tuple<int, int, int64_t, int, int, int, int, int, int, int, int, int, int,
                             int, int, int, int>,
                  std::weak_ptr<phz_ctx>>
      cache;
  std::lock_guard<std::mutex> lock(mu);
  const int64_t cap = max_points > kDefaultCapacity ? max_points : kDefaultCapacity;
  const auto key = std::make_tuple(bw, n_voxels, cap, FLAGS_phaser_core_spatial_discretize_lower,
                                   FLAGS_phaser_core_spatial_discretize_upper,
                                   FLAGS_phaser_core_spatial_low_pass_lower_bound,
                                   FLAGS_phaser_core_spatial_low_pass_upper_bound,
                                   FLAGS_bingham_peak_neighbors, FLAGS_gaussian_peak_neighbors,
                                   channels, sph_pad, method, skip_rot_axes,
                                   spat_pad, eval_mask,
                                   FLAGS_bmm_peak_neighbors, FLAGS_gmm_peak_neighbors);
  if (auto p = cache[key].lock()) return p;
  phz_config cfg;
  phz_default_config(&cfg);
  cfg.bandwidth = bw;
  cfg.n_voxels = n_voxels;
  cfg.max_points = cap;
  cfg.discretize_lower = FLAGS_phaser_core_spatial_discretize_lower;
  cfg.discretize_upper = FLAGS_phaser_core_spatial_discretize_upper;
  cfg.low_pass_lower = FLAGS_phaser_core_spatial_low_pass_lower_bound;
  cfg.low_pass_upper = FLAGS_phaser_core_spatial_low_pass_upper_bound;
  cfg.bingham_peak_neighbors = FLAGS_bingham_peak_neighbors;
  cfg.gaussian_peak_neighbors = FLAGS_gaussian_peak_neighbors;
  cfg.channels = channels;
  cfg.spherical_zero_padding = sph_pad;
  cfg.method = method;
  cfg.skip_rot_axes = skip_rot_axes;
  cfg.spatial_zero_padding = spat_pad;
  cfg.rot_eval = (eval_mask & 1) ? PHZ_EVAL_MIXTURE : PHZ_EVAL_BINGHAM;
  cfg.pos_eval = (eval_mask & 2) ? PHZ_EVAL_MIXTURE : PHZ_EVAL_GAUSSIAN;
  cfg.bmm_peak_neighbors = FLAGS_bmm_peak_neighbors;
  cfg.gmm_peak_neighbors = FLAGS_gmm_peak_neighbors;
  phz_ctx* raw = nullptr;
  check(phz_create(&cfg, &raw), "phz_create");
  std::shared_ptr<phz_ctx> p(raw, [](phz_ctx* c) { phz_destroy(c); });
  cache[key] = p;
  return p;
}

// ---------------------------------------------------------------- spherical
SphericalCorrelation::SphericalCorrelation(const uint32_t bw, const uint32_t zero_padding)
    : bw_(bw), zero_padding_(zero_padding) {
  CHECK_TRUE(zero_padding == 0, "spherical zero padding other than 0 is not supported");
  ctx_ = acquireContext((int)bw, FLAGS_phaser_core_spatial_n_voxels, kDefaultCapacity);
}

void SphericalCorrelation::correlateSignals(const std::vector<model::FunctionValue>& f1,
                                            const std::vector<model::FunctionValue>& f2) {
  SampledSignal a, b;
  for (const auto& v : f1) a.push_back(0.6 * v.getAveragedIntensity() + 0.4 * v.getAveragedRange());
  for (const auto& v : f2) b.push_back(0.6 * v.getAveragedIntensity() + 0.4 * v.getAveragedRange());
  correlateSampledSignals({a}, {b});
}

void SphericalCorrelation::correlateSampledSignals(const std::vector<SampledSignal>& f1,
                                                   const std::vector<SampledSignal>& f2) {
  CHECK_TRUE(f1.size() == f2.size(), "channel counts differ");
  CHECK_TRUE(!f1.empty(), "no channels");
  const size_t nq = 4ull * bw_ * bw_;
  CHECK_TRUE(f1[0].size() == nq && f2[0].size() == nq, "signal size is not (2 bw)^2");
  std::vector<float> grids(2 * nq);
  for (size_t i = 0; i < nq; ++i) {
    grids[i] = (float)f1[0][i];
    grids[nq + i] = (float)f2[0][i];
  }
  std::vector<float> coef(2 * 2 * (size_t)bw_ * bw_);
  check(phz_s2_transform(ctx_.get(), 2, grids.data(), coef.data()), "phz_s2_transform");
  const size_t n3 = 8ull * bw_ * bw_ * bw_;
  std::vector<float> corr(n3);
  check(phz_so3_correlate_coef(ctx_.get(), coef.data(), coef.data() + 2 * (size_t)bw_ * bw_,
                               corr.data(), nullptr, &argmax_),
        "phz_so3_correlate_coef");
  corr_.assign(corr.begin(), corr.end());
}

void SphericalCorrelationLaplace::correlateSampledSignals(const std::vector<SampledSignal>& f,
                                                          const std::vector<SampledSignal>& g) {
  CHECK_TRUE(f.size() == g.size(), "channel counts differ");
  CHECK_TRUE(f.size() > 1u, "the Laplace fusion needs at least two channels");
  const int C = (int)f.size();
  const size_t nq = 4ull * bw_ * bw_;
  std::vector<float> fg(C * nq), gg(C * nq);
  for (int c = 0; c < C; ++c) {
    CHECK_TRUE(f[c].size() == nq && g[c].size() == nq, "signal size is not (2 bw)^2");
    for (size_t i = 0; i < nq; ++i) {
      fg[c * nq + i] = (float)f[c][i];
      gg[c * nq + i] = (float)g[c][i];
    }
  }
  std::vector<float> corr(8ull * bw_ * bw_ * bw_);
  check(phz_spherical_correlate(ctx_.get(), C, fg.data(), gg.data(), corr.data(), &argmax_),
        "phz_spherical_correlate");
  corr_.assign(corr.begin(), corr.end());
}

// ---------------------------------------------------------------- spatial
// spatial-correlation.cc:11-61: the correlation of N + 2 pad voxels per
// axis, of which the first getCorrelationSize() entries are the vector
SpatialCorrelation::SpatialCorrelation(const uint32_t n_voxels, const uint32_t zero_padding)
    : n_voxels_(n_voxels), zero_padding_(zero_padding), total_(n_voxels * n_voxels * n_voxels) {
  const uint32_t per_dim = 2u * zero_padding;
  const double ratio = (n_voxels + per_dim) / static_cast<double>(n_voxels);
  total_padded_ = zero_padding ? static_cast<uint32_t>(total_ * (ratio * ratio * ratio)) : total_;
  ctx_ = acquireContext(FLAGS_phaser_core_spherical_bandwidth, (int)n_voxels, kDefaultCapacity, 2,
                        0, 0, 0, 0, (int)zero_padding);
}

double* SpatialCorrelation::correlateSignals(const std::vector<VectorXd*>& f,
                                             const std::vector<VectorXd*>& g) {
  CHECK_TRUE(!f.empty() && !g.empty(), "no channels");
  CHECK_TRUE(f[0]->size() == total_ && g[0]->size() == total_, "signal size is not N^3");
  std::vector<float> a(f[0]->begin(), f[0]->end()), b(g[0]->begin(), g[0]->end());
  const size_t np = (size_t)n_voxels_ + 2u * zero_padding_;
  std::vector<float> corr(np * np * np);
  check(phz_spatial_correlate_single_padded(ctx_.get(), (int)n_voxels_, (int)zero_padding_,
                                            a.data(), b.data(), corr.data(), nullptr),
        "phz_spatial_correlate_single_padded");
  c_.assign(corr.begin(), corr.end());
  return c_.data();
}

double* SpatialCorrelationLaplace::correlateSignals(const std::vector<VectorXd*>& f,
                                                    const std::vector<VectorXd*>& g) {
  CHECK_TRUE(f.size() == g.size() && f.size() > 1u, "the Laplace fusion needs >= 2 channels");
  const int C = (int)f.size();
  std::vector<float> a((size_t)C * total_), b((size_t)C * total_);
  for (int c = 0; c < C; ++c) {
    CHECK_TRUE(f[c]->size() == total_ && g[c]->size() == total_, "signal size is not N^3");
    for (uint32_t i = 0; i < total_; ++i) {
      a[(size_t)c * total_ + i] = (float)(*f[c])[i];
      b[(size_t)c * total_ + i] = (float)(*g[c])[i];
    }
  }
  const size_t np = (size_t)n_voxels_ + 2u * zero_padding_;
  std::vector<float> corr(np * np * np);
  check(phz_spatial_correlate(ctx_.get(), C, a.data(), b.data(), corr.data(), nullptr),
        "phz_spatial_correlate");
  c_.assign(corr.begin(), corr.end());
  return c_.data();
}

PhaseAligner::PhaseAligner()
    : n_voxels_(FLAGS_phaser_core_spatial_n_voxels),
      lower_bound_(FLAGS_phaser_core_spatial_discretize_lower),
      upper_bound_(FLAGS_phaser_core_spatial_discretize_upper) {
  ctx_ = acquireContext(FLAGS_phaser_core_spherical_bandwidth, (int)n_voxels_, kDefaultCapacity);
  spatial_correlation_.reset(
      new SpatialCorrelationLaplace(n_voxels_, FLAGS_phaser_core_spatial_zero_padding));
}

void PhaseAligner::discretizePointcloud(const model::PointCloud& cloud,
                                        std::vector<VectorXd>* f) const {
  const size_t T = (size_t)n_voxels_ * n_voxels_ * n_voxels_;
  std::shared_ptr<phz_ctx> ctx = ctx_;
  if ((int64_t)cloud.size() > kDefaultCapacity)
    ctx = acquireContext(FLAGS_phaser_core_spherical_bandwidth, (int)n_voxels_, cloud.size());
  std::vector<float> out(2 * T);
  check(phz_voxelize(ctx.get(), cloud.data(), (int64_t)cloud.size(), out.data(), nullptr),
        "phz_voxelize");
  f->assign(4, VectorXd(T, 0.0));  // {intensity, range, reflectivity, ambient}
  for (size_t v = 0; v < T; ++v) {
    (*f)[0][v] = out[v];
    (*f)[1][v] = out[T + v];
  }
}

void PhaseAligner::alignRegistered(const model::PointCloud& cloud_prev,
                                   const std::vector<model::FunctionValue>&,
                                   const model::PointCloud& cloud_reg,
                                   const std::vector<model::FunctionValue>&) {
  std::vector<VectorXd> f, g;
  discretizePointcloud(cloud_prev, &f);
  discretizePointcloud(cloud_reg, &g);
  std::vector<VectorXd*> fp, gp;
  for (int c = 0; c < 4; ++c) {
    fp.push_back(&f[c]);
    gp.push_back(&g[c]);
  }
  double* c = spatial_correlation_->correlateSignals(fp, gp);
  previous_correlation_.assign(c, c + spatial_correlation_->getCorrelationSize());
}

// ---------------------------------------------------------------- registration
SphOptRegistration::SphOptRegistration()
    : SphOptRegistration("SphOptRegistration", FLAGS_phaser_core_spherical_bandwidth,
                         PHZ_METHOD_SPH_OPT, 0, FLAGS_phaser_core_spherical_zero_padding) {}

SphOptRegistration::SphOptRegistration(const std::string& name, int bandwidth, int method,
                                       int skip_rot_axes, int sph_pad)
    : BaseRegistration(name),
      bandwidth_(bandwidth),
      method_(method),
      skip_rot_axes_(skip_rot_axes),
      sph_pad_(sph_pad) {
  CHECK_TRUE(sph_pad_ >= 0, "negative spherical zero padding");
  ctx_ = acquireContext(bandwidth_, FLAGS_phaser_core_spatial_n_voxels, kDefaultCapacity, 2,
                        sph_pad_, method_, skip_rot_axes_, eval_mask_);
  check(phz_set_profiling(ctx_.get(), 1), "phz_set_profiling");
}

void SphOptRegistration::setBandwith(const int bandwith) {
  bandwidth_ = bandwith;
  ctx_ = acquireContext(bandwidth_, FLAGS_phaser_core_spatial_n_voxels, kDefaultCapacity, 2,
                        sph_pad_, method_, skip_rot_axes_);
}

static int refine_mask() {
  return (FLAGS_refine_rot_x ? 0 : 1) | (FLAGS_refine_rot_y ? 0 : 2) |
         (FLAGS_refine_rot_z ? 0 : 4);
}

SphRegistration::SphRegistration()
    : SphRegistration(std::string(FLAGS_alignment_algorithm),
                      std::string(FLAGS_rot_evaluation_algorithm),
                      std::string(FLAGS_pos_evaluation_algorithm)) {}

SphRegistration::SphRegistration(std::string&& alignment_algorithm,
                                 std::string&& rot_evaluation_algorithm,
                                 std::string&& pos_evaluation_algorithm)
    : SphOptRegistration("SphRegistration", FLAGS_spherical_bandwith, PHZ_METHOD_SPH,
                         refine_mask(), 0) {
  // sph-registration.cc:56-79
  CHECK_TRUE(alignment_algorithm == "phase",
             "unknown alignment algorithm '" + alignment_algorithm + "' (phase only)");
  CHECK_TRUE(rot_evaluation_algorithm == "bingham" || rot_evaluation_algorithm == "bmm",
             "Unknown rot evaluation algorithm specificed: " + rot_evaluation_algorithm);
  CHECK_TRUE(pos_evaluation_algorithm == "gaussian" || pos_evaluation_algorithm == "gmm",
             "Unknown pos evaluation algorithm specificed: " + pos_evaluation_algorithm);
  const int mask = (rot_evaluation_algorithm == "bmm" ? 1 : 0) |
                   (pos_evaluation_algorithm == "gmm" ? 2 : 0);
  if (mask != eval_mask_) {
    eval_mask_ = mask;
    ctx_ = acquireContext(bandwidth_, FLAGS_phaser_core_spatial_n_voxels, kDefaultCapacity, 2,
                          sph_pad_, method_, skip_rot_axes_, eval_mask_);
    check(phz_set_profiling(ctx_.get(), 1), "phz_set_profiling");
  }
}

model::RegistrationResult SphRegistration::registerPointCloud(model::PointCloudPtr cloud_prev,
                                                              model::PointCloudPtr cloud_cur) {
  // sph-registration.cc:89-96: no rotation estimate is "not yet implemented"
  CHECK_TRUE(FLAGS_estimate_rotation, "not yet implemented.");
  // the refine_rot flags are read on every registration (sph-registration.cc:
  // 108-113): a change since the last call selects the context of the new mask
  const int mask = refine_mask();
  if (mask != skip_rot_axes_) {
    skip_rot_axes_ = mask;
    ctx_ = acquireContext(bandwidth_, FLAGS_phaser_core_spatial_n_voxels, kDefaultCapacity, 2,
                          sph_pad_, method_, skip_rot_axes_, eval_mask_);
    check(phz_set_profiling(ctx_.get(), 1), "phz_set_profiling");
  }
  if (FLAGS_estimate_translation)
    return SphOptRegistration::registerPointCloud(cloud_prev, cloud_cur);
  // rotation only (sph-registration.cc:91-93): the registered cloud is the
  // source rotated by the estimate, with no positional estimate.  The device
  // pipeline always runs both stages; its translation is not used.
  model::PointCloud rotated = cloud_cur->clone();
  model::RegistrationResult full = SphOptRegistration::registerPointCloud(cloud_prev, cloud_cur);
  rotated.transformPointCloud(full.rotation_transform);
  model::RegistrationResult result(std::move(rotated));
  result.setRotation(full.getRotation());
  result.setRotationCorrelation(full.getRotationCorrelation());
  result.rotation_index = full.rotation_index;
  result.rotation_transform = full.rotation_transform;
  result.setRotUncertaintyEstimate(full.getRotUncertaintyEstimate());
  return result;
}

model::RegistrationResult SphOptRegistration::registerPointCloud(model::PointCloudPtr cloud_prev,
                                                                 model::PointCloudPtr cloud_cur) {
  CHECK_TRUE(cloud_prev != nullptr, "cloud_prev is null");
  CHECK_TRUE(cloud_cur != nullptr, "cloud_cur is null");
  const int64_t need = (int64_t)std::max(cloud_prev->size(), cloud_cur->size());
  // clouds read from PLY files may carry reflectivity and ambient: then all
  // four channels of the fusions are live
  auto has_extra = [](const model::PointCloud& c) {
    return c.hasReflectivityPoints() || c.hasAmbientNoisePoints();
  };
  const bool four = has_extra(*cloud_prev) || has_extra(*cloud_cur);
  std::shared_ptr<phz_ctx> ctx = ctx_;
  if (need > kDefaultCapacity || four)
    ctx = acquireContext(bandwidth_, FLAGS_phaser_core_spatial_n_voxels, need, four ? 4 : 2,
                         sph_pad_, method_, skip_rot_axes_, eval_mask_);
  auto extra_of = [](model::PointCloud& c) {
    std::vector<float> e(2 * c.size(), 0.f);
    for (size_t i = 0; i < c.size(); ++i) {
      if (c.hasReflectivityPoints()) e[2 * i] = (float)c.getReflectivity(i);
      if (c.hasAmbientNoisePoints()) e[2 * i + 1] = (float)c.getAmbientNoise(i);
    }
    return e;
  };
  std::vector<float> ep, ec;
  if (four) {
    ep = extra_of(*cloud_prev);
    ec = extra_of(*cloud_cur);
  }
  phz_result r;
  // the SO(3) grid of bandwidth + spherical zero padding (getBandwidth()
  // returns bw_out, spherical-correlation.cc:118-120)
  const size_t bo = (size_t)bandwidth_ + (size_t)sph_pad_;
  const size_t n3 = 8ull * bo * bo * bo;
  std::vector<float> rot(n3);
  model::PointCloud reg = cloud_cur->clone();
  check(phz_register_channels(ctx.get(), cloud_prev->data(), four ? ep.data() : nullptr,
                              (int64_t)cloud_prev->size(), cloud_cur->data(),
                              four ? ec.data() : nullptr, (int64_t)cloud_cur->size(), &r,
                              rot.data(), nullptr, reg.data()),
        "phz_register_channels");
  model::RegistrationResult result(std::move(reg));
  result.setRotation({r.quat[0], r.quat[1], r.quat[2], r.quat[3]});
  result.setTranslation({r.translation[0], r.translation[1], r.translation[2]});
  result.setRotationCorrelation(std::vector<double>(rot.begin(), rot.end()));
  result.rotation_index = r.rot_index;
  result.translation_index = r.trans_index;
  for (int i = 0; i < 16; ++i) result.rotation_transform[i] = r.T_rot[i];
  std::array<double, 4> mode, omega;
  std::array<double, 16> M;
  std::array<double, 3> mu;
  std::array<double, 9> cov;
  for (int i = 0; i < 4; ++i) mode[i] = r.quat[i], omega[i] = r.rot_omega[i];
  for (int i = 0; i < 16; ++i) M[i] = r.rot_axes[i];
  for (int i = 0; i < 3; ++i) mu[i] = r.translation[i];
  for (int i = 0; i < 9; ++i) cov[i] = r.trans_cov[i];
  result.setRotUncertaintyEstimate(std::make_shared<common::Bingham>(mode, omega, M));
  result.setPosUncertaintyEstimate(std::make_shared<common::Gaussian>(mu, cov));
  return result;
}

std::vector<std::pair<std::string, double>> SphOptRegistration::getStatistics() const {
  phz_stage_time t[32];
  int n = 0;
  check(phz_get_stage_times(ctx_.get(), t, 32, &n), "phz_get_stage_times");
  std::vector<std::pair<std::string, double>> out;
  for (int i = 0; i < n && i < 32; ++i) out.emplace_back(t[i].name, t[i].total_ms);
  return out;
}

CloudController::CloudController(const std::string& method) {
  if (method == "sph")
    registrator_.reset(new SphRegistration());
  else if (method == "sph-opt")
    registrator_.reset(new SphOptRegistration());
  else
    fatal("Unknown registration method specified: " + method);
}

model::RegistrationResult CloudController::registerPointCloud(model::PointCloudPtr cloud_prev,
                                                              model::PointCloudPtr cloud_cur) {
  CHECK_TRUE(cloud_prev && cloud_cur, "CloudController: null cloud");
  preprocessor_.process(cloud_prev);
  preprocessor_.process(cloud_cur);
  return registrator_->registerPointCloud(cloud_prev, cloud_cur);
}

}  // namespace phaser_core

bool FLAGS_phaser_pre_enable_voxel_grid_downsampling = false;
bool FLAGS_phaser_pre_enable_pass_through_gnd_filtering = false;
bool FLAGS_phaser_pre_enable_geometric_cloud_segmentation = false;
double FLAGS_phaser_voxel_grid_leaf_size = 0.25;
double FLAGS_phaser_pass_through_gnd_filter_limit_min = -5.0;
double FLAGS_phaser_pass_through_gnd_filter_limit_max = 3.0;

namespace preproc {
namespace {

// the filters need only a stream on the current device: one small context
// shared by every command of the process
phz_ctx* filterContext() {
  static std::mutex mu;
  static std::shared_ptr<phz_ctx> ctx;
  std::lock_guard<std::mutex> lock(mu);
  if (!ctx) {
    phz_config cfg;
    phz_default_config(&cfg);
    cfg.bandwidth = 8;
    cfg.n_voxels = 8;
    cfg.max_points = 16;
    phz_ctx* c = nullptr;
    check(phz_create(&cfg, &c), "phz_create");
    ctx.reset(c, phz_destroy);
  }
  return ctx.get();
}

void resizeAfterFilter(model::PointCloud& cloud, int64_t kept) {
  if ((size_t)kept == cloud.size()) return;
  cloud.points().resize((size_t)kept);
  cloud.reflectivities().clear();
  cloud.ambientPoints().clear();
}

}  // namespace

CloudPreProcessorSettings CloudPreProcessorSettings::fromGflags() {
  CloudPreProcessorSettings s;
  s.enable_voxel_grid_downsampling = FLAGS_phaser_pre_enable_voxel_grid_downsampling;
  s.enable_pass_through_gnd_filtering = FLAGS_phaser_pre_enable_pass_through_gnd_filtering;
  s.enable_geometric_cloud_segmentation = FLAGS_phaser_pre_enable_geometric_cloud_segmentation;
  return s;
}

void VoxelGridCmd::execute(model::PointCloudPtr cloud) {
  CHECK_TRUE(cloud != nullptr, "VoxelGridCmd: null cloud");
  int64_t kept = 0;
  check(phz_voxel_grid_filter(filterContext(), cloud->data(), (int64_t)cloud->size(),
                              (float)FLAGS_phaser_voxel_grid_leaf_size, &kept),
        "phz_voxel_grid_filter");
  resizeAfterFilter(*cloud, kept);
}

void PassThroughGndFilterCmd::execute(model::PointCloudPtr cloud) {
  CHECK_TRUE(cloud != nullptr, "PassThroughGndFilterCmd: null cloud");
  int64_t kept = 0;
  check(phz_pass_through_filter(filterContext(), cloud->data(), (int64_t)cloud->size(), 2,
                                (float)FLAGS_phaser_pass_through_gnd_filter_limit_min,
                                (float)FLAGS_phaser_pass_through_gnd_filter_limit_max, 1, &kept),
        "phz_pass_through_filter");
  resizeAfterFilter(*cloud, kept);
}

CloudPreProcessor::CloudPreProcessor()
    : CloudPreProcessor(CloudPreProcessorSettings::fromGflags()) {}

CloudPreProcessor::CloudPreProcessor(const CloudPreProcessorSettings& settings) {
  initializeCommandFromSettings(settings);
}

void CloudPreProcessor::initializeCommandFromSettings(const CloudPreProcessorSettings& settings) {
  if (settings.enable_voxel_grid_downsampling) processors_.emplace_back(new VoxelGridCmd());
  if (settings.enable_pass_through_gnd_filtering)
    processors_.emplace_back(new PassThroughGndFilterCmd());
  if (settings.enable_geometric_cloud_segmentation)
    fatal("SegmentCloudCmd (geometric cloud segmentation) is not part of this build");
}

void CloudPreProcessor::process(model::PointCloudPtr cloud) {
  for (auto& p : processors_) p->execute(cloud);
}

}  // namespace preproc

// ---------------------------------------------------------------- model
namespace model {

void PointCloud::transformPointCloud(const std::array<float, 16>& T) {
  for (auto& p : points_) {
    const float x = T[0] * p.x + T[1] * p.y + T[2] * p.z + T[3];
    const float y = T[4] * p.x + T[5] * p.y + T[6] * p.z + T[7];
    const float z = T[8] * p.x + T[9] * p.y + T[10] * p.z + T[11];
    p.x = x;
    p.y = y;
    p.z = z;
  }
}

void PointCloud::readFromFile(const std::string& ply) {
  CHECK_TRUE(!ply.empty(), "empty PLY path");
  phz_ply_fields f;
  f.intensity = FLAGS_phaser_ply_intensity_str.c_str();
  f.reflectivity = FLAGS_phaser_ply_reflectivity_str.c_str();
  f.ambient = FLAGS_phaser_ply_ambient_str.c_str();
  f.range = FLAGS_phaser_ply_range_str.c_str();
  phz_ply_cloud c;
  if (phz_ply_read(ply.c_str(), &f, &c) != 0)
    fatal(std::string("reading ") + ply + ": " + phz_ply_last_error());
  ply_read_directory_ = ply;
  // parsePlyPointCloud appends to the cloud (point-cloud.cc:330-349)
  for (int64_t i = 0; i < c.n; ++i) {
    common::Point_t p;
    p.x = c.xyzi[4 * i];
    p.y = c.xyzi[4 * i + 1];
    p.z = c.xyzi[4 * i + 2];
    p.intensity = c.xyzi[4 * i + 3];
    points_.push_back(p);
  }
  reflectivities_.clear();
  ambient_points_.clear();
  ranges_.clear();
  if (c.has_reflectivity)
    for (int64_t i = 0; i < c.n; ++i) reflectivities_.push_back(c.extra[2 * i]);
  if (c.has_ambient)
    for (int64_t i = 0; i < c.n; ++i) ambient_points_.push_back(c.extra[2 * i + 1]);
  if (c.has_range) ranges_.assign(c.range, c.range + c.n);
  phz_ply_free(&c);
}

std::array<double, 8> RegistrationResult::getStateAsVec() const {
  return {quat_[0], quat_[1], quat_[2], quat_[3], 0.0, trans_[0], trans_[1], trans_[2]};
}

}  // namespace model

namespace common {

SphericalSampler::SphericalSampler(const int bandwith) { initialize(bandwith); }

void SphericalSampler::initialize(const int bandwith) {
  bandwith_ = bandwith;
  ctx_ = phaser_core::acquireContext(bandwith, phaser_core::FLAGS_phaser_core_spatial_n_voxels,
                                     kDefaultCapacity);
}

void SphericalSampler::sampleUniformly(const model::PointCloud& cloud,
                                       std::vector<model::FunctionValue>* grid) {
  CHECK_TRUE(grid != nullptr, "grid is null");
  std::shared_ptr<phz_ctx> ctx = ctx_;
  if ((int64_t)cloud.size() > kDefaultCapacity)
    ctx = phaser_core::acquireContext(bandwith_, phaser_core::FLAGS_phaser_core_spatial_n_voxels,
                                      cloud.size());
  const size_t nq = 4ull * bandwith_ * bandwith_;
  std::vector<float> g(2 * nq);
  check(phz_sample_s2(ctx.get(), cloud.data(), (int64_t)cloud.size(), g.data()), "phz_sample_s2");
  grid->clear();
  grid->reserve(nq);
  for (size_t i = 0; i < nq; ++i) grid->emplace_back(g[i], g[nq + i]);
}

}  // namespace common
