// phaser_core.hpp -- C++ host mirror of the reference's phaser_core
// interfaces for the correlation path, implemented over the C ABI of
// phaser_amd.h (libphaser_host.so).
//
// Names, argument meaning and error behaviour follow the reference:
//   BaseRegistration / SphOptRegistration  base-registration.h:11-27,
//                                          sph-opt-registration.h
//   SphericalCorrelation(+Laplace)         spherical-correlation.h:19-33
//   BaseSpatialCorrelation / SpatialCorrelation(+Laplace)
//                                          base-spatial-correlation.h:10-15,
//                                          spatial-correlation.h
//   PhaseAligner                           phase-aligner.h / base-aligner.h:12-21
//   common::SphericalSampler               spherical-sampler.h
//   CloudController                        cloud-controller.h:14-34
//   preproc::CloudPreProcessor / VoxelGridCmd / PassThroughGndFilterCmd
//                                          cloud-pre-processor.h,
//                                          voxel-grid-cmd.h,
//                                          pass-through-gnd-filter-cmd.h
//   model::PointCloud / FunctionValue / RegistrationResult (the fields the
//   correlation path reads and writes)
// Flags are the gflags of core-gflags.cc:6-39, as plain globals read by the
// constructors.  Errors abort the process with a message, as glog CHECK does
// in the reference.  Eigen types in the reference signatures are replaced by
// std::vector<double> (Eigen is not a dependency of this build).
#ifndef PHASER_CORE_HPP_
#define PHASER_CORE_HPP_

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

struct phz_ctx;

namespace phaser_core {
// core-gflags.cc:6-39
extern int32_t FLAGS_phaser_core_spherical_bandwidth;
extern int32_t FLAGS_phaser_core_spherical_zero_padding;
extern int32_t FLAGS_phaser_core_spherical_low_pass_lower_bound;
extern int32_t FLAGS_phaser_core_spherical_low_pass_upper_bound;
extern int32_t FLAGS_phaser_core_spatial_n_voxels;
extern int32_t FLAGS_phaser_core_spatial_discretize_lower;
extern int32_t FLAGS_phaser_core_spatial_discretize_upper;
extern int32_t FLAGS_phaser_core_spatial_zero_padding;
extern int32_t FLAGS_phaser_core_spatial_low_pass_lower_bound;
extern int32_t FLAGS_phaser_core_spatial_low_pass_upper_bound;
}  // namespace phaser_core

// Defined at file scope in the reference: bingham-peak-based-eval.cc:8-10 and
// gaussian-peak-based-eval.cc:9-11 (default 0: a single-sample fit at the
// argmax).
// sph-registration.cc:16-32 (SphRegistration, CloudController("sph"))
extern int32_t FLAGS_spherical_bandwith;               // 75
extern std::string FLAGS_alignment_algorithm;          // "phase"
extern std::string FLAGS_rot_evaluation_algorithm;     // "bingham"
extern std::string FLAGS_pos_evaluation_algorithm;     // "gaussian"
extern bool FLAGS_refine_rot_x, FLAGS_refine_rot_y, FLAGS_refine_rot_z;  // true
extern bool FLAGS_estimate_rotation, FLAGS_estimate_translation;         // true
extern int32_t FLAGS_bingham_peak_neighbors;
extern int32_t FLAGS_gaussian_peak_neighbors;
extern int32_t FLAGS_bmm_peak_neighbors;               // 2
extern int32_t FLAGS_gmm_peak_neighbors;               // 2
// PLY field names (ply-helper.cc:8-22)
extern std::string FLAGS_phaser_ply_intensity_str;     // "intensity"
extern std::string FLAGS_phaser_ply_reflectivity_str;  // "reflectivity"
extern std::string FLAGS_phaser_ply_ambient_str;       // "noise"
extern std::string FLAGS_phaser_ply_range_str;         // "range"

namespace common {
// pcl::PointXYZI fields used by the path (point-types.h:10-14)
struct Point_t {
  float x = 0, y = 0, z = 0, intensity = 0;
};
}  // namespace common

namespace common {
// The fitted distributions of the registration's uncertainty estimates
// (BaseDistribution::getEstimate, base-distribution.h).
// Bingham (bingham.h): the mode, the moment matrix's eigenvalues over their
// sum in ascending order (the statistic fitToMoment fits the concentrations
// to, bingham.cc:239-246) and M, whose columns are the matching eigenvectors.
class Bingham {
 public:
  Bingham() = default;
  Bingham(const std::array<double, 4>& mode, const std::array<double, 4>& omega,
          const std::array<double, 16>& M)
      : mode_(mode), omega_(omega), M_(M) {}
  std::vector<double> getEstimate() const { return {mode_.begin(), mode_.end()}; }
  std::array<double, 4> mode() const { return mode_; }
  std::array<double, 4> getOmega() const { return omega_; }
  std::array<double, 16> getM() const { return M_; }  // row-major

 private:
  std::array<double, 4> mode_{1, 0, 0, 0}, omega_{0, 0, 0, 1};
  std::array<double, 16> M_{};
};
using BinghamPtr = std::shared_ptr<Bingham>;

// Gaussian (gaussian.h): weighted mean and covariance of the peak window.
class Gaussian {
 public:
  Gaussian() = default;
  Gaussian(const std::array<double, 3>& mu, const std::array<double, 9>& cov)
      : mu_(mu), cov_(cov) {}
  std::vector<double> getEstimate() const { return {mu_.begin(), mu_.end()}; }
  std::array<double, 3> getMean() const { return mu_; }
  std::array<double, 9> getCov() const { return cov_; }  // row-major

 private:
  std::array<double, 3> mu_{0, 0, 0};
  std::array<double, 9> cov_{};
};
using GaussianPtr = std::shared_ptr<Gaussian>;
}  // namespace common

namespace model {

class PointCloud {
 public:
  PointCloud() = default;
  explicit PointCloud(std::vector<common::Point_t> points) : points_(std::move(points)) {}
  std::size_t size() const { return points_.size(); }
  common::Point_t& pointAt(std::size_t i) { return points_.at(i); }
  const common::Point_t& pointAt(std::size_t i) const { return points_.at(i); }
  const float* data() const { return reinterpret_cast<const float*>(points_.data()); }
  float* data() { return reinterpret_cast<float*>(points_.data()); }
  std::vector<common::Point_t>& points() { return points_; }
  // row-major 4x4, PCL scalar order (point-cloud.cc:183-185)
  void transformPointCloud(const std::array<float, 16>& T);
  PointCloud clone() const { return *this; }
  // PointCloud::readFromFile (point-cloud.cc:313-349): x, y, z and intensity
  // into the points, reflectivity and ambient (when the PLY has them) per
  // point; field names from the FLAGS_phaser_ply_* flags
  void readFromFile(const std::string& ply);
  bool hasReflectivityPoints() const { return !reflectivities_.empty(); }
  bool hasAmbientNoisePoints() const { return !ambient_points_.empty(); }
  double getReflectivity(std::size_t i) const { return reflectivities_.at(i); }
  double getAmbientNoise(std::size_t i) const { return ambient_points_.at(i); }
  std::vector<double>& reflectivities() { return reflectivities_; }
  std::vector<double>& ambientPoints() { return ambient_points_; }
  std::string getPlyReadDirectory() const { return ply_read_directory_; }

 private:
  std::vector<common::Point_t> points_;
  std::vector<double> reflectivities_, ambient_points_, ranges_;
  std::string ply_read_directory_;
};
using PointCloudPtr = std::shared_ptr<PointCloud>;

// function-value.h: one sample per grid point (sampling_neighbors = 1)
class FunctionValue {
 public:
  FunctionValue() = default;
  FunctionValue(double range, double intensity) : range_(range), intensity_(intensity) {}
  double getAveragedRange() const noexcept { return range_; }
  double getAveragedIntensity() const noexcept { return intensity_; }
  double getAveragedReflectivity() const noexcept { return 0.0; }
  double getAveragedAmbientNoise() const noexcept { return 0.0; }

 private:
  double range_ = 0.0, intensity_ = 0.0;
};
using FunctionValueVec = std::vector<FunctionValue>;

// registration-result.cc:13-99 and state.cc:10-20 (pose as an 8-vector)
class RegistrationResult {
 public:
  RegistrationResult() = default;
  explicit RegistrationResult(PointCloud&& reg) : reg_cloud_(std::make_shared<PointCloud>(std::move(reg))) {}
  PointCloudPtr getRegisteredCloud() const { return reg_cloud_; }
  void setRegisteredCloud(PointCloudPtr c) { reg_cloud_ = std::move(c); }
  // [qw, qx, qy, qz, 0, tx, ty, tz]
  std::array<double, 8> getStateAsVec() const;
  std::array<double, 4> getRotation() const { return quat_; }
  std::array<double, 3> getTranslation() const { return trans_; }
  void setRotation(const std::array<double, 4>& q) { quat_ = q; }
  void setTranslation(const std::array<double, 3>& t) { trans_ = t; }
  void setRotationCorrelation(std::vector<double> c) { rot_corr_ = std::move(c); }
  const std::vector<double>& getRotationCorrelation() const { return rot_corr_; }
  // registration-result.cc: the Bingham fit of the rotation peak and the
  // Gaussian fit of the translation peak (sph-opt-registration.cc:68,93)
  void setRotUncertaintyEstimate(common::BinghamPtr b) { rot_dist_ = std::move(b); }
  void setPosUncertaintyEstimate(common::GaussianPtr g) { pos_dist_ = std::move(g); }
  common::BinghamPtr getRotUncertaintyEstimate() const { return rot_dist_; }
  common::GaussianPtr getPosUncertaintyEstimate() const { return pos_dist_; }
  uint32_t rotation_index = 0, translation_index = 0;
  // the row-major float transform the rotation estimate applied to the source
  // (RotateAroundXYZ, rotation-utils.cc:17-23)
  std::array<float, 16> rotation_transform{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

 private:
  PointCloudPtr reg_cloud_;
  common::BinghamPtr rot_dist_;
  common::GaussianPtr pos_dist_;
  std::array<double, 4> quat_{1, 0, 0, 0};
  std::array<double, 3> trans_{0, 0, 0};
  std::vector<double> rot_corr_;
};

}  // namespace model

namespace phaser_core {

using SampledSignal = std::vector<double>;
using VectorXd = std::vector<double>;  // stands in for Eigen::VectorXd

// One context per (bandwidth, voxel count, capacity); shared by the objects
// below and released with the last of them.
// channels: 2, or 4 for clouds that carry reflectivity and ambient;
// sph_pad: spherical zero padding of the SO(3) correlation; method and
// skip_rot_axes: phz_config's; eval_mask bit 0 / 1: rot / pos evaluation
// PHZ_EVAL_MIXTURE; spat_pad: spatial zero padding (-1: the flag's).
std::shared_ptr<phz_ctx> acquireContext(int bw, int n_voxels, int64_t max_points,
                                        int channels = 2, int sph_pad = 0,
                                        int method = 0 /* PHZ_METHOD_SPH_OPT */,
                                        int skip_rot_axes = 0, int eval_mask = 0,
                                        int spat_pad = -1);

class SphericalCorrelation {
 public:
  explicit SphericalCorrelation(const uint32_t bw = 100, const uint32_t zero_padding = 0);
  virtual ~SphericalCorrelation() = default;
  // single blended channel 0.6 I + 0.4 R (spherical-correlation.cc:54-67,103-112)
  void correlateSignals(const std::vector<model::FunctionValue>& f1,
                        const std::vector<model::FunctionValue>& f2);
  // base class: first channel only (spherical-correlation.cc:69-80)
  virtual void correlateSampledSignals(const std::vector<SampledSignal>& f1,
                                       const std::vector<SampledSignal>& f2);
  std::vector<double> getCorrelation() const noexcept { return corr_; }
  uint32_t getBandwidth() const noexcept { return bw_ + zero_padding_; }
  uint32_t getArgmax() const noexcept { return argmax_; }
  void shutdown() {}

 protected:
  uint32_t bw_, zero_padding_;
  std::shared_ptr<phz_ctx> ctx_;
  std::vector<double> corr_;
  uint32_t argmax_ = 0;
};

// spherical-correlation-laplace.cc:14-40 (>= 2 channels, fused)
class SphericalCorrelationLaplace : public SphericalCorrelation {
 public:
  explicit SphericalCorrelationLaplace(const uint32_t bw, const uint32_t zero_padding = 0)
      : SphericalCorrelation(bw, zero_padding) {}
  void correlateSampledSignals(const std::vector<SampledSignal>& f,
                               const std::vector<SampledSignal>& g) override;
};

// base-spatial-correlation.h:10-15; the returned buffer is owned by the
// object and valid until the next call (spatial-correlation.cc:47,73)
class BaseSpatialCorrelation {
 public:
  virtual ~BaseSpatialCorrelation() = default;
  virtual double* correlateSignals(const std::vector<VectorXd*>& f,
                                   const std::vector<VectorXd*>& g) = 0;
  virtual uint32_t getZeroPadding() const = 0;
};

class SpatialCorrelation : public BaseSpatialCorrelation {
 public:
  explicit SpatialCorrelation(const uint32_t n_voxels, const uint32_t zero_padding = 0);
  double* correlateSignals(const std::vector<VectorXd*>& f,
                           const std::vector<VectorXd*>& g) override;
  uint32_t getZeroPadding() const override { return zero_padding_; }
  // spatial-correlation.cc:219-221: (N + 2 pad)^3 in double, truncated
  uint32_t getCorrelationSize() const { return total_padded_; }

 protected:
  uint32_t n_voxels_, zero_padding_, total_, total_padded_;
  std::shared_ptr<phz_ctx> ctx_;
  std::vector<double> c_;
};

class SpatialCorrelationLaplace : public SpatialCorrelation {
 public:
  explicit SpatialCorrelationLaplace(const uint32_t n_voxels, const uint32_t zero_padding = 0)
      : SpatialCorrelation(n_voxels, zero_padding) {}
  double* correlateSignals(const std::vector<VectorXd*>& f,
                           const std::vector<VectorXd*>& g) override;
};

// base-aligner.h:12-21 / phase-aligner.cc:19-148
class PhaseAligner {
 public:
  PhaseAligner();
  void alignRegistered(const model::PointCloud& cloud_prev,
                       const std::vector<model::FunctionValue>& f_prev,
                       const model::PointCloud& cloud_reg,
                       const std::vector<model::FunctionValue>& f_reg);
  std::vector<double> getCorrelation() const { return previous_correlation_; }
  // phase-aligner.cc:137-139: the padded voxel count
  uint32_t getNumberOfVoxels() const noexcept {
    return n_voxels_ + 2 * spatial_correlation_->getZeroPadding();
  }
  int32_t getLowerBound() const noexcept { return lower_bound_; }
  int32_t getUpperBound() const noexcept { return upper_bound_; }
  // {intensity, range, reflectivity, ambient} means per voxel
  void discretizePointcloud(const model::PointCloud& cloud, std::vector<VectorXd>* f) const;

 private:
  uint32_t n_voxels_;
  int32_t lower_bound_, upper_bound_;
  std::shared_ptr<phz_ctx> ctx_;
  std::unique_ptr<SpatialCorrelationLaplace> spatial_correlation_;
  std::vector<double> previous_correlation_;
};

}  // namespace phaser_core

namespace common {
// spherical-sampler.cc:8-64
class SphericalSampler {
 public:
  explicit SphericalSampler(const int bandwith);
  void initialize(const int bandwith);
  void sampleUniformly(const model::PointCloud& cloud, std::vector<model::FunctionValue>* grid);
  int getInitializedBandwith() const noexcept { return bandwith_; }

 private:
  int bandwith_;
  std::shared_ptr<phz_ctx> ctx_;
};
}  // namespace common

namespace phaser_core {

class BaseRegistration {
 public:
  explicit BaseRegistration(const std::string& name) : name_(name) {}
  virtual ~BaseRegistration() = default;
  virtual model::RegistrationResult registerPointCloud(model::PointCloudPtr cloud_prev,
                                                       model::PointCloudPtr cloud_cur) = 0;
  const std::string& getName() const { return name_; }

 protected:
  std::string name_;
};

// sph-opt-registration.cc:21-146 on the device: one fused GPU pipeline per
// call (sampling, S^2 transforms, fusion, SO(3) correlation, argmax, pose,
// voxelisation, spatial correlation, argmax, translation).
class SphOptRegistration : public BaseRegistration {
 public:
  SphOptRegistration();
  model::RegistrationResult registerPointCloud(model::PointCloudPtr cloud_prev,
                                               model::PointCloudPtr cloud_cur) override;
  void setBandwith(const int bandwith);
  // per-stage device time (ms, calls) since the last call
  std::vector<std::pair<std::string, double>> getStatistics() const;

 protected:
  // method: PHZ_METHOD_SPH_OPT or PHZ_METHOD_SPH; skip_rot_axes as phz_config
  SphOptRegistration(const std::string& name, int bandwidth, int method, int skip_rot_axes,
                     int sph_pad);
  int bandwidth_;
  int method_, skip_rot_axes_, sph_pad_;
  int eval_mask_ = 0;  // bit 0: rot evaluation "bmm", bit 1: pos evaluation "gmm"
  std::shared_ptr<phz_ctx> ctx_;
};

// sph-registration.cc:35-158 on the device: the same pipeline with one
// sampled channel 0.6 intensity + 0.4 range instead of the Laplace fusion
// (spherical-correlation.cc:54-112), bandwidth FLAGS_spherical_bandwith, no
// spherical zero padding, and the refine_rot_x/y/z masks on the applied
// rotation.  Evaluations "bingham" and "gaussian" (the defaults), or "bmm"
// and "gmm" (phz_config.rot_eval / pos_eval = PHZ_EVAL_MIXTURE, with
// FLAGS_bmm_peak_neighbors / FLAGS_gmm_peak_neighbors): with ZScoreEval's
// single argmax signal the Gaussian mixture is the GMM window's Gaussian and
// the Bingham mixture's one component is the BMM window's Bingham, whose mode
// is returned (the reference refits random Metropolis-Hastings samples of
// it, bingham-mixture.cc:55-78).  Alignment "phase" only (the reference has
// no other).
class SphRegistration : public SphOptRegistration {
 public:
  SphRegistration();
  SphRegistration(std::string&& alignment_algorithm, std::string&& rot_evaluation_algorithm,
                  std::string&& pos_evaluation_algorithm);
  model::RegistrationResult registerPointCloud(model::PointCloudPtr cloud_prev,
                                               model::PointCloudPtr cloud_cur) override;
};

}  // namespace phaser_core

// pre-processor-gflags.cc, voxel-grid-cmd.cc:5-7, pass-through-gnd-filter-cmd.cc:5-10
extern bool FLAGS_phaser_pre_enable_voxel_grid_downsampling;       // false
extern bool FLAGS_phaser_pre_enable_pass_through_gnd_filtering;    // false
extern bool FLAGS_phaser_pre_enable_geometric_cloud_segmentation;  // false
extern double FLAGS_phaser_voxel_grid_leaf_size;                   // 0.25
extern double FLAGS_phaser_pass_through_gnd_filter_limit_min;      // -5
extern double FLAGS_phaser_pass_through_gnd_filter_limit_max;      // 3

namespace preproc {

// cloud-pre-processor.h: which commands process() runs, in this order
struct CloudPreProcessorSettings {
  bool enable_voxel_grid_downsampling = false;
  bool enable_pass_through_gnd_filtering = false;
  bool enable_geometric_cloud_segmentation = false;
  static CloudPreProcessorSettings fromGflags();
};

// base-command.h: one filter applied to the cloud in place
class BaseCommand {
 public:
  virtual ~BaseCommand() = default;
  virtual void execute(model::PointCloudPtr cloud) = 0;
};

// voxel-grid-cmd.cc:11-24: phz_voxel_grid_filter with leaf
// FLAGS_phaser_voxel_grid_leaf_size on all three axes
class VoxelGridCmd : public BaseCommand {
 public:
  void execute(model::PointCloudPtr cloud) override;
};

// pass-through-gnd-filter-cmd.cc:14-26: phz_pass_through_filter on z with
// the flags' limits, negative (the points inside [min, max] are removed)
class PassThroughGndFilterCmd : public BaseCommand {
 public:
  void execute(model::PointCloudPtr cloud) override;
};

// cloud-pre-processor.cc:20-44.  The geometric segmentation command
// (SegmentCloudCmd, a PCL region-growing segmentation) is not part of this
// build: enabling it aborts at construction.  A filter that changes the
// point count drops the cloud's per-point reflectivity and ambient values.
class CloudPreProcessor {
 public:
  CloudPreProcessor();
  explicit CloudPreProcessor(const CloudPreProcessorSettings& settings);
  void process(model::PointCloudPtr cloud);

 private:
  void initializeCommandFromSettings(const CloudPreProcessorSettings& settings);
  std::vector<std::unique_ptr<BaseCommand>> processors_;
};

}  // namespace preproc

namespace phaser_core {

// cloud-controller.cc:10-34: both clouds go through the CloudPreProcessor
// (the pre-processor flags; every command off by default), then the
// registration.  "sph" (the reference's default) or "sph-opt"
class CloudController {
 public:
  explicit CloudController(const std::string& method = "sph");
  model::RegistrationResult registerPointCloud(model::PointCloudPtr cloud_prev,
                                               model::PointCloudPtr cloud_cur);

 private:
  std::unique_ptr<BaseRegistration> registrator_;
  preproc::CloudPreProcessor preprocessor_;
};

}  // namespace phaser_core

#endif  // PHASER_CORE_HPP_
