// host_driver.cc -- registers two point clouds through the C++ mirror of the
// reference API (phaser_core::CloudController("sph-opt"), the path of
// phaser_core/test/driver.cc:30-57) and prints the result as JSON.
// usage: host_driver target source bandwidth n_voxels [bingham gaussian
// [method refine_rot_xyz [refine_rot_xyz2 estimate_translation2]]]: with the
// last two, the same controller registers the pair a second time after the
// refine_rot flags and FLAGS_estimate_translation were changed (one JSON
// line per registration); a .bin file is float32 (x, y, z, intensity) quadruples and a .ply
// file is read by model::PointCloud::readFromFile (with reflectivity and
// ambient when it has them).  PHZ_TEST_PRE="vg", "pt" or "vg,pt" switches the
// controller's CloudPreProcessor commands on (the pre-processor flags).
// usage: host_driver preprocess in.bin out.bin voxel_grid pass_through [leaf]:
// a preproc::CloudPreProcessor with those settings over one cloud, the
// filtered cloud written as .bin.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <vector>

#include "phaser_core.hpp"

static void print(const model::RegistrationResult& r);

static model::PointCloudPtr load(const char* path) {
  const size_t len = std::strlen(path);
  if (len > 4 && std::strcmp(path + len - 4, ".ply") == 0) {
    auto c = std::make_shared<model::PointCloud>();
    c->readFromFile(path);
    return c;
  }
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", path);
    std::exit(2);
  }
  const size_t bytes = (size_t)f.tellg();
  f.seekg(0);
  std::vector<common::Point_t> pts(bytes / sizeof(common::Point_t));
  f.read(reinterpret_cast<char*>(pts.data()), (std::streamsize)(pts.size() * sizeof(common::Point_t)));
  return std::make_shared<model::PointCloud>(std::move(pts));
}

static int preprocess(int argc, char** argv) {
  if (argc != 6 && argc != 7) {
    std::fprintf(stderr, "usage: %s preprocess in.bin out.bin voxel_grid pass_through [leaf]\n",
                 argv[0]);
    return 2;
  }
  preproc::CloudPreProcessorSettings settings;
  settings.enable_voxel_grid_downsampling = argv[4][0] == '1';
  settings.enable_pass_through_gnd_filtering = argv[5][0] == '1';
  if (argc == 7) FLAGS_phaser_voxel_grid_leaf_size = std::atof(argv[6]);
  auto cloud = load(argv[2]);
  preproc::CloudPreProcessor pre(settings);
  pre.process(cloud);
  std::ofstream f(argv[3], std::ios::binary);
  f.write(reinterpret_cast<const char*>(cloud->data()),
          (std::streamsize)(cloud->size() * sizeof(common::Point_t)));
  std::printf("{\"points\": %zu}\n", cloud->size());
  return f ? 0 : 1;
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::strcmp(argv[1], "preprocess") == 0) return preprocess(argc, argv);
  if (const char* pre = std::getenv("PHZ_TEST_PRE")) {
    FLAGS_phaser_pre_enable_voxel_grid_downsampling = std::strstr(pre, "vg") != nullptr;
    FLAGS_phaser_pre_enable_pass_through_gnd_filtering = std::strstr(pre, "pt") != nullptr;
  }
  if (argc != 5 && argc != 7 && argc != 9 && argc != 11) {
    std::fprintf(stderr,
                 "usage: %s target.bin source.bin bandwidth n_voxels "
                 "[bingham_peak_neighbors gaussian_peak_neighbors [method refine_rot_xyz]]\n",
                 argv[0]);
    return 2;
  }
  phaser_core::FLAGS_phaser_core_spherical_bandwidth = std::atoi(argv[3]);
  FLAGS_spherical_bandwith = std::atoi(argv[3]);
  phaser_core::FLAGS_phaser_core_spatial_n_voxels = std::atoi(argv[4]);
  if (argc >= 7) {
    FLAGS_bingham_peak_neighbors = std::atoi(argv[5]);
    FLAGS_gaussian_peak_neighbors = std::atoi(argv[6]);
  }
  std::string method = "sph-opt";
  if (argc >= 9) {  // method ("sph" or "sph-opt") and refine_rot_x/y/z as "101"
    method = argv[7];
    FLAGS_refine_rot_x = argv[8][0] == '1';
    FLAGS_refine_rot_y = argv[8][1] == '1';
    FLAGS_refine_rot_z = argv[8][2] == '1';
  }
  auto target = load(argv[1]);
  auto source = load(argv[2]);
  phaser_core::CloudController controller(std::move(method));
  print(controller.registerPointCloud(target, source));
  if (argc == 11) {
    FLAGS_refine_rot_x = argv[9][0] == '1';
    FLAGS_refine_rot_y = argv[9][1] == '1';
    FLAGS_refine_rot_z = argv[9][2] == '1';
    FLAGS_estimate_translation = argv[10][0] == '1';
    print(controller.registerPointCloud(target, source));
  }
  return 0;
}

static void print(const model::RegistrationResult& r) {
  const auto s = r.getStateAsVec();
  const auto& corr = r.getRotationCorrelation();
  double cmax = 0;
  for (double v : corr) cmax = v > cmax ? v : cmax;
  const auto omega = r.getRotUncertaintyEstimate()->getOmega();
  // no positional estimate when the translation was not estimated
  std::array<double, 9> cov{};
  if (r.getPosUncertaintyEstimate()) cov = r.getPosUncertaintyEstimate()->getCov();
  double reg_sum = 0.0;
  const auto& reg = *r.getRegisteredCloud();
  for (size_t i = 0; i < reg.size(); ++i) {
    const auto& p = reg.pointAt(i);
    reg_sum += (double)p.x + (double)p.y + (double)p.z;
  }
  // the covariance is NaN where a window weight is negative (as in the
  // reference's Gaussian); JSON as Python reads it
  auto num = [](double v) {
    char b[32];
    if (std::isnan(v)) return std::string("NaN");
    std::snprintf(b, sizeof(b), "%.17g", v);
    return std::string(b);
  };
  std::printf("{\"rot_index\": %u, \"trans_index\": %u, \"state\": [%.17g, %.17g, %.17g, %.17g, %.17g, %.17g, %.17g, %.17g], "
              "\"registered_points\": %zu, \"rot_corr_size\": %zu, \"rot_corr_max\": %.9g, "
              "\"rot_omega\": [%.17g, %.17g, %.17g, %.17g], \"trans_cov_diag\": [%s, %s, %s], "
              "\"has_pos_estimate\": %d, \"registered_sum\": %.17g}\n",
              r.rotation_index, r.translation_index, s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7],
              r.getRegisteredCloud()->size(), corr.size(), cmax, omega[0], omega[1], omega[2],
              omega[3], num(cov[0]).c_str(), num(cov[4]).c_str(), num(cov[8]).c_str(),
              r.getPosUncertaintyEstimate() ? 1 : 0, reg_sum);
  std::fflush(stdout);
}
