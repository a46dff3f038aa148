// host_driver.cc -- registers two point clouds through the C++ mirror of the
// reference API (phaser_core::CloudController("sph-opt"), the path of
// phaser_core/test/driver.cc:30-57) and prints the result as JSON.
// usage: host_driver target.bin source.bin bandwidth n_voxels
// where a .bin file is float32 (x, y, z, intensity) quadruples.
#include <cstdio>
#include <fstream>
#include <vector>

#include "phaser_core.hpp"

static model::PointCloudPtr load(const char* path) {
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

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s target.bin source.bin bandwidth n_voxels\n", argv[0]);
    return 2;
  }
  phaser_core::FLAGS_phaser_core_spherical_bandwidth = std::atoi(argv[3]);
  phaser_core::FLAGS_phaser_core_spatial_n_voxels = std::atoi(argv[4]);
  auto target = load(argv[1]);
  auto source = load(argv[2]);
  phaser_core::CloudController controller("sph-opt");
  model::RegistrationResult r = controller.registerPointCloud(target, source);
  const auto s = r.getStateAsVec();
  const auto& corr = r.getRotationCorrelation();
  double cmax = 0;
  for (double v : corr) cmax = v > cmax ? v : cmax;
  std::printf("{\"rot_index\": %u, \"trans_index\": %u, \"state\": [%.17g, %.17g, %.17g, %.17g, %.17g, %.17g, %.17g, %.17g], "
              "\"registered_points\": %zu, \"rot_corr_size\": %zu, \"rot_corr_max\": %.9g}\n",
              r.rotation_index, r.translation_index, s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7],
              r.getRegisteredCloud()->size(), corr.size(), cmax);
  return 0;
}
