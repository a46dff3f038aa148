// phaser_core_driver -- the reference's command-line driver
// (phaser_core/test/driver.cc:1-57) over the C++ mirror: reads the target and
// source clouds from PLY files, registers them with
// CloudController("sph-opt") and writes the registered source cloud as an
// ASCII PLY; prints the state vector [qw qx qy qz 0 tx ty tz].
//
// Flags (--name=value or --name value): target_cloud, source_cloud,
// reg_cloud, and the core flags phaser_core_spherical_bandwidth,
// phaser_core_spherical_zero_padding, phaser_core_spatial_n_voxels,
// phaser_core_spatial_discretize_lower / _upper,
// phaser_core_spatial_low_pass_lower_bound / _upper_bound,
// bingham_peak_neighbors, gaussian_peak_neighbors and
// phaser_ply_intensity_str / _reflectivity_str / _ambient_str / _range_str.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>

#include "phaser_core.hpp"
#include "phaser_ply.h"

int main(int argc, char** argv) {
  std::map<std::string, std::string> flags;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.rfind("--", 0) != 0) {
      std::fprintf(stderr, "unexpected argument %s\n", argv[i]);
      return 2;
    }
    a = a.substr(2);
    const size_t eq = a.find('=');
    if (eq != std::string::npos) {
      flags[a.substr(0, eq)] = a.substr(eq + 1);
    } else if (i + 1 < argc) {
      flags[a] = argv[++i];
    } else {
      std::fprintf(stderr, "flag --%s needs a value\n", a.c_str());
      return 2;
    }
  }
  auto take = [&](const char* name, std::string def) {
    auto it = flags.find(name);
    if (it == flags.end()) return def;
    std::string v = it->second;
    flags.erase(it);
    return v;
  };
  auto take_int = [&](const char* name, int32_t* dst) {
    auto it = flags.find(name);
    if (it == flags.end()) return;
    *dst = (int32_t)std::atoi(it->second.c_str());
    flags.erase(it);
  };
  const std::string target = take("target_cloud", ""), source = take("source_cloud", "");
  const std::string reg = take("reg_cloud", "");
  take_int("phaser_core_spherical_bandwidth", &phaser_core::FLAGS_phaser_core_spherical_bandwidth);
  take_int("phaser_core_spherical_zero_padding",
           &phaser_core::FLAGS_phaser_core_spherical_zero_padding);
  take_int("phaser_core_spatial_n_voxels", &phaser_core::FLAGS_phaser_core_spatial_n_voxels);
  take_int("phaser_core_spatial_discretize_lower",
           &phaser_core::FLAGS_phaser_core_spatial_discretize_lower);
  take_int("phaser_core_spatial_discretize_upper",
           &phaser_core::FLAGS_phaser_core_spatial_discretize_upper);
  take_int("phaser_core_spatial_low_pass_lower_bound",
           &phaser_core::FLAGS_phaser_core_spatial_low_pass_lower_bound);
  take_int("phaser_core_spatial_low_pass_upper_bound",
           &phaser_core::FLAGS_phaser_core_spatial_low_pass_upper_bound);
  take_int("bingham_peak_neighbors", &FLAGS_bingham_peak_neighbors);
  take_int("gaussian_peak_neighbors", &FLAGS_gaussian_peak_neighbors);
  FLAGS_phaser_ply_intensity_str = take("phaser_ply_intensity_str", FLAGS_phaser_ply_intensity_str);
  FLAGS_phaser_ply_reflectivity_str =
      take("phaser_ply_reflectivity_str", FLAGS_phaser_ply_reflectivity_str);
  FLAGS_phaser_ply_ambient_str = take("phaser_ply_ambient_str", FLAGS_phaser_ply_ambient_str);
  FLAGS_phaser_ply_range_str = take("phaser_ply_range_str", FLAGS_phaser_ply_range_str);
  if (!flags.empty()) {
    std::fprintf(stderr, "unknown flag --%s\n", flags.begin()->first.c_str());
    return 2;
  }
  // driver.cc:27-33: all three paths are required (CHECKs there)
  if (target.empty() || source.empty() || reg.empty()) {
    std::fprintf(stderr,
                 "usage: %s --target_cloud=a.ply --source_cloud=b.ply --reg_cloud=out.ply "
                 "[core flags]\n",
                 argv[0]);
    return 2;
  }
  auto target_cloud = std::make_shared<model::PointCloud>();
  target_cloud->readFromFile(target);
  auto source_cloud = std::make_shared<model::PointCloud>();
  source_cloud->readFromFile(source);
  phaser_core::CloudController ctrl("sph-opt");
  model::RegistrationResult result = ctrl.registerPointCloud(target_cloud, source_cloud);
  const auto s = result.getStateAsVec();
  std::printf("Registration result: %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", s[0], s[1],
              s[2], s[3], s[4], s[5], s[6], s[7]);
  const model::PointCloudPtr r = result.getRegisteredCloud();
  if (phz_ply_write_xyzi(reg.c_str(), (int64_t)r->size(), r->data()) != 0) {
    std::fprintf(stderr, "%s\n", phz_ply_last_error());
    return 1;
  }
  std::printf("Wrote registered cloud to %s\n", reg.c_str());
  return 0;
}
