// Host build of the per-pair pose math of phaser_amd/csrc/pose.hpp (the
// same inline functions the pose kernels call), exported for the CPU tests.
#include "pose.hpp"

extern "C" {

void pose_eigen_sym4(double* A, double* d, double* V) { phz::eigen_sym4(A, d, V); }

void pose_bingham_from_moment(const double* C, const double* qp, int multi, double* mode,
                              double* omega, double* axes) {
  phz::bingham_from_moment(C, qp, multi != 0, mode, omega, axes);
}

void pose_quat_of_index(uint32_t idx, int bw, double* q) {
  double zyz[3];
  phz::zyz_of_index(idx, bw, zyz);
  phz::quat_of_zyz(zyz, q);
}

void pose_translation_of_index(uint32_t idx, int N, double width, double* t) {
  phz::translation_of_index(idx, N, width, t);
}

void pose_euler_and_transform(const double* q, double* euler, float* T) {
  phz::euler_and_transform(q, euler, T);
}
}
