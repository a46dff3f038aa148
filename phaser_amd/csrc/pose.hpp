// pose.hpp -- per-pair pose math on the device: quaternions of SO(3) grid
// indices, translations of voxel grid indices, the eigen-decomposition of a
// Bingham moment matrix, and the XYZ Euler angles and float transform of a
// quaternion.  All in FP64 (a handful of values per pair).
#pragma once

#include "common.hpp"

namespace phz {

// rotation-utils.cc:190-207: index [beta][alpha][gamma] of the (2 bw)^3 grid
// -> ZYZ Euler angles
__host__ __device__ inline void zyz_of_index(uint32_t idx, int bw, double zyz[3]) {
  const uint32_t n = 2 * bw;
  const uint32_t ii = idx / (n * n), jj = (idx / n) % n, kk = idx % n;
  const double bd = (double)bw;
  zyz[0] = M_PI * jj / bd;
  zyz[1] = M_PI * (2 * ii + 1) / (4. * bw);
  zyz[2] = M_PI * kk / bd;
}

__host__ __device__ inline void qmul(const double* x, const double* y, double* r) {
  r[0] = x[0] * y[0] - x[1] * y[1] - x[2] * y[2] - x[3] * y[3];
  r[1] = x[0] * y[1] + x[1] * y[0] + x[2] * y[3] - x[3] * y[2];
  r[2] = x[0] * y[2] + x[2] * y[0] + x[3] * y[1] - x[1] * y[3];
  r[3] = x[0] * y[3] + x[3] * y[0] + x[1] * y[2] - x[2] * y[1];
}

// rotation-utils.cc:101-107: q = Rz(a) Ry(b) Rz(c) as (w, x, y, z)
__host__ __device__ inline void quat_of_zyz(const double zyz[3], double q[4]) {
  const double qa[4] = {cos(0.5 * zyz[0]), 0, 0, sin(0.5 * zyz[0])};
  const double qb[4] = {cos(0.5 * zyz[1]), 0, sin(0.5 * zyz[1]), 0};
  const double qc[4] = {cos(0.5 * zyz[2]), 0, 0, sin(0.5 * zyz[2])};
  double t[4];
  qmul(qa, qb, t);
  qmul(t, qc, q);
}

// translation-utils.cc:40-52 of the three subscripts of a linear index of
// the N^3 grid (signal-utils.cc:11-19: x*N + y + N^2*z)
__host__ __device__ inline void translation_of_index(uint32_t idx, int N, double width, double t[3]) {
  uint32_t xyz[3];
  xyz[1] = idx % (uint32_t)N;
  const uint32_t u = idx / (uint32_t)N;
  xyz[0] = u % (uint32_t)N;
  xyz[2] = u / (uint32_t)N;
  const double half = N / 2.0;
  for (int d = 0; d < 3; ++d) {
    const double index = (double)xyz[d];
    if (index == 0)
      t[d] = 0.0;
    else if (index <= half)
      t[d] = index * width / N;
    else
      t[d] = (index - N) * width / N;
  }
}

__host__ __device__ inline void mat4_mul_nc(const float* A, const float* B, float* R) {
#pragma clang fp contract(off)
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      float s = A[4 * i] * B[j];
      s = s + A[4 * i + 1] * B[4 + j];
      s = s + A[4 * i + 2] * B[8 + j];
      s = s + A[4 * i + 3] * B[12 + j];
      R[4 * i + j] = s;
    }
}

// rotation-utils.cc:86-99 (XYZ Euler angles of q) and :17-23,109-143
// (T = Tx Ty Tz in float, as RotateAroundXYZ builds it).  skip bits 0/1/2:
// the x/y/z angle is left out of T (refine_rot_x/y/z false,
// sph-registration.cc:114-119); euler keeps all three.
__host__ __device__ inline void euler_and_transform(const double q[4], double euler[3], float T[16],
                                                    int skip = 0) {
  const double qw = q[0], qx = q[1], qy = q[2], qz = q[3];
  const double r11 = -2 * (qy * qz - qw * qx);
  const double r12 = qw * qw - qx * qx - qy * qy + qz * qz;
  const double r21 = 2 * (qx * qz + qw * qy);
  const double r31 = -2 * (qx * qy - qw * qz);
  const double r32 = qw * qw + qx * qx - qy * qy - qz * qz;
  euler[2] = -atan2(r31, r32);
  euler[1] = asin(r21);
  euler[0] = atan2(r11, r12);
  const float ea = (skip & 1) ? 0.f : (float)euler[0];
  const float eb = (skip & 2) ? 0.f : (float)euler[1];
  const float ec = (skip & 4) ? 0.f : (float)euler[2];
  float X[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  float Y[16], Z[16], XY[16];
  for (int e = 0; e < 16; ++e) Y[e] = Z[e] = X[e];
  X[5] = cosf(ea);
  X[6] = -sinf(ea);
  X[9] = sinf(ea);
  X[10] = cosf(ea);
  Y[0] = cosf(eb);
  Y[2] = sinf(eb);
  Y[8] = -sinf(eb);
  Y[10] = cosf(eb);
  Z[0] = cosf(ec);
  Z[1] = sinf(ec);
  Z[4] = -sinf(ec);
  Z[5] = cosf(ec);
  mat4_mul_nc(X, Y, XY);
  mat4_mul_nc(XY, Z, T);
}

// Eigen-decomposition of the symmetric 4x4 A (row-major, destroyed) by
// cyclic Jacobi rotations: on return d holds the eigenvalues in ascending
// order and column i of V (row-major) the unit eigenvector of d[i].
__host__ __device__ inline void eigen_sym4(double A[16], double d[4], double V[16]) {
  for (int i = 0; i < 16; ++i) V[i] = (i % 5 == 0) ? 1.0 : 0.0;
  double norm2 = 0.0;
  for (int i = 0; i < 16; ++i) norm2 += A[i] * A[i];
  for (int sweep = 0; sweep < 24; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 3; ++p)
      for (int q = p + 1; q < 4; ++q) off += A[p * 4 + q] * A[p * 4 + q];
    if (!(off > 1e-34 * norm2)) break;
    for (int p = 0; p < 3; ++p)
      for (int q = p + 1; q < 4; ++q) {
        const double apq = A[p * 4 + q];
        if (apq == 0.0) continue;
        // rotation in the (p, q) plane that zeroes A[p][q]: with
        // theta = cot(2 phi) = (A_qq - A_pp) / (2 A_pq), t = tan(phi) is the
        // smaller root of t^2 + 2 t theta - 1 = 0
        const double theta = (A[q * 4 + q] - A[p * 4 + p]) / (2.0 * apq);
        const double at = fabs(theta);
        double t = at > 1e150 ? 0.5 / at : 1.0 / (at + sqrt(theta * theta + 1.0));
        if (theta < 0) t = -t;
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 4; ++k) {  // A <- A J
          const double akp = A[k * 4 + p], akq = A[k * 4 + q];
          A[k * 4 + p] = c * akp - s * akq;
          A[k * 4 + q] = s * akp + c * akq;
        }
        for (int k = 0; k < 4; ++k) {  // A <- J^T A
          const double apk = A[p * 4 + k], aqk = A[q * 4 + k];
          A[p * 4 + k] = c * apk - s * aqk;
          A[q * 4 + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 4; ++k) {  // V <- V J
          const double vkp = V[k * 4 + p], vkq = V[k * 4 + q];
          V[k * 4 + p] = c * vkp - s * vkq;
          V[k * 4 + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < 4; ++i) d[i] = A[i * 5];
  for (int i = 0; i < 3; ++i) {  // selection sort, ascending
    int m = i;
    for (int j = i + 1; j < 4; ++j)
      if (d[j] < d[m]) m = j;
    if (m != i) {
      const double td = d[i];
      d[i] = d[m];
      d[m] = td;
      for (int k = 0; k < 4; ++k) {
        const double tv = V[k * 4 + i];
        V[k * 4 + i] = V[k * 4 + m];
        V[k * 4 + m] = tv;
      }
    }
  }
}

// The Bingham fit from the moment matrix C = sum w q q^T (upper triangle,
// row by row) of a window with more than one sample (multi), or from the
// peak quaternion qp alone: mode (the eigenvector of the largest
// eigenvalue), omega (the eigenvalues over their sum, ascending) and axes
// (row-major; column i the unit eigenvector of omega[i], the last = mode).
// An eigenvector's sign is free: the mode agrees with qp, the other columns
// have their largest component positive.  One sample: C = w q q^T, so the
// mode is qp itself, omega = (0, 0, 0, 1), and the unit quaternions
// qp i, qp j, qp k span the rest.
__host__ __device__ inline void bingham_from_moment(const double C[10], const double qp[4],
                                                    bool multi, double mode[4],
                                                    double omega[4], double axes[16]) {
  if (multi) {
    double A[16], ev[4], V[16];
    int e = 0;
    for (int r = 0; r < 4; ++r)
      for (int c = r; c < 4; ++c) A[r * 4 + c] = A[c * 4 + r] = C[e++];
    eigen_sym4(A, ev, V);
    const double tr = ev[0] + ev[1] + ev[2] + ev[3];
    for (int c = 0; c < 4; ++c) {
      double dot = 0.0, big = 0.0;
      for (int r = 0; r < 4; ++r) {
        dot += V[r * 4 + c] * qp[r];
        if (fabs(V[r * 4 + c]) > fabs(big)) big = V[r * 4 + c];
      }
      const bool flip = c == 3 ? dot < 0.0 : big < 0.0;
      for (int r = 0; r < 4; ++r) axes[r * 4 + c] = flip ? -V[r * 4 + c] : V[r * 4 + c];
      omega[c] = ev[c] / tr;
    }
    for (int r = 0; r < 4; ++r) mode[r] = axes[r * 4 + 3];
  } else {
    const double w = qp[0], x = qp[1], y = qp[2], z = qp[3];
    const double ax[4][4] = {{-x, w, z, -y}, {-y, -z, w, x}, {-z, y, -x, w}, {w, x, y, z}};
    for (int c = 0; c < 4; ++c) {
      for (int r = 0; r < 4; ++r) axes[r * 4 + c] = ax[c][r];
      omega[c] = c == 3 ? 1.0 : 0.0;
    }
    for (int r = 0; r < 4; ++r) mode[r] = qp[r];
  }
}

// Sum over the 64 lanes of a wavefront (fixed butterfly order, so the
// result does not depend on scheduling); every lane gets the sum.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

}  // namespace phz
