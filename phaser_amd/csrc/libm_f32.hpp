// libm_f32.hpp -- acosf and atan2f rounded exactly as the reference's host
// libm rounds them, for the S^2 projection of every point
// (spherical-projection.cc:19-41 calls std::acos and std::atan2 on floats).
//
// The projection decides which point is a query's nearest neighbour, and
// neither function is correctly rounded in glibc (up to 2.40): on scan
// clouds, 6-8% of acosf and 12-16% of atan2f results differ in the last bit
// from the correctly rounded value, and the device library's last bits
// differ again.  Points that then sit a bit apart flip near-equidistant
// nearest-neighbour picks, which move whole correlation grids.  These are
// the float algorithms glibc uses (Sun's fdlibm float versions, freely
// redistributable: argument reduction, the rational and odd/even
// polynomial kernels and their constants), restated so that the device
// performs the same float operations in the same order, without FMA
// contraction, with IEEE division and square root.  Built for the host, the
// same code equals glibc's acosf on every float in [-1, 1] and its atan2f on
// the sampled domains of tests/test_libm_f32.py.
#pragma once

#include <cstdint>
#include <cstring>

#if defined(__HIP__)
#define PHZ_HD __host__ __device__
#else
#define PHZ_HD
#endif

namespace phz {
namespace libm {

PHZ_HD inline float from_bits(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

PHZ_HD inline uint32_t to_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// R(z) = z * P(z) / Q(z), the rational kernel of asin / acos on [0, 1/4]
PHZ_HD inline float asin_kernel(float z) {
#pragma clang fp contract(off)
  const float p0 = from_bits(0x3e2aaaabu), p1 = from_bits(0xbea6b090u),
              p2 = from_bits(0x3e4e0aa8u), p3 = from_bits(0xbd241146u),
              p4 = from_bits(0x3a4f7f04u), p5 = from_bits(0x3811ef08u);
  const float q1 = from_bits(0xc019d139u), q2 = from_bits(0x4001572du),
              q3 = from_bits(0xbf303361u), q4 = from_bits(0x3d9dc62eu);
  const float p = z * (p0 + z * (p1 + z * (p2 + z * (p3 + z * (p4 + z * p5)))));
  const float q = 1.0f + z * (q1 + z * (q2 + z * (q3 + z * q4)));
  return p / q;
}

// acosf: |x| < 1/2 as pi/2 - (x + x R(x^2)); x < -1/2 as pi - 2 (s + s R(z))
// with z = (1 + x) / 2, s = sqrt(z); x > 1/2 as 2 (s_hi + c + s R(z)) with z
// = (1 - x) / 2, s_hi = s with its 12 low bits cleared, c = (z - s_hi^2) /
// (s + s_hi) the correction of the truncation.
PHZ_HD inline float acosf_host_exact(float x) {
#pragma clang fp contract(off)
  const float pi = from_bits(0x40490fdau), pio2_hi = from_bits(0x3fc90fdau),
              pio2_lo = from_bits(0x33a22168u);
  const uint32_t hx = to_bits(x), ix = hx & 0x7fffffffu;
  if (ix == 0x3f800000u) return (hx >> 31) ? pi + 2.0f * pio2_lo : 0.0f;
  if (ix > 0x3f800000u) return (x - x) / (x - x);
  if (ix < 0x3f000000u) {
    if (ix <= 0x32800000u) return pio2_hi + pio2_lo;
    const float r = asin_kernel(x * x);
    return pio2_hi - (x - (pio2_lo - x * r));
  }
  if (hx >> 31) {
    const float z = (1.0f + x) * 0.5f;
    const float r = asin_kernel(z);
    const float s = sqrtf(z);
    const float w = r * s - pio2_lo;
    return pi - 2.0f * (s + w);
  }
  const float z = (1.0f - x) * 0.5f;
  const float s = sqrtf(z);
  const float df = from_bits(to_bits(s) & 0xfffff000u);
  const float c = (z - df * df) / (s + df);
  const float r = asin_kernel(z);
  const float w = r * s + c;
  return 2.0f * (df + w);
}

// atanf: reduction to |t| < 7/16 around 0, 1/2, 1, 3/2 or infinity (atan of
// the breakpoint in two parts), then t - t (s1 + s2) with the degree-22 odd
// polynomial split into its even and odd halves in w = t^4.
PHZ_HD inline float atanf_host_exact(float x) {
#pragma clang fp contract(off)
  const float hi[4] = {from_bits(0x3eed6338u), from_bits(0x3f490fdau), from_bits(0x3f7b985eu),
                       from_bits(0x3fc90fdau)};
  const float lo[4] = {from_bits(0x31ac3769u), from_bits(0x33222168u), from_bits(0x33140fb4u),
                       from_bits(0x33a22168u)};
  const uint32_t hx = to_bits(x), ix = hx & 0x7fffffffu;
  if (ix >= 0x4c000000u) {  // |x| >= 2^25
    if (ix > 0x7f800000u) return x + x;
    return (hx >> 31) ? -hi[3] - lo[3] : hi[3] + lo[3];
  }
  int id;
  float t = x;
  if (ix < 0x3ee00000u) {  // |x| < 7/16
    if (ix < 0x39800000u) return x;  // |x| < 2^-12
    id = -1;
  } else {
    t = (hx >> 31) ? -x : x;
    if (ix < 0x3f980000u) {    // |x| < 19/16
      if (ix < 0x3f300000u) {  // 7/16 <= |x| < 11/16
        id = 0;
        t = (2.0f * t - 1.0f) / (2.0f + t);
      } else {
        id = 1;
        t = (t - 1.0f) / (t + 1.0f);
      }
    } else if (ix < 0x401c0000u) {  // |x| < 39/16
      id = 2;
      t = (t - 1.5f) / (1.0f + 1.5f * t);
    } else {
      id = 3;
      t = -1.0f / t;
    }
  }
  const float a0 = from_bits(0x3eaaaaabu), a1 = from_bits(0xbe4ccccdu),
              a2 = from_bits(0x3e124925u), a3 = from_bits(0xbde38e38u),
              a4 = from_bits(0x3dba2e6eu), a5 = from_bits(0xbd9d8795u),
              a6 = from_bits(0x3d886b35u), a7 = from_bits(0xbd6ef16bu),
              a8 = from_bits(0x3d4bda59u), a9 = from_bits(0xbd15a221u),
              a10 = from_bits(0x3c8569d7u);
  const float z = t * t;
  const float w = z * z;
  const float s1 = z * (a0 + w * (a2 + w * (a4 + w * (a6 + w * (a8 + w * a10)))));
  const float s2 = w * (a1 + w * (a3 + w * (a5 + w * (a7 + w * a9))));
  if (id < 0) return t - t * (s1 + s2);
  const float r = hi[id] - ((t * (s1 + s2) - lo[id]) - t);
  return (hx >> 31) ? -r : r;
}

// atan2f: the quadrant from the signs, atanf(|y / x|), then pi - (z - pi_lo)
// or (z - pi_lo) - pi for x < 0.
PHZ_HD inline float atan2f_host_exact(float y, float x) {
#pragma clang fp contract(off)
  const float pi = from_bits(0x40490fdbu), pi_o_2 = from_bits(0x3fc90fdbu),
              pi_o_4 = from_bits(0x3f490fdbu), pi_lo = from_bits(0xb3bbbd2eu);
  const float tiny = 1.0e-30f;
  const uint32_t hx = to_bits(x), ix = hx & 0x7fffffffu;
  const uint32_t hy = to_bits(y), iy = hy & 0x7fffffffu;
  if (ix > 0x7f800000u || iy > 0x7f800000u) return x + y;
  if (hx == 0x3f800000u) return atanf_host_exact(y);
  const int m = (int)((hy >> 31) & 1u) | (int)((hx >> 30) & 2u);
  if (iy == 0) {
    if (m < 2) return y;
    return m == 2 ? pi + tiny : -pi - tiny;
  }
  if (ix == 0) return (hy >> 31) ? -pi_o_2 - tiny : pi_o_2 + tiny;
  if (ix == 0x7f800000u) {
    if (iy == 0x7f800000u) {
      switch (m) {
        case 0: return pi_o_4 + tiny;
        case 1: return -pi_o_4 - tiny;
        case 2: return 3.0f * pi_o_4 + tiny;
        default: return -3.0f * pi_o_4 - tiny;
      }
    }
    switch (m) {
      case 0: return 0.0f;
      case 1: return -0.0f;
      case 2: return pi + tiny;
      default: return -pi - tiny;
    }
  }
  if (iy == 0x7f800000u) return (hy >> 31) ? -pi_o_2 - tiny : pi_o_2 + tiny;
  const int k = ((int)iy - (int)ix) >> 23;
  float z;
  if (k > 26) {
    z = pi_o_2 + 0.5f * pi_lo;
  } else if ((hx >> 31) && k < -26) {
    z = 0.0f;
  } else {
    z = atanf_host_exact(fabsf(y / x));
  }
  switch (m) {
    case 0: return z;
    case 1: return -z;
    case 2: return pi - (z - pi_lo);
    default: return (z - pi_lo) - pi;
  }
}

}  // namespace libm
}  // namespace phz
