// libm_f32_check.cc -- phz::libm::acosf_host_exact / atanf_host_exact /
// atan2f_host_exact (phaser_amd/csrc/libm_f32.hpp, built for the host)
// against this host's libm: acosf on every float in [-1, 1] (and beyond:
// NaN), atanf on every float, atan2f on random and structured pairs.
// usage: libm_f32_check [stride] -- stride 1 is exhaustive.  Prints the
// mismatch counts and the first mismatches; exit status 1 on any mismatch.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../phaser_amd/csrc/libm_f32.hpp"

using phz::libm::from_bits;
using phz::libm::to_bits;

static std::atomic<long> bad_acos{0}, bad_atan{0}, bad_atan2{0}, checked{0};

static bool same(float a, float b) {
  return to_bits(a) == to_bits(b) || (std::isnan(a) && std::isnan(b));
}

static void report(const char* what, uint32_t u, uint32_t v, float got, float want) {
  std::printf("%s(0x%08x, 0x%08x): 0x%08x, libm 0x%08x\n", what, u, v, to_bits(got), to_bits(want));
}

int main(int argc, char** argv) {
  const uint64_t stride = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1;
  const int T = 8;
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([t, stride] {
      long ba = 0, bt = 0, bt2 = 0, n = 0;
      // every bit pattern (both signs) with the given stride, offset per thread
      for (uint64_t u = (uint64_t)t * stride; u < (1ull << 32); u += (uint64_t)T * stride) {
        const float x = from_bits((uint32_t)u);
        if ((u & 0x7fffffffu) <= 0x3f800001u || (u & 0x7fffffffu) > 0x7f800000u) {
          const float a = phz::libm::acosf_host_exact(x), b = acosf(x);
          if (!same(a, b) && ba++ < 4) report("acosf", (uint32_t)u, 0, a, b);
        }
        const float c = phz::libm::atanf_host_exact(x), d = atanf(x);
        if (!same(c, d) && bt++ < 4) report("atanf", (uint32_t)u, 0, c, d);
        ++n;
      }
      // atan2f: random pairs, half over all finite bit patterns, half with
      // the exponents of coordinates in metres (2^-10 .. 2^10)
      uint64_t s = 0x9e3779b97f4a7c15ull * (t + 1);
      auto next = [&s] {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
      };
      const long pairs = (long)((1ull << 28) / stride / T);
      for (long i = 0; i < pairs; ++i) {
        uint32_t uy = (uint32_t)next(), ux = (uint32_t)next();
        if (i & 1) {
          uy = (uy & 0x807fffffu) | ((117u + (uint32_t)(next() % 21)) << 23);
          ux = (ux & 0x807fffffu) | ((117u + (uint32_t)(next() % 21)) << 23);
          if ((i & 15) == 1) ux = (ux & 0x80000000u) | (uy & 0x7fffffffu);  // |x| = |y|
        }
        const float y = from_bits(uy), x = from_bits(ux);
        const float a = phz::libm::atan2f_host_exact(y, x), b = atan2f(y, x);
        if (!same(a, b) && bt2++ < 4) report("atan2f", uy, ux, a, b);
      }
      // special values
      const float sp[] = {0.0f, -0.0f, 1.0f, -1.0f, INFINITY, -INFINITY, NAN, 1e-40f, -1e-40f,
                          3.0e38f, -3.0e38f, 1e-30f, 2.0f, 0.5f};
      if (t == 0)
        for (float y : sp)
          for (float x : sp) {
            const float a = phz::libm::atan2f_host_exact(y, x), b = atan2f(y, x);
            if (!same(a, b) && bt2++ < 8) report("atan2f", to_bits(y), to_bits(x), a, b);
          }
      bad_acos += ba;
      bad_atan += bt;
      bad_atan2 += bt2;
      checked += n;
    });
  for (auto& x : th) x.join();
  std::printf("checked %ld bit patterns (stride %llu): acosf %ld, atanf %ld, atan2f %ld mismatches\n",
              checked.load(), (unsigned long long)stride, bad_acos.load(), bad_atan.load(),
              bad_atan2.load());
  return (bad_acos || bad_atan || bad_atan2) ? 1 : 0;
}
