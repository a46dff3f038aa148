// fft_reg.hpp -- register-resident FFTs of power-of-two lengths 64..256 for
// gfx950: each line is split as L = R1 * R2 (four-step Cooley-Tukey,
// n = n2 + R2 n1, k = k1 + R1 k2):
//   X[k1 + R1 k2] = sum_n2 W_R2^(n2 k2) W_L^(n2 k1) sum_n1 x[n2 + R2 n1] W_R1^(n1 k1)
// R2 threads per line each load R1 elements straight from HBM, transform
// them in registers, apply the W_L twiddles, exchange once through LDS and
// finish with R1/R2 transforms of length R2 each, storing straight to HBM.
// One LDS round trip per line instead of one per radix stage.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace phz {

// cos(2 pi m / 16), m = 0..15
constexpr float kC16[16] = {
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
    0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
    0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};

constexpr int bitrev(int x, int bits) {
  int r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}
constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

// exp(DIR * 2 pi i j / len) for len | 16, as a compile-time constant
template <int DIR>
__host__ __device__ __forceinline__ float2 w16(int j, int len) {
  const int m = (j * (16 / len)) & 15;
  const int ms = (m + 12) & 15;  // sin(2 pi m / 16) = cos(2 pi (m - 4) / 16)
  return make_float2(kC16[m], DIR * kC16[ms]);
}

// In-register DFT of length R (2, 4, 8 or 16): v[k] <- sum_n v[n] exp(DIR 2 pi i nk / R),
// natural order in and out (radix-2 decimation in frequency, then the
// bit-reversal permutation, both resolved at compile time).
template <int R, int DIR>
__host__ __device__ __forceinline__ void dft_reg(float2 (&v)[R]) {
#pragma unroll
  for (int len = R; len >= 2; len >>= 1) {
#pragma unroll
    for (int b = 0; b < R; b += len) {
#pragma unroll
      for (int j = 0; j < len / 2; ++j) {
        const float2 u = v[b + j], w = v[b + j + len / 2];
        v[b + j] = cadd(u, w);
        const float2 d = csub(u, w);
        if (j == 0) {
          v[b + j + len / 2] = d;
        } else if (4 * j == len) {  // times exp(DIR i pi / 2) = DIR i
          v[b + j + len / 2] = DIR < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
        } else {
          v[b + j + len / 2] = cmul(d, w16<DIR>(j, len));
        }
      }
    }
  }
  float2 t[R];
#pragma unroll
  for (int k = 0; k < R; ++k) t[k] = v[bitrev(k, ilog2(R))];
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = t[k];
}

template <int L>
struct RegSplit {
  static constexpr int R1 = L >= 128 ? 16 : 8;
  static constexpr int R2 = L / R1;
};

// The twiddles W_L^(n2 k1), k1 < R1, of thread n2 (w[0] unused), read
// from tw[e] = exp(-2 pi i e / L).  A kernel loads them once, ahead of its
// data, so that both are in flight together and several transforms of the
// same thread share them.
template <int L>
struct Twiddles {
  float2 w[RegSplit<L>::R1];
  __device__ __forceinline__ float2 operator()(int k1) const { return w[k1]; }
};

// The same twiddles for a wave whose lanes 0-31 are thread n2 = a and lanes
// 32-63 thread n2 = a + 1, with a uniform over the wave: each W_L^(n2 k1) is
// a select between two scalar loads, so no vector registers hold the table
// between transforms.
template <int L>
struct PairTwiddles {
  const float2* __restrict__ tw;
  int a;
  bool hi;
  __device__ __forceinline__ float2 operator()(int k1) const {
    const float2 wa = tw[a * k1], wb = tw[(a + 1) * k1];
    return hi ? wb : wa;
  }
};
template <int L>
__device__ __forceinline__ Twiddles<L> load_twiddles(int n2, const float2* __restrict__ tw) {
  Twiddles<L> t;
  t.w[0] = make_float2(1.f, 0.f);
#pragma unroll
  for (int k1 = 1; k1 < RegSplit<L>::R1; ++k1) t.w[k1] = tw[n2 * k1];  // n2 * k1 < L
  return t;
}

// A kernel whose lines are LINES = 32 per workgroup, line = thread % 32 and
// n2 = thread / 32, has two threads n2 per wave (lanes 0-31 and 32-63): it
// can take PairTwiddles; otherwise Twiddles.
template <int L>
__device__ __forceinline__ void init_twiddles(Twiddles<L>& w, int n2,
                                              const float2* __restrict__ tw) {
  w = load_twiddles<L>(n2, tw);
}
template <int L>
__device__ __forceinline__ void init_twiddles(PairTwiddles<L>& w, int n2,
                                              const float2* __restrict__ tw) {
  w.tw = tw;
  w.a = __builtin_amdgcn_readfirstlane(n2 & ~1);
  w.hi = (n2 & 1) != 0;
}

// As fft_four_step, with the results kept in registers:
// out[j * R2 + k2] = X[k1 + R1 k2], k1 = n2 + R2 j.
template <int DIR, int L, class TW>
__device__ __forceinline__ void fft_four_step_reg(float2 (&v)[RegSplit<L>::R1], int n2,
                                                  float2* lbuf, const TW& tw,
                                                  float2 (&out)[RegSplit<L>::R1]) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2;
  dft_reg<R1, DIR>(v);
#pragma unroll
  for (int k1 = 1; k1 < R1; ++k1) {
    float2 w = tw(k1);
    if (DIR > 0) w.y = -w.y;
    v[k1] = cmul(v[k1], w);
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) lbuf[k1 * R2 + n2] = v[k1];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j) {
    const int k1 = n2 + R2 * j;
    float2 u[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) u[m] = lbuf[k1 * R2 + m];
    dft_reg<R2, DIR>(u);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) out[j * R2 + k2] = u[k2];
  }
}

// As fft_four_step_reg with the line's exchange buffer strided: entry i at
// lbuf[i * stride] (lets a line exchange through its own storage in a
// larger LDS array).
template <int DIR, int L>
__device__ __forceinline__ void fft_four_step_reg_strided(float2 (&v)[RegSplit<L>::R1], int n2,
                                                          float2* lbuf, int stride,
                                                          const Twiddles<L>& tw,
                                                          float2 (&out)[RegSplit<L>::R1]) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2;
  dft_reg<R1, DIR>(v);
#pragma unroll
  for (int k1 = 1; k1 < R1; ++k1) {
    float2 w = tw.w[k1];
    if (DIR > 0) w.y = -w.y;
    v[k1] = cmul(v[k1], w);
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) lbuf[(k1 * R2 + n2) * stride] = v[k1];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j) {
    const int k1 = n2 + R2 * j;
    float2 u[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) u[m] = lbuf[(k1 * R2 + m) * stride];
    dft_reg<R2, DIR>(u);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) out[j * R2 + k2] = u[k2];
  }
}

// The four-step transform of one line held by R2 threads of a workgroup.
// On entry thread n2 holds v[n1] = x[n2 + R2 n1]; lbuf is the line's LDS
// scratch (>= L entries).  store(k, X[k]) is called for k = k1 + R1 k2 with
// k1 = n2 + R2 j, j < R1 / R2.  Contains one __syncthreads, so every thread
// of the workgroup must call it.
template <int DIR, int L, class Store, class TW>
__device__ __forceinline__ void fft_four_step(float2 (&v)[RegSplit<L>::R1], int n2, float2* lbuf,
                                              const TW& tw, Store store) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2;
  dft_reg<R1, DIR>(v);
#pragma unroll
  for (int k1 = 1; k1 < R1; ++k1) {
    float2 w = tw(k1);
    if (DIR > 0) w.y = -w.y;
    v[k1] = cmul(v[k1], w);
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) lbuf[k1 * R2 + n2] = v[k1];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j) {
    const int k1 = n2 + R2 * j;
    float2 u[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) u[m] = lbuf[k1 * R2 + m];
    dft_reg<R2, DIR>(u);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) store(k1 + R1 * k2, u[k2]);
  }
}

// Lines of length L (64, 128 or 256) of complex data: line q of batch z
// starts at data + z*batch_stride + (q / inner) * L * inner + q % inner and
// has stride `inner`.  kLines lines per workgroup, R2 threads per line (the
// line index runs fastest across threads, so strided lines load coalesced).
// tw: exp(-2 pi i e / L), e < L (conjugated for DIR = +1).
template <int DIR, int L, int kLines>
__global__ void __launch_bounds__(kLines * RegSplit<L>::R2) k_fft_lines(
    float2* __restrict__ data, int64_t batch_stride, int nlines, int inner,
    const float2* __restrict__ tw) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2;
  constexpr int S = L + 1;  // LDS stride per line (odd: spreads banks)
  __shared__ float2 buf[kLines * S];
  const int c = threadIdx.x % kLines, n2 = threadIdx.x / kLines;
  const int q = blockIdx.x * kLines + c;
  const bool live = q < nlines;
  const Twiddles<L> w = load_twiddles<L>(n2, tw);
  float2* base = data + blockIdx.y * batch_stride;
  if (live) base += (int64_t)(q / inner) * L * inner + (q % inner);
  float2 v[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1)
    v[n1] = live ? base[(int64_t)(n2 + R2 * n1) * inner] : make_float2(0.f, 0.f);
  fft_four_step<DIR, L>(v, n2, buf + c * S, w, [&](int k, float2 x) {
    if (live) base[(int64_t)k * inner] = x;
  });
}

}  // namespace phz
