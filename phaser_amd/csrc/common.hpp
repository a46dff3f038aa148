// common.hpp -- shared device helpers for the phaser_amd HIP kernels (gfx950).
//
// * LDS Stockham FFT (radix 4/2/3/5, natural-order output, Govindaraju et al.
//   2008 formulation) used by every FFT stage;
// * order-preserving 64-bit argmax keys (first-maximum semantics of
//   std::max_element, z-score-eval.cc:46) for atomicMax reductions;
// * status/error plumbing for the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace phz {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);

#define PHZ_HIP(expr)                                                        \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::phz::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));   \
      return -2;                                                             \
    }                                                                        \
  } while (0)

#define PHZ_CHECK(cond, msg)                                                 \
  do {                                                                       \
    if (!(cond)) {                                                           \
      ::phz::set_error(msg);                                                 \
      return -1;                                                             \
    }                                                                        \
  } while (0)

// ---------------------------------------------------------------- complex
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) {
  return make_float2(a.x * s, a.y * s);
}

// ---------------------------------------------------------------- wave runs
// Runs of equal keys in adjacent lanes of a 64-lane wave (every lane must
// call; a negative key marks a lane that is its own run).  head_lane is the
// run's first lane, len its length; lets one lane per run do the memory
// atomic for the whole run.
struct WaveRun {
  int head_lane, len;
  bool head;
};
__device__ __forceinline__ WaveRun wave_run(long long key) {
  const int lane = threadIdx.x & 63;
  const long long prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key || key < 0;
  const unsigned long long heads = __ballot(head);
  const unsigned long long le = (2ull << lane) - 1ull;  // lanes <= this one
  const int hl = 63 - __clzll((long long)(heads & le));
  const unsigned long long above = heads & ~le;
  const int next = above ? __ffsll((long long)above) - 1 : 64;
  return WaveRun{hl, next - hl, head};
}

// Maximum of an unsigned value over the 64 lanes of a wave (every lane must
// call), by DPP row operations instead of LDS-routed shuffles: within quads
// ([1,0,3,2], [2,3,0,1]), half rows (row_half_mirror), rows (row_mirror),
// then row 1 and row 3 take the last lane of the row below (row_bcast:15)
// and rows 2 and 3 that of row 1 (row_bcast:31), so that lane 63 holds the
// maximum; lanes a step does not write keep their value (max is idempotent).
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ unsigned dpp_u32(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, ROWS, 0xF, false);
}
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  v = max(v, dpp_u32<0xB1>(v));
  v = max(v, dpp_u32<0x4E>(v));
  v = max(v, dpp_u32<0x141>(v));
  v = max(v, dpp_u32<0x140>(v));
  v = max(v, dpp_u32<0x142, 0xA>(v));
  v = max(v, dpp_u32<0x143, 0xC>(v));
  return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// ---------------------------------------------------------------- argmax keys
// key = (ordered value bits << 32) | (0xffffffff - index): atomicMax keeps the
// largest value and, among equal values, the smallest index.
__device__ __forceinline__ uint32_t ordered_bits(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float unordered_value(uint32_t o) {
  const uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
#ifdef __HIP_DEVICE_COMPILE__
  return __uint_as_float(u);
#else
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
#endif
}
__device__ __forceinline__ unsigned long long make_key(float v, uint32_t idx) {
  return ((unsigned long long)ordered_bits(v) << 32) | (unsigned long long)(0xffffffffu - idx);
}
__host__ __device__ inline uint32_t key_index(unsigned long long k) {
  return 0xffffffffu - (uint32_t)(k & 0xffffffffull);
}

// wave-level max of keys, then one atomic per wave
__device__ __forceinline__ unsigned long long wave_max_key(unsigned long long k) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(k, off, 64);
    k = o > k ? o : k;
  }
  return k;
}

// workgroup-level max of keys (blockDim a multiple of 64, <= 1024); the
// result is valid in thread 0.  Contains barriers: call from every thread.
__device__ __forceinline__ unsigned long long block_max_key(unsigned long long k) {
  __shared__ unsigned long long wk[16];
  k = wave_max_key(k);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) wk[w] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int i = 1; i < nw; ++i) k = wk[i] > k ? wk[i] : k;
  }
  __syncthreads();
  return k;
}

// ---------------------------------------------------------------- FFT plan
struct FftPlan {
  int L;          // transform length
  int nst;        // stages
  int radix[12];  // radices in stage order: 4, 2, 3, 5, then other primes
};

// host: factor L into radix-4, 2, 3 and 5 stages followed by one stage per
// other prime factor (fft_stage_prime); nst = -1 if L < 1 or L needs more
// than 12 stages (never for L <= 4096)
FftPlan make_plan(int L);
// host: true when L factors into 2, 3 and 5 only (the register engines and
// the radix-specialised stages)
bool smooth235(int L);

// One Stockham stage of radix R for nb lines of length L stored line-major
// in `in` (line stride ls >= L; ls = L + 1 keeps transposed accesses free of
// LDS bank conflicts), written to `out`.  tw[e] = exp(-2 pi i e / L).
// DIR = -1 forward (e^{-i}), +1 backward.
template <int R, int DIR>
__device__ __forceinline__ void fft_stage(const float2* __restrict__ in, float2* __restrict__ out,
                                          int L, int ls, int nb, int Ns,
                                          const float2* __restrict__ tw) {
  const int Lr = L / R;
  const int tstep = L / (Ns * R);
  for (int w = threadIdx.x; w < nb * Lr; w += blockDim.x) {
    const int line = w / Lr;
    const int j = w - line * Lr;
    const float2* src = in + line * ls;
    float2* dst = out + line * ls;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[j + r * Lr];
    const int k = j % Ns;
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 t = tw[(k * r * tstep) % L];
        if (DIR > 0) t.y = -t.y;
        v[r] = cmul(v[r], t);
      }
    }
    if (R == 2) {
      const float2 a = v[0], b = v[1];
      v[0] = cadd(a, b);
      v[1] = csub(a, b);
    } else if (R == 4) {
      const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
      const float2 a2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
      // forward: -i*d ; backward: +i*d
      const float2 a3 = DIR < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
      v[0] = cadd(a0, a2);
      v[2] = csub(a0, a2);
      v[1] = cadd(a1, a3);
      v[3] = csub(a1, a3);
    } else if (R == 3) {
      const float c = -0.5f, s = DIR * 0.86602540378443865f;
      const float2 a = v[0], b = v[1], cc = v[2];
      const float2 t1 = cadd(b, cc), t2 = csub(b, cc);
      v[0] = cadd(a, t1);
      const float2 m1 = make_float2(a.x + c * t1.x, a.y + c * t1.y);
      const float2 m2 = make_float2(-s * t2.y, s * t2.x);  // i*s*t2
      v[1] = cadd(m1, m2);
      v[2] = csub(m1, m2);
    } else if (R == 5) {
      const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
      const float s1 = DIR * 0.95105651629515357f, s2 = DIR * 0.58778525229247313f;
      const float2 a = v[0];
      const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
      const float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
      v[0] = cadd(a, cadd(t1, t2));
      const float2 b1 = make_float2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
      const float2 b2 = make_float2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
      const float2 e1 = make_float2(-(s1 * t3.y + s2 * t4.y), s1 * t3.x + s2 * t4.x);
      const float2 e2 = make_float2(-(s2 * t3.y - s1 * t4.y), s2 * t3.x - s1 * t4.x);
      v[1] = cadd(b1, e1);
      v[4] = csub(b1, e1);
      v[2] = cadd(b2, e2);
      v[3] = csub(b2, e2);
    }
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
  }
}

// One Stockham stage of a prime radix R > 5 (lengths with other factors
// than 2, 3 and 5, e.g. 201 = 3 * 67): one thread per output of a butterfly,
// the length-R DFT summed directly.  With W = exp(-2 pi i / L) (conjugated
// for DIR = +1), butterfly j (k = j mod Ns) has inputs x_r = in[j + r L/R],
// and output q is sum_r x_r W^(r (k tstep + q L/R)): the stage twiddle
// W^(r k tstep) and the DFT factor W_R^(q r) = W^(q r L/R) in one lookup.
// Same data layout and output order as fft_stage.
template <int DIR>
__device__ __forceinline__ void fft_stage_prime(const float2* __restrict__ in,
                                                float2* __restrict__ out, int L, int ls, int nb,
                                                int Ns, int R, const float2* __restrict__ tw) {
  const int Lr = L / R;
  const int tstep = L / (Ns * R);
  for (int w = threadIdx.x; w < nb * L; w += blockDim.x) {
    const int line = w / L;
    const int rem = w - line * L;  // q * Lr + j
    const int q = rem / Lr, j = rem - q * Lr;
    const float2* src = in + line * ls;
    const int k = j % Ns;
    const int e = k * tstep + q * Lr;  // < L/R + (R - 1) L/R = L
    float2 acc = src[j];
    int ex = 0;
    for (int r = 1; r < R; ++r) {
      ex += e;
      if (ex >= L) ex -= L;
      float2 t = tw[ex];
      if (DIR > 0) t.y = -t.y;
      acc = cadd(acc, cmul(src[j + r * Lr], t));
    }
    out[line * ls + (j / Ns) * Ns * R + k + q * Ns] = acc;
  }
}

// nb FFTs of length plan.L on line-major LDS data in `a` (line stride ls,
// scratch `b` of the same size).  Returns the buffer holding the result.
// Ends with a barrier.
template <int DIR>
__device__ float2* lds_fft(float2* a, float2* b, int nb, const FftPlan& plan,
                           const float2* __restrict__ tw, int ls = 0) {
  if (ls == 0) ls = plan.L;
  int Ns = 1;
  for (int s = 0; s < plan.nst; ++s) {
    const int R = plan.radix[s];
    if (R == 4)
      fft_stage<4, DIR>(a, b, plan.L, ls, nb, Ns, tw);
    else if (R == 2)
      fft_stage<2, DIR>(a, b, plan.L, ls, nb, Ns, tw);
    else if (R == 3)
      fft_stage<3, DIR>(a, b, plan.L, ls, nb, Ns, tw);
    else if (R == 5)
      fft_stage<5, DIR>(a, b, plan.L, ls, nb, Ns, tw);
    else
      fft_stage_prime<DIR>(a, b, plan.L, ls, nb, Ns, R, tw);
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

}  // namespace phz
