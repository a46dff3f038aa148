// kernels_mfma.hip -- the two table contractions of the path on gfx950
// matrix cores, in split FP16 with FP32 accumulation:
//   * the Legendre contraction of the S^2 transform (s2_semi_memo.c:123-375):
//     coef[s][l,m] = sum_i (w_i P~_l^m(cos th_i)) F[s][m][i];
//   * the Wigner-d synthesis of the inverse SO(3) transform
//     (wignerTransforms_fftw.c:407-668): S(m1,m2,beta_k) = sum_l c^l d~^l(beta_k).
// Both are batched GEMMs: one table block against every signal (or pair) of
// a batch, so each table block is read once per batch.
//
// Split FP16: x = hi + lo with hi = fp16(x), lo = fp16(x - hi); a product
// keeps hi*hi + hi*lo + lo*hi (the lo*lo term is below FP32 rounding of the
// sum).  The coefficient rows are first scaled by a power of two so that
// their largest magnitude is <= 1 (FP16 range), and the scale is undone on
// the FP32 result.  Relative to the row maximum this keeps ~22 bits, close
// to FP32; the error against the FP64 oracle is stated by the tests.
//
// MFMA used: v_mfma_f32_32x32x16_f16 (wave64; A 32x16, B 16x32, D 32x32).
// Lane l holds A[l%32][8*(l/32) .. +7] and B[8*(l/32) .. +7][l%32]; D
// register r of lane l is D[8*(r/4) + 4*(l/32) + r%4][l%32].
#include "kernels.hpp"

namespace phz {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kKPad = 8;  // LDS row padding (halves) against bank conflicts
constexpr size_t kMfmaLdsMax = 160 * 1024;

__device__ __forceinline__ void split_store(float x, _Float16* hi, _Float16* lo) {
  const _Float16 h = (_Float16)x;
  *hi = h;
  *lo = (_Float16)(x - (float)h);
}

// power-of-two scale that brings a row maximum m into (0.5, 1]
__device__ __forceinline__ float row_scale(float m) {
  if (!(m > 0.f) || !isfinite(m)) return 1.f;
  int e;
  frexpf(m, &e);  // m = f * 2^e, f in [0.5, 1)
  return ldexpf(1.f, -e);
}

// C[M x N] = A[M x K] B[K x N]; A rows and B columns are K-contiguous in LDS
// (leading dimensions lda, ldb in halves); M, N multiples of 32, K of 16.
// store(i, j, value) receives every element once.
template <class Store>
__device__ void split_gemm(const _Float16* __restrict__ Ahi, const _Float16* __restrict__ Alo,
                           int lda, const _Float16* __restrict__ Bhi,
                           const _Float16* __restrict__ Blo, int ldb, int M, int N, int K,
                           Store store) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int tn = N / 32, tiles = (M / 32) * tn;
  const int r32 = lane & 31, kh = 8 * (lane >> 5);
  for (int tile = wave; tile < tiles; tile += nwaves) {
    const int i0 = (tile / tn) * 32, j0 = (tile % tn) * 32;
    floatx16 acc = {};
    const _Float16* ah = Ahi + (i0 + r32) * lda + kh;
    const _Float16* al = Alo + (i0 + r32) * lda + kh;
    const _Float16* bh = Bhi + (j0 + r32) * ldb + kh;
    const _Float16* bl = Blo + (j0 + r32) * ldb + kh;
    for (int k0 = 0; k0 < K; k0 += 16) {
      const half8 a_h = *(const half8*)(ah + k0);
      const half8 a_l = *(const half8*)(al + k0);
      const half8 b_h = *(const half8*)(bh + k0);
      const half8 b_l = *(const half8*)(bl + k0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_l, b_h, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_h, b_l, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_h, b_h, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r)
      store(i0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), j0 + r32, acc[r]);
  }
}

__device__ __forceinline__ int round_up(int x, int m) { return (x + m - 1) / m * m; }

__device__ __forceinline__ int sean_idx(int m, int l, int bw) {
  const int L = bw - 1;
  if (m >= 0) return m * (L + 1) - ((m * (m - 1)) / 2) + (l - m);
  return ((L * (L + 3)) / 2) + 1 + ((L + m) * (L + m + 1) / 2) + (l + m);
}

// c^l_{m1,m2} = wigNorm (-1)^{m1+m2} f_{l,-m1} conj(h_{l,-m2})
// (so3_correlate_fftw.c:120-157)
__device__ __forceinline__ float2 combine_c(const float2* f, const float2* h, float wn, int l,
                                            int m1, int m2, int bw) {
  const float2 s = f[sean_idx(-m1, l, bw)];
  const float2 p = h[sean_idx(-m2, l, bw)];
  const float fudge = ((m1 + m2) & 1) ? -1.f : 1.f;
  const float pr = fudge * p.x, pi = fudge * (-p.y);
  return make_float2(wn * (s.x * pr - s.y * pi), wn * (s.x * pi + s.y * pr));
}

}  // namespace

// ------------------------------------------------------------ Legendre
// One workgroup per order m, chunk of `sc` signals and chunk of `dc`
// degrees.  A rows: (signal, re/im) of F[s][m][.] over theta_i (K = n);
// B columns: degrees l = m + t0 .. of the weighted table.
struct LegChunks {
  int sig, deg;
};

static size_t leg_lds(int bw, LegChunks c) {
  const int ld = ((2 * bw + 15) / 16) * 16 + kKPad;
  return (size_t)2 * (2 * c.sig + c.deg) * ld * sizeof(_Float16) + 2 * c.sig * sizeof(float);
}

static LegChunks leg_chunks(int bw) {
  const int full = ((bw + 31) / 32) * 32;
  const int degs[3] = {full, 64, 32};
  const int sigs[3] = {64, 32, 16};
  for (int d : degs)
    for (int sg : sigs)
      if (d <= full && leg_lds(bw, LegChunks{sg, d}) <= kMfmaLdsMax) return LegChunks{sg, d};
  return LegChunks{16, 32};
}

__global__ void __launch_bounds__(256) k_legendre_mfma(const float2* __restrict__ F, int S,
                                                       int n, int bw, LegChunks ch,
                                                       const float* __restrict__ leg,
                                                       const int* __restrict__ leg_off,
                                                       float2* __restrict__ coef) {
  const int m = blockIdx.x, s0 = blockIdx.y * ch.sig, t0 = blockIdx.z * ch.deg;
  const int L = bw - m;
  if (t0 >= L) return;  // the whole workgroup leaves before any barrier
  const int ns = min(ch.sig, S - s0), nt = min(ch.deg, L - t0);
  const int Kp = round_up(n, 16), ld = Kp + kKPad;
  const int Mp = round_up(2 * ns, 32), Np = round_up(nt, 32);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  _Float16* Ahi = (_Float16*)lds_raw;
  _Float16* Alo = Ahi + (size_t)2 * ch.sig * ld;
  _Float16* Bhi = Alo + (size_t)2 * ch.sig * ld;
  _Float16* Blo = Bhi + (size_t)ch.deg * ld;
  float* scale = (float*)(Blo + (size_t)ch.deg * ld);
  // row scales: one wave per row, max |re| or |im| over theta
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  for (int r = wave; r < Mp; r += nwaves) {
    float mx = 0.f;
    if (r < 2 * ns) {
      const float2* src = F + ((int64_t)(s0 + (r >> 1)) * bw + m) * n;
      for (int i = lane; i < n; i += 64) mx = fmaxf(mx, fabsf((r & 1) ? src[i].y : src[i].x));
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (lane == 0) scale[r] = row_scale(mx);
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < Mp * Kp; idx += blockDim.x) {
    const int r = idx / Kp, i = idx - r * Kp;
    float v = 0.f;
    if (r < 2 * ns && i < n) {
      const float2 x = F[((int64_t)(s0 + (r >> 1)) * bw + m) * n + i];
      v = ((r & 1) ? x.y : x.x) * scale[r];
    }
    split_store(v, &Ahi[r * ld + i], &Alo[r * ld + i]);
  }
  const float* T = leg + leg_off[m];  // [i][t], t = l - m
  for (int idx = threadIdx.x; idx < Np * Kp; idx += blockDim.x) {
    const int i = idx / Np, t = idx - i * Np;
    const float v = (i < n && t < nt) ? T[(int64_t)i * L + t0 + t] : 0.f;
    split_store(v, &Bhi[t * ld + i], &Blo[t * ld + i]);
  }
  __syncthreads();
  const int L1 = bw - 1;
  const float sg = (m & 1) ? -1.f : 1.f;
  split_gemm(Ahi, Alo, ld, Bhi, Blo, ld, Mp, Np, Kp, [&](int r, int t, float v) {
    if (r >= 2 * ns || t >= nt) return;
    const int s = s0 + (r >> 1), l = m + t0 + t;
    float* out = (float*)(coef + (int64_t)s * bw * bw);
    const float x = v / scale[r];
    const int ip = m * (L1 + 1) - ((m * (m - 1)) / 2) + (l - m);
    out[2 * ip + (r & 1)] = x;
    if (m > 0) {  // f_{l,-m} = (-1)^m conj f_{l,m} (s2_semi_memo.c:349-373)
      const int im = ((L1 * (L1 + 3)) / 2) + 1 + ((L1 - m) * (L1 - m + 1) / 2) + (l - m);
      out[2 * im + (r & 1)] = (r & 1) ? -sg * x : sg * x;
    }
  });
}

void launch_legendre_mfma(const float2* F, int S, int bw, const float* legendre,
                          const int* leg_off, float2* coef, hipStream_t s) {
  const int n = 2 * bw;
  const LegChunks ch = leg_chunks(bw);
  dim3 g(bw, (S + ch.sig - 1) / ch.sig, (bw + ch.deg - 1) / ch.deg);
  hipLaunchKernelGGL(k_legendre_mfma, g, dim3(256), leg_lds(bw, ch), s, F, S, n, bw, ch,
                     legendre, leg_off, coef);
}

// ------------------------------------------------------------ Wigner-d
// One workgroup per table block (a <= b), chunk of `pairs` pairs and chunk
// of `cols` beta columns.  A rows: (pair, case, re/im) combined coefficients
// over l (K = bw - b); B columns: beta_k.  Cases as in k_so3_synth: 0 (a,b);
// 1 (b,0) or (b,a); 2 (-a,a) or (a,-b) at pi - beta; 3 (b,-a) at pi - beta.
struct SynChunks {
  int pairs, cols;
};

// LDS of one workgroup: split A (8 rows per pair) and B (one row per beta
// column) at the largest block's leading dimension, the combined
// coefficients (4 cases per pair, FP32 complex) and the row scales
static size_t syn_lds(int bw, SynChunks c) {
  const int kp = ((bw + 15) / 16) * 16, ld = kp + kKPad;
  return (size_t)2 * (8 * c.pairs + c.cols) * ld * sizeof(_Float16) +
         (size_t)4 * c.pairs * kp * sizeof(float2) + 8 * c.pairs * sizeof(float);
}

// the largest chunk whose LDS leaves room for two workgroups per CU (a
// workgroup's phases are separated by barriers: a second one hides them)
static SynChunks syn_chunks(int bw) {
  const int n = 2 * bw, full = ((n + 31) / 32) * 32;
  const int cols[4] = {full, 128, 64, 32};
  const int pairs[3] = {16, 8, 4};
  SynChunks best{4, 32};
  for (int cc : cols)
    for (int p : pairs)
      if (cc <= full && syn_lds(bw, SynChunks{p, cc}) <= kMfmaLdsMax / 3 &&
          p * cc > best.pairs * best.cols)
        best = SynChunks{p, cc};
  return best;
}

// combined coefficient of case c for degree l (zero where the case is unused)
__device__ __forceinline__ float2 synth_coef(const float2* f, const float2* h, float wn, int l,
                                             int a, int b, int cs, int c, int bw) {
  float2 v = make_float2(0.f, 0.f);
  if (c == 0) {
    v = combine_c(f, h, wn, l, a, b, bw);
  } else if (c == 1 && cs == 1) {
    v = combine_c(f, h, wn, l, b, 0, bw);
    if (b & 1) v = make_float2(-v.x, -v.y);
  } else if (c == 1 && cs == 3) {
    v = combine_c(f, h, wn, l, b, a, bw);
    if ((b - a) & 1) v = make_float2(-v.x, -v.y);
  } else if (c == 2 && cs == 2) {
    v = combine_c(f, h, wn, l, -a, a, bw);
    const float sg = ((l - a) & 1) ? -1.f : 1.f;
    v = make_float2(sg * v.x, sg * v.y);
  } else if (c == 2 && cs == 3) {
    v = combine_c(f, h, wn, l, a, -b, bw);
    const float sg = ((l + a) & 1) ? -1.f : 1.f;
    v = make_float2(sg * v.x, sg * v.y);
  } else if (c == 3 && cs == 3) {
    v = combine_c(f, h, wn, l, b, -a, bw);
    const float sg = ((l + a) & 1) ? -1.f : 1.f;
    v = make_float2(sg * v.x, sg * v.y);
  }
  return v;
}

// The Wigner table split once per context: per block (a <= b), the hi
// halves of d~^l_{a,b}(beta_k) as rows k of ld = round16(bw - b) + kKPad
// halves (degrees t = l - b along the row, zero past bw - b), then the lo
// halves likewise: the layout of the GEMM's B operand in LDS.
__global__ void k_wig_split(int bw, const float* __restrict__ wig,
                            const int64_t* __restrict__ wig_off,
                            const int64_t* __restrict__ hl_off, _Float16* __restrict__ hl) {
  const int n = 2 * bw, blk = blockIdx.x;
  int b = (int)((sqrtf(8.f * blk + 1.f) - 1.f) * 0.5f);
  while ((b + 1) * (b + 2) / 2 <= blk) ++b;
  while (b * (b + 1) / 2 > blk) --b;
  const int L = bw - b, ld = round_up(L, 16) + kKPad;
  const float* T = wig + wig_off[blk];
  _Float16* hi = hl + hl_off[blk];
  _Float16* lo = hi + (int64_t)n * ld;
  for (int idx = threadIdx.x; idx < n * ld; idx += blockDim.x) {
    const int k = idx / ld, t = idx - k * ld;
    split_store(t < L ? T[(int64_t)t * n + k] : 0.f, &hi[idx], &lo[idx]);
  }
}

// One workgroup per table block (a <= b), chunk of `pairs` pairs and chunk
// of `cols` beta columns.  A rows: (pair, case, re/im) of the combined
// coefficients over l (each computed once, kept in LDS for the row scales
// and the split); B: the block's pre-split table rows, copied as they lie.
// The GEMM's re/im rows sit in neighbouring accumulator registers, so each
// lane stores whole complex values: consecutive beta in consecutive lanes.
__global__ void __launch_bounds__(256) k_so3_synth_mfma(
    int bw, const float2* __restrict__ fc, int64_t f_stride, const float2* __restrict__ hc,
    int64_t h_stride, int npairs, SynChunks ch, const _Float16* __restrict__ hl,
    const int64_t* __restrict__ hl_off, const float* __restrict__ wnorm,
    float2* __restrict__ S) {
  const int n = 2 * bw, w1 = 2 * bw - 1;
  const int blk = blockIdx.x, p0 = blockIdx.y * ch.pairs, k0 = blockIdx.z * ch.cols;
  const int np = min(ch.pairs, npairs - p0), nk = min(ch.cols, n - k0);
  int b = (int)((sqrtf(8.f * blk + 1.f) - 1.f) * 0.5f);
  while ((b + 1) * (b + 2) / 2 <= blk) ++b;
  while (b * (b + 1) / 2 > blk) --b;
  const int a = blk - b * (b + 1) / 2;
  const int L = bw - b;
  const int cs = (b == 0) ? 0 : (a == 0 ? 1 : (a == b ? 2 : 3));
  const int Kp = round_up(L, 16), ld = Kp + kKPad;
  const int kpw = round_up(bw, 16), ldw = kpw + kKPad;  // the LDS carve-up's sizes
  const int Mp = round_up(np * 8, 32), Np = round_up(nk, 32);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  _Float16* Ahi = (_Float16*)lds_raw;
  _Float16* Alo = Ahi + (size_t)8 * ch.pairs * ldw;
  _Float16* Bhi = Alo + (size_t)8 * ch.pairs * ldw;
  _Float16* Blo = Bhi + (size_t)ch.cols * ldw;
  float2* cf = (float2*)(Blo + (size_t)ch.cols * ldw);  // [pair*4 + case][t], stride kpw
  float* scale = (float*)(cf + (size_t)4 * ch.pairs * kpw);
  int* scale_bits = (int*)scale;
  for (int r = threadIdx.x; r < Mp; r += blockDim.x) scale_bits[r] = 0;
  __syncthreads();
  // 1. the combined coefficients, once each, and the row maxima
  // (non-negative floats order as ints)
  for (int idx = threadIdx.x; idx < np * 4 * L; idx += blockDim.x) {
    const int row4 = idx / L, t = idx - row4 * L;
    const int p = row4 >> 2, c = row4 & 3, l = b + t;
    const float2 v = synth_coef(fc + (p0 + p) * f_stride, hc + (p0 + p) * h_stride, wnorm[l], l,
                                a, b, cs, c, bw);
    cf[row4 * kpw + t] = v;
    atomicMax(&scale_bits[2 * row4], __float_as_int(fabsf(v.x)));
    atomicMax(&scale_bits[2 * row4 + 1], __float_as_int(fabsf(v.y)));
  }
  __syncthreads();
  for (int r = threadIdx.x; r < Mp; r += blockDim.x)
    scale[r] = row_scale(__int_as_float(scale_bits[r]));
  __syncthreads();
  // 2. scaled split A from LDS
  for (int idx = threadIdx.x; idx < Mp / 2 * Kp; idx += blockDim.x) {
    const int row4 = idx / Kp, t = idx - row4 * Kp;
    const float2 v = (row4 < np * 4 && t < L) ? cf[row4 * kpw + t] : make_float2(0.f, 0.f);
    split_store(v.x * scale[2 * row4], &Ahi[(2 * row4) * ld + t], &Alo[(2 * row4) * ld + t]);
    split_store(v.y * scale[2 * row4 + 1], &Ahi[(2 * row4 + 1) * ld + t],
                &Alo[(2 * row4 + 1) * ld + t]);
  }
  // 3. the block's pre-split rows k0 .. k0 + Np (zero past the chunk), 16 bytes at a time
  {
    const _Float16* ghi = hl + hl_off[blk] + (int64_t)k0 * ld;
    const _Float16* glo = ghi + (int64_t)n * ld;
    const int vec = Np * ld / 8;
    for (int v = threadIdx.x; v < vec; v += blockDim.x) {
      const int kk = (v * 8) / ld;
      uint4 x = make_uint4(0u, 0u, 0u, 0u), y = x;
      if (kk < nk) {
        x = ((const uint4*)ghi)[v];
        y = ((const uint4*)glo)[v];
      }
      ((uint4*)Bhi)[v] = x;
      ((uint4*)Blo)[v] = y;
    }
  }
  __syncthreads();
  // 4. GEMM; rows (pair, case, re/im), column k -> S[pair][m2][m1][beta]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int tn = Np / 32, tiles = (Mp / 32) * tn;
  const int r32 = lane & 31, kh = 8 * (lane >> 5);
  for (int tile = wave; tile < tiles; tile += nwaves) {
    const int i0 = (tile / tn) * 32, j0 = (tile % tn) * 32;
    floatx16 acc = {};
    const _Float16* ah = Ahi + (i0 + r32) * ld + kh;
    const _Float16* al = Alo + (i0 + r32) * ld + kh;
    const _Float16* bh = Bhi + (j0 + r32) * ld + kh;
    const _Float16* bl = Blo + (j0 + r32) * ld + kh;
    for (int kq = 0; kq < Kp; kq += 16) {
      const half8 a_h = *(const half8*)(ah + kq);
      const half8 a_l = *(const half8*)(al + kq);
      const half8 b_h = *(const half8*)(bh + kq);
      const half8 b_l = *(const half8*)(bl + kq);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_l, b_h, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_h, b_l, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_h, b_h, acc, 0, 0, 0);
    }
    const int kc = j0 + r32;
    if (kc >= nk) continue;
    const int k = k0 + kc, kf = n - 1 - k;  // the beta-flipped cases read d(pi - beta)
#pragma unroll
    for (int r = 0; r < 16; r += 2) {
      // rows re = i0 + 8 (r / 4) + 4 (lane / 32) + r % 4 (even) and im = re + 1
      const int row = i0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (row >= np * 8) continue;
      const int c = (row >> 1) & 3, p = row >> 3;
      const float2 z = make_float2(acc[r] / scale[row], acc[r + 1] / scale[row + 1]);
      float2* Sp = S + (int64_t)(p0 + p) * bw * w1 * n;
      auto put = [&](int m1, int m2, int kk, bool conj) {
        Sp[((int64_t)m2 * w1 + (m1 + bw - 1)) * n + kk] = conj ? make_float2(z.x, -z.y) : z;
      };
      if (c == 0) {
        put(a, b, k, false);
      } else if (c == 1) {
        if (cs == 1) {
          put(b, 0, k, false);
          put(-b, 0, k, true);
        } else if (cs == 3) {
          put(b, a, k, false);
        }
      } else if (c == 2) {
        if (cs == 2) put(-a, a, kf, false);
        else if (cs == 3) put(-a, b, kf, true);
      } else if (cs == 3) {
        put(-b, a, kf, true);
      }
    }
  }
}

int64_t wig_split_offsets(int bw, std::vector<int64_t>* off) {
  const int n = 2 * bw, nblk = bw * (bw + 1) / 2;
  off->assign(nblk, 0);
  int64_t total = 0;
  for (int b = 0; b < bw; ++b)
    for (int a = 0; a <= b; ++a) {
      const int ld = ((bw - b + 15) / 16) * 16 + kKPad;
      (*off)[b * (b + 1) / 2 + a] = total;
      total += (int64_t)2 * n * ld;
    }
  return total;
}

void launch_wig_split(int bw, const float* wig, const int64_t* wig_off, const int64_t* hl_off,
                      uint16_t* hl, hipStream_t s) {
  hipLaunchKernelGGL(k_wig_split, dim3(bw * (bw + 1) / 2), dim3(256), 0, s, bw, wig, wig_off,
                     hl_off, (_Float16*)hl);
}

// ------------------------------------------------------------ MFMA peak
// Dense MFMA issue rate: every wave runs `iters` rounds of four independent
// accumulator chains of one matrix instruction on register data (no memory
// traffic) and stores one value per lane so that nothing is dead.
// DT 0: v_mfma_f32_32x32x16_f16, 1: v_mfma_f32_32x32x16_bf16,
// 2: v_mfma_f32_32x32x2_f32.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int DT>
__global__ void __launch_bounds__(256) k_mfma_peak(int iters, float* __restrict__ out) {
  floatx16 c0 = {}, c1 = {}, c2 = {}, c3 = {};
  const float x = 1e-3f * (float)(threadIdx.x & 63);
  if constexpr (DT == 0) {
    half8 a, b;
    for (int i = 0; i < 8; ++i) {
      a[i] = (_Float16)(x + i);
      b[i] = (_Float16)(x - i);
    }
    for (int it = 0; it < iters; ++it) {
      c0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c1, 0, 0, 0);
      c2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c2, 0, 0, 0);
      c3 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c3, 0, 0, 0);
    }
  } else if constexpr (DT == 1) {
    bf16x8 a, b;
    for (int i = 0; i < 8; ++i) {
      a[i] = (__bf16)(x + i);
      b[i] = (__bf16)(x - i);
    }
    for (int it = 0; it < iters; ++it) {
      c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c1, 0, 0, 0);
      c2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c2, 0, 0, 0);
      c3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c3, 0, 0, 0);
    }
  } else {
    const float a = x + 1.f, b = x - 1.f;
    for (int it = 0; it < iters; ++it) {
      c0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c1, 0, 0, 0);
      c2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c2, 0, 0, 0);
      c3 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c3, 0, 0, 0);
    }
  }
  float s = 0.f;
  for (int r = 0; r < 16; ++r) s += c0[r] + c1[r] + c2[r] + c3[r];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

double launch_mfma_peak(int dtype, int iters, int blocks, float* out, hipStream_t s) {
  switch (dtype) {
    case 1: hipLaunchKernelGGL(k_mfma_peak<1>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 2: hipLaunchKernelGGL(k_mfma_peak<2>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    default: hipLaunchKernelGGL(k_mfma_peak<0>, dim3(blocks), dim3(256), 0, s, iters, out); break;
  }
  // 2 M N K flops per instruction, four per round, four waves per block
  const double per = dtype == 2 ? 2.0 * 32 * 32 * 2 : 2.0 * 32 * 32 * 16;
  return per * 4.0 * iters * 4.0 * blocks;
}

void launch_so3_synth_mfma(int bw, const float2* f, int64_t f_stride, const float2* h,
                           int64_t h_stride, int npairs, const uint16_t* hl, const int64_t* hl_off,
                           const float* wig_norm, float2* S, hipStream_t s) {
  const int n = 2 * bw;
  const SynChunks ch = syn_chunks(bw);
  dim3 g(bw * (bw + 1) / 2, (npairs + ch.pairs - 1) / ch.pairs, (n + ch.cols - 1) / ch.cols);
  hipLaunchKernelGGL(k_so3_synth_mfma, g, dim3(256), syn_lds(bw, ch), s, bw, f, f_stride, h,
                     h_stride, npairs, ch, (const _Float16*)hl, hl_off, wig_norm, S);
}

}  // namespace phz
