// kernels_so3.hip -- SO(3) correlation for gfx950: coefficient combination
// fused into the Wigner-d synthesis, the 2-D inverse FFT over (m1, m2) ->
// (alpha, gamma) with |.|^2 and a first-maximum argmax, and the pose.
//
// Restates so3CombineCoef_fftw (so3_correlate_fftw.c:102-161) and
// Inverse_SO3_Naive_fftw with flag = 1 (soft_fftw.c:1036-1912): the
// reference synthesises S(m1,m2,beta) for one half of the (m1,m2) plane and
// fills the other half with conjugates, so the 2-D transform is Hermitian and
// its result real; here the half m2 >= 0 is synthesised (with the reference's
// own choice of which entries are direct and which are conjugate copies), the
// transform over m1 is a complex FFT and the one over m2 a real-output FFT.
#include "fft_reg.hpp"
#include "kernels.hpp"
#include "pose.hpp"

namespace phz {

namespace {

__device__ __forceinline__ int sean(int m, int l, int bw) {
  const int L = bw - 1;
  if (m >= 0) return m * (L + 1) - ((m * (m - 1)) / 2) + (l - m);
  return ((L * (L + 3)) / 2) + 1 + ((L + m) * (L + m + 1) / 2) + (l + m);
}

// c^l_{m1,m2} = wigNorm (-1)^{m1+m2} f_{l,-m1} conj(h_{l,-m2})
__device__ __forceinline__ float2 combine(const float2* f, const float2* h, float wn, int l,
                                          int m1, int m2, int bw) {
  const float2 s = f[sean(-m1, l, bw)];
  const float2 p = h[sean(-m2, l, bw)];
  const float fudge = ((m1 + m2) & 1) ? -1.f : 1.f;
  const float pr = fudge * p.x, pi = fudge * (-p.y);
  return make_float2(wn * (s.x * pr - s.y * pi), wn * (s.x * pi + s.y * pr));
}

// One workgroup per Wigner block (a <= b) and group of NP pairs; thread
// k = beta index.  The block's table column is read once and applied to the
// NP pairs (each pair's sums in the same order as alone, so the result does
// not depend on NP).  Columns: 0 -> (a,b) direct, 1 -> (b,a) direct, 2/3 ->
// beta-flipped (see the case table in the host launcher).
template <int NP>
__global__ void k_so3_synth(int bw, int npairs, const float2* __restrict__ fc, int64_t f_stride,
                            const float2* __restrict__ hc, int64_t h_stride,
                            const float* __restrict__ wig,
                            const int64_t* __restrict__ wig_off, const float* __restrict__ wnorm,
                            float2* __restrict__ S) {
  const int n = 2 * bw;
  const int blk = blockIdx.x, p0 = blockIdx.y * NP;
  int b = (int)((sqrtf(8.f * blk + 1.f) - 1.f) * 0.5f);
  while ((b + 1) * (b + 2) / 2 <= blk) ++b;
  while (b * (b + 1) / 2 > blk) --b;
  const int a = blk - b * (b + 1) / 2;
  const int L = bw - b;
  extern __shared__ float2 col[];  // NP x 4 x bw
  // case: 0 -> (0,0); 1 -> a=0<b; 2 -> 0<a=b; 3 -> 0<a<b
  const int cs = (b == 0) ? 0 : (a == 0 ? 1 : (a == b ? 2 : 3));
  for (int q = 0; q < NP; ++q) {
    const int pair = p0 + q < npairs ? p0 + q : npairs - 1;  // spare slots repeat the last
    const float2* f = fc + pair * f_stride;
    const float2* h = hc + pair * h_stride;
    float2* cq = col + q * 4 * bw;
    for (int t = threadIdx.x; t < L; t += blockDim.x) {
      const int l = b + t;
      const float wn = wnorm[l];
      float2 c0 = combine(f, h, wn, l, a, b, bw);
      float2 c1 = make_float2(0.f, 0.f), c2 = c1, c3 = c1;
      if (cs == 1) {
        c1 = combine(f, h, wn, l, b, 0, bw);
        if (b & 1) c1 = make_float2(-c1.x, -c1.y);
      } else if (cs == 2) {
        c2 = combine(f, h, wn, l, -a, a, bw);
        const float sg = ((l - a) & 1) ? -1.f : 1.f;
        c2 = make_float2(sg * c2.x, sg * c2.y);
      } else if (cs == 3) {
        c1 = combine(f, h, wn, l, b, a, bw);
        if ((b - a) & 1) c1 = make_float2(-c1.x, -c1.y);
        const float sg = ((l + a) & 1) ? -1.f : 1.f;
        c2 = combine(f, h, wn, l, a, -b, bw);
        c2 = make_float2(sg * c2.x, sg * c2.y);
        c3 = combine(f, h, wn, l, b, -a, bw);
        c3 = make_float2(sg * c3.x, sg * c3.y);
      }
      cq[t] = c0;
      cq[bw + t] = c1;
      cq[2 * bw + t] = c2;
      cq[3 * bw + t] = c3;
    }
  }
  __syncthreads();
  const int k = threadIdx.x;
  if (k >= n) return;
  const float* T = wig + wig_off[blk];
  float2 acc[NP][4];
#pragma unroll
  for (int q = 0; q < NP; ++q)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[q][j] = make_float2(0.f, 0.f);
  if (cs <= 1) {
    for (int t = 0; t < L; ++t) {
      const float d = T[(int64_t)t * n + k];
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const float2* cq = col + q * 4 * bw;
        acc[q][0].x += cq[t].x * d;
        acc[q][0].y += cq[t].y * d;
        acc[q][1].x += cq[bw + t].x * d;
        acc[q][1].y += cq[bw + t].y * d;
      }
    }
  } else if (cs == 2) {
    for (int t = 0; t < L; ++t) {
      const float d = T[(int64_t)t * n + k];
      const float df = T[(int64_t)t * n + (n - 1 - k)];
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const float2* cq = col + q * 4 * bw;
        acc[q][0].x += cq[t].x * d;
        acc[q][0].y += cq[t].y * d;
        acc[q][2].x += cq[2 * bw + t].x * df;
        acc[q][2].y += cq[2 * bw + t].y * df;
      }
    }
  } else {
    for (int t = 0; t < L; ++t) {
      const float d = T[(int64_t)t * n + k];
      const float df = T[(int64_t)t * n + (n - 1 - k)];
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const float2* cq = col + q * 4 * bw;
        acc[q][0].x += cq[t].x * d;
        acc[q][0].y += cq[t].y * d;
        acc[q][1].x += cq[bw + t].x * d;
        acc[q][1].y += cq[bw + t].y * d;
        acc[q][2].x += cq[2 * bw + t].x * df;
        acc[q][2].y += cq[2 * bw + t].y * df;
        acc[q][3].x += cq[3 * bw + t].x * df;
        acc[q][3].y += cq[3 * bw + t].y * df;
      }
    }
  }
  const int w1 = 2 * bw - 1;
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    if (p0 + q >= npairs) break;
    float2* Sp = S + (int64_t)(p0 + q) * bw * w1 * n;
    auto at = [&](int m1, int m2) -> float2* {
      return Sp + ((int64_t)m2 * w1 + (m1 + bw - 1)) * n + k;
    };
    *at(a, b) = acc[q][0];  // (a,b); also (0,0) and (a,a)
    if (cs == 1) {
      *at(b, 0) = acc[q][1];
      *at(-b, 0) = cconj(acc[q][1]);
    } else if (cs == 2) {
      *at(-a, a) = acc[q][2];
    } else if (cs == 3) {
      *at(b, a) = acc[q][1];
      *at(-a, b) = cconj(acc[q][2]);
      *at(-b, a) = cconj(acc[q][3]);
    }
  }
}

// FFT over m1 (M1 = m1 mod n, Nyquist row zero): U[k][m2][alpha]
template <int KT>
__global__ void k_so3_fft_alpha(int bw, const float2* __restrict__ S, FftPlan plan,
                                const float2* __restrict__ tw, float2* __restrict__ U) {
  const int n = 2 * bw, w1 = 2 * bw - 1;
  const int m2 = blockIdx.x, k0 = blockIdx.y * KT, pair = blockIdx.z;
  extern __shared__ float2 sm[];
  const int ls = n + 1;  // padded line stride for the transposed fill
  float2* A = sm;
  float2* B = sm + KT * ls;
  const float2* Sp = S + ((int64_t)pair * bw + m2) * w1 * n;
  for (int idx = threadIdx.x; idx < KT * ls; idx += blockDim.x) A[idx] = make_float2(0.f, 0.f);
  __syncthreads();
  for (int idx = threadIdx.x; idx < w1 * KT; idx += blockDim.x) {
    const int mi = idx / KT, kk = idx - mi * KT;
    const int m1 = mi - (bw - 1);
    const int M1 = m1 < 0 ? m1 + n : m1;
    if (k0 + kk < n) A[kk * ls + M1] = Sp[(int64_t)mi * n + k0 + kk];
  }
  __syncthreads();
  float2* R = lds_fft<-1>(A, B, KT, plan, tw, ls);
  float2* Up = U + (int64_t)pair * n * bw * n;
  for (int idx = threadIdx.x; idx < KT * n; idx += blockDim.x) {
    const int kk = idx / n, al = idx - kk * n;
    if (k0 + kk < n) Up[((int64_t)(k0 + kk) * bw + m2) * n + al] = R[kk * ls + al];
  }
}

// Hermitian FFT over m2 -> gamma for AT alpha values (two per complex
// line), scale (1/n)(bw/pi) (soft_fftw.c:1905-1911), |C|^2, argmax.
__global__ void k_so3_fft_gamma(int bw, const float2* __restrict__ U, FftPlan plan,
                                const float2* __restrict__ tw, float* __restrict__ corr,
                                int64_t corr_stride, unsigned long long* __restrict__ partials) {
  constexpr int AT = 16, NL = AT / 2;
  const int n = 2 * bw;
  const int a0 = blockIdx.x * AT, k = blockIdx.y, pair = blockIdx.z;
  extern __shared__ float2 sm[];
  const int xs = bw + 1;      // padded staging stride
  float2* X = sm;             // AT x (bw+1) staging
  float2* A = sm + AT * xs;   // NL x n
  float2* B = A + NL * n;
  const float2* Up = U + ((int64_t)pair * n + k) * bw * n;
  for (int idx = threadIdx.x; idx < bw * AT; idx += blockDim.x) {
    const int m2 = idx / AT, aa = idx - m2 * AT;
    X[aa * xs + m2] = (a0 + aa < n) ? Up[(int64_t)m2 * n + a0 + aa] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < NL * n; idx += blockDim.x) {
    const int p = idx / n, m = idx - p * n;
    const float2* xa = X + (2 * p) * xs;
    const float2* xb = X + (2 * p + 1) * xs;
    float2 va, vb;
    if (m == 0) {
      va = make_float2(xa[0].x, 0.f);
      vb = make_float2(xb[0].x, 0.f);
    } else if (m < bw) {
      va = xa[m];
      vb = xb[m];
    } else if (m == bw) {
      va = vb = make_float2(0.f, 0.f);
    } else {
      va = cconj(xa[n - m]);
      vb = cconj(xb[n - m]);
    }
    A[idx] = make_float2(va.x - vb.y, va.y + vb.x);  // va + i vb
  }
  __syncthreads();
  float2* R = lds_fft<-1>(A, B, NL, plan, tw);
  const float scale = (1.f / (float)n) * ((float)bw / 3.14159265358979323846f);
  unsigned long long best = 0ull;
  for (int idx = threadIdx.x; idx < AT * n; idx += blockDim.x) {
    const int aa = idx / n, g = idx - aa * n;
    if (a0 + aa >= n) continue;
    const float2 z = R[(aa >> 1) * n + g];
    const float c = ((aa & 1) ? z.y : z.x) * scale;
    const float v = c * c;
    const uint32_t gi = (uint32_t)(((int64_t)k * n + (a0 + aa)) * n + g);
    if (corr) corr[pair * corr_stride + gi] = v;
    const unsigned long long key = make_key(v, gi);
    best = key > best ? key : best;
  }
  best = block_max_key(best);
  if (threadIdx.x == 0) partials[(int64_t)pair * gridDim.x * gridDim.y + blockIdx.y * gridDim.x +
                                 blockIdx.x] = best;
}

// Register four-step forms (n = 64, 128, 256) of the two passes above.
// alpha: one line per beta index k (LINES consecutive k per workgroup), the
// result transposed through LDS so that U rows are written contiguously.
template <int L>
__global__ void __launch_bounds__(256) k_so3_fft_alpha_reg(int bw, const float2* __restrict__ S,
                                                           const float2* __restrict__ tw,
                                                           float2* __restrict__ U) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2, LINES = 256 / R2, SS = L + 1;
  __shared__ float2 buf[LINES * SS];
  const int w1 = 2 * bw - 1;
  const int m2 = blockIdx.x, k0 = blockIdx.y * LINES, pair = blockIdx.z;
  const int c = threadIdx.x % LINES, n2 = threadIdx.x / LINES;
  const float2* Sp = S + ((int64_t)pair * bw + m2) * w1 * L + k0 + c;
  const Twiddles<L> w = load_twiddles<L>(n2, tw);
  // every row loads unconditionally (the Nyquist row m1 = -bw reads row
  // 0 and is zeroed afterwards), so the R1 loads are in flight together
  float2 v[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1) {
    const int M1 = n2 + R2 * n1;  // m1 mod n
    const int m1 = M1 < bw ? M1 : M1 - L;
    v[n1] = Sp[(int64_t)(M1 == bw ? 0 : m1 + bw - 1) * L];
  }
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1)
    if (n2 + R2 * n1 == bw) v[n1] = make_float2(0.f, 0.f);
  float2 out[R1];
  fft_four_step_reg<-1, L>(v, n2, buf + c * SS, w, out);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) buf[c * SS + (n2 + R2 * j) + R1 * k2] = out[j * R2 + k2];
  __syncthreads();
  float2* Up = U + (int64_t)pair * L * bw * L;
  for (int idx = threadIdx.x; idx < LINES * L; idx += blockDim.x) {
    const int cc = idx / L, al = idx - cc * L;
    Up[((int64_t)(k0 + cc) * bw + m2) * L + al] = buf[cc * SS + al];
  }
}

// gamma: one line per pair of alpha values (Hermitian packing), |C|^2 and
// the argmax from registers, the grid transposed through LDS for the write.
template <int L>
__global__ void __launch_bounds__(256) k_so3_fft_gamma_reg(int bw, const float2* __restrict__ U,
                                                           const float2* __restrict__ tw,
                                                           float* __restrict__ corr,
                                                           int64_t corr_stride,
                                                           unsigned long long* __restrict__ partials) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2, LINES = 256 / R2, SS = L + 1;
  __shared__ float2 buf[LINES * SS];
  const int a0 = blockIdx.x * 2 * LINES, k = blockIdx.y, pair = blockIdx.z;
  const int c = threadIdx.x % LINES, n2 = threadIdx.x / LINES;
  const int aa = a0 + 2 * c;  // alpha values aa and aa + 1
  const float2* Up = U + ((int64_t)pair * L + k) * bw * L + aa;
  const Twiddles<L> w = load_twiddles<L>(n2, tw);
  // every row loads unconditionally (the Nyquist row m = bw reads row
  // bw - 1 and is zeroed afterwards), so the R1 loads are in flight together
  float4 ab[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1) {
    const int m = n2 + R2 * n1;
    const int mm = m < bw ? m : L - m;
    ab[n1] = *(const float4*)(Up + (int64_t)(mm < bw ? mm : bw - 1) * L);
  }
  float2 v[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1) {
    const int m = n2 + R2 * n1;
    float2 va = make_float2(ab[n1].x, ab[n1].y), vb = make_float2(ab[n1].z, ab[n1].w);
    // m = 0: real parts only; m > bw: the conjugate half; m = bw: zero
    const float s = m > bw ? -1.f : (m == 0 ? 0.f : 1.f);
    va.y *= s;
    vb.y *= s;
    if (m == bw) va = vb = make_float2(0.f, 0.f);
    v[n1] = make_float2(va.x - vb.y, va.y + vb.x);  // va + i vb
  }
  float2 out[R1];
  fft_four_step_reg<-1, L>(v, n2, buf + c * SS, w, out);
  const float scale = (1.f / (float)L) * ((float)bw / 3.14159265358979323846f);
  unsigned long long best = 0ull;
  __syncthreads();
  float* g2 = (float*)buf;  // 2 * LINES rows of L values, stride SS
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      const int g = (n2 + R2 * j) + R1 * k2;
      const float2 z = out[j * R2 + k2];
      const float ca = z.x * scale, cb = z.y * scale;
      const float va = ca * ca, vb = cb * cb;
      const uint32_t ga = (uint32_t)(((int64_t)k * L + aa) * L + g);
      const unsigned long long ka = make_key(va, ga), kb = make_key(vb, ga + (uint32_t)L);
      best = ka > best ? ka : best;
      best = kb > best ? kb : best;
      g2[(2 * c) * SS + g] = va;
      g2[(2 * c + 1) * SS + g] = vb;
    }
  best = block_max_key(best);  // contains a barrier
  if (threadIdx.x == 0)
    partials[(int64_t)pair * gridDim.x * gridDim.y + blockIdx.y * gridDim.x + blockIdx.x] = best;
  if (corr) {
    float* cp = corr + pair * corr_stride + ((int64_t)k * L + a0) * L;
    for (int idx = threadIdx.x; idx < 2 * LINES * L; idx += blockDim.x) {
      const int r = idx / L, g = idx - r * L;
      cp[(int64_t)r * L + g] = g2[r * SS + g];
    }
  }
}

// The rotation pose of pair blockIdx.x (one wavefront): the argmax's ZYZ
// angles (rotation-utils.cc:190-207) and the Bingham fit to the
// +-k linear-index window around it (BinghamPeakBasedEval,
// bingham-peak-based-eval.cc:22-90; Bingham::fit / fitToMoment,
// bingham.cc:239-296): weights = the grid values, normalised to sum 1 when
// the window holds more than one index; moment C = sum w q q^T; mode = the
// eigenvector of C's largest eigenvalue (the maximum-likelihood
// concentrations order like the eigenvalues).  Then the XYZ Euler angles
// and float transform of the mode (rotation-utils.cc:17-23,86-143).
__global__ void __launch_bounds__(64) k_rot_pose(int bw, const unsigned long long* __restrict__ keys,
                                                 const float* __restrict__ corr,
                                                 int64_t corr_stride, int k,
                                                 DevPose* __restrict__ pose, int skip_axes,
                                                 int mixture) {
  const int p = blockIdx.x, lane = threadIdx.x;
  const unsigned long long key = keys[p];
  const uint32_t idx = key_index(key);
  const uint32_t n = 2 * bw, n_corr = n * n * n;
  double zyz[3], qp[4];
  zyz_of_index(idx, bw, zyz);
  quat_of_zyz(zyz, qp);
  // bingham-peak-based-eval.cc:47-58; an end of n_corr fails CHECK_LT (:69)
  const uint32_t uk = (uint32_t)k;
  const uint32_t start = uk > idx ? idx : idx - uk;
  const uint32_t end = uk + idx > n_corr ? idx : idx + uk;
  int32_t status = 0;
  bool multi = false;
  double C[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // upper triangle, row by row
  if (end >= n_corr) {
    status |= kPoseRotWindow;
  } else if (mixture && end == start && idx < n_corr) {
    // BmmPeakBasedEval normalises one sample too (bmm-peak-based-eval.cc:
    // 83-85): its weight must be positive; the fit of one sample is its
    // quaternion either way
    if (!(unordered_value((uint32_t)(key >> 32)) > 0.f)) status |= kPoseRotWeights;
  } else if (end > start && idx < n_corr) {
    const float* g = corr + p * corr_stride;
    double sum = 0.0;
    for (uint32_t i = start + lane; i <= end; i += 64) sum += (double)g[i];
    sum = wave_sum(sum);
    if (!(sum > 0.0)) {
      status |= kPoseRotWeights;
    } else {
      multi = true;
      for (uint32_t i = start + lane; i <= end; i += 64) {
        const double w = (double)g[i] / sum;
        double z[3], q[4];
        zyz_of_index(i, bw, z);
        quat_of_zyz(z, q);
        int e = 0;
        for (int r = 0; r < 4; ++r)
          for (int c = r; c < 4; ++c) C[e++] += w * q[r] * q[c];
      }
      for (int e = 0; e < 10; ++e) C[e] = wave_sum(C[e]);
    }
  }
  if (lane != 0) return;
  DevPose& P = pose[p];
  P.rot_index = idx;
  P.rot_value = unordered_value((uint32_t)(key >> 32));
  for (int d = 0; d < 3; ++d) P.zyz[d] = zyz[d];
  bingham_from_moment(C, qp, multi, P.quat, P.rot_omega, P.rot_axes);
  euler_and_transform(P.quat, P.euler, P.T, skip_axes);
  P.status = status;
}

// Coefficients of each set in the Bo layout: out[j] = in[map[j]] (map: the
// B-layout index of the same (m, l), or -1 for degrees the input lacks).
__global__ void k_pad_coef(const float2* __restrict__ in, int64_t in_stride,
                           const int* __restrict__ map, int64_t count,
                           float2* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int src = map[j];
  out[blockIdx.y * count + j] =
      src >= 0 ? in[blockIdx.y * in_stride + src] : make_float2(0.f, 0.f);
}

}  // namespace

// coefficient q kept when its FFTShift position (q - s) mod count, s =
// ceil(count / 2) (signal-utils.h:29-37), lies in [lower, upper)
__global__ void k_coef_band(float2* __restrict__ coef, int64_t stride, uint32_t count,
                            uint32_t lower, uint32_t upper) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const uint32_t s = count % 2 ? count / 2 + 1 : count / 2;
  const uint32_t pos = q >= s ? q - s : q + count - s;
  if (pos < lower || pos >= upper) coef[blockIdx.y * stride + q] = make_float2(0.f, 0.f);
}

void launch_coef_band(float2* coef, int64_t stride, int nsets, uint32_t count, uint32_t lower,
                      uint32_t upper, hipStream_t s) {
  hipLaunchKernelGGL(k_coef_band, dim3((count + 255) / 256, nsets), dim3(256), 0, s, coef, stride,
                     count, lower, upper);
}

void launch_pad_coef(const float2* in, int64_t in_stride, const int* map, int64_t count,
                     float2* out, int nsets, hipStream_t s) {
  hipLaunchKernelGGL(k_pad_coef, dim3((unsigned)((count + 255) / 256), nsets), dim3(256), 0, s,
                     in, in_stride, map, count, out);
}

void launch_so3_synth(int bw, const float2* f_coef, int64_t f_stride, const float2* h_coef,
                      int64_t h_stride, int npairs, const float* wigner,
                      const int64_t* wig_off_dev, const float* wig_norm, float2* S,
                      int precision, hipStream_t s, const uint16_t* wig_hl,
                      const int64_t* hl_off) {
  if (precision == 1) {
    launch_so3_synth_mfma(bw, f_coef, f_stride, h_coef, h_stride, npairs, wig_hl, hl_off,
                          wig_norm, S, s);
    return;
  }
  const int n = 2 * bw;
  const int nblk = bw * (bw + 1) / 2;
  const int th = ((n + 63) / 64) * 64;
  // table columns shared by four pairs per workgroup once there are enough
  // pairs to keep the device busy (eight pairs per workgroup measured 0.35
  // against 0.22 ms at 64 pairs: their accumulators cost occupancy)
  if (npairs >= 8) {
    constexpr int NP = 4;
    hipLaunchKernelGGL(k_so3_synth<NP>, dim3(nblk, (npairs + NP - 1) / NP), dim3(th),
                       NP * 4 * bw * sizeof(float2), s, bw, npairs, f_coef, f_stride, h_coef,
                       h_stride, wigner, wig_off_dev, wig_norm, S);
  } else {
    hipLaunchKernelGGL(k_so3_synth<1>, dim3(nblk, npairs), dim3(th), 4 * bw * sizeof(float2), s,
                       bw, npairs, f_coef, f_stride, h_coef, h_stride, wigner, wig_off_dev,
                       wig_norm, S);
  }
}

bool so3_reg_ok(int bw) { return bw == 32 || bw == 64 || bw == 128; }

template <int L>
static void so3_alpha_reg_t(int bw, int npairs, const float2* S, const float2* tw, float2* U,
                            hipStream_t s) {
  constexpr int LINES = 256 / RegSplit<L>::R2;
  hipLaunchKernelGGL(k_so3_fft_alpha_reg<L>, dim3(bw, L / LINES, npairs), dim3(256), 0, s, bw, S,
                     tw, U);
}

template <int L>
static int so3_gamma_reg_blocks_t() {
  constexpr int LINES = 256 / RegSplit<L>::R2;
  return L / (2 * LINES) * L;
}

template <int L>
static void so3_gamma_reg_t(int bw, int npairs, const float2* U, const float2* tw, float* corr,
                            int64_t corr_stride, unsigned long long* partials, hipStream_t s) {
  constexpr int LINES = 256 / RegSplit<L>::R2;
  hipLaunchKernelGGL(k_so3_fft_gamma_reg<L>, dim3(L / (2 * LINES), L, npairs), dim3(256), 0, s,
                     bw, U, tw, corr, corr_stride, partials);
}

void launch_so3_fft_alpha(int bw, int npairs, const float2* S, const FftPlan& plan,
                          const float2* tw, float2* U, hipStream_t s, bool reg) {
  const int n = 2 * bw;
  if (reg && so3_reg_ok(bw)) {
    if (n == 64) so3_alpha_reg_t<64>(bw, npairs, S, tw, U, s);
    else if (n == 128) so3_alpha_reg_t<128>(bw, npairs, S, tw, U, s);
    else so3_alpha_reg_t<256>(bw, npairs, S, tw, U, s);
    return;
  }
  if (n <= 128) {
    dim3 g(bw, (n + 15) / 16, npairs);
    hipLaunchKernelGGL(k_so3_fft_alpha<16>, g, dim3(256), 2 * 16 * (n + 1) * sizeof(float2), s,
                       bw, S, plan, tw, U);
  } else {
    dim3 g(bw, (n + 7) / 8, npairs);
    hipLaunchKernelGGL(k_so3_fft_alpha<8>, g, dim3(256), 2 * 8 * (n + 1) * sizeof(float2), s, bw,
                       S, plan, tw, U);
  }
}

int so3_gamma_blocks(int bw) {
  const int n = 2 * bw;
  return ((n + 15) / 16) * n;
}

void launch_so3_fft_gamma(int bw, int npairs, const float2* U, const FftPlan& plan,
                          const float2* tw, float* corr, int64_t corr_stride,
                          unsigned long long* partials, unsigned long long* keys,
                          hipStream_t s, bool reg) {
  const int n = 2 * bw;
  if (reg && so3_reg_ok(bw)) {
    int blocks;
    if (n == 64) {
      so3_gamma_reg_t<64>(bw, npairs, U, tw, corr, corr_stride, partials, s);
      blocks = so3_gamma_reg_blocks_t<64>();
    } else if (n == 128) {
      so3_gamma_reg_t<128>(bw, npairs, U, tw, corr, corr_stride, partials, s);
      blocks = so3_gamma_reg_blocks_t<128>();
    } else {
      so3_gamma_reg_t<256>(bw, npairs, U, tw, corr, corr_stride, partials, s);
      blocks = so3_gamma_reg_blocks_t<256>();
    }
    launch_reduce_keys(partials, blocks, npairs, keys, s);
    return;
  }
  dim3 g((n + 15) / 16, n, npairs);
  const size_t sh = (16 * (bw + 1) + 2 * 8 * n) * sizeof(float2);
  hipLaunchKernelGGL(k_so3_fft_gamma, g, dim3(256), sh, s, bw, U, plan, tw, corr, corr_stride,
                     partials);
  launch_reduce_keys(partials, so3_gamma_blocks(bw), npairs, keys, s);
}

void launch_rot_pose(int bw, int npairs, const unsigned long long* keys, const float* corr,
                     int64_t corr_stride, int k, DevPose* pose, hipStream_t s, int skip_axes,
                     bool mixture) {
  hipLaunchKernelGGL(k_rot_pose, dim3(npairs), dim3(64), 0, s, bw, keys, corr, corr_stride, k,
                     pose, skip_axes, (int)mixture);
}

}  // namespace phz
