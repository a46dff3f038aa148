// kernels_spatial.hip -- translation stage for gfx950: voxelisation,
// 3-D real FFT of the voxel means, Laplace channel fusion fused with the
// cross-power spectrum, inverse 3-D FFT and a signed first-maximum argmax.
//
// Restates PhaseAligner::discretizePointcloud (phase-aligner.cc:71-131,
// include/igl/histc.cpp:34-73) and SpatialCorrelationLaplace::
// correlateSignals (spatial-correlation-laplace.cc:17-89,
// spatial-correlation.cc:142-161): C = F conj(G) with no normalisation on the
// r2c half spectrum, c2r of the first N^2(N/2+1) entries.  Grid layout is
// the reference's: linear index x*N + y + N^2*z, i.e. [z][x][y].
#include "fft_reg.hpp"
#include "kernels.hpp"
#include "pose.hpp"

namespace phz {

namespace {

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// streaming 16-byte copy: the HBM reference (bandwidth, traffic counters)
__global__ void k_copy_f4(float4* __restrict__ dst, const float4* __restrict__ src, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// igl::histc bin of x against strictly increasing float edges E[0..N),
// N >= 2.  The reference's binary search returns, for E[0] <= x <= E[N-1],
// the largest k <= N-2 with E[k] <= x, or N-1 when x == E[N-1] (and N-2 for
// a NaN, which passes both range tests); -1 outside.  The same k is found
// from the uniform spacing (inv = (N-1) / (E[N-1] - E[0])) and corrected
// against the actual edges, one or two lookups instead of log2 N dependent
// ones.
__device__ __forceinline__ int histc_bin(float x, const float* E, int N, float inv) {
  if (x < E[0] || x > E[N - 1]) return -1;
  if (x != x) return N - 2;
  if (x == E[N - 1]) return N - 1;
  int k = (int)((x - E[0]) * inv);
  k = k < N - 2 ? k : N - 2;
  while (k > 0 && E[k] > x) --k;
  while (k < N - 2 && E[k + 1] <= x) ++k;
  return k;
}

// kClear = false: accumulate (phase-aligner.cc:98-120); true: zero the same
// voxels again after use.  Scans arrive in ring order, so neighbouring
// points of a wave often share a voxel: runs of equal voxels among adjacent
// lanes are summed with a segmented wave reduction and only the run's first
// lane touches memory (one atomic per run instead of per point).  The same
// lanes set (or reset) the occupancy bytes of the voxel's row and z plane.
// kExtra: the clouds also carry (reflectivity, ambient) per point, summed
// into sumF / sumA like the intensity (phase-aligner.cc:113-118).
template <bool kClear, bool kExtra>
__global__ void k_voxel_scatter(const float4* __restrict__ pts, const int* __restrict__ npts,
                                int64_t cap, const DevPose* __restrict__ pose, int pose_stride,
                                int N, const float* __restrict__ edges, float* __restrict__ sumI,
                                float* __restrict__ sumR, uint32_t* __restrict__ cnt,
                                VoxelOcc occ, const float2* __restrict__ extra,
                                float* __restrict__ sumF, float* __restrict__ sumA) {
#pragma clang fp contract(off)
  // the grid covers cap points per cloud: workgroups past the cloud's
  // points leave before staging the edges (uniform over the workgroup)
  if ((int64_t)blockIdx.x * blockDim.x >= npts[blockIdx.y]) return;
  extern __shared__ float E[];
  for (int i = threadIdx.x; i < N; i += blockDim.x) E[i] = edges[i];
  __syncthreads();
  const float inv = (float)(N - 1) / (E[N - 1] - E[0]);
  const int c = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  // every lane of the wave takes part in the reduction; lanes without a
  // voxel carry a key no other lane has
  int64_t key = -1 - (int64_t)lane;
  float vi = 0.f, vr = 0.f, vf = 0.f, va = 0.f;
  if (i < npts[c]) {
    float4 p = pts[c * cap + i];
  if (pose) {
    const float* T = pose[c * pose_stride].T;
    const float x = T[0] * p.x + T[1] * p.y + T[2] * p.z + T[3];
    const float y = T[4] * p.x + T[5] * p.y + T[6] * p.z + T[7];
    const float z = T[8] * p.x + T[9] * p.y + T[10] * p.z + T[11];
    p.x = x;
    p.y = y;
    p.z = z;
  }
    const uint32_t bx = (uint32_t)histc_bin(p.x, E, N, inv);
    const uint32_t by = (uint32_t)histc_bin(p.y, E, N, inv);
    const uint32_t bz = (uint32_t)histc_bin(p.z, E, N, inv);
    const uint32_t NN = (uint32_t)N;
    const uint32_t lin = (bx * NN + by) + (NN * NN * bz);  // Sub2Ind in uint32
    // the reference skips lin > nf and overruns at nf
    if (lin < NN * NN * NN) {
      key = (int64_t)c * NN * NN * NN + lin;
      vi = p.w;
      vr = sqrtf(p.x * p.x + p.y * p.y + p.z * p.z);
      if (kExtra) {
        const float2 e = extra[c * cap + i];
        vf = e.x;
        va = e.y;
      }
    }
  }
  // runs of equal keys in adjacent lanes: head flags, segment ids, then a
  // suffix sum within each segment that leaves the total in the head lane
  const int64_t prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const int seg = __popcll(heads & ((2ull << lane) - 1ull));
  float n = (key >= 0) ? 1.f : 0.f;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const float oi = __shfl_down(vi, d, 64), orr = __shfl_down(vr, d, 64),
                on = __shfl_down(n, d, 64);
    const int os = __shfl_down(seg, d, 64);
    float of = 0.f, oa = 0.f;
    if (kExtra) {
      of = __shfl_down(vf, d, 64);
      oa = __shfl_down(va, d, 64);
    }
    if (!kClear && lane + d < 64 && os == seg) {
      vi += oi;
      vr += orr;
      n += on;
      if (kExtra) {
        vf += of;
        va += oa;
      }
    }
  }
  if (!head || key < 0) return;
  const size_t off = (size_t)key;
  // key = cloud N^3 + z N^2 + x N + y: row (cloud, z, x) = key / N
  if (occ.row) occ.row[off / (size_t)N] = kClear ? 0 : 1;
  if (occ.plane) occ.plane[off / ((size_t)N * N)] = kClear ? 0 : 1;
  if (kClear) {
    sumI[off] = 0.f;
    sumR[off] = 0.f;
    cnt[off] = 0u;
    if (kExtra) {
      sumF[off] = 0.f;
      sumA[off] = 0.f;
    }
    return;
  }
  atomicAdd(&sumI[off], vi);
  atomicAdd(&sumR[off], vr);
  atomicAdd(&cnt[off], (uint32_t)n);
  if (kExtra) {
    atomicAdd(&sumF[off], vf);
    atomicAdd(&sumA[off], va);
  }
}

__global__ void k_reduce_keys(const unsigned long long* __restrict__ partials, int count,
                              unsigned long long* __restrict__ keys) {
  const int p = blockIdx.x;
  unsigned long long best = 0ull;
  for (int i = threadIdx.x; i < count; i += blockDim.x) {
    const unsigned long long k = partials[(int64_t)p * count + i];
    best = k > best ? k : best;
  }
  best = block_max_key(best);
  if (threadIdx.x == 0) keys[p] = best;
}

// pass 1: r2c along y of (mean I + i mean R) rows, split into the two
// half spectra.  spec[cloud][ch][row][k], row = z*N + x.
template <bool kDense>
__global__ void k_spat_r2c(int N, const float* __restrict__ a0, const float* __restrict__ a1,
                           const uint32_t* __restrict__ cnt, int64_t cloud_stride, FftPlan plan,
                           const float2* __restrict__ tw, float2* __restrict__ spec,
                           int64_t spec_stride0, int64_t spec_stride1) {
  constexpr int RT = 16;
  const int row0 = blockIdx.x * RT, c = blockIdx.y;
  const int rows = N * N, h = N / 2 + 1;
  extern __shared__ float2 sm[];
  float2* A = sm;
  float2* B = sm + RT * N;
  for (int idx = threadIdx.x; idx < RT * N; idx += blockDim.x) {
    const int r = idx / N, y = idx - r * N;
    float2 v = make_float2(0.f, 0.f);
    if (row0 + r < rows) {
      const int64_t off = c * cloud_stride + (int64_t)(row0 + r) * N + y;
      if (kDense) {
        v.x = a0[off];
        v.y = a1 ? a1[off] : 0.f;
      } else {
        // empty voxels (the vast majority) never touch the sum arrays
        const uint32_t k = cnt[off];
        if (k) {
          v.x = a0[off] / (float)k;
          v.y = a1[off] / (float)k;
        }
      }
    }
    A[idx] = v;
  }
  __syncthreads();
  float2* R = lds_fft<-1>(A, B, RT, plan, tw);
  for (int idx = threadIdx.x; idx < RT * h; idx += blockDim.x) {
    const int r = idx / h, k = idx - r * h;
    if (row0 + r >= rows) continue;
    const float2 zk = R[r * N + k];
    const float2 zc = cconj(R[r * N + ((N - k) % N)]);
    const float2 s = cadd(zk, zc), d = csub(zk, zc);
    const int64_t o = (int64_t)(row0 + r) * h + k;
    spec[c * spec_stride0 + o] = make_float2(0.5f * s.x, 0.5f * s.y);
    if (spec_stride1 >= 0) spec[c * spec_stride0 + spec_stride1 + o] = make_float2(0.5f * d.y, -0.5f * d.x);
  }
}

// pass 1, plane form (N <= kPlaneMaxN): one workgroup per (z plane, cloud)
// keeps the whole complex plane Z = mean I + i mean R in LDS, transforms it
// along y and then x, and splits the 2-D spectrum into the two half spectra
// with Z(-k) = conj of the real parts' transforms:
//   I^(k) = (Z(k) + conj Z(-k)) / 2,  R^(k) = (Z(k) - conj Z(-k)) / 2i.
// Replaces k_spat_r2c plus the x column pass (one HBM round trip less).
constexpr int kPlaneMaxN = 128;
constexpr int kPlaneChunk = 8;  // lines per LDS FFT batch

size_t spat_plane_lds_bytes(int N) {
  return ((size_t)N * (N + 1) + 2 * (size_t)kPlaneChunk * (N + 1)) * sizeof(float2);
}

__global__ void __launch_bounds__(256) k_spat_plane_fwd(
    int N, const float* __restrict__ sumI, const float* __restrict__ sumR,
    const uint32_t* __restrict__ cnt, int64_t cloud_stride, FftPlan plan,
    const float2* __restrict__ tw, float2* __restrict__ spec, int64_t spec_stride0,
    int64_t spec_stride1) {
  const int z = blockIdx.x, c = blockIdx.y;
  const int h = N / 2 + 1, ls = N + 1;
  extern __shared__ float2 sm[];
  float2* P = sm;                     // N x ls: the plane, row x, column y
  float2* A = sm + (size_t)N * ls;    // kPlaneChunk x ls
  float2* Bf = A + kPlaneChunk * ls;  // kPlaneChunk x ls
  const int64_t base = c * cloud_stride + (int64_t)z * N * N;
  // rows (x) in chunks: load means, FFT along y, park in P
  for (int x0 = 0; x0 < N; x0 += kPlaneChunk) {
    for (int idx = threadIdx.x; idx < kPlaneChunk * N; idx += blockDim.x) {
      const int r = idx / N, y = idx - r * N;
      float2 v = make_float2(0.f, 0.f);
      if (x0 + r < N) {
        const int64_t off = base + (int64_t)(x0 + r) * N + y;
        const uint32_t k = cnt[off];  // empty voxels never touch the sums
        if (k) {
          v.x = sumI[off] / (float)k;
          v.y = sumR[off] / (float)k;
        }
      }
      A[r * ls + y] = v;
    }
    __syncthreads();
    const float2* R = lds_fft<-1>(A, Bf, kPlaneChunk, plan, tw, ls);
    for (int idx = threadIdx.x; idx < kPlaneChunk * N; idx += blockDim.x) {
      const int r = idx / N, y = idx - r * N;
      if (x0 + r < N) P[(x0 + r) * ls + y] = R[r * ls + y];
    }
    __syncthreads();
  }
  // columns (ky) in chunks: FFT along x in place through A/Bf
  for (int k0 = 0; k0 < N; k0 += kPlaneChunk) {
    for (int idx = threadIdx.x; idx < kPlaneChunk * N; idx += blockDim.x) {
      const int x = idx / kPlaneChunk, j = idx - x * kPlaneChunk;
      A[j * ls + x] = (k0 + j < N) ? P[x * ls + k0 + j] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const float2* R = lds_fft<-1>(A, Bf, kPlaneChunk, plan, tw, ls);
    for (int idx = threadIdx.x; idx < kPlaneChunk * N; idx += blockDim.x) {
      const int x = idx / kPlaneChunk, j = idx - x * kPlaneChunk;
      if (k0 + j < N) P[x * ls + k0 + j] = R[j * ls + x];
    }
    __syncthreads();
  }
  // split into the two half spectra, spec[cloud][ch][z][kx][ky], ky < h
  float2* o0 = spec + c * spec_stride0 + (int64_t)z * N * h;
  float2* o1 = o0 + spec_stride1;
  for (int idx = threadIdx.x; idx < N * h; idx += blockDim.x) {
    const int kx = idx / h, ky = idx - kx * h;
    const float2 zk = P[kx * ls + ky];
    const float2 zc = cconj(P[((N - kx) % N) * ls + (N - ky) % N]);
    const float2 sm2 = cadd(zk, zc), d = csub(zk, zc);
    o0[idx] = make_float2(0.5f * sm2.x, 0.5f * sm2.y);
    o1[idx] = make_float2(0.5f * d.y, -0.5f * d.x);
  }
}

// in-place FFT along the middle axis of [outer][L][inner] arrays; one
// workgroup per (inner tile, outer, batch array)
template <int DIR, int IT>
__global__ void k_fft_cols(float2* __restrict__ data, int64_t batch_stride, int L, int inner,
                           FftPlan plan, const float2* __restrict__ tw) {
  const int i0 = blockIdx.x * IT, o = blockIdx.y;
  float2* base = data + blockIdx.z * batch_stride + (int64_t)o * L * inner;
  extern __shared__ float2 sm[];
  const int ls = L + 1;  // padded line stride: the transposed accesses hit distinct banks
  float2* A = sm;
  float2* B = sm + IT * ls;
  for (int idx = threadIdx.x; idx < IT * L; idx += blockDim.x) {
    const int e = idx / IT, ii = idx - e * IT;
    A[ii * ls + e] = (i0 + ii < inner) ? base[(int64_t)e * inner + i0 + ii] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2* R = lds_fft<DIR>(A, B, IT, plan, tw, ls);
  for (int idx = threadIdx.x; idx < IT * L; idx += blockDim.x) {
    const int e = idx / IT, ii = idx - e * IT;
    if (i0 + ii < inner) base[(int64_t)e * inner + i0 + ii] = R[ii * ls + e];
  }
}

__device__ __forceinline__ float2 fuse_at(const ChannelSet& cs, int mode, int64_t q) {
  float2 v = make_float2(0.f, 0.f);
  if (mode == 0) {
    float be = -1.f;
    for (int c = 0; c < cs.C; ++c) {
      const float2 x = cs.ch[c] ? cs.ch[c][q] : make_float2(0.f, 0.f);
      const float e = x.x * x.x + x.y * x.y;
      if (e > be) {
        be = e;
        v = x;
      }
    }
  } else if (mode == 1) {
    float ar = 0.f, ai = 0.f;
    for (int c = 0; c < cs.C; ++c)
      if (cs.ch[c]) {
        ar += cs.ch[c][q].x;
        ai += cs.ch[c][q].y;
      }
    v = make_float2(ar / (float)cs.C, ai / (float)cs.C);
  }
  return v;
}

// inverse pass A: fused F, G -> C = F conj(G) (spatial-correlation.cc:158-159)
// -> inverse FFT along z.  work[z][x][k2].
template <int IT>
__global__ void k_spat_cross_z(int N, ChannelSet F, int64_t f_stride, ChannelSet G,
                               int64_t g_stride, FusionRuns runs, int fuse, FftPlan plan,
                               const float2* __restrict__ tw, float2* __restrict__ work) {
  const int h = N / 2 + 1, inner = N * h;
  const int i0 = blockIdx.x * IT;
  // pair blockIdx.y: its F and G channels and its work array
  const int64_t fo = blockIdx.y * f_stride, go = blockIdx.y * g_stride;
  work += blockIdx.y * (int64_t)N * inner;
  extern __shared__ float2 sm[];
  const int ls = N + 1;
  float2* A = sm;
  float2* B = sm + IT * ls;
  for (int idx = threadIdx.x; idx < IT * N; idx += blockDim.x) {
    const int z = idx / IT, ii = idx - z * IT;
    float2 c = make_float2(0.f, 0.f);
    if (i0 + ii < inner) {
      const int64_t q = (int64_t)z * inner + i0 + ii;
      float2 f, g;
      if (fuse) {
        int mode = 0;
        for (int r = 0; r < runs.count; ++r)
          if ((uint64_t)q >= runs.start[r]) mode = (int)runs.mode[r];
        f = fuse_at(F, mode, q + fo);
        g = fuse_at(G, mode, q + go);
      } else {
        f = F.ch[0][q + fo];
        g = G.ch[0][q + go];
      }
      c.x = f.x * g.x - f.y * (-g.y);
      c.y = f.x * (-g.y) + f.y * g.x;
    }
    A[ii * ls + z] = c;
  }
  __syncthreads();
  float2* R = lds_fft<+1>(A, B, IT, plan, tw, ls);
  for (int idx = threadIdx.x; idx < IT * N; idx += blockDim.x) {
    const int z = idx / IT, ii = idx - z * IT;
    if (i0 + ii < inner) work[(int64_t)z * inner + i0 + ii] = R[ii * ls + z];
  }
}

// last pass: c2r along y (real parts of k=0 and k=N/2, FFTW hc2r), two rows
// per complex line, signed argmax over x*N + y + N^2*z.
__global__ void k_spat_c2r(int N, const float2* __restrict__ work, FftPlan plan,
                           const float2* __restrict__ tw, float* __restrict__ corr,
                           unsigned long long* __restrict__ partials, uint32_t limit) {
  constexpr int RT = 16, NL = RT / 2;
  const int h = N / 2 + 1, rows = N * N;
  const int row0 = blockIdx.x * RT;
  // pair blockIdx.y: its half spectrum, output grid and argmax partials
  work += blockIdx.y * (int64_t)rows * h;
  if (corr) corr += blockIdx.y * (int64_t)rows * N;
  partials += blockIdx.y * (int64_t)gridDim.x;
  extern __shared__ float2 sm[];
  float2* X = sm;  // RT x h
  float2* A = sm + RT * h;
  float2* B = A + NL * N;
  for (int idx = threadIdx.x; idx < RT * h; idx += blockDim.x) {
    const int r = idx / h, k = idx - r * h;
    X[idx] = (row0 + r < rows) ? work[(int64_t)(row0 + r) * h + k] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < NL * N; idx += blockDim.x) {
    const int p = idx / N, m = idx - p * N;
    const float2* xa = X + (2 * p) * h;
    const float2* xb = X + (2 * p + 1) * h;
    float2 va, vb;
    if (m == 0 || 2 * m == N) {
      va = make_float2(xa[m].x, 0.f);
      vb = make_float2(xb[m].x, 0.f);
    } else if (m < h) {
      va = xa[m];
      vb = xb[m];
    } else {
      va = cconj(xa[N - m]);
      vb = cconj(xb[N - m]);
    }
    A[idx] = make_float2(va.x - vb.y, va.y + vb.x);
  }
  __syncthreads();
  float2* R = lds_fft<+1>(A, B, NL, plan, tw);
  unsigned long long best = 0ull;
  for (int idx = threadIdx.x; idx < RT * N; idx += blockDim.x) {
    const int r = idx / N, y = idx - r * N;
    if (row0 + r >= rows) continue;
    const float2 z = R[(r >> 1) * N + y];
    const float v = (r & 1) ? z.y : z.x;
    const uint32_t lin = (uint32_t)((int64_t)(row0 + r) * N + y);
    if (corr) corr[lin] = v;
    if (lin >= limit) continue;  // past the correlation vector's length
    const unsigned long long kk = make_key(v, lin);
    best = kk > best ? kk : best;
  }
  best = block_max_key(best);
  if (threadIdx.x == 0) partials[blockIdx.x] = best;
}

// SpatialCorrelation::complexMulSeq (spatial-correlation.cc:142-161) on the
// fused (or, fuse = 0, single-channel) spectra without the inverse
// transform: work[pair][q] = F[q] conj(G[q]) over the H unshifted
// half-spectrum entries (the rest of the N^3 buffer is zero)
__global__ void k_spat_cross_product(int64_t H, ChannelSet F, int64_t f_stride, ChannelSet G,
                                     int64_t g_stride, FusionRuns runs, int fuse,
                                     float2* __restrict__ work) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= H) return;
  const int64_t fo = blockIdx.y * f_stride, go = blockIdx.y * g_stride;
  float2 f, g;
  if (fuse) {
    int mode = 0;
    for (int r = 0; r < runs.count; ++r)
      if ((uint64_t)q >= runs.start[r]) mode = (int)runs.mode[r];
    f = fuse_at(F, mode, q + fo);
    g = fuse_at(G, mode, q + go);
  } else {
    f = F.ch[0][q + fo];
    g = G.ch[0][q + go];
  }
  float2 c;
  c.x = f.x * g.x - f.y * (-g.y);
  c.y = f.x * (-g.y) + f.y * g.x;
  work[blockIdx.y * H + q] = c;
}

// The c2r input of the zero-padded correlation (spatial-correlation.cc:
// 142-161,221-230): the reference writes product entry i (i < T = N^3, in
// the product's own order) to Sub2Ind(x + pad, y + pad, z + pad, Np, Np) of a
// zeroed buffer, (x, y, z) = Ind2Sub(i, N, N); the Laplace class (shifted)
// then rotates the buffer's first T entries left by T/2 (IFFTShift,
// spatial-correlation-laplace.cc:36) and the c2r of size Np reads its first
// Hp = Np^2 (Np/2 + 1) entries.  Gathered per output q < Hp: the buffer
// entry r it holds, r's padded coordinates, the product index i, and, for
// the shifted order of the Laplace class, the unshifted half-spectrum entry
// u = (i + s) mod T (FFTShift rotates left by s = ceil(T/2), signal-utils.h:
// 29-37) that work holds when u < H.
// cube_lo <= cube_hi: SpatialCorrelationLowPass (spatial-correlation-low-pass.
// cc:39-94) multiplies only the product entries whose shifted index is
// Sub2Ind(i, j, k, N, N) with i, j, k in [cube_lo, cube_hi]; the rest stay zero
__global__ void k_spat_pad_gather(int N, int pad, int shifted, const float2* __restrict__ work,
                                  int64_t H, float2* __restrict__ out, int64_t Hp, int cube_lo,
                                  int cube_hi) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= Hp) return;
  const int64_t T = (int64_t)N * N * N;
  const int64_t Np = N + 2 * pad;
  int64_t r = q;
  if (shifted && q < T) {
    r = q + T / 2;
    if (r >= T) r -= T;
  }
  const int64_t zp = r / (Np * Np), rem = r - zp * Np * Np;
  const int64_t xp = rem / Np, yp = rem - xp * Np;
  float2 v = make_float2(0.f, 0.f);
  bool in = zp >= pad && zp < pad + N && xp >= pad && xp < pad + N && yp >= pad && yp < pad + N;
  if (in && cube_lo <= cube_hi)
    in = xp - pad >= cube_lo && xp - pad <= cube_hi && yp - pad >= cube_lo &&
         yp - pad <= cube_hi && zp - pad >= cube_lo && zp - pad <= cube_hi;
  if (in) {
    const int64_t i = (xp - pad) * N + (yp - pad) + (int64_t)N * N * (zp - pad);
    int64_t u = i;
    if (shifted) {
      u = i + (T % 2 ? T / 2 + 1 : T / 2);
      if (u >= T) u -= T;
    }
    if (u < H) v = work[blockIdx.y * H + u];
  }
  out[blockIdx.y * Hp + q] = v;
}

// The translation pose of pair blockIdx.x (one wavefront): the Gaussian fit
// to the +-k linear-index window around the argmax (GaussianPeakBasedEval,
// gaussian-peak-based-eval.cc:26-102; gaussian.cc:14-56): weights = the grid
// values normalised to sum 1 (CHECK_GT(sum, 0), :100), samples = the
// translations of the window's indices (translation-utils.cc:40-52,
// signal-utils.cc:11-19), mean = sum w t, covariance = sum w (t-mean)(t-mean)^T,
// NaN where the reference takes the root of a negative weight.
__global__ void __launch_bounds__(64) k_trans_pose(int N, int lower, int upper,
                                                   const unsigned long long* __restrict__ keys,
                                                   const float* __restrict__ corr,
                                                   int64_t corr_stride, int k,
                                                   DevPose* __restrict__ pose, uint32_t n_corr,
                                                   int mixture) {
  const int p = blockIdx.x, lane = threadIdx.x;
  const unsigned long long key = keys[p];
  const uint32_t idx = key_index(key);
  const float value = unordered_value((uint32_t)(key >> 32));
  const double width = (double)(abs(lower) + abs(upper) + 1);
  // gaussian-peak-based-eval.cc:26-37
  const uint32_t uk = (uint32_t)k;
  const uint32_t start = uk > idx ? idx : idx - uk;
  // GmmPeakBasedEval (gmm-peak-based-eval.cc:58-68): end = idx only when
  // idx + k > n_corr, so an end of n_corr fails its CHECK_LT (:81)
  const uint32_t end = mixture ? (uk + idx > n_corr ? idx : idx + uk)
                               : (uk + idx >= n_corr ? idx : idx + uk);
  int32_t status = 0;
  double mean[3], cov[6] = {0, 0, 0, 0, 0, 0};  // upper triangle, row by row
  bool negative = false;
  if (end >= n_corr) {
    status |= kPoseTransWindow;
    translation_of_index(idx, N, width, mean);
  } else if (end == start || idx >= n_corr) {  // one sample (the default 0 neighbours): value / value
    if (!(value > 0.f)) status |= kPoseTransWeights;
    translation_of_index(idx, N, width, mean);
  } else {
    const float* g = corr + p * corr_stride;
    double sum = 0.0;
    for (uint32_t i = start + lane; i <= end; i += 64) sum += (double)g[i];
    sum = wave_sum(sum);
    if (!(sum > 0.0)) status |= kPoseTransWeights;
    double m[3] = {0, 0, 0};
    for (uint32_t i = start + lane; i <= end; i += 64) {
      const double w = (double)g[i] / sum;
      negative |= w < 0.0;
      double t[3];
      translation_of_index(i, N, width, t);
      for (int d = 0; d < 3; ++d) m[d] += w * t[d];
    }
    for (int d = 0; d < 3; ++d) mean[d] = wave_sum(m[d]);
    for (uint32_t i = start + lane; i <= end; i += 64) {
      const double w = (double)g[i] / sum;
      double t[3];
      translation_of_index(i, N, width, t);
      for (int d = 0; d < 3; ++d) t[d] -= mean[d];
      int e = 0;
      for (int r = 0; r < 3; ++r)
        for (int c = r; c < 3; ++c) cov[e++] += w * t[r] * t[c];
    }
    for (int e = 0; e < 6; ++e) cov[e] = wave_sum(cov[e]);
    negative = __any(negative);
  }
  if (lane != 0) return;
  DevPose& P = pose[p];
  P.trans_index = idx;
  P.trans_value = value;
  for (int d = 0; d < 3; ++d) P.translation[d] = mean[d];
  int e = 0;
  for (int r = 0; r < 3; ++r)
    for (int c = r; c < 3; ++c) {
      const double v = cov[e++];
      P.trans_cov[r * 3 + c] = P.trans_cov[c * 3 + r] = negative ? __builtin_nan("") : v;
    }
  P.status |= status;
}

__global__ void k_apply_pose(const float4* __restrict__ pts, int64_t P,
                             const DevPose* __restrict__ pose, float4* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float* T = pose->T;
  const float4 p = pts[i];
  const float x = T[0] * p.x + T[1] * p.y + T[2] * p.z + T[3];
  const float y = T[4] * p.x + T[5] * p.y + T[6] * p.z + T[7];
  const float z = T[8] * p.x + T[9] * p.y + T[10] * p.z + T[11];
  out[i] = make_float4(x + (float)pose->translation[0], y + (float)pose->translation[1],
                       z + (float)pose->translation[2], p.w);
}

}  // namespace

void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_copy_f4(float4* dst, const float4* src, size_t n, hipStream_t s) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) return;
  hipLaunchKernelGGL(k_copy_f4, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
}

template <bool kClear>
static void voxel_pass(const float4* pts, const int* npts, int64_t cap, int nclouds,
                       const DevPose* pose, int pose_stride, int N, const float* edges,
                       float* sumI, float* sumR, uint32_t* cnt, VoxelOcc occ, VoxelExtra ex,
                       hipStream_t s) {
  dim3 g((unsigned)((cap + 255) / 256), nclouds);
  if (ex.extra)
    hipLaunchKernelGGL((k_voxel_scatter<kClear, true>), g, dim3(256), N * sizeof(float), s, pts,
                       npts, cap, pose, pose_stride, N, edges, sumI, sumR, cnt, occ, ex.extra,
                       ex.sumF, ex.sumA);
  else
    hipLaunchKernelGGL((k_voxel_scatter<kClear, false>), g, dim3(256), N * sizeof(float), s, pts,
                       npts, cap, pose, pose_stride, N, edges, sumI, sumR, cnt, occ,
                       (const float2*)nullptr, (float*)nullptr, (float*)nullptr);
}

void launch_voxel_scatter(const float4* pts, const int* npts, int64_t cap, int nclouds,
                          const DevPose* pose, int pose_stride, int N, const float* edges,
                          float* sumI, float* sumR, uint32_t* cnt, VoxelOcc occ, hipStream_t s,
                          VoxelExtra ex) {
  voxel_pass<false>(pts, npts, cap, nclouds, pose, pose_stride, N, edges, sumI, sumR, cnt, occ,
                    ex, s);
}

void launch_voxel_clear(const float4* pts, const int* npts, int64_t cap, int nclouds,
                        const DevPose* pose, int pose_stride, int N, const float* edges,
                        float* sumI, float* sumR, uint32_t* cnt, VoxelOcc occ, hipStream_t s,
                        VoxelExtra ex) {
  voxel_pass<true>(pts, npts, cap, nclouds, pose, pose_stride, N, edges, sumI, sumR, cnt, occ,
                   ex, s);
}

void launch_reduce_keys(const unsigned long long* partials, int count, int npairs,
                        unsigned long long* keys, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_keys, dim3(npairs), dim3(256), 0, s, partials, count, keys);
}

int spatial_c2r_blocks(int N) { return (N * N + 15) / 16; }

// ---------------------------------------------------------------- register
// engine versions of the three fused passes (L = N in {64, 128, 256}); the
// line transforms use fft_reg.hpp.

// pass 1: rows (z, x) along y, contiguous; R2 threads per row, n2 fastest.
// A row without points (row_occ 0) is not read: its transform is zero.  With
// sparse_out it is not written either, and a workgroup whose rows are all
// empty ends at once.
template <int L>
__global__ void __launch_bounds__(256) k_spat_r2c_reg(
    const float* __restrict__ sumI, const float* __restrict__ sumR,
    const uint32_t* __restrict__ cnt, int64_t cloud_stride, const uint8_t* __restrict__ row_occ,
    bool sparse_out, const float2* __restrict__ tw, float2* __restrict__ spec,
    int64_t spec_stride0, int64_t spec_stride1) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2, LINES = 256 / R2, S = L + 1;
  constexpr int h = L / 2 + 1;
  __shared__ float2 buf[LINES * S];
  __shared__ uint8_t row_live[LINES];
  const int ln = threadIdx.x / R2, n2 = threadIdx.x % R2;
  const int row0 = blockIdx.x * LINES, c = blockIdx.y;  // L*L rows, a multiple of LINES
  const uint8_t* occ = row_occ ? row_occ + (int64_t)c * L * L + row0 : nullptr;
  const bool live = !occ || occ[ln];
  const Twiddles<L> w = load_twiddles<L>(n2, tw);
  if (n2 == 0) row_live[ln] = live;  // read after the transform's barriers
  if (sparse_out && !__syncthreads_or(live)) return;  // uniform over the workgroup
  const int64_t off = c * cloud_stride + (int64_t)(row0 + ln) * L;
  float2 v[R1];
  if (live) {
    // count and sums of every voxel of the row in one batch of loads (the
    // sums of empty voxels are zero), then the means
    uint32_t k[R1];
    float si[R1], sr[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      const int64_t o = off + n2 + R2 * n1;
      k[n1] = cnt[o];
      si[n1] = sumI[o];
      sr[n1] = sumR[o];
    }
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1)
      v[n1] = k[n1] ? make_float2(si[n1] / (float)k[n1], sr[n1] / (float)k[n1])
                    : make_float2(0.f, 0.f);
  } else {
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) v[n1] = make_float2(0.f, 0.f);
  }
  float2 out[R1];
  fft_four_step_reg<-1, L>(v, n2, buf + ln * S, w, out);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) buf[ln * S + (n2 + R2 * j) + R1 * k2] = out[j * R2 + k2];
  __syncthreads();
  // split Z = I^ + i R^ into the two half spectra, coalesced over the rows
  float2* o0 = spec + c * spec_stride0 + (int64_t)row0 * h;
  float2* o1 = o0 + spec_stride1;
  for (int idx = threadIdx.x; idx < LINES * h; idx += blockDim.x) {
    const int r = idx / h, k = idx - r * h;
    if (sparse_out && !row_live[r]) continue;
    const float2 zk = buf[r * S + k];
    const float2 zc = cconj(buf[r * S + (L - k) % L]);
    const float2 sm = cadd(zk, zc), d = csub(zk, zc);
    o0[idx] = make_float2(0.5f * sm.x, 0.5f * sm.y);
    o1[idx] = make_float2(0.5f * d.y, -0.5f * d.x);
  }
}

// pass 1 over the occupied rows only: one workgroup per (z plane, cloud).
// A plane without points ends at once; otherwise its occupied rows are
// listed (ascending) in LDS and transformed LINES at a time, so rows from
// different parts of the plane share a batch.  Rows without points are
// neither read nor written (their spectrum is zero; only
// launch_spatial_fft_x_occ reads the result).
template <int L>
__global__ void __launch_bounds__(256) k_spat_r2c_rows(
    const float* __restrict__ sumI, const float* __restrict__ sumR,
    const uint32_t* __restrict__ cnt, int64_t cloud_stride, const uint8_t* __restrict__ row_occ,
    const uint8_t* __restrict__ plane_occ, const float2* __restrict__ tw,
    float2* __restrict__ spec, int64_t spec_stride0, int64_t spec_stride1) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2, LINES = 256 / R2, S = L + 1;
  constexpr int h = L / 2 + 1;
  static_assert(L <= 256, "one row per thread when listing");
  __shared__ float2 buf[LINES * S];
  __shared__ int rows[L];
  __shared__ int nrows;
  const int z = blockIdx.x, c = blockIdx.y;
  if (plane_occ[(int64_t)c * L + z] == 0) return;  // uniform over the workgroup
  const int ln = threadIdx.x / R2, n2 = threadIdx.x % R2;
  const Twiddles<L> w = load_twiddles<L>(n2, tw);
  // list the occupied rows x of the plane in ascending order
  {
    const int x = threadIdx.x;
    const bool occ = x < L && row_occ[((int64_t)c * L + z) * L + x] != 0;
    const unsigned long long m = __ballot(occ);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __shared__ int wcount[4];
    if (lane == 0) wcount[wv] = __popcll(m);
    __syncthreads();
    int base = 0;
    for (int i = 0; i < wv; ++i) base += wcount[i];
    if (occ) rows[base + __popcll(m & ((1ull << lane) - 1ull))] = x;
    if (threadIdx.x == 0) nrows = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    __syncthreads();
  }
  const int nr = nrows;
  const int64_t plane_off = (int64_t)z * L;  // first row of the plane
  for (int r0 = 0; r0 < nr; r0 += LINES) {
    const bool live = r0 + ln < nr;
    const int x = live ? rows[r0 + ln] : 0;
    const int64_t off = c * cloud_stride + (plane_off + x) * L;
    float2 v[R1];
    if (live) {
      uint32_t k[R1];
      float si[R1], sr[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) {
        const int64_t o = off + n2 + R2 * n1;
        k[n1] = cnt[o];
        si[n1] = sumI[o];
        sr[n1] = sumR[o];
      }
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1)
        v[n1] = k[n1] ? make_float2(si[n1] / (float)k[n1], sr[n1] / (float)k[n1])
                      : make_float2(0.f, 0.f);
    } else {
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) v[n1] = make_float2(0.f, 0.f);
    }
    float2 out[R1];
    fft_four_step_reg<-1, L>(v, n2, buf + ln * S, w, out);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) buf[ln * S + (n2 + R2 * j) + R1 * k2] = out[j * R2 + k2];
    __syncthreads();
    // split Z = I^ + i R^ into the two half spectra, one row of h at a time
    float2* o0 = spec + c * spec_stride0 + plane_off * h;
    float2* o1 = o0 + spec_stride1;
    const int nlive = nr - r0 < LINES ? nr - r0 : LINES;
    for (int idx = threadIdx.x; idx < nlive * h; idx += blockDim.x) {
      const int r = idx / h, k = idx - r * h;
      const int64_t o = (int64_t)rows[r0 + r] * h + k;
      const float2 zk = buf[r * S + k];
      const float2 zc = cconj(buf[r * S + (L - k) % L]);
      const float2 sm = cadd(zk, zc), d = csub(zk, zc);
      o0[o] = make_float2(0.5f * sm.x, 0.5f * sm.y);
      o1[o] = make_float2(0.5f * d.y, -0.5f * d.x);
    }
    __syncthreads();  // buf is reused by the next batch
  }
}

// inverse pass A: fused F, G -> C = F conj(G) -> inverse FFT along z, for
// lines q = (x, k2) of stride inner = N h; pair blockIdx.y.
template <int L>
__global__ void __launch_bounds__(256) k_spat_cross_z_reg(
    ChannelSet F, int64_t f_stride, ChannelSet G, int64_t g_stride, FusionRuns runs, int fuse,
    const float2* __restrict__ tw, float2* __restrict__ work) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2, LINES = 256 / R2, S = L + 1;
  constexpr int inner = L * (L / 2 + 1);
  __shared__ float2 buf[LINES * S];
  const int cl = threadIdx.x % LINES, n2 = threadIdx.x / LINES;
  const int q = blockIdx.x * LINES + cl;
  const bool live = q < inner;
  const int64_t fo = blockIdx.y * f_stride, go = blockIdx.y * g_stride;
  float2* wp = work + blockIdx.y * (int64_t)L * inner + q;
  const Twiddles<L> w = load_twiddles<L>(n2, tw);
  float2 v[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1) {
    float2 cz = make_float2(0.f, 0.f);
    if (live) {
      const int64_t idx = (int64_t)(n2 + R2 * n1) * inner + q;
      float2 f, g;
      if (fuse) {
        int mode = 0;
        for (int r = 0; r < runs.count; ++r)
          if ((uint64_t)idx >= runs.start[r]) mode = (int)runs.mode[r];
        f = fuse_at(F, mode, idx + fo);
        g = fuse_at(G, mode, idx + go);
      } else {
        f = F.ch[0][idx + fo];
        g = G.ch[0][idx + go];
      }
      cz.x = f.x * g.x - f.y * (-g.y);
      cz.y = f.x * (-g.y) + f.y * g.x;
    }
    v[n1] = cz;
  }
  fft_four_step<+1, L>(v, n2, buf + cl * S, w, [&](int k, float2 x) {
    if (live) wp[(int64_t)k * inner] = x;
  });
}

// fuse_at over the channels {x0, x1} followed by C - 2 absent (zero) ones:
// mode 0 keeps the first channel of largest energy (an absent channel never
// wins, its energy 0 does not exceed a present one's), 1 averages over C,
// 2 gives zero.
__device__ __forceinline__ float2 fuse_pair(float2 x0, float2 x1, int C, int mode) {
  if (mode == 0) {
    const float e0 = x0.x * x0.x + x0.y * x0.y, e1 = x1.x * x1.x + x1.y * x1.y;
    return e1 > e0 ? x1 : x0;
  }
  if (mode == 1) return make_float2((x0.x + x1.x) / (float)C, (x0.y + x1.y) / (float)C);
  return make_float2(0.f, 0.f);
}

// Forward FFT along z of the two channels of every x-transformed half
// spectrum of a pair, Laplace fusion, C = F conj(G) and the inverse FFT
// along z, all in registers (one LDS exchange per transform): the z-pass
// output is never stored.  Lines q = (x, k2) with stride inner = L h; pair
// blockIdx.y reads F.ch[0..1] + p f_stride and G.ch[0..1] + p g_stride (the
// other C - 2 channels are absent) and, where f_occ / g_occ are given (the
// clouds' z-plane occupancy, advanced by p f_occ_stride / p g_occ_stride),
// only the planes that hold points; the others are zero and were never
// written by the x pass.  A thread's forward output (k = n2 + R2 j + R1 k2)
// is the inverse transform's input of n1 = j + (R1 / R2) k2.
template <int L>
__global__ void __launch_bounds__(256, L <= 128 ? 3 : 2) k_spat_cross_zz_reg(
    ChannelSet F, int64_t f_stride, ChannelSet G, int64_t g_stride,
    const uint8_t* __restrict__ f_occ, int64_t f_occ_stride, const uint8_t* __restrict__ g_occ,
    int64_t g_occ_stride, FusionRuns runs, const float2* __restrict__ tw,
    float2* __restrict__ work) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2, LINES = 256 / R2, S = L + 1;
  constexpr int inner = L * (L / 2 + 1);
  __shared__ float2 buf[LINES * S];
  __shared__ uint32_t occ_w[2 * L / 4];
  const int cl = threadIdx.x % LINES, n2 = threadIdx.x / LINES;
  const int q = blockIdx.x * LINES + cl;
  const bool live = q < inner;
  const int p = blockIdx.y;
  // twiddles from scalar loads where a wave holds two n2 (L = 64, 128):
  // 30 vector registers fewer across the transforms
  std::conditional_t<LINES == 32, PairTwiddles<L>, Twiddles<L>> w;
  init_twiddles<L>(w, n2, tw);
  for (int i = threadIdx.x; i < 2 * L / 4; i += blockDim.x) {
    const uint8_t* o = i < L / 4 ? f_occ : g_occ;
    const int64_t st = i < L / 4 ? f_occ_stride : g_occ_stride;
    occ_w[i] = o ? ((const uint32_t*)(o + p * st))[i % (L / 4)] : 0x01010101u;
  }
  __syncthreads();
  const uint8_t* fo = (const uint8_t*)occ_w;
  const uint8_t* go = fo + L;
  float2* lb = buf + cl * S;
  // fused F over this thread's output frequencies
  float2 ff[R1], gf[R1];
  {
    float2 a[R1], b[R1], xa[R1], xb[R1];
    const float2* c0 = F.ch[0] + p * f_stride + q;
    const float2* c1 = F.ch[1] + p * f_stride + q;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      const int z = n2 + R2 * n1;
      const bool ld = live && fo[z];
      a[n1] = ld ? c0[(int64_t)z * inner] : make_float2(0.f, 0.f);
      b[n1] = ld ? c1[(int64_t)z * inner] : make_float2(0.f, 0.f);
    }
    fft_four_step_reg<-1, L>(a, n2, lb, w, xa);
    __syncthreads();
    fft_four_step_reg<-1, L>(b, n2, lb, w, xb);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R1; ++i) {
      ff[i] = xa[i];
      gf[i] = xb[i];
    }
  }
  int mode[R1];
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      const uint64_t idx = (uint64_t)(n2 + R2 * j + R1 * k2) * inner + q;
      int m = 0;
      for (int r = 0; r < runs.count; ++r)
        if (idx >= runs.start[r]) m = (int)runs.mode[r];
      mode[j * R2 + k2] = m;
    }
#pragma unroll
  for (int i = 0; i < R1; ++i) ff[i] = fuse_pair(ff[i], gf[i], F.C, mode[i]);
  {
    float2 a[R1], b[R1], xa[R1], xb[R1];
    const float2* c0 = G.ch[0] + p * g_stride + q;
    const float2* c1 = G.ch[1] + p * g_stride + q;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      const int z = n2 + R2 * n1;
      const bool ld = live && go[z];
      a[n1] = ld ? c0[(int64_t)z * inner] : make_float2(0.f, 0.f);
      b[n1] = ld ? c1[(int64_t)z * inner] : make_float2(0.f, 0.f);
    }
    fft_four_step_reg<-1, L>(a, n2, lb, w, xa);
    __syncthreads();
    fft_four_step_reg<-1, L>(b, n2, lb, w, xb);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R1; ++i) gf[i] = fuse_pair(xa[i], xb[i], G.C, mode[i]);
  }
  float2 v[R1];
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      const float2 f = ff[j * R2 + k2], g = gf[j * R2 + k2];
      float2 cz;
      cz.x = f.x * g.x - f.y * (-g.y);
      cz.y = f.x * (-g.y) + f.y * g.x;
      v[j + (R1 / R2) * k2] = cz;
    }
  float2* wp = work + p * (int64_t)L * inner + q;
  fft_four_step<+1, L>(v, n2, lb, w, [&](int k, float2 x) {
    if (live) wp[(int64_t)k * inner] = x;
  });
}

// The forward x and z passes of both clouds' two channels, their Laplace
// fusion, C = F conj(G) and the inverse z pass for one pair at one y
// frequency k2, with the (z, x) plane of that frequency in LDS: the
// x-transformed spectra never go to memory (k_fft_x_occ followed by
// k_spat_cross_zz_reg store and reload them, two HBM round trips of every
// occupied plane).  Workgroup (pair, k2) of L R2 threads: line = thread % L,
// n2 = thread / L, so each wave holds 64 lines at one n2.
// Input: the half spectra of k_spat_r2c_rows (layout [z][x][k2]; only rows
// (z, x) that hold points were written), read with buffer loads whose
// masked lanes address past the buffer (no memory access, zero returned);
// the row bits come from the clouds' occupancy bytes.  x pass: over the z
// planes that hold points, listed ascending one per line (a wave whose
// lines are all past the list idles); the z pass reads every other plane
// row as zero.  Each line exchanges through its own row (x pass) or
// column (z pass) of the plane.  A thread's z-pass output k = n2 + R2 j +
// R1 k2' is the inverse transform's input n1 = j + (R1 / R2) k2' (as in
// k_spat_cross_zz_reg).  Output: work[pair][k2][z][x], one contiguous
// slab per workgroup (k_spat_inv_plane<L, TPB, true> reads it).
// f_shared: every pair reads target cloud 0.  xcd: workgroup b runs pair
// (b mod 8) + 8 (b / 8h) at k2 = (b / 8) mod h, so that the h workgroups of
// a pair run on one XCD and share its L2 (a 128-byte line of a spectrum
// row holds 16 consecutive k2); needs a multiple of 8 pairs.
// The fused F side waits in global memory (the workgroup's own output
// slab) while the G side runs, which keeps the registers of 1024 threads
// within 128.  Measured at L = 128, 64 pairs: 1.08 ms against 0.85 ms for
// k_fft_x_occ + k_spat_cross_zz_reg: with 138 KB of LDS one workgroup fills
// a CU, and its load latency, LDS phases and barriers have no other
// workgroup to overlap with; an option (phz_tuning.xz_slab), off by default.
template <int L>
__global__ void __launch_bounds__(L * RegSplit<L>::R2) k_spat_cross_slab(
    ChannelSet F, int64_t f_stride, ChannelSet G, int64_t g_stride,
    const uint8_t* __restrict__ f_rows, const uint8_t* __restrict__ f_planes, int f_shared,
    const uint8_t* __restrict__ g_rows, const uint8_t* __restrict__ g_planes, FusionRuns runs,
    int xcd, const float2* __restrict__ tw, float2* __restrict__ work) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2, S = L + 1, h = L / 2 + 1;
  constexpr int TPB = L * R2, W = L / 32;  // W: bit words per plane row
  constexpr int64_t inner = (int64_t)L * h, H = inner * L;
  static_assert(32 % R2 == 0 && TPB >= 2 * L && L % 64 == 0, "thread layout");
  static_assert(H * 8 < 0x7fffffffLL, "one buffer resource per spectrum");
  extern __shared__ float2 P[];  // L rows (z) of S entries (x, padded)
  __shared__ uint32_t rowbits[2][L * W];
  __shared__ uint32_t planebits[2][W];
  __shared__ int zlist[2][L];
  __shared__ int nzs[2];
  __shared__ int wcnt[2 * L / 64];
  // n2 is uniform over a wave (L >= 64): the twiddles W_L^(n2 k1) are
  // scalar loads, kept out of the vector registers
  const int t = threadIdx.x, line = t % L, n2 = __builtin_amdgcn_readfirstlane(t / L);
  int p, k2;
  if (xcd) {
    const int b = blockIdx.x;
    p = (b & 7) + 8 * (b / (8 * h));
    k2 = (b >> 3) % h;
  } else {
    p = blockIdx.x / h;
    k2 = blockIdx.x - p * h;
  }
  const int fc = f_shared ? 0 : p;
  // row occupancy bytes -> bits: word r of a side covers bytes 32 r .. 32 r + 31
  for (int wi = t; wi < 2 * L * W; wi += TPB) {
    const int side = wi / (L * W), r = wi - side * (L * W);
    const uint4* src = (const uint4*)((side ? g_rows + (int64_t)p * L * L
                                            : f_rows + (int64_t)fc * L * L) + 32 * r);
    const uint4 a = src[0], b = src[1];
    const uint32_t wd[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t bits = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j) bits |= (uint32_t)(((wd[k] >> (8 * j)) & 0xffu) != 0u) << (4 * k + j);
    rowbits[side][r] = bits;
  }
  // occupied planes of both sides: bits and ascending lists
  {
    const bool in = t < 2 * L;
    const int side = t / L, z = t % L, lane = t & 63, wv = t >> 6;
    const bool occ = in && (side ? g_planes[(int64_t)p * L + z] : f_planes[(int64_t)fc * L + z]) != 0;
    const unsigned long long m = __ballot(occ);
    if (in && lane == 0) {
      wcnt[wv] = __popcll(m);
      planebits[side][z >> 5] = (uint32_t)m;
      planebits[side][(z >> 5) + 1] = (uint32_t)(m >> 32);
    }
    __syncthreads();
    if (in) {
      int base = 0;
      for (int i = side * (L / 64); i < wv; ++i) base += wcnt[i];
      if (occ) zlist[side][base + __popcll(m & ((1ull << lane) - 1ull))] = z;
      if (z == 0) {
        int tot = 0;
        for (int i = 0; i < L / 64; ++i) tot += wcnt[side * (L / 64) + i];
        nzs[side] = tot;
      }
    }
    __syncthreads();
  }
  // forward x and z passes of one channel of one side; out[j R2 + k2'] =
  // X(kz = n2 + R2 j + R1 k2', kx = line)
  auto fwd2 = [&](int side, const float2* X, float2(&out)[R1]) {
    const int nz = nzs[side];
    const bool wave_on = (line & ~63) < nz, act = line < nz;
    const int z = act ? zlist[side][line] : 0;
    float2* row = P + z * S;
    float2 v[R1];
    if (wave_on) {
      // this thread's 16 row bits (x = n2 + R2 n1) in one register
      uint32_t xm = 0;
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const uint32_t rw = act ? rowbits[side][z * W + i] : 0u;
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1)
          if (((R2 * n1) >> 5) == i) xm |= ((rw >> ((n2 + R2 * n1) & 31)) & 1u) << n1;
      }
      const __amdgpu_buffer_rsrc_t rs =
          __builtin_amdgcn_make_buffer_rsrc((void*)X, (short)0, (int)(H * 8), 0x00020000);
      // lane offset of x = n2 plus a scalar offset per n1; a masked lane's
      // offset lies past the buffer with or without the scalar part.  Two
      // 32-bit loads per entry (merged into one 64-bit load): this
      // compiler's 64-bit buffer-load builtin fetches one dword and returns
      // it in both halves.
      const uint32_t base = (uint32_t)(((z * L + n2) * h + k2) * 8);
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) {
        const uint32_t o = ((xm >> n1) & 1u) ? base : 0x80000000u;
        const int so = R2 * n1 * h * 8;
        v[n1] = make_float2(
            __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, o, so, 0)),
            __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, o + 4u, so, 0)));
      }
      dft_reg<R1, -1>(v);
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], tw[n2 * k1]);
    }
    __syncthreads();  // the previous readers of the plane are done
    if (wave_on && act) {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) row[k1 * R2 + n2] = v[k1];
    }
    __syncthreads();
    float2 r[R1];
    if (wave_on && act) {
#pragma unroll
      for (int j = 0; j < R1 / R2; ++j) {
        const int k1 = n2 + R2 * j;
        float2 u[R2];
#pragma unroll
        for (int m = 0; m < R2; ++m) u[m] = row[k1 * R2 + m];
        dft_reg<R2, -1>(u);
#pragma unroll
        for (int kk = 0; kk < R2; ++kk) r[j * R2 + kk] = u[kk];
      }
    }
    __syncthreads();  // the exchange is read before the results go back
    if (wave_on && act) {
#pragma unroll
      for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
        for (int kk = 0; kk < R2; ++kk) row[(n2 + R2 * j) + R1 * kk] = r[j * R2 + kk];
    }
    __syncthreads();
    // z pass along column x = line; planes without points are zero
    uint32_t pw[W];  // uniform: scalar registers
#pragma unroll
    for (int i = 0; i < W; ++i) pw[i] = __builtin_amdgcn_readfirstlane(planebits[side][i]);
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      const int zz = n2 + R2 * n1;
      const float2 e = P[zz * S + line];
      v[n1] = ((pw[(R2 * n1) >> 5] >> (zz & 31)) & 1u) ? e : make_float2(0.f, 0.f);
    }
    dft_reg<R1, -1>(v);
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], tw[n2 * k1]);
    __syncthreads();  // the columns are read before the exchanges overwrite them
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) P[(k1 * R2 + n2) * S + line] = v[k1];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R1 / R2; ++j) {
      const int k1 = n2 + R2 * j;
      float2 u[R2];
#pragma unroll
      for (int m = 0; m < R2; ++m) u[m] = P[(k1 * R2 + m) * S + line];
      dft_reg<R2, -1>(u);
#pragma unroll
      for (int kk = 0; kk < R2; ++kk) out[j * R2 + kk] = u[kk];
    }
  };
  // fusion mode of this thread's frequencies (2 bits each)
  uint32_t modes = 0;
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
    for (int kk = 0; kk < R2; ++kk) {
      const uint64_t idx = (uint64_t)(n2 + R2 * j + R1 * kk) * inner + (uint64_t)line * h + k2;
      uint32_t m = 0;
      for (int r = 0; r < runs.count; ++r)
        if (idx >= runs.start[r]) m = runs.mode[r];
      modes |= (m & 3u) << (2 * (j * R2 + kk));
    }
  // the fused F side waits in this workgroup's own output slab (rows kz,
  // column line) while the G side is transformed: every thread rereads and
  // finally overwrites exactly the entries it parked there, so no other
  // thread's accesses intervene
  float2* slab = work + ((int64_t)p * h + k2) * L * L + line;
  {
    float2 fa[R1], fb[R1];
    fwd2(0, F.ch[0] + p * f_stride, fa);
    fwd2(0, F.ch[1] + p * f_stride, fb);
#pragma unroll
    for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
      for (int kk = 0; kk < R2; ++kk) {
        const int i = j * R2 + kk;
        slab[(int64_t)(n2 + R2 * j + R1 * kk) * L] =
            fuse_pair(fa[i], fb[i], F.C, (modes >> (2 * i)) & 3u);
      }
  }
  float2 ga[R1], xb[R1];
  fwd2(1, G.ch[0] + p * g_stride, ga);
  fwd2(1, G.ch[1] + p * g_stride, xb);
#pragma unroll
  for (int i = 0; i < R1; ++i) ga[i] = fuse_pair(ga[i], xb[i], G.C, (modes >> (2 * i)) & 3u);
  // the last z pass's exchange is read before the inverse one is written;
  // the parked F entries are loaded only after the G side has been fused
  __syncthreads();
  float2 v[R1];
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
    for (int kk = 0; kk < R2; ++kk) {
      const int i = j * R2 + kk;
      const float2 f = slab[(int64_t)(n2 + R2 * j + R1 * kk) * L], g = ga[i];
      float2 cz;
      cz.x = f.x * g.x - f.y * (-g.y);
      cz.y = f.x * (-g.y) + f.y * g.x;
      v[j + (R1 / R2) * kk] = cz;
    }
  // inverse z pass along column line, straight to the slab
  dft_reg<R1, +1>(v);
#pragma unroll
  for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], cconj(tw[n2 * k1]));
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) P[(k1 * R2 + n2) * S + line] = v[k1];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j) {
    const int k1 = n2 + R2 * j;
    float2 u[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) u[m] = P[(k1 * R2 + m) * S + line];
    dft_reg<R2, +1>(u);
#pragma unroll
    for (int kk = 0; kk < R2; ++kk) slab[(int64_t)(k1 + R1 * kk) * L] = u[kk];
  }
}

// One side (F or G) of the four-channel cross spectrum: each channel's
// forward FFT along z (occupied planes only), folded into the fusion state
// in channel order as fuse_at does: mode 0 keeps the first channel of
// largest energy, 1 sums ((c0 + c1) + c2) + c3 and divides by C, 2 is zero.
template <int L, class TW>
__device__ __forceinline__ void fused_side4(const ChannelSet& X, int64_t off, const uint8_t* occ,
                                            int q, bool live, int n2, float2* lb, const TW& w,
                                            const int (&mode)[RegSplit<L>::R1],
                                            float2 (&acc)[RegSplit<L>::R1]) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2;
  constexpr int inner = L * (L / 2 + 1);
  float best[R1];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float2 a[R1], xa[R1];
    const float2* src = X.ch[c] + off + q;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      const int z = n2 + R2 * n1;
      const bool ld = live && occ[z];
      a[n1] = ld ? src[(int64_t)z * inner] : make_float2(0.f, 0.f);
    }
    fft_four_step_reg<-1, L>(a, n2, lb, w, xa);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R1; ++i) {
      const float e = xa[i].x * xa[i].x + xa[i].y * xa[i].y;
      if (c == 0) {
        acc[i] = xa[i];
        best[i] = e;
      } else if (mode[i] == 0) {
        if (e > best[i]) {
          acc[i] = xa[i];
          best[i] = e;
        }
      } else {
        acc[i].x += xa[i].x;
        acc[i].y += xa[i].y;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < R1; ++i) {
    if (mode[i] == 1)
      acc[i] = make_float2(acc[i].x / (float)X.C, acc[i].y / (float)X.C);
    else if (mode[i] != 0)
      acc[i] = make_float2(0.f, 0.f);
  }
}

// k_spat_cross_zz_reg for four live channels (clouds with reflectivity and
// ambient): F.ch[0..3] and G.ch[0..3] present; the channels of a side are
// transformed one at a time to bound the registers held.
template <int L>
__global__ void __launch_bounds__(256, 2) k_spat_cross_zz4_reg(
    ChannelSet F, int64_t f_stride, ChannelSet G, int64_t g_stride,
    const uint8_t* __restrict__ f_occ, int64_t f_occ_stride, const uint8_t* __restrict__ g_occ,
    int64_t g_occ_stride, FusionRuns runs, const float2* __restrict__ tw,
    float2* __restrict__ work) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2, LINES = 256 / R2, S = L + 1;
  constexpr int inner = L * (L / 2 + 1);
  __shared__ float2 buf[LINES * S];
  __shared__ uint32_t occ_w[2 * L / 4];
  const int cl = threadIdx.x % LINES, n2 = threadIdx.x / LINES;
  const int q = blockIdx.x * LINES + cl;
  const bool live = q < inner;
  const int p = blockIdx.y;
  // twiddles from scalar loads where a wave holds two n2 (L = 64, 128):
  // 30 vector registers fewer across the transforms
  std::conditional_t<LINES == 32, PairTwiddles<L>, Twiddles<L>> w;
  init_twiddles<L>(w, n2, tw);
  for (int i = threadIdx.x; i < 2 * L / 4; i += blockDim.x) {
    const uint8_t* o = i < L / 4 ? f_occ : g_occ;
    const int64_t st = i < L / 4 ? f_occ_stride : g_occ_stride;
    occ_w[i] = o ? ((const uint32_t*)(o + p * st))[i % (L / 4)] : 0x01010101u;
  }
  __syncthreads();
  const uint8_t* fo = (const uint8_t*)occ_w;
  const uint8_t* go = fo + L;
  float2* lb = buf + cl * S;
  int mode[R1];
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      const uint64_t idx = (uint64_t)(n2 + R2 * j + R1 * k2) * inner + q;
      int m = 0;
      for (int r = 0; r < runs.count; ++r)
        if (idx >= runs.start[r]) m = (int)runs.mode[r];
      mode[j * R2 + k2] = m;
    }
  float2 ff[R1], gf[R1];
  fused_side4<L>(F, p * f_stride, fo, q, live, n2, lb, w, mode, ff);
  fused_side4<L>(G, p * g_stride, go, q, live, n2, lb, w, mode, gf);
  float2 v[R1];
#pragma unroll
  for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      const float2 f = ff[j * R2 + k2], g = gf[j * R2 + k2];
      float2 cz;
      cz.x = f.x * g.x - f.y * (-g.y);
      cz.y = f.x * (-g.y) + f.y * g.x;
      v[j + (R1 / R2) * k2] = cz;
    }
  float2* wp = work + p * (int64_t)L * inner + q;
  fft_four_step<+1, L>(v, n2, lb, w, [&](int k, float2 x) {
    if (live) wp[(int64_t)k * inner] = x;
  });
}

// last pass: c2r along y (real parts of k = 0, N/2, FFTW hc2r), two rows per
// complex line, signed first-maximum argmax over x*N + y + N^2*z.
template <int L>
__global__ void __launch_bounds__(256) k_spat_c2r_reg(const float2* __restrict__ work,
                                                      const float2* __restrict__ tw,
                                                      float* __restrict__ corr,
                                                      unsigned long long* __restrict__ partials) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2, LINES = 256 / R2, S = L + 1;
  constexpr int h = L / 2 + 1, rows = L * L;
  // the workgroup's 2 LINES rows are contiguous: staged through LDS with
  // 16-byte loads, then the transform's line buffers reuse the space
  constexpr int kStage = 2 * LINES * h, kLds = kStage > LINES * S ? kStage : LINES * S;
  __shared__ __align__(16) float2 buf[kLds];
  const int ln = threadIdx.x / R2, n2 = threadIdx.x % R2;
  const int ra = (blockIdx.x * LINES + ln) * 2;  // rows ra, ra + 1 (rows is a multiple)
  work += blockIdx.y * (int64_t)rows * h;
  if (corr) corr += blockIdx.y * (int64_t)rows * L;
  partials += blockIdx.y * (int64_t)gridDim.x;
  const Twiddles<L> w = load_twiddles<L>(n2, tw);
  {
    static_assert(kStage % 2 == 0, "whole float4s");
    const float4* src = (const float4*)(work + (int64_t)blockIdx.x * 2 * LINES * h);
    for (int i = threadIdx.x; i < kStage / 2; i += blockDim.x) ((float4*)buf)[i] = src[i];
  }
  __syncthreads();
  const float2* xa = buf + 2 * ln * h;
  const float2* xb = xa + h;
  float2 v[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1) {
    const int m = n2 + R2 * n1;
    float2 va, vb;
    if (m == 0 || 2 * m == L) {
      va = make_float2(xa[m].x, 0.f);
      vb = make_float2(xb[m].x, 0.f);
    } else if (m < h) {
      va = xa[m];
      vb = xb[m];
    } else {
      va = cconj(xa[L - m]);
      vb = cconj(xb[L - m]);
    }
    v[n1] = make_float2(va.x - vb.y, va.y + vb.x);
  }
  __syncthreads();  // the staged rows are read; the line buffers overwrite them
  unsigned long long best = 0ull;
  fft_four_step<+1, L>(v, n2, buf + ln * S, w, [&](int y, float2 z) {
    const uint32_t la = (uint32_t)((int64_t)ra * L + y), lb = la + (uint32_t)L;
    if (corr) {
      corr[la] = z.x;
      corr[lb] = z.y;
    }
    const unsigned long long ka = make_key(z.x, la), kb = make_key(z.y, lb);
    best = ka > best ? ka : best;
    best = kb > best ? kb : best;
  });
  best = block_max_key(best);
  if (threadIdx.x == 0) partials[blockIdx.x] = best;
}

// Inverse FFT along x and c2r along y of one z plane of a pair, with the
// plane in LDS (L h complex, 66.5 KB at L = 128): the x-transformed plane is
// never stored.  Workgroup (z, pair) of 256 threads: lines of R2 threads,
// LINES = TPB / R2 at a time (line = thread % LINES fastest, n2 = thread /
// LINES).  Pass x: L / 2 lines, one per column k2 = 1 .. L/2 - 1 and line 0
// for columns 0 and L/2 together: c2r reads only the real parts of those two,
// and Re IFFT(a) + i Re IFFT(b) = IFFT(a_h + i b_h) with the Hermitian parts
// a_h[k] = (a[k] + conj a[-k]) / 2.  Each line exchanges through its own
// column (stride h) and writes its results back there.  Pass y: row pairs
// (2r, 2r + 1) packed as one complex line (real parts of k = 0, L/2, FFTW
// hc2r), exchanging through the pair's own two rows; signed first-maximum
// argmax over x N + y + N^2 z into partials[pair][z]; the grid is written
// when corr is given.  kSlab: work is laid out [pair][k2][z][x] (the output
// of k_spat_cross_slab) instead of [pair][z][x][k2]; the plane's h rows of
// L entries are transposed into the same LDS layout as they are stored.
template <int L, int TPB, bool kSlab>
__global__ void __launch_bounds__(TPB) k_spat_inv_plane(const float2* __restrict__ work,
                                                        const float2* __restrict__ tw,
                                                        float* __restrict__ corr,
                                                        unsigned long long* __restrict__ partials) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2, LINES = TPB / R2;
  constexpr int h = L / 2 + 1;
  static_assert((L / 2) % LINES == 0, "whole passes");
  __shared__ __align__(16) float2 P[L * h];
  const int c = threadIdx.x % LINES, n2 = threadIdx.x / LINES;
  const int z = blockIdx.x, pair = blockIdx.y;
  const Twiddles<L> w = load_twiddles<L>(n2, tw);
  if (kSlab) {
    // h rows (k2) of L contiguous entries (x), two per float4, all loads
    // of a thread in flight before its LDS stores
    constexpr int NV = h * L / 2, PER = (NV + TPB - 1) / TPB;
    const float4* src = (const float4*)(work + (int64_t)pair * h * L * L + (int64_t)z * L);
    float4 t[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int idx = threadIdx.x + TPB * i, ic = idx < NV ? idx : NV - 1;
      const int k2 = ic / (L / 2), x2 = ic - k2 * (L / 2);
      t[i] = src[(int64_t)k2 * (L * L / 2) + x2];  // in bounds; the stores are predicated
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int idx = threadIdx.x + TPB * i;
      if (idx < NV) {
        const int k2 = idx / (L / 2), x = 2 * (idx - k2 * (L / 2));
        P[x * h + k2] = make_float2(t[i].x, t[i].y);
        P[(x + 1) * h + k2] = make_float2(t[i].z, t[i].w);
      }
    }
  } else {
    // the plane (L h complex) in one batch of loads per thread, then into
    // LDS: a strided copy loop waits for each load before its LDS store,
    // one HBM round trip per iteration
    static_assert((L * h) % 2 == 0, "whole float4s");
    constexpr int NV = L * h / 2, PER = (NV + TPB - 1) / TPB;
    const float4* src = (const float4*)(work + ((int64_t)pair * L + z) * L * h);
    float4 t[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int idx = threadIdx.x + TPB * i;
      t[i] = src[idx < NV ? idx : NV - 1];  // in bounds; the store below is predicated
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int idx = threadIdx.x + TPB * i;
      if (idx < NV) ((float4*)P)[idx] = t[i];
    }
  }
  __syncthreads();
  // inverse FFT along x
  for (int g0 = 0; g0 < L / 2; g0 += LINES) {
    const int k2 = g0 + c;  // line 0: columns 0 and L/2
    float2 v[R1], out[R1];
    if (k2 == 0) {
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) {
        const int x = n2 + R2 * n1, xm = (L - x) & (L - 1);
        const float2 a = P[x * h], ar = P[xm * h], b = P[x * h + L / 2], br = P[xm * h + L / 2];
        const float2 ah = make_float2(0.5f * (a.x + ar.x), 0.5f * (a.y - ar.y));
        const float2 bh = make_float2(0.5f * (b.x + br.x), 0.5f * (b.y - br.y));
        v[n1] = make_float2(ah.x - bh.y, ah.y + bh.x);
      }
    } else {
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) v[n1] = P[(n2 + R2 * n1) * h + k2];
    }
    __syncthreads();  // the columns are read before the exchanges overwrite them
    fft_four_step_reg_strided<+1, L>(v, n2, P + k2, h, w, out);
    __syncthreads();  // the exchanges are read before the results go back
#pragma unroll
    for (int j = 0; j < R1 / R2; ++j)
#pragma unroll
      for (int kk = 0; kk < R2; ++kk) {
        const int x = (n2 + R2 * j) + R1 * kk;
        const float2 o = out[j * R2 + kk];
        if (k2 == 0) {
          P[x * h] = make_float2(o.x, 0.f);
          P[x * h + L / 2] = make_float2(o.y, 0.f);
        } else {
          P[x * h + k2] = o;
        }
      }
    __syncthreads();
  }
  // c2r along y, two rows per complex line; here the R2 threads of a line
  // are adjacent (n2 fastest) so that a line's grid stores are contiguous
  const int ln = threadIdx.x / R2, m2 = threadIdx.x % R2;
  const Twiddles<L> wy = load_twiddles<L>(m2, tw);
  unsigned long long best = 0ull;
  for (int r0 = 0; r0 < L / 2; r0 += LINES) {
    const int ra = 2 * (r0 + ln);
    float2* xa = P + ra * h;
    const float2* xb = xa + h;
    float2 v[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      const int m = m2 + R2 * n1;
      float2 va, vb;
      if (m == 0 || 2 * m == L) {
        va = make_float2(xa[m].x, 0.f);
        vb = make_float2(xb[m].x, 0.f);
      } else if (m < h) {
        va = xa[m];
        vb = xb[m];
      } else {
        va = cconj(xa[L - m]);
        vb = cconj(xb[L - m]);
      }
      v[n1] = make_float2(va.x - vb.y, va.y + vb.x);
    }
    __syncthreads();  // the rows are read before the exchange overwrites them
    const uint32_t rowa = (uint32_t)(z * L + ra);
    fft_four_step<+1, L>(v, m2, xa, wy, [&](int y, float2 zz) {
      const uint32_t la = rowa * (uint32_t)L + (uint32_t)y, lb = la + (uint32_t)L;
      if (corr) {
        corr[(int64_t)pair * L * L * L + la] = zz.x;
        corr[(int64_t)pair * L * L * L + lb] = zz.y;
      }
      const unsigned long long ka = make_key(zz.x, la), kb = make_key(zz.y, lb);
      best = ka > best ? ka : best;
      best = kb > best ? kb : best;
    });
  }
  best = block_max_key(best);
  if (threadIdx.x == 0) partials[(int64_t)pair * L + z] = best;
}

bool spatial_inv_plane_ok(int N) { return N == 64 || N == 128; }

void launch_spatial_inv_plane(int N, int npairs, const float2* work, const float2* tw,
                              float* corr, unsigned long long* partials,
                              unsigned long long* keys, hipStream_t s, bool slab) {
  const dim3 g(N, npairs);
  // 512 threads at N = 128: one pass over the columns and one over the row
  // pairs, and four waves per SIMD beside the plane's 66.5 KB of LDS
  if (N == 64) {
    if (slab)
      hipLaunchKernelGGL((k_spat_inv_plane<64, 256, true>), g, dim3(256), 0, s, work, tw, corr,
                         partials);
    else
      hipLaunchKernelGGL((k_spat_inv_plane<64, 256, false>), g, dim3(256), 0, s, work, tw, corr,
                         partials);
  } else {
    if (slab)
      hipLaunchKernelGGL((k_spat_inv_plane<128, 512, true>), g, dim3(512), 0, s, work, tw, corr,
                         partials);
    else
      hipLaunchKernelGGL((k_spat_inv_plane<128, 512, false>), g, dim3(512), 0, s, work, tw, corr,
                         partials);
  }
  launch_reduce_keys(partials, N, npairs, keys, s);
}

bool spatial_cross_slab_ok(int N, const ChannelSet& F, const ChannelSet& G) {
  return (N == 64 || N == 128) && spatial_cross_zz_ok(N, F, G);
}

template <int L>
static void launch_cross_slab_t(int npairs, ChannelSet F, int64_t fs, ChannelSet G, int64_t gs,
                                const uint8_t* f_rows, const uint8_t* f_planes, bool f_shared,
                                const uint8_t* g_rows, const uint8_t* g_planes,
                                const FusionRuns& runs, const float2* tw, float2* work,
                                hipStream_t s) {
  constexpr int h = L / 2 + 1, TPB = L * RegSplit<L>::R2;
  const int xcd = npairs % 8 == 0;
  const size_t lds = sizeof(float2) * L * (L + 1);
  hipLaunchKernelGGL(k_spat_cross_slab<L>, dim3(npairs * h), dim3(TPB), lds, s, F, fs, G, gs,
                     f_rows, f_planes, (int)f_shared, g_rows, g_planes, runs, xcd, tw, work);
}

void launch_spatial_cross_slab(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                               int64_t g_stride, const uint8_t* f_rows, const uint8_t* f_planes,
                               bool f_shared, const uint8_t* g_rows, const uint8_t* g_planes,
                               const FusionRuns& runs, const float2* tw, float2* work,
                               hipStream_t s) {
  if (N == 64)
    launch_cross_slab_t<64>(npairs, F, f_stride, G, g_stride, f_rows, f_planes, f_shared, g_rows,
                            g_planes, runs, tw, work, s);
  else
    launch_cross_slab_t<128>(npairs, F, f_stride, G, g_stride, f_rows, f_planes, f_shared,
                             g_rows, g_planes, runs, tw, work, s);
}

bool spatial_reg_ok(int N) { return N == 64 || N == 128 || N == 256; }

template <int L>
static void launch_reg_passes_r2c(int nclouds, const float* sumI, const float* sumR,
                                  const uint32_t* cnt, const uint8_t* row_occ, bool sparse_out,
                                  const float2* tw, float2* spec, hipStream_t s) {
  constexpr int LINES = 256 / RegSplit<L>::R2;
  static_assert((L * L) % LINES == 0 && L % LINES == 0, "workgroups must not straddle planes");
  const int64_t H = (int64_t)L * L * (L / 2 + 1), T = (int64_t)L * L * L;
  hipLaunchKernelGGL(k_spat_r2c_reg<L>, dim3(L * L / LINES, nclouds), dim3(256), 0, s, sumI, sumR,
                     cnt, T, row_occ, sparse_out, tw, spec, 2 * H, H);
}

template <int L>
static void launch_r2c_rows_t(int nclouds, const float* sumI, const float* sumR,
                              const uint32_t* cnt, VoxelOcc occ, const float2* tw, float2* spec,
                              hipStream_t s) {
  const int64_t H = (int64_t)L * L * (L / 2 + 1), T = (int64_t)L * L * L;
  hipLaunchKernelGGL(k_spat_r2c_rows<L>, dim3(L, nclouds), dim3(256), 0, s, sumI, sumR, cnt, T,
                     occ.row, occ.plane, tw, spec, 2 * H, H);
}

void launch_spatial_r2c_rows(int N, int nclouds, const float* sumI, const float* sumR,
                             const uint32_t* cnt, VoxelOcc occ, const float2* tw, float2* spec,
                             hipStream_t s) {
  if (N == 64) launch_r2c_rows_t<64>(nclouds, sumI, sumR, cnt, occ, tw, spec, s);
  else if (N == 128) launch_r2c_rows_t<128>(nclouds, sumI, sumR, cnt, occ, tw, spec, s);
  else launch_r2c_rows_t<256>(nclouds, sumI, sumR, cnt, occ, tw, spec, s);
}

void launch_spatial_r2c_reg(int N, int nclouds, const float* sumI, const float* sumR,
                            const uint32_t* cnt, const uint8_t* row_occ, bool sparse_out,
                            const float2* tw, float2* spec, hipStream_t s) {
  if (sparse_out && !row_occ) sparse_out = false;
  if (N == 64)
    launch_reg_passes_r2c<64>(nclouds, sumI, sumR, cnt, row_occ, sparse_out, tw, spec, s);
  else if (N == 128)
    launch_reg_passes_r2c<128>(nclouds, sumI, sumR, cnt, row_occ, sparse_out, tw, spec, s);
  else
    launch_reg_passes_r2c<256>(nclouds, sumI, sumR, cnt, row_occ, sparse_out, tw, spec, s);
}

// Forward x pass over the spectra a sparse_out r2c left (spectrum b of
// cloud b / 2, layout [z][x][k2], k2 < h = L/2+1): line q = (z, k2) along x
// (stride h).  When plane z holds no points the line is zero and is neither
// read nor written; otherwise row x is read only where row (z, x) holds
// points.  The row bytes of the (at most two) planes a workgroup's lines
// touch are staged in LDS first, so that the predicated data loads are all
// in flight together.
template <int L, int kLines>
__global__ void __launch_bounds__(kLines * RegSplit<L>::R2) k_fft_x_occ(
    float2* __restrict__ data, int64_t batch_stride, const uint8_t* __restrict__ row_occ,
    const uint8_t* __restrict__ plane_occ, const float2* __restrict__ tw) {
  constexpr int R1 = RegSplit<L>::R1, R2 = RegSplit<L>::R2, h = L / 2 + 1, S = L + 1;
  constexpr int nlines = L * h;
  static_assert(kLines <= h, "a workgroup's lines span at most two planes");
  __shared__ float2 buf[kLines * S];
  __shared__ uint32_t occ_w[2 * L / 4 + 1];
  const int c = threadIdx.x % kLines, n2 = threadIdx.x / kLines;
  const int q = blockIdx.x * kLines + c;
  const int cloud = blockIdx.y >> 1;
  const int z0 = (blockIdx.x * kLines) / h;
  const Twiddles<L> w = load_twiddles<L>(n2, tw);
  for (int i = threadIdx.x; i < 2 * L / 4; i += blockDim.x) {
    const int zz = z0 + (4 * i) / L;
    occ_w[i] = zz < L ? ((const uint32_t*)(row_occ + ((int64_t)cloud * L + zz) * L))[i % (L / 4)]
                      : 0u;
  }
  if (threadIdx.x == 0) {  // bytes 0, 1: planes z0, z0 + 1
    const uint8_t* po = plane_occ + (int64_t)cloud * L;
    occ_w[2 * L / 4] = po[z0] | (z0 + 1 < L ? (uint32_t)po[z0 + 1] << 8 : 0u);
  }
  __syncthreads();
  const uint8_t* ob = (const uint8_t*)occ_w;
  float2* base = data + blockIdx.y * batch_stride;
  bool live = q < nlines;
  const uint8_t* rb = ob;
  if (live) {
    const int z = q / h;
    live = ob[2 * L + (z - z0)] != 0;
    base += (int64_t)z * L * h + (q - z * h);
    rb = ob + (z - z0) * L;
  }
  if (!__syncthreads_or(live)) return;  // uniform over the workgroup
  float2 v[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1) {
    const int e = n2 + R2 * n1;
    v[n1] = (live && rb[e]) ? base[(int64_t)e * h] : make_float2(0.f, 0.f);
  }
  fft_four_step<-1, L>(v, n2, buf + c * S, w, [&](int k, float2 x) {
    if (live) base[(int64_t)k * h] = x;
  });
}

template <int L>
static void fft_x_occ_t(int nbatch, float2* spec, VoxelOcc occ, const float2* tw, hipStream_t s) {
  constexpr int lines = 256 / RegSplit<L>::R2, nl = L * (L / 2 + 1);
  const int64_t H = (int64_t)nl * L;
  dim3 g((nl + lines - 1) / lines, nbatch);
  hipLaunchKernelGGL((k_fft_x_occ<L, lines>), g, dim3(256), 0, s, spec, H, occ.row, occ.plane,
                     tw);
}

void launch_spatial_fft_x_occ(int N, int nbatch, float2* spec, VoxelOcc occ, const float2* tw,
                              hipStream_t s) {
  if (N == 64) fft_x_occ_t<64>(nbatch, spec, occ, tw, s);
  else if (N == 128) fft_x_occ_t<128>(nbatch, spec, occ, tw, s);
  else fft_x_occ_t<256>(nbatch, spec, occ, tw, s);
}

template <int L>
static void launch_cross_z_reg_t(int npairs, ChannelSet F, int64_t fs, ChannelSet G, int64_t gs,
                                 const FusionRuns& runs, bool fuse, const float2* tw,
                                 float2* work, hipStream_t s) {
  constexpr int LINES = 256 / RegSplit<L>::R2, inner = L * (L / 2 + 1);
  hipLaunchKernelGGL(k_spat_cross_z_reg<L>, dim3((inner + LINES - 1) / LINES, npairs), dim3(256),
                     0, s, F, fs, G, gs, runs, (int)fuse, tw, work);
}

template <int L>
static void launch_cross_zz_t(int npairs, ChannelSet F, int64_t fs, ChannelSet G, int64_t gs,
                              const uint8_t* f_occ, int64_t f_occ_stride, const uint8_t* g_occ,
                              int64_t g_occ_stride, const FusionRuns& runs, const float2* tw,
                              float2* work, hipStream_t s) {
  constexpr int LINES = 256 / RegSplit<L>::R2, inner = L * (L / 2 + 1);
  hipLaunchKernelGGL(k_spat_cross_zz_reg<L>, dim3((inner + LINES - 1) / LINES, npairs), dim3(256),
                     0, s, F, fs, G, gs, f_occ, f_occ_stride, g_occ, g_occ_stride, runs, tw, work);
}

template <int L>
static void launch_cross_zz4_t(int npairs, ChannelSet F, int64_t fs, ChannelSet G, int64_t gs,
                               const uint8_t* f_occ, int64_t f_occ_stride, const uint8_t* g_occ,
                               int64_t g_occ_stride, const FusionRuns& runs, const float2* tw,
                               float2* work, hipStream_t s) {
  constexpr int LINES = 256 / RegSplit<L>::R2, inner = L * (L / 2 + 1);
  hipLaunchKernelGGL(k_spat_cross_zz4_reg<L>, dim3((inner + LINES - 1) / LINES, npairs),
                     dim3(256), 0, s, F, fs, G, gs, f_occ, f_occ_stride, g_occ, g_occ_stride,
                     runs, tw, work);
}

bool spatial_cross_zz4_ok(int N, const ChannelSet& F, const ChannelSet& G) {
  if (!spatial_reg_ok(N)) return false;
  for (const ChannelSet* cs : {&F, &G}) {
    if (cs->C != 4) return false;
    for (int c = 0; c < 4; ++c)
      if (!cs->ch[c]) return false;
  }
  return true;
}

void launch_spatial_cross_zz4_reg(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                                  int64_t g_stride, const uint8_t* f_occ, int64_t f_occ_stride,
                                  const uint8_t* g_occ, int64_t g_occ_stride,
                                  const FusionRuns& runs, const float2* tw, float2* work,
                                  hipStream_t s) {
  if (N == 64)
    launch_cross_zz4_t<64>(npairs, F, f_stride, G, g_stride, f_occ, f_occ_stride, g_occ,
                           g_occ_stride, runs, tw, work, s);
  else if (N == 128)
    launch_cross_zz4_t<128>(npairs, F, f_stride, G, g_stride, f_occ, f_occ_stride, g_occ,
                            g_occ_stride, runs, tw, work, s);
  else
    launch_cross_zz4_t<256>(npairs, F, f_stride, G, g_stride, f_occ, f_occ_stride, g_occ,
                            g_occ_stride, runs, tw, work, s);
}

bool spatial_cross_zz_ok(int N, const ChannelSet& F, const ChannelSet& G) {
  if (!spatial_reg_ok(N)) return false;
  for (const ChannelSet* cs : {&F, &G}) {
    if (cs->C < 2 || cs->C > 8 || !cs->ch[0] || !cs->ch[1]) return false;
    for (int c = 2; c < cs->C; ++c)
      if (cs->ch[c]) return false;
  }
  return true;
}

void launch_spatial_cross_zz_reg(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                                 int64_t g_stride, const uint8_t* f_occ, int64_t f_occ_stride,
                                 const uint8_t* g_occ, int64_t g_occ_stride,
                                 const FusionRuns& runs, const float2* tw, float2* work,
                                 hipStream_t s) {
  if (N == 64)
    launch_cross_zz_t<64>(npairs, F, f_stride, G, g_stride, f_occ, f_occ_stride, g_occ,
                          g_occ_stride, runs, tw, work, s);
  else if (N == 128)
    launch_cross_zz_t<128>(npairs, F, f_stride, G, g_stride, f_occ, f_occ_stride, g_occ,
                           g_occ_stride, runs, tw, work, s);
  else
    launch_cross_zz_t<256>(npairs, F, f_stride, G, g_stride, f_occ, f_occ_stride, g_occ,
                           g_occ_stride, runs, tw, work, s);
}

void launch_spatial_cross_z_reg(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                                int64_t g_stride, const FusionRuns& runs, bool fuse,
                                const float2* tw, float2* work, hipStream_t s) {
  if (N == 64) launch_cross_z_reg_t<64>(npairs, F, f_stride, G, g_stride, runs, fuse, tw, work, s);
  else if (N == 128) launch_cross_z_reg_t<128>(npairs, F, f_stride, G, g_stride, runs, fuse, tw, work, s);
  else launch_cross_z_reg_t<256>(npairs, F, f_stride, G, g_stride, runs, fuse, tw, work, s);
}

// rows / (2 rows per line * lines per workgroup), lines = 256 / R2
int spatial_c2r_reg_blocks(int N) {
  const int r2 = N / (N >= 128 ? 16 : 8);
  return N * N / (2 * (256 / r2));
}

void launch_spatial_c2r_reg(int N, int npairs, const float2* work, const float2* tw, float* corr,
                            unsigned long long* partials, unsigned long long* keys,
                            hipStream_t s) {
  const dim3 g(spatial_c2r_reg_blocks(N), npairs);
  if (N == 64) hipLaunchKernelGGL(k_spat_c2r_reg<64>, g, dim3(256), 0, s, work, tw, corr, partials);
  else if (N == 128) hipLaunchKernelGGL(k_spat_c2r_reg<128>, g, dim3(256), 0, s, work, tw, corr, partials);
  else hipLaunchKernelGGL(k_spat_c2r_reg<256>, g, dim3(256), 0, s, work, tw, corr, partials);
  launch_reduce_keys(partials, spatial_c2r_reg_blocks(N), npairs, keys, s);
}

// column tiles: IT columns per workgroup, two padded LDS line buffers
int fft_cols_default_tile(int L) { return L <= 128 ? 16 : 8; }

bool fft_cols_tile_ok(int L, int it) {
  if (it == kRegisterEngine) return L == 64 || L == 128 || L == 256;
  return (it == 8 || it == 16 || it == 32 || it == 64) &&
         2 * (size_t)it * (L + 1) * sizeof(float2) <= 160 * 1024;
}

// register four-step engine: 256 threads, 256 / R2 lines per workgroup
template <int DIR, int L>
static void fft_lines_t(float2* data, int nbatch, int64_t batch_stride, int outer, int inner,
                        const float2* tw, hipStream_t s) {
  constexpr int lines = 256 / RegSplit<L>::R2;
  const int nl = outer * inner;
  dim3 g((nl + lines - 1) / lines, nbatch);
  hipLaunchKernelGGL((k_fft_lines<DIR, L, lines>), g, dim3(256), 0, s, data, batch_stride, nl,
                     inner, tw);
}

template <int DIR>
static void fft_lines(float2* data, int nbatch, int64_t batch_stride, int outer, int L, int inner,
                      const float2* tw, hipStream_t s) {
  if (L == 64)
    fft_lines_t<DIR, 64>(data, nbatch, batch_stride, outer, inner, tw, s);
  else if (L == 128)
    fft_lines_t<DIR, 128>(data, nbatch, batch_stride, outer, inner, tw, s);
  else
    fft_lines_t<DIR, 256>(data, nbatch, batch_stride, outer, inner, tw, s);
}

template <int DIR, int IT>
static void fft_cols_it(float2* data, int nbatch, int64_t batch_stride, int outer, int L,
                        int inner, const FftPlan& plan, const float2* tw, hipStream_t s) {
  dim3 g((inner + IT - 1) / IT, outer, nbatch);
  const int threads = IT <= 16 ? 256 : (IT == 32 ? 512 : 1024);
  hipLaunchKernelGGL((k_fft_cols<DIR, IT>), g, dim3(threads), 2 * IT * (L + 1) * sizeof(float2),
                     s, data, batch_stride, L, inner, plan, tw);
}

template <int DIR>
static void fft_cols_t(float2* data, int nbatch, int64_t batch_stride, int outer, int L,
                       int inner, int it, const FftPlan& plan, const float2* tw, hipStream_t s) {
  if (!fft_cols_tile_ok(L, it)) it = fft_cols_default_tile(L);
  if (it == kRegisterEngine) {
    fft_lines<DIR>(data, nbatch, batch_stride, outer, L, inner, tw, s);
    return;
  }
  switch (it) {
    case 8: fft_cols_it<DIR, 8>(data, nbatch, batch_stride, outer, L, inner, plan, tw, s); break;
    case 16: fft_cols_it<DIR, 16>(data, nbatch, batch_stride, outer, L, inner, plan, tw, s); break;
    case 32: fft_cols_it<DIR, 32>(data, nbatch, batch_stride, outer, L, inner, plan, tw, s); break;
    default: fft_cols_it<DIR, 64>(data, nbatch, batch_stride, outer, L, inner, plan, tw, s); break;
  }
}

static void fft_cols(float2* data, int nbatch, int64_t batch_stride, int outer, int L, int inner,
                     int dir, int it, const FftPlan& plan, const float2* tw, hipStream_t s) {
  if (dir < 0)
    fft_cols_t<-1>(data, nbatch, batch_stride, outer, L, inner, it, plan, tw, s);
  else
    fft_cols_t<+1>(data, nbatch, batch_stride, outer, L, inner, it, plan, tw, s);
}

void launch_spatial_r2c(int N, int nclouds, const float* sumI, const float* sumR,
                        const uint32_t* cnt, const FftPlan& plan, const float2* tw, float2* spec,
                        hipStream_t s) {
  const int h = N / 2 + 1;
  const int64_t H = (int64_t)N * N * h, T = (int64_t)N * N * N;
  dim3 g((N * N + 15) / 16, nclouds);
  hipLaunchKernelGGL(k_spat_r2c<false>, g, dim3(256), 2 * 16 * N * sizeof(float2), s, N, sumI,
                     sumR, cnt, T, plan, tw, spec, 2 * H, H);
}

bool spatial_plane_ok(int N) { return N <= kPlaneMaxN && spat_plane_lds_bytes(N) <= 160 * 1024; }

void launch_spatial_plane_fwd(int N, int nclouds, const float* sumI, const float* sumR,
                              const uint32_t* cnt, const FftPlan& plan, const float2* tw,
                              float2* spec, hipStream_t s) {
  const int h = N / 2 + 1;
  const int64_t H = (int64_t)N * N * h, T = (int64_t)N * N * N;
  hipLaunchKernelGGL(k_spat_plane_fwd, dim3(N, nclouds), dim3(256), spat_plane_lds_bytes(N), s, N,
                     sumI, sumR, cnt, T, plan, tw, spec, 2 * H, H);
}

void launch_spatial_fft_pass(int N, int nbatch, float2* buf, int which, int it,
                             const FftPlan& plan, const float2* tw, hipStream_t s) {
  const int h = N / 2 + 1;
  const int64_t H = (int64_t)N * N * h;
  if (which == 0)  // forward along x: [z][x][k2], columns over (z, k2)
    fft_cols(buf, nbatch, H, N, N, h, -1, it, plan, tw, s);
  else if (which == 1)  // forward along z: columns over (x, k2)
    fft_cols(buf, nbatch, H, 1, N, N * h, -1, it, plan, tw, s);
  else if (which == 2)  // inverse along x
    fft_cols(buf, nbatch, H, N, N, h, +1, it, plan, tw, s);
  else  // inverse along z
    fft_cols(buf, nbatch, H, 1, N, N * h, +1, it, plan, tw, s);
}

void launch_spatial_fft_fwd(int N, int nspec, float2* spec, const FftPlan& plan,
                            const float2* tw, const SpatialTiles& tiles, hipStream_t s) {
  // x then z, over all nspec half spectra
  launch_spatial_fft_pass(N, nspec, spec, 0, tiles.fwd_x, plan, tw, s);
  launch_spatial_fft_pass(N, nspec, spec, 1, tiles.fwd_z, plan, tw, s);
}

void launch_spatial_forward(int N, int nclouds, const float* sumI, const float* sumR,
                            const uint32_t* cnt, const FftPlan& plan, const float2* tw,
                            float2* spec, hipStream_t s) {
  launch_spatial_r2c(N, nclouds, sumI, sumR, cnt, plan, tw, spec, s);
  launch_spatial_fft_fwd(N, 2 * nclouds, spec, plan, tw, SpatialTiles{}, s);
}

void launch_spatial_forward_dense(int N, int nch, const float* chans, const FftPlan& plan,
                                  const float2* tw, float2* spec, const SpatialTiles& tiles,
                                  hipStream_t s) {
  const int h = N / 2 + 1;
  const int64_t H = (int64_t)N * N * h, T = (int64_t)N * N * N;
  for (int c = 0; c < nch; c += 2) {
    const bool pairc = c + 1 < nch;
    dim3 g((N * N + 15) / 16, 1);
    hipLaunchKernelGGL(k_spat_r2c<true>, g, dim3(256), 2 * 16 * N * sizeof(float2), s, N,
                       chans + c * T, pairc ? chans + (c + 1) * T : (const float*)nullptr,
                       (const uint32_t*)nullptr, T, plan, tw, spec + c * H, 0,
                       pairc ? H : (int64_t)-1);
  }
  launch_spatial_fft_fwd(N, nch, spec, plan, tw, tiles, s);
}

void launch_spatial_forward_two(int N, const float* a, const float* b, const FftPlan& plan,
                                const float2* tw, float2* spec, hipStream_t s) {
  const int64_t H = (int64_t)N * N * (N / 2 + 1), T = (int64_t)N * N * N;
  hipLaunchKernelGGL(k_spat_r2c<true>, dim3((N * N + 15) / 16, 1), dim3(256),
                     2 * 16 * N * sizeof(float2), s, N, a, b, (const uint32_t*)nullptr, T, plan,
                     tw, spec, 0, H);
  launch_spatial_fft_fwd(N, 2, spec, plan, tw, SpatialTiles{}, s);
}

void launch_spatial_cross_z(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                            int64_t g_stride, const FusionRuns& runs, bool fuse,
                            const FftPlan& plan, const float2* tw, float2* work, hipStream_t s) {
  const int inner = N * (N / 2 + 1);
  if (N <= 128)
    hipLaunchKernelGGL(k_spat_cross_z<16>, dim3((inner + 15) / 16, npairs), dim3(256),
                       2 * 16 * (N + 1) * sizeof(float2), s, N, F, f_stride, G, g_stride, runs,
                       (int)fuse, plan, tw, work);
  else
    hipLaunchKernelGGL(k_spat_cross_z<8>, dim3((inner + 7) / 8, npairs), dim3(256),
                       2 * 8 * (N + 1) * sizeof(float2), s, N, F, f_stride, G, g_stride, runs,
                       (int)fuse, plan, tw, work);
}

void launch_spatial_fft_inv(int N, int npairs, float2* work, const FftPlan& plan,
                            const float2* tw, const SpatialTiles& tiles, hipStream_t s) {
  launch_spatial_fft_pass(N, npairs, work, 2, tiles.inv_x, plan, tw, s);
}

void launch_spatial_c2r(int N, int npairs, const float2* work, const FftPlan& plan,
                        const float2* tw, float* corr, unsigned long long* partials,
                        unsigned long long* keys, hipStream_t s, uint32_t limit) {
  const int h = N / 2 + 1;
  const size_t sh = (16 * h + 2 * 8 * N) * sizeof(float2);
  hipLaunchKernelGGL(k_spat_c2r, dim3(spatial_c2r_blocks(N), npairs), dim3(256), sh, s, N, work,
                     plan, tw, corr, partials, limit);
  launch_reduce_keys(partials, spatial_c2r_blocks(N), npairs, keys, s);
}

void launch_spatial_inverse(int N, int npairs, ChannelSet F, int64_t f_stride, ChannelSet G,
                            int64_t g_stride, const FusionRuns& runs, bool fuse,
                            const FftPlan& plan, const float2* tw, float2* work, float* corr,
                            unsigned long long* partials, unsigned long long* keys,
                            const SpatialTiles& tiles, hipStream_t s) {
  launch_spatial_cross_z(N, npairs, F, f_stride, G, g_stride, runs, fuse, plan, tw, work, s);
  launch_spatial_fft_inv(N, npairs, work, plan, tw, tiles, s);
  launch_spatial_c2r(N, npairs, work, plan, tw, corr, partials, keys, s);
}

void launch_trans_pose(int N, int lower, int upper, int npairs, const unsigned long long* keys,
                       const float* corr, int64_t corr_stride, int k, DevPose* pose,
                       hipStream_t s, uint32_t n_corr, bool mixture) {
  if (n_corr == 0) n_corr = (uint32_t)N * N * N;
  hipLaunchKernelGGL(k_trans_pose, dim3(npairs), dim3(64), 0, s, N, lower, upper, keys, corr,
                     corr_stride, k, pose, n_corr, (int)mixture);
}

void launch_spatial_cross_product(int N, int npairs, ChannelSet F, int64_t f_stride,
                                  ChannelSet G, int64_t g_stride, const FusionRuns& runs,
                                  bool fuse, float2* work, hipStream_t s) {
  const int64_t H = (int64_t)N * N * (N / 2 + 1);
  hipLaunchKernelGGL(k_spat_cross_product, dim3((unsigned)((H + 255) / 256), npairs), dim3(256),
                     0, s, H, F, f_stride, G, g_stride, runs, (int)fuse, work);
}

void launch_spatial_pad_gather(int N, int pad, bool shifted, int npairs, const float2* work,
                               float2* out, hipStream_t s, int cube_lo, int cube_hi) {
  const int64_t H = (int64_t)N * N * (N / 2 + 1);
  const int64_t Np = N + 2 * pad, Hp = Np * Np * (Np / 2 + 1);
  hipLaunchKernelGGL(k_spat_pad_gather, dim3((unsigned)((Hp + 255) / 256), npairs), dim3(256),
                     0, s, N, pad, (int)shifted, work, H, out, Hp, cube_lo, cube_hi);
}

void launch_apply_pose(const float4* pts, int64_t P, const DevPose* pose, float4* out,
                       hipStream_t s) {
  if (P <= 0) return;
  hipLaunchKernelGGL(k_apply_pose, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, pts, P,
                     pose, out);
}

}  // namespace phz
