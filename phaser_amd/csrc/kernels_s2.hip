// kernels_s2.hip -- S^2 sampling, forward S^2 transform and coefficient
// fusion for gfx950.
//
// Sampling restates SphericalSampler::sampleUniformly
// (spherical-sampler.cc:20-27): project every point onto the unit sphere
// (spherical-projection.cc:19-41), then take the exact nearest projected
// point of each Driscoll-Healy query direction (point-cloud.cc:93-157, FLANN
// L2_Simple float distance).  Instead of a kd-tree the points are counting-
// sorted into a (theta band x phi sector) grid aligned with the queries, and
// each query searches cells outward from its own with exact lower bounds on
// the chord distance, so the answer equals a brute-force scan (ties go to the
// lowest point index).
#include "kernels.hpp"
#include "libm_f32.hpp"

namespace phz {

namespace {

constexpr float kPi = 3.14159265358979323846f;

// a point's distance from the sensor as spherical-projection.cc:22-24 forms
// it (float xy norm, double sum, no FMA contraction); the search recomputes a
// sample's range with this same function, so it equals the projection's
__device__ __forceinline__ double point_dist(float x, float y, float z) {
#pragma clang fp contract(off)
  const float dxy = sqrtf(x * x + y * y);
  const double dxyd = (double)dxy;
  const double zz = (double)(z * z);
  return sqrt(dxyd * dxyd + zz);
}

// spherical-projection.cc:19-41 with the reference's float/double mix and no
// FMA contraction; acosf and atan2f round as the host libm rounds them, so
// that the projected points are the reference's to the bit.
__device__ __forceinline__ void project_point(float x, float y, float z, float* px, float* py,
                                              float* pz) {
#pragma clang fp contract(off)
  const double dist = point_dist(x, y, z);
  const float pzf = (float)((double)z / dist);
  const double tx = (double)libm::acosf_host_exact(pzf);
  const double ty = fmod((double)libm::atan2f_host_exact(y, x) + 2.0 * M_PI, 2.0 * M_PI);
  const double stx = sin(tx);
  double sty, cty;
  sincos(ty, &sty, &cty);  // one argument reduction for both
  *px = (float)(stx * cty);
  *py = (float)(stx * sty);
  *pz = pzf;
}

// The S^2 search grid has at most kMaxSectors bands and sectors (n = 2B,
// B <= 160).
constexpr int kMaxSectors = 320;
constexpr int kMaskWords = (kMaxSectors + 63) / 64;

// k_project's workgroup: the band ranges of its 4 waves are merged in LDS
// before the memory atomics (one pair per band per workgroup; measured at
// 64 pairs: 256, 512, 1024 threads 0.124, 0.125, 0.130 ms)
constexpr int kProjectThreads = 256;

__global__ void __launch_bounds__(kProjectThreads)
k_project(const float4* __restrict__ pts, const int* __restrict__ npts, int64_t cap, int nb, int ns,
          float4* __restrict__ proj, int* __restrict__ cell,
          int* __restrict__ cell_count, unsigned* __restrict__ band_th) {
  const int c = blockIdx.y;
  // the grid covers cap points per cloud: workgroups past the cloud's
  // points leave at once (uniformly) instead of projecting dummy points
  if ((int64_t)blockIdx.x * blockDim.x >= npts[c]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < npts[c];
  const float4 p = valid ? pts[c * cap + i] : make_float4(0.f, 0.f, 1.f, 0.f);
  float px, py, pz;
  project_point(p.x, p.y, p.z, &px, &py, &pz);
  // cell of the stored float point: nb colatitude bands, ns azimuth sectors
  // centred on phi_s = 2 pi s / ns
  const float th = atan2f(sqrtf(px * px + py * py), pz);
  int band = (int)(th * ((float)nb / kPi));
  band = band < 0 ? 0 : (band >= nb ? nb - 1 : band);
  float ph = atan2f(py, px);
  if (ph < 0.f) ph += 2.f * kPi;
  int sec = (int)floorf(ph * ((float)ns / (2.f * kPi)) + 0.5f);
  sec = sec >= ns ? sec - ns : (sec < 0 ? 0 : sec);
  const int cid = band * ns + sec;
  // scans arrive in ring order: one count atomic per run of equal cells
  const WaveRun run = wave_run(valid ? (long long)cid : -1ll - (threadIdx.x & 63));
  if (valid) {
    proj[c * cap + i] = make_float4(px, py, pz, __int_as_float(i));
    cell[c * cap + i] = cid;
    if (run.head) atomicAdd(&cell_count[(int64_t)c * nb * ns + cid], run.len);
  }
  // each band's colatitude range (the search bounds a band by it rather than
  // by its edges), kept as the complement of the smallest and the largest
  // colatitude bits of its points (th >= 0 orders as its bits; both are
  // maxima, 0 = no point yet): per band present in a wave (a ring's points
  // share one or two), wave-wide maxima into the workgroup's LDS table, then
  // one memory atomic pair per band the workgroup holds
  __shared__ unsigned sth[2 * kMaxSectors];
  for (int k = threadIdx.x; k < 2 * nb; k += blockDim.x) sth[k] = 0u;
  __syncthreads();
  {
    const unsigned tb = __float_as_uint(th);
    unsigned long long todo = __ballot(valid);
    while (todo) {
      const int lead = __builtin_ctzll(todo);
      const int bl = __builtin_amdgcn_readlane(band, lead);
      const bool mine = valid && band == bl;
      todo &= ~__ballot(mine);
      const unsigned cmn = wave_max_u32(mine ? ~tb : 0u);
      const unsigned mx = wave_max_u32(mine ? tb : 0u);
      if ((threadIdx.x & 63) == lead) {
        atomicMax(&sth[2 * bl], cmn);
        atomicMax(&sth[2 * bl + 1], mx);
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 2 * nb; k += blockDim.x) {
    const unsigned v = sth[k];
    if (v != 0u) atomicMax(&band_th[(int64_t)c * nb * 2 + k], v);
  }
}

// exclusive scan of the n*n cell counts of a cloud (one 1024-thread
// workgroup per cloud, ncell <= 1024 * PER): chunk k holds cells 1024 k +
// thread (coalesced); wave scans per chunk, one wave scans the PER x 16 wave
// totals in (chunk, wave) order, and the starts are written back, also
// coalesced.  PER <= 16: every count is loaded up front and the within-wave
// sums stay in registers; larger PER (B >= 128): 16 chunks at a time, the
// within-wave sums parked in cell_start and the wave offsets added in a
// second pass (registers for all of them would spill).  The counts are
// zeroed for reuse as the scatter cursors; band_start[b] = start of cell b ns.
template <int PER>
__global__ void __launch_bounds__(1024) k_cell_scan(int nb, int ns, int* __restrict__ cell_count,
                                                    int* __restrict__ cell_start,
                                                    int* __restrict__ band_start) {
  const int c = blockIdx.x;
  const int ncell = nb * ns;
  int* cnt = cell_count + (int64_t)c * ncell;
  int* st = cell_start + (int64_t)c * (ncell + 1);
  int* bs = band_start + (int64_t)c * (nb + 1);
  constexpr int kW = 16;  // waves per workgroup
  constexpr bool kHold = PER <= 16;
  __shared__ int wsum[PER * kW];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // within-wave exclusive sum of one chunk; the wave's total into wsum
  auto wave_scan = [&](int k, int own) {
    int y = own;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(y, off, 64);
      if (lane >= off) y += t;
    }
    if (lane == 63) wsum[k * kW + wv] = y;
    return y - own;
  };
  int v[kHold ? PER : 1];  // the counts, then their exclusive sums within the wave
  if constexpr (kHold) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = k * 1024 + threadIdx.x;
      v[k] = i < ncell ? cnt[i] : 0;
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) v[k] = wave_scan(k, v[k]);
  } else {
#pragma unroll 1
    for (int k0 = 0; k0 < PER; k0 += 16) {
      int u[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int i = (k0 + j) * 1024 + threadIdx.x;
        u[j] = (k0 + j < PER && i < ncell) ? cnt[i] : 0;
      }
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        if (k0 + j >= PER) break;
        const int i = (k0 + j) * 1024 + threadIdx.x;
        const int e = wave_scan(k0 + j, u[j]);
        if (i < ncell) st[i] = e;
      }
    }
  }
  __syncthreads();
  if (wv == 0) {
    constexpr int E = (PER * kW + 63) / 64;  // totals per lane
    int sum = 0;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const int idx = lane * E + j;
      sum += idx < PER * kW ? wsum[idx] : 0;
    }
    int incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int u = __shfl_up(incl, off, 64);
      if (lane >= off) incl += u;
    }
    int acc = incl - sum;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const int idx = lane * E + j;
      if (idx < PER * kW) {
        const int t = wsum[idx];
        wsum[idx] = acc;
        acc += t;
      }
    }
    if (lane == 63) {
      st[ncell] = incl;
      bs[nb] = incl;
    }
  }
  __syncthreads();
  auto finish = [&](int k, int i, int excl) {
    const int start = wsum[k * kW + wv] + excl;
    st[i] = start;
    cnt[i] = 0;  // reused as the scatter cursor
    if (i % ns == 0) bs[i / ns] = start;
  };
  if constexpr (kHold) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = k * 1024 + threadIdx.x;
      if (i < ncell) finish(k, i, v[k]);
    }
  } else {
#pragma unroll 4
    for (int k = 0; k < PER; ++k) {
      const int i = k * 1024 + threadIdx.x;
      if (i < ncell) finish(k, i, st[i]);  // the parked within-wave sum
    }
  }
}

__global__ void k_cell_scatter(const int* __restrict__ npts, int64_t cap, int ncell,
                               const float4* __restrict__ proj,
                               const int* __restrict__ cell, int* __restrict__ cursor,
                               const int* __restrict__ cell_start, float4* __restrict__ sorted) {
  const int c = blockIdx.y;
  if ((int64_t)blockIdx.x * blockDim.x >= npts[c]) return;  // past the cloud (uniform)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = i < npts[c];
  const int cid = valid ? cell[c * cap + i] : 0;
  // one cursor atomic per run of equal cells; members take consecutive slots
  const WaveRun run = wave_run(valid ? (long long)cid : -1ll - lane);
  int base = 0;
  if (valid && run.head) base = atomicAdd(&cursor[(int64_t)c * ncell + cid], run.len);
  base = __shfl(base, run.head_lane, 64);
  if (!valid) return;
  const int pos = base + (lane - run.head_lane);
  const int dst = cell_start[(int64_t)c * (ncell + 1) + cid] + pos;
  sorted[c * cap + dst] = proj[c * cap + i];
}


// per cell: nearest non-empty sector of the same band to the right (>= s)
// and to the left (<= s), cyclic; -1 in an empty band.  One wave per band:
// the occupancy of the band's n <= kMaxSectors sectors is a ballot mask of
// kMaskWords 64-bit words.  Per band also its longest cyclic run of empty
// sectors, band_gap = (length << 16) | first sector (0: none, or an empty
// band), from which the search bounds a query inside the run.
__global__ void k_jump_tables(int nb, int n, const int* __restrict__ cell_start,
                              int4* __restrict__ nr, int4* __restrict__ nl,
                              unsigned long long* __restrict__ band_mask,
                              unsigned* __restrict__ band_gap) {
  // n = sectors per band here
  const int b = blockIdx.x, c = blockIdx.y;
  const int lane = threadIdx.x;  // 64 lanes
  const int* st = cell_start + (int64_t)c * (nb * n + 1) + (int64_t)b * n;
  int4* r = nr + (int64_t)c * nb * n + (int64_t)b * n;
  int4* l = nl + (int64_t)c * nb * n + (int64_t)b * n;
  unsigned long long m[kMaskWords];
  const int nw = (n + 63) / 64;
  unsigned long long all = 0ull;
#pragma unroll
  for (int w = 0; w < kMaskWords; ++w) {
    const int s = w * 64 + lane;
    const bool ne = w < nw && s < n && st[s + 1] > st[s];
    m[w] = __ballot(ne);
    all |= m[w];
  }
  // the band's occupancy words, read by the search to skip bands in LDS
#pragma unroll
  for (int w = 0; w < kMaskWords; ++w)
    if (lane == w && w < nw) band_mask[((int64_t)c * nb + b) * nw + w] = m[w];
  const bool any = all != 0ull;
  // first set bit at or after x, cyclically (-1: none)
  auto next_set = [&](int x) {
    int r = -1;
    for (int k = 0; k <= nw && r < 0; ++k) {
      const int w = (x / 64 + k) % nw;
      unsigned long long mm = m[w];
      if (k == 0) mm &= (~0ull) << (x % 64);
      if (k == nw) mm = m[w] & ((x % 64) ? ((1ull << (x % 64)) - 1ull) : 0ull);
      if (mm) r = w * 64 + __builtin_ctzll(mm);
    }
    return r;
  };
  unsigned gap = 0u;
  for (int s = lane; s < n; s += 64) {
    int right = -1, left = -1;
    if (any) {
      right = next_set(s);
      // the empty run after an occupied sector: up to the next occupied one
      if (right == s) {
        const int x = s + 1 < n ? s + 1 : 0, r1 = next_set(x);
        const int len = (r1 - s - 1 + n) % n;
        gap = max(gap, ((unsigned)len << 16) | (unsigned)x);
      }
      // last set bit at or before s, cyclically
      for (int k = 0; k <= nw && left < 0; ++k) {
        const int w = ((s / 64 - k) % nw + nw) % nw;
        unsigned long long mm = m[w];
        if (k == 0) mm &= (s % 64 == 63) ? ~0ull : ((1ull << (s % 64 + 1)) - 1ull);
        if (k == nw) mm = m[w] & ((s % 64 == 63) ? 0ull : ((~0ull) << (s % 64 + 1)));
        if (mm) left = w * 64 + 63 - __builtin_clzll(mm);
      }
    }
    // the target sector's point range rides along: one dependent load less
    r[s] = right < 0 ? make_int4(-1, 0, 0, 0) : make_int4(right, st[right], st[right + 1], 0);
    l[s] = left < 0 ? make_int4(-1, 0, 0, 0) : make_int4(left, st[left], st[left + 1], 0);
  }
  gap = wave_max_u32(gap);
  if (lane == 0) band_gap[(int64_t)c * nb + b] = gap;
}

// lane exchange within aligned groups of four: DPP quad permutes
// (xor 1: [1,0,3,2], xor 2: [2,3,0,1]) instead of an LDS-routed shuffle
template <int CTRL>
__device__ __forceinline__ int quad_xor(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}

// the LANES lanes of a query agree on the smallest key (squared distance
// bits above, original point index below: non-negative floats order as
// their bits, so the key orders by distance, then by the lower index)
template <int LANES>
__device__ __forceinline__ void agree_best(unsigned long long& key) {
  auto step = [&](unsigned lo, unsigned hi) {
    const unsigned long long o = ((unsigned long long)hi << 32) | lo;
    key = o < key ? o : key;
  };
  if constexpr (LANES >= 16)
    step(__shfl_xor((unsigned)key, 8, 64), __shfl_xor((unsigned)(key >> 32), 8, 64));
  if constexpr (LANES >= 8)
    step(__shfl_xor((unsigned)key, 4, 64), __shfl_xor((unsigned)(key >> 32), 4, 64));
  if constexpr (LANES >= 4)
    step((unsigned)quad_xor<0x4E>((int)(unsigned)key), (unsigned)quad_xor<0x4E>((int)(key >> 32)));
  if constexpr (LANES >= 2)
    step((unsigned)quad_xor<0xB1>((int)(unsigned)key), (unsigned)quad_xor<0xB1>((int)(key >> 32)));
}

__device__ __forceinline__ float l2_simple(float4 q, float4 p) {
#pragma clang fp contract(off)
  const float d0 = q.x - p.x, d1 = q.y - p.y, d2 = q.z - p.z;
  float r = 0.0f;
  r += d0 * d0;
  r += d1 * d1;
  r += d2 * d2;
  return r;
}

// Seed points of the search: per coarse cell (F x F cells of the search
// grid, cnb x cns of them), a point of the cloud near the coarse cell's
// centre, by jump flooding (each pass takes the nearest of the candidates of
// the 3 x 3 coarse cells at distance k, k = 2^j ... 1, then 2, 1).  Only the
// seeds' quality depends on this: any point of the cloud bounds a query's
// nearest distance from above, and the search stays exact.  One workgroup
// per cloud, both candidate generations in LDS (2 x kSeedCells float4: the
// point and its sorted position).
__global__ void __launch_bounds__(1024) k_cell_seeds(int nb, int ns, int F, int cnb, int cns,
                                                     const int* __restrict__ cell_start,
                                                     const float4* __restrict__ sorted,
                                                     int64_t cap, int* __restrict__ seed_pos) {
  extern __shared__ float4 gen[];  // 2 * kSeedCells
  const int c = blockIdx.x, ncc = cnb * cns;
  const int* st = cell_start + (int64_t)c * (nb * ns + 1);
  const float4* sp = sorted + c * cap;
  float4* cur = gen;
  float4* nxt = gen + kSeedCells;
  // generation 0: the first point of the coarse cell's first occupied cell
  for (int cc = threadIdx.x; cc < ncc; cc += blockDim.x) {
    const int cb = cc / cns, cs = cc - cb * cns;
    int pos = -1;
    for (int i = 0; i < F && pos < 0; ++i) {
      const int b = cb * F + i;
      if (b >= nb) break;
      for (int j = 0; j < F; ++j) {
        const int sct = cs * F + j;
        if (sct >= ns) break;
        if (st[b * ns + sct + 1] > st[b * ns + sct]) {
          pos = st[b * ns + sct];
          break;
        }
      }
    }
    float4 v = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
    if (pos >= 0) {
      v = sp[pos];
      v.w = __int_as_float(pos);
    }
    cur[cc] = v;
  }
  __syncthreads();
  // steps kmax, kmax / 2, ..., 1, then 2, 1 (two refining passes)
  int kmax = 1, lg = 0;
  while (2 * kmax < (cnb > cns ? cnb : cns)) {
    kmax *= 2;
    ++lg;
  }
  const int npass = lg + 1 + (kmax >= 2 ? 2 : 0);
  for (int pass = 0; pass < npass; ++pass) {
    const int k = pass <= lg ? kmax >> pass : (pass == lg + 1 ? 2 : 1);
    for (int cc = threadIdx.x; cc < ncc; cc += blockDim.x) {
      const int cb = cc / cns, cs = cc - cb * cns;
      const float th = fminf((cb + 0.5f) * F * (kPi / nb), kPi);
      const float ph = (cs * F + 0.5f * (F - 1)) * (2.f * kPi / ns);
      const float4 q = make_float4(sinf(th) * cosf(ph), sinf(th) * sinf(ph), cosf(th), 0.f);
      float4 best = cur[cc];
      float bd = __float_as_int(best.w) >= 0 ? l2_simple(q, best) : INFINITY;
      for (int db = -1; db <= 1; ++db) {
        const int bb = cb + db * k;
        if (bb < 0 || bb >= cnb) continue;
        for (int ds = -1; ds <= 1; ++ds) {
          int ss = cs + ds * k;
          ss = ((ss % cns) + cns) % cns;
          const float4 v = cur[bb * cns + ss];
          if (__float_as_int(v.w) < 0) continue;
          const float d = l2_simple(q, v);
          if (d < bd) {
            bd = d;
            best = v;
          }
        }
      }
      nxt[cc] = best;
    }
    __syncthreads();
    float4* t = cur;
    cur = nxt;
    nxt = t;
  }
  for (int cc = threadIdx.x; cc < ncc; cc += blockDim.x)
    seed_pos[(int64_t)c * kSeedCells + cc] = __float_as_int(cur[cc].w);
}

// first set bit at or after s (cyclically) of a band's occupancy words, or -1
__device__ __forceinline__ int mask_next(const unsigned long long* m, int nw, int s) {
  for (int k = 0; k <= nw; ++k) {
    int w = (s >> 6) + k;
    w = w >= nw ? w - nw : w;
    unsigned long long mm = m[w];
    if (k == 0) mm &= ~0ull << (s & 63);
    if (k == nw) mm &= (s & 63) ? ((1ull << (s & 63)) - 1ull) : 0ull;
    if (mm) return w * 64 + __builtin_ctzll(mm);
  }
  return -1;
}

// last set bit at or before s (cyclically), or -1
__device__ __forceinline__ int mask_prev(const unsigned long long* m, int nw, int s) {
  for (int k = 0; k <= nw; ++k) {
    int w = (s >> 6) - k;
    w = w < 0 ? w + nw : w;
    unsigned long long mm = m[w];
    if (k == 0) mm &= (s & 63) == 63 ? ~0ull : ((1ull << ((s & 63) + 1)) - 1ull);
    if (k == nw) mm &= (s & 63) == 63 ? 0ull : (~0ull << ((s & 63) + 1));
    if (mm) return w * 64 + 63 - __builtin_clzll(mm);
  }
  return -1;
}

// Exact nearest neighbour of every query, LANES lanes per query.  Bands
// are visited in order of colatitude distance (own, +1, -1, +2, ...),
// sectors of a band outward from the query's own, and a cell is skipped once
// its lower bound on the squared chord distance,
//   |q - p|^2 >= 2(1 - cos dth) + 2 sin(th_q) min sin(th_p) (1 - cos dphi),
// exceeds the best distance found (with a float slack): points outside it are
// provably farther.  dth and min sin(th_p) are taken over the colatitudes of
// the band's points (their range, from the projection), so a band whose
// points lie far from the query's colatitude within it is skipped whole; a
// direction's walk over bands ends on the bound from the band's edges, which
// grows with the band offset.  A query whose sector lies in the band's
// longest run of empty sectors skips the band when the run's nearer end is
// out of reach.  Empty sectors are jumped over with the per-cell tables,
// empty bands with the cloud's band table (in LDS).  The LANES lanes of a
// query split each cell's points and agree on the best after it.  Loads are
// batched so that a walk step costs about one memory round trip: the points
// of a cell in groups of four per lane, the next jump-table entries while the
// current cells are scanned (the walk then decides with a best that may lag
// by one step, which only prunes less).  With seeds, a query starts from its
// coarse cell's seed point (a point of the cloud: its distance bounds the
// nearest one from above, so the walk prunes from its first band and still
// finds the exact nearest point with the same tie rule); with the mask, a
// band whose nearest occupied sectors on both sides are already out of reach
// is skipped on its occupancy bits in LDS (the decision the walk would take
// after loading its first entries).
template <int LANES>
__global__ void k_nn(const float4* __restrict__ grid, const float* __restrict__ theta, int n,
                     int nb, int ns, float max_sq, const int* __restrict__ cell_start,
                     const int* __restrict__ band_start, const int4* __restrict__ next_right,
                     const int4* __restrict__ next_left, const float4* __restrict__ sorted,
                     const float4* __restrict__ pts, int64_t cap,
                     float* __restrict__ grids, const float2* __restrict__ extra,
                     float* __restrict__ xgrids, int xcd_blocks, int blend, NnSeeds seeds,
                     const unsigned* __restrict__ band_th,
                     const unsigned* __restrict__ band_gap) {
  constexpr int kBatch = 2;  // point loads in flight per lane (measured on this search: 1, 2, 4 -> 0.46, 0.42, 0.43 ms at 64 pairs)
  // xcd_blocks > 0: a 1-D grid of xcd_blocks query blocks per cloud, where
  // consecutive workgroups go to the 8 XCDs in turn: XCD x takes clouds
  // x, x + 8, ... one after the other, so the clouds whose points an XCD's
  // L2 holds are the few it is working on (instead of every cloud on every
  // XCD).  0: blockIdx.y is the cloud.
  int c, qb;
  if (xcd_blocks > 0) {
    const int b = blockIdx.x, k = b >> 3, j = k / xcd_blocks;
    c = (b & 7) + 8 * j;
    qb = k - j * xcd_blocks;
  } else {
    c = blockIdx.y;
    qb = blockIdx.x;
  }
  const int sub = threadIdx.x & (LANES - 1);
  const int q = qb * (blockDim.x / LANES) + threadIdx.x / LANES;
  const int nq = n * n, ncell = nb * ns;
  const int nw = (ns + 63) >> 6;
  // min sin(theta) over each band's points, the cloud's band table (nb <=
  // kMaxSectors) and the bands' occupancy words, once per workgroup
  __shared__ float band_sin_min[kMaxSectors];
  __shared__ int bs[kMaxSectors + 1];
  // per band its longest cyclic run of empty sectors, (length << 16) | first
  __shared__ unsigned bgap[kMaxSectors];
  // per band the nearest non-empty band at or after it (nb: none) and at or
  // before it (-1: none), so that the walk over bands skips empty runs
  __shared__ short ne_up[kMaxSectors], ne_dn[kMaxSectors];
  __shared__ unsigned long long bandbits[kMaskWords];
  // the azimuth term of the cell bound by sector distance u = 0 .. ns / 2:
  // 2 sin^2(dphi / 2), dphi = the azimuth distance from a sector's centre to
  // the nearest edge of the sector u away (queries sit on sector centres)
  __shared__ float wsec[kMaxSectors / 2 + 1];
  // the colatitude term of the cell bound, 4 sin^2(dth / 2), per band for
  // each query row of this workgroup (rows r0 .. r0 + nrows - 1; at most
  // blockDim / n + 2 rows of nb bands, n >= 16): .x from the band's edges,
  // which grows with the band offset and so ends a direction's walk; .y
  // from the colatitude range of the band's points, which skips a band (one
  // LDS read for both)
  __shared__ float2 atab[2 * kMaxSectors + 256];
  const int qpb = blockDim.x / LANES, q0 = qb * qpb;
  const int r0 = q0 / n, nrows = min(q0 + qpb, nq) > q0 ? (min(q0 + qpb, nq) - 1) / n - r0 + 1 : 0;
  extern __shared__ unsigned long long bmask[];  // nb * nw words with the mask, else none
  // band bb's colatitude range: its edges, or within them the range of its
  // points' colatitudes, each widened by 1e-5 (the error of a point's
  // computed colatitude, and of the band it is sorted into)
  auto band_edges = [&](int bb, float& lo, float& hi) {
    lo = bb * (kPi / nb) - 1e-5f;
    hi = (bb + 1) * (kPi / nb) + 1e-5f;
  };
  auto band_range = [&](int bb, float& lo, float& hi) {
    band_edges(bb, lo, hi);
    const unsigned* bt = band_th + ((int64_t)c * nb + bb) * 2;
    const unsigned mnc = bt[0], mx = bt[1];
    if (mnc != 0u) {
      lo = fmaxf(lo, __uint_as_float(~mnc) - 1e-5f);
      hi = fminf(hi, __uint_as_float(mx) + 1e-5f);
    }
  };
  for (int bb = threadIdx.x; bb < nb; bb += blockDim.x) {
    float lo, hi;
    band_range(bb, lo, hi);
    band_sin_min[bb] = fmaxf(0.f, fminf(sinf(fmaxf(lo, 0.f)), sinf(fminf(hi, kPi))));
  }
  for (int i = threadIdx.x; i < nrows * nb; i += blockDim.x) {
    const int rr = i / nb, bb = i - rr * nb;
    const float th = theta[r0 + rr];
    float lo, hi;
    band_edges(bb, lo, hi);
    const float s1 = __sinf(0.5f * fmaxf(0.f, fmaxf(lo - th, th - hi)));
    band_range(bb, lo, hi);
    const float s2 = __sinf(0.5f * fmaxf(0.f, fmaxf(lo - th, th - hi)));
    atab[i] = make_float2(4.f * s1 * s1, 4.f * s2 * s2);
  }
  for (int u = threadIdx.x; u <= ns / 2; u += blockDim.x) {
    const float dsec_ = 2.f * kPi / ns;
    const float dphi = fmaxf(u * dsec_ - 0.5f * dsec_ - 1e-5f, 0.f);
    const float s2 = __sinf(0.5f * dphi);
    wsec[u] = 2.f * s2 * s2;
  }
  for (int bb = threadIdx.x; bb <= nb; bb += blockDim.x) {
    bs[bb] = band_start[(int64_t)c * (nb + 1) + bb];
    if (bb < nb) bgap[bb] = band_gap[(int64_t)c * nb + bb];
  }
  __syncthreads();
  // band occupancy bits (one ballot per 64 bands), then the two tables
  for (int w0 = threadIdx.x & ~63; w0 < nb; w0 += blockDim.x) {
    const int bb = w0 + (threadIdx.x & 63);
    const unsigned long long m = __ballot(bb < nb && bs[bb + 1] != bs[bb]);
    if ((threadIdx.x & 63) == 0) bandbits[w0 >> 6] = m;
  }
  __syncthreads();
  for (int bb = threadIdx.x; bb < nb; bb += blockDim.x) {
    const int nwb = (nb + 63) >> 6, w = bb >> 6, o = bb & 63;
    int up_b = nb, dn_b = -1;
    const unsigned long long hi = bandbits[w] & (~0ull << o);
    if (hi) {
      up_b = (w << 6) + __builtin_ctzll(hi);
    } else {
      for (int k = w + 1; k < nwb; ++k)
        if (bandbits[k]) {
          up_b = (k << 6) + __builtin_ctzll(bandbits[k]);
          break;
        }
    }
    const unsigned long long lo = bandbits[w] & (o == 63 ? ~0ull : ((1ull << (o + 1)) - 1ull));
    if (lo) {
      dn_b = (w << 6) + 63 - __builtin_clzll(lo);
    } else {
      for (int k = w - 1; k >= 0; --k)
        if (bandbits[k]) {
          dn_b = (k << 6) + 63 - __builtin_clzll(bandbits[k]);
          break;
        }
    }
    ne_up[bb] = (short)up_b;
    ne_dn[bb] = (short)dn_b;
  }
  if (seeds.use_mask)
    for (int i = threadIdx.x; i < nb * nw; i += blockDim.x)
      bmask[i] = seeds.mask[(int64_t)c * nb * nw + i];
  __syncthreads();
  if (q >= nq) return;  // whole groups leave together
  const int qrow = q / n, qcol = q - qrow * n;
  const float4 qv = grid[q];
  const float thq = theta[qrow];
  const float phq = (2.f * kPi / n) * qcol;
  const float sq = sinf(thq);
  const int* st = cell_start + (int64_t)c * (ncell + 1);
  const int4* nr = next_right + (int64_t)c * ncell;
  const int4* nl = next_left + (int64_t)c * ncell;
  const float4* sp = sorted + c * cap;
  const float dband = kPi / nb, dsec = 2.f * kPi / ns;
  int qi = (int)(thq / dband);
  qi = qi < 0 ? 0 : (qi >= nb ? nb - 1 : qi);
  int qj = (int)floorf(phq / dsec + 0.5f);
  qj = qj >= ns ? qj - ns : qj;
  const float eps = 1e-5f, slack = 1e-5f;
  // best so far: (squared distance bits, original index); ~0 = none
  unsigned long long bkey = ~0ull;
  bool up = true, down = true;
  // pruning threshold: squared (distance + slack) plus kBoundErr, the
  // largest error of the bounds below as computed with the hardware sine
  // (v_sin_f32, absolute error below 2^-20), so pruning stays exact;
  // recomputed only when the best changes.  The square root is the
  // hardware's (v_sqrt_f32, within 1 ulp: below 2.4e-7 for the chord
  // distances < 2 here) plus kSqrtErr = 1e-6, so the threshold is never below
  // the one a correctly rounded root gives and pruning stays exact; the
  // correctly rounded sqrtf costs about 14 VALU instructions per update.
  constexpr float kBoundErr = 3e-5f, kSqrtErr = 1e-6f;
  auto thr_of = [&](float d) {
    const float t = __builtin_amdgcn_sqrtf(d) + (slack + kSqrtErr);
    return t * t + kBoundErr;
  };
  float thr = INFINITY;
  // the coarse cell's seed point: a point of the cloud, so its distance
  // bounds the exact nearest one from above and the walk below (unchanged,
  // with the same tie rule) prunes from its first band
  if (seeds.use_seeds) {
    const int pos = seeds.pos[(int64_t)c * kSeedCells + (qi / seeds.F) * seeds.cns + qj / seeds.F];
    if (pos >= 0) {
      const float4 pv = sp[pos];
      const float d = l2_simple(qv, pv);
      bkey = ((unsigned long long)__float_as_uint(d) << 32) | (unsigned)__float_as_int(pv.w);
      thr = thr_of(d);
    }
  }
  // x in [0, 2 ns) reduced mod ns without an integer division
  auto wrap = [&](int x) { return x >= ns ? x - ns : x; };
  auto scan_range = [&](int s0, int e) {
    // branch-free batches: indices past the range still load (the sorted
    // array carries kNnScanPad points of padding past its last cloud, so
    // every load stays inside the allocation) and are masked out of the
    // comparison
    for (int p0 = s0 + sub; p0 < e; p0 += kBatch * LANES) {
      const float4* base = sp + p0;
      float4 pv[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) pv[u] = base[u * LANES];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int p = p0 + u * LANES;
        const float d = l2_simple(qv, pv[u]);
        const unsigned long long k =
            ((unsigned long long)__float_as_uint(d) << 32) | (unsigned)__float_as_int(pv[u].w);
        const bool take = (p < e) & (k < bkey);
        bkey = take ? k : bkey;
      }
    }
    agree_best<LANES>(bkey);
    if (bkey != ~0ull) thr = thr_of(__uint_as_float((unsigned)(bkey >> 32)));
  };
  // offset of the next non-empty band up (increasing colatitude index) and
  // down at or beyond offset h; past the grid when there is none
  auto up_off = [&](int h) { return (qi + h < nb ? (int)ne_up[qi + h] : nb) - qi; };
  auto dn_off = [&](int h) { return qi - (qi - h >= 0 ? (int)ne_dn[qi - h] : -1); };
  int hu = up_off(1), hd = dn_off(1);
  for (int t = 0; t < 2 * nb && (up || down); ++t) {
    // band of step t: the own band, then the nearer of the next non-empty
    // bands up and down (up on a tie) while both directions walk, else the
    // next one of the direction still walking.  Empty bands take no steps:
    // the bands searched, their order and every result are those of a
    // strict up / down alternation over all bands.  A direction's flag
    // clears past the grid or once the band's bound exceeds the threshold
    // (the bound grows with the band offset, so every farther band that
    // way is farther still).  Decided with selects, then one branch around
    // the band's work.
    const bool goup = t != 0 && up && (!down || hu <= hd);
    const int braw = t == 0 ? qi : (goup ? qi + hu : qi - hd);
    if (t != 0) {
      if (goup)
        hu = up_off(hu + 1);
      else
        hd = dn_off(hd + 1);
    }
    const bool inr = braw >= 0 && braw < nb;
    const int b = inr ? braw : qi;
    // the band's table entries in one round of LDS reads
    const float2 Ab = atab[(qrow - r0) * nb + b];
    const int bs0 = bs[b], bs1 = bs[b + 1];
    const bool far = !inr || Ab.x > thr;
    up = (far && (t == 0 || goup)) ? false : up;
    down = (far && (t == 0 || !goup)) ? false : down;
    // out of reach, an empty band, or the band's points out of reach (their
    // colatitudes, not its edges: skipped while the walk goes on, as a
    // farther band's points may lie nearer its edge)
    if (far || bs1 == bs0 || Ab.y > thr) continue;
    const float A = Ab.y;
    const float spm = band_sin_min[b];
    const float Bc = 2.f * sq * spm;
    const int cb = b * ns;
    auto bound = [&](int sec) {
      // cyclic sector distance from the query's own sector (the query's
      // azimuth is that sector's centre), in [0, ns / 2], where the term
      // grows: the bound grows with it
      int u = sec > qj ? sec - qj : qj - sec;
      u = u > ns - u ? ns - u : u;
      return A + Bc * wsec[u];
    };
    {
      // the query's sector inside the band's longest run of empty sectors:
      // every occupied sector is at least as far (cyclically) as the
      // nearer end of the run, u <= ns / 2, so the bound there covers every
      // point of the band (read in LDS, without the walk's round trip)
      const unsigned gp = bgap[b];
      const int glen = (int)(gp >> 16);
      int off = qj - (int)(gp & 0xffffu);
      off += off < 0 ? ns : 0;
      const int u = min(off + 1, glen - off);
      if (off < glen && A + Bc * wsec[u] > thr) continue;
    }
    if (seeds.use_mask) {
      // own cell empty and the nearest occupied sector on each side (right
      // at distances 1..ns/2, left at 1..(ns-1)/2, as the walks below) out
      // of reach: the band holds no point within the threshold
      const unsigned long long* m = bmask + b * nw;
      if (!((m[qj >> 6] >> (qj & 63)) & 1ull)) {
        const int pr = mask_next(m, nw, wrap(qj + 1));
        const int pl = mask_prev(m, nw, wrap(qj - 1 + ns));
        const int dr = pr < 0 ? ns : wrap(pr - qj + ns);
        const int dl = pl < 0 ? ns : wrap(qj - pl + ns);
        const bool reach_r = dr >= 1 && dr <= ns / 2 && !(bound(pr) > thr);
        const bool reach_l = dl >= 1 && dl <= (ns - 1) / 2 && !(bound(pl) > thr);
        if (!reach_r && !reach_l) continue;
      }
    }
    // the own cell's range and the first entries of both walks in one round
    const int own0 = st[cb + qj], own1 = st[cb + qj + 1];
    int4 jr = nr[cb + wrap(qj + 1)];
    int4 jl = nl[cb + wrap(qj - 1 + ns)];
    scan_range(own0, own1);
    // right: sector distances 1..ns/2; left: 1..(ns-1)/2 (each sector once)
    int ur = 0, ul = 0;
    bool rdone = false, ldone = false;
    // the steps are predicated rather than branched: a walk that is done
    // keeps loading its (in-range) next entry and scans an empty range, so
    // the lanes of a wave share one instruction stream (bound() is safe for
    // any sector in [-1, ns))
    while (!rdone || !ldone) {
      const int4 cr = jr, cl = jl;
      const float th = thr;
      const int uR = wrap(jr.x - qj + ns);
      const int uL = wrap(qj - jl.x + ns);
      const bool scan_r = !rdone && jr.x >= 0 && uR > ur && uR <= ns / 2 && !(bound(jr.x) > th);
      const bool scan_l =
          !ldone && jl.x >= 0 && uL > ul && uL <= (ns - 1) / 2 && !(bound(jl.x) > th);
      rdone = !scan_r;
      ldone = !scan_l;
      ur = scan_r ? uR : ur;
      ul = scan_l ? uL : ul;
      // next entries first, so they are in flight during the scans
      jr = nr[cb + wrap(qj + ur + 1)];
      jl = nl[cb + wrap(qj - ul - 1 + ns)];
      scan_range(scan_r ? cr.y : 0, scan_r ? cr.z : 0);
      scan_range(scan_l ? cl.y : 0, scan_l ? cl.z : 0);
    }
  }
  if (sub == 0) {
    float r = 0.f, it = 0.f;
    // the nearest point's values from the cloud by its original index, where
    // the squared distance is accepted: its range as the projection forms it
    // (the same function) and its intensity
    const bool hit = bkey != ~0ull && __uint_as_float((unsigned)(bkey >> 32)) <= max_sq;
    const int oi = (int)(unsigned)bkey;
    if (hit) {
      const float4 p = pts[c * cap + oi];
      r = (float)point_dist(p.x, p.y, p.z);
      it = p.w;
    }
    if (blend) {  // SphRegistration's one channel, spherical-correlation.cc:103-112
      grids[(int64_t)c * nq + q] = 0.6f * it + 0.4f * r;
    } else {
      grids[(int64_t)c * 2 * nq + q] = r;
      grids[(int64_t)c * 2 * nq + nq + q] = it;
    }
    if (extra) {  // the nearest point's (reflectivity, ambient), point-cloud.cc:152-155
      float2 e = make_float2(0.f, 0.f);
      if (hit) e = extra[c * cap + oi];
      xgrids[(int64_t)c * 2 * nq + q] = e.x;
      xgrids[(int64_t)c * 2 * nq + nq + q] = e.y;
    }
  }
}

// ---------------------------------------------------------------- FST
// F[s][m][i] = (sqrt(2 pi)/n) sum_j x[i][j] e^{-2 pi i m j / n}, m < bw
// (s2_semi_memo.c:156-177)
__global__ void k_phi_dft(const float* __restrict__ sig, int n, int bw,
                          const float2* __restrict__ tw, float2* __restrict__ F) {
  const int i = blockIdx.x, s = blockIdx.y;
  extern __shared__ float smem[];
  float* row = smem;
  float2* t = (float2*)(smem + n);
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    row[j] = sig[((int64_t)s * n + i) * n + j];
    t[j] = tw[j];
  }
  __syncthreads();
  const float scale = 2.5066282746310002f / (float)n;  // sqrt(2 pi) / n
  for (int m = threadIdx.x; m < bw; m += blockDim.x) {
    float ar = 0.f, ai = 0.f;
    int e = 0;
    for (int j = 0; j < n; ++j) {
      const float2 w = t[e];
      ar += row[j] * w.x;
      ai += row[j] * w.y;
      e += m;
      if (e >= n) e -= n;
    }
    F[((int64_t)s * bw + m) * n + i] = make_float2(ar * scale, ai * scale);
  }
}

// coef[s][sean(m,l)] = sum_i (w_i P~_l^m(cos th_i)) F[s][m][i]; negative
// orders by real-data symmetry (s2_semi_memo.c:349-373).  One workgroup per
// (m, signal), one thread per degree l.
__global__ void k_legendre(const float2* __restrict__ F, int n, int bw,
                           const float* __restrict__ leg, const int* __restrict__ leg_off,
                           float2* __restrict__ coef) {
  const int m = blockIdx.x, s = blockIdx.y;
  extern __shared__ float2 fs[];
  for (int i = threadIdx.x; i < n; i += blockDim.x) fs[i] = F[((int64_t)s * bw + m) * n + i];
  __syncthreads();
  const int L = bw - m;
  const float* T = leg + leg_off[m];
  for (int t = threadIdx.x; t < L; t += blockDim.x) {
    float ar = 0.f, ai = 0.f;
    for (int i = 0; i < n; ++i) {
      const float w = T[(int64_t)i * L + t];
      ar += w * fs[i].x;
      ai += w * fs[i].y;
    }
    const int l = m + t;
    const int L1 = bw - 1;
    const int ip = m * (L1 + 1) - ((m * (m - 1)) / 2) + (l - m);
    const int im = ((L1 * (L1 + 3)) / 2) + 1 + ((L1 - m) * (L1 - m + 1) / 2) + (l - m);
    const float sg = (m & 1) ? -1.f : 1.f;
    float2* out = coef + (int64_t)s * bw * bw;
    out[ip] = make_float2(ar, ai);
    if (m > 0) out[im] = make_float2(sg * ar, -sg * ai);
  }
}

// ---------------------------------------------------------------- fusion
// LaplacePyramid::fuseChannels (laplace-pyramid.cc:60-151): per index, the
// channel of largest |.|^2 (first on ties), the mean over all channels in the
// innermost low-pass band, or zero where the reference's expand leaves one.
// blockIdx.y selects one of `nsets` channel sets spaced by set_stride.
__global__ void k_fuse(ChannelSet cs, int64_t set_stride, uint32_t count, FusionRuns runs,
                       float2* __restrict__ out) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const int64_t so = (int64_t)blockIdx.y * set_stride;
  int mode = 0;
  for (int r = 0; r < runs.count; ++r)
    if (q >= runs.start[r]) mode = (int)runs.mode[r];
  float2 v = make_float2(0.f, 0.f);
  if (mode == 0) {
    float be = -1.f;
    for (int c = 0; c < cs.C; ++c) {
      const float2 x = cs.ch[c] ? cs.ch[c][so + q] : make_float2(0.f, 0.f);
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
        ar += cs.ch[c][so + q].x;
        ai += cs.ch[c][so + q].y;
      }
    v = make_float2(ar / (float)cs.C, ai / (float)cs.C);
  }
  out[(int64_t)blockIdx.y * count + q] = v;
}

}  // namespace

void launch_sample_s2(const float4* pts, const int* npts, int64_t cap, int nclouds, int n,
                      const float4* grid, const float* theta, float max_sq, CellBuffers cb,
                      float* grids, hipStream_t s) {
  launch_s2_cells(pts, npts, cap, nclouds, n, cb, s);
  cb.pts = pts;
  launch_s2_nn(cap, nclouds, n, grid, theta, max_sq, cb, grids, 4, s, nullptr, nullptr, false,
               false, kNnSeedDefault);
}

void launch_s2_cells(const float4* pts, const int* npts, int64_t cap, int nclouds, int n,
                     CellBuffers cb, hipStream_t s) {
  (void)n;
  const int ncell = cb.nb * cb.ns;
  launch_fill_u32((uint32_t*)cb.cell_count, (size_t)nclouds * ncell, 0u, s);
  launch_fill_u32(cb.band_th, (size_t)nclouds * cb.nb * 2, 0u, s);
  dim3 g((unsigned)((cap + 255) / 256), nclouds);
  dim3 gp((unsigned)((cap + kProjectThreads - 1) / kProjectThreads), nclouds);
  hipLaunchKernelGGL(k_project, gp, dim3(kProjectThreads), 0, s, pts, npts, cap, cb.nb, cb.ns,
                     cb.proj, cb.cell, cb.cell_count, cb.band_th);
  if (ncell <= 4 * 1024)
    hipLaunchKernelGGL(k_cell_scan<4>, dim3(nclouds), dim3(1024), 0, s, cb.nb, cb.ns,
                       cb.cell_count, cb.cell_start, cb.band_start);
  else if (ncell <= 16 * 1024)
    hipLaunchKernelGGL(k_cell_scan<16>, dim3(nclouds), dim3(1024), 0, s, cb.nb, cb.ns,
                       cb.cell_count, cb.cell_start, cb.band_start);
  else if (ncell <= 64 * 1024)
    hipLaunchKernelGGL(k_cell_scan<64>, dim3(nclouds), dim3(1024), 0, s, cb.nb, cb.ns,
                       cb.cell_count, cb.cell_start, cb.band_start);
  else  // up to kMaxSectors^2 = 100 * 1024 cells
    hipLaunchKernelGGL(k_cell_scan<100>, dim3(nclouds), dim3(1024), 0, s, cb.nb, cb.ns,
                       cb.cell_count, cb.cell_start, cb.band_start);
  hipLaunchKernelGGL(k_cell_scatter, g, dim3(256), 0, s, npts, cap, ncell, cb.proj,
                     cb.cell, cb.cell_count, cb.cell_start, cb.sorted);
  hipLaunchKernelGGL(k_jump_tables, dim3(cb.nb, nclouds), dim3(64), 0, s, cb.nb, cb.ns,
                     cb.cell_start, cb.next_right, cb.next_left, cb.band_mask, cb.band_gap);
}

template <int L>
static void launch_nn_t(int64_t cap, int nclouds, int n, const float4* grid, const float* theta,
                        float max_sq, const CellBuffers& cb, float* grids, const float2* extra,
                        float* xgrids, bool xcd, bool blend, NnSeeds seeds, hipStream_t s) {
  const int nq = n * n, qpb = 256 / L;
  const int nqb = (nq + qpb - 1) / qpb;
  // the XCD-ordered grid needs whole groups of 8 clouds
  const bool by_xcd = xcd && nclouds % 8 == 0;
  const dim3 gq = by_xcd ? dim3((unsigned)(nqb * nclouds)) : dim3((unsigned)nqb, nclouds);
  const size_t lds = seeds.use_mask ? (size_t)cb.nb * ((cb.ns + 63) / 64) * 8 : 0;
  hipLaunchKernelGGL(k_nn<L>, gq, dim3(256), lds, s, grid, theta, n, cb.nb, cb.ns, max_sq,
                     cb.cell_start, cb.band_start, cb.next_right, cb.next_left, cb.sorted,
                     cb.pts, cap, grids, extra, xgrids, by_xcd ? nqb : 0, (int)blend,
                     seeds, cb.band_th, cb.band_gap);
}

int nn_auto_lanes(int nclouds) { return nclouds >= 64 ? 2 : 4; }

static void launch_nn_lanes(int lanes, int64_t cap, int nclouds, int n, const float4* grid,
                            const float* theta, float max_sq, const CellBuffers& cb, float* grids,
                            const float2* extra, float* xgrids, bool xcd, bool blend,
                            NnSeeds seeds, hipStream_t s) {
  switch (lanes) {
    case 1: launch_nn_t<1>(cap, nclouds, n, grid, theta, max_sq, cb, grids, extra, xgrids, xcd, blend, seeds, s); break;
    case 2: launch_nn_t<2>(cap, nclouds, n, grid, theta, max_sq, cb, grids, extra, xgrids, xcd, blend, seeds, s); break;
    case 8: launch_nn_t<8>(cap, nclouds, n, grid, theta, max_sq, cb, grids, extra, xgrids, xcd, blend, seeds, s); break;
    case 16: launch_nn_t<16>(cap, nclouds, n, grid, theta, max_sq, cb, grids, extra, xgrids, xcd, blend, seeds, s); break;
    default: launch_nn_t<4>(cap, nclouds, n, grid, theta, max_sq, cb, grids, extra, xgrids, xcd, blend, seeds, s); break;
  }
}

void launch_s2_nn(int64_t cap, int nclouds, int n, const float4* grid, const float* theta,
                  float max_sq, CellBuffers cb, float* grids, int lanes, hipStream_t s,
                  const float2* extra, float* xgrids, bool xcd, bool blend, int seed_mode) {
  if (lanes == 0) lanes = nn_auto_lanes(nclouds);
  NnSeeds sd{};
  sd.use_seeds = (seed_mode & 1) && cb.seed_pos;
  sd.use_mask = (seed_mode & 2) && cb.band_mask;
  sd.pos = cb.seed_pos;
  sd.mask = cb.band_mask;
  sd.F = seed_factor(cb.nb > cb.ns ? cb.nb : cb.ns);
  sd.cns = (cb.ns + sd.F - 1) / sd.F;
  if (sd.use_seeds) {
    const int cnb = (cb.nb + sd.F - 1) / sd.F;
    hipLaunchKernelGGL(k_cell_seeds, dim3(nclouds), dim3(1024), 2 * kSeedCells * sizeof(float4), s,
                       cb.nb, cb.ns, sd.F, cnb, sd.cns, cb.cell_start, cb.sorted, cap,
                       cb.seed_pos);
  }
  launch_nn_lanes(lanes, cap, nclouds, n, grid, theta, max_sq, cb, grids, extra, xgrids, xcd,
                  blend, sd, s);
}

void launch_fst(const float* sig, int S, int bw, const float* legendre, const int* leg_off_dev,
                const float2* tw_n, float2* F, float2* coef, int precision, hipStream_t s) {
  const int n = 2 * bw;
  const int th = ((bw + 63) / 64) * 64;
  hipLaunchKernelGGL(k_phi_dft, dim3(n, S), dim3(th), n * sizeof(float) + n * sizeof(float2), s,
                     sig, n, bw, tw_n, F);
  if (precision == 1) {
    launch_legendre_mfma(F, S, bw, legendre, leg_off_dev, coef, s);
    return;
  }
  hipLaunchKernelGGL(k_legendre, dim3(bw, S), dim3(th), n * sizeof(float2), s, F, n, bw, legendre,
                     leg_off_dev, coef);
}

void launch_fuse(ChannelSet cs, uint32_t count, const FusionRuns& runs, float2* out,
                 hipStream_t s, int nsets, int64_t set_stride) {
  hipLaunchKernelGGL(k_fuse, dim3((count + 255) / 256, nsets), dim3(256), 0, s, cs, set_stride,
                     count, runs, out);
}

}  // namespace phz
