// kernels_peaks.hip -- ZScorePeakExtraction (z-score-peak-extraction.cc:
// 42-112) on the device: the smoothed z-score detector over a correlation
// vector, one wavefront per vector.
//
// The detector is a recurrence: an input that deviates from the mean of the
// window of the lag values before the last processed index is replaced by a
// blend with its predecessor, which changes later windows.  The wave
// therefore advances in speculative steps of 64 indices: every lane tests
// its index against the window it would see if no index of the step before
// it were replaced; up to and including the first lane that replaces its
// input the speculation is exact, so those indices are committed (the
// replaced value written to a ring of committed values in LDS) and the next
// step starts after it.  Steps without a replacement commit all 64 lanes,
// so the work is about n / 64 steps plus one per replaced input.
#include "kernels.hpp"

namespace phz {

namespace {

constexpr int kRing = 128;  // committed values [b - kRing, b); lag <= 64 needs b - lag

__global__ void __launch_bounds__(64) k_zscore_peaks(const float* __restrict__ corr, int64_t stride,
                                                     uint32_t n, int lag, double threshold,
                                                     double influence, double filter,
                                                     uint32_t* __restrict__ peaks,
                                                     int64_t max_peaks,
                                                     unsigned long long* __restrict__ counts) {
  __shared__ double ring[kRing];
  __shared__ double chunk[64];
  const int lane = threadIdx.x;
  const float* g = corr + blockIdx.x * stride;
  uint32_t* out = peaks + blockIdx.x * max_peaks;
  unsigned long long count = 0;
  if ((double)n <= (double)lag + 2.0) {  // z-score-peak-extraction.cc:57-58
    if (lane == 0) counts[blockIdx.x] = 0;
    return;
  }
  const double L = (double)lag, lagm1 = L - 1.0, inv_influence = 1.0 - influence;
  // committed window statistics: the initial window [0, lag) (:68-72)
  for (int j = lane; j <= lag; j += 64) ring[j % kRing] = (double)g[j];
  __syncthreads();
  auto window_stats = [&](uint32_t p, double* avg, double* dev) {
    // the lag committed values [p - lag, p), summed in index order
    double s = 0.0;
    for (int j = 0; j < lag; ++j) s += ring[(p - lag + j) % kRing];
    const double a = s / L;
    double q = 0.0;
    for (int j = 0; j < lag; ++j) {
      const double d = ring[(p - lag + j) % kRing] - a;
      q += d * d;
    }
    *avg = a;
    *dev = sqrt(q / lagm1);
  };
  double cavg, cdev;  // the statistics the next processed index sees
  window_stats((uint32_t)lag, &cavg, &cdev);
  uint32_t b = (uint32_t)lag + 1u;  // :76
  while (b < n) {
    const uint32_t i = b + lane;
    const double x = i < n ? (double)g[i] : 0.0;
    chunk[lane] = x;
    const bool proc = i < n && x > filter;  // :77-79
    const unsigned long long pm = __ballot(proc);
    __syncthreads();
    // the last processed lane before this one: its window is this lane's
    const unsigned long long below = pm & ((1ull << lane) - 1ull);
    double avg = cavg, dev = cdev;
    if (below) {
      const uint32_t p = b + (63 - __clzll((long long)below));
      double s = 0.0;
      for (int j = 0; j < lag; ++j) {
        const uint32_t v = p - lag + j;
        s += v >= b ? chunk[v - b] : ring[v % kRing];
      }
      avg = s / L;
      double q = 0.0;
      for (int j = 0; j < lag; ++j) {
        const uint32_t v = p - lag + j;
        const double d = (v >= b ? chunk[v - b] : ring[v % kRing]) - avg;
        q += d * d;
      }
      dev = sqrt(q / lagm1);
    }
    const bool flag = proc && fabs(x - avg) > threshold * dev;  // :81
    const unsigned long long fm = __ballot(flag);
    const int t0 = fm ? __ffsll((long long)fm) - 1 : 64;
    const int last = t0 < 64 ? t0 : 63;
    const uint32_t end = b + (uint32_t)last;  // last committed index
    // commit lanes <= last: raw values, then the flagged lane's blend
    if (lane < t0 && i < n) ring[i % kRing] = x;
    __syncthreads();
    if (t0 < 64 && lane == t0) {
      if (x >= avg) {  // :82-83
        if ((int64_t)count < max_peaks) out[count] = i;
      }
      ring[i % kRing] = influence * x + inv_influence * ring[(i - 1) % kRing];  // :86
    }
    if (t0 < 64) count += __shfl(x >= avg ? 1 : 0, t0, 64);
    __syncthreads();
    // statistics after the last processed committed index (:107-110)
    const unsigned long long done = t0 < 64 ? ((2ull << t0) - 1ull) : ~0ull;
    const unsigned long long pc = pm & done;
    if (pc) window_stats(b + (63 - __clzll((long long)pc)), &cavg, &cdev);
    b = end + 1u;
    if (end >= n - 1) break;
    __syncthreads();
  }
  if (lane == 0) counts[blockIdx.x] = count;
}

}  // namespace

void launch_zscore_peaks(const float* corr, int64_t stride, int nvec, uint32_t n, int lag,
                         double threshold, double influence, double filter, uint32_t* peaks,
                         int64_t max_peaks, unsigned long long* counts, hipStream_t s) {
  hipLaunchKernelGGL(k_zscore_peaks, dim3(nvec), dim3(64), 0, s, corr, stride, n, lag, threshold,
                     influence, filter, peaks, max_peaks, counts);
}

}  // namespace phz
