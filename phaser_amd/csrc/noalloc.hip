// noalloc.hip -- SpatialCorrelation::correlateSignals (spatial-correlation.cc:
// 163-206, one channel, no padding) on a file-scope device array instead of
// allocated buffers: phz_spatial_correlate_single_noalloc.
//
// `scratch` holds, per device, the FFT twiddle tables of every n_voxels up to
// kNoallocMaxN (written once at set-up), the two r2c spectra, the cross
// spectrum / inverse workspace and the argmax keys.  Host code never passes
// the array's name to a kernel or a copy: the name is a host-side stand-in.
// Set-up (phz_create, once per device, with that device current) takes the
// device address with hipGetSymbolAddress and keeps it per device; the entry
// point uses only that address.  Every call on one device shares that
// device's array (the last call's contents stay): calls that overlap on one
// device race, so callers serialise them (one stream, or their own lock).
#include <mutex>
#include <vector>

#include "../../include/phaser_amd.h"
#include "kernels.hpp"

namespace phz {

namespace {

constexpr int kNoallocMaxN = 128;
constexpr int64_t kAlign2 = 32;  // 256 bytes in float2 units
constexpr int64_t round_up(int64_t v) { return (v + kAlign2 - 1) / kAlign2 * kAlign2; }
// twiddles of every N in [1, kNoallocMaxN] back to back: N at N(N-1)/2
constexpr int64_t tw_offset(int N) { return (int64_t)N * (N - 1) / 2; }
constexpr int64_t kTwEntries = round_up(tw_offset(kNoallocMaxN + 1));
constexpr int64_t kHMax = (int64_t)kNoallocMaxN * kNoallocMaxN * (kNoallocMaxN / 2 + 1);
constexpr int64_t kSpecOff = kTwEntries;
constexpr int64_t kWorkOff = kSpecOff + round_up(2 * kHMax);
constexpr int64_t kKeysOff = kWorkOff + round_up(kHMax);
constexpr int64_t kKeyEntries = 1 + (kNoallocMaxN * kNoallocMaxN + 15) / 16;  // spatial_c2r_blocks
constexpr int64_t kScratchEntries = kKeysOff + round_up(kKeyEntries);  // float2 (8-byte) units

static_assert(sizeof(unsigned long long) == sizeof(float2), "keys share the float2 units");

// 25.6 MB per device (kScratchEntries * 8 bytes), held until the process ends
static __device__ __align__(256) float2 scratch[kScratchEntries];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float2* g_ptr[kMaxDevices] = {};  // per device: the address of `scratch`

}  // namespace

size_t noalloc_scratch_bytes() { return sizeof(float2) * (size_t)kScratchEntries; }
int noalloc_max_n() { return kNoallocMaxN; }

// Once per device, with that device current (phz_create calls it after
// hipSetDevice): the device address of `scratch` and the twiddle tables.
int noalloc_setup(int device, hipStream_t s) {
  PHZ_CHECK(device >= 0 && device < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[device]) return 0;
  void* p = nullptr;
  PHZ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  PHZ_CHECK(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
  std::vector<float2> tw((size_t)kTwEntries, make_float2(0.f, 0.f));
  for (int N = 1; N <= kNoallocMaxN; ++N) {
    const std::vector<float> t = twiddles(N);  // the same table as the allocating path
    for (int e = 0; e < N; ++e)
      tw[(size_t)(tw_offset(N) + e)] = make_float2(t[2 * e], t[2 * e + 1]);
  }
  PHZ_HIP(hipMemcpyAsync(p, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice, s));
  PHZ_HIP(hipStreamSynchronize(s));
  g_ptr[device] = (float2*)p;
  return 0;
}

// The same kernels as phz_spatial_correlate_single, on the device's scratch
// array: r2c of f and g, cross spectrum and inverse z, inverse x, c2r and the
// first-max argmax.  No allocation, no host-side table.
int spatial_single_noalloc(int device, int N, const float* d_f, const float* d_g, float* d_corr,
                           uint32_t* argmax, hipStream_t s) {
  PHZ_CHECK(device >= 0 && device < kMaxDevices, "device ordinal out of range");
  float2* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    base = g_ptr[device];
  }
  PHZ_CHECK(base != nullptr, "the scratch array of this device is not set up (phz_create on it first)");
  PHZ_CHECK(N >= 2 && N <= kNoallocMaxN, "n_voxels must be in [2, 128] for the no-allocation entry point");
  const FftPlan plan = make_plan(N);
  PHZ_CHECK(d_f && d_g, "null argument");
  const float2* tw = base + tw_offset(N);
  float2* spec = base + kSpecOff;
  float2* work = base + kWorkOff;
  unsigned long long* keys = (unsigned long long*)(base + kKeysOff);
  const int64_t H = (int64_t)N * N * (N / 2 + 1);
  launch_spatial_forward_two(N, d_f, d_g, plan, tw, spec, s);
  ChannelSet F{}, G{};
  F.C = G.C = 1;
  F.ch[0] = spec;
  G.ch[0] = spec + H;
  FusionRuns none{};
  launch_spatial_inverse(N, 1, F, 0, G, 0, none, false, plan, tw, work, d_corr, keys + 1, keys,
                         SpatialTiles{}, s);
  PHZ_HIP(hipGetLastError());
  if (argmax) {
    unsigned long long key = 0;
    PHZ_HIP(hipMemcpyAsync(&key, keys, sizeof(key), hipMemcpyDeviceToHost, s));
    PHZ_HIP(hipStreamSynchronize(s));
    *argmax = key_index(key);
  }
  return 0;
}

}  // namespace phz
