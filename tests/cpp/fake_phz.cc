// fake_phz.cc -- a host-only stand-in for the phaser_amd entry points that
// the request loop calls, so that its scheduling (batching, admission by
// memory budget, retirement, overlap) is tested without a GPU.  Each
// "registration" takes a fixed time and returns indices derived from the
// clouds' contents, so a test can check which target each source met.
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>

#include "phaser_amd.h"

struct phz_ctx {
  phz_config cfg;
  size_t bytes;
};

namespace fake {
std::atomic<int> in_flight{0}, max_in_flight{0};
std::atomic<long long> live_bytes{0}, peak_live_bytes{0};
std::atomic<int> creates{0}, batches{0}, pairs{0}, shared_batches{0};
std::atomic<int> max_batch_seen{0};
int fail_create_bandwidth = -1;  // phz_create fails for this bandwidth
double us_per_batch = 1500, us_per_pair = 200;

uint32_t checksum(const float* p, int64_t n) {
  uint32_t h = 2166136261u;
  for (int64_t i = 0; i < 4 * n; ++i) {
    uint32_t b;
    std::memcpy(&b, &p[i], 4);
    h = (h ^ b) * 16777619u;
  }
  return h;
}
}  // namespace fake

namespace {
thread_local std::string err;
void bump_max(std::atomic<int>& m, int v) {
  int cur = m.load();
  while (v > cur && !m.compare_exchange_weak(cur, v)) {
  }
}
void bump_max64(std::atomic<long long>& m, long long v) {
  long long cur = m.load();
  while (v > cur && !m.compare_exchange_weak(cur, v)) {
  }
}
bool factors235(int n) {
  for (int f : {2, 3, 5})
    while (n % f == 0) n /= f;
  return n == 1;
}
}  // namespace

extern "C" {

void phz_default_config(phz_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->bandwidth = 64;
  c->n_voxels = 128;
  c->discretize_lower = -50;
  c->discretize_upper = 50;
  c->low_pass_upper = 100000;
  c->max_points = 262144;
  c->max_batch = 1;
}

const char* phz_last_error(void) { return err.c_str(); }

int phz_estimate_memory(const phz_config* c, size_t* bytes) {
  if (c->bandwidth < 2 || c->bandwidth > 128 || !factors235(2 * c->bandwidth)) {
    err = "2*bandwidth must factor into 2, 3 and 5";
    return -1;
  }
  if (c->n_voxels < 2 || c->n_voxels > 256 || !factors235(c->n_voxels)) {
    err = "n_voxels must factor into 2, 3 and 5";
    return -1;
  }
  *bytes = (size_t)c->bandwidth * c->n_voxels * 1024;
  return 0;
}

int phz_create(const phz_config* c, phz_ctx** out) {
  size_t b = 0;
  if (int rc = phz_estimate_memory(c, &b)) return rc;
  std::this_thread::sleep_for(std::chrono::milliseconds(3));
  if (c->bandwidth == fake::fail_create_bandwidth) {
    err = "hipMalloc failed: out of memory";
    return -2;
  }
  bump_max64(fake::peak_live_bytes, fake::live_bytes += (long long)b);
  ++fake::creates;
  *out = new phz_ctx{*c, b};
  return 0;
}

int phz_autotune(phz_ctx* c, int repeats, phz_tuning* out) {
  (void)c;
  (void)repeats;
  if (out) std::memset(out, 0, sizeof(*out));
  return 0;
}

int phz_set_graphs(phz_ctx* c, int enable) {
  (void)c;
  (void)enable;
  return 0;
}

void phz_destroy(phz_ctx* c) {
  if (!c) return;
  fake::live_bytes -= (long long)c->bytes;
  delete c;
}

int phz_register_pairs(phz_ctx* c, int n, int shared, const float* const* targets,
                       const int64_t* Pt, const float* const* sources, const int64_t* Ps,
                       phz_result* results) {
  if (n < 1 || n > c->cfg.max_batch) {
    err = "pair count exceeds max_batch";
    return -1;
  }
  bump_max(fake::max_in_flight, ++fake::in_flight);
  bump_max(fake::max_batch_seen, n);
  ++fake::batches;
  fake::pairs += n;
  if (shared) ++fake::shared_batches;
  std::this_thread::sleep_for(
      std::chrono::microseconds((long long)(fake::us_per_batch + fake::us_per_pair * n)));
  for (int p = 0; p < n; ++p) {
    const int t = shared ? 0 : p;
    std::memset(&results[p], 0, sizeof(phz_result));
    results[p].rot_index = fake::checksum(sources[p], Ps[p]) ^ (uint32_t)c->cfg.bandwidth;
    results[p].trans_index = fake::checksum(targets[t], Pt[t]) ^ (uint32_t)c->cfg.n_voxels;
  }
  --fake::in_flight;
  return 0;
}
}
