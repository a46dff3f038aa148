// serve_sched_test.cc -- scheduling tests of the request loop
// (phaser_amd/csrc/host/phaser_serve.cc) against the host-only fake backend
// in fake_phz.cc.  Prints one line per scenario; exits 1 on the first
// failed check.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "phaser_serve.h"

namespace fake {
extern std::atomic<int> in_flight, max_in_flight;
extern std::atomic<long long> live_bytes, peak_live_bytes;
extern std::atomic<int> creates, batches, pairs, shared_batches, max_batch_seen;
extern int fail_create_bandwidth;
extern double us_per_batch, us_per_pair;
uint32_t checksum(const float* p, int64_t n);
}  // namespace fake

#define CHECK(cond, ...)                                    \
  do {                                                      \
    if (!(cond)) {                                          \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);      \
      std::printf(__VA_ARGS__);                             \
      std::printf("\n");                                    \
      std::exit(1);                                         \
    }                                                       \
  } while (0)

namespace {

struct Cloud {
  std::vector<float> v;
  int64_t n() const { return (int64_t)v.size() / 4; }
};

Cloud make_cloud(std::mt19937& rng, int npts) {
  Cloud c;
  c.v.resize(4 * (size_t)npts);
  std::uniform_real_distribution<float> u(-40.f, 40.f);
  for (float& x : c.v) x = u(rng);
  return c;
}

void reset_fake() {
  fake::max_in_flight = 0;
  fake::peak_live_bytes = fake::live_bytes.load();
  fake::creates = fake::batches = fake::pairs = fake::shared_batches = 0;
  fake::max_batch_seen = 0;
  fake::fail_create_bandwidth = -1;
}

struct Shape {
  int B, N;
};

void expect_result(const phs_response& r, const Cloud& tgt, const Cloud& src, Shape sh) {
  CHECK(r.status == 0, "ticket %llu failed: %s", (unsigned long long)r.ticket, r.error);
  const uint32_t want_rot = fake::checksum(src.v.data(), src.n()) ^ (uint32_t)sh.B;
  const uint32_t want_tr = fake::checksum(tgt.v.data(), tgt.n()) ^ (uint32_t)sh.N;
  CHECK(r.result.rot_index == want_rot, "ticket %llu: source mismatch",
        (unsigned long long)r.ticket);
  CHECK(r.result.trans_index == want_tr, "ticket %llu: target mismatch",
        (unsigned long long)r.ticket);
  CHECK(r.latency_ms >= r.service_ms && r.queue_ms >= 0, "timings");
}

// Mixed shapes arriving while batches run: every result belongs to its own
// pair, batches hold several pairs and contexts overlap.
void mixed_arrivals() {
  reset_fake();
  std::mt19937 rng(1);
  const Shape shapes[3] = {{32, 64}, {64, 128}, {96, 192}};
  const int R = 90;
  std::vector<Cloud> tg, sr;
  std::vector<int> which;
  for (int i = 0; i < R; ++i) {
    tg.push_back(make_cloud(rng, 50 + i));
    sr.push_back(make_cloud(rng, 70 + i));
    which.push_back((int)(rng() % 3));
  }
  phs_config cfg;
  phs_default_config(&cfg);
  cfg.max_batch = 8;
  cfg.contexts_per_shape = 2;
  phs_server* s = nullptr;
  CHECK(phs_create(&cfg, &s) == 0, "create: %s", phs_last_error());
  std::vector<uint64_t> tickets(R);
  for (int i = 0; i < R; ++i) {
    phs_request q{};
    q.bandwidth = shapes[which[i]].B;
    q.n_voxels = shapes[which[i]].N;
    q.target = tg[i].v.data();
    q.Pt = tg[i].n();
    q.source = sr[i].v.data();
    q.Ps = sr[i].n();
    q.tag = 1000 + i;
    CHECK(phs_submit(s, &q, &tickets[i]) == 0, "submit: %s", phs_last_error());
    if (i % 7 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 2500));
  }
  for (int i = 0; i < R; ++i) {
    phs_response r;
    CHECK(phs_wait(s, tickets[i], &r) == 0, "wait: %s", phs_last_error());
    CHECK(r.tag == (uint64_t)(1000 + i), "tag");
    CHECK(r.shared_target == 0, "no keys were given");
    expect_result(r, tg[i], sr[i], shapes[which[i]]);
  }
  phs_stats st;
  CHECK(phs_get_stats(s, &st) == 0, "stats");
  CHECK(st.completed == (uint64_t)R && st.failed == 0, "completed %llu",
        (unsigned long long)st.completed);
  CHECK(st.pairs_in_batches == (uint64_t)R, "pairs");
  CHECK(st.batches < (uint64_t)R, "continuous batching formed no multi-pair batch");
  CHECK(fake::max_batch_seen <= 8, "batch above max_batch");
  CHECK(fake::max_in_flight >= 2, "contexts never overlapped");
  CHECK(st.contexts_created <= 6, "more than two contexts per shape");
  phs_response again;
  CHECK(phs_wait(s, tickets[0], &again) == -1, "a response is returned once");
  phs_destroy(s);
  CHECK(fake::live_bytes == 0, "contexts leaked");
  std::printf("mixed_arrivals ok: %d requests, %llu batches, max %d in flight, %llu contexts\n", R,
              (unsigned long long)st.batches, fake::max_in_flight.load(),
              (unsigned long long)st.contexts_created);
}

// A budget that holds one context: shapes take turns, nothing is lost and
// the budget is never exceeded.
void tight_budget() {
  reset_fake();
  std::mt19937 rng(2);
  const Shape shapes[3] = {{32, 64}, {48, 96}, {64, 128}};
  const int R = 40;
  std::vector<Cloud> tg, sr;
  for (int i = 0; i < R; ++i) {
    tg.push_back(make_cloud(rng, 30));
    sr.push_back(make_cloud(rng, 31));
  }
  phs_config cfg;
  phs_default_config(&cfg);
  cfg.max_batch = 4;
  cfg.contexts_per_shape = 2;
  cfg.hbm_budget = (uint64_t)64 * 128 * 1024;  // the largest single context
  phs_server* s = nullptr;
  CHECK(phs_create(&cfg, &s) == 0, "create");
  std::vector<uint64_t> tickets(R);
  for (int i = 0; i < R; ++i) {
    const Shape sh = shapes[i % 3];
    phs_request q{};
    q.bandwidth = sh.B;
    q.n_voxels = sh.N;
    q.target = tg[i].v.data();
    q.Pt = tg[i].n();
    q.source = sr[i].v.data();
    q.Ps = sr[i].n();
    CHECK(phs_submit(s, &q, &tickets[i]) == 0, "submit: %s", phs_last_error());
  }
  for (int i = 0; i < R; ++i) {
    phs_response r;
    CHECK(phs_wait(s, tickets[i], &r) == 0, "wait");
    expect_result(r, tg[i], sr[i], shapes[i % 3]);
  }
  phs_stats st;
  phs_get_stats(s, &st);
  CHECK(st.bytes_peak <= cfg.hbm_budget, "reserved bytes above budget");
  CHECK(fake::peak_live_bytes <= (long long)cfg.hbm_budget, "live bytes above budget");
  CHECK(st.contexts_released >= 2, "no context was released");
  phs_destroy(s);
  CHECK(fake::live_bytes == 0, "contexts leaked");
  std::printf("tight_budget ok: %llu contexts created, %llu released, peak %llu bytes\n",
              (unsigned long long)st.contexts_created, (unsigned long long)st.contexts_released,
              (unsigned long long)st.bytes_peak);
}

// One target against many sources: requests with equal keys form shared
// batches; others keep their own targets.
void shared_target() {
  reset_fake();
  fake::us_per_batch = 4000;  // a slow first batch lets the rest queue up
  std::mt19937 rng(3);
  const Shape sh{64, 128};
  Cloud target = make_cloud(rng, 200);
  const int R = 24;
  std::vector<Cloud> sr, own;
  for (int i = 0; i < R; ++i) {
    sr.push_back(make_cloud(rng, 100 + i));
    own.push_back(make_cloud(rng, 90 + i));
  }
  phs_config cfg;
  phs_default_config(&cfg);
  cfg.max_batch = 16;
  cfg.contexts_per_shape = 1;
  phs_server* s = nullptr;
  CHECK(phs_create(&cfg, &s) == 0, "create");
  std::vector<uint64_t> tickets(R);
  for (int i = 0; i < R; ++i) {
    const bool keyed = i % 4 != 0;  // the oldest request has its own target
    phs_request q{};
    q.bandwidth = sh.B;
    q.n_voxels = sh.N;
    q.target_key = keyed ? 77 : 0;
    q.target = keyed ? target.v.data() : own[i].v.data();
    q.Pt = keyed ? target.n() : own[i].n();
    q.source = sr[i].v.data();
    q.Ps = sr[i].n();
    CHECK(phs_submit(s, &q, &tickets[i]) == 0, "submit");
  }
  int shared_seen = 0;
  for (int i = 0; i < R; ++i) {
    phs_response r;
    CHECK(phs_wait(s, tickets[i], &r) == 0, "wait");
    const bool keyed = i % 4 != 0;
    expect_result(r, keyed ? target : own[i], sr[i], sh);
    if (!keyed) CHECK(r.shared_target == 0, "an unkeyed request was served in a shared batch");
    shared_seen += r.shared_target;
  }
  CHECK(shared_seen >= 12 && fake::shared_batches >= 1, "keyed requests were not grouped");
  phs_destroy(s);
  fake::us_per_batch = 1500;
  std::printf("shared_target ok: %d requests served with a shared target, %d shared batches\n",
              shared_seen, fake::shared_batches.load());
}

// Rejections at submit, failed context creation, poll and shutdown.
void errors_and_poll() {
  reset_fake();
  std::mt19937 rng(4);
  Cloud a = make_cloud(rng, 10), b = make_cloud(rng, 12);
  phs_config cfg;
  phs_default_config(&cfg);
  cfg.max_batch = 4;
  cfg.max_points = 1000;
  cfg.hbm_budget = (uint64_t)96 * 192 * 1024 - 1;
  phs_server* s = nullptr;
  CHECK(phs_create(&cfg, &s) == 0, "create");
  phs_request q{};
  q.bandwidth = 7;  // 2B = 14 does not factor into 2, 3, 5
  q.n_voxels = 64;
  q.target = a.v.data();
  q.Pt = a.n();
  q.source = b.v.data();
  q.Ps = b.n();
  uint64_t t;
  CHECK(phs_submit(s, &q, &t) == -1, "bad bandwidth accepted");
  q.bandwidth = 32;
  q.Ps = 1001;
  CHECK(phs_submit(s, &q, &t) == -3, "oversized cloud accepted");
  q.Ps = b.n();
  q.bandwidth = 96;
  q.n_voxels = 192;
  CHECK(phs_submit(s, &q, &t) == -3, "context above budget accepted");
  CHECK(std::strstr(phs_last_error(), "hbm_budget") != nullptr, "message: %s", phs_last_error());
  CHECK(phs_wait(s, 12345, nullptr) == -1, "null response accepted");
  // a shape whose context cannot be built fails its requests
  fake::fail_create_bandwidth = 40;
  q.bandwidth = 40;
  q.n_voxels = 80;
  uint64_t tf;
  CHECK(phs_submit(s, &q, &tf) == 0, "submit");
  phs_response rf;
  CHECK(phs_wait(s, tf, &rf) == 0, "wait");
  CHECK(rf.status == -2 && std::strstr(rf.error, "out of memory"), "status %d %s", rf.status,
        rf.error);
  fake::fail_create_bandwidth = -1;
  // poll collects everything exactly once
  q.bandwidth = 32;
  q.n_voxels = 64;
  const int R = 10;
  for (int i = 0; i < R; ++i) CHECK(phs_submit(s, &q, &t) == 0, "submit");
  int got = 0;
  phs_response buf[4];
  for (int iter = 0; iter < 1000 && got < R; ++iter) {
    int k = 0;
    CHECK(phs_poll(s, buf, 4, &k, 100) == 0, "poll");
    for (int i = 0; i < k; ++i) expect_result(buf[i], a, b, Shape{32, 64});
    got += k;
  }
  CHECK(got == R, "poll returned %d of %d", got, R);
  int k = -1;
  CHECK(phs_poll(s, buf, 4, &k, -1) == 0 && k == 0, "poll with nothing outstanding must return");
  // shutdown finishes queued work
  const int before = fake::pairs;
  for (int i = 0; i < R; ++i) CHECK(phs_submit(s, &q, &t) == 0, "submit");
  phs_destroy(s);
  CHECK(fake::pairs - before == R, "destroy dropped queued requests");
  CHECK(fake::live_bytes == 0, "contexts leaked");
  std::printf("errors_and_poll ok\n");
}

}  // namespace

int main() {
  mixed_arrivals();
  tight_budget();
  shared_target();
  errors_and_poll();
  std::printf("all ok\n");
  return 0;
}
