// The serving queue (mdsctk_amd/csrc/serve_queue.h) with a host-only device,
// for the sanitizer builds: client threads submit, poll and wait on requests
// of two shapes and many sizes while the serving thread packs them into
// batches over two slots; batches "run" for a few polls and then write each
// row's outputs from that row's input value, so every request's outputs can
// be checked whatever batches it was split over.  Then: injected launch
// failures (the requests of the failing batch fail with the device's error
// text, the others complete), and a stop with requests still queued (they
// fail with MDSCTK_ERR_STATE).
//
// Usage: serve_driver [quick]   prints "<n> failures" and exits non-zero on any.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../mdsctk_amd/csrc/serve_queue.h"

namespace {

std::atomic<int> g_failures{0};
std::mutex g_print;

void check(bool ok, const std::string& what) {
  if (ok) return;
  ++g_failures;
  std::lock_guard<std::mutex> g(g_print);
  std::fprintf(stderr, "FAIL: %s\n", what.c_str());
}

struct TestShape {
  int k_out = 0;
};

struct TestSlot {
  int countdown = 0;
};

// One "frame" is one float here; a row's outputs are (x + j, 7 row + j).
class HostDevice {
 public:
  using Shape = TestShape;
  using SlotState = TestSlot;

  static bool same_shape(const TestShape& a, const TestShape& b) { return a.k_out == b.k_out; }
  double now() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
  }
  void thread_init() {}
  int64_t frame_limit(const TestShape&) { return limit; }
  int launch(TestSlot& s, const mdsctk::ServeBatch<TestShape>& b) {
    ++launches;
    if (fail_every && launches % fail_every == 0) {
      err = "host device: injected launch failure";
      return MDSCTK_ERR_HIP;
    }
    check(b.n > 0 && b.n <= limit, "batch size within the frame limit");
    // Whole short requests (at most 64 rows left) are packed before the
    // rest: a chunk that completes its request with at most 64 rows comes
    // only from that first pass, so none follows any other chunk.
    bool other_seen = false;
    for (const auto& ch : b.chunks) {
      const bool short_whole = ch.nrows <= 64 && ch.row0 + ch.nrows == ch.r->n;
      if (short_whole) check(!other_seen, "short requests packed first");
      else other_seen = true;
    }
    s.countdown = 1 + (int)(b.n % 3);
    return 0;
  }
  int poll(TestSlot& s) { return --s.countdown > 0 ? 1 : 0; }
  int finish(TestSlot&, const mdsctk::ServeBatch<TestShape>& b, std::vector<int64_t>* fb) {
    fb->assign(b.chunks.size(), 0);
    int64_t brow = 0;
    for (const auto& ch : b.chunks) {
      check(ch.brow0 == brow, "chunks packed back to back");
      brow += ch.nrows;
      if (ch.r->status != MDSCTK_ERR_PENDING) continue;
      const int k = b.q.k_out;
      for (int64_t i = 0; i < ch.nrows; ++i) {
        const int64_t row = ch.row0 + i;
        for (int j = 0; j < k; ++j) {
          ch.r->out_dist[row * k + j] = ch.r->xyz[row] + j;
          ch.r->out_idx[row * k + j] = (int32_t)(7 * row + j);
        }
      }
    }
    check(brow == b.n, "batch rows add up");
    return 0;
  }
  void drain(TestSlot&) {}
  const std::string& error_text() { return err; }
  int64_t ref_frames() { return 1000; }

  int64_t limit = 100;
  int fail_every = 0;
  int launches = 0;  // serving thread only
  std::string err;   // serving thread only
};

using Queue = mdsctk::ServeQueue<HostDevice>;

struct Job {
  std::vector<float> x;
  std::vector<double> d;
  std::vector<int32_t> i;
  int k = 0;
  int64_t id = 0;
};

Job make_job(std::mt19937& rng, int max_rows) {
  Job j;
  const int n = (int)(rng() % (unsigned)(max_rows + 1));
  j.k = (rng() & 1) ? 3 : 5;
  j.x.resize(n);
  for (int r = 0; r < n; ++r) j.x[r] = (float)(rng() % 100000) * 0.5f;
  j.d.assign((size_t)n * j.k, -1.0);
  j.i.assign((size_t)n * j.k, -1);
  return j;
}

bool outputs_ok(const Job& j) {
  for (size_t r = 0; r < j.x.size(); ++r)
    for (int c = 0; c < j.k; ++c)
      if (j.d[r * j.k + c] != (double)j.x[r] + c || j.i[r * j.k + c] != (int32_t)(7 * r + c))
        return false;
  return true;
}

// Each client submits its jobs in bursts, then collects them by polling or
// waiting.
void client(Queue* q, int seed, int n_jobs, bool expect_failures, std::atomic<int>* n_failed) {
  std::mt19937 rng(seed);
  std::vector<Job> jobs;
  jobs.reserve(n_jobs);
  for (int b = 0; b < n_jobs; b += 5) {
    const size_t first = jobs.size();
    for (int t = b; t < std::min(n_jobs, b + 5); ++t) {
      jobs.push_back(make_job(rng, 250));
      Job& j = jobs.back();
      TestShape s;
      s.k_out = j.k;
      q->submit(j.x.data(), (int64_t)j.x.size(), s, j.d.data(), j.i.data(), &j.id);
    }
    for (size_t t = first; t < jobs.size(); ++t) {
      Job& j = jobs[t];
      mdsctk_knn_request_info info;
      std::string err;
      int rc;
      if (rng() & 1) {
        rc = q->poll(j.id, &info, true, &err);
      } else {
        while ((rc = q->poll(j.id, &info, false, &err)) == MDSCTK_ERR_PENDING)
          std::this_thread::yield();
      }
      if (rc == MDSCTK_ERR_HIP && expect_failures) {
        check(err == "host device: injected launch failure", "failed request carries the error text");
        ++*n_failed;
      } else {
        check(rc == j.k, "request completes with status k");
        check(outputs_ok(j), "request outputs");
        check(info.n_batches >= (j.x.empty() ? 0 : 1), "batches counted");
      }
      std::string e2;
      check(q->poll(j.id, nullptr, false, &e2) == MDSCTK_ERR_STATE, "a retired id is unknown");
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  const bool quick = argc > 1 && std::strcmp(argv[1], "quick") == 0;
  const int n_threads = 6, n_jobs = quick ? 40 : 200;

  {  // many clients, two shapes, requests split over batches
    HostDevice dev;
    TestSlot s0, s1;
    Queue q(dev, {&s0, &s1});
    q.start();
    std::atomic<int> n_failed{0};
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads; ++t) th.emplace_back(client, &q, 100 + t, n_jobs, false, &n_failed);
    for (auto& t : th) t.join();
    check(q.idle(), "queue idle after all requests");
    mdsctk_knn_serve_stats st;
    q.stats(&st);
    check(st.requests_done > 0 && st.slots == 2, "serve stats");
    check(st.max_batch_frames <= dev.limit, "max batch within the limit");
    q.stop();
  }
  {  // injected launch failures
    HostDevice dev;
    dev.fail_every = 7;
    TestSlot s0, s1;
    Queue q(dev, {&s0, &s1});
    q.start();
    std::atomic<int> n_failed{0};
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads; ++t) th.emplace_back(client, &q, 200 + t, n_jobs, true, &n_failed);
    for (auto& t : th) t.join();
    check(n_failed.load() > 0, "some requests failed");
    q.stop();
  }
  {  // stop with requests queued: each fails with MDSCTK_ERR_STATE or completed
    HostDevice dev;
    dev.limit = 10;
    TestSlot s0;
    Queue q(dev, {&s0});
    std::mt19937 rng(7);
    std::vector<Job> jobs;
    for (int t = 0; t < 50; ++t) jobs.push_back(make_job(rng, 40));
    for (auto& j : jobs) {
      TestShape s;
      s.k_out = j.k;
      q.submit(j.x.data(), (int64_t)j.x.size(), s, j.d.data(), j.i.data(), &j.id);
    }
    q.start();
    q.stop();
    int stopped = 0;
    for (auto& j : jobs) {
      std::string err;
      const int rc = q.poll(j.id, nullptr, true, &err);
      if (rc == MDSCTK_ERR_STATE) {
        ++stopped;
        check(err == "context destroyed or reference set replaced", "stopped request's error text");
      } else {
        check(rc == j.k && outputs_ok(j), "request finished before the stop");
      }
    }
    std::printf("stop: %d of %zu requests failed as stopped\n", stopped, jobs.size());
  }
  std::printf("%d failures\n", g_failures.load());
  return g_failures.load() ? 1 : 0;
}
