// Serving queue of the knn_rms engine: requests, row-wise batch packing,
// batch slots and the scheduling loop of the serving thread.
//
// It makes no device call.  A Device class enqueues a batch on a slot,
// reports whether the batch has finished and copies its outputs out: the
// HIP side in serve.cpp, or a host-only one (tests/sanitize/serve_driver.cpp,
// which runs this queue under ThreadSanitizer and AddressSanitizer).
//
// Device interface:
//   using Shape;      per-request query shape, with an int member k_out
//   using SlotState;  per-slot device state
//   static bool same_shape(const Shape&, const Shape&);
//   double now();                                    milliseconds
//   void thread_init();                              on the serving thread
//   int64_t frame_limit(const Shape&);               largest batch, frames
//   int launch(SlotState&, const ServeBatch<Shape>&);   enqueue: 0 or error
//   int poll(SlotState&);                            0 done, 1 running, else error
//   int finish(SlotState&, const ServeBatch<Shape>&, std::vector<int64_t>* fallback);
//   void drain(SlotState&);                          after a failed launch
//   const std::string& error_text();                 serving thread's last failure
//   int64_t ref_frames();                            reference frames (stats)
// frame_limit is called with the queue lock held, the others without it,
// all on the serving thread.
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mdsctk_knn.h"

namespace mdsctk {

template <class Shape>
struct ServeRequest {
  int64_t id = 0;
  const float* xyz = nullptr;
  int64_t n = 0;
  Shape q;
  double* out_dist = nullptr;
  int32_t* out_idx = nullptr;
  int64_t next_row = 0;   // rows placed in batches so far
  int64_t rows_done = 0;  // rows whose outputs are written
  int status = MDSCTK_ERR_PENDING;
  std::string err;
  double t_submit = 0, t_admit = -1, t_done = -1;
  int n_batches = 0;
  int64_t fallback_rows = 0;
  int64_t batch_frames = 0;
};

template <class Shape>
struct ServeChunk {
  std::shared_ptr<ServeRequest<Shape>> r;
  int64_t row0 = 0;   // first row of the request in this batch
  int64_t nrows = 0;
  int64_t brow0 = 0;  // its first row in the batch
};

// One batch: rows of one or more requests of the same shape.
template <class Shape>
struct ServeBatch {
  std::vector<ServeChunk<Shape>> chunks;
  Shape q;
  int64_t n = 0;      // frames
  int64_t limit = 0;  // the frame limit it was packed under
};

template <class Device>
class ServeQueue {
 public:
  using Shape = typename Device::Shape;
  using Request = ServeRequest<Shape>;

  // slot_state[i]: the device state of slot i, owned by the caller and
  // alive until stop() has returned.
  ServeQueue(Device& dev, std::vector<typename Device::SlotState*> slot_state) : dev_(dev) {
    slots_.resize(slot_state.size());
    for (size_t i = 0; i < slots_.size(); ++i) slots_[i].dev = slot_state[i];
  }
  ~ServeQueue() { stop(); }

  void start() { th_ = std::thread([this] { run(); }); }

  // Ends the serving thread once the batches on the device have finished;
  // requests never started fail with MDSCTK_ERR_STATE.
  void stop() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    if (th_.joinable()) th_.join();
  }

  int submit(const float* xyz, int64_t n, const Shape& q, double* od, int32_t* oi, int64_t* id) {
    auto r = std::make_shared<Request>();
    r->xyz = xyz;
    r->n = n;
    r->q = q;
    r->out_dist = od;
    r->out_idx = oi;
    r->t_submit = dev_.now();
    {
      std::lock_guard<std::mutex> g(mu_);
      r->id = next_id_++;
      *id = r->id;
      live_[r->id] = r;
      if (n == 0 || q.k_out == 0) {
        r->status = q.k_out;
        r->t_admit = r->t_done = r->t_submit;
      } else {
        pending_.push_back(r);
      }
    }
    cv_work_.notify_one();
    return 0;
  }

  int poll(int64_t id, mdsctk_knn_request_info* info, bool wait, std::string* err) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = live_.find(id);
    if (it == live_.end()) {
      *err = "request id unknown (never submitted or already retired)";
      return MDSCTK_ERR_STATE;
    }
    std::shared_ptr<Request> r = it->second;
    if (wait) cv_done_.wait(lk, [&] { return r->status != MDSCTK_ERR_PENDING; });
    if (info) {
      info->status = r->status;
      info->n_batches = r->n_batches;
      info->n_fallback_rows = r->fallback_rows;
      info->t_submit_ms = r->t_submit;
      info->t_admit_ms = r->t_admit;
      info->t_done_ms = r->t_done;
      info->batch_frames = r->batch_frames;
    }
    if (r->status != MDSCTK_ERR_PENDING) {
      live_.erase(it);
      if (r->status < 0) *err = r->err;
    }
    return r->status;
  }

  // Queue-side statistics (the device adds its own fields).
  void stats(mdsctk_knn_serve_stats* out) {
    std::lock_guard<std::mutex> g(mu_);
    *out = st_;
    out->slots = (int32_t)slots_.size();
    out->busy_ms = busy_ms_ + (busy_since_ >= 0 ? dev_.now() - busy_since_ : 0.0);
  }

  bool idle() {
    std::lock_guard<std::mutex> g(mu_);
    if (!pending_.empty()) return false;
    for (auto& kv : live_)
      if (kv.second->status == MDSCTK_ERR_PENDING) return false;
    return true;
  }

 private:
  struct Slot {
    bool busy = false;
    ServeBatch<Shape> b;
    typename Device::SlotState* dev = nullptr;
  };

  // Requests of at most this many rows (left) go first into a batch.
  static constexpr int64_t kShortRows = 64;

  // Under mu_: pack pending requests of the oldest request's shape into the
  // slot's batch: first every whole pending request of at most kShortRows
  // rows that fits, so that a short request waits for the batches already
  // in flight and not also for the rows of larger requests queued ahead of
  // it; then the rest in arrival order (a request larger than the room left
  // is split by rows and its remainder goes first in the next batch).  A
  // row's outputs do not depend on its batch.  Returns false when nothing is
  // pending.
  bool form_batch(Slot& sl) {
    if (pending_.empty()) return false;
    ServeBatch<Shape>& b = sl.b;
    b.chunks.clear();
    b.q = pending_.front()->q;
    const int64_t limit = dev_.frame_limit(b.q);
    int64_t n = 0;
    const double now = dev_.now();
    auto place = [&](const std::shared_ptr<Request>& rp, int64_t take) {
      Request& r = *rp;
      b.chunks.push_back(ServeChunk<Shape>{rp, r.next_row, take, n});
      if (r.t_admit < 0) r.t_admit = now;
      ++r.n_batches;
      r.next_row += take;
      n += take;
    };
    for (auto it = pending_.begin(); it != pending_.end() && n < limit;) {
      Request& r = **it;
      const int64_t left = r.n - r.next_row;
      if (Device::same_shape(r.q, b.q) && left <= kShortRows && left <= limit - n) {
        place(*it, left);
        it = pending_.erase(it);
      } else {
        ++it;
      }
    }
    for (auto it = pending_.begin(); it != pending_.end() && n < limit;) {
      Request& r = **it;
      if (!Device::same_shape(r.q, b.q)) {
        ++it;
        continue;
      }
      const int64_t take = std::min(r.n - r.next_row, limit - n);
      place(*it, take);
      if (r.next_row == r.n)
        it = pending_.erase(it);
      else
        break;  // the batch is full; the rest of this request goes next
    }
    b.n = n;
    b.limit = limit;
    st_.max_batch_frames = std::max(st_.max_batch_frames, n);
    st_.slot_frames = (int32_t)limit;
    return true;
  }

  // Under mu_: record the outcome of a slot's batch.
  void retire(Slot& sl, int rc, const std::vector<int64_t>& fb_per_chunk) {
    const double now = dev_.now();
    ServeBatch<Shape>& b = sl.b;
    for (size_t j = 0; j < b.chunks.size(); ++j) {
      Request& r = *b.chunks[j].r;
      r.batch_frames = std::max(r.batch_frames, b.n);
      if (r.status != MDSCTK_ERR_PENDING) continue;  // already failed
      if (rc) {
        r.status = rc;
        r.err = dev_.error_text();
        r.t_done = now;
        if (r.next_row < r.n) {
          r.next_row = r.n;  // place no further rows
          pending_.erase(std::remove(pending_.begin(), pending_.end(), b.chunks[j].r),
                         pending_.end());
        }
        continue;
      }
      r.rows_done += b.chunks[j].nrows;
      r.fallback_rows += j < fb_per_chunk.size() ? fb_per_chunk[j] : 0;
      if (r.rows_done == r.n) {
        r.status = r.q.k_out;
        r.t_done = now;
        ++st_.requests_done;
      }
    }
    if (!rc) {
      ++st_.batches;
      st_.frames += b.n;
      st_.pairs += b.n * dev_.ref_frames();
      for (int64_t v : fb_per_chunk) st_.fallback_rows += v;
    }
    b.chunks.clear();
    sl.busy = false;
    cv_done_.notify_all();
  }

  void note_busy_change() {
    bool any = false;
    for (auto& sl : slots_) any |= sl.busy;
    if (any && busy_since_ < 0) busy_since_ = dev_.now();
    if (!any && busy_since_ >= 0) {
      busy_ms_ += dev_.now() - busy_since_;
      busy_since_ = -1;
    }
  }

  // Wait up to `us` microseconds for work while batches run.  A deadline
  // on the system clock: the timed wait on it is pthread_cond_timedwait,
  // which thread sanitizers follow (the steady-clock form is
  // pthread_cond_clockwait, which some do not, and then report the queue
  // lock as never released); a clock step only shortens or lengthens one
  // poll interval.
  void nap(std::unique_lock<std::mutex>& lk, int us) {
    cv_work_.wait_until(lk, std::chrono::system_clock::now() + std::chrono::microseconds(us));
  }

  void run() {
    dev_.thread_init();
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      bool any_busy = false;
      for (auto& sl : slots_) {
        if (!sl.busy) continue;
        lk.unlock();
        int rc = dev_.poll(*sl.dev);
        std::vector<int64_t> fb;
        if (rc == 0) rc = dev_.finish(*sl.dev, sl.b, &fb);
        lk.lock();
        if (rc == 1) {
          any_busy = true;
          continue;
        }
        retire(sl, rc, fb);
        note_busy_change();
      }
      if (stop_) {
        if (any_busy) {
          nap(lk, 50);
          continue;
        }
        // Requests never started fail; callers are told the context closed.
        for (auto& r : pending_) {
          if (r->status == MDSCTK_ERR_PENDING) {
            r->status = MDSCTK_ERR_STATE;
            r->err = "context destroyed or reference set replaced";
          }
        }
        pending_.clear();
        cv_done_.notify_all();
        return;
      }
      bool launched = false;
      for (auto& sl : slots_) {
        if (sl.busy || !form_batch(sl)) continue;
        sl.busy = true;
        note_busy_change();
        lk.unlock();
        const int rc = dev_.launch(*sl.dev, sl.b);
        if (rc) dev_.drain(*sl.dev);
        lk.lock();
        if (rc) {
          retire(sl, rc, {});
          note_busy_change();
        } else {
          any_busy = true;
          launched = true;
        }
      }
      if (launched) continue;
      if (any_busy)
        nap(lk, 20);
      else
        cv_work_.wait(lk, [&] { return stop_ || !pending_.empty(); });
    }
  }

  Device& dev_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<std::shared_ptr<Request>> pending_;
  std::unordered_map<int64_t, std::shared_ptr<Request>> live_;
  int64_t next_id_ = 1;
  bool stop_ = false;
  std::vector<Slot> slots_;
  mdsctk_knn_serve_stats st_ = {};
  double busy_ms_ = 0, busy_since_ = -1;
  std::thread th_;
};

}  // namespace mdsctk
