// Serving loop of the knn_rms engine: asynchronous out-of-sample queries
// against the resident reference set with continuous batching.
//
// One host thread per context owns `serve_slots` batch slots, each with its
// own HIP stream and grow-only device workspace.  Whenever a slot is free and
// requests are pending, the pending requests (arrival order, same k and
// flags as the oldest one) are packed row-wise into one batch of at most
// `slot_frames` frames and launched as one fused sequence: H2D of the
// frames, prep, contraction, merge, D2H of the outputs into pinned buffers.
// While it runs, later arrivals fill the other slot(s), so their copies,
// prep and contraction overlap the running batch on the device.  A request
// larger than a batch is split by rows over consecutive batches.
//
// Output values are independent of batch composition: each is the FP64
// re-score of a certified candidate, sorted by (distance, index), computed by
// the same per-pair code whatever rows share the batch.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>

#include "engine.h"

namespace mdsctk {

namespace {

struct Request {
  int64_t id = 0;
  const float* xyz = nullptr;
  int64_t n = 0;
  QueryShape q;
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

struct Chunk {
  std::shared_ptr<Request> r;
  int64_t row0 = 0;   // first row of the request in this batch
  int64_t nrows = 0;
  int64_t brow0 = 0;  // its first row in the batch
};

struct ServeSlot {
  Slot s;
  bool busy = false;
  std::vector<Chunk> chunks;
  QueryShape q;
  int64_t n = 0;
  double* d_dist = nullptr;
  int* d_idx = nullptr;
  int* d_fb = nullptr;
  // pinned host mirrors of the outputs
  double* h_dist = nullptr;
  int* h_idx = nullptr;
  int* h_fb = nullptr;
  size_t h_elems = 0, h_rows = 0;
  hipEvent_t done = nullptr;
  hipGraphExec_t exec = nullptr;  // the slot's batch kernels, updated per batch
  // the captured graph exec was last instantiated or updated from; kept
  // until a later capture has replaced it in exec (and that exec has been
  // launched), so nothing the executable may refer to is released while it
  // can still run
  hipGraph_t graph = nullptr;
  int64_t limit = 0;  // frame limit of the batch in the slot
  // shape the workspace was last sized for
  int64_t sized_limit = 0;
  int sized_k_out = -1, sized_k1 = -1;
};

bool same_shape(const QueryShape& a, const QueryShape& b) {
  return a.k_out == b.k_out && a.k1 == b.k1 && a.rank_off == b.rank_off && a.nofit == b.nofit;
}

}  // namespace

class Server {
 public:
  explicit Server(mdsctk_knn_ctx* c) : c_(c) {}

  int start() {
    slots_.resize(c_->serve_slots);
    for (auto& sl : slots_) {
      int rc = slot_init(c_, sl.s);
      if (rc) return rc;
      if (hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
        return fail(c_, MDSCTK_ERR_HIP, "serve: event creation failed");
    }
    {
      std::lock_guard<std::mutex> g(c_->mem_mu);
      avail_ = c_->budget ? (c_->budget > c_->held ? c_->budget - c_->held : 0) : 0;
    }
    th_ = std::thread([this] { run(); });
    return 0;
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    if (th_.joinable()) th_.join();
    for (auto& sl : slots_) {
      if (sl.s.st) (void)hipStreamSynchronize(sl.s.st);
      if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
      if (sl.graph) (void)hipGraphDestroy(sl.graph);
      slot_release(c_, sl.s);
      if (sl.done) (void)hipEventDestroy(sl.done);
      free_host(sl);
    }
  }

  int submit(const float* xyz, int64_t n, const QueryShape& q, double* od, int32_t* oi,
             int64_t* id) {
    auto r = std::make_shared<Request>();
    r->xyz = xyz;
    r->n = n;
    r->q = q;
    r->out_dist = od;
    r->out_idx = oi;
    r->t_submit = clock_ms(c_);
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

  void stats(mdsctk_knn_serve_stats* out) {
    std::lock_guard<std::mutex> g(mu_);
    *out = st_;
    out->slots = (int32_t)slots_.size();
    out->busy_ms = busy_ms_ + (busy_since_ >= 0 ? clock_ms(c_) - busy_since_ : 0.0);
    out->graph_launches = graph_launches_.load();
    out->graph_builds = graph_builds_.load();
  }

  bool idle() {
    std::lock_guard<std::mutex> g(mu_);
    if (!pending_.empty()) return false;
    for (auto& kv : live_)
      if (kv.second->status == MDSCTK_ERR_PENDING) return false;
    return true;
  }

 private:
  // Largest list workspace (entries, list headers) any batch of up to
  // `limit` frames needs: the phase count grows as batches shrink, so take
  // the maximum over batch sizes (it changes only at 128-frame steps).
  void list_need(int64_t limit, const QueryShape& q, size_t* entries, size_t* headers) const {
    *entries = *headers = 0;
    const size_t cap = (size_t)list_capacity_for(q.k1);
    for (int64_t n = std::min<int64_t>(limit, 128);; n = std::min<int64_t>(limit, n + 128)) {
      const size_t e = fused_list_entries(c_, n, q.k1);
      *entries = std::max(*entries, e);
      *headers = std::max(*headers, e / cap);
      if (n == limit) break;
    }
  }

  // Device bytes one slot holds for batches of up to `limit` frames.
  size_t slot_bytes(int64_t limit, const QueryShape& q) const {
    const double natoms = c_->ref.natoms;
    const double frames = (double)((limit + 127) / 128 * 128);
    // xyz + operand (12 B per atom each) + per-frame terms + outputs
    double b = frames * (24.0 * natoms + 64 + 12.0 * q.k_out);
    if (q.fused) {
      size_t ent, hdr;
      list_need(limit, q, &ent, &hdr);
      b += 8.0 * ent + 8.0 * hdr;
    }
    return (size_t)b;
  }

  // Largest batch (frames) whose workspace fits, for every slot at once,
  // in the HBM budget left when serving started.
  int64_t frame_limit(const QueryShape& q) const {
    int64_t f = std::max<int64_t>(1, c_->serve_max_frames);
    if (!avail_) return f;
    while (f > 32 && (double)slot_bytes(f, q) * slots_.size() > 0.9 * (double)avail_) f /= 2;
    return f;
  }

  // Size the slot's buffers once for its largest batch of this shape, so
  // that steady-state batches allocate nothing.
  int presize(ServeSlot& sl, int64_t limit) {
    Slot& s = sl.s;
    const QueryShape& q = sl.q;
    int rc;
    if ((rc = slot_frames(c_, s, limit))) return rc;
    if (q.fused) {
      size_t ent, hdr;
      list_need(limit, q, &ent, &hdr);
      if ((rc = ensure(c_, &s.ws_lists, &s.ws_lists_bytes, ent * 8, "candidate lists"))) return rc;
      const size_t small = hdr * 8 + (size_t)limit * 12 + 256;
      if ((rc = ensure(c_, &s.ws_small, &s.ws_small_bytes, small, "list headers"))) return rc;
    }
    const size_t elems = (size_t)limit * q.k_out;
    if ((rc = ensure(c_, &s.ws_out, &s.ws_out_bytes, elems * (sizeof(double) + sizeof(int)),
                     "serve outputs")))
      return rc;
    if ((rc = ensure_host(sl, elems, (size_t)limit))) return rc;
    sl.sized_limit = limit;
    sl.sized_k_out = q.k_out;
    sl.sized_k1 = q.k1;
    return 0;
  }

  static void free_host(ServeSlot& sl) {
    if (sl.h_dist) (void)hipHostFree(sl.h_dist);
    if (sl.h_idx) (void)hipHostFree(sl.h_idx);
    if (sl.h_fb) (void)hipHostFree(sl.h_fb);
    sl.h_dist = nullptr;
    sl.h_idx = nullptr;
    sl.h_fb = nullptr;
    sl.h_elems = sl.h_rows = 0;
  }

  int ensure_host(ServeSlot& sl, size_t elems, size_t rows) {
    if (elems > sl.h_elems) {
      if (sl.h_dist) (void)hipHostFree(sl.h_dist);
      if (sl.h_idx) (void)hipHostFree(sl.h_idx);
      sl.h_dist = nullptr;
      sl.h_idx = nullptr;
      sl.h_elems = 0;
      if (hipHostMalloc((void**)&sl.h_dist, elems * sizeof(double), hipHostMallocDefault) !=
              hipSuccess ||
          hipHostMalloc((void**)&sl.h_idx, elems * sizeof(int), hipHostMallocDefault) != hipSuccess)
        return fail(c_, MDSCTK_ERR_NOMEM, "serve: pinned output buffers");
      sl.h_elems = elems;
    }
    if (rows > sl.h_rows) {
      if (sl.h_fb) (void)hipHostFree(sl.h_fb);
      sl.h_fb = nullptr;
      sl.h_rows = 0;
      if (hipHostMalloc((void**)&sl.h_fb, rows * sizeof(int), hipHostMallocDefault) != hipSuccess)
        return fail(c_, MDSCTK_ERR_NOMEM, "serve: pinned flag buffer");
      sl.h_rows = rows;
    }
    return 0;
  }

  // Under mu_: pack pending requests into the slot.  Returns false when
  // nothing is pending.
  bool form_batch(ServeSlot& sl) {
    if (pending_.empty()) return false;
    sl.chunks.clear();
    sl.q = pending_.front()->q;
    const int64_t limit = frame_limit(sl.q);
    int64_t n = 0;
    const double now = clock_ms(c_);
    for (auto it = pending_.begin(); it != pending_.end() && n < limit;) {
      Request& r = **it;
      if (!same_shape(r.q, sl.q)) {
        ++it;
        continue;
      }
      const int64_t take = std::min(r.n - r.next_row, limit - n);
      sl.chunks.push_back(Chunk{*it, r.next_row, take, n});
      if (r.t_admit < 0) r.t_admit = now;
      ++r.n_batches;
      r.next_row += take;
      n += take;
      if (r.next_row == r.n)
        it = pending_.erase(it);
      else
        break;  // the batch is full; the rest of this request goes next
    }
    sl.n = n;
    sl.limit = limit;
    st_.max_batch_frames = std::max(st_.max_batch_frames, n);
    st_.slot_frames = (int32_t)limit;
    return true;
  }

  // Serving thread, mu_ not held: enqueue the whole batch on the slot.
  int launch(ServeSlot& sl) {
    Slot& s = sl.s;
    const QueryShape& q = sl.q;
    const int natoms = c_->ref.natoms;
    int rc;
    if (sl.sized_limit < sl.limit || sl.sized_k_out != q.k_out || sl.sized_k1 != q.k1)
      if ((rc = presize(sl, sl.limit))) return rc;
    if ((rc = slot_frames(c_, s, sl.n))) return rc;
    const size_t elems = (size_t)sl.n * q.k_out;
    if ((rc = ensure(c_, &s.ws_out, &s.ws_out_bytes, elems * (sizeof(double) + sizeof(int)),
                     "serve outputs")))
      return rc;
    sl.d_dist = (double*)s.ws_out;
    sl.d_idx = (int*)((char*)s.ws_out + elems * sizeof(double));
    if ((rc = ensure_host(sl, elems, (size_t)sl.n))) return rc;
    const size_t frame_bytes = (size_t)natoms * 3 * sizeof(float);
    for (const Chunk& ch : sl.chunks) {
      hipError_t e = hipMemcpyAsync((char*)s.fit.xyz + ch.brow0 * frame_bytes,
                                    ch.r->xyz + ch.row0 * natoms * 3, ch.nrows * frame_bytes,
                                    hipMemcpyHostToDevice, s.st);
      if (e != hipSuccess) return hip_fail(c_, e, "serve: query upload");
    }
    if (q.fused) {
      FusedPlan plan;
      if ((rc = plan_fused(c_, s, q, &plan))) return rc;
      if (c_->serve_share_cus && slots_.size() > 1)
        plan.grid = std::max(8, (plan.grid / (int)slots_.size()) & ~7);
      if (c_->serve_graphs) {
        if ((rc = replay_graph(sl, plan))) return rc;
      } else {
        if ((rc = prep_slot(c_, s))) return rc;
        if ((rc = enqueue_fused(c_, s, q, plan, sl.d_dist, sl.d_idx, false))) return rc;
      }
      sl.d_fb = plan.d_fallback;
      if (hipMemcpyAsync(sl.h_fb, sl.d_fb, sl.n * sizeof(int), hipMemcpyDeviceToHost, s.st) !=
          hipSuccess)
        return fail(c_, MDSCTK_ERR_HIP, "serve: flag copy");
    } else {
      if ((rc = prep_slot(c_, s))) return rc;
      if ((rc = exact_rows_select(c_, s, s.fit, nullptr, (int)sl.n, q, sl.d_dist, sl.d_idx)))
        return rc;
      std::memset(sl.h_fb, 0, sl.n * sizeof(int));
    }
    if (hipMemcpyAsync(sl.h_dist, sl.d_dist, elems * sizeof(double), hipMemcpyDeviceToHost,
                       s.st) != hipSuccess ||
        hipMemcpyAsync(sl.h_idx, sl.d_idx, elems * sizeof(int), hipMemcpyDeviceToHost, s.st) !=
            hipSuccess ||
        hipEventRecord(sl.done, s.st) != hipSuccess)
      return fail(c_, MDSCTK_ERR_HIP, "serve: output copy");
    return 0;
  }

  // The batch's kernels (prep, list init, window, contraction, merge) are
  // captured from the slot's stream into a graph.  The slot keeps one
  // executable graph and updates it in place with each new capture (same
  // topology, new sizes and launch shapes); it is rebuilt only when the
  // update is refused.  Nothing in the sequence is a memset or an
  // allocation: plan_fused has sized the workspace before the capture.
  // The captured graph stays alive until the next capture has replaced it:
  // the executable is only ever launched while the graph it came from
  // exists (a tracer reading the launched kernels' node data at launch time
  // then never reads released memory).
  int replay_graph(ServeSlot& sl, const FusedPlan& plan) {
    Slot& s = sl.s;
    hipError_t e = hipStreamBeginCapture(s.st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(c_, e, "serve: graph capture");
    int rc = prep_slot(c_, s);
    if (!rc) rc = enqueue_fused(c_, s, sl.q, plan, sl.d_dist, sl.d_idx, false);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s.st, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : hip_fail(c_, e, "serve: graph capture");
    }
    if (sl.exec) {
      hipGraphNode_t err_node = nullptr;
      hipGraphExecUpdateResult res;
      if (hipGraphExecUpdate(sl.exec, g, &err_node, &res) != hipSuccess ||
          res != hipGraphExecUpdateSuccess) {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(sl.exec);
        sl.exec = nullptr;
      }
    }
    if (!sl.exec) {
      e = hipGraphInstantiate(&sl.exec, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        sl.exec = nullptr;
        (void)hipGraphDestroy(g);
        return hip_fail(c_, e, "serve: graph instantiate");
      }
      ++graph_builds_;
    }
    e = hipGraphLaunch(sl.exec, s.st);
    // exec now holds g's parameters and has been launched: the graph it
    // replaced can go (the slot's stream orders every earlier launch of
    // this exec before this one)
    if (sl.graph) (void)hipGraphDestroy(sl.graph);
    sl.graph = g;
    if (e != hipSuccess) return hip_fail(c_, e, "serve: graph launch");
    ++graph_launches_;
    return 0;
  }

  // Serving thread, mu_ not held: the slot's work has finished.  Rows the
  // merge could not certify take the exact path; then outputs go to the
  // callers' buffers.
  int finish(ServeSlot& sl, std::vector<int64_t>* fb_per_chunk) {
    Slot& s = sl.s;
    const QueryShape& q = sl.q;
    std::vector<int> rows;
    for (int64_t i = 0; i < sl.n; ++i)
      if (sl.h_fb[i]) rows.push_back((int)i);
    const size_t elems = (size_t)sl.n * q.k_out;
    if (!rows.empty()) {
      int rc = exact_rows_select(c_, s, s.fit, rows.data(), (int)rows.size(), q, sl.d_dist,
                                 sl.d_idx);
      if (rc) return rc;
      if (hipMemcpyAsync(sl.h_dist, sl.d_dist, elems * sizeof(double), hipMemcpyDeviceToHost,
                         s.st) != hipSuccess ||
          hipMemcpyAsync(sl.h_idx, sl.d_idx, elems * sizeof(int), hipMemcpyDeviceToHost, s.st) !=
              hipSuccess ||
          hipStreamSynchronize(s.st) != hipSuccess)
        return fail(c_, MDSCTK_ERR_HIP, "serve: output copy");
    }
    fb_per_chunk->assign(sl.chunks.size(), 0);
    for (size_t j = 0; j < sl.chunks.size(); ++j) {
      const Chunk& ch = sl.chunks[j];
      // A request that already failed in another batch has been handed back
      // to its caller: write nothing more into its buffers.  (Only this
      // thread changes a request's status once it is queued.)
      if (ch.r->status != MDSCTK_ERR_PENDING) continue;
      const size_t off_b = (size_t)ch.brow0 * q.k_out, off_r = (size_t)ch.row0 * q.k_out;
      const size_t cnt = (size_t)ch.nrows * q.k_out;
      std::memcpy(ch.r->out_dist + off_r, sl.h_dist + off_b, cnt * sizeof(double));
      std::memcpy(ch.r->out_idx + off_r, sl.h_idx + off_b, cnt * sizeof(int));
      for (int64_t i = 0; i < ch.nrows; ++i) (*fb_per_chunk)[j] += sl.h_fb[ch.brow0 + i] ? 1 : 0;
    }
    return 0;
  }

  // Under mu_: record the outcome of a slot's batch.
  void retire(ServeSlot& sl, int rc, const std::vector<int64_t>& fb_per_chunk) {
    const double now = clock_ms(c_);
    for (size_t j = 0; j < sl.chunks.size(); ++j) {
      Request& r = *sl.chunks[j].r;
      r.batch_frames = std::max(r.batch_frames, sl.n);
      if (r.status != MDSCTK_ERR_PENDING) continue;  // already failed
      if (rc) {
        r.status = rc;
        r.err = err_;
        r.t_done = now;
        if (r.next_row < r.n) {
          r.next_row = r.n;  // place no further rows
          pending_.erase(std::remove(pending_.begin(), pending_.end(), sl.chunks[j].r),
                         pending_.end());
        }
        continue;
      }
      r.rows_done += sl.chunks[j].nrows;
      r.fallback_rows += j < fb_per_chunk.size() ? fb_per_chunk[j] : 0;
      if (r.rows_done == r.n) {
        r.status = r.q.k_out;
        r.t_done = now;
        ++st_.requests_done;
      }
    }
    if (!rc) {
      ++st_.batches;
      st_.frames += sl.n;
      st_.pairs += sl.n * c_->ref.n;
      for (int64_t v : fb_per_chunk) st_.fallback_rows += v;
    }
    sl.chunks.clear();
    sl.busy = false;
    cv_done_.notify_all();
  }

  void note_busy_change() {
    bool any = false;
    for (auto& sl : slots_) any |= sl.busy;
    if (any && busy_since_ < 0) busy_since_ = clock_ms(c_);
    if (!any && busy_since_ >= 0) {
      busy_ms_ += clock_ms(c_) - busy_since_;
      busy_since_ = -1;
    }
  }

  void run() {
    (void)hipSetDevice(c_->device);
    set_error_sink(&err_);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      bool any_busy = false;
      for (auto& sl : slots_) {
        if (!sl.busy) continue;
        lk.unlock();
        const hipError_t e = hipEventQuery(sl.done);
        int rc = 0;
        std::vector<int64_t> fb;
        if (e == hipSuccess)
          rc = finish(sl, &fb);
        else if (e != hipErrorNotReady)
          rc = hip_fail(c_, e, "serve: batch");
        lk.lock();
        if (e == hipErrorNotReady) {
          any_busy = true;
          continue;
        }
        retire(sl, rc, fb);
        note_busy_change();
      }
      if (stop_) {
        if (any_busy) {
          cv_work_.wait_for(lk, std::chrono::microseconds(50));
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
        const int rc = launch(sl);
        if (rc) (void)hipStreamSynchronize(sl.s.st);
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
        cv_work_.wait_for(lk, std::chrono::microseconds(20));
      else
        cv_work_.wait(lk, [&] { return stop_ || !pending_.empty(); });
    }
  }

  mdsctk_knn_ctx* c_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<std::shared_ptr<Request>> pending_;
  std::unordered_map<int64_t, std::shared_ptr<Request>> live_;
  int64_t next_id_ = 1;
  bool stop_ = false;
  std::vector<ServeSlot> slots_;
  size_t avail_ = 0;  // HBM budget left when serving started (0 = no budget)
  mdsctk_knn_serve_stats st_ = {};
  double busy_ms_ = 0, busy_since_ = -1;
  std::atomic<int64_t> graph_launches_{0}, graph_builds_{0};
  std::string err_;  // error text of the serving thread's last failure
  std::thread th_;
};

int server_submit(mdsctk_knn_ctx* c, const float* xyz, int64_t n_fit, int32_t k, uint32_t flags,
                  double* out_dist, int32_t* out_idx, int64_t* id) {
  if (!id) return MDSCTK_ERR_ARG;
  *id = 0;
  std::lock_guard<std::mutex> g(c->server_mu);
  if (!c->has_ref) return fail(c, MDSCTK_ERR_STATE, "submit: no reference set");
  if (c->w.metric != kMetricRMSD)
    return fail(c, MDSCTK_ERR_STATE,
                "submit: serving takes frames; data vectors use mdsctk_data_query");
  if (!xyz || n_fit < 0 || n_fit > INT32_MAX || (n_fit > 0 && (!out_dist || !out_idx)))
    return fail(c, MDSCTK_ERR_ARG, "submit: need query frames and outputs (n_fit=%lld)",
                (long long)n_fit);
  if (flags & MDSCTK_KNN_FULL_ROWS)
    return fail(c, MDSCTK_ERR_ARG, "submit: full rows are served by mdsctk_knn_query only");
  QueryShape q;
  int rc = resolve_shape(c, k, flags, &q);
  if (rc) return rc;
  if (!c->server) {
    auto* s = new Server(c);
    rc = s->start();
    if (rc) {
      delete s;
      return rc;
    }
    c->server = s;
  }
  return c->server->submit(xyz, n_fit, q, out_dist, out_idx, id);
}

int server_poll(mdsctk_knn_ctx* c, int64_t id, mdsctk_knn_request_info* info, bool wait) {
  Server* s;
  {
    std::lock_guard<std::mutex> g(c->server_mu);
    s = c->server;
  }
  if (!s) return fail(c, MDSCTK_ERR_STATE, "poll: no request was submitted");
  std::string err;
  const int rc = s->poll(id, info, wait, &err);
  if (rc < 0 && rc != MDSCTK_ERR_PENDING) set_ctx_error(c, err);  // the polling thread's text
  return rc;
}

int server_stats(mdsctk_knn_ctx* c, mdsctk_knn_serve_stats* out) {
  std::lock_guard<std::mutex> g(c->server_mu);
  *out = mdsctk_knn_serve_stats();
  if (c->server) c->server->stats(out);
  return 0;
}

bool server_idle(mdsctk_knn_ctx* c) {
  std::lock_guard<std::mutex> g(c->server_mu);
  return !c->server || c->server->idle();
}

void server_stop(mdsctk_knn_ctx* c) {
  std::lock_guard<std::mutex> g(c->server_mu);
  delete c->server;
  c->server = nullptr;
}

}  // namespace mdsctk
