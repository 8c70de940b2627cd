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
//
// The queue, the batch packing and the serving thread's loop are in
// serve_queue.h, free of device calls; this file is their HIP side
// (HipServeDevice) and the C entry points.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>

#include "engine.h"
#include "serve_queue.h"

namespace mdsctk {

namespace {

// Device state of one batch slot.
struct HipSlot {
  Slot s;
  double* d_dist = nullptr;
  int* d_idx = nullptr;
  int* d_fb = nullptr;
  // pinned host mirrors of the outputs
  double* h_dist = nullptr;
  int* h_idx = nullptr;
  int* h_fb = nullptr;
  bool all_exact = false;  // the batch took the exact rows, not the contraction
  size_t h_elems = 0, h_rows = 0;
  hipEvent_t done = nullptr;
  hipGraphExec_t exec = nullptr;  // the slot's batch kernels, updated per batch
  // the captured graph exec was last instantiated or updated from; kept
  // until a later capture has replaced it in exec (and that exec has been
  // launched), so nothing the executable may refer to is released while it
  // can still run
  hipGraph_t graph = nullptr;
  // shape the workspace was last sized for
  int64_t sized_limit = 0;
  int sized_k_out = -1, sized_k1 = -1;
};

}  // namespace

// The HIP side of the serving queue (the Device of serve_queue.h): a batch
// is H2D of its frames, prep, contraction and merge on the slot's stream
// (replayed as a graph), D2H of the outputs into pinned buffers.
class HipServeDevice {
 public:
  using Shape = QueryShape;
  using SlotState = HipSlot;
  using Batch = ServeBatch<QueryShape>;

  explicit HipServeDevice(mdsctk_knn_ctx* c) : c_(c) {}

  static bool same_shape(const QueryShape& a, const QueryShape& b) {
    return a.k_out == b.k_out && a.k1 == b.k1 && a.rank_off == b.rank_off && a.nofit == b.nofit;
  }

  int init(int n_slots) {
    slots_.resize(n_slots);
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
    return 0;
  }

  void release() {
    for (auto& sl : slots_) {
      if (sl.s.st) (void)hipStreamSynchronize(sl.s.st);
      if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
      if (sl.graph) (void)hipGraphDestroy(sl.graph);
      slot_release(c_, sl.s);
      if (sl.done) (void)hipEventDestroy(sl.done);
      free_host(sl);
    }
    slots_.clear();
  }

  std::vector<HipSlot*> slot_states() {
    std::vector<HipSlot*> v;
    for (auto& sl : slots_) v.push_back(&sl);
    return v;
  }

  double now() { return clock_ms(c_); }

  void thread_init() {
    (void)hipSetDevice(c_->device);
    set_error_sink(&err_);
  }

  const std::string& error_text() { return err_; }
  int64_t ref_frames() { return c_->ref.n; }
  int64_t graph_launches() const { return graph_launches_.load(); }
  int64_t graph_builds() const { return graph_builds_.load(); }

  // Largest batch (frames) whose workspace fits, for every slot at once,
  // in the HBM budget left when serving started.
  int64_t frame_limit(const QueryShape& q) const {
    int64_t f = std::max<int64_t>(1, c_->serve_max_frames);
    if (!avail_) return f;
    while (f > 32 && (double)slot_bytes(f, q) * slots_.size() > 0.9 * (double)avail_) f /= 2;
    return f;
  }

  // Enqueue the whole batch on the slot.
  int launch(HipSlot& sl, const Batch& b) {
    Slot& s = sl.s;
    const QueryShape& q = b.q;
    const int natoms = c_->ref.natoms;
    int rc;
    if (sl.sized_limit < b.limit || sl.sized_k_out != q.k_out || sl.sized_k1 != q.k1)
      if ((rc = presize(sl, q, b.limit))) return rc;
    if ((rc = slot_frames(c_, s, b.n))) return rc;
    const size_t elems = (size_t)b.n * q.k_out;
    if ((rc = ensure(c_, &s.ws_out, &s.ws_out_bytes, elems * (sizeof(double) + sizeof(int)),
                     "serve outputs")))
      return rc;
    sl.d_dist = (double*)s.ws_out;
    sl.d_idx = (int*)((char*)s.ws_out + elems * sizeof(double));
    if ((rc = ensure_host(sl, elems, (size_t)b.n))) return rc;
    const size_t frame_bytes = (size_t)natoms * 3 * sizeof(float);
    for (const auto& ch : b.chunks) {
      hipError_t e = hipMemcpyAsync((char*)s.fit.xyz + ch.brow0 * frame_bytes,
                                    ch.r->xyz + ch.row0 * natoms * 3, ch.nrows * frame_bytes,
                                    hipMemcpyHostToDevice, s.st);
      if (e != hipSuccess) return hip_fail(c_, e, "serve: query upload");
    }
    // --nofit through the contraction needs the resident set's operands as
    // given; without room for them the batch takes the exact rows
    bool fused = q.fused;
    if (fused && q.nofit && (rc = ensure_plain(c_, s.st))) {
      if (rc != MDSCTK_ERR_NOMEM) return rc;
      fused = false;
      err_.clear();  // not a failure of the batch
    }
    sl.all_exact = !fused;
    if (fused) {
      FusedPlan plan;
      if ((rc = plan_fused(c_, s, q, &plan))) return rc;
      if (c_->serve_share_cus && slots_.size() > 1)
        plan.grid = std::max(8, (plan.grid / (int)slots_.size()) & ~7);
      if (c_->serve_graphs) {
        if ((rc = replay_graph(sl, q, plan))) return rc;
      } else {
        if ((rc = prep_slot(c_, s, q.nofit))) return rc;
        if ((rc = enqueue_fused(c_, s, q, plan, sl.d_dist, sl.d_idx, false))) return rc;
      }
      sl.d_fb = plan.d_fallback;
      if (hipMemcpyAsync(sl.h_fb, sl.d_fb, b.n * sizeof(int), hipMemcpyDeviceToHost, s.st) !=
          hipSuccess)
        return fail(c_, MDSCTK_ERR_HIP, "serve: flag copy");
    } else {
      if ((rc = prep_slot(c_, s, q.nofit))) return rc;
      if ((rc = exact_rows_select(c_, s, s.fit, nullptr, (int)b.n, q, sl.d_dist, sl.d_idx)))
        return rc;
      std::memset(sl.h_fb, 0, b.n * sizeof(int));
    }
    if (hipMemcpyAsync(sl.h_dist, sl.d_dist, elems * sizeof(double), hipMemcpyDeviceToHost,
                       s.st) != hipSuccess ||
        hipMemcpyAsync(sl.h_idx, sl.d_idx, elems * sizeof(int), hipMemcpyDeviceToHost, s.st) !=
            hipSuccess ||
        hipEventRecord(sl.done, s.st) != hipSuccess)
      return fail(c_, MDSCTK_ERR_HIP, "serve: output copy");
    return 0;
  }

  int poll(HipSlot& sl) {
    const hipError_t e = hipEventQuery(sl.done);
    if (e == hipSuccess) return 0;
    if (e == hipErrorNotReady) return 1;
    return hip_fail(c_, e, "serve: batch");
  }

  void drain(HipSlot& sl) { (void)hipStreamSynchronize(sl.s.st); }

  // The slot's work has finished.  Rows the merge could not certify take the
  // exact path; then outputs go to the callers' buffers.
  int finish(HipSlot& sl, const Batch& b, std::vector<int64_t>* fb_per_chunk) {
    Slot& s = sl.s;
    const QueryShape& q = b.q;
    std::vector<int> rows;
    for (int64_t i = 0; i < b.n; ++i)
      if (sl.h_fb[i]) rows.push_back((int)i);
    const size_t elems = (size_t)b.n * q.k_out;
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
    fb_per_chunk->assign(b.chunks.size(), 0);
    for (size_t j = 0; j < b.chunks.size(); ++j) {
      const auto& ch = b.chunks[j];
      // A request that already failed in another batch has been handed back
      // to its caller: write nothing more into its buffers.  (Only the
      // serving thread changes a request's status once it is queued.)
      if (ch.r->status != MDSCTK_ERR_PENDING) continue;
      const size_t off_b = (size_t)ch.brow0 * q.k_out, off_r = (size_t)ch.row0 * q.k_out;
      const size_t cnt = (size_t)ch.nrows * q.k_out;
      std::memcpy(ch.r->out_dist + off_r, sl.h_dist + off_b, cnt * sizeof(double));
      std::memcpy(ch.r->out_idx + off_r, sl.h_idx + off_b, cnt * sizeof(int));
      for (int64_t i = 0; i < ch.nrows; ++i)
        (*fb_per_chunk)[j] += (sl.all_exact || sl.h_fb[ch.brow0 + i]) ? 1 : 0;
    }
    return 0;
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
    // xyz + operand (12 B per atom each) + per-frame terms (incl. the 72-byte
    // principal axes) + outputs
    double b = frames * (24.0 * natoms + 136 + 12.0 * q.k_out);
    if (q.fused) {
      size_t ent, hdr;
      list_need(limit, q, &ent, &hdr);
      b += 8.0 * ent + 8.0 * hdr;
    }
    return (size_t)b;
  }

  // Size the slot's buffers once for its largest batch of this shape, so
  // that steady-state batches allocate nothing.
  int presize(HipSlot& sl, const QueryShape& q, int64_t limit) {
    Slot& s = sl.s;
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

  static void free_host(HipSlot& sl) {
    if (sl.h_dist) (void)hipHostFree(sl.h_dist);
    if (sl.h_idx) (void)hipHostFree(sl.h_idx);
    if (sl.h_fb) (void)hipHostFree(sl.h_fb);
    sl.h_dist = nullptr;
    sl.h_idx = nullptr;
    sl.h_fb = nullptr;
    sl.h_elems = sl.h_rows = 0;
  }

  int ensure_host(HipSlot& sl, size_t elems, size_t rows) {
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
  int replay_graph(HipSlot& sl, const QueryShape& q, const FusedPlan& plan) {
    Slot& s = sl.s;
    hipError_t e = hipStreamBeginCapture(s.st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(c_, e, "serve: graph capture");
    int rc = prep_slot(c_, s, q.nofit);
    if (!rc) rc = enqueue_fused(c_, s, q, plan, sl.d_dist, sl.d_idx, false);
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

  mdsctk_knn_ctx* c_;
  std::vector<HipSlot> slots_;
  size_t avail_ = 0;  // HBM budget left when serving started (0 = no budget)
  std::atomic<int64_t> graph_launches_{0}, graph_builds_{0};
  std::string err_;  // error text of the serving thread's last failure
};

class Server {
 public:
  explicit Server(mdsctk_knn_ctx* c) : dev_(c) {}

  int start(int n_slots) {
    int rc = dev_.init(n_slots);
    if (rc) {
      dev_.release();
      return rc;
    }
    q_ = std::make_unique<ServeQueue<HipServeDevice>>(dev_, dev_.slot_states());
    q_->start();
    return 0;
  }

  ~Server() {
    if (q_) q_->stop();
    q_.reset();
    dev_.release();
  }

  int submit(const float* xyz, int64_t n, const QueryShape& q, double* od, int32_t* oi,
             int64_t* id) {
    return q_->submit(xyz, n, q, od, oi, id);
  }
  int poll(int64_t id, mdsctk_knn_request_info* info, bool wait, std::string* err) {
    return q_->poll(id, info, wait, err);
  }
  void stats(mdsctk_knn_serve_stats* out) {
    q_->stats(out);
    out->graph_launches = dev_.graph_launches();
    out->graph_builds = dev_.graph_builds();
  }
  bool idle() { return q_->idle(); }

 private:
  HipServeDevice dev_;
  std::unique_ptr<ServeQueue<HipServeDevice>> q_;
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
    rc = s->start(c->serve_slots);
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
