// Internal host-side engine shared by the synchronous C ABI (api.cpp) and
// the serving loop (serve.cpp): the context object, device memory
// accounting and the per-batch query pipeline.  Not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mdsctk_knn.h"
#include "knn_device.h"

namespace mdsctk {

// Device workspace of one query batch in flight: its stream, the query
// frames and their derived arrays, candidate lists and device outputs.
// Every buffer grows on demand and is reused by later batches.
struct Slot {
  hipStream_t st = nullptr;
  DevFrames fit;
  size_t fit_cap_frames = 0;
  int fit_natoms = 0;
  int fit_vdim = 0;    // data vectors: length of the staged vectors
  int fit_metric = 0;
  void* ws_lists = nullptr;
  size_t ws_lists_bytes = 0;
  void* ws_small = nullptr;
  size_t ws_small_bytes = 0;
  int* n_cand = nullptr;
  void* ws_out = nullptr;
  size_t ws_out_bytes = 0;
  void* ws_rows = nullptr;
  size_t ws_rows_bytes = 0;
  void* ws_sym = nullptr;  // mirrored lists of a symmetric self query
  size_t ws_sym_bytes = 0;
  void* ws_relist = nullptr;  // row ids and outputs of rows recomputed through the lists
  size_t ws_relist_bytes = 0;
  void* ws_sort = nullptr;  // exact rows with k+1 beyond the LDS select: compaction + sort
  size_t ws_sort_bytes = 0;
  // ev[0] start, ev[1] query prepared, ev[2] before the contraction, ev[3]
  // after it, ev[4] after the merge, ev[5] end; ev[6] after a symmetric self
  // query's landmark pass (so ev[2]..ev[6] is that pass).
  hipEvent_t ev[8] = {};
};

// Shape of one query: output width, ranks kept, and the route it takes.
struct QueryShape {
  int k_out = 0;     // entries written per row
  int k1 = 0;        // ranks selected (k_out + rank_off)
  int rank_off = 1;  // 1 = drop the nearest (knn_rms.cpp:246-255)
  bool nofit = false;
  bool full = false;   // --sort false: full rows
  bool fused = false;  // contraction + merge (else exact rows)
};

// Workspace layout and launch shape of one fused batch, fixed before any
// work is enqueued (so that the enqueue step allocates nothing and can be
// captured into a graph).
struct FusedPlan {
  Lists L;
  float* eps2 = nullptr;
  int* d_fallback = nullptr;
  int phases = 0;
  int grid = 0;
  double kappa = 0;
};

// What the fused pipeline did for a batch.
struct FusedInfo {
  int phases = 0;
  int cap = 0;
  double kappa = 0;
  int* d_fallback = nullptr;  // [n_fit] device flags (in the slot's workspace)
  int64_t pairs_computed = 0;  // pairs the main contraction scored
  // symmetric self query only
  bool symmetric = false;
  int64_t pairs_landmark = 0;  // pairs of its landmark pass (0: none)
  unsigned long long* d_mirror_sum = nullptr;  // [2]: mirrored appends, most per frame
};

// Pieces of caller-provided device memory handed out in order, 256-byte
// aligned: the allocation-free query carves its whole workspace from the
// per-GPU fixed region this way.
struct Carve {
  char* base = nullptr;
  size_t size = 0;
  size_t used = 0;
  static size_t round(size_t bytes) { return (bytes + 255) & ~size_t(255); }
  size_t left() const { return size - used; }
  void* take(size_t bytes) {
    const size_t b = round(bytes ? bytes : 1);
    if (!base || b > left()) return nullptr;
    void* p = base + used;
    used += b;
    return p;
  }
};

class Server;  // serve.cpp

}  // namespace mdsctk

struct mdsctk_knn_ctx {
  int device = 0;
  int prec = mdsctk::kFP32;
  int n_cu = 256;
  // Device memory accounting; the serving thread allocates too.
  mutable std::mutex mem_mu;
  size_t budget = 0;  // 0 = unlimited
  size_t held = 0;
  std::unordered_map<void*, size_t> allocs;
  double eps_scale = 1.0;
  int forced_phases = 0;
  int serve_slots = 2;
  int64_t serve_max_frames = 2048;
  bool serve_graphs = true;  // replay each served batch's kernels as a hipGraph
  bool serve_share_cus = true;  // each batch's contraction takes 1/serve_slots of the CUs
  bool self_sym = true;  // self queries score each unordered pair once when the lists fit
  int mirror_cap = mdsctk::kMirrorCap;  // entries of each frame's mirrored list

  mdsctk::DevWeights w;  // masses, or unit masses plus the metric for data vectors
  // The resident set.  For fitted RMSD scores in a contraction form its
  // operands hold every frame in its own principal axes (ref_axes), as do
  // the query frames of fitted queries: the scores do not change, and the
  // contraction's lambda_max bound becomes nearly exact.  --nofit scores
  // need the frames as given: ref_plain holds those operands (and its own
  // centroids, G and G~), built when a --nofit query first needs them.
  mdsctk::DevFrames ref;
  mdsctk::DevFrames ref_plain;
  bool ref_axes = false;
  std::mutex plain_mu;
  double* data_shift = nullptr;  // data vectors, Euclidean: the reference set's mean vector
  bool has_ref = false;
  mdsctk::Slot main;  // workspace of the synchronous entry points
  mdsctk_knn_stats stats = {};
  mdsctk::Server* server = nullptr;
  std::mutex server_mu;  // guards creating / stopping the serving thread
  // This device's fixed workspace of the allocation-free query (looked up
  // at create; null with the reason in scratch_err when unavailable).
  void* scratch = nullptr;
  std::string scratch_err;
};

namespace mdsctk {

// Error reporting: sets the context's error text (or, on the serving
// thread, that thread's error sink) and returns `code`.
int fail(mdsctk_knn_ctx* c, int code, const char* fmt, ...);
int hip_fail(mdsctk_knn_ctx* c, hipError_t e, const char* what);
// Directs fail() on the calling thread to `sink` (nullptr = the context).
void set_error_sink(std::string* sink);
// The calling thread's error text for context c (mdsctk_knn_last_error).
void set_ctx_error(const mdsctk_knn_ctx* c, const std::string& text);

int dalloc(mdsctk_knn_ctx* c, void** p, size_t bytes, const char* what);
void dfree(mdsctk_knn_ctx* c, void* p);
int ensure(mdsctk_knn_ctx* c, void** p, size_t* have, size_t need, const char* what);

int slot_init(mdsctk_knn_ctx* c, Slot& s);
void slot_release(mdsctk_knn_ctx* c, Slot& s);
// Free the symmetric self query's mirrored lists, which the context
// otherwise keeps for the next self query.  Called by the synchronous entry
// points only (a new reference set; a query that is not a self query; the
// end of a self query while serving runs), so it never races a query that
// uses them.
void release_mirror_lists(mdsctk_knn_ctx* c);

// Validate k/flags against the resident set and fill `q`.
int resolve_shape(mdsctk_knn_ctx* c, int32_t k, uint32_t flags, QueryShape* q);

// Make the slot's query frame set hold n_fit frames of the resident
// set's atom count (grow-only), with its tile count trimmed to n_fit.
int slot_frames(mdsctk_knn_ctx* c, Slot& s, int64_t n_fit);
// Centring / operand pass over the slot's query frames (async); fitted
// queries (nofit false) on a resident set in principal axes take the
// frames in theirs.
int prep_slot(mdsctk_knn_ctx* c, Slot& s, bool nofit);
// The resident set's --nofit operands, built on first use (synchronous).
int ensure_plain(mdsctk_knn_ctx* c, hipStream_t st);
// Contraction grid (workgroups) and reference phases for a batch of
// `fit_tiles` query tiles; candidate-list entries a batch of n_fit frames
// needs (each entry 8 bytes; the headers take 8 bytes per list).
int contract_grid(const mdsctk_knn_ctx* c);
int fused_phases(const mdsctk_knn_ctx* c, int fit_tiles);
size_t fused_list_entries(const mdsctk_knn_ctx* c, int64_t n_fit, int k1);
// The fused pipeline (lists, contraction, merge) for the slot's query
// frames: plan_fused sizes the slot's workspace (it may allocate);
// enqueue_fused only enqueues kernels on the slot's stream (with timing,
// also events ev[2..4]), outputs to d_dist/d_idx; launch_fused does both
// with timing.
// With `cv`, plan_fused takes the workspace from it instead of the slot's
// buffers (allocating nothing), lowering the reference phases until the
// candidate lists fit; it fails with MDSCTK_ERR_ARG when even one phase does
// not.
int plan_fused(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, FusedPlan* p,
               Carve* cv = nullptr);
int enqueue_fused(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, const FusedPlan& p,
                  double* d_dist, int* d_idx, bool timing);
int launch_fused(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, double* d_dist, int* d_idx,
                 FusedInfo* info, Carve* cv = nullptr);
// A self query (the slot's frames are the resident set's query view)
// scoring each unordered pair once: a landmark pass bounds every row's
// threshold, the contraction scores the upper block triangle into each fit
// frame's lists and each reference frame's mirrored list, and the merge
// re-scores both as for the lists, with timing.  No score matrix is
// written.  Returns 1 with nothing enqueued when the lists do not fit the
// free device memory (or the context's budget), or k+1 exceeds what the
// lists hold.
int launch_fused_sym(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, double* d_dist, int* d_idx,
                     FusedInfo* info);
// Rows of a symmetric self query that it could not certify (a mirrored
// list overflowed) again through the candidate lists, their frames gathered
// from the resident set; `rows` keeps those still uncertified.
int relist_rows(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, std::vector<int>& rows,
                double* d_dist, int* d_idx);
// Exact FP64 rows for the listed fit rows (nullptr = rows [0,nrows)) of
// the slot's query frames, synchronously on the slot's stream.  Its row
// scratch comes from `cv` when given (else the slot's grow-only buffer).
int exact_rows_select(mdsctk_knn_ctx* c, Slot& s, const DevFrames& fit, const int* h_rows,
                      int nrows, const QueryShape& q, double* d_dist, int* d_idx,
                      Carve* cv = nullptr);

double clock_ms(const mdsctk_knn_ctx* c);

// Serving loop (serve.cpp).
int server_submit(mdsctk_knn_ctx* c, const float* xyz, int64_t n_fit, int32_t k, uint32_t flags,
                  double* out_dist, int32_t* out_idx, int64_t* id);
int server_poll(mdsctk_knn_ctx* c, int64_t id, mdsctk_knn_request_info* info, bool wait);
int server_stats(mdsctk_knn_ctx* c, mdsctk_knn_serve_stats* out);
bool server_idle(mdsctk_knn_ctx* c);
void server_stop(mdsctk_knn_ctx* c);

}  // namespace mdsctk
