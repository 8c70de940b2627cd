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
  void* ws_sym = nullptr;  // score matrix of a symmetric self query
  size_t ws_sym_bytes = 0;
  hipEvent_t ev[6] = {};
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
  int64_t pairs_computed = 0;  // pairs the contraction scored
};

class Server;  // serve.cpp

}  // namespace mdsctk

struct mdsctk_knn_ctx {
  int device = 0;
  int prec = mdsctk::kFP32;
  int n_cu = 256;
  std::string err;
  // Device memory accounting; the serving thread allocates too.
  std::mutex mem_mu;
  size_t budget = 0;  // 0 = unlimited
  size_t held = 0;
  std::unordered_map<void*, size_t> allocs;
  double eps_scale = 1.0;
  int forced_phases = 0;
  int serve_slots = 2;
  int64_t serve_max_frames = 8192;
  bool serve_graphs = true;  // replay each served batch's kernels as a hipGraph
  bool serve_share_cus = true;  // each batch's contraction takes 1/serve_slots of the CUs
  bool self_sym = true;  // self queries through the score matrix when it fits

  mdsctk::DevWeights w;  // masses, or unit masses plus the metric for data vectors
  mdsctk::DevFrames ref;
  double* data_shift = nullptr;  // data vectors, Euclidean: the reference set's mean vector
  bool has_ref = false;
  mdsctk::Slot main;  // workspace of the synchronous entry points
  mdsctk_knn_stats stats = {};
  mdsctk::Server* server = nullptr;
  std::mutex server_mu;  // guards creating / stopping the serving thread
};

namespace mdsctk {

// Error reporting: sets the context's error text (or, on the serving
// thread, that thread's error sink) and returns `code`.
int fail(mdsctk_knn_ctx* c, int code, const char* fmt, ...);
int hip_fail(mdsctk_knn_ctx* c, hipError_t e, const char* what);
// Directs fail() on the calling thread to `sink` (nullptr = the context).
void set_error_sink(std::string* sink);

int dalloc(mdsctk_knn_ctx* c, void** p, size_t bytes, const char* what);
void dfree(mdsctk_knn_ctx* c, void* p);
int ensure(mdsctk_knn_ctx* c, void** p, size_t* have, size_t need, const char* what);

int slot_init(mdsctk_knn_ctx* c, Slot& s);
void slot_release(mdsctk_knn_ctx* c, Slot& s);
// Free the self query's score matrix, which the context otherwise keeps
// for the next self query.  Called by the synchronous entry points only (a
// new reference set; the end of a self query while serving runs), so it
// never races a query that uses the matrix.
void release_score_matrix(mdsctk_knn_ctx* c);

// Validate k/flags against the resident set and fill `q`.
int resolve_shape(mdsctk_knn_ctx* c, int32_t k, uint32_t flags, QueryShape* q);

// Make the slot's query frame set hold n_fit frames of the resident
// set's atom count (grow-only), with its tile count trimmed to n_fit.
int slot_frames(mdsctk_knn_ctx* c, Slot& s, int64_t n_fit);
// Centring / operand pass over the slot's query frames (async).
int prep_slot(mdsctk_knn_ctx* c, Slot& s);
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
int plan_fused(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, FusedPlan* p);
int enqueue_fused(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, const FusedPlan& p,
                  double* d_dist, int* d_idx, bool timing);
int launch_fused(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, double* d_dist, int* d_idx,
                 FusedInfo* info);
// A self query (the slot's frames are the resident set's query view)
// through the n x n FP32 score matrix: the contraction scores the upper block
// triangle only, each row's candidates are selected from the matrix, and
// the merge re-scores them as for the lists, with timing.  Returns 1 with
// nothing enqueued when the matrix and its lists do not fit the free device
// memory (or the context's budget).
int launch_fused_sym(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, double* d_dist, int* d_idx,
                     FusedInfo* info);
// Exact FP64 rows for the listed fit rows (nullptr = rows [0,nrows)) of
// the slot's query frames, synchronously on the slot's stream.
int exact_rows_select(mdsctk_knn_ctx* c, Slot& s, const DevFrames& fit, const int* h_rows,
                      int nrows, const QueryShape& q, double* d_dist, int* d_idx);

double clock_ms(const mdsctk_knn_ctx* c);

// Serving loop (serve.cpp).
int server_submit(mdsctk_knn_ctx* c, const float* xyz, int64_t n_fit, int32_t k, uint32_t flags,
                  double* out_dist, int32_t* out_idx, int64_t* id);
int server_poll(mdsctk_knn_ctx* c, int64_t id, mdsctk_knn_request_info* info, bool wait);
int server_stats(mdsctk_knn_ctx* c, mdsctk_knn_serve_stats* out);
bool server_idle(mdsctk_knn_ctx* c);
void server_stop(mdsctk_knn_ctx* c);

}  // namespace mdsctk
