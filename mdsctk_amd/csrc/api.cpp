// C ABI of the knn_rms engine (include/mdsctk_knn.h): context, device
// memory accounting, and the per-query orchestration of the gfx950 kernels.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "engine.h"

namespace mdsctk {

namespace {

thread_local std::string* tl_err_sink = nullptr;
// Error text of failed calls, per calling thread and context: calls on one
// context may come from several threads (the serving calls may overlap a
// synchronous query), and none of them writes another thread's text.
thread_local std::unordered_map<const mdsctk_knn_ctx*, std::string> tl_ctx_err;

// Candidate window: |s~ - s| <= kappa (G_q + max G_ref) for the approximate
// score s~ of each contraction form (s = G_q + G_r - 2 lambda).  kappa is an
// envelope fitted to the largest relative deviation measured on the
// synthetic generators (random, metastable with near-duplicates, and
// pre-aligned frames with 30 % near-duplicates; uniform and backbone masses;
// fit and --nofit; cross and self queries) for N = 16..2000 atoms, times a
// safety factor of 2.  FP32 accumulation contributes the a*sqrt(N) term (one
// rounding per MFMA step).  The split forms add a representation term:
//   FP16  a constant b = 1e-7 (11-bit hi part);
//   BF16  c / sqrt(N) with the dithered split (operand_kernel's atom_scale),
//         whose residues no longer align between frames, so the term
//         averages over atoms; c = 3.4e-5 covers the largest deviation,
//         8.5e-6 at N = 16 with --nofit (without dithering the same term was
//         a constant 4e-6 from aligned frames, 2.4x wider at N = 1000).
// It is a measured statistical envelope, not a worst-case proof.
double kappa_for(int prec, int natoms) {
  const double rn = std::sqrt((double)std::max(natoms, 16));
  switch (prec) {
    case kFP32: return 2.0 * (5.5e-8 * rn);
    case kBF16: return 2.0 * (3.0e-8 * rn + 3.4e-5 / rn);
    case kFP16: return 2.0 * (3.0e-8 * rn + 1.0e-7);
    default: return 0.0;
  }
}

// Window of data vectors.  Their value distributions are arbitrary (few
// elements, no averaging over atoms), so the split forms' representation
// term is the proven worst case instead of the fitted constant: with
// A = hi + lo and r = x - A (|lo| <= 2^-8 |x|, |r| <= 2^-16 |x|),
// s~ = sum (A_a - A_b)^2 + 2 sum lo_a lo_b, and Cauchy-Schwarz bounds
// |s~ - |a - b|^2| by 5 * 2^-16 (G_a + G_b).  Rounding the FP64 vectors to
// the FP32 source values adds at most 2^-22 (G_a + G_b) (|delta| <= 2^-24
// |x|, s <= 2 (G_a + G_b)).  FP32 accumulation keeps the fitted sqrt(N) term.
double kappa_data(int prec, int natoms) {
  const double rn = std::sqrt((double)std::max(natoms, 16));
  const double input = 2.4e-7;
  switch (prec) {
    case kFP32: return 2.0 * (5.5e-8 * rn) + input;
    case kBF16: return 2.0 * (3.0e-8 * rn) + 5.0 * 1.5259e-5 + input;
    default: return 0.0;
  }
}

size_t operand_bytes(int prec, int n_tiles, int n_groups) {
  return ((size_t)n_tiles * n_groups + kOperandTailGroups) * blocks_per_group(prec) * 64 * 16;
}

int padded_tiles(int64_t n) {
  const int64_t tiles = (n + kTile - 1) / kTile;
  return (int)((tiles + kTileAlign - 1) / kTileAlign * kTileAlign);
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

}  // namespace

void set_error_sink(std::string* sink) { tl_err_sink = sink; }

void set_ctx_error(const mdsctk_knn_ctx* c, const std::string& text) { tl_ctx_err[c] = text; }

int fail(mdsctk_knn_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (tl_err_sink)
    *tl_err_sink = buf;
  else if (c)
    tl_ctx_err[c] = buf;
  return code;
}

int hip_fail(mdsctk_knn_ctx* c, hipError_t e, const char* what) {
  return fail(c, e == hipErrorOutOfMemory ? MDSCTK_ERR_NOMEM : MDSCTK_ERR_HIP, "%s: %s", what,
              hipGetErrorString(e));
}

#define HIP_TRY(c, expr)                                   \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail((c), e_, #expr); \
  } while (0)

int dalloc(mdsctk_knn_ctx* c, void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  bytes = (bytes + 255) & ~size_t(255);
  {
    std::lock_guard<std::mutex> g(c->mem_mu);
    if (c->budget && c->held + bytes > c->budget)
      return fail(c, MDSCTK_ERR_NOMEM,
                  "%s: %zu bytes would exceed the HBM budget (%zu held of %zu)", what, bytes,
                  c->held, c->budget);
    c->held += bytes;  // reserved before the allocation so that threads cannot overshoot
  }
  hipError_t e = hipMalloc(p, bytes);
  std::lock_guard<std::mutex> g(c->mem_mu);
  if (e != hipSuccess) {
    *p = nullptr;
    c->held -= bytes;
    return fail(c, MDSCTK_ERR_NOMEM, "%s: hipMalloc(%zu) failed: %s", what, bytes,
                hipGetErrorString(e));
  }
  c->allocs[*p] = bytes;
  return 0;
}

void dfree(mdsctk_knn_ctx* c, void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> g(c->mem_mu);
    auto it = c->allocs.find(p);
    if (it != c->allocs.end()) {
      c->held -= it->second;
      c->allocs.erase(it);
    }
  }
  (void)hipFree(p);
}

// Grow-only workspace buffer.
int ensure(mdsctk_knn_ctx* c, void** p, size_t* have, size_t need, const char* what) {
  if (*have >= need && *p) return 0;
  dfree(c, *p);
  *p = nullptr;
  *have = 0;
  int rc = dalloc(c, p, need, what);
  if (rc) return rc;
  *have = need;
  return 0;
}

namespace {

void free_frames(mdsctk_knn_ctx* c, DevFrames& f) {
  if (f.owns_xyz) dfree(c, f.xyz);
  dfree(c, f.x64);
  dfree(c, f.aux);
  dfree(c, f.gmax);
  dfree(c, f.cen);
  dfree(c, f.G);
  dfree(c, f.Gt);
  dfree(c, f.op);
  dfree(c, f.Gt_q);
  dfree(c, f.op_q);
  dfree(c, f.rot);
  f = DevFrames();
}

// Allocate the derived arrays of a frame set (xyz handled by the caller).
int alloc_frames(mdsctk_knn_ctx* c, DevFrames& f, int64_t n, int natoms, int prec) {
  f.n = n;
  f.natoms = natoms;
  f.prec = prec == kExact ? kFP32 : prec;
  f.n_tiles = padded_tiles(n);
  f.cap_tiles = f.n_tiles;
  f.n_groups = operand_groups(f.prec, natoms);
  const size_t rows = (size_t)f.n_tiles * kTile;
  int rc;
  if ((rc = dalloc(c, (void**)&f.cen, rows * 3 * sizeof(double), "centroids"))) return rc;
  if ((rc = dalloc(c, (void**)&f.G, rows * sizeof(double), "G terms"))) return rc;
  if ((rc = dalloc(c, (void**)&f.Gt, rows * sizeof(double), "G~ terms"))) return rc;
  if (prec != kExact &&
      (rc = dalloc(c, (void**)&f.op, operand_bytes(f.prec, f.n_tiles, f.n_groups), "operand")))
    return rc;
  if (prec != kExact && c->w.metric == kMetricRMSD &&
      (rc = dalloc(c, (void**)&f.rot, rows * 9 * sizeof(double), "principal axes")))
    return rc;
  return 0;
}

// The split-BF16 form of frames dithers its operands (operand_kernel's
// atom_scale): its window then no longer pays for the coherent error of
// aligned frames.
bool dithered(const mdsctk_knn_ctx* c) { return c->prec == kBF16 && c->w.metric == kMetricRMSD; }

// resident: the reference set (reference-side operand, plus the query-side
// one for self queries when dithered); otherwise a query batch.  axes: the
// operands hold each frame in its principal axes (fitted scores; f.rot).
int prep_frames(mdsctk_knn_ctx* c, DevFrames& f, hipStream_t st, bool resident, bool axes) {
  double* rot = axes ? f.rot : nullptr;
  if (axes && !rot) return fail(c, MDSCTK_ERR_ARG, "prep: no principal-axes buffer");
  // The exact form needs only centroids and G (f.op is not allocated).
  if (c->prec == kExact || !dithered(c)) {
    HIP_TRY(c, launch_prep(f, c->w, 0, nullptr, 0, st, nullptr, nullptr,
                           c->prec == kExact ? nullptr : rot));
    return 0;
  }
  if (!resident) {
    HIP_TRY(c, launch_prep(f, c->w, -1, nullptr, 0, st, nullptr, nullptr, rot));
    return 0;
  }
  const size_t rows = (size_t)f.n_tiles * kTile;
  int rc;
  if (!f.op_q &&
      (rc = dalloc(c, (void**)&f.op_q, operand_bytes(f.prec, f.n_tiles, f.n_groups),
                   "query-side operand")))
    return rc;
  if (!f.Gt_q && (rc = dalloc(c, (void**)&f.Gt_q, rows * sizeof(double), "query-side G~")))
    return rc;
  const DevFrames q = query_view(f);
  HIP_TRY(c, launch_prep(f, c->w, +1, &q, -1, st, nullptr, nullptr, rot));
  return 0;
}

// max Gt of a resident set, kept on the device for every later window.
int resident_gmax(mdsctk_knn_ctx* c, DevFrames& f, hipStream_t st) {
  if (c->prec == kExact) return 0;  // no contraction, no window
  int rc = dalloc(c, (void**)&f.gmax, sizeof(double), "max G term");
  if (rc) return rc;
  HIP_TRY(c, launch_max_gt(f, f.gmax, st));
  return 0;
}

// Fitted RMSD scores in a contraction form take the resident operands in
// principal axes, and query frames are prepared the same way.
bool axes_form(const mdsctk_knn_ctx* c) { return c->w.metric == kMetricRMSD && c->prec != kExact; }

// The resident operands a query with these flags contracts against.
const DevFrames& ref_ops(const mdsctk_knn_ctx* c, bool nofit) {
  return (c->ref_axes && nofit) ? c->ref_plain : c->ref;
}

void free_plain(mdsctk_knn_ctx* c) {
  DevFrames& p = c->ref_plain;
  dfree(c, p.cen);
  dfree(c, p.G);
  dfree(c, p.Gt);
  dfree(c, p.op);
  dfree(c, p.op_q);
  dfree(c, p.Gt_q);
  dfree(c, p.gmax);
  p = DevFrames();
}

}  // namespace

// The --nofit operands of the resident set (frames as given), built on first
// use: their own centroids, G, G~, operands and max G~; the coordinates are
// the resident set's.  Synchronous, so that a batch on any stream may use
// them once this returns.
int ensure_plain(mdsctk_knn_ctx* c, hipStream_t st) {
  if (!c->ref_axes) return 0;
  std::lock_guard<std::mutex> g(c->plain_mu);
  if (c->ref_plain.op) return 0;
  DevFrames p = c->ref;
  p.owns_xyz = false;
  p.cen = p.G = p.Gt = p.Gt_q = p.gmax = p.rot = nullptr;
  p.op = p.op_q = nullptr;
  const size_t rows = (size_t)p.n_tiles * kTile;
  const size_t opb = operand_bytes(p.prec, p.n_tiles, p.n_groups);
  int rc = dalloc(c, (void**)&p.cen, rows * 3 * sizeof(double), "plain centroids");
  if (!rc) rc = dalloc(c, (void**)&p.G, rows * sizeof(double), "plain G terms");
  if (!rc) rc = dalloc(c, (void**)&p.Gt, rows * sizeof(double), "plain G~ terms");
  if (!rc) rc = dalloc(c, (void**)&p.op, opb, "plain operand");
  if (!rc && c->ref.op_q) rc = dalloc(c, (void**)&p.op_q, opb, "plain query-side operand");
  if (!rc && c->ref.op_q) rc = dalloc(c, (void**)&p.Gt_q, rows * sizeof(double), "plain query-side G~");
  c->ref_plain = p;
  if (rc) {
    free_plain(c);
    return rc;
  }
  if ((rc = prep_frames(c, c->ref_plain, st, true, false)) ||
      (rc = resident_gmax(c, c->ref_plain, st))) {
    free_plain(c);
    return rc;
  }
  HIP_TRY(c, hipStreamSynchronize(st));
  return 0;
}

namespace {

int set_weights(mdsctk_knn_ctx* c, const float* weights, int natoms) {
  std::vector<float> w(natoms), sw(natoms);
  double W = 0;
  for (int i = 0; i < natoms; ++i) {
    const float wi = weights ? weights[i] : 1.0f;
    if (!(wi >= 0.0f) || !std::isfinite(wi))
      return fail(c, MDSCTK_ERR_ARG, "weight %d is negative or not finite", i);
    w[i] = wi;
    sw[i] = (float)std::sqrt((double)wi);
    W += wi;
  }
  if (!(W > 0)) return fail(c, MDSCTK_ERR_ARG, "weights sum to zero");
  dfree(c, c->w.w);
  dfree(c, c->w.sw);
  c->w = DevWeights();
  int rc;
  if ((rc = dalloc(c, (void**)&c->w.w, natoms * sizeof(float), "weights"))) return rc;
  if ((rc = dalloc(c, (void**)&c->w.sw, natoms * sizeof(float), "weights"))) return rc;
  hipStream_t st = c->main.st;
  HIP_TRY(c, hipMemcpyAsync(c->w.w, w.data(), natoms * sizeof(float), hipMemcpyHostToDevice, st));
  HIP_TRY(c, hipMemcpyAsync(c->w.sw, sw.data(), natoms * sizeof(float), hipMemcpyHostToDevice, st));
  HIP_TRY(c, hipStreamSynchronize(st));
  c->w.natoms = natoms;
  c->w.W = W;
  return 0;
}

int set_reference_impl(mdsctk_knn_ctx* c, const float* xyz, bool on_device, int64_t n_ref,
                       int32_t natoms, const float* weights) {
  if (!c) return MDSCTK_ERR_ARG;
  if (!xyz || n_ref <= 0 || natoms <= 0)
    return fail(c, MDSCTK_ERR_ARG, "set_reference: need frames (n_ref=%lld natoms=%d)",
                (long long)n_ref, natoms);
  if (n_ref > INT32_MAX) return fail(c, MDSCTK_ERR_ARG, "n_ref exceeds the int32 index range");
  if (!server_idle(c))
    return fail(c, MDSCTK_ERR_STATE, "set_reference: served requests are still outstanding");
  server_stop(c);
  HIP_TRY(c, hipSetDevice(c->device));
  free_frames(c, c->ref);
  free_plain(c);
  c->ref_axes = false;
  release_mirror_lists(c);
  dfree(c, c->data_shift);
  c->data_shift = nullptr;
  c->has_ref = false;
  int rc = set_weights(c, weights, natoms);
  if (rc) return rc;
  const size_t xbytes = (size_t)n_ref * natoms * 3 * sizeof(float);
  if ((rc = dalloc(c, (void**)&c->ref.xyz, xbytes, "reference coordinates"))) return rc;
  HIP_TRY(c, hipMemcpyAsync(c->ref.xyz, xyz, xbytes,
                            on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                            c->main.st));
  c->ref.owns_xyz = true;
  if ((rc = alloc_frames(c, c->ref, n_ref, natoms, c->prec))) return rc;
  c->ref_axes = axes_form(c);
  if ((rc = prep_frames(c, c->ref, c->main.st, true, c->ref_axes))) return rc;
  if ((rc = resident_gmax(c, c->ref, c->main.st))) return rc;
  HIP_TRY(c, hipStreamSynchronize(c->main.st));
  c->has_ref = true;
  return 0;
}

// Host-side checks of data vectors: finite values, and for the correlation
// metric no constant vector (its standard deviation is zero and
// correlation_distance divides by it).
int check_data(mdsctk_knn_ctx* c, const double* x, int64_t n, int vdim, int metric,
               const char* who) {
  for (int64_t i = 0; i < n; ++i) {
    const double* r = x + i * (int64_t)vdim;
    double lo = r[0], hi = r[0];
    for (int e = 0; e < vdim; ++e) {
      if (!std::isfinite(r[e]))
        return fail(c, MDSCTK_ERR_ARG, "%s: vector %lld holds a non-finite value", who,
                    (long long)i);
      lo = std::min(lo, r[e]);
      hi = std::max(hi, r[e]);
    }
    if (metric == kMetricCorrelation && !(hi > lo))
      return fail(c, MDSCTK_ERR_ARG,
                  "%s: vector %lld is constant, so its correlation distance is undefined", who,
                  (long long)i);
  }
  return 0;
}

int set_data_reference_impl(mdsctk_knn_ctx* c, const double* x, int64_t n_ref, int32_t vdim,
                            int metric) {
  if (!c) return MDSCTK_ERR_ARG;
  if (metric != kMetricEuclidean && metric != kMetricCorrelation)
    return fail(c, MDSCTK_ERR_ARG, "set_data_reference: unknown metric %d", metric);
  if (!x || n_ref <= 0 || vdim <= 0 || (metric == kMetricCorrelation && vdim < 2))
    return fail(c, MDSCTK_ERR_ARG, "set_data_reference: need vectors (n_ref=%lld vdim=%d)",
                (long long)n_ref, vdim);
  if (n_ref > INT32_MAX) return fail(c, MDSCTK_ERR_ARG, "n_ref exceeds the int32 index range");
  if (c->prec == kFP16)
    return fail(c, MDSCTK_ERR_ARG,
                "set_data_reference: data vectors take the split-BF16, FP32 or exact form "
                "(FP16 cannot hold arbitrary value ranges)");
  int rc = check_data(c, x, n_ref, vdim, metric, "set_data_reference");
  if (rc) return rc;
  if (!server_idle(c))
    return fail(c, MDSCTK_ERR_STATE, "set_data_reference: served requests are still outstanding");
  server_stop(c);
  HIP_TRY(c, hipSetDevice(c->device));
  free_frames(c, c->ref);
  free_plain(c);
  c->ref_axes = false;
  release_mirror_lists(c);
  dfree(c, c->data_shift);
  c->data_shift = nullptr;
  c->has_ref = false;
  const int natoms = (vdim + 2) / 3;
  if ((rc = set_weights(c, nullptr, natoms))) return rc;
  c->w.metric = metric;
  c->w.out_scale = 1.0;
  c->w.W = metric == kMetricCorrelation ? 4.0 : 1.0;
  DevFrames& f = c->ref;
  hipStream_t st = c->main.st;
  if ((rc = dalloc(c, (void**)&f.x64, (size_t)n_ref * vdim * sizeof(double), "reference vectors")))
    return rc;
  HIP_TRY(c, hipMemcpyAsync(f.x64, x, (size_t)n_ref * vdim * sizeof(double),
                            hipMemcpyHostToDevice, st));
  if (metric == kMetricCorrelation &&
      (rc = dalloc(c, (void**)&f.aux, (size_t)n_ref * 2 * sizeof(double), "vector statistics")))
    return rc;
  if ((rc = dalloc(c, (void**)&f.xyz, (size_t)n_ref * natoms * 3 * sizeof(float),
                   "reference operand source")))
    return rc;
  f.owns_xyz = true;
  if ((rc = alloc_frames(c, f, n_ref, natoms, c->prec))) return rc;
  f.vdim = vdim;
  if (metric == kMetricEuclidean) {
    if ((rc = dalloc(c, (void**)&c->data_shift, (size_t)vdim * sizeof(double), "mean vector")))
      return rc;
    HIP_TRY(c, launch_column_mean(f.x64, n_ref, vdim, c->data_shift, st));
  }
  HIP_TRY(c, launch_data_prep(f, metric, c->data_shift, st));
  if ((rc = prep_frames(c, f, st, true, false))) return rc;
  if ((rc = resident_gmax(c, f, st))) return rc;
  HIP_TRY(c, hipStreamSynchronize(st));
  c->has_ref = true;
  return 0;
}

}  // namespace

int slot_init(mdsctk_knn_ctx* c, Slot& s) {
  HIP_TRY(c, hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
  for (auto& e : s.ev) HIP_TRY(c, hipEventCreate(&e));
  int rc;
  if ((rc = dalloc(c, (void**)&s.n_cand, 256, "scratch"))) return rc;
  return 0;
}

void release_mirror_lists(mdsctk_knn_ctx* c) {
  Slot& s = c->main;
  if (!s.ws_sym) return;
  if (s.st) (void)hipStreamSynchronize(s.st);
  dfree(c, s.ws_sym);
  s.ws_sym = nullptr;
  s.ws_sym_bytes = 0;
}

void slot_release(mdsctk_knn_ctx* c, Slot& s) {
  if (s.st) (void)hipStreamSynchronize(s.st);
  free_frames(c, s.fit);
  dfree(c, s.ws_lists);
  dfree(c, s.ws_small);
  dfree(c, s.n_cand);
  dfree(c, s.ws_out);
  dfree(c, s.ws_rows);
  dfree(c, s.ws_sym);
  dfree(c, s.ws_relist);
  dfree(c, s.ws_sort);
  for (auto& e : s.ev)
    if (e) (void)hipEventDestroy(e);
  if (s.st) (void)hipStreamDestroy(s.st);
  s = Slot();
}

namespace {
// One epoch for the whole process, so that request times of different
// contexts compare directly.
const std::chrono::steady_clock::time_point g_clock_epoch = std::chrono::steady_clock::now();
}  // namespace

double clock_ms(const mdsctk_knn_ctx*) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                   g_clock_epoch)
      .count();
}

int resolve_shape(mdsctk_knn_ctx* c, int32_t k, uint32_t flags, QueryShape* q) {
  const int64_t n_ref = c->ref.n;
  // data vectors take the trace of the 3x3 block (their dot product)
  q->nofit = (flags & MDSCTK_KNN_NOFIT) || c->w.metric != kMetricRMSD;
  q->full = flags & MDSCTK_KNN_FULL_ROWS;
  q->rank_off = (flags & MDSCTK_KNN_KEEP_RANK0) ? 0 : 1;
  if (!q->full && k < 0) return fail(c, MDSCTK_ERR_ARG, "query: k must be >= 0");
  // knn_rms.cpp:224-225 clamps k to n_ref-1; k+1 follows the clamp here.
  q->k_out = q->full ? (int)n_ref : (int)std::min<int64_t>(k, n_ref - 1);
  if (!q->full && q->rank_off == 0) q->k_out = (int)std::min<int64_t>(k, n_ref);
  q->k1 = q->full ? 0 : q->k_out + q->rank_off;
  q->fused = !q->full && c->prec != kExact && list_capacity_for(q->k1) > 0;
  return 0;
}

int slot_frames(mdsctk_knn_ctx* c, Slot& s, int64_t n_fit) {
  DevFrames& f = s.fit;
  const int natoms = c->ref.natoms;
  const int want_prec = c->prec == kExact ? kFP32 : c->prec;
  const int metric = c->w.metric;
  const int vdim = metric == kMetricRMSD ? 0 : c->ref.vdim;
  if (f.cap_tiles == 0 || (size_t)n_fit > s.fit_cap_frames || s.fit_natoms != natoms ||
      f.prec != want_prec || s.fit_vdim != vdim || s.fit_metric != metric) {
    free_frames(c, f);
    s.fit_cap_frames = 0;
    int rc;
    if ((rc = dalloc(c, (void**)&f.xyz, (size_t)std::max<int64_t>(n_fit, 1) * natoms * 3 * sizeof(float),
                     "query coordinates")))
      return rc;
    f.owns_xyz = true;
    if ((rc = alloc_frames(c, f, std::max<int64_t>(n_fit, 1), natoms, c->prec))) return rc;
    const size_t cap = (size_t)std::max<int64_t>(n_fit, 1);
    if (vdim) {
      if ((rc = dalloc(c, (void**)&f.x64, cap * vdim * sizeof(double), "query vectors"))) return rc;
      if (metric == kMetricCorrelation &&
          (rc = dalloc(c, (void**)&f.aux, cap * 2 * sizeof(double), "vector statistics")))
        return rc;
    }
    f.vdim = vdim;
    s.fit_cap_frames = cap;
    s.fit_natoms = natoms;
    s.fit_vdim = vdim;
    s.fit_metric = metric;
  }
  // Reuse the allocation: a smaller batch only changes the row and tile
  // counts (the kernels walk n_tiles; the slack beyond it is never read).
  f.n = n_fit;
  f.n_tiles = std::max(kTileAlign, padded_tiles(n_fit));
  return 0;
}

int prep_slot(mdsctk_knn_ctx* c, Slot& s, bool nofit) {
  return prep_frames(c, s.fit, s.st, false, c->ref_axes && !nofit);
}

int stage_input(mdsctk_knn_ctx* c, Slot& s, const void* in, bool on_device, int64_t n_fit,
                bool nofit) {
  int rc = slot_frames(c, s, n_fit);
  if (rc) return rc;
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  if (c->w.metric == kMetricRMSD) {
    HIP_TRY(c, hipMemcpyAsync(s.fit.xyz, in, (size_t)n_fit * c->ref.natoms * 3 * sizeof(float),
                              kind, s.st));
  } else {
    HIP_TRY(c, hipMemcpyAsync(s.fit.x64, in, (size_t)n_fit * c->ref.vdim * sizeof(double), kind,
                              s.st));
    HIP_TRY(c, launch_data_prep(s.fit, c->w.metric, c->data_shift, s.st));
  }
  return prep_slot(c, s, nofit);
}

int contract_grid(const mdsctk_knn_ctx* c) {
  // workgroups resident at once: two 4-wave or one 8-wave per CU
  const int ft = contract_fit_tiles(c->prec == kExact ? kFP32 : c->prec);
  return std::max(8, (c->n_cu * (ft == 2 ? 2 : 1)) & ~7);
}

int fused_phases(const mdsctk_knn_ctx* c, int fit_tiles) {
  const int ft = contract_fit_tiles(c->prec == kExact ? kFP32 : c->prec);
  const int fit_blocks = std::max(1, fit_tiles / ft);
  const int ref_pairs = c->ref.n_tiles / kRefTilesPerWG;
  const int grid = contract_grid(c);
  int phases = c->forced_phases;
  if (phases <= 0) {
    // enough (fit block, phase) items to fill the grid, at least 8 phases
    phases = std::max(8, (grid + fit_blocks - 1) / fit_blocks);
    int p2 = 1;
    while (p2 < phases) p2 <<= 1;
    phases = p2;
  }
  return std::max(1, std::min(phases, ref_pairs));
}

size_t fused_list_entries(const mdsctk_knn_ctx* c, int64_t n_fit, int k1) {
  const int phases = fused_phases(c, std::max(kTileAlign, padded_tiles(n_fit)));
  return (size_t)n_fit * phases * 4 * list_capacity_for(k1);
}

int plan_fused(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, FusedPlan* p, Carve* cv) {
  const DevFrames& fit = s.fit;
  const int64_t n_fit = fit.n;
  const int cap = list_capacity_for(q.k1);
  p->grid = contract_grid(c);
  p->phases = fused_phases(c, fit.n_tiles);
  auto ent_bytes_for = [&](int phases) { return (size_t)n_fit * phases * 4 * cap * sizeof(uint2); };
  auto small_bytes_for = [&](int phases) {
    return (size_t)n_fit * phases * 4 * (sizeof(int) + sizeof(float)) +
           (size_t)n_fit * (sizeof(unsigned) + sizeof(float) + sizeof(int)) + 256;
  };
  void* ent = nullptr;
  void* small_p = nullptr;
  if (cv) {
    // Fewer phases mean fewer lists per query (and fewer, longer work
    // items); one exact row's scratch stays in reserve for the fallback.
    const size_t reserve = Carve::round((size_t)c->ref.n * sizeof(double) + sizeof(int));
    auto need = [&](int ph) {
      return Carve::round(ent_bytes_for(ph)) + Carve::round(small_bytes_for(ph)) + reserve;
    };
    while (p->phases > 1 && need(p->phases) > cv->left()) p->phases /= 2;
    if (need(p->phases) > cv->left())
      return fail(c, MDSCTK_ERR_ARG,
                  "query (fixed workspace): %lld query frames with k+1=%d need %zu bytes of workspace, "
                  "more than the %zu left of the %zu-byte fixed region",
                  (long long)n_fit, q.k1, need(p->phases), cv->left(), cv->size);
    ent = cv->take(ent_bytes_for(p->phases));
    small_p = cv->take(small_bytes_for(p->phases));
  } else {
    // Under an HBM budget, fewer phases (fewer lists per query) until the
    // lists fit what the rest of the context leaves of it (these buffers
    // are regrown, so their current size counts as free)
    std::lock_guard<std::mutex> lk(c->mem_mu);
    if (c->budget) {
      const size_t mine = s.ws_lists_bytes + s.ws_small_bytes;
      const size_t other = c->held - std::min(c->held, mine);
      const size_t room = c->budget > other ? c->budget - other : 0;
      while (p->phases > 1 && ent_bytes_for(p->phases) + small_bytes_for(p->phases) > room)
        p->phases /= 2;
    }
  }
  // The phase count is final here: the list layout below follows it.
  Lists& L = p->L;
  L = Lists();
  L.nl = p->phases * 4;
  L.cap = cap;
  if (!cv) {
    int rc = ensure(c, &s.ws_lists, &s.ws_lists_bytes, ent_bytes_for(p->phases),
                    "candidate lists");
    if (rc) return rc;
    rc = ensure(c, &s.ws_small, &s.ws_small_bytes, small_bytes_for(p->phases), "list headers");
    if (rc) return rc;
    ent = s.ws_lists;
    small_p = s.ws_small;
  }
  const size_t n_hdr = (size_t)n_fit * L.nl;
  L.ent = (uint2*)ent;
  char* sp = (char*)small_p;
  L.cnt = (int*)sp; sp += n_hdr * sizeof(int);
  L.ovf = (float*)sp; sp += n_hdr * sizeof(float);
  L.thr = (unsigned*)sp; sp += n_fit * sizeof(unsigned);
  p->eps2 = (float*)sp; sp += n_fit * sizeof(float);
  p->d_fallback = (int*)sp; sp += n_fit * sizeof(int);
  if ((size_t)(sp - (char*)small_p) > small_bytes_for(p->phases))
    return fail(c, MDSCTK_ERR_ARG, "internal: list headers for %d lists per query exceed their %zu bytes",
                L.nl, small_bytes_for(p->phases));
  p->kappa = (c->w.metric == kMetricRMSD ? kappa_for(c->prec, c->ref.natoms)
                                         : kappa_data(c->prec, c->ref.natoms)) *
             c->eps_scale;
  return 0;
}

int enqueue_fused(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, const FusedPlan& p,
                  double* d_dist, int* d_idx, bool timing) {
  const DevFrames& fit = s.fit;
  HIP_TRY(c, launch_init_lists(p.L, fit.n, s.st));
  const DevFrames& R = ref_ops(c, q.nofit);
  HIP_TRY(c, launch_eps(fit, R, p.kappa, p.eps2, s.n_cand, s.st));
  if (timing) HIP_TRY(c, hipEventRecord(s.ev[2], s.st));
  HIP_TRY(c, launch_contract(R, fit, p.L, p.eps2, q.k1, p.phases, q.nofit, nullptr, p.grid,
                             s.st));
  if (timing) HIP_TRY(c, hipEventRecord(s.ev[3], s.st));
  HIP_TRY(c, launch_merge(c->ref, fit, c->w, p.L, p.eps2, q.k1, q.k_out, q.rank_off, q.nofit,
                          d_dist, d_idx, p.d_fallback, s.n_cand, nullptr, s.st));
  if (timing) HIP_TRY(c, hipEventRecord(s.ev[4], s.st));
  return 0;
}

int launch_fused(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, double* d_dist, int* d_idx,
                 FusedInfo* info, Carve* cv) {
  FusedPlan p;
  int rc = plan_fused(c, s, q, &p, cv);
  if (rc) return rc;
  if ((rc = enqueue_fused(c, s, q, p, d_dist, d_idx, true))) return rc;
  info->phases = p.phases;
  info->cap = p.L.cap;
  info->kappa = p.kappa;
  info->d_fallback = p.d_fallback;
  info->pairs_computed = s.fit.n * c->ref.n;
  return 0;
}

int launch_fused_sym(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, double* d_dist, int* d_idx,
                     FusedInfo* info) {
  const DevFrames& fit = s.fit;
  const int64_t n = fit.n;
  if (n != c->ref.n || fit.n_tiles != c->ref.n_tiles || n < 1) return 1;
  const int cap = list_capacity_for(q.k1);
  if (cap == 0) return 1;
  const int ref_pairs = c->ref.n_tiles / kRefTilesPerWG;
  // Each fit frame sees about half of its pairs, so half the list form's
  // phases (longer lists per phase, tighter list thresholds).
  int phases = fused_phases(c, fit.n_tiles);
  if (c->forced_phases <= 0) phases = std::max(1, std::min(ref_pairs, std::max(4, phases / 2)));
  const int nl = phases * 4;
  // Landmark pass: every lstride-th reference tile pair.  The (k+1)-th
  // smallest key of L landmarks ranks about (k+1) n / L among a row's n
  // keys, and a row may receive up to that many mirrored scores before its
  // own lists tighten its threshold: L >= 2 (k+1) n / cap keeps that at half
  // the mirrored capacity (3200 frames, 3 % of the pairs, at 200k frames
  // and k = 64).  At least 4 (k+1) landmarks; stride kLandmarkStride at
  // most, 8 at least (the pass then costs at most a quarter of the
  // symmetric contraction, and overflowing rows are recomputed).
  const int64_t lm_target = std::max<int64_t>(
      4 * (int64_t)q.k1, 2 * (int64_t)q.k1 * n / std::max(1, c->mirror_cap));
  const int lstride = std::max(
      1, std::min<int>(kLandmarkStride,
                       std::max<int>(std::min(8, ref_pairs),
                                     (int)((int64_t)ref_pairs * 2 * kTile / lm_target))));
  const int lm_pairs = (ref_pairs + lstride - 1) / lstride;
  const int nlab = lm_pairs * 2 * kTile;
  const size_t ent_bytes = (size_t)n * nl * cap * sizeof(uint2);
  // Canonical form (fitted scores of frames, n >= 512): the main pass runs
  // over a copy of the set in landmark order, every frame in its principal
  // axes.  Frames near each other in the order are near each other in
  // space, so most 32 x 32 tiles hold only far pairs, which the
  // contraction's lambda_max bound rejects without the QCP; the axes make
  // that bound nearly exact for similar frames.  The copy: operands (both
  // dither sides), centroids, G, G~, rotations, windows, thresholds.
  const bool canon = !q.nofit && c->w.metric == kMetricRMSD && n >= 512;
  const bool dith = c->ref.op_q != nullptr;
  const size_t rows = (size_t)c->ref.n_tiles * kTile;
  const size_t opb = operand_bytes(c->ref.prec, c->ref.n_tiles, c->ref.n_groups);
  const size_t perm_bytes =
      canon ? Carve::round(rows * 3 * sizeof(double)) + Carve::round(rows * sizeof(double)) * 2 +
                  (dith ? Carve::round(rows * sizeof(double)) + Carve::round(opb) : 0) +
                  Carve::round(rows * 9 * sizeof(double)) + Carve::round(opb) +
                  Carve::round((size_t)n * sizeof(float)) +
                  Carve::round((size_t)n * sizeof(unsigned)) + 512
            : 0;
  const size_t mir_bytes = (size_t)n * c->mirror_cap * sizeof(uint2);
  const size_t sym_bytes = Carve::round(mir_bytes) + perm_bytes;
  const size_t small = (size_t)n * nl * (sizeof(int) + sizeof(float)) +
                       (size_t)n * 7 * sizeof(int) + (size_t)nlab * sizeof(int) + 1024;
  // Room: the free device memory plus what the buffers already hold, less a
  // reserve for everything else; and the context's budget.
  size_t free_b = 0, total_b = 0;
  HIP_TRY(c, hipMemGetInfo(&free_b, &total_b));
  const size_t have = s.ws_sym_bytes + s.ws_lists_bytes + s.ws_small_bytes;
  const size_t need = ent_bytes + sym_bytes + small;
  const size_t reserve = size_t(2) << 30;
  if (need + reserve > free_b + have) return 1;
  {
    std::lock_guard<std::mutex> g(c->mem_mu);
    if (c->budget && c->held - std::min(c->held, have) + need > c->budget) return 1;
  }
  int rc = ensure(c, &s.ws_sym, &s.ws_sym_bytes, sym_bytes, "mirrored lists");
  if (!rc) rc = ensure(c, &s.ws_lists, &s.ws_lists_bytes, ent_bytes, "candidate lists");
  if (!rc) rc = ensure(c, &s.ws_small, &s.ws_small_bytes, small, "list headers");
  if (rc == MDSCTK_ERR_NOMEM) {  // the list path instead, without the mirrored lists
    release_mirror_lists(c);
    tl_ctx_err.erase(c);
    return 1;
  }
  if (rc) return rc;
  Lists L;
  L.nl = nl;
  L.cap = cap;
  L.ent = (uint2*)s.ws_lists;
  char* sp = (char*)s.ws_small;
  double* gmax_q = (double*)sp; sp += 256;
  unsigned long long* msum = (unsigned long long*)sp; sp += 256;
  L.cnt = (int*)sp; sp += (size_t)n * nl * sizeof(int);
  L.ovf = (float*)sp; sp += (size_t)n * nl * sizeof(float);
  L.thr = (unsigned*)sp; sp += n * sizeof(unsigned);
  float* eps2 = (float*)sp; sp += n * sizeof(float);
  int* label = (int*)sp; sp += n * sizeof(int);
  int* order = (int*)sp; sp += n * sizeof(int);
  int* d_fallback = (int*)sp; sp += n * sizeof(int);
  int* counts = (int*)sp; sp += (size_t)nlab * sizeof(int);
  Mirror M;
  M.ent = (uint2*)s.ws_sym;
  M.cap = c->mirror_cap;
  M.cnt = (int*)sp; sp += n * sizeof(int);
  M.ovf = (float*)sp;
  const double kappa = (c->w.metric == kMetricRMSD ? kappa_for(c->prec, c->ref.natoms)
                                                   : kappa_data(c->prec, c->ref.natoms)) *
                       c->eps_scale;
  const int grid = contract_grid(c);
  const DevFrames& R = ref_ops(c, q.nofit);
  HIP_TRY(c, launch_eps_sym(fit, R, kappa, gmax_q, eps2, s.n_cand, s.st));
  HIP_TRY(c, hipEventRecord(s.ev[2], s.st));
  // landmark pass into the first 4 lists' worth of the list workspace
  Lists LL = L;
  LL.nl = 4;
  HIP_TRY(c, launch_init_lists(LL, n, s.st));
  HIP_TRY(c, launch_contract(R, fit, LL, eps2, q.k1, lstride, q.nofit, nullptr, grid, s.st,
                             nullptr, 1));
  HIP_TRY(c, launch_landmark_cap(LL, n, q.k1, eps2, lstride, L.thr, label, s.st));
  // rows grouped by nearest landmark: the merge takes them in this order (it
  // pays where candidate frames no longer fit the caches), and the canonical
  // form contracts the frames in it
  const bool ordered = n >= 512;
  if (ordered) HIP_TRY(c, launch_row_order(label, n, nlab, counts, order, s.st));
  // the main pass's frames, windows and thresholds
  DevFrames mref = R, mfit = fit;
  Lists LM = L;
  float* meps2 = eps2;
  if (canon) {
    Carve cp;
    cp.base = (char*)s.ws_sym + Carve::round(mir_bytes);
    cp.size = perm_bytes;
    mref.xyz = nullptr;
    mref.owns_xyz = false;
    mref.cen = (double*)cp.take(rows * 3 * sizeof(double));
    mref.G = (double*)cp.take(rows * sizeof(double));
    mref.Gt = (double*)cp.take(rows * sizeof(double));
    mref.op = (uint4*)cp.take(opb);
    mref.op_q = dith ? (uint4*)cp.take(opb) : nullptr;
    mref.Gt_q = dith ? (double*)cp.take(rows * sizeof(double)) : nullptr;
    double* rot = (double*)cp.take(rows * 9 * sizeof(double));
    meps2 = (float*)cp.take((size_t)n * sizeof(float));
    LM.thr = (unsigned*)cp.take((size_t)n * sizeof(unsigned));
    double* gm = (double*)cp.take(2 * sizeof(double));
    mref.gmax = gm;
    if (!mref.cen || !mref.G || !mref.Gt || !mref.op || !rot || !meps2 || !LM.thr || !gm ||
        (dith && (!mref.op_q || !mref.Gt_q)))
      return fail(c, MDSCTK_ERR_ARG, "symmetric self query: canonical workspace carve");
    const DevFrames mq = query_view(mref);
    HIP_TRY(c, launch_prep(mref, c->w, dith ? +1 : 0, dith ? &mq : nullptr, -1, s.st,
                           c->ref.xyz, order, rot));
    HIP_TRY(c, launch_max_gt(mref, gm, s.st));
    mfit = mq;
    HIP_TRY(c, launch_eps_sym(mfit, mref, kappa, gm + 1, meps2, nullptr, s.st));
    // the landmark thresholds in the new order, widened by the new window
    // (the landmark keys came from the original operand, the main pass's
    // from the rotated one: both within their windows of the exact scores)
    HIP_TRY(c, launch_permute_thresholds(L.thr, order, meps2, n, LM.thr, s.st));
  }
  HIP_TRY(c, launch_init_lists(LM, n, s.st, false));
  HIP_TRY(c, launch_mirror_init(M, n, s.st));
  HIP_TRY(c, hipEventRecord(s.ev[6], s.st));
  HIP_TRY(c, launch_contract(mref, mfit, LM, meps2, q.k1, phases, q.nofit, nullptr, grid, s.st,
                             &M));
  HIP_TRY(c, hipEventRecord(s.ev[3], s.st));
  HIP_TRY(c, launch_mirror_sum(M, n, msum, s.st));
  HIP_TRY(c, launch_merge(c->ref, fit, c->w, LM, meps2, q.k1, q.k_out, q.rank_off, q.nofit, d_dist,
                          d_idx, d_fallback, s.n_cand, (ordered && !canon) ? order : nullptr, s.st,
                          &M, canon ? order : nullptr));
  HIP_TRY(c, hipEventRecord(s.ev[4], s.st));
  // pairs of the upper block triangle: fit block fb against frames from
  // its own first frame on
  const int64_t blk = (int64_t)kTile * contract_fit_tiles(c->prec);
  int64_t pairs = 0;
  for (int64_t f0 = 0; f0 < n; f0 += blk) pairs += std::min(blk, n - f0) * (n - f0);
  info->phases = phases;
  info->cap = cap;
  info->kappa = kappa;
  info->d_fallback = d_fallback;
  info->pairs_computed = pairs;
  info->symmetric = true;
  info->pairs_landmark = n * std::min<int64_t>(n, (int64_t)lm_pairs * 2 * kTile);
  info->d_mirror_sum = msum;
  return 0;
}

int relist_rows(mdsctk_knn_ctx* c, Slot& s, const QueryShape& q, std::vector<int>& rows,
                double* d_dist, int* d_idx) {
  const int nr = (int)rows.size();
  if (nr == 0) return 0;
  const size_t out_elems = (size_t)nr * q.k_out;
  int rc = ensure(c, &s.ws_relist, &s.ws_relist_bytes,
                  Carve::round(nr * sizeof(int)) + out_elems * (sizeof(double) + sizeof(int)),
                  "relisted rows");
  if (rc) return rc;
  int* d_rows = (int*)s.ws_relist;
  double* t_dist = (double*)((char*)s.ws_relist + Carve::round(nr * sizeof(int)));
  int* t_idx = (int*)(t_dist + out_elems);
  HIP_TRY(c, hipMemcpyAsync(d_rows, rows.data(), nr * sizeof(int), hipMemcpyHostToDevice, s.st));
  if ((rc = slot_frames(c, s, nr))) return rc;
  HIP_TRY(c, launch_gather_frames(c->ref.xyz, d_rows, nr, c->ref.natoms, s.fit.xyz, s.st));
  if ((rc = prep_slot(c, s, q.nofit))) return rc;
  FusedInfo info;
  if ((rc = launch_fused(c, s, q, t_dist, t_idx, &info))) return rc;
  HIP_TRY(c, launch_scatter_rows(t_dist, t_idx, d_rows, nr, q.k_out, d_dist, d_idx, s.st));
  std::vector<int> fb(nr);
  HIP_TRY(c, hipMemcpyAsync(fb.data(), info.d_fallback, nr * sizeof(int), hipMemcpyDeviceToHost,
                            s.st));
  HIP_TRY(c, hipStreamSynchronize(s.st));
  std::vector<int> left;
  for (int i = 0; i < nr; ++i)
    if (fb[i]) left.push_back(rows[i]);
  rows.swap(left);
  return 0;
}

// Exact FP64 path for the listed fit rows: scores of whole rows in batches,
// then (s, index) selection.  rows == nullptr means rows [0, nrows).
int exact_rows_select(mdsctk_knn_ctx* c, Slot& s, const DevFrames& fit, const int* h_rows,
                      int nrows, const QueryShape& q, double* d_dist, int* d_idx, Carve* cv) {
  const int64_t n_ref = c->ref.n;
  const size_t row_bytes = (size_t)n_ref * sizeof(double);
  int batch = (int)std::max<int64_t>(1, std::min<int64_t>(nrows, (int64_t)(1ull << 30) / row_bytes));
  // k+1 beyond the LDS select: compaction and a segmented sort in device
  // memory (the reference has no bound on k, knn_rms.cpp:131, 242)
  const bool sorted_path = q.k1 > kSelectLdsMaxK1;
  void* ws = nullptr;
  std::unique_lock<std::mutex> lk(c->mem_mu);
  if (!cv && c->budget) {
    // Under an HBM budget the rows go through in batches that fit what the
    // rest of the context leaves of it (this scratch is regrown, so its
    // current size counts as free); the sort's own scratch is estimated
    const size_t mine = s.ws_rows_bytes + (sorted_path ? s.ws_sort_bytes : 0);
    const size_t other = c->held - std::min(c->held, mine);
    const size_t room = c->budget > other ? c->budget - other : 0;
    const size_t per_row = row_bytes + sizeof(int) + (sorted_path ? 24 * (size_t)q.k1 + 4 : 0);
    batch = (int)std::max<int64_t>(1, std::min<int64_t>(batch, (int64_t)(room / per_row)));
    // the sort's scratch also has a part the estimate leaves out (the radix
    // sort's own storage): shrink until rows and scratch fit together
    while (sorted_path && batch > 1 &&
           (row_bytes + sizeof(int)) * batch + sort_rows_scratch(batch, q.k1) > room)
      batch = std::max(1, batch * 3 / 4);
  }
  lk.unlock();
  if (cv) {
    const size_t room = cv->left() > 256 ? cv->left() - 256 : 0;
    batch = (int)std::min<int64_t>(batch, (int64_t)(room / (row_bytes + sizeof(int))));
    if (batch < 1)
      return fail(c, MDSCTK_ERR_ARG,
                  "query (fixed workspace): one exact row of %lld references does not fit the %zu bytes "
                  "left of the fixed region",
                  (long long)n_ref, cv->left());
    ws = cv->take(row_bytes * batch + sizeof(int) * batch);
  } else {
    int rc = ensure(c, &s.ws_rows, &s.ws_rows_bytes, row_bytes * batch + sizeof(int) * batch,
                    "exact-row scratch");
    if (rc) return rc;
    ws = s.ws_rows;
  }
  double* srow = (double*)ws;
  int* drows = (int*)((char*)ws + row_bytes * batch);
  size_t sort_bytes = 0;
  if (sorted_path) {
    if (cv)
      return fail(c, MDSCTK_ERR_ARG,
                  "query (fixed workspace): k+1=%d exceeds %d, which needs a sorting workspace",
                  q.k1, kSelectLdsMaxK1);
    sort_bytes = sort_rows_scratch(batch, q.k1);
    int rc = ensure(c, &s.ws_sort, &s.ws_sort_bytes, sort_bytes, "row sort scratch");
    if (rc) return rc;
  }
  std::vector<int> ids(batch);
  for (int b0 = 0; b0 < nrows; b0 += batch) {
    const int nb = std::min(batch, nrows - b0);
    for (int i = 0; i < nb; ++i) ids[i] = h_rows ? h_rows[b0 + i] : b0 + i;
    HIP_TRY(c, hipMemcpyAsync(drows, ids.data(), nb * sizeof(int), hipMemcpyHostToDevice, s.st));
    HIP_TRY(c, launch_exact_rows(c->ref, fit, c->w, drows, nb, q.nofit, srow, nullptr, 0, s.st));
    if (sorted_path)
      HIP_TRY(c, launch_select_rows_sorted(srow, n_ref, drows, nb, c->w, q.k1, q.k_out,
                                           q.rank_off, d_dist, d_idx, s.ws_sort, sort_bytes,
                                           s.st));
    else
      HIP_TRY(c, launch_select_rows(srow, n_ref, drows, nb, c->w, q.k1, q.k_out, q.rank_off,
                                    d_dist, d_idx, s.st));
    HIP_TRY(c, hipStreamSynchronize(s.st));
  }
  return 0;
}

namespace {

// in: query frames (float xyz, nm) or, for a context holding data vectors,
// FP64 vectors; nullptr = the resident set itself.
int query_impl(mdsctk_knn_ctx* c, const void* xyz, bool data, bool on_device, int64_t n_fit,
               int32_t k, uint32_t flags, double* out_dist, int32_t* out_idx,
               bool out_on_device, Carve* cv = nullptr) {
  if (!c) return MDSCTK_ERR_ARG;
  if (!c->has_ref) return fail(c, MDSCTK_ERR_STATE, "query: no reference set");
  if (data != (c->w.metric != kMetricRMSD))
    return fail(c, MDSCTK_ERR_STATE,
                data ? "data query: the context holds frames (use mdsctk_knn_query)"
                     : "query: the context holds data vectors (use mdsctk_data_query)");
  if (data && xyz && !on_device) {
    const int rc0 = check_data(c, (const double*)xyz, n_fit, c->ref.vdim, c->w.metric, "query");
    if (rc0) return rc0;
  }
  if (xyz == nullptr) n_fit = c->ref.n;
  if (n_fit < 0) return fail(c, MDSCTK_ERR_ARG, "query: negative frame count");
  if (n_fit > INT32_MAX) return fail(c, MDSCTK_ERR_ARG, "query: too many frames");
  QueryShape q;
  int rc = resolve_shape(c, k, flags, &q);
  if (rc) return rc;
  if (n_fit == 0 || q.k_out == 0) {
    c->stats = mdsctk_knn_stats();
    return q.k_out;
  }
  HIP_TRY(c, hipSetDevice(c->device));
  // A query that is not a self query does not need the self query's
  // mirrored lists; their memory goes back to the lists it does need.
  if (xyz && !cv) release_mirror_lists(c);
  // --nofit through the contraction: the resident operands as given (built
  // once).  Without them, in the allocation-free form (cv) or when they do
  // not fit, the query takes the exact rows instead.
  bool plain_missing = false;
  if (q.nofit && c->ref_axes) {
    std::lock_guard<std::mutex> g(c->plain_mu);
    plain_missing = !c->ref_plain.op;
  }
  if (plain_missing && q.fused && !q.full && !cv) {
    rc = ensure_plain(c, c->main.st);
    if (rc == 0) {
      plain_missing = false;
    } else if (rc == MDSCTK_ERR_NOMEM) {
      tl_ctx_err.erase(c);  // no room for them: the exact rows instead
    } else {
      return rc;
    }
  }
  Slot& s = c->main;
  mdsctk_knn_stats S = {};
  S.n_pairs = n_fit * c->ref.n;
  S.n_pairs_computed = S.n_pairs;
  HIP_TRY(c, hipEventRecord(s.ev[0], s.st));

  // Stage (or alias) the query frames and run their prep pass.  Without
  // allocation (cv) the query frames are read where the caller holds them
  // and their derived arrays are carved from the fixed region; the slot's
  // own frame set is left untouched.
  const DevFrames* fitp = &c->ref;
  Slot cs;
  if (cv) {
    cs = s;
    cs.fit = DevFrames();
    cs.n_cand = (int*)cv->take(sizeof(int));
    if (xyz) {
      DevFrames& f = cs.fit;
      f.n = n_fit;
      f.natoms = c->ref.natoms;
      f.prec = c->ref.prec;
      f.n_tiles = std::max(kTileAlign, padded_tiles(n_fit));
      f.cap_tiles = f.n_tiles;
      f.n_groups = c->ref.n_groups;
      f.xyz = (float*)xyz;
      f.owns_xyz = false;
      const size_t rows = (size_t)f.n_tiles * kTile;
      f.cen = (double*)cv->take(rows * 3 * sizeof(double));
      f.G = (double*)cv->take(rows * sizeof(double));
      f.Gt = (double*)cv->take(rows * sizeof(double));
      if (c->prec != kExact) f.op = (uint4*)cv->take(operand_bytes(f.prec, f.n_tiles, f.n_groups));
      const bool axes = c->ref_axes && !q.nofit;
      if (axes) f.rot = (double*)cv->take(rows * 9 * sizeof(double));
      if (!cs.n_cand || !f.cen || !f.G || !f.Gt || (c->prec != kExact && !f.op) ||
          (axes && !f.rot))
        return fail(c, MDSCTK_ERR_ARG,
                    "query (fixed workspace): %lld query frames of %d atoms do not fit the %zu-byte "
                    "fixed region",
                    (long long)n_fit, f.natoms, cv->size);
      if ((rc = prep_frames(c, f, cs.st, false, axes))) return rc;
      fitp = &cs.fit;
    }
  } else if (xyz) {
    if ((rc = stage_input(c, s, xyz, on_device, n_fit, q.nofit))) return rc;
    fitp = &s.fit;
  }
  Slot& ws = cv ? cs : s;
  HIP_TRY(c, hipEventRecord(s.ev[1], s.st));

  const size_t out_elems = (size_t)n_fit * q.k_out;
  double* d_dist = out_dist;
  int* d_idx = out_idx;
  if (!out_on_device) {
    rc = ensure(c, &s.ws_out, &s.ws_out_bytes, out_elems * (sizeof(double) + sizeof(int)),
                "output buffers");
    if (rc) return rc;
    d_dist = (double*)s.ws_out;
    d_idx = (int*)((char*)s.ws_out + out_elems * sizeof(double));
  }

  float ms_contract = 0, ms_merge = 0, ms_fallback = 0;
  // The fused pipeline runs on the slot's frames; a self query points the
  // slot at the resident set for its duration and, when it fits, scores
  // each unordered pair once (the symmetric form).  Without allocation a
  // self query takes the candidate lists.
  int fused_rc = 1;  // 1: the fused pipeline did not run
  FusedInfo info;
  if (!q.full && q.fused && !plain_missing) {
    DevFrames saved = ws.fit;
    if (!xyz) ws.fit = query_view(ref_ops(c, q.nofit));
    if (!xyz && c->self_sym && !cv) fused_rc = launch_fused_sym(c, ws, q, d_dist, d_idx, &info);
    if (fused_rc == 1) fused_rc = launch_fused(c, ws, q, d_dist, d_idx, &info, cv);
    if (!xyz) ws.fit = saved;
    if (fused_rc == MDSCTK_ERR_NOMEM && !cv) {
      // even one phase's lists do not fit: the exact rows, in batches that do
      tl_ctx_err.erase(c);
      fused_rc = 1;
    }
    if (fused_rc < 0) return fused_rc;
  }
  if (q.full) {
    HIP_TRY(c, launch_exact_rows(c->ref, *fitp, c->w, nullptr, (int)n_fit, q.nofit, nullptr,
                                 d_dist, c->ref.n, s.st));
    HIP_TRY(c, launch_identity_rows(n_fit, c->ref.n, d_idx, s.st));
    HIP_TRY(c, hipEventRecord(s.ev[2], s.st));
    HIP_TRY(c, hipEventSynchronize(s.ev[2]));
    ms_contract = elapsed(s.ev[1], s.ev[2]);
  } else if (fused_rc == 1) {
    HIP_TRY(c, hipEventRecord(s.ev[2], s.st));
    rc = exact_rows_select(c, ws, *fitp, nullptr, (int)n_fit, q, d_dist, d_idx, cv);
    if (rc) return rc;
    HIP_TRY(c, hipEventRecord(s.ev[3], s.st));
    HIP_TRY(c, hipEventSynchronize(s.ev[3]));
    ms_fallback = elapsed(s.ev[2], s.ev[3]);
    S.n_fallback_rows = n_fit;
  } else {
    std::vector<int> fb(n_fit);
    int ncand = 0;
    unsigned long long msum[2] = {0, 0};
    HIP_TRY(c, hipMemcpyAsync(fb.data(), info.d_fallback, n_fit * sizeof(int),
                              hipMemcpyDeviceToHost, s.st));
    HIP_TRY(c, hipMemcpyAsync(&ncand, ws.n_cand, sizeof(int), hipMemcpyDeviceToHost, s.st));
    if (info.d_mirror_sum)
      HIP_TRY(c, hipMemcpyAsync(msum, info.d_mirror_sum, sizeof msum, hipMemcpyDeviceToHost, s.st));
    HIP_TRY(c, hipStreamSynchronize(s.st));
    if (info.symmetric) {
      S.ms_landmark = elapsed(s.ev[2], s.ev[6]);
      ms_contract = elapsed(s.ev[6], s.ev[3]);
    } else {
      ms_contract = elapsed(s.ev[2], s.ev[3]);
    }
    ms_merge = elapsed(s.ev[3], s.ev[4]);
    S.n_candidates = ncand;
    S.ref_splits = info.phases;
    S.list_capacity = info.cap;
    S.eps_rel = info.kappa;
    S.n_pairs_computed = info.pairs_computed;
    S.n_pairs_landmark = info.pairs_landmark;
    S.n_mirrored = (int64_t)msum[0];
    S.n_mirrored_max = (int64_t)msum[1];
    std::vector<int> rows;
    for (int64_t i = 0; i < n_fit; ++i)
      if (fb[i]) rows.push_back((int)i);
    if (!rows.empty()) {
      HIP_TRY(c, hipEventRecord(s.ev[2], s.st));
      // rows of a symmetric self query whose mirrored list overflowed go
      // through the candidate lists first, the rest (and any they flag) to
      // the exact path
      if (info.symmetric) {
        S.n_relisted_rows = (int64_t)rows.size();
        if ((rc = relist_rows(c, s, q, rows, d_dist, d_idx))) return rc;
      }
      if (!rows.empty()) {
        rc = exact_rows_select(c, ws, *fitp, rows.data(), (int)rows.size(), q, d_dist, d_idx, cv);
        if (rc) return rc;
      }
      HIP_TRY(c, hipEventRecord(s.ev[3], s.st));
      HIP_TRY(c, hipEventSynchronize(s.ev[3]));
      ms_fallback = elapsed(s.ev[2], s.ev[3]);
      S.n_fallback_rows = (int64_t)rows.size();
    }
  }
  if (!out_on_device) {
    HIP_TRY(c, hipMemcpyAsync(out_dist, d_dist, out_elems * sizeof(double), hipMemcpyDeviceToHost,
                              s.st));
    HIP_TRY(c, hipMemcpyAsync(out_idx, d_idx, out_elems * sizeof(int), hipMemcpyDeviceToHost,
                              s.st));
  }
  HIP_TRY(c, hipEventRecord(s.ev[5], s.st));
  HIP_TRY(c, hipEventSynchronize(s.ev[5]));
  // While serving runs, a self query does not keep its mirrored lists: the
  // memory goes back to the serving slots' workspace.
  bool serving;
  {
    std::lock_guard<std::mutex> g(c->server_mu);
    serving = c->server != nullptr;
  }
  if (serving && !cv) release_mirror_lists(c);
  S.ms_prep = elapsed(s.ev[0], s.ev[1]);
  S.ms_contract = ms_contract;
  S.ms_merge = ms_merge;
  S.ms_fallback = ms_fallback;
  S.ms_total = elapsed(s.ev[0], s.ev[5]);
  c->stats = S;
  return q.k_out;
}

}  // namespace
}  // namespace mdsctk

namespace mdsctk {
namespace {

int debug_scores_impl(mdsctk_knn_ctx* c, const void* xyz_nm, bool data, int64_t n_fit,
                      uint32_t flags, double* approx_s, double* exact_s) {
  if (!c) return MDSCTK_ERR_ARG;
  if (!c->has_ref) return fail(c, MDSCTK_ERR_STATE, "debug_scores: no reference set");
  if (data != (c->w.metric != kMetricRMSD))
    return fail(c, MDSCTK_ERR_STATE, "debug_scores: wrong entry point for the resident set");
  if (data) flags |= MDSCTK_KNN_NOFIT;
  if (c->prec == kExact) return fail(c, MDSCTK_ERR_STATE, "debug_scores: no contraction form");
  if (!approx_s) return fail(c, MDSCTK_ERR_ARG, "debug_scores: null output");
  if (xyz_nm == nullptr) n_fit = c->ref.n;
  if (n_fit <= 0 || (double)n_fit * c->ref.n > 6.0e7)
    return fail(c, MDSCTK_ERR_ARG, "debug_scores: n_fit*n_ref must be in (0, 6e7]");
  HIP_TRY(c, hipSetDevice(c->device));
  Slot& s = c->main;
  const bool nofit = flags & MDSCTK_KNN_NOFIT;
  int rc;
  if (nofit && (rc = ensure_plain(c, s.st))) return rc;
  const DevFrames& R = ref_ops(c, nofit);
  const DevFrames self = query_view(R);
  const DevFrames* fit = &self;
  if (xyz_nm) {
    if ((rc = stage_input(c, s, xyz_nm, false, n_fit, nofit))) return rc;
    fit = &s.fit;
  }
  const size_t n = (size_t)n_fit * c->ref.n;
  void* buf = nullptr;
  if ((rc = dalloc(c, &buf, n * sizeof(double), "debug scores"))) return rc;
  Lists L;
  L.nl = 4;
  L.cap = 256;
  hipError_t e = launch_contract(R, *fit, L, nullptr, 1, 1, nofit, (double*)buf,
                                 std::max(8, c->n_cu & ~7), s.st);  // any grid is valid
  if (e == hipSuccess)
    e = hipMemcpyAsync(approx_s, buf, n * sizeof(double), hipMemcpyDeviceToHost, s.st);
  if (e == hipSuccess && exact_s) {
    e = launch_exact_rows(c->ref, *fit, c->w, nullptr, (int)n_fit, nofit, (double*)buf, nullptr, 0,
                          s.st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(exact_s, buf, n * sizeof(double), hipMemcpyDeviceToHost, s.st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s.st);
  dfree(c, buf);
  if (e != hipSuccess) return hip_fail(c, e, "debug_scores");
  return 0;
}
}  // namespace
}  // namespace mdsctk

using namespace mdsctk;

extern "C" {

int mdsctk_knn_create(mdsctk_knn_ctx** out, int device, int precision) {
  if (!out) return MDSCTK_ERR_ARG;
  *out = nullptr;
  if (precision < MDSCTK_PREC_FP32 || precision > MDSCTK_PREC_EXACT_FP64) return MDSCTK_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MDSCTK_ERR_HIP;
  if (device < 0 || device >= ndev) return MDSCTK_ERR_ARG;
  auto* c = new mdsctk_knn_ctx();
  c->device = device;
  c->prec = precision;
  if (hipSetDevice(device) != hipSuccess) { delete c; return MDSCTK_ERR_HIP; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    c->n_cu = prop.multiProcessorCount;
  {
    // the allocation-free query's fixed region on this device (its address
    // is looked up once per device, with the device current)
    const hipError_t e = scratch_region(device, &c->scratch);
    if (e != hipSuccess) {
      c->scratch = nullptr;
      c->scratch_err = hipGetErrorString(e);
    }
  }
  const int rc = slot_init(c, c->main);
  if (rc) {
    mdsctk_knn_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int mdsctk_knn_destroy(mdsctk_knn_ctx* c) {
  if (!c) return MDSCTK_ERR_ARG;
  (void)hipSetDevice(c->device);
  server_stop(c);
  slot_release(c, c->main);
  for (auto& kv : c->allocs) (void)hipFree(kv.first);
  c->allocs.clear();
  tl_ctx_err.erase(c);
  delete c;
  return 0;
}

const char* mdsctk_knn_last_error(const mdsctk_knn_ctx* c) {
  if (!c) return "null context";
  const auto it = tl_ctx_err.find(c);
  return it == tl_ctx_err.end() ? "" : it->second.c_str();
}

int mdsctk_knn_set_option(mdsctk_knn_ctx* c, int option, double value) {
  if (!c) return MDSCTK_ERR_ARG;
  switch (option) {
    case MDSCTK_OPT_HBM_BUDGET:
      if (value < 0) return fail(c, MDSCTK_ERR_ARG, "budget must be >= 0");
      {
        std::lock_guard<std::mutex> lk(c->mem_mu);  // the serving thread allocates too
        c->budget = (size_t)value;
      }
      return 0;
    case MDSCTK_OPT_EPS_SCALE:
      if (!(value > 0)) return fail(c, MDSCTK_ERR_ARG, "eps scale must be > 0");
      c->eps_scale = value;
      return 0;
    case MDSCTK_OPT_REF_SPLITS:
      if (value < 0) return fail(c, MDSCTK_ERR_ARG, "phases must be >= 0");
      c->forced_phases = (int)value;
      return 0;
    case MDSCTK_OPT_SERVE_SLOTS:
      if (!(value >= 1 && value <= 4)) return fail(c, MDSCTK_ERR_ARG, "serve slots must be 1..4");
      if (c->server) return fail(c, MDSCTK_ERR_STATE, "serving already started");
      c->serve_slots = (int)value;
      return 0;
    case MDSCTK_OPT_SERVE_SHARE_CUS:
      if (c->server) return fail(c, MDSCTK_ERR_STATE, "serving already started");
      c->serve_share_cus = value != 0;
      return 0;
    case MDSCTK_OPT_SELF_SYMMETRIC:
      c->self_sym = value != 0;
      return 0;
    case MDSCTK_OPT_MIRROR_CAP:
      if (!(value >= 1 && value <= (1 << 20)))
        return fail(c, MDSCTK_ERR_ARG, "mirrored list capacity must be 1..1048576");
      release_mirror_lists(c);
      c->mirror_cap = (int)value;
      return 0;
    case MDSCTK_OPT_SERVE_GRAPHS:
      if (c->server) return fail(c, MDSCTK_ERR_STATE, "serving already started");
      c->serve_graphs = value != 0;
      return 0;
    case MDSCTK_OPT_SERVE_MAX_FRAMES:
      if (!(value >= 1 && value <= (double)INT32_MAX))
        return fail(c, MDSCTK_ERR_ARG, "serve max frames must be >= 1");
      if (c->server) return fail(c, MDSCTK_ERR_STATE, "serving already started");
      c->serve_max_frames = (int64_t)value;
      return 0;
    default:
      return fail(c, MDSCTK_ERR_ARG, "unknown option %d", option);
  }
}

int mdsctk_knn_set_reference(mdsctk_knn_ctx* c, const float* xyz_nm, int64_t n_ref,
                             int32_t natoms, const float* weights) {
  return set_reference_impl(c, xyz_nm, false, n_ref, natoms, weights);
}

int mdsctk_knn_set_reference_device(mdsctk_knn_ctx* c, const float* d_xyz_nm, int64_t n_ref,
                                    int32_t natoms, const float* weights) {
  return set_reference_impl(c, d_xyz_nm, true, n_ref, natoms, weights);
}

int mdsctk_knn_query(mdsctk_knn_ctx* c, const float* xyz_nm, int64_t n_fit, int32_t k,
                     uint32_t flags, double* out_dist, int32_t* out_idx) {
  if (c && (!out_dist || !out_idx)) return fail(c, MDSCTK_ERR_ARG, "query: null output");
  return query_impl(c, xyz_nm, false, false, n_fit, k, flags, out_dist, out_idx, false);
}

int mdsctk_knn_query_device(mdsctk_knn_ctx* c, const float* d_xyz_nm, int64_t n_fit, int32_t k,
                            uint32_t flags, double* d_out_dist, int32_t* d_out_idx) {
  if (c && (!d_out_dist || !d_out_idx)) return fail(c, MDSCTK_ERR_ARG, "query: null output");
  return query_impl(c, d_xyz_nm, false, true, n_fit, k, flags, d_out_dist, d_out_idx, true);
}


int mdsctk_knn_query_device_noalloc(mdsctk_knn_ctx* c, const float* d_xyz_nm, int64_t n_fit,
                                    int32_t k, uint32_t flags, double* d_out_dist,
                                    int32_t* d_out_idx) {
  if (!c) return MDSCTK_ERR_ARG;
  if (!d_out_dist || !d_out_idx) return fail(c, MDSCTK_ERR_ARG, "query (fixed workspace): null output");
  if (c->has_ref && c->w.metric != kMetricRMSD)
    return fail(c, MDSCTK_ERR_STATE, "query (fixed workspace): the context holds data vectors");
  if (!c->scratch)
    return fail(c, MDSCTK_ERR_HIP, "query (fixed workspace): the fixed region is unavailable on device %d: %s",
                c->device, c->scratch_err.c_str());
  Carve cv;
  cv.base = (char*)c->scratch;
  cv.size = kScratchBytes;
  return query_impl(c, d_xyz_nm, false, true, n_fit, k, flags, d_out_dist, d_out_idx, true, &cv);
}

int mdsctk_knn_debug_scores(mdsctk_knn_ctx* c, const float* xyz_nm, int64_t n_fit,
                            uint32_t flags, double* approx_s, double* exact_s) {
  return debug_scores_impl(c, xyz_nm, false, n_fit, flags, approx_s, exact_s);
}

int mdsctk_data_set_reference(mdsctk_knn_ctx* c, const double* x, int64_t n_ref, int32_t vdim,
                              int metric) {
  return set_data_reference_impl(c, x, n_ref, vdim, metric);
}

int mdsctk_data_query(mdsctk_knn_ctx* c, const double* x, int64_t n_fit, int32_t k,
                      uint32_t flags, double* out_dist, int32_t* out_idx) {
  if (c && (!out_dist || !out_idx)) return fail(c, MDSCTK_ERR_ARG, "data query: null output");
  if (c && (flags & MDSCTK_KNN_NOFIT))
    return fail(c, MDSCTK_ERR_ARG, "data query: --nofit applies to frames only");
  return query_impl(c, x, true, false, n_fit, k, flags, out_dist, out_idx, false);
}

int mdsctk_data_debug_scores(mdsctk_knn_ctx* c, const double* x, int64_t n_fit,
                             double* approx_s, double* exact_s) {
  if (c && x && n_fit > 0 && c->has_ref && c->w.metric != kMetricRMSD) {
    const int rc = check_data(c, x, n_fit, c->ref.vdim, c->w.metric, "debug_scores");
    if (rc) return rc;
  }
  return debug_scores_impl(c, x, true, n_fit, 0, approx_s, exact_s);
}

int mdsctk_knn_get_stats(const mdsctk_knn_ctx* c, mdsctk_knn_stats* out) {
  if (!c || !out) return MDSCTK_ERR_ARG;
  *out = c->stats;
  std::lock_guard<std::mutex> lk(c->mem_mu);
  out->device_bytes = (int64_t)c->held;
  return 0;
}

int mdsctk_knn_row_width(const mdsctk_knn_ctx* c, int32_t k, uint32_t flags) {
  if (!c) return MDSCTK_ERR_ARG;
  if (!c->has_ref) return MDSCTK_ERR_STATE;
  if (flags & MDSCTK_KNN_FULL_ROWS) return (int)c->ref.n;
  if (k < 0) return MDSCTK_ERR_ARG;
  if (flags & MDSCTK_KNN_KEEP_RANK0) return (int)std::min<int64_t>(k, c->ref.n);
  return (int)std::min<int64_t>(k, c->ref.n - 1);
}

int mdsctk_knn_submit(mdsctk_knn_ctx* c, const float* xyz_nm, int64_t n_fit, int32_t k,
                      uint32_t flags, double* out_dist, int32_t* out_idx, int64_t* request_id) {
  if (!c) return MDSCTK_ERR_ARG;
  return server_submit(c, xyz_nm, n_fit, k, flags, out_dist, out_idx, request_id);
}

int mdsctk_knn_poll(mdsctk_knn_ctx* c, int64_t request_id, mdsctk_knn_request_info* info) {
  if (!c) return MDSCTK_ERR_ARG;
  return server_poll(c, request_id, info, false);
}

int mdsctk_knn_wait(mdsctk_knn_ctx* c, int64_t request_id, mdsctk_knn_request_info* info) {
  if (!c) return MDSCTK_ERR_ARG;
  return server_poll(c, request_id, info, true);
}

int mdsctk_knn_serve_stats_get(mdsctk_knn_ctx* c, mdsctk_knn_serve_stats* out) {
  if (!c || !out) return MDSCTK_ERR_ARG;
  return server_stats(c, out);
}

double mdsctk_knn_clock_ms(const mdsctk_knn_ctx* c) { return c ? clock_ms(c) : 0.0; }

}  // extern "C"
