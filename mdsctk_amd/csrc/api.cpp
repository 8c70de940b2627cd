// C ABI of the knn_rms engine (include/mdsctk_knn.h): context, device
// memory accounting, and the per-query orchestration of the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mdsctk_knn.h"
#include "knn_device.h"

using namespace mdsctk;

namespace {

// Candidate window: |s~ - s| <= kappa (G_q + max G_ref) for the approximate
// score s~ of each contraction form (s = G_q + G_r - 2 lambda).  kappa is an
// envelope a*sqrt(N) + b fitted to the largest relative deviation measured
// on the synthetic generators (random, metastable with near-duplicates, and
// pre-aligned frames) for N = 20..2000 atoms, times a safety factor of 2.
// FP32 accumulation contributes the sqrt(N) term (one rounding per MFMA
// step); the split forms add a constant term from the dropped lo*lo'
// product.  It is a measured statistical envelope, not a worst-case proof.
double kappa_for(int prec, int natoms) {
  const double rn = std::sqrt((double)std::max(natoms, 16));
  switch (prec) {
    case kFP32: return 2.0 * (5.5e-8 * rn);
    case kBF16: return 2.0 * (3.0e-8 * rn + 4.0e-6);
    case kFP16: return 2.0 * (3.0e-8 * rn + 1.0e-7);
    default: return 0.0;
  }
}

}  // namespace

struct mdsctk_knn_ctx {
  int device = 0;
  int prec = kFP32;
  int n_cu = 256;
  hipStream_t st = nullptr;
  std::string err;
  size_t budget = 0;  // 0 = unlimited
  size_t held = 0;
  std::unordered_map<void*, size_t> allocs;
  double eps_scale = 1.0;
  int forced_phases = 0;

  DevWeights w;
  DevFrames ref;
  bool has_ref = false;
  DevFrames fit;  // query frames (grown on demand)
  size_t fit_cap_frames = 0;
  int fit_natoms = 0;

  // query workspace (grown on demand)
  void* ws_lists = nullptr;
  size_t ws_lists_bytes = 0;
  void* ws_small = nullptr;  // per-query arrays
  size_t ws_small_bytes = 0;
  double* gmax = nullptr;
  int* n_cand = nullptr;
  void* ws_out = nullptr;  // device outputs for the host API
  size_t ws_out_bytes = 0;
  void* ws_rows = nullptr;  // exact-row scratch
  size_t ws_rows_bytes = 0;

  hipEvent_t ev[6] = {};
  mdsctk_knn_stats stats = {};
};

namespace {

int fail(mdsctk_knn_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
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
  if (c->budget && c->held + bytes > c->budget)
    return fail(c, MDSCTK_ERR_NOMEM,
                "%s: %zu bytes would exceed the HBM budget (%zu held of %zu)", what, bytes,
                c->held, c->budget);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, MDSCTK_ERR_NOMEM, "%s: hipMalloc(%zu) failed: %s", what, bytes,
                hipGetErrorString(e));
  }
  c->held += bytes;
  c->allocs[*p] = bytes;
  return 0;
}

void dfree(mdsctk_knn_ctx* c, void* p) {
  if (!p) return;
  auto it = c->allocs.find(p);
  if (it != c->allocs.end()) {
    c->held -= it->second;
    c->allocs.erase(it);
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

void free_frames(mdsctk_knn_ctx* c, DevFrames& f) {
  if (f.owns_xyz) dfree(c, f.xyz);
  dfree(c, f.cen);
  dfree(c, f.G);
  dfree(c, f.Gt);
  dfree(c, f.op);
  f = DevFrames();
}

size_t operand_bytes(int prec, int n_tiles, int n_groups) {
  return ((size_t)n_tiles * n_groups + kOperandTailGroups) * blocks_per_group(prec) * 64 * 16;
}

// Allocate the derived arrays of a frame set (xyz handled by the caller).
int alloc_frames(mdsctk_knn_ctx* c, DevFrames& f, int64_t n, int natoms, int prec) {
  f.n = n;
  f.natoms = natoms;
  f.prec = prec == kExact ? kFP32 : prec;
  const int64_t tiles = (n + kTile - 1) / kTile;
  f.n_tiles = (int)((tiles + kTileAlign - 1) / kTileAlign * kTileAlign);
  const int ga = group_atoms(f.prec);
  f.n_groups = (natoms + ga - 1) / ga;
  const size_t rows = (size_t)f.n_tiles * kTile;
  int rc;
  if ((rc = dalloc(c, (void**)&f.cen, rows * 3 * sizeof(double), "centroids"))) return rc;
  if ((rc = dalloc(c, (void**)&f.G, rows * sizeof(double), "G terms"))) return rc;
  if ((rc = dalloc(c, (void**)&f.Gt, rows * sizeof(double), "G~ terms"))) return rc;
  if (prec != kExact &&
      (rc = dalloc(c, (void**)&f.op, operand_bytes(f.prec, f.n_tiles, f.n_groups), "operand")))
    return rc;
  return 0;
}

int prep_frames(mdsctk_knn_ctx* c, DevFrames& f) {
  if (c->prec == kExact) {
    // The exact form needs only centroids and G; run the FP32 prep without
    // an operand buffer is not possible, so compute them through a
    // temporary operand.
    void* tmp = nullptr;
    int rc = dalloc(c, &tmp, operand_bytes(kFP32, f.n_tiles, f.n_groups), "prep scratch");
    if (rc) return rc;
    f.op = (uint4*)tmp;
    hipError_t e = launch_prep(f, c->w, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    f.op = nullptr;
    dfree(c, tmp);
    if (e != hipSuccess) return hip_fail(c, e, "prep");
    return 0;
  }
  HIP_TRY(c, launch_prep(f, c->w, c->st));
  return 0;
}

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
  HIP_TRY(c, hipMemcpyAsync(c->w.w, w.data(), natoms * sizeof(float), hipMemcpyHostToDevice, c->st));
  HIP_TRY(c, hipMemcpyAsync(c->w.sw, sw.data(), natoms * sizeof(float), hipMemcpyHostToDevice, c->st));
  HIP_TRY(c, hipStreamSynchronize(c->st));
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
  HIP_TRY(c, hipSetDevice(c->device));
  free_frames(c, c->ref);
  c->has_ref = false;
  int rc = set_weights(c, weights, natoms);
  if (rc) return rc;
  const size_t xbytes = (size_t)n_ref * natoms * 3 * sizeof(float);
  if ((rc = dalloc(c, (void**)&c->ref.xyz, xbytes, "reference coordinates"))) return rc;
  HIP_TRY(c, hipMemcpyAsync(c->ref.xyz, xyz, xbytes,
                            on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->st));
  c->ref.owns_xyz = true;
  if ((rc = alloc_frames(c, c->ref, n_ref, natoms, c->prec))) return rc;
  if ((rc = prep_frames(c, c->ref))) return rc;
  HIP_TRY(c, hipStreamSynchronize(c->st));
  c->has_ref = true;
  return 0;
}

// Upload (or alias) the query frames and run their prep pass.
int stage_queries(mdsctk_knn_ctx* c, const float* xyz, bool on_device, int64_t n_fit,
                  DevFrames** out) {
  if (!xyz) {
    *out = &c->ref;
    return 0;
  }
  DevFrames& f = c->fit;
  const int natoms = c->ref.natoms;
  if (f.n_tiles == 0 || (size_t)n_fit > c->fit_cap_frames || c->fit_natoms != natoms ||
      f.prec != (c->prec == kExact ? kFP32 : c->prec)) {
    free_frames(c, f);
    int rc;
    if ((rc = dalloc(c, (void**)&f.xyz, (size_t)n_fit * natoms * 3 * sizeof(float),
                     "query coordinates")))
      return rc;
    f.owns_xyz = true;
    if ((rc = alloc_frames(c, f, n_fit, natoms, c->prec))) return rc;
    c->fit_cap_frames = (size_t)n_fit;
    c->fit_natoms = natoms;
  }
  // Reuse the allocation: a smaller batch only changes the row count.
  f.n = n_fit;
  const size_t xbytes = (size_t)n_fit * natoms * 3 * sizeof(float);
  HIP_TRY(c, hipMemcpyAsync(f.xyz, xyz, xbytes,
                            on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->st));
  int rc = prep_frames(c, f);
  if (rc) return rc;
  *out = &f;
  return 0;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

// Exact FP64 path for the listed fit rows: scores of whole rows in batches,
// then (s, index) selection.  rows == nullptr means rows [0, nrows).
int exact_rows_select(mdsctk_knn_ctx* c, DevFrames& fit, const int* h_rows, int nrows, int k1,
                      int k_out, int rank_off, bool nofit, double* d_dist, int* d_idx) {
  const int64_t n_ref = c->ref.n;
  const size_t row_bytes = (size_t)n_ref * sizeof(double);
  int batch = (int)std::max<int64_t>(1, std::min<int64_t>(nrows, (int64_t)(1ull << 30) / row_bytes));
  int rc = ensure(c, &c->ws_rows, &c->ws_rows_bytes, row_bytes * batch + sizeof(int) * batch,
                  "exact-row scratch");
  if (rc) return rc;
  double* srow = (double*)c->ws_rows;
  int* drows = (int*)((char*)c->ws_rows + row_bytes * batch);
  std::vector<int> ids(batch);
  for (int b0 = 0; b0 < nrows; b0 += batch) {
    const int nb = std::min(batch, nrows - b0);
    for (int i = 0; i < nb; ++i) ids[i] = h_rows ? h_rows[b0 + i] : b0 + i;
    HIP_TRY(c, hipMemcpyAsync(drows, ids.data(), nb * sizeof(int), hipMemcpyHostToDevice, c->st));
    HIP_TRY(c, launch_exact_rows(c->ref, fit, c->w, drows, nb, nofit, srow, nullptr, 0, c->st));
    HIP_TRY(c, launch_select_rows(srow, n_ref, drows, nb, c->w.W, k1, k_out, rank_off, d_dist,
                                  d_idx, c->st));
    HIP_TRY(c, hipStreamSynchronize(c->st));
  }
  return 0;
}

int query_impl(mdsctk_knn_ctx* c, const float* xyz, bool on_device, int64_t n_fit, int32_t k,
               uint32_t flags, double* out_dist, int32_t* out_idx, bool out_on_device) {
  if (!c) return MDSCTK_ERR_ARG;
  if (!c->has_ref) return fail(c, MDSCTK_ERR_STATE, "query: no reference set");
  if (xyz == nullptr) n_fit = c->ref.n;
  if (n_fit < 0) return fail(c, MDSCTK_ERR_ARG, "query: negative frame count");
  if (n_fit > INT32_MAX) return fail(c, MDSCTK_ERR_ARG, "query: too many frames");
  const bool nofit = flags & MDSCTK_KNN_NOFIT;
  const bool full = flags & MDSCTK_KNN_FULL_ROWS;
  const int rank_off = (flags & MDSCTK_KNN_KEEP_RANK0) ? 0 : 1;
  const int64_t n_ref = c->ref.n;
  if (!full && k < 0) return fail(c, MDSCTK_ERR_ARG, "query: k must be >= 0");
  // knn_rms.cpp:224-225 clamps k to n_ref-1; k+1 follows the clamp here.
  int k_out = full ? (int)n_ref : (int)std::min<int64_t>(k, n_ref - 1);
  if (!full && rank_off == 0) k_out = (int)std::min<int64_t>(k, n_ref);
  const int k1 = full ? 0 : k_out + rank_off;
  if (n_fit == 0 || k_out == 0) {
    c->stats = mdsctk_knn_stats();
    return k_out;
  }
  if (!full && k1 > 2048)
    return fail(c, MDSCTK_ERR_ARG, "query: k=%d exceeds the supported maximum 2047", k);
  HIP_TRY(c, hipSetDevice(c->device));
  mdsctk_knn_stats S = {};
  S.n_pairs = n_fit * n_ref;
  HIP_TRY(c, hipEventRecord(c->ev[0], c->st));

  DevFrames* fit = nullptr;
  int rc = stage_queries(c, xyz, on_device, n_fit, &fit);
  if (rc) return rc;
  HIP_TRY(c, hipEventRecord(c->ev[1], c->st));

  const size_t out_elems = (size_t)n_fit * k_out;
  double* d_dist = out_dist;
  int* d_idx = out_idx;
  if (!out_on_device) {
    rc = ensure(c, &c->ws_out, &c->ws_out_bytes, out_elems * (sizeof(double) + sizeof(int)),
                "output buffers");
    if (rc) return rc;
    d_dist = (double*)c->ws_out;
    d_idx = (int*)((char*)c->ws_out + out_elems * sizeof(double));
  }

  float ms_contract = 0, ms_merge = 0, ms_fallback = 0;
  if (full) {
    HIP_TRY(c, launch_exact_rows(c->ref, *fit, c->w, nullptr, (int)n_fit, nofit, nullptr, d_dist,
                                 n_ref, c->st));
    HIP_TRY(c, launch_identity_rows(n_fit, n_ref, d_idx, c->st));
    HIP_TRY(c, hipEventRecord(c->ev[2], c->st));
    HIP_TRY(c, hipEventRecord(c->ev[3], c->st));
    HIP_TRY(c, hipEventSynchronize(c->ev[3]));
    ms_contract = elapsed(c->ev[1], c->ev[2]);
  } else if (c->prec == kExact || list_capacity_for(k1) == 0) {
    HIP_TRY(c, hipEventRecord(c->ev[2], c->st));
    rc = exact_rows_select(c, *fit, nullptr, (int)n_fit, k1, k_out, rank_off, nofit, d_dist, d_idx);
    if (rc) return rc;
    HIP_TRY(c, hipEventRecord(c->ev[3], c->st));
    HIP_TRY(c, hipEventSynchronize(c->ev[3]));
    ms_fallback = elapsed(c->ev[2], c->ev[3]);
    S.n_fallback_rows = n_fit;
  } else {
    // Fused path.
    const int cap = list_capacity_for(k1);
    const int ft = contract_fit_tiles(fit->prec);
    const int fit_blocks = fit->n_tiles / ft;
    const int ref_pairs = c->ref.n_tiles / kRefTilesPerWG;
    // workgroups resident at once: two 4-wave or one 8-wave per CU
    const int grid = std::max(8, (c->n_cu * (ft == 2 ? 2 : 1)) & ~7);
    int phases = c->forced_phases;
    if (phases <= 0) {
      phases = std::max(8, (grid + fit_blocks - 1) / fit_blocks);
      int p2 = 1;
      while (p2 < phases) p2 <<= 1;
      phases = p2;
    }
    phases = std::max(1, std::min(phases, ref_pairs));
    Lists L;
    L.nl = phases * 4;
    L.cap = cap;
    const size_t n_hdr = (size_t)n_fit * L.nl;
    const size_t ent_bytes = n_hdr * cap * sizeof(uint2);
    rc = ensure(c, &c->ws_lists, &c->ws_lists_bytes, ent_bytes, "candidate lists");
    if (rc) return rc;
    L.ent = (uint2*)c->ws_lists;
    const size_t small = n_hdr * (sizeof(int) + sizeof(float)) +
                         (size_t)n_fit * (sizeof(unsigned) + sizeof(float) + sizeof(int)) + 1024;
    rc = ensure(c, &c->ws_small, &c->ws_small_bytes, small, "list headers");
    if (rc) return rc;
    char* sp = (char*)c->ws_small;
    L.cnt = (int*)sp; sp += n_hdr * sizeof(int);
    L.ovf = (float*)sp; sp += n_hdr * sizeof(float);
    L.thr = (unsigned*)sp; sp += n_fit * sizeof(unsigned);
    float* eps2 = (float*)sp; sp += n_fit * sizeof(float);
    int* d_fb = (int*)sp;
    const double kappa = kappa_for(c->prec, c->ref.natoms) * c->eps_scale;
    HIP_TRY(c, launch_init_lists(L, n_fit, c->st));
    HIP_TRY(c, launch_eps(*fit, c->ref, kappa, eps2, c->gmax, c->st));
    HIP_TRY(c, hipMemsetAsync(c->n_cand, 0, sizeof(int), c->st));
    HIP_TRY(c, hipEventRecord(c->ev[2], c->st));
    HIP_TRY(c, launch_contract(c->ref, *fit, L, eps2, k1, phases, nofit, nullptr, grid, c->st));
    HIP_TRY(c, hipEventRecord(c->ev[3], c->st));
    HIP_TRY(c, launch_merge(c->ref, *fit, c->w, L, eps2, k1, k_out, rank_off, nofit, d_dist,
                            d_idx, d_fb, c->n_cand, c->st));
    HIP_TRY(c, hipEventRecord(c->ev[4], c->st));
    std::vector<int> fb(n_fit);
    int ncand = 0;
    HIP_TRY(c, hipMemcpyAsync(fb.data(), d_fb, n_fit * sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIP_TRY(c, hipMemcpyAsync(&ncand, c->n_cand, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIP_TRY(c, hipStreamSynchronize(c->st));
    ms_contract = elapsed(c->ev[2], c->ev[3]);
    ms_merge = elapsed(c->ev[3], c->ev[4]);
    S.n_candidates = ncand;
    S.ref_splits = phases;
    S.list_capacity = cap;
    S.eps_rel = kappa;
    std::vector<int> rows;
    for (int64_t i = 0; i < n_fit; ++i)
      if (fb[i]) rows.push_back((int)i);
    if (!rows.empty()) {
      HIP_TRY(c, hipEventRecord(c->ev[2], c->st));
      rc = exact_rows_select(c, *fit, rows.data(), (int)rows.size(), k1, k_out, rank_off, nofit,
                             d_dist, d_idx);
      if (rc) return rc;
      HIP_TRY(c, hipEventRecord(c->ev[3], c->st));
      HIP_TRY(c, hipEventSynchronize(c->ev[3]));
      ms_fallback = elapsed(c->ev[2], c->ev[3]);
      S.n_fallback_rows = (int64_t)rows.size();
    }
  }
  if (!out_on_device) {
    HIP_TRY(c, hipMemcpyAsync(out_dist, d_dist, out_elems * sizeof(double), hipMemcpyDeviceToHost,
                              c->st));
    HIP_TRY(c, hipMemcpyAsync(out_idx, d_idx, out_elems * sizeof(int), hipMemcpyDeviceToHost,
                              c->st));
  }
  HIP_TRY(c, hipEventRecord(c->ev[5], c->st));
  HIP_TRY(c, hipEventSynchronize(c->ev[5]));
  S.ms_prep = elapsed(c->ev[0], c->ev[1]);
  S.ms_contract = ms_contract;
  S.ms_merge = ms_merge;
  S.ms_fallback = ms_fallback;
  S.ms_total = elapsed(c->ev[0], c->ev[5]);
  c->stats = S;
  return k_out;
}

}  // namespace

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
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MDSCTK_ERR_HIP;
  }
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) { mdsctk_knn_destroy(c); return MDSCTK_ERR_HIP; }
  if (dalloc(c, (void**)&c->gmax, 256, "scratch") || dalloc(c, (void**)&c->n_cand, 256, "scratch")) {
    mdsctk_knn_destroy(c);
    return MDSCTK_ERR_NOMEM;
  }
  *out = c;
  return 0;
}

int mdsctk_knn_destroy(mdsctk_knn_ctx* c) {
  if (!c) return MDSCTK_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (auto& kv : c->allocs) (void)hipFree(kv.first);
  c->allocs.clear();
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return 0;
}

const char* mdsctk_knn_last_error(const mdsctk_knn_ctx* c) {
  if (!c) return "null context";
  return c->err.c_str();
}

int mdsctk_knn_set_option(mdsctk_knn_ctx* c, int option, double value) {
  if (!c) return MDSCTK_ERR_ARG;
  switch (option) {
    case MDSCTK_OPT_HBM_BUDGET:
      if (value < 0) return fail(c, MDSCTK_ERR_ARG, "budget must be >= 0");
      c->budget = (size_t)value;
      return 0;
    case MDSCTK_OPT_EPS_SCALE:
      if (!(value > 0)) return fail(c, MDSCTK_ERR_ARG, "eps scale must be > 0");
      c->eps_scale = value;
      return 0;
    case MDSCTK_OPT_REF_SPLITS:
      if (value < 0) return fail(c, MDSCTK_ERR_ARG, "phases must be >= 0");
      c->forced_phases = (int)value;
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
  return query_impl(c, xyz_nm, false, n_fit, k, flags, out_dist, out_idx, false);
}

int mdsctk_knn_query_device(mdsctk_knn_ctx* c, const float* d_xyz_nm, int64_t n_fit, int32_t k,
                            uint32_t flags, double* d_out_dist, int32_t* d_out_idx) {
  if (c && (!d_out_dist || !d_out_idx)) return fail(c, MDSCTK_ERR_ARG, "query: null output");
  return query_impl(c, d_xyz_nm, true, n_fit, k, flags, d_out_dist, d_out_idx, true);
}

int mdsctk_knn_debug_scores(mdsctk_knn_ctx* c, const float* xyz_nm, int64_t n_fit,
                            uint32_t flags, double* approx_s, double* exact_s) {
  if (!c) return MDSCTK_ERR_ARG;
  if (!c->has_ref) return fail(c, MDSCTK_ERR_STATE, "debug_scores: no reference set");
  if (c->prec == kExact) return fail(c, MDSCTK_ERR_STATE, "debug_scores: no contraction form");
  if (xyz_nm == nullptr) n_fit = c->ref.n;
  if (n_fit <= 0 || (double)n_fit * c->ref.n > 6.0e7)
    return fail(c, MDSCTK_ERR_ARG, "debug_scores: n_fit*n_ref must be in (0, 6e7]");
  HIP_TRY(c, hipSetDevice(c->device));
  const bool nofit = flags & MDSCTK_KNN_NOFIT;
  DevFrames* fit = nullptr;
  int rc = stage_queries(c, xyz_nm, false, n_fit, &fit);
  if (rc) return rc;
  const size_t n = (size_t)n_fit * c->ref.n;
  void* buf = nullptr;
  if ((rc = dalloc(c, &buf, n * sizeof(double), "debug scores"))) return rc;
  Lists L;
  L.nl = 4;
  L.cap = 256;
  hipError_t e = launch_contract(c->ref, *fit, L, nullptr, 1, 1, nofit, (double*)buf,
                                 std::max(8, c->n_cu & ~7), c->st);  // any grid is valid
  if (e == hipSuccess)
    e = hipMemcpyAsync(approx_s, buf, n * sizeof(double), hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess && exact_s) {
    std::vector<int> rows(n_fit);
    for (int64_t i = 0; i < n_fit; ++i) rows[i] = (int)i;
    e = launch_exact_rows(c->ref, *fit, c->w, nullptr, (int)n_fit, nofit, (double*)buf, nullptr, 0,
                          c->st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(exact_s, buf, n * sizeof(double), hipMemcpyDeviceToHost, c->st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  dfree(c, buf);
  if (e != hipSuccess) return hip_fail(c, e, "debug_scores");
  return 0;
}

int mdsctk_knn_get_stats(const mdsctk_knn_ctx* c, mdsctk_knn_stats* out) {
  if (!c || !out) return MDSCTK_ERR_ARG;
  *out = c->stats;
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

}  // extern "C"
