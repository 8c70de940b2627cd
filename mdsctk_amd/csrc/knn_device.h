// Internal interface between the C-ABI layer (api.cpp) and the gfx950
// kernels (knn_kernels.hip).  Not installed; include/mdsctk_knn.h is the
// public surface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mdsctk {

constexpr int kTile = 32;  // frames per MFMA tile edge (v_mfma_*_32x32xK)
// Contraction workgroup: one wave per 32x32 (ref x fit) frame tile,
// 2 reference tiles x contract_fit_tiles(prec) fit tiles.
constexpr int kRefTilesPerWG = 2;
// Frame sets are padded to a multiple of this many tiles.
constexpr int kTileAlign = 4;
// Symmetric self query: capacity of each frame's mirrored list (the scores
// it receives as the reference side of a pair), and the stride of the
// landmark pass (every kLandmarkStride-th reference tile pair).
constexpr int kMirrorCap = 8192;
constexpr int kLandmarkStride = 128;

enum Prec : int { kFP32 = 0, kBF16 = 1, kFP16 = 2, kExact = 3 };

// What a frame is and which distance the engine ranks by.
//   kMetricRMSD         frames of atoms: optimal-superposition RMSD (knn_rms)
//   kMetricEuclidean    data vectors: Euclidean distance (knn_data)
//   kMetricCorrelation  data vectors: correlation distance sqrt((1-r)/2)
//                       (knn_data -c)
// Data vectors run through the same contraction as frames with the QCP
// step replaced by the --nofit trace: a vector of length V is laid out as
// ceil(V/3) "atoms" of 3 values with unit masses, uncentred, so that the
// trace of the 3x3 block is the dot product.
enum Metric : int { kMetricRMSD = 0, kMetricEuclidean = 1, kMetricCorrelation = 2 };

// Atoms per operand K-group: split forms hold 16 atoms (one 32x32x16 MFMA
// step) per group, FP32 holds 8 atoms (four 32x32x2 MFMA steps).
inline int group_atoms(int prec) { return prec == kFP32 ? 8 : 16; }
// The split forms' K-loop walks K-groups four at a time (its four stage
// buffers), so their operand holds a multiple of four groups (the padding
// atoms are zero and contribute nothing).
constexpr int kSplitGroupAlign = 4;
inline int operand_groups(int prec, int natoms) {
  const int ga = group_atoms(prec);
  const int g = (natoms + ga - 1) / ga;
  return prec == kFP32 ? g : (g + kSplitGroupAlign - 1) / kSplitGroupAlign * kSplitGroupAlign;
}
// 16-byte operand blocks of 64 lanes per (tile, K-group): 3 planes x parts.
inline int blocks_per_group(int prec) { return prec == kFP32 ? 3 : 6; }
// K-groups of zero slack after a frame set's operand: the contraction's
// copy pipeline reads up to four groups past the end of the last tile.
constexpr int kOperandTailGroups = 4;

// A frame set resident in HBM.  Row counts are padded to a multiple of
// kTileAlign*kTile (the contraction walks reference tiles in pairs and fit
// tiles in fours); padded rows are zero.
struct DevFrames {
  int64_t n = 0;        // real frames
  int natoms = 0;
  int n_tiles = 0;      // padded tile count (multiple of kTileAlign)
  int cap_tiles = 0;    // tiles the derived arrays are allocated for (>= n_tiles)
  int n_groups = 0;     // operand K-groups
  int prec = kFP32;
  float* xyz = nullptr;   // [n][natoms][3] as given (nm)
  double* cen = nullptr;  // [n_tiles*32][3] FP64 mass-weighted centroid
  double* G = nullptr;    // [n_tiles*32] sum w |x-c|^2, FP64
  double* Gt = nullptr;   // [n_tiles*32] same sum over the rounded operand
  uint4* op = nullptr;    // contraction operand, fragment-major (see operand_kernel)
  bool owns_xyz = true;
  // data vectors only: the FP64 vectors [n][vdim] the exact scores use and,
  // for the correlation metric, [n][2] (mean, sample standard deviation)
  int vdim = 0;
  double* x64 = nullptr;
  double* aux = nullptr;
  // resident set only: one device double, max Gt over the set (the window's
  // reference term), computed once when the set is made resident
  double* gmax = nullptr;
  // resident set of a dithered split form only: its operand and G~ in the
  // query-side scaling, which a self query uses as its query operand
  uint4* op_q = nullptr;
  double* Gt_q = nullptr;
  // frames of fitted RMSD scores: each frame's principal axes, [rows][9],
  // written by the prep pass when it builds operands in those axes
  double* rot = nullptr;
};

// The resident set seen as a query set (self query): the query-side operand
// where the set has one.
inline DevFrames query_view(const DevFrames& ref) {
  DevFrames v = ref;
  if (ref.op_q) {
    v.op = ref.op_q;
    v.Gt = ref.Gt_q;
  }
  return v;
}

struct DevWeights {
  int natoms = 0;
  float* w = nullptr;   // masses
  float* sw = nullptr;  // sqrt(mass)
  // Output distance = out_scale * sqrt(s / W): for frames W is the sum of
  // masses and out_scale 10 (nm to Angstrom); for data vectors W is 1
  // (Euclidean) or 4 (correlation, s = 2(1-r)) and out_scale 1.
  double W = 0;
  double out_scale = 10.0;
  int metric = kMetricRMSD;
};

// Per-query candidate lists written by the contraction kernel.
struct Lists {
  int nl = 0;          // lists per query
  int cap = 0;         // entries per list
  uint2* ent = nullptr;       // [n_fit][nl][cap] (float key bits, ref index)
  int* cnt = nullptr;         // [n_fit][nl]
  float* ovf = nullptr;       // [n_fit][nl] smallest key dropped on overflow
  unsigned* thr = nullptr;    // [n_fit] shared threshold (float bits)
};

// The symmetric self query's second list per frame: the scores a frame
// receives as the reference side of a pair contracted only the other way
// round ([n][cap] entries, counters, smallest key dropped past cap).
struct Mirror {
  uint2* ent = nullptr;
  int* cnt = nullptr;
  float* ovf = nullptr;
  int cap = 0;
};

// Launchers (all asynchronous on `st`).
// The prep pass: the FP64 centroids, G and G~ (one wave per frame), then,
// if f.op is set, the operand (one workgroup per tile and 4 K-groups).
// side 0: plain operand; +1 / -1: the dithered split's reference / query
// side (split forms only).  second (may be null): a view of the same frames
// whose op / Gt receive the operand and G~ of side2 as well (the resident
// set's query-side operand).  src_xyz (if given): the frames are read from
// it instead of f.xyz, row r from frame perm[r] when perm is given.  rot (if
// given, frames only): every frame is expressed in its principal axes,
// whose rotation is stored at rot[9 r] (frame_stats_kernel).
hipError_t launch_prep(const DevFrames& f, const DevWeights& w, int side,
                       const DevFrames* second, int side2, hipStream_t st,
                       const float* src_xyz = nullptr, const int* perm = nullptr,
                       double* rot = nullptr);
// reset_thr false keeps the shared thresholds (a landmark pass set them).
hipError_t launch_init_lists(const Lists& L, int64_t n_fit, hipStream_t st,
                             bool reset_thr = true);
// eps2[q] = 2 kappa (Gt_fit[q] + *ref.gmax); *n_cand (if given) is set to 0.
hipError_t launch_eps(const DevFrames& fit, const DevFrames& ref, double kappa, float* eps2,
                      int* n_cand, hipStream_t st);
// Approximate contraction into candidate lists.  When dbg_scores != nullptr
// it instead writes the approximate score of every pair to
// dbg_scores[fit*n_ref+ref] (small sizes).  With `mirror` (a self query: fit
// is the reference set's query view) it contracts the upper block triangle
// only: each pair joins the fit frame's lists and, contracted one way round
// only, the reference frame's mirrored list when it lies below that
// frame's shared threshold L.thr.  ph_run < ref_phases walks phases
// 0..ph_run-1 only (every ref_phases-th reference tile pair: the landmark
// pass); 0 = all.
hipError_t launch_contract(const DevFrames& ref, const DevFrames& fit, const Lists& L,
                           const float* eps2, int k1, int ref_phases, bool nofit,
                           double* dbg_scores, int grid_wgs, hipStream_t st,
                           const Mirror* mirror = nullptr, int ph_run = 0);
// The window of a symmetric self query (fit: the reference set's query
// view): eps2[r] = 2 kappa max(G~q[r] + max G~ref, G~ref[r] + max G~q);
// *gmax_q receives max G~q.
hipError_t launch_eps_sym(const DevFrames& fit, const DevFrames& ref, double kappa,
                          double* gmax_q, float* eps2, int* n_cand, hipStream_t st);
// After the landmark pass (lists L of n rows, reference phases `phases`):
// thr[q] = the certified bound of row q's k1-th smallest key in the full
// query, label[q] = its nearest landmark (dense, < 64 ceil(pairs / phases)).
hipError_t launch_landmark_cap(const Lists& L, int64_t n, int k1, const float* eps2,
                               int phases, unsigned* thr, int* label, hipStream_t st);
// order[0..n) = rows grouped by label (counting sort; counts holds nlab ints).
hipError_t launch_row_order(const int* label, int64_t n, int nlab, int* counts, int* order,
                            hipStream_t st);
hipError_t launch_mirror_init(const Mirror& m, int64_t n, hipStream_t st);
// out[0] = total mirrored appends (including those dropped past cap), out[1] = max per frame.
hipError_t launch_mirror_sum(const Mirror& m, int64_t n, unsigned long long* out,
                             hipStream_t st);
// dst[i] = frame rows[i] of src (natoms x 3 floats each).
hipError_t launch_gather_frames(const float* src, const int* rows, int nrows, int natoms,
                                float* dst, hipStream_t st);
// Output row i of (sd, si) to row rows[i] of (dd, di), k_out entries each.
hipError_t launch_scatter_rows(const double* sd, const int* si, const int* rows, int nrows,
                               int k_out, double* dd, int* di, hipStream_t st);
// order: the query rows in the order the merge workgroups take them
// (nullptr: 0..n_fit-1).  mirror: a symmetric self query's second list per
// row (nullptr: none).  perm (replaces order): the lists, mirrored lists and
// eps2 were built over the frames in the order perm (list row i is frame
// perm[i], entry index j frame perm[j]); ref/fit and the outputs and
// fallback flags use the original indices.
hipError_t launch_merge(const DevFrames& ref, const DevFrames& fit, const DevWeights& w,
                        const Lists& L, const float* eps2, int k1, int k_out, int rank_off,
                        bool nofit, double* out_dist, int* out_idx, int* fallback,
                        int* n_cand, const int* order, hipStream_t st,
                        const Mirror* mirror = nullptr, const int* perm = nullptr);
// dst[i] = the float bits src[perm[i]] raised by add[i] (and one unit in the
// last place): the shared thresholds of a landmark pass moved to the
// permuted order, widened by the permuted frames' window.
hipError_t launch_permute_thresholds(const unsigned* src, const int* perm, const float* add,
                                     int64_t n, unsigned* dst, hipStream_t st);
// Exact FP64 scores s = G_a + G_b - 2 lambda for fit rows [row0,row0+nrows)
// against all refs.  Writes s into srow[(row-row0)*n_ref + ref] unless
// out_full is given, in which case it writes Angstrom distances into
// out_full[row*n_ref + ref] (the --sort false layout).
hipError_t launch_exact_rows(const DevFrames& ref, const DevFrames& fit, const DevWeights& w,
                             const int* rows, int nrows, bool nofit, double* srow,
                             double* out_full, int64_t out_ld, hipStream_t st);
// Top-k1 by (s, index) over exact score rows; writes ranks
// [rank_off, rank_off+k_out) of each listed fit row.
hipError_t launch_select_rows(const double* srow, int64_t n_ref, const int* rows, int nrows,
                              const DevWeights& w, int k1, int k_out, int rank_off,
                              double* out_dist, int* out_idx, hipStream_t st);
// Any k1 (the LDS select takes k1 <= 2048): the rows' k1 smallest (s, index)
// compacted in index order, then a stable segmented radix sort; `scratch`
// holds sort_rows_scratch(nrows, k1) bytes.
size_t sort_rows_scratch(int nrows, int k1);
hipError_t launch_select_rows_sorted(const double* srow, int64_t n_ref, const int* rows,
                                     int nrows, const DevWeights& w, int k1, int k_out,
                                     int rank_off, double* out_dist, int* out_idx, void* scratch,
                                     size_t scratch_bytes, hipStream_t st);
constexpr int kSelectLdsMaxK1 = 2048;
hipError_t launch_identity_rows(int64_t n_fit, int64_t n_ref, int* out_idx, hipStream_t st);
// Data vectors: mean[e] over the n vectors x[n][vdim] (FP64).
hipError_t launch_column_mean(const double* x, int64_t n, int vdim, double* mean, hipStream_t st);
// Data vectors to the frame layout of f (f.xyz, f.natoms * 3 >= f.vdim):
// Euclidean x - shift, correlation (x - mean) / |x - mean| with f.aux.
hipError_t launch_data_prep(const DevFrames& f, int metric, const double* shift, hipStream_t st);
hipError_t launch_max_gt(const DevFrames& f, double* out, hipStream_t st);

// Size of the per-GPU fixed workspace of the allocation-free query entry
// point, and its device address on `device` (which must be the current
// device on the first call for that device; later calls return the cached
// address).  Calls that overlap on one GPU share it.
constexpr size_t kScratchBytes = size_t(128) << 20;
hipError_t scratch_region(int device, void** addr);

// Fit tiles per contraction workgroup for a precision (2 or 4).
int contract_fit_tiles(int prec);
// Choose the per-list capacity for k1 (0 when the fused path cannot serve it).
int list_capacity_for(int k1);

}  // namespace mdsctk
