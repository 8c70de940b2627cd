/*
 * mdsctk_knn.h - C ABI of the MI355X knn_rms engine.
 *
 * The reference (MDSCTK 1.2.5) has no library interface for this path: the
 * knn_rms executable (knn_rms.cpp) is the contract.  Inside it, two pieces
 * are replaced by this library:
 *   - double distance(int natoms, real w[], coord_array ref,
 *                     coord_array fit, bool dofit)       knn_rms.cpp:38-41
 *     called for every (fit, ref) pair from the OpenMP loops at
 *     knn_rms.cpp:231-243 and 256-279, and
 *   - permutation<double>::sort(k1)                       mdsctk.h:177-199
 *     with the "drop rank 0, write ranks 1..k" row writes at
 *     knn_rms.cpp:246-255 and 281-291.
 * The frame loading and centring (knn_rms.cpp:149-206) maps to
 * mdsctk_knn_set_reference / the query upload; file formats are in
 * mdsctk_io.h.
 *
 * Conventions
 *   - Coordinates are float32 nm, frame-major [n][natoms][3], uncentred; the
 *     library centres them with the mass weights (knn_rms.cpp:189,199).
 *   - Distances are returned in Angstrom (knn_rms.cpp:40 multiplies by 10),
 *     indices are 0-based reference frame numbers, rows in query order.
 *   - Return value: >= 0 on success, a negative MDSCTK_ERR_* code on
 *     failure; mdsctk_knn_last_error() describes the failure.  The library
 *     never calls exit().
 *   - A context is used from one host thread at a time, except for the
 *     serving calls (submit / poll / wait / serve_stats), which any number
 *     of host threads may make concurrently, also while a synchronous query
 *     runs on another thread.  mdsctk_knn_last_error() reports the calling
 *     thread's last failure on the context, so concurrent calls never share
 *     an error text.  A context owns its HIP streams and all device memory
 *     it allocates.  Several contexts may exist at once (e.g. one per
 *     resident reference set).
 */
#ifndef MDSCTK_KNN_H
#define MDSCTK_KNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. */
#define MDSCTK_OK 0
#define MDSCTK_ERR_ARG (-1)      /* invalid argument */
#define MDSCTK_ERR_HIP (-2)      /* HIP runtime failure */
#define MDSCTK_ERR_NOMEM (-3)    /* device allocation failed or HBM budget exceeded */
#define MDSCTK_ERR_IO (-4)       /* file could not be opened / read / written */
#define MDSCTK_ERR_FORMAT (-5)   /* malformed input file */
#define MDSCTK_ERR_NATOMS (-6)   /* atom count mismatch (knn_rms.cpp:158-178 exits 4) */
#define MDSCTK_ERR_STATE (-7)    /* no reference set, or request unknown */
#define MDSCTK_ERR_PENDING (-8)  /* request still running (mdsctk_knn_poll) */

/* Precision of the cross-covariance contraction.  Every form re-scores its
 * candidate shortlist in FP64, so each returned distance is the FP64
 * distance of its returned index.  The form decides how candidates are found
 * (and the speed): the index set is exact within the candidate window's
 * envelope, a measured bound (2x margin) on the approximate scores'
 * deviation that the merge checks at run time on the re-scored candidates
 * and the keys just above each row's cut (rows where it fails are recomputed
 * exactly; a detector, not a proof).  MDSCTK_PREC_EXACT_FP64 scores every
 * pair in FP64. */
#define MDSCTK_PREC_FP32 0        /* FP32 MFMA (v_mfma_f32_32x32x2_f32) */
#define MDSCTK_PREC_SPLIT_BF16 1  /* hi/lo BF16, 3 MFMA products, FP32 accumulate */
#define MDSCTK_PREC_SPLIT_FP16 2  /* hi/lo FP16, 3 MFMA products, FP32 accumulate */
#define MDSCTK_PREC_EXACT_FP64 3  /* FP64 for every pair (no MFMA); slow, for checks */

/* Query flags. */
#define MDSCTK_KNN_NOFIT 1u       /* --nofit: centre only, no rotation (knn_rms.cpp:39) */
#define MDSCTK_KNN_FULL_ROWS 2u   /* --sort false: full unsorted rows, identity indices */
#define MDSCTK_KNN_KEEP_RANK0 4u  /* keep the nearest neighbour instead of dropping it */

/* Options for mdsctk_knn_set_option. */
#define MDSCTK_OPT_HBM_BUDGET 1   /* bytes of device memory the context may hold (0 = no limit);
                                     a query that would exceed it takes fewer candidate lists,
                                     or exact rows in batches that fit, with the same outputs */
#define MDSCTK_OPT_EPS_SCALE 2    /* multiplier on the candidate window (default 1); rows whose
                                     re-scored candidates, or the keys just above their cut,
                                     deviate by more than the scaled window's bound are
                                     recomputed exactly.  The check sees only those keys: a
                                     scale below 1 costs time, and exactness only as far as a
                                     pair outside them breaks the bound unseen */
#define MDSCTK_OPT_REF_SPLITS 3   /* force the number of reference phases (0 = automatic) */
#define MDSCTK_OPT_SERVE_SLOTS 4  /* batches in flight at once when serving (1..4, default 2) */
#define MDSCTK_OPT_SERVE_MAX_FRAMES 5 /* query frames per served batch (default 2048) */
#define MDSCTK_OPT_SERVE_GRAPHS 6   /* 1 (default): each served batch's kernels are captured
                                       into a hipGraph and replayed; 0: launched directly */
#define MDSCTK_OPT_SERVE_SHARE_CUS 7 /* 1 (default): each served batch's persistent
                                        contraction takes 1/slots of the CUs, so batches in
                                        flight run side by side; 0: each takes every CU and
                                        batches follow one another */
#define MDSCTK_OPT_SELF_SYMMETRIC 8 /* 1 (default): a self query (mdsctk_knn_query with
                                       xyz NULL) scores each unordered pair once: the
                                       contraction covers the upper block triangle, and a
                                       pair contracted one way round also joins the other
                                       frame's mirrored list (64 KiB per frame by default,
                                       MDSCTK_OPT_MIRROR_CAP, plus the
                                       candidate lists).  No score matrix is written.  Used
                                       when k+1 <= 896 and the lists fit the free device
                                       memory and the HBM budget; outputs are identical
                                       either way.  0: the candidate-list path */
#define MDSCTK_OPT_MIRROR_CAP 9     /* entries of each frame's mirrored list in the symmetric
                                       self query (default 8192, 8 bytes each).  A frame
                                       whose list overflows is recomputed through the
                                       candidate lists, with the same outputs */

typedef struct mdsctk_knn_ctx mdsctk_knn_ctx;

/* Timings and counters of the last completed query (device times from HIP
 * events on the context's stream). */
typedef struct mdsctk_knn_stats {
  double ms_prep;          /* centring/operand pass over the query frames */
  double ms_contract;      /* fused contraction + QCP + top-k list kernel */
  double ms_merge;         /* candidate selection + FP64 re-score + output */
  double ms_fallback;      /* exact rows for queries whose shortlist overflowed */
  double ms_total;         /* device time of the whole query (incl. copies) */
  int64_t n_pairs;         /* n_fit * n_ref */
  int64_t n_candidates;    /* candidates re-scored in FP64 */
  int64_t n_fallback_rows; /* rows recomputed exactly */
  int32_t ref_splits;      /* reference phases used by the contraction */
  int32_t list_capacity;   /* per-list candidate capacity */
  double eps_rel;          /* candidate window relative to (G_q + max G_ref) */
  int64_t n_pairs_computed; /* pairs the contraction scored: n_pairs, or about half of
                               them for a symmetric self query */
  double ms_landmark;       /* symmetric self query: its landmark pass (thresholds) */
  int64_t n_pairs_landmark; /* pairs the landmark pass scored */
  int64_t n_mirrored;       /* scores appended to frames' mirrored lists */
  int64_t n_mirrored_max;   /* most appended to one frame's mirrored list */
  int64_t n_relisted_rows;  /* rows recomputed through the candidate lists (a mirrored
                               list overflowed) */
  int64_t device_bytes;     /* device memory the context holds when the stats are read
                               (the amount MDSCTK_OPT_HBM_BUDGET caps) */
} mdsctk_knn_stats;

/* Create a context on HIP device `device` with the given MDSCTK_PREC_*.
 * Replaces nothing in the reference (the reference has no device). */
int mdsctk_knn_create(mdsctk_knn_ctx** ctx, int device, int precision);

/* Release all device memory and the stream. */
int mdsctk_knn_destroy(mdsctk_knn_ctx* ctx);

/* Text of the last failure of a call the calling thread made on this
 * context (never NULL; empty when there was none). */
const char* mdsctk_knn_last_error(const mdsctk_knn_ctx* ctx);

int mdsctk_knn_set_option(mdsctk_knn_ctx* ctx, int option, double value);

/* Upload and centre the reference set; it stays resident in HBM.
 * Replaces the reference-frame load/centre loop knn_rms.cpp:181-193
 * (weights = topology masses, knn_rms.cpp:181-182). */
int mdsctk_knn_set_reference(mdsctk_knn_ctx* ctx, const float* xyz_nm, int64_t n_ref,
                             int32_t natoms, const float* weights);

/* Same, from device memory (already in HBM). */
int mdsctk_knn_set_reference_device(mdsctk_knn_ctx* ctx, const float* d_xyz_nm, int64_t n_ref,
                                    int32_t natoms, const float* weights);

/* k nearest reference frames of every query frame.
 * Replaces the pair loops + permutation::sort + row writes
 * (knn_rms.cpp:231-291, mdsctk.h:177-199).
 *   xyz_nm   query frames [n_fit][natoms][3]; NULL queries the reference set
 *            against itself (knn_rms without -f, knn_rms.cpp:103-104), each
 *            unordered pair scored once when MDSCTK_OPT_SELF_SYMMETRIC allows
 *            (same outputs).
 *   k        neighbours per row; clamped to n_ref-1 (knn_rms.cpp:224-225).
 *   out_dist [n_fit][k] Angstrom, ascending; out_idx [n_fit][k] 0-based.
 *            With MDSCTK_KNN_FULL_ROWS the outputs are [n_fit][n_ref] in
 *            reference order and out_idx is the identity.
 * Returns the number of entries written per row. */
int mdsctk_knn_query(mdsctk_knn_ctx* ctx, const float* xyz_nm, int64_t n_fit, int32_t k,
                     uint32_t flags, double* out_dist, int32_t* out_idx);

/* Same with query frames and outputs in device memory (d_xyz_nm NULL = self);
 * work is ordered on the context's stream and complete on return. */
int mdsctk_knn_query_device(mdsctk_knn_ctx* ctx, const float* d_xyz_nm, int64_t n_fit,
                            int32_t k, uint32_t flags, double* d_out_dist, int32_t* d_out_idx);

/* Same as mdsctk_knn_query_device, but the call allocates nothing: no
 * hipMalloc or hipFree, and no other allocating routine.  No counterpart in
 * the reference; it is for callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * it).  Its workspace is a fixed per-GPU region of 128 MiB inside the
 * library (held by every GPU the library runs on until the process ends).
 * Sizes whose workspace does not fit it are refused with MDSCTK_ERR_ARG:
 * the query frames' derived arrays (about 12 natoms + 64 bytes each), their
 * candidate lists (the reference phases are lowered until they fit), and
 * one exact FP64 row of 8 n_ref bytes.  A self query (d_xyz_nm NULL) takes
 * the candidate lists, never the symmetric form.  Frames only (a context
 * holding data vectors is refused).  Every call on one GPU shares that
 * GPU's region, whatever the context: calls that overlap on one GPU race,
 * so callers serialise them. */
int mdsctk_knn_query_device_noalloc(mdsctk_knn_ctx* ctx, const float* d_xyz_nm, int64_t n_fit,
                                    int32_t k, uint32_t flags, double* d_out_dist,
                                    int32_t* d_out_idx);

/* Diagnostic (no counterpart in the reference): the approximate score
 * s~ = G_r + G_f - 2 lambda~ (mass nm^2) that the contraction form computes
 * for every pair, and optionally the exact FP64 score, both [n_fit][n_ref].
 * Used to measure the contraction error that sizes the candidate window.
 * Limited to n_fit*n_ref <= 6e7. */
int mdsctk_knn_debug_scores(mdsctk_knn_ctx* ctx, const float* xyz_nm, int64_t n_fit,
                            uint32_t flags, double* approx_s, double* exact_s);

/* ---- Serving: asynchronous requests against the resident reference set.
 *
 * The reference tool has no counterpart: this is the out-of-sample
 * (Nystrom) query form of knn_rms.cpp:231-291 with the reference set kept
 * resident.  The first submit starts a serving thread that owns
 * MDSCTK_OPT_SERVE_SLOTS batch slots, each with its own HIP stream.
 * Pending requests are admitted into the next free slot, whole requests of
 * at most 64 frames first and then the rest in arrival order, up to
 * MDSCTK_OPT_SERVE_MAX_FRAMES frames per batch (clamped so that the
 * slots fit the HBM budget); a request larger than that is split over
 * consecutive batches by rows.  A batch is one fused launch sequence over
 * the concatenated frames of its requests, so requests that arrive while
 * others run join the next batch and overlap the running one on the device.
 * Only requests with the same k and flags share a batch.
 *
 * Results do not depend on what a request was batched with: every output
 * value is the FP64 re-score of the certified candidate window, sorted by
 * (distance, index), and a pair's FP64 score is computed the same way in
 * every batch.
 *
 * Buffers stay owned by the caller and must stay valid until the request
 * is reported done (xyz_nm may be released once done as well). */

typedef struct mdsctk_knn_request_info {
  int32_t status;          /* entries per row (>= 0) when done, MDSCTK_ERR_PENDING, or an error */
  int32_t n_batches;       /* batches the request's rows were spread over */
  int64_t n_fallback_rows; /* rows that took the exact path */
  double t_submit_ms;      /* mdsctk_knn_clock_ms() at submit */
  double t_admit_ms;       /* first rows placed in a batch */
  double t_done_ms;        /* outputs written */
  int64_t batch_frames;    /* frames of the largest batch it ran in */
} mdsctk_knn_request_info;

typedef struct mdsctk_knn_serve_stats {
  int64_t requests_done;
  int64_t batches;
  int64_t frames;          /* query frames served */
  int64_t pairs;           /* frames * n_ref */
  int64_t max_batch_frames;
  int64_t fallback_rows;
  int32_t slots;
  int32_t slot_frames;     /* frames per batch after the HBM budget clamp */
  double busy_ms;          /* wall time with at least one batch in flight */
  int64_t graph_launches;  /* batches replayed as a hipGraph */
  int64_t graph_builds;    /* executable graphs instantiated (the others were updated) */
} mdsctk_knn_serve_stats;

/* Queue a query (same arguments and output layout as mdsctk_knn_query;
 * NULL xyz_nm and MDSCTK_KNN_FULL_ROWS are not accepted).  Returns 0 and
 * the request id, or a negative error. */
int mdsctk_knn_submit(mdsctk_knn_ctx* ctx, const float* xyz_nm, int64_t n_fit, int32_t k,
                      uint32_t flags, double* out_dist, int32_t* out_idx, int64_t* request_id);

/* State of a request: entries per row when done, MDSCTK_ERR_PENDING while
 * queued or running, or the request's error code.  A request reported done
 * or failed is retired (its id is then unknown: MDSCTK_ERR_STATE).  `info`
 * may be NULL. */
int mdsctk_knn_poll(mdsctk_knn_ctx* ctx, int64_t request_id, mdsctk_knn_request_info* info);

/* Block until the request is done or failed; then as mdsctk_knn_poll. */
int mdsctk_knn_wait(mdsctk_knn_ctx* ctx, int64_t request_id, mdsctk_knn_request_info* info);

/* Counters of the serving loop since the first submit. */
int mdsctk_knn_serve_stats_get(mdsctk_knn_ctx* ctx, mdsctk_knn_serve_stats* out);

/* The host clock the request times use: milliseconds of the steady clock
 * since the library was loaded, the same for every context. */
double mdsctk_knn_clock_ms(const mdsctk_knn_ctx* ctx);

/* Statistics of the last completed query. */
int mdsctk_knn_get_stats(const mdsctk_knn_ctx* ctx, mdsctk_knn_stats* out);

/* Number of entries per row a query with this k and flags returns. */
int mdsctk_knn_row_width(const mdsctk_knn_ctx* ctx, int32_t k, uint32_t flags);

/* ---- knn_data: data vectors instead of frames.
 *
 * Replaces knn_data's pair loop (knn_data.cpp:196-237) with its distance
 * functions euclidean_distance / correlation_distance (mdsctk.cpp:329-356,
 * selected by -c at knn_data.cpp:103-105) on the same engine: contraction
 * for candidates, FP64 re-score of the certified window, (distance, index)
 * order, k clamp, drop-rank-0 and --sort false layout as mdsctk_knn_query.
 * A context holds either frames or data vectors; the frame entry points
 * (and serving) refuse a context holding data vectors and the reverse.
 * Precision forms: split-BF16, FP32 or exact (FP16 is refused: arbitrary
 * data ranges do not fit it). */
#define MDSCTK_METRIC_EUCLIDEAN 1   /* sqrt(sum (a-b)^2) */
#define MDSCTK_METRIC_CORRELATION 2 /* sqrt(max(0, (1 - r) / 2)), r the Pearson correlation */

/* Resident set of n_ref vectors x[n_ref][vdim] (FP64, as the .pts file
 * holds them).  Non-finite values, and constant vectors under the
 * correlation metric, are rejected with MDSCTK_ERR_ARG. */
int mdsctk_data_set_reference(mdsctk_knn_ctx* ctx, const double* x, int64_t n_ref, int32_t vdim,
                              int metric);

/* k nearest of each query vector x[n_fit][vdim] (NULL = self query) in the
 * resident set; flags MDSCTK_KNN_FULL_ROWS / MDSCTK_KNN_KEEP_RANK0 as for
 * mdsctk_knn_query (MDSCTK_KNN_NOFIT is refused).  Distances are in the data's
 * units.  Returns the number of entries written per row. */
int mdsctk_data_query(mdsctk_knn_ctx* ctx, const double* x, int64_t n_fit, int32_t k,
                      uint32_t flags, double* out_dist, int32_t* out_idx);

/* Diagnostic (no counterpart in the reference): approximate and exact
 * scores [n_fit][n_ref] of data vectors, as mdsctk_knn_debug_scores
 * (s = squared Euclidean distance, or 2 (1 - r) for the correlation
 * metric). */
int mdsctk_data_debug_scores(mdsctk_knn_ctx* ctx, const double* x, int64_t n_fit,
                             double* approx_s, double* exact_s);

#ifdef __cplusplus
}
#endif

#endif /* MDSCTK_KNN_H */
