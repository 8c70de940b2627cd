// gfx950 kernels of the knn_rms engine.
//
// Pipeline for one query batch (reference: knn_rms.cpp:231-291):
//   frame_stats_kernel
//                    one wave per frame: FP64 mass-weighted centroid,
//                    G = sum w|x-c|^2 (FP64) and G~ of the operand.
//   operand_kernel   per (32-frame tile, 4 K-groups): the contraction operand
//                    sqrt(w)(x-c) rounded to FP32 or split into hi/lo 16-bit
//                    halves (dithered for BF16), written fragment-major
//                    (below).  HBM-bound.
//   contract_kernel  all-pairs 3x3 cross-covariance blocks as one dense MFMA
//                    contraction (9 accumulator tiles per wave: planes c,d of
//                    32 ref x 32 fit frames), FP64 QCP Newton epilogue giving
//                    an approximate score s~ = G_r + G_f - 2 lambda_max, and
//                    per-lane candidate lists (threshold + append, ballot
//                    select compaction); no distance matrix is written.
//                    For a self query it may instead score the upper block
//                    triangle only and write the n x n FP32 score matrix,
//                    from which sym_select_kernel lists each row's
//                    candidates.
//   merge_kernel     per query: radix-select the (k+1)-th smallest s~ over
//                    its lists, take every candidate within the certified
//                    window, re-score them exactly in FP64 from the raw
//                    coordinates, sort by (s, index), write ranks 1..k.
//   exact_rows / select_rows
//                    FP64 scores of whole rows: overflow fallback, the
//                    MDSCTK_PREC_EXACT form, and --sort false full rows.
//
// Operand layout ("fragment-major"): for tile T (32 frames), K-group g and
// block b (plane c, part), 64 lanes x 16 bytes are contiguous, laid out
// exactly as the MFMA A/B operand registers want them:
//   split (32x32x16 bf16/f16): lane l holds frame l%32, atoms 16g+8(l/32)+j,
//                              j=0..7; b = 2c + part (part 0 = hi, 1 = lo)
//   fp32  (32x32x2 f32):       lane l holds frame l%32, atoms 8g+2j+(l/32),
//                              j=0..3 (one per MFMA step); b = c
// so every global load and every LDS read of a fragment is one contiguous
// 1 KB access per wave.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cfloat>
#include <climits>
#include <cmath>
#include <mutex>

#include "knn_device.h"

namespace mdsctk {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr unsigned kInfBits = 0x7f800000u;
constexpr int kCandMax = 2048;
constexpr int kCandSmall = 1024;   // merge instance for k1 <= kMergeSmallK1
constexpr int kMergeSmallK1 = 256;

// ---------------------------------------------------------------------------
// small helpers

__device__ __forceinline__ unsigned short bf16_rne(float v) {
  unsigned u = __float_as_uint(v);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ double bf16_val(unsigned short h) {
  return (double)__uint_as_float(((unsigned)h) << 16);
}
__device__ __forceinline__ unsigned short f16_rne(float v) {
  _Float16 h = (_Float16)v;
  return __builtin_bit_cast(unsigned short, h);
}
__device__ __forceinline__ double f16_val(unsigned short h) {
  return (double)(float)__builtin_bit_cast(_Float16, h);
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ double det3(const double* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Horn key matrix entries of S (row-major 3x3) and the coefficients of its
// characteristic polynomial l^4 + c2 l^2 + c1 l + c0.
__device__ __forceinline__ void qcp_coeffs(const double* S, double& c2, double& c1, double& c0) {
  const double xx = S[0], xy = S[1], xz = S[2];
  const double yx = S[3], yy = S[4], yz = S[5];
  const double zx = S[6], zy = S[7], zz = S[8];
  const double k00 = xx + yy + zz, k01 = yz - zy, k02 = zx - xz, k03 = xy - yx;
  const double k11 = xx - yy - zz, k12 = xy + yx, k13 = zx + xz;
  const double k22 = -xx + yy - zz, k23 = yz + zy;
  const double k33 = -xx - yy + zz;
  // det(K) by 2x2 minors of rows (0,1) and (2,3); K symmetric.
  const double s0 = k00 * k11 - k01 * k01;
  const double s1 = k00 * k12 - k01 * k02;
  const double s2 = k00 * k13 - k01 * k03;
  const double s3 = k01 * k12 - k11 * k02;
  const double s4 = k01 * k13 - k11 * k03;
  const double s5 = k02 * k13 - k12 * k03;
  const double t5 = k22 * k33 - k23 * k23;
  const double t4 = k12 * k33 - k13 * k23;
  const double t3 = k12 * k23 - k13 * k22;
  const double t2 = k02 * k33 - k03 * k23;
  const double t1 = k02 * k23 - k03 * k22;
  const double t0 = k02 * k13 - k03 * k12;
  c0 = s0 * t5 - s1 * t4 + s2 * t3 + s3 * t2 - s4 * t1 + s5 * t0;
  double fro = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) fro += S[i] * S[i];
  c2 = -2.0 * fro;
  c1 = -8.0 * det3(S);
}

// Largest eigenvalue of a symmetric 4x4 matrix by cyclic Jacobi (used only
// when Newton does not converge, e.g. near-linear frames).
__device__ double jacobi4_max(const double* S) {
  double a[4][4];
  {
    const double xx = S[0], xy = S[1], xz = S[2], yx = S[3], yy = S[4], yz = S[5];
    const double zx = S[6], zy = S[7], zz = S[8];
    a[0][0] = xx + yy + zz; a[0][1] = yz - zy; a[0][2] = zx - xz; a[0][3] = xy - yx;
    a[1][1] = xx - yy - zz; a[1][2] = xy + yx; a[1][3] = zx + xz;
    a[2][2] = -xx + yy - zz; a[2][3] = yz + zy; a[3][3] = -xx - yy + zz;
    a[1][0] = a[0][1]; a[2][0] = a[0][2]; a[3][0] = a[0][3];
    a[2][1] = a[1][2]; a[3][1] = a[1][3]; a[3][2] = a[2][3];
  }
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 3; ++p)
      for (int q = p + 1; q < 4; ++q) off += fabs(a[p][q]);
    if (off < 1e-300) break;
    for (int p = 0; p < 3; ++p)
      for (int q = p + 1; q < 4; ++q) {
        const double apq = a[p][q];
        if (apq == 0.0) continue;
        const double th = (a[q][q] - a[p][p]) / (2.0 * apq);
        double t = 1.0 / (fabs(th) + sqrt(th * th + 1.0));
        if (th < 0.0) t = -t;
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 4; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 4; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
      }
  }
  return fmax(fmax(a[0][0], a[1][1]), fmax(a[2][2], a[3][3]));
}

// Exact lambda_max of K(S) for the FP64 re-score.
__device__ double qcp_lambda_exact(const double* S_in, double ga, double gb) {
  const double g = 0.5 * (ga + gb);
  if (!(g > 0.0)) return 0.0;
  double S[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) S[i] = S_in[i] / g;
  double c2, c1, c0;
  qcp_coeffs(S, c2, c1, c0);
  double l = 1.0, dp = 0.0;
  bool conv = false;
  for (int it = 0; it < 100; ++it) {
    const double l2 = l * l;
    const double p = (l2 + c2) * l2 + c1 * l + c0;
    dp = (4.0 * l2 + 2.0 * c2) * l + c1;
    if (dp == 0.0) break;
    const double step = p / dp;
    l -= step;
    if (fabs(step) <= 1e-15 * fabs(l)) { conv = true; break; }
  }
  // A nearly double largest root (near-collinear frames) leaves Newton with
  // about half the digits; the Jacobi eigen-solve resolves it.
  if (!conv || fabs(dp) < 1e-4) l = jacobi4_max(S);
  return l * g;
}

// Approximate score for candidate selection.  Newton from an upper bound of
// lambda_max decreases monotonically to it (the key matrix is symmetric, so
// the quartic's roots are real and it is convex and increasing right of the
// largest), so every iterate gives a lower bound of s~; once that bound
// exceeds `reject` the pair cannot be a candidate and iteration stops.  The
// start is l = 1, i.e. (G_a+G_b)/2, which bounds lambda_max for exact
// sums; approximate m can exceed it (then s~ < 0: large G terms against
// small scores), so the start is raised to the column-norm bound of m when
// that is larger -- starting below the largest root, an iterate could pass
// `reject` on its way and drop the pair.
__device__ __forceinline__ float lambda_upper(const float* m);
__device__ __forceinline__ double qcp_score_approx(const float* m, double ga, double gb,
                                                   double reject, bool nofit) {
  if (nofit) return ga + gb - 2.0 * ((double)m[0] + (double)m[4] + (double)m[8]);
  const double g = 0.5 * (ga + gb);
  if (!(g > 0.0)) return 0.0;
  // Any consistent scale works here: s = 2 g (1 - l) is evaluated with the
  // same g, so the reciprocal's last-bit error only rescales l.
  const double ig = __builtin_amdgcn_rcp(g);
  double S[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) S[i] = (double)m[i] * ig;
  double c2, c1, c0;
  qcp_coeffs(S, c2, c1, c0);
  const double rej = 0.5 * reject * ig;
  double l = fmax(1.0, (double)lambda_upper(m) * ig);
  for (int it = 0; it < 24; ++it) {
    const double l2 = l * l;
    const double p = (l2 + c2) * l2 + c1 * l + c0;
    const double dp = (4.0 * l2 + 2.0 * c2) * l + c1;
    const double step = p * __builtin_amdgcn_rcp(dp);
    l -= step;
    if (1.0 - l > rej) break;
    if (fabs(step) < 1e-11) break;
  }
  const double s = 2.0 * g * (1.0 - l);
  // A non-finite iterate (degenerate input) must not hide a pair: report 0
  // so that it becomes a candidate and is re-scored exactly.
  return (s == s && s < INFINITY) ? s : 0.0;
}

// An upper bound of lambda_max(K(m)), the largest eigenvalue of the key
// matrix of the 3x3 cross-covariance m: lambda_max = max over rotations R of
// tr(R m) = sum_j <row j of R, column j of m> <= the sum of the column norms
// of m, since every row of R has norm 1.  Evaluated in FP32 and raised by
// 1e-5 relative, which covers the rounding of the norms and their sum (a
// few units of 2^-24 each).  For two similar frames, each expressed in its
// principal axes, m is nearly diagonal and the bound nearly exact.
__device__ __forceinline__ float lambda_upper(const float* m) {
  float cs = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) cs += sqrtf(fmaf(m[i], m[i], fmaf(m[3 + i], m[3 + i], m[6 + i] * m[6 + i])));
  return cs * (1.0f + 1e-5f);
}

// ---------------------------------------------------------------------------
// prep: one 192-thread workgroup per 32-frame tile.  Thread t -> plane
// c = t/64, lane l = t%64 of that plane's operand block.

// One atom's coordinates: a single 12-byte load per lane.
struct F3 {
  float x, y, z;
};

// Per-atom scale of the dithered split (a golden-ratio sequence in [1, 2)).
// Reference operands carry sqrt(w) x * s_a and query operands sqrt(w) x / s_a,
// so every product is unchanged, but the two sides round on different
// grids: the hi/lo residues of the same coordinate no longer coincide, and
// the dropped lo*lo' term of aligned or duplicate frames stops adding up
// coherently.
__device__ __forceinline__ double atom_scale(int a) {
  const double t = (double)a * 0.6180339887498949;
  return 1.0 + (t - floor(t));
}

// The operand value of sqrt(w) (x - c) (already multiplied by the side's
// dither scale) as the contraction form represents it: FP32, or hi + lo.
template <int PREC>
__device__ __forceinline__ double represented(double v) {
  if constexpr (PREC == kFP32) {
    return (double)(float)v;
  } else if constexpr (PREC == kBF16) {
    const double vh = bf16_val(bf16_rne((float)v));
    return vh + bf16_val(bf16_rne((float)(v - vh)));
  } else {
    const double vh = f16_val(f16_rne((float)v));
    return vh + f16_val(f16_rne((float)(v - vh)));
  }
}

__device__ __forceinline__ double side_scale(int side, int a) {
  return side == 0 ? 1.0 : side > 0 ? atom_scale(a) : 1.0 / atom_scale(a);
}

// Principal axes of a frame: the eigenvectors of T = sum w d d^T (d the
// centred coordinates; T symmetric 3x3, upper triangle t = (xx, xy, xz, yy,
// yz, zz)) by cyclic Jacobi in FP64, as the columns of a proper rotation Q
// (row-major q[3*i + j] = component i of axis j), axes by decreasing
// eigenvalue, the first two signed so that their largest component is
// positive and the third their cross product.  Every lane of a wave computes
// the same Q from the same sums.  The RMSD after optimal superposition does
// not depend on the orientation of either frame, so expressing every frame
// in its own axes changes no score; it makes the cross-covariance of two
// similar frames nearly diagonal, which is what the contraction's rejection
// bound needs.
__device__ void canonical_axes(const double* t, double* q) {
  double a[3][3] = {{t[0], t[1], t[2]}, {t[1], t[3], t[4]}, {t[2], t[4], t[5]}};
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 12; ++sweep) {
    const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    const double scale = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
    if (off <= 1e-15 * scale || off == 0.0) break;
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0, r = pq == 0 ? 1 : 2;
      const double apr = a[p][r];
      if (apr == 0.0) continue;
      const double th = (a[r][r] - a[p][p]) / (2.0 * apr);
      double tt = 1.0 / (fabs(th) + sqrt(th * th + 1.0));
      if (th < 0.0) tt = -tt;
      const double cs = 1.0 / sqrt(tt * tt + 1.0), sn = tt * cs;
      for (int k = 0; k < 3; ++k) {
        const double akp = a[k][p], akr = a[k][r];
        a[k][p] = cs * akp - sn * akr;
        a[k][r] = sn * akp + cs * akr;
      }
      for (int k = 0; k < 3; ++k) {
        const double apk = a[p][k], ark = a[r][k];
        a[p][k] = cs * apk - sn * ark;
        a[r][k] = sn * apk + cs * ark;
      }
      for (int k = 0; k < 3; ++k) {
        const double vkp = v[k][p], vkr = v[k][r];
        v[k][p] = cs * vkp - sn * vkr;
        v[k][r] = sn * vkp + cs * vkr;
      }
    }
  }
  // order by decreasing eigenvalue
  int o[3] = {0, 1, 2};
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2 - i; ++j)
      if (a[o[j]][o[j]] < a[o[j + 1]][o[j + 1]]) {
        const int x = o[j];
        o[j] = o[j + 1];
        o[j + 1] = x;
      }
  // A frame that is a line to within far better than FP32 resolves
  // (second moment under 1e-4 of the first) keeps its own orientation: in
  // their axes such frames align along one axis, and the contraction's
  // approximate scores of aligned, nearly collinear frames of 1000 atoms and
  // more were measured at up to 66 times the window (as given: under half
  // of it; every other generator in its axes: under 0.6).  A plane to within
  // the same ratio (third moment) keeps it too, as the same kind of
  // degenerate frame.  Any proper rotation per frame leaves the scores
  // unchanged.
  if (!(a[o[1]][o[1]] >= 1e-4 * a[o[0]][o[0]] && a[o[2]][o[2]] >= 1e-4 * a[o[0]][o[0]])) {
#pragma unroll
    for (int i = 0; i < 9; ++i) q[i] = (i % 4 == 0) ? 1.0 : 0.0;
    return;
  }
  double e[2][3];
  for (int c = 0; c < 2; ++c) {
    int big = 0;
    for (int i = 1; i < 3; ++i)
      if (fabs(v[i][o[c]]) > fabs(v[big][o[c]])) big = i;
    const double sg = v[big][o[c]] < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < 3; ++i) e[c][i] = sg * v[i][o[c]];
  }
  const double e2[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                        e[0][0] * e[1][1] - e[0][1] * e[1][0]};
  for (int i = 0; i < 3; ++i) {
    q[3 * i + 0] = e[0][i];
    q[3 * i + 1] = e[1][i];
    q[3 * i + 2] = e2[i];
  }
}

// The centred coordinate e along axis (q0, q1, q2), one rounding order
// wherever it is formed (G~ and the operand must see the same value).
__device__ __forceinline__ double axis_coord(double e0, double e1, double e2, double q0, double q1,
                                             double q2) {
  return fma(e2, q2, fma(e1, q1, e0 * q0));
}

// One wave per frame row (rows up to n_tiles*32; padded rows get zeros),
// reading the frame as one contiguous run of 12-byte atoms:
//   cen  FP64 mass-weighted centroid (center == 0, data vectors: 0)
//   G    sum w |x - c|^2 in FP64
//   Gt   G~ = sum of the squared operand values as the contraction form
//        represents them (scale of the dither side taken back out); Gt2,
//        when given, the same for a second side (the resident set's
//        query-side operand).
// perm: row f holds frame perm[f] of xyz (else frame f).  rot: the frame's
// coordinates are taken in its principal axes (canonical_axes, Q stored at
// rot[9 f]), and G~ is that of the rotated operand.
template <int PREC>
__global__ __launch_bounds__(256) void frame_stats_kernel(
    const float* __restrict__ xyz, int64_t n, int64_t rows, int natoms,
    const float* __restrict__ w, const float* __restrict__ sw, double Wtot, int center,
    int side, int side2, double* __restrict__ cen, double* __restrict__ G,
    double* __restrict__ Gt, double* __restrict__ Gt2, const int* __restrict__ perm,
    double* __restrict__ rot) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t f = (int64_t)blockIdx.x * 4 + wave;
  if (f >= rows) return;
  double c0 = 0.0, c1 = 0.0, c2 = 0.0, g = 0.0, gt = 0.0, gt2 = 0.0;
  double q[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (f < n) {
    const int64_t src = perm ? (int64_t)perm[f] : f;
    const F3* x = reinterpret_cast<const F3*>(xyz) + src * (int64_t)natoms;
    // with rot, the second moments in the same pass: T = sum w x x^T - W c c^T
    // (its cancellation only perturbs the axes, which need not be exact)
    double t[6] = {0, 0, 0, 0, 0, 0};
    if (center) {
      for (int a = lane; a < natoms; a += 64) {
        const double wa = (double)w[a];
        const F3 v = x[a];
        const double x0 = (double)v.x, x1 = (double)v.y, x2 = (double)v.z;
        c0 += wa * x0;
        c1 += wa * x1;
        c2 += wa * x2;
        if (rot) {
          t[0] += wa * x0 * x0; t[1] += wa * x0 * x1; t[2] += wa * x0 * x2;
          t[3] += wa * x1 * x1; t[4] += wa * x1 * x2; t[5] += wa * x2 * x2;
        }
      }
      c0 = wave_sum_d(c0) / Wtot;
      c1 = wave_sum_d(c1) / Wtot;
      c2 = wave_sum_d(c2) / Wtot;
    }
    if (rot) {
#pragma unroll
      for (int i = 0; i < 6; ++i) t[i] = wave_sum_d(t[i]);
      t[0] -= Wtot * c0 * c0; t[1] -= Wtot * c0 * c1; t[2] -= Wtot * c0 * c2;
      t[3] -= Wtot * c1 * c1; t[4] -= Wtot * c1 * c2; t[5] -= Wtot * c2 * c2;
      canonical_axes(t, q);
    }
    for (int a = lane; a < natoms; a += 64) {
      const F3 v = x[a];
      const double e0 = (double)v.x - c0, e1 = (double)v.y - c1, e2 = (double)v.z - c2;
      g += (double)w[a] * (e0 * e0 + e1 * e1 + e2 * e2);
      // the coordinates in the frame's axes (the identity without rot)
      const double d0 = rot ? axis_coord(e0, e1, e2, q[0], q[3], q[6]) : e0;
      const double d1 = rot ? axis_coord(e0, e1, e2, q[1], q[4], q[7]) : e1;
      const double d2 = rot ? axis_coord(e0, e1, e2, q[2], q[5], q[8]) : e2;
      if (Gt) {
        const double s = (double)sw[a], sc = side_scale(side, a);
        const double r0 = represented<PREC>(s * d0 * sc) / sc;
        const double r1 = represented<PREC>(s * d1 * sc) / sc;
        const double r2 = represented<PREC>(s * d2 * sc) / sc;
        gt += r0 * r0 + r1 * r1 + r2 * r2;
        if (Gt2) {
          const double sc2 = side_scale(side2, a);
          const double q0 = represented<PREC>(s * d0 * sc2) / sc2;
          const double q1 = represented<PREC>(s * d1 * sc2) / sc2;
          const double q2 = represented<PREC>(s * d2 * sc2) / sc2;
          gt2 += q0 * q0 + q1 * q1 + q2 * q2;
        }
      }
    }
    g = wave_sum_d(g);
    gt = wave_sum_d(gt);
    gt2 = wave_sum_d(gt2);
  }
  if (lane == 0) {
    cen[f * 3 + 0] = c0;
    cen[f * 3 + 1] = c1;
    cen[f * 3 + 2] = c2;
    G[f] = g;
    if (Gt) Gt[f] = gt;
    if (Gt2) Gt2[f] = gt2;
  }
  if (rot && lane < 9) rot[f * 9 + lane] = q[lane];
}

// The contraction operand in fragment order, from frame_stats_kernel's
// centroids: one workgroup per (32-frame tile, kGroupsPerStage K-groups),
// staged through LDS once.  Thread t works on plane c = t/64, lane
// l = t%64 of that plane's operand block.  The arithmetic of each value is
// represented<PREC>'s.
constexpr int kGroupsPerStage = 4;

// perm and rot as frame_stats_kernel (rot: the Q it stored for each row).
// op2 (split forms): the operand of a second dither side, side2, from the
// same staged frames.
template <int PREC>
__global__ __launch_bounds__(192) void operand_kernel(const float* __restrict__ xyz, int64_t n,
                                                      int natoms, int n_groups,
                                                      const float* __restrict__ sw,
                                                      const double* __restrict__ cen, int side,
                                                      uint4* __restrict__ op,
                                                      const int* __restrict__ perm,
                                                      const double* __restrict__ rot, int side2,
                                                      uint4* __restrict__ op2) {
  constexpr int KG = (PREC == kFP32) ? 8 : 16;
  constexpr int CH = KG * 3;                     // floats of one frame in one K-group
  constexpr int SW = CH * kGroupsPerStage;       // floats of one frame in one stage
  constexpr int SP = SW + 1;                     // LDS row pitch: odd, so the 32 frames
                                                 // of a plane's read fall in 32 banks
  __shared__ float xs[kTile * SP];

  const int tile = blockIdx.x;
  const int g0 = blockIdx.y * kGroupsPerStage;
  const int t = threadIdx.x;
  const int c = t >> 6;
  const int lane = t & 63;
  const int row = lane & 31;
  const int h = lane >> 5;
  const int64_t frow = (int64_t)tile * kTile + row;
  const double cm0 = cen[frow * 3 + 0], cm1 = cen[frow * 3 + 1], cm2 = cen[frow * 3 + 2];
  // plane c of the operand: column c of the frame's Q (or axis c)
  double qc0 = c == 0, qc1 = c == 1, qc2 = c == 2;
  if (rot && frow < n) {
    qc0 = rot[frow * 9 + 0 + c];
    qc1 = rot[frow * 9 + 3 + c];
    qc2 = rot[frow * 9 + 6 + c];
  }
  // the centred coordinate of local atom la along plane c
  auto coord = [&](const float* xr, int la) -> double {
    const double e0 = (double)xr[la * 3 + 0] - cm0, e1 = (double)xr[la * 3 + 1] - cm1,
                 e2 = (double)xr[la * 3 + 2] - cm2;
    if (!rot) return c == 0 ? e0 : c == 1 ? e1 : e2;
    return axis_coord(e0, e1, e2, qc0, qc1, qc2);
  };

  for (int i = t; i < kTile * SW; i += 192) {
    const int fr = i / SW, e = i - fr * SW;
    const int64_t ff = (int64_t)tile * kTile + fr;
    const int a = g0 * KG + e / 3;
    float v = 0.f;
    if (ff < n && a < natoms) {
      const int64_t src = perm ? (int64_t)perm[ff] : ff;
      v = xyz[(src * natoms + (int64_t)g0 * KG) * 3 + e];
    }
    xs[fr * SP + e] = v;
  }
  __syncthreads();
  auto local_atom = [&](int j) { return (PREC == kFP32) ? (2 * j + h) : (8 * h + j); };

  for (int gg = 0; gg < kGroupsPerStage && g0 + gg < n_groups; ++gg) {
    const int g = g0 + gg;
    const float* xr = xs + row * SP + gg * CH;
    if constexpr (PREC == kFP32) {
      float v4[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int la = local_atom(j);
        const int a = g * KG + la;
        v4[j] = a < natoms ? (float)((double)sw[a] * coord(xr, la)) : 0.f;
      }
      uint4 o;
      o.x = __float_as_uint(v4[0]);
      o.y = __float_as_uint(v4[1]);
      o.z = __float_as_uint(v4[2]);
      o.w = __float_as_uint(v4[3]);
      op[(((int64_t)tile * n_groups + g) * 3 + c) * 64 + lane] = o;
      if (op2) op2[(((int64_t)tile * n_groups + g) * 3 + c) * 64 + lane] = o;  // no dither sides
    } else {
      double base[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int la = local_atom(j);
        const int a = g * KG + la;
        base[j] = a < natoms ? (double)sw[a] * coord(xr, la) : 0.0;
      }
      const int64_t blk = (((int64_t)tile * n_groups + g) * 6 + 2 * c);
      auto emit = [&](int sd, uint4* __restrict__ dst) {
        unsigned short hi[8], lo[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int a = g * KG + local_atom(j);
          const double v = a < natoms ? base[j] * side_scale(sd, a) : 0.0;
          if constexpr (PREC == kBF16) {
            hi[j] = bf16_rne((float)v);
            lo[j] = bf16_rne((float)(v - bf16_val(hi[j])));
          } else {
            hi[j] = f16_rne((float)v);
            lo[j] = f16_rne((float)(v - f16_val(hi[j])));
          }
        }
        uint4 oh, ol;
        oh.x = hi[0] | ((unsigned)hi[1] << 16); oh.y = hi[2] | ((unsigned)hi[3] << 16);
        oh.z = hi[4] | ((unsigned)hi[5] << 16); oh.w = hi[6] | ((unsigned)hi[7] << 16);
        ol.x = lo[0] | ((unsigned)lo[1] << 16); ol.y = lo[2] | ((unsigned)lo[3] << 16);
        ol.z = lo[4] | ((unsigned)lo[5] << 16); ol.w = lo[6] | ((unsigned)lo[7] << 16);
        dst[blk * 64 + lane] = oh;
        dst[(blk + 1) * 64 + lane] = ol;
      };
      emit(side, op);
      if (op2) emit(side2, op2);
    }
  }
}

// ---------------------------------------------------------------------------
// contraction + QCP + candidate lists

struct ContractParams {
  const uint4* rop;
  const uint4* fop;
  const double* rGt;
  const double* fGt;
  int64_t n_ref, n_fit;
  int ref_pairs, fit_tiles, n_groups, phases, k1;
  const float* eps2;
  uint2* ent;
  int* cnt;
  float* ovf;
  unsigned* thr;
  int nl, cap;
  int nofit;
  double* dbg;
  // Reference phases actually walked: items are (fit block, phase) for
  // phases 0..ph_run-1 (ph_run < phases: every phases-th reference tile
  // pair only, the landmark pass of a symmetric self query).
  int ph_run;
  // kOutSym: one list per reference frame for the scores it receives as
  // the reference side of a pair (mcap entries each, appended through the
  // atomic counter mcnt; the smallest key dropped past mcap goes to movf).
  uint2* ment;
  int* mcnt;
  float* movf;
  int mcap;
};

// What the contraction's epilogue produces.
constexpr int kOutLists = 0;  // per-query candidate lists (the fused pipeline)
constexpr int kOutDebug = 1;  // every score in FP64 (dbg, n_fit x n_ref)
// A self query (the query set is the reference set, tile for tile): only
// the upper block triangle is contracted -- reference pairs that start at
// or after the workgroup's fit block.  Each score joins the fit frame's own
// candidate list as in kOutLists and, off the diagonal blocks (where the
// pair is contracted both ways round), the reference frame's list as well
// when it lies below that frame's shared threshold.  No score matrix is
// written.
constexpr int kOutSym = 2;

// One K-group of MFMA work for a wave: acc[c*3+d] += A_c * B_d^T over the
// group's atoms, reading the wave's fragments from LDS.  For the split forms
// the three products hi*hi + hi*lo + lo*hi accumulate into the same FP32
// accumulators (the lo*lo term is dropped).
// TRACE: only the trace of the block is needed (--nofit and data vectors):
// the diagonal products A_c * B_c accumulate into acc[0], acc[4], acc[8]
// exactly as in the full product, a third of the MFMA work, and the other
// accumulators stay zero.  (One accumulator for all three would triple the
// FP32 accumulation chain, and with it the error of aligned frames.)
template <int PREC, bool TRACE>
__device__ __forceinline__ void mma_group(f32x16 (&acc)[9], const uint4* __restrict__ As,
                                          const uint4* __restrict__ Bs, int lane) {
  if constexpr (TRACE) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      f32x16& d = acc[4 * c];
      if constexpr (PREC == kFP32) {
        const uint4 a = As[c * 64 + lane], b = Bs[c * 64 + lane];
        d = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.x), __uint_as_float(b.x), d,
                                                 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.y), __uint_as_float(b.y), d,
                                                 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.z), __uint_as_float(b.z), d,
                                                 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.w), __uint_as_float(b.w), d,
                                                 0, 0, 0);
      } else {
        const uint4 ah = As[(2 * c) * 64 + lane], al = As[(2 * c + 1) * 64 + lane];
        const uint4 bh = Bs[(2 * c) * 64 + lane], bl = Bs[(2 * c + 1) * 64 + lane];
#pragma unroll
        for (int prod = 0; prod < 3; ++prod) {
          const uint4 av = (prod == 2) ? al : ah;
          const uint4 bv = (prod == 1) ? bl : bh;
          if constexpr (PREC == kBF16)
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                        __builtin_bit_cast(bf16x8, bv), d, 0, 0, 0);
          else
            d = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, av),
                                                       __builtin_bit_cast(f16x8, bv), d, 0, 0, 0);
        }
      }
    }
  } else if constexpr (PREC == kFP32) {
    uint4 b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = Bs[d * 64 + lane];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const uint4 a = As[c * 64 + lane];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float av = __uint_as_float(j == 0 ? a.x : j == 1 ? a.y : j == 2 ? a.z : a.w);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float bv =
              __uint_as_float(j == 0 ? b[d].x : j == 1 ? b[d].y : j == 2 ? b[d].z : b[d].w);
          acc[c * 3 + d] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc[c * 3 + d], 0, 0, 0);
        }
      }
    }
  } else {
    uint4 bh[3], bl[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      bh[d] = Bs[(2 * d) * 64 + lane];
      bl[d] = Bs[(2 * d + 1) * 64 + lane];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const uint4 ah = As[(2 * c) * 64 + lane];
      const uint4 al = As[(2 * c + 1) * 64 + lane];
#pragma unroll
      for (int prod = 0; prod < 3; ++prod) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const uint4 av = (prod == 2) ? al : ah;
          const uint4 bv = (prod == 1) ? bl[d] : bh[d];
          if constexpr (PREC == kBF16) {
            acc[c * 3 + d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                __builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv), acc[c * 3 + d], 0,
                0, 0);
          } else {
            acc[c * 3 + d] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                __builtin_bit_cast(f16x8, av), __builtin_bit_cast(f16x8, bv), acc[c * 3 + d], 0,
                0, 0);
          }
        }
      }
    }
  }
}

// One plane c of a split-form K-group: acc[c*3+d] += the three products of
// A_c (hi, lo) and B_d (hi, lo), d = 0..2.
template <int PREC>
__device__ __forceinline__ void mma_plane(f32x16 (&acc)[9], int c, const uint4 ah, const uint4 al,
                                          const uint4 (&bh)[3], const uint4 (&bl)[3]) {
#pragma unroll
  for (int prod = 0; prod < 3; ++prod) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const uint4 av = (prod == 2) ? al : ah;
      const uint4 bv = (prod == 1) ? bl[d] : bh[d];
      if constexpr (PREC == kBF16)
        acc[c * 3 + d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
            __builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv), acc[c * 3 + d], 0, 0, 0);
      else
        acc[c * 3 + d] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
            __builtin_bit_cast(f16x8, av), __builtin_bit_cast(f16x8, bv), acc[c * 3 + d], 0, 0, 0);
    }
  }
}

// One stage of the register-pipelined split-form K-loop.  On entry the
// stage's B fragments (bh, bl) and plane 0's A fragment (ah, al) are in
// registers (loaded during the previous stage); planes 1 and 2 of A are
// read from the current buffer (ca) one plane ahead of their MFMAs, and
// when NEXT the following stage's B fragments (into nbh, nbl) and plane-0
// A fragment (into ah, al) are read from the next buffer (na, nb) while this
// stage's MFMAs run.  MMA false: a padding stage (no MFMAs, no reads).
template <int PREC, bool MMA, bool NEXT>
__device__ __forceinline__ void pipe_stage(f32x16 (&acc)[9], const uint4* __restrict__ ca,
                                           const uint4* __restrict__ na,
                                           const uint4* __restrict__ nb, const uint4 (&bh)[3],
                                           const uint4 (&bl)[3], uint4 (&nbh)[3], uint4 (&nbl)[3],
                                           uint4& ah, uint4& al, int lane) {
  if constexpr (!MMA) return;
  const uint4 a1h = ca[2 * 64 + lane], a1l = ca[3 * 64 + lane];
  if constexpr (NEXT) {
    nbh[0] = nb[0 * 64 + lane];
    nbl[0] = nb[1 * 64 + lane];
  }
  mma_plane<PREC>(acc, 0, ah, al, bh, bl);
  const uint4 a2h = ca[4 * 64 + lane], a2l = ca[5 * 64 + lane];
  if constexpr (NEXT) {
    nbh[1] = nb[2 * 64 + lane];
    nbl[1] = nb[3 * 64 + lane];
  }
  mma_plane<PREC>(acc, 1, a1h, a1l, bh, bl);
  if constexpr (NEXT) {
    nbh[2] = nb[4 * 64 + lane];
    nbl[2] = nb[5 * 64 + lane];
    ah = na[0 * 64 + lane];
    al = na[1 * 64 + lane];
  }
  mma_plane<PREC>(acc, 2, a2h, a2l, bh, bl);
}

typedef __attribute__((address_space(3))) void lds_void_t;

// Wave `slot` copies its tile's K-group g (BPG blocks of 64 x 16 B) from HBM
// straight into LDS (global_load_lds_dwordx4: no VGPR staging).
template <int BPG>
__device__ __forceinline__ void dma_group(const uint4* __restrict__ src_tile, int g,
                                          uint4* dst_slot, int lane) {
  const uint4* src = src_tile + (int64_t)g * (BPG * 64) + lane;
#pragma unroll
  for (int j = 0; j < BPG; ++j)
    __builtin_amdgcn_global_load_lds((const void*)(src + j * 64), (lds_void_t*)(dst_slot + j * 64),
                                     16, 0, 0);
}

// Number of keys below `t` over the wave's PER x 64 slots.
template <int PER>
__device__ __forceinline__ int wave_count_below(const unsigned (&key)[PER], unsigned t) {
  int c = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) c += __popcll(__ballot(key[j] < t));
  return c;
}

// Smallest value t of the form (prefix << low) | ((1 << low) - 1) such that
// more than `rank` keys are <= t: with low = 0 the rank-th smallest key
// (0-based) exactly, with low > 0 an upper bound within 2^low units of it.
// Bitwise search from the top bit, one ballot round per bit.
template <int PER>
__device__ __forceinline__ unsigned wave_select(const unsigned (&key)[PER], int rank, int low) {
  unsigned t = 0;
  for (int bit = 31; bit >= low; --bit) {
    const unsigned cand = t | (1u << bit);
    if (wave_count_below<PER>(key, cand) <= rank) t = cand;
  }
  return low ? (t | ((1u << low) - 1u)) : t;
}

// Compact the lists of every lane whose list is nearly full, or holds 2 k1
// entries while the lane has no threshold yet (a query without a bound from
// elsewhere: the first compaction sets one early, so the row's other lists
// stop taking every pair).  The whole wave
// takes each such list in turn: it finds (an upper bound, within 2^-7
// relative, of) the list's k1-th smallest key tau, keeps the entries
// <= tau + 2 eps -- at most cap-64 of them; when more qualify, only those
// below the (cap-64)-th smallest key d stay and d is remembered in `ovf` --
// and lowers the lane's threshold to the cut.  Float keys are non-negative,
// so their bit patterns order like the values.  Kept entries are written
// back unsorted (the merge selects, it does not need order).
template <int PER>
__device__ __attribute__((noinline)) void compact_lists(uint2* __restrict__ ent, int64_t my_base, int& cnt,
                                              float& thr, float& ovf, float eps2, int k1,
                                              int cap, bool active) {
  const int lane = threadIdx.x & 63;
  uint64_t need = __ballot(active && (cnt > cap - 16 || (cnt >= 2 * k1 && thr == INFINITY)));
  while (need) {
    const int L = __builtin_ctzll(need);
    need &= need - 1;
    const int64_t base = __shfl(my_base, L);
    const int n = __shfl(cnt, L);
    const float e2 = __shfl(eps2, L);
    unsigned key[PER];
    unsigned val[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int e = lane * PER + j;
      if (e < n) {
        const uint2 x = ent[base + e];
        key[j] = x.x;
        val[j] = x.y;
      } else {
        key[j] = kInfBits;
        val[j] = 0u;
      }
    }
    float cut = INFINITY;
    if (n >= k1) {
      const float tau = __uint_as_float(wave_select<PER>(key, k1 - 1, 16));
      cut = tau * (1.0f + 2.4e-7f) + e2;
    }
    unsigned limit = __float_as_uint(cut);  // keep key <= limit
    const int kcap = cap - 64;
    float dropped = INFINITY;
    int keep = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) keep += __popcll(__ballot(key[j] <= limit && key[j] != kInfBits));
    if (keep > kcap) {
      const unsigned d = wave_select<PER>(key, kcap, 0);  // the kcap-th smallest, exactly
      dropped = __uint_as_float(d);
      // keep key < d; with d == 0 (more than kcap keys are exactly 0) keep
      // the zeros and let the slot cap below drop all but kcap of them
      limit = d == 0u ? 0u : d - 1u;
    }
    // Stream compaction in slot order e = lane*PER + j.
    const uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int pos = 0;
    bool kp[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      kp[j] = key[j] <= limit && key[j] != kInfBits;
      pos += __popcll(__ballot(kp[j]) & lanes_below);
    }
    // At most kcap entries survive (only the d == 0 case has more to drop):
    // the list then has room for the next tile's 16 appends per lane.
    int wrote = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (kp[j]) {
        if (pos + wrote < kcap) ent[base + pos + wrote] = make_uint2(key[j], val[j]);
        ++wrote;
      }
    const int total = min(wave_sum_i(wrote), kcap);
    if (lane == L) {
      cnt = total;
      thr = fminf(thr, cut);
      if (dropped < INFINITY) {
        ovf = fminf(ovf, dropped);
        thr = fminf(thr, dropped);
      }
    }
  }
}

// End of one pipeline step: this wave's copies of the stage to be read next
// have landed (at most N newer copies -- the stage just issued -- stay in
// flight), its LDS reads are done, then the workgroup barrier.  The empty
// asm statements keep the compiler from moving memory accesses across the
// barrier (s_barrier itself is not a memory fence).
template <int N>
__device__ __forceinline__ void stage_barrier() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  asm volatile("" ::: "memory");
  // s_waitcnt: vmcnt = N (bits 3:0 and 15:14), expcnt = 7 (no wait),
  // lgkmcnt = 0
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4));
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// Append this lane's mirrored score (pass: the key is below the reference
// frame's threshold) to the reference frame's list.  For a fixed accumulator
// element, the lanes of a half-wave hold the same reference frame (row) and
// 32 different fit frames, so each half-wave reserves its slots with one
// atomic add on that row's counter and its passing lanes write in lane
// order.  b: the wave's ballot of pass.
__device__ __forceinline__ void mirror_append(uint2* __restrict__ ment, int* __restrict__ mcnt,
                                              float* __restrict__ movf, int mcap, int64_t row,
                                              bool pass, uint64_t b, float key,
                                              unsigned other, int lane) {
  const int h = lane >> 5;
  const unsigned hb = h ? (unsigned)(b >> 32) : (unsigned)b;
  const int lead = hb ? (int)__builtin_ctz(hb) + 32 * h : lane;
  int base = 0;
  if (hb && lane == lead) base = atomicAdd(&mcnt[row], __builtin_popcount(hb));
  base = __shfl(base, lead);
  if (pass) {
    const int pos = base + __builtin_popcount(hb & ((1u << (lane & 31)) - 1u));
    if (pos < mcap)
      ment[row * (int64_t)mcap + pos] = make_uint2(__float_as_uint(key), other);
    else
      atomicMin(reinterpret_cast<unsigned*>(&movf[row]), __float_as_uint(key));
  }
}

// The pipelined K-loop's stage end: this wave's copies older than the N
// most recent have landed, then the workgroup barrier.  LDS reads stay in
// flight (they read the next stage's buffer, which no copy touches before
// the following barrier; each is waited for before its MFMAs).
template <int N>
__device__ __forceinline__ void stage_barrier_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  asm volatile("" ::: "memory");
  // vmcnt = N (bits 3:0 and 15:14), expcnt = 7, lgkmcnt = 15 (no wait)
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// FT = fit tiles per workgroup (2 or 4); the workgroup has 2*FT waves, one
// per (reference tile, fit tile) pair.  MODE: kOutLists, kOutDebug or
// kOutSym.
template <int PREC, int PER, int FT, int MODE, bool TRACE>
__global__ __launch_bounds__(64 * kRefTilesPerWG * FT, 2) void contract_kernel(ContractParams P) {
  constexpr int BPG = (PREC == kFP32) ? 3 : 6;
  constexpr int TILE_VEC = BPG * 64;
  constexpr int STAGE_VEC = (kRefTilesPerWG + FT) * TILE_VEC;
  // PIPE (the split forms' full product): a register-pipelined K-loop over
  // four stage buffers -- stage g computes from fragments read during stage
  // g-1, reads stage g+1's, and copies stage g+3.  Otherwise three buffers
  // (two K-groups in flight while one is consumed), each stage read just
  // before its MFMAs.
  constexpr bool PIPE = PREC != kFP32 && !TRACE;
  // Stage buffers, declared as distinct LDS objects so that the compiler does
  // not serialise reads of one behind copies into another.  Stage layout:
  // [ref tile 0][ref tile 1][fit tile 0] ... [fit tile FT-1].
  __shared__ uint4 sbuf0[STAGE_VEC];
  __shared__ uint4 sbuf1[STAGE_VEC];
  __shared__ uint4 sbuf2[STAGE_VEC];
  __shared__ uint4 sbuf3[PIPE ? STAGE_VEC : 1];
  __shared__ double sGr[2][2 * kTile];  // G~ of the current / next reference tile pair
  // kOutSym: shared thresholds of the current / next reference tile pair's
  // frames (the filter of their mirrored scores)
  __shared__ float sThr[2][2 * kTile];

  const int tid = threadIdx.x;
  // wave-uniform, so that everything derived from it (tile roles, copy
  // sources, LDS slots) lives in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int ri = wave & 1;   // reference tile of this wave
  const int fi = wave >> 1;  // fit tile of this wave
  const int ng = P.n_groups;
  // Stages per tile, a multiple of the buffer count so that the buffer
  // rotation lines up with tile boundaries; stages g >= ng copy padding and
  // skip the MFMAs.
  const int ngs = PIPE ? ng : (ng + 2) / 3 * 3;

  // Waves 0..kSlots-1 each copy one stage slot (slot = wave); the others
  // only compute.
  constexpr int kSlots = kRefTilesPerWG + FT;
  const bool copier = wave < kSlots;
  uint4* const d0 = sbuf0 + (copier ? wave : 0) * TILE_VEC;
  uint4* const d1 = sbuf1 + (copier ? wave : 0) * TILE_VEC;
  uint4* const d2 = sbuf2 + (copier ? wave : 0) * TILE_VEC;
  const uint4* const a0 = sbuf0 + ri * TILE_VEC;
  const uint4* const a1 = sbuf1 + ri * TILE_VEC;
  const uint4* const a2 = sbuf2 + ri * TILE_VEC;
  const uint4* const b0 = sbuf0 + (kRefTilesPerWG + fi) * TILE_VEC;
  const uint4* const b1 = sbuf1 + (kRefTilesPerWG + fi) * TILE_VEC;
  const uint4* const b2 = sbuf2 + (kRefTilesPerWG + fi) * TILE_VEC;
  uint4* const d3 = sbuf3 + (PIPE && copier ? wave : 0) * TILE_VEC;
  const uint4* const a3 = sbuf3 + (PIPE ? ri : 0) * TILE_VEC;
  const uint4* const b3 = sbuf3 + (PIPE ? kRefTilesPerWG + fi : 0) * TILE_VEC;
  const int fit_blocks = P.fit_tiles / FT;

  // Workgroups are dispatched round-robin over the 8 XCDs; renumber them so
  // that consecutive work items (same fit block, neighbouring reference
  // phases) land on the same XCD and share its L2.
  const int nwg = gridDim.x;
  const int logical = (int)(blockIdx.x & 7) * (nwg >> 3) + (int)(blockIdx.x >> 3);
  const int n_items = fit_blocks * P.ph_run;

  // The shared threshold of reference frame r (kOutSym), or +inf past the set.
  auto ref_thr = [&](int64_t r) -> float {
    return r < P.n_ref ? __uint_as_float(P.thr[r]) : INFINITY;
  };

  const int rounds = (n_items + nwg - 1) / nwg;
  for (int round = 0; round < rounds; ++round) {
    const int item = logical + round * nwg;
    if (item >= n_items) break;
    const int fb = item / P.ph_run;
    const int ph = item - fb * P.ph_run;
    const int n_tiles = (P.ref_pairs - ph + P.phases - 1) / P.phases;  // ref pairs ph, ph+R, ...
    // First tile of the item: 0, or for kOutSym the first reference pair at
    // or after the fit block (pair p holds frames 64p.., the block 32*FT*fb..).
    int t0 = 0;
    if constexpr (MODE == kOutSym) {
      const int p_min = fb * FT / kRefTilesPerWG;
      if (p_min > ph) t0 = (p_min - ph + P.phases - 1) / P.phases;
    }
    if (t0 >= n_tiles) continue;
    const int64_t fit_idx = (int64_t)(FT * fb + fi) * kTile + (lane & 31);
    const bool fvalid = fit_idx < P.n_fit;
    const int list_id = (ph * 2 + ri) * 2 + (lane >> 5);
    const int64_t hdr = fit_idx * P.nl + list_id;
    const int64_t my_base = hdr * P.cap;
    const double gf = P.fGt[fit_idx];
    const float e2 = (fvalid && P.eps2) ? P.eps2[fit_idx] : 0.f;
    int cnt = 0;
    float thr = INFINITY, ovf = INFINITY;

    // Stage s of this item is K-group g = s % ngs of tile t = t0 + s / ngs.  Wave
    // w < 2 copies ref tile 2p+w (p = ph + t*R), wave 2..FT+1 fit tile
    // FT*fb+w-2.  Stages past the item's last tile re-read its last tile
    // (never used).
    const uint4* const fit_base =
        P.fop + (int64_t)(FT * fb + (copier ? wave : 2) - kRefTilesPerWG) * ng * TILE_VEC;
    auto src_of = [&](int t, int g) -> const uint4* {
      if (t >= n_tiles) t = n_tiles - 1;
      if (wave >= kRefTilesPerWG) return fit_base + (int64_t)g * TILE_VEC;
      const int p = ph + t * P.phases;
      return P.rop + ((int64_t)(kRefTilesPerWG * p + wave) * ng + g) * TILE_VEC;
    };
    auto dma_stage = [&](int t, int g, uint4* dst) {
      if (!copier) return;
      if (g >= ngs) { g -= ngs; ++t; }
      const uint4* src = src_of(t, g) + lane;
#pragma unroll
      for (int j = 0; j < BPG; ++j)
        __builtin_amdgcn_global_load_lds((const void*)(src + j * 64),
                                         (lds_void_t*)(dst + j * 64), 16, 0, 0);
    };

    __syncthreads();  // previous item's readers of sbuf*/sGr/sThr are done
    if (tid < 2 * kTile) {
      const int64_t f0 = (int64_t)(ph + t0 * P.phases) * 2 * kTile + tid;
      sGr[t0 & 1][tid] = P.rGt[f0];
      if constexpr (MODE == kOutSym) sThr[t0 & 1][tid] = ref_thr(f0);
    }
    dma_stage(t0, 0, d0);
    dma_stage(t0, 1, d1);
    if constexpr (PIPE) dma_stage(t0, 2, d2);
    stage_barrier<BPG>();  // stage 0 (and with PIPE stage 1) landed
    // PIPE: the fragments of the stage about to be computed (B double-buffered)
    uint4 bh0[3], bl0[3], bh1[3], bl1[3], ah, al;
    if constexpr (PIPE) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        bh0[d] = b0[(2 * d) * 64 + lane];
        bl0[d] = b0[(2 * d + 1) * 64 + lane];
      }
      ah = a0[0 * 64 + lane];
      al = a0[1 * 64 + lane];
    }

    for (int t = t0; t < n_tiles; ++t) {
      const int p = ph + t * P.phases;
      // Loads consumed only at the end of the tile: the next tile's G~ (and
      // for kOutSym its frames' thresholds) and the shared per-query
      // threshold.
      double g_next = 0.0;
      float thr_next = INFINITY;
      if (tid < 2 * kTile && t + 1 < n_tiles) {
        const int64_t f1 = (int64_t)(p + P.phases) * 2 * kTile + tid;
        g_next = P.rGt[f1];
        if constexpr (MODE == kOutSym) thr_next = ref_thr(f1);
      }
      unsigned thr_shared = kInfBits;
      if (fvalid && P.thr) thr_shared = __builtin_nontemporal_load(&P.thr[fit_idx]);

      f32x16 acc[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) acc[q] = (f32x16){0};
      if constexpr (PIPE) {
        // Stage g lives in buffer g % 4 (ng is a multiple of 4 here: the
        // operand is padded with zero atoms).  At the start of stage g its
        // fragments are in registers (read during stage g-1) and stage g+1
        // has landed; stage g reads stage g+1's fragments while its MFMAs
        // run, and copies stage g+3 into the buffer of stage g-1, which every
        // wave finished reading before the previous barrier.  The tile's last
        // stage reads the next tile's first one, so the pipeline runs through
        // the epilogue (past the item's last tile it reads a re-copied tile,
        // never used).
        for (int g = 0; g < ng; g += 4) {
          pipe_stage<PREC, true, true>(acc, a0, a1, b1, bh0, bl0, bh1, bl1, ah, al, lane);
          dma_stage(t, g + 3, d3);
          stage_barrier_vm<BPG>();
          pipe_stage<PREC, true, true>(acc, a1, a2, b2, bh1, bl1, bh0, bl0, ah, al, lane);
          dma_stage(t, g + 4, d0);
          stage_barrier_vm<BPG>();
          pipe_stage<PREC, true, true>(acc, a2, a3, b3, bh0, bl0, bh1, bl1, ah, al, lane);
          dma_stage(t, g + 5, d1);
          stage_barrier_vm<BPG>();
          pipe_stage<PREC, true, true>(acc, a3, a0, b0, bh1, bl1, bh0, bl0, ah, al, lane);
          dma_stage(t, g + 6, d2);
          stage_barrier_vm<BPG>();
        }
      } else {
        // Stage g lives in buffer g % 3; while it is consumed, stage g+2 (in
        // this tile or the next) is being copied.
        // The copy of stage g+2 is issued after stage g's MFMAs: the MFMAs
        // start right after the barrier, and the copy (into the buffer every
        // wave finished reading before that barrier) overlaps their tail.
        // Measured 6 % faster than issuing the copy first.
        for (int g = 0; g < ngs; g += 3) {
          mma_group<PREC, TRACE>(acc, a0, b0, lane);
          dma_stage(t, g + 2, d2);
          stage_barrier<BPG>();
          if (g + 1 < ng) mma_group<PREC, TRACE>(acc, a1, b1, lane);
          dma_stage(t, g + 3, d0);
          stage_barrier<BPG>();
          if (g + 2 < ng) mma_group<PREC, TRACE>(acc, a2, b2, lane);
          dma_stage(t, g + 4, d1);
          stage_barrier<BPG>();
        }
      }
      // Publish the next tile's G~ and thresholds (read by its epilogue,
      // after at least one more barrier).
      if (tid < 2 * kTile && t + 1 < n_tiles) {
        sGr[(t + 1) & 1][tid] = g_next;
        if constexpr (MODE == kOutSym) sThr[(t + 1) & 1][tid] = thr_next;
      }
      if (fvalid && P.thr) thr = fminf(thr, __uint_as_float(thr_shared));

      // Epilogue: lane holds fit frame (lane%32) against 16 ref frames.
      const double* const gr_tile = sGr[t & 1];
      const int64_t ref_base = (int64_t)(kRefTilesPerWG * p + ri) * kTile;
      // kOutSym: a reference pair wholly past the fit block is contracted
      // only this way round, so its scores also go to the reference frames'
      // lists.
      bool mirror = false;
      if constexpr (MODE == kOutSym) {
        // opaque to the optimiser: the condition is monotone in t, and
        // splitting the tile loop at its switch would duplicate the whole
        // K-loop and epilogue (instruction-cache pressure)
        int mir = p >= (fb + 1) * FT / kRefTilesPerWG;
        asm volatile("" : "+v"(mir));
        mirror = __builtin_amdgcn_readfirstlane(mir) != 0;
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int rl = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int64_t ref_idx = ref_base + rl;
        const bool valid = fvalid && ref_idx < P.n_ref;
        float m[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = acc[q][r];
        const double gr = gr_tile[ri * kTile + rl];
        if constexpr (MODE == kOutDebug) {
          if (valid)
            P.dbg[fit_idx * P.n_ref + ref_idx] = qcp_score_approx(m, gr, gf, INFINITY, P.nofit);
          continue;
        }
        // Newton stops once the pair is out of every list it could join.
        // tr: the reference frame's threshold, for a tile whose scores are
        // mirrored; otherwise a value below every threshold, which only
        // feeds the rejection bound (an approximate score can be negative,
        // so the mirrored append tests `mirror` itself).
        float tr = -1.0f;
        if constexpr (MODE == kOutSym)
          if (mirror) tr = sThr[t & 1][ri * kTile + rl];
        bool own = false, mir = false;
        float key = 0.f;
        if (valid) {
          const double rej = fmax((double)thr, (double)tr);
          // A pair whose score is above every list it could join, by the
          // bound on lambda_max, skips the QCP: the QCP would reject it too.
          // Its Newton iterate approaches the score from below, to within
          // far less than 1e-9 (G_r + G_f) (the tolerance of its last step and
          // its rounding), which the test allows for.  (The trace score of
          // --nofit has no QCP to skip.)
          const bool far =
              !P.nofit && gr + gf - 2.0 * (double)lambda_upper(m) - 1e-9 * (gr + gf) > rej;
          if (!far) {
            const double s = qcp_score_approx(m, gr, gf, rej, P.nofit);
            key = (float)fmax(s, 0.0);
            own = s <= (double)thr;
            mir = mirror && s <= (double)tr;
          }
        }
        if (own) {
          // The compaction keeps cnt <= cap - 16 between tiles, so the guard
          // never drops; if it did, ovf would send the row to the exact path.
          if (cnt < P.cap) {
            P.ent[my_base + cnt] = make_uint2(__float_as_uint(key), (unsigned)ref_idx);
            ++cnt;
          } else {
            ovf = fminf(ovf, key);
          }
        }
        if constexpr (MODE == kOutSym) {
          const uint64_t b = __ballot(mir);
          if (b)
            mirror_append(P.ment, P.mcnt, P.movf, P.mcap, ref_idx, mir, b, key,
                          (unsigned)fit_idx, lane);
        }
      }
      if constexpr (MODE == kOutDebug) continue;
      const float thr_before = thr;
      compact_lists<PER>(P.ent, my_base, cnt, thr, ovf, e2, P.k1, P.cap, fvalid);
      thr = fminf(thr, __shfl_xor(thr, 32));
      if (fvalid && P.thr && thr < thr_before) atomicMin(&P.thr[fit_idx], __float_as_uint(thr));
    }
    if (fvalid && MODE != kOutDebug) {
      P.cnt[hdr] = cnt;
      P.ovf[hdr] = ovf;
    }
  }
}

// ---------------------------------------------------------------------------
// exact FP64 score of one pair, by a 16-lane group (lanes over atoms).

// A merge workgroup's query frame, staged once in LDS as its weighted
// centred coordinates w_i (b_i - c_b) in FP64 (3 per atom).
extern __shared__ double qw_stage[];

__device__ __forceinline__ double group_sum16(double v) {
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// The weighted centred query coordinate w (x - c), formed the same way
// wherever it is needed (staged or on the fly), so both give the same bits.
__device__ __forceinline__ double weighted_centred(float x, double c, double w) {
  return w * ((double)x - c);
}

// Exact FP64 score of the pair (reference frame a, query frame b) by the
// 16 lanes of a group, lane j of the group taking atoms j, j + 16, ...:
//   S = sum_i a_i (w_i (b_i - c_b))^T
// with the reference frame uncentred (sum_i w_i (b_i - c_b) = 0, so S is
// the weighted cross-covariance of the centred frames), then lambda_max of
// the key matrix and s = G_a + G_b - 2 lambda_max.  Every lane of the group
// returns the same value.  QS: the query's weighted centred coordinates
// are in qw_stage (the pointer b is then unused).
template <bool QS>
__device__ __forceinline__ double group_rmsd_score(const float* __restrict__ a, double ga,
                                                   const float* __restrict__ b,
                                                   const double* __restrict__ cb, double gb,
                                                   const float* __restrict__ w, int natoms,
                                                   bool nofit) {
  const int sub = threadIdx.x & 15;
  double S[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) S[i] = 0.0;
  const double cb0 = cb[0], cb1 = cb[1], cb2 = cb[2];
  const F3* __restrict__ a3 = reinterpret_cast<const F3*>(a);
  const F3* __restrict__ b3 = reinterpret_cast<const F3*>(b);
#pragma unroll 8
  for (int i = sub; i < natoms; i += 16) {
    const F3 av = a3[i];
    double bx, by, bz;
    if constexpr (QS) {
      bx = qw_stage[3 * i];
      by = qw_stage[3 * i + 1];
      bz = qw_stage[3 * i + 2];
    } else {
      const F3 bv = b3[i];
      const double wi = (double)w[i];
      bx = weighted_centred(bv.x, cb0, wi);
      by = weighted_centred(bv.y, cb1, wi);
      bz = weighted_centred(bv.z, cb2, wi);
    }
    const double ax = (double)av.x, ay = (double)av.y, az = (double)av.z;
    S[0] = fma(ax, bx, S[0]); S[1] = fma(ax, by, S[1]); S[2] = fma(ax, bz, S[2]);
    S[3] = fma(ay, bx, S[3]); S[4] = fma(ay, by, S[4]); S[5] = fma(ay, bz, S[5]);
    S[6] = fma(az, bx, S[6]); S[7] = fma(az, by, S[7]); S[8] = fma(az, bz, S[8]);
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) S[i] = group_sum16(S[i]);
  const double lam = nofit ? (S[0] + S[4] + S[8]) : qcp_lambda_exact(S, ga, gb);
  return fmax(ga + gb - 2.0 * lam, 0.0);
}

// Exact FP64 score of one pair of data vectors (a: reference, b: query),
// lanes over elements: the squared Euclidean distance
// (euclidean_distance, mdsctk.cpp:329-334), or for the correlation metric
// s = 4 max(v, 0) with v = (1 - cov / ((V-1) sd_b sd_a)) / 2 as
// correlation_distance forms it (mdsctk.cpp:335-356; its first argument is
// the query vector).
// The output distance is then sqrt(s) or sqrt(s / 4) = sqrt(v).
__device__ __forceinline__ double wave_data_score(const double* __restrict__ a,
                                                  const double* __restrict__ a_aux,
                                                  const double* __restrict__ b,
                                                  const double* __restrict__ b_aux, int vdim,
                                                  int metric) {
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  if (metric == kMetricEuclidean) {
    for (int e = lane; e < vdim; e += 64) {
      const double d = b[e] - a[e];
      acc += d * d;
    }
    return wave_sum_d(acc);
  }
  const double ma = a_aux[0], mb = b_aux[0];
  for (int e = lane; e < vdim; e += 64) acc += (b[e] - mb) * (a[e] - ma);
  acc = wave_sum_d(acc);
  const double v = (1.0 - (acc / ((((double)vdim - 1.0) * b_aux[1]) * a_aux[1]))) / 2.0;
  // clamped as the reference clamps v: scores are never negative, which the
  // exact-row selection's 64-bit key order relies on
  return fmax(4.0 * v, 0.0);
}

// The exact score of reference r against query q under the set's metric.
struct PairSource {
  const float* rxyz;
  const double* rcen;
  const double* rG;
  const float* fxyz;
  const double* fcen;
  const double* fG;
  const float* w;
  int natoms;
  int metric, vdim;
  const double* r64;
  const double* raux;
  const double* f64;
  const double* faux;
};

// RMSD: the exact score of reference r against query q by a 16-lane group.
template <bool QS = false>
__device__ __forceinline__ double group_pair_score(const PairSource& P, int64_t r, int64_t q,
                                                   bool nofit) {
  return group_rmsd_score<QS>(P.rxyz + r * (int64_t)P.natoms * 3, P.rG[r],
                              P.fxyz + q * (int64_t)P.natoms * 3, P.fcen + q * 3, P.fG[q], P.w,
                              P.natoms, nofit);
}

// Data vectors: the exact score of reference r against query q by a wave.
__device__ __forceinline__ double wave_data_pair_score(const PairSource& P, int64_t r, int64_t q) {
  return wave_data_score(P.r64 + r * (int64_t)P.vdim, P.raux + r * 2, P.f64 + q * (int64_t)P.vdim,
                         P.faux + q * 2, P.vdim, P.metric);
}

// ---------------------------------------------------------------------------
// merge: radix select over the lists, certified window, FP64 re-score.

struct MergeParams {
  PairSource src;
  const int* order;  // query rows in processing order (nullptr: 0..n_fit-1)
  // the lists were built over a permutation of the frames (the symmetric
  // self query in landmark order): list row i and list entry index j are
  // frames perm[i] and perm[j]; the frames' data and the outputs use the
  // original indices.  Replaces order (the lists are in processing order).
  const int* perm;
  double W, out_scale;
  const uint2* ent;
  const int* cnt;
  const float* ovf;
  int nl, cap;
  // a symmetric self query's second list per row (the scores the row
  // received as the reference side of a pair); null otherwise
  const uint2* ment;
  const int* mcnt;
  const float* movf;
  int mcap;
  const float* eps2;
  int k1, k_out, rank_off;
  int nofit;
  double* out_dist;
  int* out_idx;
  int* fallback;
  int* n_cand;
};

__device__ void block_bitonic_sort(double* s, int* idx, int n_pow2) {
  const int tid = threadIdx.x;
  for (int size = 2; size <= n_pow2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (int i = tid; i < n_pow2 / 2; i += blockDim.x) {
        const int lo = 2 * stride * (i / stride) + (i % stride);
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const bool gt = (s[lo] > s[hi]) || (s[lo] == s[hi] && idx[lo] > idx[hi]);
        if (gt == up) {
          const double ts = s[lo]; s[lo] = s[hi]; s[hi] = ts;
          const int ti = idx[lo]; idx[lo] = idx[hi]; idx[hi] = ti;
        }
      }
    }
  }
  __syncthreads();
}

// Keys just above a row's cut that the merge re-scores to check the window.
constexpr int kMergeGuard = 32;

// STAGE: for frames of up to kMergeStageAtoms atoms the query frame's
// weighted centred coordinates are formed once, in FP64 in LDS (24 KB at
// 1024 atoms), so each candidate reads only its reference frame from memory
// and spends 3 conversions and 9 FMAs per atom.
constexpr int kMergeStageAtoms = 1024;

// CAND: the most candidates a row's window may hold (more: exact path).
// The smaller instance leaves room for more workgroups per CU; the
// waves-per-EU hint keeps it within 96 VGPRs (5 workgroups per CU).
template <bool STAGE, int CAND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void merge_kernel(MergeParams P) {
  const int64_t qi = P.perm ? (int64_t)blockIdx.x : (P.order ? P.order[blockIdx.x] : blockIdx.x);
  const int64_t q = P.perm ? P.perm[qi] : qi;  // the frame: data and output row
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  __shared__ unsigned hist[256];
  __shared__ unsigned s_prefix;
  __shared__ int s_kk, s_nc, s_flag, s_total;
  __shared__ int cand_idx[CAND];
  __shared__ double cand_s[CAND];
  __shared__ float cand_key[CAND];
  __shared__ int s_bad;
  // the guard: keys just above the cut, re-scored to check the window
  __shared__ int guard_idx[kMergeGuard];
  __shared__ float guard_key[kMergeGuard];
  __shared__ double guard_s[kMergeGuard];
  __shared__ int s_ng;

  const int* cnt = P.cnt + qi * P.nl;
  const uint2* ent = P.ent + qi * P.nl * (int64_t)P.cap;
  const float* ovf = P.ovf + qi * P.nl;
  const int mc = P.ment ? min(P.mcnt[qi], P.mcap) : 0;
  const uint2* ment = P.ment ? P.ment + qi * (int64_t)P.mcap : nullptr;
  auto frame_of = [&](unsigned j) -> int { return P.perm ? P.perm[j] : (int)j; };
  // f(entry) for every entry of the row's lists, and the whole workgroup
  // over the mirrored list.  Few lists: waves over the lists, lanes over
  // entries (coalesced).  Many lists (a small query batch spread over many
  // reference phases): a thread per list, streaming its entries with four
  // loads in flight -- a wave per list would wait on one list's count and
  // entries at a time, hundreds of times per pass.
  const bool list_per_thread = P.nl >= 256;
  auto visit = [&](auto&& f) {
    if (list_per_thread) {
      for (int l = tid; l < P.nl; l += 256) {
        const int c = cnt[l];
        const uint2* el = ent + (int64_t)l * P.cap;
        int i = 0;
        for (; i + 4 <= c; i += 4) {
          const uint2 x0 = el[i], x1 = el[i + 1], x2 = el[i + 2], x3 = el[i + 3];
          f(x0);
          f(x1);
          f(x2);
          f(x3);
        }
        for (; i < c; ++i) f(el[i]);
      }
    } else {
      for (int l = wave; l < P.nl; l += 4) {
        const int c = cnt[l];
        for (int i = lane; i < c; i += 64) f(ent[(int64_t)l * P.cap + i]);
      }
    }
    for (int i = tid; i < mc; i += 256) f(ment[i]);
  };

  if (tid == 0) { s_total = mc; s_nc = 0; s_flag = 0; s_bad = 0; s_ng = 0; }
  __syncthreads();
  {
    int part = 0;
    for (int l = tid; l < P.nl; l += 256) part += cnt[l];
    atomicAdd(&s_total, part);
  }
  __syncthreads();
  if (s_total < P.k1) {
    if (tid == 0) P.fallback[q] = 1;
    return;
  }

  unsigned prefix = 0;
  int kk = P.k1 - 1;
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[tid] = 0;
    __syncthreads();
    const unsigned mhi = (shift == 24) ? 0u : (0xFFFFFFFFu << (shift + 8));
    visit([&](const uint2 e) {
      const unsigned k = e.x;
      if ((k & mhi) == (prefix & mhi)) atomicAdd(&hist[(k >> shift) & 255u], 1u);
    });
    __syncthreads();
    if (tid == 0) {
      unsigned cum = 0;
      int b = 0;
      for (; b < 255; ++b) {
        if (cum + hist[b] > (unsigned)kk) break;
        cum += hist[b];
      }
      s_prefix = prefix | ((unsigned)b << shift);
      s_kk = kk - (int)cum;
    }
    __syncthreads();
    prefix = s_prefix;
    kk = s_kk;
  }
  const float tau = __uint_as_float(prefix);
  const float e2q = P.eps2[qi];
  const float cut = tau * (1.0f + 2.4e-7f) + e2q;
  // The guard range: one more window width above the cut.  A window too
  // narrow for its row would drop pairs whose keys lie just above the cut,
  // so up to kMergeGuard of those are re-scored as well.
  const float guard_cut = cut + e2q;
  for (int l = tid; l < P.nl; l += 256)
    if (ovf[l] <= cut) s_flag = 1;
  if (tid == 0 && P.ment && P.movf[qi] <= cut) s_flag = 1;
  visit([&](const uint2 e) {
    const float key = __uint_as_float(e.x);
    if (key <= cut) {
      const int pos = atomicAdd(&s_nc, 1);
      if (pos < CAND) {
        cand_idx[pos] = frame_of(e.y);
        cand_key[pos] = key;
      }
    } else if (key <= guard_cut) {
      const int pos = atomicAdd(&s_ng, 1);
      if (pos < kMergeGuard) {
        guard_idx[pos] = frame_of(e.y);
        guard_key[pos] = key;
      }
    }
  });
  __syncthreads();
  const int nc = s_nc;
  const int ng = min(s_ng, kMergeGuard);
  if (s_flag || nc > CAND) {
    if (tid == 0) P.fallback[q] = 1;
    return;
  }
  if (tid == 0) {
    P.fallback[q] = 0;
    atomicAdd(P.n_cand, nc);
  }

  if constexpr (STAGE) {
    const F3* fx = reinterpret_cast<const F3*>(P.src.fxyz) + q * (int64_t)P.src.natoms;
    const double c0 = P.src.fcen[q * 3], c1 = P.src.fcen[q * 3 + 1], c2 = P.src.fcen[q * 3 + 2];
    for (int i = tid; i < P.src.natoms; i += 256) {
      const F3 b = fx[i];
      const double wi = (double)P.src.w[i];
      qw_stage[3 * i] = weighted_centred(b.x, c0, wi);
      qw_stage[3 * i + 1] = weighted_centred(b.y, c1, wi);
      qw_stage[3 * i + 2] = weighted_centred(b.z, c2, wi);
    }
    __syncthreads();
  }
  auto put = [&](int c, double s) {
    if (c < nc)
      cand_s[c] = s;
    else
      guard_s[c - nc] = s;
  };
  if (P.src.metric == kMetricRMSD) {
    // four pairs per wave, one per 16-lane group
    const int grp = lane >> 4;
    for (int c0 = 4 * wave; c0 < nc + ng; c0 += 16) {
      const int c = c0 + grp;
      if (c < nc + ng) {
        const int r = c < nc ? cand_idx[c] : guard_idx[c - nc];
        const double s = group_pair_score<STAGE>(P.src, r, q, P.nofit);
        if ((lane & 15) == 0) put(c, s);
      }
    }
  } else {
    for (int c = wave; c < nc + ng; c += 4) {
      const int r = c < nc ? cand_idx[c] : guard_idx[c - nc];
      const double s = wave_data_pair_score(P.src, r, q);
      if (lane == 0) put(c, s);
    }
  }
  __syncthreads();
  // The window assumes that every approximate score lies within
  // eps = eps2 / 2 of the exact one.  The exact scores of the candidates and
  // of the guard keys are known now: a row where one deviates further takes
  // the exact path.  The keys are rounded to FP32, hence the relative
  // allowance.
  {
    const double eps = 0.5 * (double)e2q;
    for (int c = tid; c < nc + ng; c += 256) {
      const double s = c < nc ? cand_s[c] : guard_s[c - nc];
      const double kq = (double)(c < nc ? cand_key[c] : guard_key[c - nc]);
      if (fabs(kq - s) > eps + 1.2e-7 * fmax(kq, s)) s_bad = 1;
    }
  }
  __syncthreads();
  if (s_bad) {
    if (tid == 0) P.fallback[q] = 1;
    return;
  }
  int np2 = 64;
  while (np2 < nc) np2 <<= 1;
  for (int i = nc + tid; i < np2; i += 256) {
    cand_s[i] = INFINITY;
    cand_idx[i] = INT_MAX;
  }
  block_bitonic_sort(cand_s, cand_idx, np2);
  for (int j = tid; j < P.k_out; j += 256) {
    const int src = j + P.rank_off;
    P.out_dist[q * P.k_out + j] = P.out_scale * sqrt(fmax(cand_s[src], 0.0) / P.W);
    P.out_idx[q * P.k_out + j] = cand_idx[src];
  }
}

// ---------------------------------------------------------------------------
// Symmetric self query: thresholds from the landmark pass, merge order.

// One 256-thread workgroup per row q of the landmark pass (the contraction
// of every frame against every phases-th reference tile pair, kOutLists
// into nl lists per row).  Its k1-th smallest key tau_L bounds the row's
// k1-th smallest key of the full query: k1 landmark pairs have keys <= tau_L,
// so their exact scores are <= tau_L + eps and their keys in the full query
// (computed either way round) <= tau_L + 2 eps = tau_L + e2.  thr[q] becomes
// the window of that bound, (tau_L + e2) (1 + 2.4e-7) + e2 rounded up with
// slack; with fewer than k1 landmark keys it is +inf.  label[q]: the nearest
// landmark (smallest key, then smallest index) as a dense number
// 64 j + o for the landmark frame 64 (j R) + o (R = phases).
__global__ __launch_bounds__(256) void landmark_cap_kernel(const uint2* __restrict__ ent,
                                                           const int* __restrict__ cnt, int nl,
                                                           int cap, int k1,
                                                           const float* __restrict__ eps2,
                                                           int phases, unsigned* __restrict__ thr,
                                                           int* __restrict__ label) {
  const int64_t q = blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  __shared__ unsigned hist[256];
  __shared__ unsigned s_prefix;
  __shared__ int s_kk, s_total;
  __shared__ unsigned long long s_best;
  const int* c = cnt + q * nl;
  const uint2* e = ent + q * (int64_t)nl * cap;
  auto visit = [&](auto&& f) {
    for (int l = wave; l < nl; l += 4) {
      const int cl = c[l];
      for (int i = lane; i < cl; i += 64) f(e[(int64_t)l * cap + i]);
    }
  };
  if (tid == 0) {
    s_total = 0;
    s_best = ~0ull;
  }
  __syncthreads();
  {
    int part = 0;
    unsigned long long best = ~0ull;
    for (int l = tid; l < nl; l += 256) part += c[l];
    atomicAdd(&s_total, part);
    visit([&](const uint2 x) {
      const unsigned long long kx = ((unsigned long long)x.x << 32) | x.y;
      best = kx < best ? kx : best;
    });
    atomicMin(&s_best, best);
  }
  __syncthreads();
  if (tid == 0) {
    int lab = 0;
    if (s_best != ~0ull) {
      const unsigned r = (unsigned)(s_best & 0xffffffffu);
      const unsigned span = 2u * kTile * (unsigned)phases;
      lab = (int)((r / span) * (2u * kTile) + r % (2u * kTile));
    }
    label[q] = lab;
  }
  if (s_total < k1) {
    if (tid == 0) thr[q] = kInfBits;
    return;
  }
  unsigned prefix = 0;
  int kk = k1 - 1;
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[tid] = 0;
    __syncthreads();
    const unsigned mhi = (shift == 24) ? 0u : (0xFFFFFFFFu << (shift + 8));
    visit([&](const uint2 x) {
      if ((x.x & mhi) == (prefix & mhi)) atomicAdd(&hist[(x.x >> shift) & 255u], 1u);
    });
    __syncthreads();
    if (tid == 0) {
      unsigned cum = 0;
      int b = 0;
      for (; b < 255; ++b) {
        if (cum + hist[b] > (unsigned)kk) break;
        cum += hist[b];
      }
      s_prefix = prefix | ((unsigned)b << shift);
      s_kk = kk - (int)cum;
    }
    __syncthreads();
    prefix = s_prefix;
    kk = s_kk;
  }
  if (tid == 0) {
    const double tau = (double)__uint_as_float(prefix), e2 = (double)eps2[q];
    const float bound = (float)((tau + e2) * (1.0 + 1e-6) + e2);
    thr[q] = bound < INFINITY ? __float_as_uint(bound) + 1u : kInfBits;
  }
}

// Merge order of a symmetric self query: rows grouped by their nearest
// landmark (a counting sort over nlab labels), so that rows merged at the
// same time share candidate frames in L2 and MALL.  The order only decides
// which workgroup takes which row, never a result.
__global__ void order_count_kernel(const int* __restrict__ label, int64_t n,
                                   int* __restrict__ counts) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) atomicAdd(&counts[label[q]], 1);
}

// counts -> exclusive offsets in place; one block of 1024 threads, each
// over a run of consecutive labels.
__global__ __launch_bounds__(1024) void order_scan_kernel(int* __restrict__ counts, int nlab) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (nlab + 1023) / 1024;
  const int lo = min(t * per, nlab), hi = min(lo + per, nlab);
  int sum = 0;
  for (int j = lo; j < hi; ++j) sum += counts[j];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int acc = part[t] - sum;
  for (int j = lo; j < hi; ++j) {
    const int v = counts[j];
    counts[j] = acc;
    acc += v;
  }
}

__global__ void order_scatter_kernel(const int* __restrict__ label, int64_t n,
                                     int* __restrict__ offsets, int* __restrict__ order) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) order[atomicAdd(&offsets[label[q]], 1)] = (int)q;
}

__global__ void fill_int_kernel(int* __restrict__ p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// The mirrored lists' counters and overflow keys (set before the symmetric
// contraction) and, after it, their total fill.
__global__ void mirror_init_kernel(int* __restrict__ mcnt, float* __restrict__ movf, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    mcnt[i] = 0;
    movf[i] = INFINITY;
  }
}

__global__ void mirror_sum_kernel(const int* __restrict__ mcnt, int64_t n, int mcap,
                                  unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[256];
  unsigned long long s = 0, mx = 0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const unsigned long long v = (unsigned long long)mcnt[i];
    s += v;
    mx = v > mx ? v : mx;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
  __syncthreads();
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k)
      red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + k] ? red[threadIdx.x]
                                                                 : red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[1] = red[0];
  (void)mcap;
}

// Rows of a frame set gathered into a batch (rows[i] -> dst row i), and
// output rows scattered back (src row i -> dst row rows[i]).
__global__ void permute_thresholds_kernel(const unsigned* __restrict__ src,
                                          const int* __restrict__ perm,
                                          const float* __restrict__ add, int64_t n,
                                          unsigned* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float t = __uint_as_float(src[perm[i]]);
  const float v = t + add[i];
  dst[i] = v < INFINITY ? __float_as_uint(v) + 1u : kInfBits;
}

__global__ void gather_frames_kernel(const float* __restrict__ src, const int* __restrict__ rows,
                                     int natoms, float* __restrict__ dst) {
  const int64_t i = blockIdx.x;
  const int64_t w = (int64_t)natoms * 3;
  const float* s = src + (int64_t)rows[i] * w;
  float* d = dst + i * w;
  for (int64_t e = threadIdx.x; e < w; e += blockDim.x) d[e] = s[e];
}

__global__ void scatter_rows_kernel(const double* __restrict__ sd, const int* __restrict__ si,
                                    const int* __restrict__ rows, int k_out,
                                    double* __restrict__ dd, int* __restrict__ di) {
  const int64_t i = blockIdx.x;
  const int64_t r = rows[i];
  for (int j = threadIdx.x; j < k_out; j += blockDim.x) {
    dd[r * k_out + j] = sd[i * k_out + j];
    di[r * k_out + j] = si[i * k_out + j];
  }
}

// ---------------------------------------------------------------------------
// exact rows

struct ExactParams {
  PairSource src;
  double W, out_scale;
  int64_t n_ref;
  const int* rows;
  int64_t row0;
  int nofit;
  double* srow;
  double* out_full;
  int64_t out_ld;
};

constexpr int kExactRefsPerWG = 64;

__global__ __launch_bounds__(256) void exact_rows_kernel(ExactParams P) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t ri = blockIdx.y;
  const int64_t q = P.rows ? P.rows[ri] : P.row0 + ri;
  const int64_t r0 = (int64_t)blockIdx.x * kExactRefsPerWG;
  auto put = [&](int64_t r, double s) {
    if (P.out_full)
      P.out_full[q * P.out_ld + r] = P.out_scale * sqrt(fmax(s, 0.0) / P.W);
    else
      P.srow[ri * P.n_ref + r] = s;
  };
  if (P.src.metric == kMetricRMSD) {
    // four pairs per wave, one per 16-lane group
    const int grp = lane >> 4;
    for (int j = 4 * wave + grp; j < kExactRefsPerWG; j += 16) {
      const int64_t r = r0 + j;
      if (r >= P.n_ref) break;
      const double s = group_pair_score(P.src, r, q, P.nofit);
      if ((lane & 15) == 0) put(r, s);
    }
  } else {
    for (int j = wave; j < kExactRefsPerWG; j += 4) {
      const int64_t r = r0 + j;
      if (r >= P.n_ref) break;
      const double s = wave_data_pair_score(P.src, r, q);
      if (lane == 0) put(r, s);
    }
  }
}

// The k1-th smallest (s, index) of one exact score row, by radix select on
// the 64-bit keys (s >= 0, so the bit patterns order like the values) and,
// among keys equal to it, on the index: every entry with key < prefix, or
// key == prefix and index <= ip, is among the k1 smallest (exactly k1 of
// them).  256 threads.
__device__ void row_threshold(const double* __restrict__ s, int64_t n_ref, int k1,
                              unsigned long long* prefix_out, unsigned* ip_out) {
  const int tid = threadIdx.x;
  __shared__ unsigned hist[256];
  __shared__ unsigned long long s_prefix;
  __shared__ int s_kk;
  __shared__ unsigned s_iprefix;
  unsigned long long prefix = 0;
  int kk = k1 - 1;
  for (int shift = 56; shift >= 0; shift -= 8) {
    hist[tid] = 0;
    __syncthreads();
    const unsigned long long mhi = (shift == 56) ? 0ull : (~0ull << (shift + 8));
    for (int64_t i = tid; i < n_ref; i += 256) {
      const unsigned long long k = (unsigned long long)__double_as_longlong(s[i]);
      if ((k & mhi) == (prefix & mhi)) atomicAdd(&hist[(k >> shift) & 255ull], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      unsigned cum = 0;
      int b = 0;
      for (; b < 255; ++b) {
        if (cum + hist[b] > (unsigned)kk) break;
        cum += hist[b];
      }
      s_prefix = prefix | ((unsigned long long)b << shift);
      s_kk = kk - (int)cum;
    }
    __syncthreads();
    prefix = s_prefix;
    kk = s_kk;
  }
  // Among keys equal to tau take the kk+1 smallest indices: radix select on
  // the index.
  unsigned ip = 0;
  int ik = kk;
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[tid] = 0;
    __syncthreads();
    const unsigned mhi = (shift == 24) ? 0u : (0xFFFFFFFFu << (shift + 8));
    for (int64_t i = tid; i < n_ref; i += 256) {
      if ((unsigned long long)__double_as_longlong(s[i]) != prefix) continue;
      const unsigned k = (unsigned)i;
      if ((k & mhi) == (ip & mhi)) atomicAdd(&hist[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      unsigned cum = 0;
      int b = 0;
      for (; b < 255; ++b) {
        if (cum + hist[b] > (unsigned)ik) break;
        cum += hist[b];
      }
      s_iprefix = ip | ((unsigned)b << shift);
      s_kk = ik - (int)cum;
    }
    __syncthreads();
    ip = s_iprefix;
    ik = s_kk;
  }
  *prefix_out = prefix;
  *ip_out = ip;
}

__device__ __forceinline__ bool among_smallest(double v, int64_t i, unsigned long long prefix,
                                               unsigned ip) {
  const unsigned long long k = (unsigned long long)__double_as_longlong(v);
  return k < prefix || (k == prefix && (unsigned)i <= ip);
}

// Top-k1 by (s, index) over one exact score row per workgroup (k1 <= kCandMax).
__global__ __launch_bounds__(256) void select_rows_kernel(const double* __restrict__ srow,
                                                          int64_t n_ref,
                                                          const int* __restrict__ rows,
                                                          double W, double out_scale,
                                                          int k1, int k_out,
                                                          int rank_off, double* out_dist,
                                                          int* out_idx) {
  const int64_t ri = blockIdx.x;
  const int64_t q = rows ? rows[ri] : ri;
  const double* s = srow + ri * n_ref;
  const int tid = threadIdx.x;
  __shared__ int s_nc;
  __shared__ int cand_idx[kCandMax];
  __shared__ double cand_s[kCandMax];
  unsigned long long prefix;
  unsigned ip;
  row_threshold(s, n_ref, k1, &prefix, &ip);
  if (tid == 0) s_nc = 0;
  __syncthreads();
  for (int64_t i = tid; i < n_ref; i += 256) {
    if (among_smallest(s[i], i, prefix, ip)) {
      const int pos = atomicAdd(&s_nc, 1);
      if (pos < kCandMax) {
        cand_idx[pos] = (int)i;
        cand_s[pos] = s[i];
      }
    }
  }
  __syncthreads();
  const int nc = min(s_nc, kCandMax);
  int np2 = 64;
  while (np2 < nc) np2 <<= 1;
  for (int i = nc + tid; i < np2; i += 256) {
    cand_s[i] = INFINITY;
    cand_idx[i] = INT_MAX;
  }
  block_bitonic_sort(cand_s, cand_idx, np2);
  for (int j = tid; j < k_out; j += 256) {
    const int src = j + rank_off;
    out_dist[q * k_out + j] = out_scale * sqrt(fmax(cand_s[src], 0.0) / W);
    out_idx[q * k_out + j] = cand_idx[src];
  }
}

// Any k1 (beyond the LDS candidate arrays): the k1 smallest (s, index) of
// each row, in index order, to [ri][k1] key / index arrays (the keys are the
// scores' bit patterns).  Thread t takes a contiguous run of the row, so a
// block-wide scan of the per-thread counts gives every entry its slot.
__global__ __launch_bounds__(256) void compact_rows_kernel(const double* __restrict__ srow,
                                                           int64_t n_ref, int k1,
                                                           unsigned long long* __restrict__ keys,
                                                           int* __restrict__ idx) {
  const int64_t ri = blockIdx.x;
  const double* s = srow + ri * n_ref;
  const int tid = threadIdx.x;
  __shared__ int part[256];
  unsigned long long prefix;
  unsigned ip;
  row_threshold(s, n_ref, k1, &prefix, &ip);
  const int64_t per = (n_ref + 255) / 256;
  const int64_t lo = min(n_ref, tid * per), hi = min(n_ref, lo + per);
  int cnt = 0;
  for (int64_t i = lo; i < hi; ++i) cnt += among_smallest(s[i], i, prefix, ip) ? 1 : 0;
  part[tid] = cnt;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {  // inclusive Hillis-Steele scan
    const int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int pos = part[tid] - cnt;
  unsigned long long* kr = keys + ri * (int64_t)k1;
  int* ir = idx + ri * (int64_t)k1;
  for (int64_t i = lo; i < hi && pos < k1; ++i)
    if (among_smallest(s[i], i, prefix, ip)) {
      kr[pos] = (unsigned long long)__double_as_longlong(s[i]);
      ir[pos] = (int)i;
      ++pos;
    }
}

// Sorted [ri][k1] rows to the outputs: ranks rank_off .. rank_off+k_out-1.
__global__ void write_sorted_rows_kernel(const unsigned long long* __restrict__ keys,
                                         const int* __restrict__ idx, const int* __restrict__ rows,
                                         int k1, int k_out, int rank_off, double W,
                                         double out_scale, double* __restrict__ out_dist,
                                         int* __restrict__ out_idx) {
  const int64_t ri = blockIdx.x;
  const int64_t q = rows ? rows[ri] : ri;
  for (int j = threadIdx.x; j < k_out; j += blockDim.x) {
    const int64_t src = ri * (int64_t)k1 + j + rank_off;
    const double sv = __longlong_as_double((long long)keys[src]);
    out_dist[q * k_out + j] = out_scale * sqrt(fmax(sv, 0.0) / W);
    out_idx[q * k_out + j] = idx[src];
  }
}

__global__ void segment_offsets_kernel(int* off, int n, int k1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) off[i] = i * k1;
}

// thr may be null: the shared thresholds are kept (set by a landmark pass).
__global__ void init_lists_kernel(int* cnt, float* ovf, unsigned* thr, int64_t n_hdr,
                                  int64_t n_fit) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_hdr) {
    cnt[i] = 0;
    ovf[i] = INFINITY;
  }
  if (thr && i < n_fit) thr[i] = kInfBits;
}

__global__ void max_kernel(const double* __restrict__ v, int64_t n, double* out) {
  __shared__ double red[256];
  double m = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) m = fmax(m, v[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Also clears the merge's candidate counter: a kernel store instead of a
// memset node keeps the per-batch sequence free of memsets when it is
// captured into a graph.
__global__ void eps_kernel(const double* __restrict__ fGt, int64_t n_fit,
                           const double* __restrict__ gmax, double kappa, float* eps2,
                           int* n_cand) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && n_cand) *n_cand = 0;
  if (i < n_fit) eps2[i] = (float)(2.0 * kappa * (fGt[i] + *gmax));
}

// Symmetric self query: row r also receives the scores of
// pairs computed the other way round (r's reference-side operand against
// f's query-side one), whose error bound is kappa (G~ref[r] + G~q[f]); the
// window covers both kinds.
__global__ void eps_sym_kernel(const double* __restrict__ qGt, const double* __restrict__ rGt,
                               int64_t n, const double* __restrict__ gmax_ref,
                               const double* __restrict__ gmax_q, double kappa, float* eps2,
                               int* n_cand) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && n_cand) *n_cand = 0;
  if (i < n)
    eps2[i] = (float)(2.0 * kappa * fmax(qGt[i] + *gmax_ref, rGt[i] + *gmax_q));
}

__global__ void identity_kernel(int64_t n_fit, int64_t n_ref, int* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_fit * n_ref) out[i] = (int)(i % n_ref);
}


// ---------------------------------------------------------------------------
// data vectors (knn_data)

// mean[e] = (sum over the n vectors of x[i][e]) / n, one thread per element.
__global__ __launch_bounds__(256) void column_mean_kernel(const double* __restrict__ x, int64_t n,
                                                          int vdim, double* __restrict__ mean) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= vdim) return;
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += x[i * vdim + e];
  mean[e] = s / (double)n;
}

// One wave per vector: the float source values of the contraction operand,
// laid out as natoms "atoms" of 3 values (zero beyond vdim).  Euclidean:
// x - shift (any common shift leaves the distances unchanged and a centred
// set has smaller G terms, so a tighter window).  Correlation: the unit
// vector (x - mean) / |x - mean|, whose squared distances are 2(1 - r); aux
// gets the mean and the sample standard deviation in the form
// correlation_distance computes it (mdsctk.cpp:335-349).
__global__ __launch_bounds__(256) void data_prep_kernel(const double* __restrict__ x, int64_t n,
                                                        int vdim, int natoms, int metric,
                                                        const double* __restrict__ shift,
                                                        float* __restrict__ out,
                                                        double* __restrict__ aux) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + wave;
  if (row >= n) return;
  const double* xr = x + row * vdim;
  float* o = out + row * (int64_t)natoms * 3;
  const int width = natoms * 3;
  if (metric == kMetricCorrelation) {
    double s = 0.0, ss = 0.0;
    for (int e = lane; e < vdim; e += 64) {
      const double v = xr[e];
      s += v;
      ss += v * v;
    }
    s = wave_sum_d(s);
    ss = wave_sum_d(ss);
    const double dn = (double)vdim;
    const double mean = s / dn;
    double c2 = 0.0;
    for (int e = lane; e < vdim; e += 64) {
      const double d = xr[e] - mean;
      c2 += d * d;
    }
    c2 = wave_sum_d(c2);
    const double inv = c2 > 0.0 ? 1.0 / sqrt(c2) : 0.0;
    for (int e = lane; e < width; e += 64) o[e] = e < vdim ? (float)((xr[e] - mean) * inv) : 0.f;
    if (lane == 0) {
      aux[row * 2 + 0] = mean;
      aux[row * 2 + 1] = sqrt(((dn * ss) - (s * s)) / (dn * (dn - 1.0)));
    }
  } else {
    for (int e = lane; e < width; e += 64) o[e] = e < vdim ? (float)(xr[e] - shift[e]) : 0.f;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers

int contract_fit_tiles(int prec) { return prec == kFP32 ? 2 : 4; }

int list_capacity_for(int k1) {
  if (k1 <= 128) return 256;
  if (k1 <= 384) return 512;
  if (k1 <= 896) return 1024;
  return 0;
}

hipError_t launch_prep(const DevFrames& f, const DevWeights& w, int side,
                       const DevFrames* second, int side2, hipStream_t st, const float* src_xyz,
                       const int* perm, double* rot) {
  const int64_t rows = (int64_t)f.n_tiles * kTile;
  const float* xyz = src_xyz ? src_xyz : f.xyz;
  if (rot && w.metric != kMetricRMSD) return hipErrorInvalidValue;
  const int center = w.metric == kMetricRMSD ? 1 : 0;
  double* gt = f.op ? f.Gt : nullptr;
  double* gt2 = (f.op && second) ? second->Gt : nullptr;
  const dim3 sgrid((unsigned)((rows + 3) / 4)), sblock(256);
  const dim3 ogrid((unsigned)f.n_tiles, (unsigned)((f.n_groups + kGroupsPerStage - 1) / kGroupsPerStage));
#define MDSCTK_PREP(PR)                                                                       \
  do {                                                                                        \
    hipLaunchKernelGGL(frame_stats_kernel<PR>, sgrid, sblock, 0, st, xyz, f.n, rows,          \
                       f.natoms, w.w, w.sw, w.W, center, side, side2, f.cen, f.G, gt, gt2,    \
                       perm, rot);                                                            \
    if (f.op)                                                                                 \
      hipLaunchKernelGGL(operand_kernel<PR>, ogrid, dim3(192), 0, st, xyz, f.n, f.natoms,     \
                         f.n_groups, w.sw, f.cen, side, f.op, perm, rot, side2,              \
                         second ? second->op : nullptr);                                      \
  } while (0)
  switch (f.prec) {
    case kFP32: MDSCTK_PREP(kFP32); break;
    case kBF16: MDSCTK_PREP(kBF16); break;
    case kFP16: MDSCTK_PREP(kFP16); break;
    default: return hipErrorInvalidValue;
  }
#undef MDSCTK_PREP
  return hipGetLastError();
}

hipError_t launch_init_lists(const Lists& L, int64_t n_fit, hipStream_t st, bool reset_thr) {
  const int64_t n_hdr = n_fit * L.nl;
  const int64_t n = n_hdr > n_fit ? n_hdr : n_fit;
  hipLaunchKernelGGL(init_lists_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     L.cnt, L.ovf, reset_thr ? L.thr : nullptr, n_hdr, n_fit);
  return hipGetLastError();
}

hipError_t launch_max_gt(const DevFrames& f, double* out, hipStream_t st) {
  hipLaunchKernelGGL(max_kernel, dim3(1), dim3(256), 0, st, f.Gt, (int64_t)f.n_tiles * kTile,
                     out);
  return hipGetLastError();
}

hipError_t launch_eps(const DevFrames& fit, const DevFrames& ref, double kappa, float* eps2,
                      int* n_cand, hipStream_t st) {
  if (!ref.gmax) return hipErrorInvalidValue;
  const int64_t n = fit.n > 0 ? fit.n : 1;
  hipLaunchKernelGGL(eps_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, fit.Gt,
                     fit.n, ref.gmax, kappa, eps2, n_cand);
  return hipGetLastError();
}

hipError_t launch_eps_sym(const DevFrames& fit, const DevFrames& ref, double kappa,
                          double* gmax_q, float* eps2, int* n_cand, hipStream_t st) {
  if (!ref.gmax || fit.n != ref.n) return hipErrorInvalidValue;
  hipError_t e = launch_max_gt(fit, gmax_q, st);
  if (e != hipSuccess) return e;
  const int64_t n = fit.n > 0 ? fit.n : 1;
  hipLaunchKernelGGL(eps_sym_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, fit.Gt,
                     ref.Gt, fit.n, ref.gmax, gmax_q, kappa, eps2, n_cand);
  return hipGetLastError();
}

hipError_t launch_contract(const DevFrames& ref, const DevFrames& fit, const Lists& L,
                           const float* eps2, int k1, int ref_phases, bool nofit,
                           double* dbg_scores, int grid_wgs, hipStream_t st, const Mirror* mirror,
                           int ph_run) {
  ContractParams P;
  P.rop = ref.op;
  P.fop = fit.op;
  P.rGt = ref.Gt;
  P.fGt = fit.Gt;
  P.n_ref = ref.n;
  P.n_fit = fit.n;
  P.ref_pairs = ref.n_tiles / kRefTilesPerWG;
  P.fit_tiles = fit.n_tiles;
  P.n_groups = ref.n_groups;
  P.phases = ref_phases;
  P.k1 = k1;
  P.eps2 = eps2;
  P.ent = L.ent;
  P.cnt = L.cnt;
  P.ovf = L.ovf;
  P.thr = L.thr;
  P.nl = L.nl;
  P.cap = L.cap;
  P.nofit = nofit ? 1 : 0;
  P.dbg = dbg_scores;
  P.ph_run = ph_run > 0 ? ph_run : ref_phases;
  P.ment = mirror ? mirror->ent : nullptr;
  P.mcnt = mirror ? mirror->cnt : nullptr;
  P.movf = mirror ? mirror->ovf : nullptr;
  P.mcap = mirror ? mirror->cap : 0;
  if (P.ph_run > ref_phases || ref_phases < 1) return hipErrorInvalidValue;
  // the split forms' pipelined K-loop walks K-groups four at a time
  if (ref.prec != kFP32 && (ref.n_groups % kSplitGroupAlign)) return hipErrorInvalidValue;
  if (mirror && (fit.n != ref.n || fit.n_tiles != ref.n_tiles || !L.thr || !mirror->ent ||
                 !mirror->cnt || !mirror->ovf || mirror->cap < 1 || P.ph_run != ref_phases))
    return hipErrorInvalidValue;
  dim3 grid(grid_wgs);
  // FP32 runs two 4-wave workgroups per CU (2 x 2 tiles each); the split
  // forms one 8-wave workgroup per CU (2 x 4 tiles: a third less L2 traffic
  // per MFMA).
#define MDSCTK_CONTRACT(PR, PER, FT, MODE, TR)                                            \
  hipLaunchKernelGGL((contract_kernel<PR, PER, FT, MODE, TR>), grid,                        \
                     dim3(64 * kRefTilesPerWG * FT), 0, st, P)
#define MDSCTK_CONTRACT_TR(PR, PER, FT, MODE)     \
  do {                                            \
    if (nofit)                                    \
      MDSCTK_CONTRACT(PR, PER, FT, MODE, true);   \
    else                                          \
      MDSCTK_CONTRACT(PR, PER, FT, MODE, false);  \
  } while (0)
#define MDSCTK_CONTRACT_CAP(PR, FT)                 \
  do {                                              \
    if (dbg_scores)                                 \
      MDSCTK_CONTRACT_TR(PR, 4, FT, kOutDebug);     \
    else if (mirror && L.cap == 256)                \
      MDSCTK_CONTRACT_TR(PR, 4, FT, kOutSym);       \
    else if (mirror && L.cap == 512)                \
      MDSCTK_CONTRACT_TR(PR, 8, FT, kOutSym);       \
    else if (mirror && L.cap == 1024)               \
      MDSCTK_CONTRACT_TR(PR, 16, FT, kOutSym);      \
    else if (mirror)                                \
      return hipErrorInvalidValue;                  \
    else if (L.cap == 256)                          \
      MDSCTK_CONTRACT_TR(PR, 4, FT, kOutLists);     \
    else if (L.cap == 512)                          \
      MDSCTK_CONTRACT_TR(PR, 8, FT, kOutLists);     \
    else if (L.cap == 1024)                         \
      MDSCTK_CONTRACT_TR(PR, 16, FT, kOutLists);    \
    else                                            \
      return hipErrorInvalidValue;                  \
  } while (0)
  switch (ref.prec) {
    case kFP32: MDSCTK_CONTRACT_CAP(kFP32, 2); break;
    case kBF16: MDSCTK_CONTRACT_CAP(kBF16, 4); break;
    case kFP16: MDSCTK_CONTRACT_CAP(kFP16, 4); break;
    default: return hipErrorInvalidValue;
  }
#undef MDSCTK_CONTRACT_CAP
#undef MDSCTK_CONTRACT_TR
#undef MDSCTK_CONTRACT
  return hipGetLastError();
}

static PairSource pair_source(const DevFrames& ref, const DevFrames& fit, const DevWeights& w) {
  PairSource P;
  P.rxyz = ref.xyz;
  P.rcen = ref.cen;
  P.rG = ref.G;
  P.fxyz = fit.xyz;
  P.fcen = fit.cen;
  P.fG = fit.G;
  P.w = w.w;
  P.natoms = ref.natoms;
  P.metric = w.metric;
  P.vdim = ref.vdim;
  P.r64 = ref.x64;
  P.raux = ref.aux;
  P.f64 = fit.x64;
  P.faux = fit.aux;
  return P;
}

hipError_t launch_merge(const DevFrames& ref, const DevFrames& fit, const DevWeights& w,
                        const Lists& L, const float* eps2, int k1, int k_out, int rank_off,
                        bool nofit, double* out_dist, int* out_idx, int* fallback, int* n_cand,
                        const int* order, hipStream_t st, const Mirror* mirror, const int* perm) {
  MergeParams P;
  P.src = pair_source(ref, fit, w);
  P.W = w.W;
  P.out_scale = w.out_scale;
  P.ent = L.ent;
  P.cnt = L.cnt;
  P.ovf = L.ovf;
  P.nl = L.nl;
  P.cap = L.cap;
  P.ment = mirror ? mirror->ent : nullptr;
  P.mcnt = mirror ? mirror->cnt : nullptr;
  P.movf = mirror ? mirror->ovf : nullptr;
  P.mcap = mirror ? mirror->cap : 0;
  P.eps2 = eps2;
  P.k1 = k1;
  P.k_out = k_out;
  P.rank_off = rank_off;
  P.nofit = nofit ? 1 : 0;
  P.out_dist = out_dist;
  P.out_idx = out_idx;
  P.fallback = fallback;
  P.n_cand = n_cand;
  P.order = order;
  P.perm = perm;
  // k1 <= 256: windows of up to 1024 candidates (C3: 184 per query), half
  // the candidate LDS of the general instance
  const bool stage = w.metric == kMetricRMSD && ref.natoms <= kMergeStageAtoms;
  const size_t dyn = stage ? (size_t)ref.natoms * 3 * sizeof(double) : 0;
  const dim3 grid((unsigned)fit.n), block(256);
  if (k1 <= kMergeSmallK1) {
    if (stage)
      hipLaunchKernelGGL((merge_kernel<true, kCandSmall>), grid, block, dyn, st, P);
    else
      hipLaunchKernelGGL((merge_kernel<false, kCandSmall>), grid, block, dyn, st, P);
  } else {
    if (stage)
      hipLaunchKernelGGL((merge_kernel<true, kCandMax>), grid, block, dyn, st, P);
    else
      hipLaunchKernelGGL((merge_kernel<false, kCandMax>), grid, block, dyn, st, P);
  }
  return hipGetLastError();
}

hipError_t launch_permute_thresholds(const unsigned* src, const int* perm, const float* add,
                                     int64_t n, unsigned* dst, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(permute_thresholds_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     src, perm, add, n, dst);
  return hipGetLastError();
}

hipError_t launch_landmark_cap(const Lists& L, int64_t n, int k1, const float* eps2,
                               int phases, unsigned* thr, int* label, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (k1 < 1 || phases < 1 || !thr || !label) return hipErrorInvalidValue;
  hipLaunchKernelGGL(landmark_cap_kernel, dim3((unsigned)n), dim3(256), 0, st, L.ent, L.cnt, L.nl,
                     L.cap, k1, eps2, phases, thr, label);
  return hipGetLastError();
}

hipError_t launch_row_order(const int* label, int64_t n, int nlab, int* counts, int* order,
                            hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (nlab < 1) return hipErrorInvalidValue;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)((nlab + 255) / 256)), dim3(256), 0, st,
                     counts, (int64_t)nlab, 0);
  hipLaunchKernelGGL(order_count_kernel, dim3(blocks), dim3(256), 0, st, label, n, counts);
  hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(1024), 0, st, counts, nlab);
  hipLaunchKernelGGL(order_scatter_kernel, dim3(blocks), dim3(256), 0, st, label, n, counts,
                     order);
  return hipGetLastError();
}

hipError_t launch_mirror_init(const Mirror& m, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mirror_init_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     m.cnt, m.ovf, n);
  return hipGetLastError();
}

hipError_t launch_mirror_sum(const Mirror& m, int64_t n, unsigned long long* out,
                             hipStream_t st) {
  hipLaunchKernelGGL(mirror_sum_kernel, dim3(1), dim3(256), 0, st, m.cnt, n, m.cap, out);
  return hipGetLastError();
}

hipError_t launch_gather_frames(const float* src, const int* rows, int nrows, int natoms,
                                float* dst, hipStream_t st) {
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_frames_kernel, dim3((unsigned)nrows), dim3(256), 0, st, src, rows,
                     natoms, dst);
  return hipGetLastError();
}

hipError_t launch_scatter_rows(const double* sd, const int* si, const int* rows, int nrows,
                               int k_out, double* dd, int* di, hipStream_t st) {
  if (nrows <= 0 || k_out <= 0) return hipSuccess;
  hipLaunchKernelGGL(scatter_rows_kernel, dim3((unsigned)nrows), dim3(64), 0, st, sd, si, rows,
                     k_out, dd, di);
  return hipGetLastError();
}

hipError_t launch_exact_rows(const DevFrames& ref, const DevFrames& fit, const DevWeights& w,
                             const int* rows, int nrows, bool nofit, double* srow,
                             double* out_full, int64_t out_ld, hipStream_t st) {
  ExactParams P;
  P.src = pair_source(ref, fit, w);
  P.W = w.W;
  P.out_scale = w.out_scale;
  P.n_ref = ref.n;
  P.nofit = nofit ? 1 : 0;
  P.out_ld = out_ld;
  // grid.y is limited to 65535; walk the rows in chunks.
  constexpr int kChunk = 65535;
  for (int r0 = 0; r0 < nrows; r0 += kChunk) {
    const int nr = nrows - r0 < kChunk ? nrows - r0 : kChunk;
    P.rows = rows ? rows + r0 : nullptr;
    P.row0 = rows ? 0 : r0;
    P.srow = srow ? srow + (int64_t)r0 * ref.n : nullptr;
    P.out_full = out_full;
    dim3 grid((unsigned)((ref.n + kExactRefsPerWG - 1) / kExactRefsPerWG), (unsigned)nr);
    hipLaunchKernelGGL(exact_rows_kernel, grid, dim3(256), 0, st, P);
  }
  return hipGetLastError();
}

hipError_t launch_select_rows(const double* srow, int64_t n_ref, const int* rows, int nrows,
                              const DevWeights& w, int k1, int k_out, int rank_off,
                              double* out_dist, int* out_idx, hipStream_t st) {
  if (k1 > kCandMax) return hipErrorInvalidValue;
  hipLaunchKernelGGL(select_rows_kernel, dim3((unsigned)nrows), dim3(256), 0, st, srow, n_ref,
                     rows, w.W, w.out_scale, k1, k_out, rank_off, out_dist, out_idx);
  return hipGetLastError();
}

size_t sort_rows_scratch(int nrows, int k1) {
  size_t temp = 0;
  (void)hipcub::DeviceSegmentedRadixSort::SortPairs<unsigned long long, int, const int*>(
      nullptr, temp, nullptr, nullptr, nullptr, nullptr, nrows * k1, nrows, nullptr, nullptr);
  const size_t items = (size_t)nrows * k1;
  return 2 * items * (sizeof(unsigned long long) + sizeof(int)) + (size_t)(nrows + 1) * 4 + temp +
         1024;
}

hipError_t launch_select_rows_sorted(const double* srow, int64_t n_ref, const int* rows,
                                     int nrows, const DevWeights& w, int k1, int k_out,
                                     int rank_off, double* out_dist, int* out_idx, void* scratch,
                                     size_t scratch_bytes, hipStream_t st) {
  if (nrows <= 0) return hipSuccess;
  if ((int64_t)nrows * k1 > INT32_MAX || k1 > n_ref) return hipErrorInvalidValue;
  const size_t items = (size_t)nrows * k1;
  char* p = (char*)scratch;
  auto take = [&](size_t b) {
    char* r = p;
    p += (b + 255) & ~size_t(255);
    return (void*)r;
  };
  auto* k_in = (unsigned long long*)take(items * sizeof(unsigned long long));
  auto* k_out_b = (unsigned long long*)take(items * sizeof(unsigned long long));
  auto* i_in = (int*)take(items * sizeof(int));
  auto* i_out = (int*)take(items * sizeof(int));
  auto* off = (int*)take((size_t)(nrows + 1) * sizeof(int));
  const size_t used = (size_t)(p - (char*)scratch);
  if (used > scratch_bytes) return hipErrorInvalidValue;
  size_t temp = scratch_bytes - used;
  hipLaunchKernelGGL(compact_rows_kernel, dim3((unsigned)nrows), dim3(256), 0, st, srow, n_ref, k1,
                     k_in, i_in);
  hipLaunchKernelGGL(segment_offsets_kernel, dim3((unsigned)((nrows + 256) / 256)), dim3(256), 0,
                     st, off, nrows, k1);
  // stable radix sort on the score bits: rows come in index order, so ties
  // keep ascending indices
  hipError_t e = hipcub::DeviceSegmentedRadixSort::SortPairs(
      (void*)p, temp, k_in, k_out_b, i_in, i_out, (int)items, nrows, (const int*)off,
      (const int*)off + 1, 0, 64, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(write_sorted_rows_kernel, dim3((unsigned)nrows), dim3(256), 0, st, k_out_b,
                     i_out, rows, k1, k_out, rank_off, w.W, w.out_scale, out_dist, out_idx);
  return hipGetLastError();
}

hipError_t launch_column_mean(const double* x, int64_t n, int vdim, double* mean, hipStream_t st) {
  hipLaunchKernelGGL(column_mean_kernel, dim3((unsigned)((vdim + 255) / 256)), dim3(256), 0, st, x,
                     n, vdim, mean);
  return hipGetLastError();
}

hipError_t launch_data_prep(const DevFrames& f, int metric, const double* shift, hipStream_t st) {
  if (f.vdim <= 0 || f.natoms * 3 < f.vdim || !f.x64 || !f.xyz) return hipErrorInvalidValue;
  if (metric == kMetricCorrelation ? !f.aux : !shift) return hipErrorInvalidValue;
  if (f.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(data_prep_kernel, dim3((unsigned)((f.n + 3) / 4)), dim3(256), 0, st, f.x64,
                     f.n, f.vdim, f.natoms, metric, shift, f.xyz, f.aux);
  return hipGetLastError();
}

hipError_t launch_identity_rows(int64_t n_fit, int64_t n_ref, int* out_idx, hipStream_t st) {
  const int64_t n = n_fit * n_ref;
  hipLaunchKernelGGL(identity_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n_fit,
                     n_ref, out_idx);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fixed workspace of the allocation-free query (mdsctk_knn_query_device_noalloc):
// a file-scope device array, one copy per GPU, held until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory, so
// it is never passed to a kernel or a copy: scratch_region() looks up each
// device's own address once (hipGetSymbolAddress, with that device current)
// and every caller uses that address.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kScratchDevices] = {};
}  // namespace

hipError_t scratch_region(int device, void** addr) {
  *addr = nullptr;
  if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    int current = -1;
    hipError_t e = hipGetDevice(&current);
    if (e != hipSuccess) return e;
    if (current != device) return hipErrorInvalidDevice;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_addr[device] = p;
  }
  *addr = g_scratch_addr[device];
  return hipSuccess;
}

}  // namespace mdsctk
