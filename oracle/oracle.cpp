// CPU oracle for the knn_rms hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (mdsctk_amd/, include/,
// the knn_rms executable) links, loads or calls this file.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and there
// only as the checker or as the timed CPU baseline, never as the measured or
// shipped path.
//
// It holds two restatements of knn_rms.cpp (the reference tool), both
// written from the published mathematics, not from GROMACS sources:
//
//  1. The FP64 oracle (orc_*_exact, orc_rmsd_nm).  Mass-weighted centring in
//     FP64 (knn_rms.cpp:189,199 call GROMACS reset_x for this), the 3x3
//     weighted cross-covariance M in FP64, and the largest eigenvalue of
//     Horn's 4x4 quaternion key matrix K(M) found by Newton's method on its
//     characteristic polynomial (QCP: Theobald 2005, Liu/Agrafiotis/Theobald
//     2010).  RMSD^2 = (G_a + G_b - 2 lambda_max) / W.  The value is the
//     minimum over proper rotations, the quantity GROMACS do_fit + rmsdev
//     approximate in float (knn_rms.cpp:38-41).  Output is in Angstrom
//     (x10, knn_rms.cpp:40); top-(k+1) by (value, index) with rank 0 dropped
//     (knn_rms.cpp:131,242,248-249).
//     Pinned by: a long-double 4x4 Jacobi eigen-solve of K (orc_jacobi4),
//     numpy SVD-Kabsch golden vectors (tests/golden), closed forms (rigid
//     copies give 0, symmetry, mirror images, 2/3-atom cases) in
//     tests/test_oracle.py.
//
//  2. The reference-equivalent "port" (orc_port_*).  Float arithmetic as the
//     reference's GROMACS calls use it: float centring (reset_x), Kabsch's
//     construction of the optimal proper rotation from the eigenvectors of
//     the 6x6 matrix [[0,u^T],[u,0]] (cyclic Jacobi in double), the third
//     axis by cross products, rotation of the fit copy IN PLACE, float
//     rmsdev, x10.  As in knn_rms.cpp:236-238/272-274 the fit frame is copied
//     once per fit frame and then rotated cumulatively by every pair, and the
//     reference frame is memcpy'd per pair.  Top-k uses the reference's two
//     independent std::partial_sort calls (mdsctk.h:177-199).  OpenMP over
//     fit frames (knn_rms.cpp:232,268).  This is the timed CPU baseline.
//     Pinned against the FP64 oracle (agreement to float accuracy).
//
//  3. knn_data (orc_knn_data, orc_euclidean, orc_correlation): the
//     reference's FP64 distance functions (mdsctk.cpp:329-356) with their
//     operations in the same order, and knn_data's top-(k+1) / drop-rank-0
//     output (knn_data.cpp:196-237).  Pinned by numpy closed forms in
//     tests/test_oracle.py: Euclidean norms of differences, and
//     sqrt((1 - r) / 2) with r from numpy.corrcoef.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (in place).
// On return d[] holds eigenvalues, v[][] the eigenvectors as columns.
template <typename T, int N>
void jacobi_sym(T a[N][N], T d[N], T v[N][N]) {
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) v[i][j] = (i == j) ? T(1) : T(0);
  for (int sweep = 0; sweep < 100; ++sweep) {
    T off = 0, diag = 0;
    for (int p = 0; p < N; ++p) {
      diag += std::fabs(a[p][p]);
      for (int q = p + 1; q < N; ++q) off += std::fabs(a[p][q]);
    }
    if (off == T(0) || off <= diag * std::numeric_limits<T>::epsilon() * T(1e-3)) break;
    for (int p = 0; p < N - 1; ++p) {
      for (int q = p + 1; q < N; ++q) {
        T apq = a[p][q];
        if (apq == T(0)) continue;
        T theta = (a[q][q] - a[p][p]) / (T(2) * apq);
        T t = T(1) / (std::fabs(theta) + std::sqrt(theta * theta + T(1)));
        if (theta < T(0)) t = -t;
        T c = T(1) / std::sqrt(t * t + T(1));
        T s = t * c;
        for (int k = 0; k < N; ++k) {  // A <- A J
          T akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < N; ++k) {  // A <- J^T A
          T apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < N; ++k) {  // V <- V J
          T vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
  for (int i = 0; i < N; ++i) d[i] = a[i][i];
}

// Horn's key matrix K(S) for S[r][c] = sum_i w_i a_ir b_ic (row-major S[9]).
template <typename T>
void key_matrix(const T* S, T K[4][4]) {
  T xx = S[0], xy = S[1], xz = S[2];
  T yx = S[3], yy = S[4], yz = S[5];
  T zx = S[6], zy = S[7], zz = S[8];
  K[0][0] = xx + yy + zz; K[0][1] = yz - zy;       K[0][2] = zx - xz;        K[0][3] = xy - yx;
  K[1][0] = K[0][1];      K[1][1] = xx - yy - zz;  K[1][2] = xy + yx;        K[1][3] = zx + xz;
  K[2][0] = K[0][2];      K[2][1] = K[1][2];       K[2][2] = -xx + yy - zz;  K[2][3] = yz + zy;
  K[3][0] = K[0][3];      K[3][1] = K[1][3];       K[3][2] = K[2][3];        K[3][3] = -xx - yy + zz;
}

// Determinant of a 4x4 matrix by Laplace expansion over the 2x2 minors of
// rows (0,1) and rows (2,3).
template <typename T>
T det4(const T a[4][4]) {
  T s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
  T s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
  T s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
  T s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
  T s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
  T s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
  T c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
  T c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
  T c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
  T c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
  T c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
  T c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
  return s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
}

template <typename T>
T det3(const T* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Largest eigenvalue of K(S) by Newton's method on
//   P(l) = l^4 + c2 l^2 + c1 l + c0,  c2 = -2|S|_F^2, c1 = -8 det S, c0 = det K,
// started at the upper bound (G_a+G_b)/2.  Works on S scaled by that bound so
// that the iteration variable is O(1).  Returns -1 in *iters on non-convergence.
double qcp_lambda(const double* S_in, double ga, double gb, int* iters) {
  double g = 0.5 * (ga + gb);
  if (!(g > 0.0)) { if (iters) *iters = 0; return 0.0; }
  double S[9];
  for (int i = 0; i < 9; ++i) S[i] = S_in[i] / g;
  double K[4][4];
  key_matrix(S, K);
  double fro = 0.0;
  for (int i = 0; i < 9; ++i) fro += S[i] * S[i];
  const double c2 = -2.0 * fro;
  const double c1 = -8.0 * det3(S);
  const double c0 = det4(K);
  double l = 1.0;
  int it = 0;
  bool conv = false;
  double dp = 0.0;
  for (; it < 200; ++it) {
    double l2 = l * l;
    double p = (l2 + c2) * l2 + c1 * l + c0;
    dp = (4.0 * l2 + 2.0 * c2) * l + c1;
    if (dp == 0.0) break;
    double step = p / dp;
    l -= step;
    if (std::fabs(step) <= 1e-15 * std::fabs(l)) { conv = true; ++it; break; }
  }
  // Near a double root (near-collinear frames, lambda_1 ~ lambda_2) Newton
  // only resolves about half the digits: report non-convergence so that the
  // caller uses the Jacobi eigen-solve instead.
  if (std::fabs(dp) < 1e-4) conv = false;
  if (iters) *iters = conv ? it : -1;
  return l * g;
}

long double jacobi4_lambda(const double* S) {
  long double Sl[9];
  for (int i = 0; i < 9; ++i) Sl[i] = S[i];
  long double K[4][4], d[4], v[4][4];
  key_matrix(Sl, K);
  jacobi_sym<long double, 4>(K, d, v);
  long double m = d[0];
  for (int i = 1; i < 4; ++i) m = std::max(m, d[i]);
  return m;
}

// FP64 centring: c = sum w x / sum w; xc = x - c; G = sum w |xc|^2.
void center_f64(int n, const float* x, const float* w, double* xc, double* G, double* W,
                double* centroid = nullptr) {
  double c[3] = {0, 0, 0}, ws = 0;
  for (int i = 0; i < n; ++i) {
    double wi = w[i];
    ws += wi;
    for (int d = 0; d < 3; ++d) c[d] += wi * (double)x[3 * i + d];
  }
  for (int d = 0; d < 3; ++d) c[d] = (ws > 0) ? c[d] / ws : 0.0;
  if (centroid)
    for (int d = 0; d < 3; ++d) centroid[d] = c[d];
  double g = 0;
  for (int i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) {
      double v = (double)x[3 * i + d] - c[d];
      xc[3 * i + d] = v;
      g += (double)w[i] * v * v;
    }
  }
  *G = g;
  *W = ws;
}

// Squared RMSD (nm^2) between two FP64-centred frames.
double rmsd2_centred(int n, const double* a, const double* b, const float* w, double ga,
                     double gb, double W, int fit) {
  double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    double wi = w[i];
    for (int r = 0; r < 3; ++r) {
      double war = wi * a[3 * i + r];
      for (int c = 0; c < 3; ++c) S[3 * r + c] += war * b[3 * i + c];
    }
  }
  double lam;
  if (fit) {
    int it = 0;
    lam = qcp_lambda(S, ga, gb, &it);
    if (it < 0) lam = (double)jacobi4_lambda(S);
  } else {
    lam = S[0] + S[4] + S[8];
  }
  double s = ga + gb - 2.0 * lam;
  if (s < 0) s = 0;
  return (W > 0) ? s / W : 0.0;
}

struct Centred {
  std::vector<double> xc;
  std::vector<double> G;
  double W = 0;
};

Centred center_all(const float* x, int64_t nf, int n, const float* w, int nthreads) {
  Centred out;
  out.xc.resize((size_t)nf * n * 3);
  out.G.resize(nf);
  double W = 0;
  for (int i = 0; i < n; ++i) W += w[i];
  out.W = W;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t f = 0; f < nf; ++f) {
    double Wf;
    center_f64(n, x + (size_t)f * n * 3, w, &out.xc[(size_t)f * n * 3], &out.G[f], &Wf);
  }
  return out;
}

// ---------------------------------------------------------------------------
// Reference-equivalent float path ("port").

// Float mass-weighted centring, as reset_x does for knn_rms.cpp:189,199.
void port_reset_x(int n, float* x, const float* w) {
  float c[3] = {0, 0, 0}, mt = 0;
  for (int i = 0; i < n; ++i) {
    mt += w[i];
    for (int d = 0; d < 3; ++d) c[d] += w[i] * x[3 * i + d];
  }
  for (int d = 0; d < 3; ++d) c[d] /= mt;
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) x[3 * i + d] -= c[d];
}

// Kabsch fit of `fit` onto `ref` in place: u = sum w ref fit^T (float),
// eigenvectors of [[0,u^T],[u,0]] give the singular vectors, the third pair
// from cross products makes the rotation proper; fit <- R fit.
void port_do_fit(int n, const float* w, const float* ref, float* fit) {
  float u[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < n; ++i) {
    float m = w[i];
    if (m == 0.0f) continue;
    for (int r = 0; r < 3; ++r) {
      double xr = ref[3 * i + r];
      for (int c = 0; c < 3; ++c) u[r][c] += (float)(m * xr * (double)fit[3 * i + c]);
    }
  }
  double om[6][6], d[6], v[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) om[i][j] = 0.0;
  // [[0, u^T], [u, 0]]: eigenvector (h; k) has u h = s k, u^T k = s h, so h is
  // a fit-space (right) and k a ref-space (left) singular vector.
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      om[3 + r][c] = u[r][c];
      om[c][3 + r] = u[r][c];
    }
  jacobi_sym<double, 6>(om, d, v);
  double h[3][3], k[3][3];
  for (int j = 0; j < 2; ++j) {
    int best = 0;
    for (int i = 1; i < 6; ++i)
      if (d[i] > d[best]) best = i;
    d[best] = -1e300;
    for (int i = 0; i < 3; ++i) {
      h[j][i] = M_SQRT2 * v[i][best];
      k[j][i] = M_SQRT2 * v[3 + i][best];
    }
  }
  h[2][0] = h[0][1] * h[1][2] - h[0][2] * h[1][1];
  h[2][1] = h[0][2] * h[1][0] - h[0][0] * h[1][2];
  h[2][2] = h[0][0] * h[1][1] - h[0][1] * h[1][0];
  k[2][0] = k[0][1] * k[1][2] - k[0][2] * k[1][1];
  k[2][1] = k[0][2] * k[1][0] - k[0][0] * k[1][2];
  k[2][2] = k[0][0] * k[1][1] - k[0][1] * k[1][0];
  float R[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int j = 0; j < 3; ++j) s += k[j][r] * h[j][c];
      R[r][c] = (float)s;
    }
  for (int i = 0; i < n; ++i) {
    float x0 = fit[3 * i], x1 = fit[3 * i + 1], x2 = fit[3 * i + 2];
    for (int r = 0; r < 3; ++r) fit[3 * i + r] = R[r][0] * x0 + R[r][1] * x1 + R[r][2] * x2;
  }
}

float port_rmsdev(int n, const float* w, const float* a, const float* b) {
  float tm = 0, rd2 = 0;
  for (int i = 0; i < n; ++i) {
    float d2 = 0;
    for (int r = 0; r < 3; ++r) {
      float dd = a[3 * i + r] - b[3 * i + r];
      d2 += dd * dd;
    }
    tm += w[i];
    rd2 += w[i] * d2;
  }
  return std::sqrt(rd2 / tm);
}

// knn_rms.cpp:38-41
double port_distance(int n, const float* w, const float* ref, float* fit, bool dofit) {
  if (dofit) port_do_fit(n, w, ref, fit);
  return (double)port_rmsdev(n, w, ref, fit) * 10.0;
}

// mdsctk.h:177-199, k > 0 branch.
struct Perm {
  std::vector<double> data;
  std::vector<int> indices;
  bool operator()(int l, int r) const { return data[l] < data[r]; }
  void sort(int k) {
    indices.resize(data.size());
    for (size_t x = 0; x < data.size(); ++x) indices[x] = (int)x;
    if (k <= 0) {
      std::sort(indices.begin(), indices.end(), *this);
      std::sort(data.begin(), data.end());
    } else {
      std::partial_sort(indices.begin(), indices.begin() + k, indices.end(), *this);
      std::partial_sort(data.begin(), data.begin() + k, data.end());
    }
  }
};

}  // namespace

extern "C" {

// ---- FP64 oracle ---------------------------------------------------------

// Largest eigenvalue of K(S), QCP Newton (returns iterations via *iters, -1
// when it did not converge).
double orc_qcp_lambda(const double* S, double ga, double gb, int* iters) {
  return qcp_lambda(S, ga, gb, iters);
}

// Largest eigenvalue of K(S) by long-double Jacobi (independent check).
double orc_jacobi4_lambda(const double* S) { return (double)jacobi4_lambda(S); }

// FP64 centring of one frame; returns G, writes the centroid and W.
double orc_center(int natoms, const float* x, const float* w, double* centroid, double* W) {
  std::vector<double> xc((size_t)natoms * 3);
  double G, Ws;
  center_f64(natoms, x, w, xc.data(), &G, &Ws, centroid);
  if (W) *W = Ws;
  return G;
}

// Cross-covariance S[r][c] = sum w (a_r - ca_r)(b_c - cb_c) of two frames,
// FP64 centring.  Also returns G_a, G_b.
void orc_cross_cov(int natoms, const float* a, const float* b, const float* w, double* S,
                   double* ga, double* gb) {
  std::vector<double> ac((size_t)natoms * 3), bc((size_t)natoms * 3);
  double W;
  center_f64(natoms, a, w, ac.data(), ga, &W);
  center_f64(natoms, b, w, bc.data(), gb, &W);
  for (int i = 0; i < 9; ++i) S[i] = 0;
  for (int i = 0; i < natoms; ++i)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) S[3 * r + c] += (double)w[i] * ac[3 * i + r] * bc[3 * i + c];
}

// Optimal proper-rotation (fit=1) or centred-only (fit=0) RMSD in nm.
double orc_rmsd_nm(int natoms, const float* a, const float* b, const float* w, int fit) {
  std::vector<double> ac((size_t)natoms * 3), bc((size_t)natoms * 3);
  double ga, gb, W;
  center_f64(natoms, a, w, ac.data(), &ga, &W);
  center_f64(natoms, b, w, bc.data(), &gb, &W);
  return std::sqrt(rmsd2_centred(natoms, ac.data(), bc.data(), w, ga, gb, W, fit));
}

// RMSD (nm) for an explicit list of (ref, fit) pairs.
void orc_pairs_nm(const float* ref, const float* fit, int natoms, const float* w, int64_t npairs,
                  const int64_t* ref_idx, const int64_t* fit_idx, int fitflag, double* out,
                  int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
  for (int64_t p = 0; p < npairs; ++p)
    out[p] = orc_rmsd_nm(natoms, ref + (size_t)ref_idx[p] * natoms * 3,
                         fit + (size_t)fit_idx[p] * natoms * 3, w, fitflag);
}

// Full FP64 distance matrix in Angstrom, row-major [n_fit][n_ref].
int orc_full_exact(const float* ref, int64_t n_ref, const float* fit, int64_t n_fit, int natoms,
                   const float* w, int nofit, double* out_A, int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  Centred R = center_all(ref, n_ref, natoms, w, nthreads);
  Centred F = center_all(fit, n_fit, natoms, w, nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
  for (int64_t f = 0; f < n_fit; ++f)
    for (int64_t r = 0; r < n_ref; ++r)
      out_A[(size_t)f * n_ref + r] =
          10.0 * std::sqrt(rmsd2_centred(natoms, &R.xc[(size_t)r * natoms * 3],
                                         &F.xc[(size_t)f * natoms * 3], w, R.G[r], F.G[f], R.W,
                                         !nofit));
  return 0;
}

// Exact kNN: per fit row the k+1 smallest FP64 distances ordered by
// (value, index); writes ranks 1..k (or 0..k-1 with keep_rank0).  k is
// clamped to n_ref-1 (knn_rms.cpp:224-225), k+1 recomputed after the clamp.
// Returns the k actually used.
int orc_knn_exact(const float* ref, int64_t n_ref, const float* fit, int64_t n_fit, int natoms,
                  const float* w, int k, int nofit, int keep_rank0, double* out_A, int* out_idx,
                  int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  if (n_ref <= 0) return -1;
  if (k > n_ref - 1) k = (int)(n_ref - 1);
  if (k < 0) k = 0;
  const int k1 = k + 1;
  Centred R = center_all(ref, n_ref, natoms, w, nthreads);
  Centred F = center_all(fit, n_fit, natoms, w, nthreads);
#pragma omp parallel num_threads(nthreads)
  {
    std::vector<std::pair<double, int>> row(n_ref);
#pragma omp for schedule(dynamic, 1)
    for (int64_t f = 0; f < n_fit; ++f) {
      for (int64_t r = 0; r < n_ref; ++r)
        row[r] = {10.0 * std::sqrt(rmsd2_centred(natoms, &R.xc[(size_t)r * natoms * 3],
                                                 &F.xc[(size_t)f * natoms * 3], w, R.G[r],
                                                 F.G[f], R.W, !nofit)),
                  (int)r};
      std::partial_sort(row.begin(), row.begin() + k1, row.end());
      int off = keep_rank0 ? 0 : 1;
      for (int j = 0; j < k; ++j) {
        out_A[(size_t)f * k + j] = row[j + off].first;
        out_idx[(size_t)f * k + j] = row[j + off].second;
      }
    }
  }
  return k;
}

// ---- knn_data: data vectors --------------------------------------------------

// euclidean_distance (mdsctk.cpp:329-334), the same operations in the same
// order; knn_data passes the query vector first (knn_data.cpp:206-209).
double orc_euclidean(int size, const double* reference, const double* fitting) {
  double value = 0.0;
  for (int x = 0; x < size; x++) value += (reference[x] - fitting[x]) * (reference[x] - fitting[x]);
  return std::sqrt(value);
}

// correlation_distance (mdsctk.cpp:335-356): sums, the sample standard
// deviations in the reference's one-pass form, the centred product, then
// sqrt(max(0, (1 - r) / 2)).
double orc_correlation(int size, const double* reference, const double* fitting) {
  double rvalue = 0.0, rsvalue = 0.0, fvalue = 0.0, fsvalue = 0.0;
  const double dsize = (double)size;
  double value = 0.0;
  for (int x = 0; x < size; x++) {
    rvalue += reference[x];
    rsvalue += reference[x] * reference[x];
    fvalue += fitting[x];
    fsvalue += fitting[x] * fitting[x];
  }
  rsvalue = std::sqrt(((dsize * rsvalue) - (rvalue * rvalue)) / (dsize * (dsize - 1.0)));
  rvalue /= dsize;
  fsvalue = std::sqrt(((dsize * fsvalue) - (fvalue * fvalue)) / (dsize * (dsize - 1.0)));
  fvalue /= dsize;
  for (int x = 0; x < size; x++) value += (reference[x] - rvalue) * (fitting[x] - fvalue);
  value = (1.0 - (value / ((dsize - 1.0) * rsvalue * fsvalue))) / 2.0;
  if (value < 0.0) value = 0.0;
  return std::sqrt(value);
}

// knn_data's pair loop and output (knn_data.cpp:196-237): per query row the
// k+1 smallest distances ordered by (value, index), ranks 1..k written (or
// 0..k-1 with keep_rank0); k clamped to n_ref-1 with k+1 recomputed.
// metric 1 = Euclidean, 2 = correlation.  Returns the k used.
int orc_knn_data(const double* ref, int64_t n_ref, const double* fit, int64_t n_fit, int vdim,
                 int metric, int k, int keep_rank0, double* out, int* out_idx, int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  if (n_ref <= 0 || (metric != 1 && metric != 2)) return -1;
  if (k > n_ref - 1) k = (int)(n_ref - 1);
  if (k < 0) k = 0;
  const int k1 = k + 1;
#pragma omp parallel num_threads(nthreads)
  {
    std::vector<std::pair<double, int>> row(n_ref);
#pragma omp for schedule(dynamic, 1)
    for (int64_t f = 0; f < n_fit; ++f) {
      const double* q = fit + (size_t)f * vdim;
      for (int64_t r = 0; r < n_ref; ++r) {
        const double* x = ref + (size_t)r * vdim;
        row[r] = {metric == 1 ? orc_euclidean(vdim, q, x) : orc_correlation(vdim, q, x), (int)r};
      }
      std::partial_sort(row.begin(), row.begin() + k1, row.end());
      const int off = keep_rank0 ? 0 : 1;
      for (int j = 0; j < k; ++j) {
        out[(size_t)f * k + j] = row[j + off].first;
        out_idx[(size_t)f * k + j] = row[j + off].second;
      }
    }
  }
  return k;
}

// ---- reference-equivalent float port ---------------------------------------

// Float centring of a frame set in place (knn_rms.cpp:186-206).
void orc_port_center(float* x, int64_t nf, int natoms, const float* w) {
  for (int64_t f = 0; f < nf; ++f) port_reset_x(natoms, x + (size_t)f * natoms * 3, w);
}

// knn_rms.cpp:38-41 for one pair of float-centred frames; `fit` is rotated
// in place exactly as the reference does.
double orc_port_distance(int natoms, const float* w, const float* ref, float* fit, int dofit) {
  return port_distance(natoms, w, ref, fit, dofit != 0);
}

// The reference pair loop over fit rows [row0, row1) against all refs, on
// float-centred frames (call orc_port_center first).  Writes rows in the
// reference's file layout: k entries per row from rank 1 (sort=1), or the
// whole unsorted row (sort=0; indices are then the identity).  k is clamped
// as knn_rms.cpp:224-225 does, with k+1 recomputed after the clamp.
int orc_port_knn_rows(const float* ref_c, int64_t n_ref, const float* fit_c, int64_t row0,
                      int64_t row1, int natoms, const float* w, int k, int sort, int nofit,
                      double* out_A, int* out_idx, int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  if (n_ref <= 0) return -1;
  if (k > n_ref - 1) k = (int)(n_ref - 1);
  const int k1 = k + 1;
#pragma omp parallel num_threads(nthreads)
  {
    Perm p;
    p.data.resize(n_ref);
    std::vector<float> rbuf((size_t)natoms * 3), fbuf((size_t)natoms * 3);
#pragma omp for schedule(dynamic, 1)
    for (int64_t f = row0; f < row1; ++f) {
      std::memcpy(fbuf.data(), fit_c + (size_t)f * natoms * 3, sizeof(float) * natoms * 3);
      for (int64_t r = 0; r < n_ref; ++r) {
        std::memcpy(rbuf.data(), ref_c + (size_t)r * natoms * 3, sizeof(float) * natoms * 3);
        p.data[r] = port_distance(natoms, w, rbuf.data(), fbuf.data(), !nofit);
      }
      const int64_t row = f - row0;
      if (sort) {
        p.sort(k1);
        for (int j = 0; j < k; ++j) {
          out_A[(size_t)row * k + j] = p.data[1 + j];
          out_idx[(size_t)row * k + j] = p.indices[1 + j];
        }
      } else {
        for (int64_t r = 0; r < n_ref; ++r) {
          out_A[(size_t)row * n_ref + r] = p.data[r];
          out_idx[(size_t)row * n_ref + r] = (int)r;
        }
      }
    }
  }
  return k;
}

}  // extern "C"
