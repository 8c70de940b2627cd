// The stages after knn_rms on the host: the symmetric sparse kNN graph
// (.ssm, make_sysparse semantics) and its leading eigenpairs after the
// normalisation of decomp_sparse.  include/mdsctk_graph.h is the interface.
//
// The eigen-solver is a thick-restart Lanczos method (Wu & Simon, SIAM J.
// Matrix Anal. Appl. 22 (2000) 602-616) with full reorthogonalisation
// (classical Gram-Schmidt applied twice) and a cyclic Jacobi eigen-solve of
// the projected matrix; it stands in for the ARPACK dsaupd/dseupd call of
// decomp_sparse.cpp:236-264 with the same subspace size and target.
#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/mdsctk_graph.h"
#include "../../include/mdsctk_knn.h"

extern "C" void mdsctk_io_set_error(const char* msg);  // defined in xtc.cpp

namespace {

int io_fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int io_fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  mdsctk_io_set_error(buf);
  return code;
}

struct Edge {
  int32_t from, to;
  int64_t order;
  double d;
};

// y = S x for the symmetric matrix whose strict lower triangle is stored
// column-wise (pcol/irow/a): column c holds rows r > c.
void sym_spmv(int32_t n, const int32_t* pcol, const int32_t* irow, const double* a,
              const double* x, double* y) {
  std::fill(y, y + n, 0.0);
  for (int32_t c = 0; c < n; ++c) {
    const double xc = x[c];
    double acc = 0.0;
    for (int32_t e = pcol[c]; e < pcol[c + 1]; ++e) {
      const int32_t r = irow[e];
      y[r] += a[e] * xc;
      acc += a[e] * x[r];
    }
    y[c] += acc;
  }
}

double dot(const double* a, const double* b, int32_t n) {
  double s = 0.0;
  for (int32_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// Eigen-decomposition of a symmetric m x m matrix (row-major in `a`, which
// is destroyed) by cyclic Jacobi rotations: eigenvalues in `w`, the
// eigenvector of w[i] in column i of `v` (row-major m x m).
void jacobi_eig(int m, std::vector<double>& a, std::vector<double>& w, std::vector<double>& v) {
  v.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) v[(size_t)i * m + i] = 1.0;
  auto A = [&](int i, int j) -> double& { return a[(size_t)i * m + j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < m; ++i) {
      diag += A(i, i) * A(i, i);
      for (int j = i + 1; j < m; ++j) off += A(i, j) * A(i, j);
    }
    if (off <= 1e-34 * (diag + 1e-300)) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = A(p, q);
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; ++k) {  // columns p, q
          const double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {  // rows p, q
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          double* vr = &v[(size_t)k * m];
          const double vkp = vr[p], vkq = vr[q];
          vr[p] = c * vkp - s * vkq;
          vr[q] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(m);
  for (int i = 0; i < m; ++i) w[i] = A(i, i);
}

// Deterministic pseudo-random start / restart vectors (splitmix64).
double rand_unit(uint64_t& s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

// w -= V[:, 0..j] (V^T w), twice; returns the first-pass coefficients plus
// the second-pass corrections in h[0..j].
// A fixed team of host threads for the Lanczos vector work: run(f) calls
// f(t) on members t = 0..size()-1 (the caller is member 0) and returns when
// all are done; barrier() synchronises the members inside run().
class Team {
 public:
  explicit Team(int size) : size_(std::max(1, size)) {
    for (int t = 1; t < size_; ++t) workers_.emplace_back([this, t] { work(t); });
  }
  ~Team() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& th : workers_) th.join();
  }
  int size() const { return size_; }
  void run(const std::function<void(int)>& f) {
    {
      std::lock_guard<std::mutex> g(mu_);
      job_ = &f;
      done_ = 0;
      ++gen_;
    }
    cv_.notify_all();
    f(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return done_ == size_ - 1; });
    job_ = nullptr;
  }
  void barrier() {
    std::unique_lock<std::mutex> lk(mu_);
    const uint64_t phase = bar_phase_;
    if (++bar_count_ == size_) {
      bar_count_ = 0;
      ++bar_phase_;
      bar_cv_.notify_all();
    } else {
      bar_cv_.wait(lk, [&] { return bar_phase_ != phase; });
    }
  }

 private:
  void work(int t) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int)>* f;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        f = job_;
      }
      (*f)(t);
      {
        std::lock_guard<std::mutex> g(mu_);
        ++done_;
      }
      done_cv_.notify_one();
    }
  }
  const int size_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_, bar_cv_;
  const std::function<void(int)>* job_ = nullptr;
  uint64_t gen_ = 0, bar_phase_ = 0;
  int done_ = 0, bar_count_ = 0;
  bool stop_ = false;
};

// Row blocks of the reorthogonalisation.  Their number is fixed, so the
// floating-point sums do not depend on how many threads take them.
constexpr int kRowBlocks = 16;

// w -= V[:, 0..j] (V[:, 0..j]^T w), twice (classical Gram-Schmidt applied
// twice); h[0..j] receives the summed coefficients.  Rows are split into
// kRowBlocks blocks: each block's partial inner products are formed by one
// member, summed over the blocks in block order, and each member then
// updates its own blocks' rows.  part holds kRowBlocks x (j+1) doubles.
void reorthogonalise(Team& team, const std::vector<double>& V, int32_t n, int j, double* w,
                     double* h, std::vector<double>& part) {
  const int nv = j + 1;
  part.resize((size_t)kRowBlocks * nv);
  team.run([&](int t) {
    const int T = team.size();
    std::vector<double> hp(nv);
    for (int pass = 0; pass < 2; ++pass) {
      for (int b = t; b < kRowBlocks; b += T) {
        const int32_t r0 = (int32_t)((int64_t)n * b / kRowBlocks);
        const int32_t r1 = (int32_t)((int64_t)n * (b + 1) / kRowBlocks);
        for (int i = 0; i < nv; ++i)
          part[(size_t)b * nv + i] = dot(&V[(size_t)i * n + r0], w + r0, r1 - r0);
      }
      team.barrier();
      for (int i = 0; i < nv; ++i) {
        double c = 0.0;
        for (int b = 0; b < kRowBlocks; ++b) c += part[(size_t)b * nv + i];
        hp[i] = c;
      }
      team.barrier();  // every member has read part before it is rewritten
      for (int b = t; b < kRowBlocks; b += T) {
        const int32_t r0 = (int32_t)((int64_t)n * b / kRowBlocks);
        const int32_t r1 = (int32_t)((int64_t)n * (b + 1) / kRowBlocks);
        for (int i = 0; i < nv; ++i) {
          const double c = hp[i];
          const double* vi = &V[(size_t)i * n];
          for (int32_t r = r0; r < r1; ++r) w[r] -= c * vi[r];
        }
      }
      if (t == 0)
        for (int i = 0; i < nv; ++i) h[i] = pass == 0 ? hp[i] : h[i] + hp[i];
    }
  });
}

// Rows [r0, r1) of row block b.
inline int32_t block_row(int32_t n, int b) { return (int32_t)((int64_t)n * b / kRowBlocks); }

// The symmetric matrix whose strict lower triangle is (pcol, irow, a), as
// full rows: row r lists (c, value) for both triangles.
struct SymRows {
  std::vector<int64_t> ptr;
  std::vector<int32_t> col;
  std::vector<double> val;
};

SymRows sym_rows(int32_t n, const int32_t* pcol, const int32_t* irow, const double* a) {
  SymRows m;
  m.ptr.assign((size_t)n + 1, 0);
  for (int32_t c = 0; c < n; ++c)
    for (int32_t e = pcol[c]; e < pcol[c + 1]; ++e) {
      ++m.ptr[(size_t)irow[e] + 1];
      ++m.ptr[(size_t)c + 1];
    }
  for (int32_t r = 0; r < n; ++r) m.ptr[r + 1] += m.ptr[r];
  m.col.resize(m.ptr[n]);
  m.val.resize(m.ptr[n]);
  std::vector<int64_t> at(m.ptr.begin(), m.ptr.end() - 1);
  for (int32_t c = 0; c < n; ++c)
    for (int32_t e = pcol[c]; e < pcol[c + 1]; ++e) {
      const int32_t r = irow[e];
      m.col[at[r]] = c;
      m.val[at[r]++] = a[e];
      m.col[at[c]] = r;
      m.val[at[c]++] = a[e];
    }
  return m;
}

// y = S x, rows in blocks over the team.
void sym_rows_spmv(Team& team, int32_t n, const SymRows& m, const double* x, double* y) {
  team.run([&](int t) {
    for (int b = t; b < kRowBlocks; b += team.size())
      for (int32_t r = block_row(n, b); r < block_row(n, b + 1); ++r) {
        double acc = 0.0;
        for (int64_t e = m.ptr[r]; e < m.ptr[r + 1]; ++e) acc += m.val[e] * x[m.col[e]];
        y[r] = acc;
      }
  });
}

}  // namespace

extern "C" {

int mdsctk_ssm_build(const double* dist, const int32_t* idx, int64_t n_rows, int32_t maxk,
                     int32_t k, int32_t* pcol, int32_t* irow, double* vals, int64_t* nnz) {
  if (!dist || !idx || !nnz || n_rows < 0 || n_rows > INT32_MAX || maxk <= 0 || k <= 0)
    return io_fail(MDSCTK_ERR_ARG, "ssm_build: need rows, maxk > 0 and k > 0");
  if (k > maxk)
    return io_fail(MDSCTK_ERR_ARG, "ssm_build: output k (%d) is larger than the input k (%d)", k,
                   maxk);
  std::vector<Edge> edges;
  edges.reserve((size_t)n_rows * k);
  int64_t order = 0;
  for (int64_t i = 0; i < n_rows; ++i)
    for (int32_t x = 0; x < k; ++x) {
      const int32_t j = idx[i * maxk + x];
      const double d = dist[i * maxk + x];
      if (j < 0 || j >= n_rows)
        return io_fail(MDSCTK_ERR_ARG, "ssm_build: row %lld holds index %d outside [0, %lld)",
                       (long long)i, j, (long long)n_rows);
      if (j == i) continue;
      edges.push_back(Edge{(int32_t)std::min<int64_t>(i, j), (int32_t)std::max<int64_t>(i, j),
                           order++, d});
    }
  // (from, to) ascending; of repeated edges the last put wins
  std::sort(edges.begin(), edges.end(), [](const Edge& a, const Edge& b) {
    if (a.from != b.from) return a.from < b.from;
    if (a.to != b.to) return a.to < b.to;
    return a.order < b.order;
  });
  size_t m = 0;
  for (size_t e = 0; e < edges.size(); ++e) {
    if (m > 0 && edges[m - 1].from == edges[e].from && edges[m - 1].to == edges[e].to)
      edges[m - 1] = edges[e];
    else
      edges[m++] = edges[e];
  }
  edges.resize(m);
  *nnz = (int64_t)m;
  if (!pcol && !irow && !vals) return 0;
  if (!pcol || !irow || !vals) return io_fail(MDSCTK_ERR_ARG, "ssm_build: null output buffer");
  if (m > (size_t)INT32_MAX) return io_fail(MDSCTK_ERR_ARG, "ssm_build: too many edges");
  std::fill(pcol, pcol + n_rows + 1, 0);
  for (const Edge& e : edges) ++pcol[e.from + 1];
  for (int64_t c = 0; c < n_rows; ++c) pcol[c + 1] += pcol[c];
  for (size_t e = 0; e < m; ++e) {
    irow[e] = edges[e].to;
    vals[e] = edges[e].d;
  }
  return 0;
}

int mdsctk_ssm_write(const char* path, int32_t n, const int32_t* pcol, const int32_t* irow,
                     const double* vals) {
  if (!path || n < 0 || !pcol) return io_fail(MDSCTK_ERR_ARG, "ssm_write: bad arguments");
  const int64_t nnz = pcol[n];
  if (nnz > 0 && (!irow || !vals)) return io_fail(MDSCTK_ERR_ARG, "ssm_write: null arrays");
  FILE* f = std::fopen(path, "wb");
  if (!f) return io_fail(MDSCTK_ERR_IO, "cannot open %s for writing", path);
  bool ok = std::fwrite(&n, sizeof n, 1, f) == 1 &&
            std::fwrite(pcol, sizeof(int32_t), (size_t)n + 1, f) == (size_t)n + 1 &&
            std::fwrite(irow, sizeof(int32_t), (size_t)nnz, f) == (size_t)nnz &&
            std::fwrite(vals, sizeof(double), (size_t)nnz, f) == (size_t)nnz;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return io_fail(MDSCTK_ERR_IO, "write to %s failed", path);
  return 0;
}

int mdsctk_ssm_read(const char* path, int32_t* n_out, int64_t* nnz_out, int32_t* pcol,
                    int32_t* irow, double* vals) {
  if (!path || !n_out || !nnz_out) return io_fail(MDSCTK_ERR_ARG, "ssm_read: bad arguments");
  FILE* f = std::fopen(path, "rb");
  if (!f) return io_fail(MDSCTK_ERR_IO, "cannot open %s", path);
  int32_t n = 0;
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (std::fread(&n, sizeof n, 1, f) != 1 || n < 0 ||
      (long)(sizeof(int32_t) * ((int64_t)n + 2)) > size) {
    std::fclose(f);
    return io_fail(MDSCTK_ERR_FORMAT, "%s is not an .ssm file (bad dimension)", path);
  }
  std::vector<int32_t> pc((size_t)n + 1);
  if (std::fread(pc.data(), sizeof(int32_t), pc.size(), f) != pc.size()) {
    std::fclose(f);
    return io_fail(MDSCTK_ERR_FORMAT, "%s: truncated column pointers", path);
  }
  const int64_t nnz = pc[n];
  bool mono = pc[0] == 0;
  for (int32_t c = 0; c < n && mono; ++c) mono = pc[c] <= pc[c + 1];
  if (!mono || size != (long)(4 + 4 * ((int64_t)n + 1) + 12 * nnz)) {
    std::fclose(f);
    return io_fail(MDSCTK_ERR_FORMAT, "%s: column pointers do not match the file size", path);
  }
  *n_out = n;
  *nnz_out = nnz;
  if (!pcol && !irow && !vals) {
    std::fclose(f);
    return 0;
  }
  if (!pcol || !irow || !vals) {
    std::fclose(f);
    return io_fail(MDSCTK_ERR_ARG, "ssm_read: null output buffer");
  }
  std::memcpy(pcol, pc.data(), pc.size() * sizeof(int32_t));
  const bool ok = std::fread(irow, sizeof(int32_t), (size_t)nnz, f) == (size_t)nnz &&
                  std::fread(vals, sizeof(double), (size_t)nnz, f) == (size_t)nnz;
  std::fclose(f);
  if (!ok) return io_fail(MDSCTK_ERR_FORMAT, "%s: truncated", path);
  return 0;
}

}  // extern "C"

namespace {

// Checks a CSC matrix of `cols` columns whose row indices must lie in
// [0, rows); returns 0 or an error code (text set).
int check_csc(const char* who, int32_t cols, int32_t rows, const int32_t* pcol,
              const int32_t* irow, const double* dist) {
  if (cols <= 0 || !pcol) return io_fail(MDSCTK_ERR_ARG, "%s: empty matrix", who);
  if (pcol[0] != 0) return io_fail(MDSCTK_ERR_FORMAT, "%s: pcol[0] must be 0", who);
  for (int32_t c = 0; c < cols; ++c)
    if (pcol[c + 1] < pcol[c]) return io_fail(MDSCTK_ERR_FORMAT, "%s: pcol decreases", who);
  if (pcol[cols] > 0 && (!irow || !dist)) return io_fail(MDSCTK_ERR_ARG, "%s: null arrays", who);
  for (int32_t c = 0; c < cols; ++c)
    for (int32_t e = pcol[c]; e < pcol[c + 1]; ++e)
      if (irow[e] < 0 || irow[e] >= rows)
        return io_fail(MDSCTK_ERR_FORMAT, "%s: row index %d out of range", who, irow[e]);
  return 0;
}

// Gaussian affinities a = exp(-d^2 / 2 sigma^2) of the stored (strict lower
// triangle) edges, normalised as D^-1/2 A D^-1/2 with the degrees of the
// full symmetric matrix (decomp_sparse.cpp:177-205).  dinv[i] = d_i^-1/2.
void normalised_affinity(int32_t n, const int32_t* pcol, const int32_t* irow, const double* dist,
                         double sigma, std::vector<double>& a, std::vector<double>& dinv) {
  const int64_t nnz = pcol[n];
  a.resize((size_t)nnz);
  dinv.assign(n, 0.0);
  for (int64_t e = 0; e < nnz; ++e) a[e] = std::exp(-(dist[e] * dist[e]) / (2.0 * sigma * sigma));
  for (int32_t c = 0; c < n; ++c)
    for (int32_t e = pcol[c]; e < pcol[c + 1]; ++e) {
      dinv[c] += a[e];
      dinv[irow[e]] += a[e];
    }
  for (int32_t i = 0; i < n; ++i) dinv[i] = 1.0 / std::sqrt(dinv[i]);
  for (int32_t c = 0; c < n; ++c)
    for (int32_t e = pcol[c]; e < pcol[c + 1]; ++e) a[e] *= dinv[irow[e]] * dinv[c];
}

// Leading nev eigenpairs (largest algebraic first) of the symmetric matrix
// held as a strict lower triangle in (pcol, irow, a).  Returns the number of
// converged pairs.
int leading_eigenpairs(int32_t n, const int32_t* pcol, const int32_t* irow,
                       const std::vector<double>& a, int32_t nev, double* evals, double* evecs,
                       double* residuals) {
  // Thick-restart Lanczos.
  const int m = (int)std::min<int64_t>((int64_t)10 * nev + 1, n);
  const double tol = 1e-13;
  std::vector<double> V((size_t)(m + 1) * n), T((size_t)m * m, 0.0), h(m + 1), w(n), part;
  // threads for the vector work; small graphs are not worth them
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  Team team(n >= 4096 ? std::min(hw, kRowBlocks) : 1);
  const SymRows S = sym_rows(n, pcol, irow, a.data());
  uint64_t seed = 0x5eed5eedull;
  auto random_start = [&](int j) {
    // a random vector orthogonal to V[:, 0..j-1], normalised into V[:, j]
    double* v = &V[(size_t)j * n];
    for (int attempt = 0; attempt < 8; ++attempt) {
      for (int32_t r = 0; r < n; ++r) v[r] = rand_unit(seed);
      if (j > 0) reorthogonalise(team, V, n, j - 1, v, h.data(), part);
      const double nv = std::sqrt(dot(v, v, n));
      if (nv > 1e-8) {
        for (int32_t r = 0; r < n; ++r) v[r] /= nv;
        return true;
      }
    }
    return false;
  };
  if (!random_start(0)) return io_fail(MDSCTK_ERR_ARG, "spectral_decomp: no start vector");
  int j0 = 0;              // vectors kept from the previous cycle
  const int max_cycles = 100 * nev;
  std::vector<double> theta, Y;
  int converged = 0;
  double beta = 0.0;
  for (int cycle = 0; cycle < max_cycles; ++cycle) {
    int mm = m;
    for (int j = j0; j < m; ++j) {
      sym_rows_spmv(team, n, S, &V[(size_t)j * n], w.data());
      reorthogonalise(team, V, n, j, w.data(), h.data(), part);
      for (int i = 0; i <= j; ++i) T[(size_t)i * m + j] = T[(size_t)j * m + i] = h[i];
      beta = std::sqrt(dot(w.data(), w.data(), n));
      if (j + 1 == m) break;
      if (beta <= 1e-14) {
        // invariant subspace: continue from a fresh direction
        if (!random_start(j + 1)) {
          mm = j + 1;
          beta = 0.0;
          break;
        }
        T[(size_t)(j + 1) * m + j] = T[(size_t)j * m + j + 1] = 0.0;
      } else {
        double* vn = &V[(size_t)(j + 1) * n];
        for (int32_t r = 0; r < n; ++r) vn[r] = w[r] / beta;
        T[(size_t)(j + 1) * m + j] = T[(size_t)j * m + j + 1] = beta;
      }
    }
    // Rayleigh-Ritz on the projected matrix, Ritz values descending.
    std::vector<double> Tm((size_t)mm * mm);
    for (int i = 0; i < mm; ++i)
      for (int j = 0; j < mm; ++j) Tm[(size_t)i * mm + j] = T[(size_t)i * m + j];
    std::vector<double> wv, vv;
    jacobi_eig(mm, Tm, wv, vv);
    std::vector<int> ord(mm);
    for (int i = 0; i < mm; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return wv[x] > wv[y]; });
    theta.assign(mm, 0.0);
    Y.assign((size_t)mm * mm, 0.0);
    for (int i = 0; i < mm; ++i) {
      theta[i] = wv[ord[i]];
      for (int r = 0; r < mm; ++r) Y[(size_t)r * mm + i] = vv[(size_t)r * mm + ord[i]];
    }
    converged = 0;
    for (int i = 0; i < std::min(nev, mm); ++i) {
      const double res = std::fabs(beta * Y[(size_t)(mm - 1) * mm + i]);
      if (res <= tol * std::max(std::fabs(theta[i]), 1e-10)) ++converged;
      else break;
    }
    if (converged >= nev || mm < m) break;
    // Thick restart: keep the leading k Ritz vectors, continue from the
    // residual direction.
    const int keep = std::min(m - 1, nev + std::max(nev, (m - nev) / 2));
    std::vector<double> U((size_t)keep * n, 0.0);
    team.run([&](int t) {
      for (int b = t; b < kRowBlocks; b += team.size()) {
        const int32_t q0 = block_row(n, b), q1 = block_row(n, b + 1);
        for (int i = 0; i < keep; ++i) {
          double* u = &U[(size_t)i * n];
          for (int r = 0; r < mm; ++r) {
            const double y = Y[(size_t)r * mm + i];
            const double* vr = &V[(size_t)r * n];
            for (int32_t q = q0; q < q1; ++q) u[q] += y * vr[q];
          }
        }
      }
    });
    std::copy(U.begin(), U.end(), V.begin());
    double* vk = &V[(size_t)keep * n];
    for (int32_t r = 0; r < n; ++r) vk[r] = w[r] / beta;
    std::fill(T.begin(), T.end(), 0.0);
    for (int i = 0; i < keep; ++i) {
      T[(size_t)i * m + i] = theta[i];
      const double s = beta * Y[(size_t)(mm - 1) * mm + i];
      T[(size_t)i * m + keep] = T[(size_t)keep * m + i] = s;
    }
    j0 = keep;
  }

  // Ritz vectors and residuals of the leading nev pairs.
  const int mm = (int)theta.size();
  std::vector<double> z(n), az(n);
  for (int i = 0; i < nev && i < mm; ++i) {
    std::fill(z.begin(), z.end(), 0.0);
    for (int r = 0; r < mm; ++r) {
      const double y = Y[(size_t)r * mm + i];
      const double* vr = &V[(size_t)r * n];
      for (int32_t q = 0; q < n; ++q) z[q] += y * vr[q];
    }
    const double nz = std::sqrt(dot(z.data(), z.data(), n));
    for (int32_t q = 0; q < n; ++q) z[q] /= nz;
    evals[i] = theta[i];
    if (evecs) std::copy(z.begin(), z.end(), evecs + (size_t)i * n);
    if (residuals) {
      sym_spmv(n, pcol, irow, a.data(), z.data(), az.data());
      double rr = 0.0;
      for (int32_t q = 0; q < n; ++q) {
        const double d = az[q] - theta[i] * z[q];
        rr += d * d;
      }
      residuals[i] = std::sqrt(rr) / std::fabs(theta[i]);
    }
  }
  return converged >= nev ? nev : converged;
}

}  // namespace

extern "C" {

int mdsctk_spectral_decomp(int32_t n, const int32_t* pcol, const int32_t* irow,
                           const double* dist, double sigma, int32_t nev, double* evals,
                           double* evecs, double* residuals) {
  if (n <= 0 || !pcol || !evals || nev <= 0 || nev > n || !(sigma > 0))
    return io_fail(MDSCTK_ERR_ARG, "spectral_decomp: need n > 0, 0 < nev <= n, sigma > 0");
  if (int rc = check_csc("spectral_decomp", n, n, pcol, irow, dist)) return rc;
  std::vector<double> a, dinv;
  normalised_affinity(n, pcol, irow, dist, sigma, a, dinv);
  return leading_eigenpairs(n, pcol, irow, a, nev, evals, evecs, residuals);
}

int mdsctk_gsm_build(const double* dist, const int32_t* idx, int64_t n_rows, int32_t maxk,
                     int32_t k, int symmetric, int32_t* pcol, int32_t* irow, double* vals,
                     int64_t* nnz) {
  if (!dist || !idx || !nnz || n_rows < 0 || n_rows > INT32_MAX || maxk <= 0 || k <= 0)
    return io_fail(MDSCTK_ERR_ARG, "gsm_build: need rows, maxk > 0 and k > 0");
  if (k > maxk)
    return io_fail(MDSCTK_ERR_ARG, "gsm_build: output k (%d) is larger than the input k (%d)", k,
                   maxk);
  // kind 0: the row's own entry (a later one replaces an earlier one);
  // kind 1: the mirrored entry of -s, which only fills an edge that has no
  // entry yet, so the row's own entry always wins and the first mirror does.
  struct GEdge {
    int32_t from, to;
    int32_t kind;
    int64_t order;
    double d;
  };
  std::vector<GEdge> edges;
  edges.reserve((size_t)n_rows * k * (symmetric ? 2 : 1));
  int64_t order = 0;
  for (int64_t i = 0; i < n_rows; ++i)
    for (int32_t x = 0; x < k; ++x) {
      const int32_t j = idx[i * maxk + x];
      const double d = dist[i * maxk + x];
      if (j < 0)
        return io_fail(MDSCTK_ERR_ARG, "gsm_build: row %lld holds a negative index %d",
                       (long long)i, j);
      edges.push_back(GEdge{(int32_t)i, j, 0, order++, d});
      // columns are query rows: a mirrored edge from a reference index
      // beyond the last row has no column to live in
      if (symmetric && j < n_rows) edges.push_back(GEdge{j, (int32_t)i, 1, order++, d});
    }
  std::sort(edges.begin(), edges.end(), [](const GEdge& a, const GEdge& b) {
    if (a.from != b.from) return a.from < b.from;
    if (a.to != b.to) return a.to < b.to;
    if (a.kind != b.kind) return a.kind < b.kind;
    return a.kind == 0 ? a.order > b.order : a.order < b.order;
  });
  size_t m = 0;
  for (size_t e = 0; e < edges.size(); ++e)
    if (m == 0 || edges[m - 1].from != edges[e].from || edges[m - 1].to != edges[e].to)
      edges[m++] = edges[e];
  edges.resize(m);
  *nnz = (int64_t)m;
  if (!pcol && !irow && !vals) return 0;
  if (!pcol || !irow || !vals) return io_fail(MDSCTK_ERR_ARG, "gsm_build: null output buffer");
  if (m > (size_t)INT32_MAX) return io_fail(MDSCTK_ERR_ARG, "gsm_build: too many edges");
  std::fill(pcol, pcol + n_rows + 1, 0);
  for (const GEdge& e : edges) ++pcol[e.from + 1];
  for (int64_t c = 0; c < n_rows; ++c) pcol[c + 1] += pcol[c];
  for (size_t e = 0; e < m; ++e) {
    irow[e] = edges[e].to;
    vals[e] = edges[e].d;
  }
  return 0;
}

int mdsctk_nystrom_decomp(int32_t n, const int32_t* pcol_a, const int32_t* irow_a,
                          const double* dist_a, int32_t m, const int32_t* pcol_b,
                          const int32_t* irow_b, const double* dist_b, double sigma, int32_t nev,
                          double* evals, double* evecs, double* ext, double* residuals) {
  if (n <= 0 || m <= 0 || !evals || !ext || nev <= 0 || nev > n || !(sigma > 0))
    return io_fail(MDSCTK_ERR_ARG,
                   "nystrom_decomp: need n > 0, m > 0, 0 < nev <= n, sigma > 0, ext buffer");
  if (int rc = check_csc("nystrom_decomp (ssm)", n, n, pcol_a, irow_a, dist_a)) return rc;
  if (int rc = check_csc("nystrom_decomp (gsm)", m, n, pcol_b, irow_b, dist_b)) return rc;
  std::vector<double> a, dinv;
  normalised_affinity(n, pcol_a, irow_a, dist_a, sigma, a, dinv);
  std::vector<double> z_own;
  if (!evecs) {
    z_own.resize((size_t)nev * n);
    evecs = z_own.data();
  }
  const int got = leading_eigenpairs(n, pcol_a, irow_a, a, nev, evals, evecs, residuals);
  // Out-of-sample affinities, normalised by the in-sample degrees of their
  // reference rows and their own column sums (decomp_sparse_nystrom.cpp:
  // 182-215), then ext = B^T (z / l) column by column (:297-304).
  const int64_t nnz_b = pcol_b[m];
  std::vector<double> b((size_t)nnz_b);
  for (int64_t e = 0; e < nnz_b; ++e)
    b[e] = std::exp(-(dist_b[e] * dist_b[e]) / (2.0 * sigma * sigma));
  for (int32_t c = 0; c < m; ++c) {
    double sum = 0.0;
    for (int32_t e = pcol_b[c]; e < pcol_b[c + 1]; ++e) sum += b[e];
    const double dc = 1.0 / std::sqrt(sum);
    for (int32_t e = pcol_b[c]; e < pcol_b[c + 1]; ++e) b[e] *= dinv[irow_b[e]] * dc;
  }
  std::vector<double> zl(n);
  for (int x = 0; x < nev && x < got; ++x) {
    const double* z = evecs + (size_t)x * n;
    for (int32_t r = 0; r < n; ++r) zl[r] = z[r] / evals[x];
    double* y = ext + (size_t)x * m;
    for (int32_t c = 0; c < m; ++c) {
      double acc = 0.0;
      for (int32_t e = pcol_b[c]; e < pcol_b[c + 1]; ++e) acc += zl[irow_b[e]] * b[e];
      y[c] = acc;
    }
  }
  return got;
}

}  // extern "C"
