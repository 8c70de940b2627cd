// decomp_sparse_nystrom: leading eigenpairs of the normalised Gaussian
// affinity of an .ssm kNN graph and their Nystrom extension to the
// out-of-sample rows of a .gsm matrix.  Options, defaults, outputs (text,
// largest eigenvalue first; each eigenvector line holds the n in-sample
// values then the m extended ones) and exit codes of
// decomp_sparse_nystrom.cpp:60-320; the work is mdsctk_nystrom_decomp
// (include/mdsctk_graph.h).
#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "cli_options.h"
#include "mdsctk_graph.h"
#include "mdsctk_io.h"

namespace {
const std::vector<mdsctk_cli::Opt> kOpts = {
    {"help", 'h', false, "show this help message and exit"},
    {"sigma", 'q', true, "Input:  Standard deviation of gaussian kernel (real)"},
    {"nevals", 'n', true, "Input:  Number of eigenvalues/vectors (int)"},
    {"ssm-file", 's', true, "Input:  Symmetric sparse matrix file (string:filename)"},
    {"gsm-file", 'g', true, "Input:  General sparse matrix file (string:filename)"},
    {"evals-file", 'v', true, "Output:  Eigenvalues file (string:filename)"},
    {"evecs-file", 'e', true, "Output: Eigenvectors file (string:filename)"},
    {"residuals-file", 'r', true, "Output: Residuals file (string:filename)"},
};

bool read_csc(const std::string& path, int32_t& n, std::vector<int32_t>& pcol,
              std::vector<int32_t>& irow, std::vector<double>& vals) {
  int64_t nnz = 0;
  if (mdsctk_ssm_read(path.c_str(), &n, &nnz, nullptr, nullptr, nullptr) < 0) return false;
  pcol.resize((size_t)n + 1);
  irow.resize((size_t)nnz);
  vals.resize((size_t)nnz);
  return mdsctk_ssm_read(path.c_str(), &n, &nnz, pcol.data(), irow.data(), vals.data()) == 0;
}
}  // namespace

int main(int argc, char** argv) {
  std::printf("\n   decomp_sparse_nystrom (MDSCTK decomp_sparse_nystrom)\n");
  std::printf("   Computes the eigenvalues and eigenvectors of the normalised\n");
  std::printf("   affinity of a symmetric sparse distance matrix and projects the\n");
  std::printf("   general sparse matrix onto them for out-of-sample prediction.\n\n");
  std::map<std::string, std::string> v;
  std::string err;
  if (!mdsctk_cli::parse(argc, argv, kOpts, v, err)) {
    std::fprintf(stderr, "ERROR: %s\n", err.c_str());
    return 1;
  }
  if (v.count("help")) {
    mdsctk_cli::usage("decomp_sparse_nystrom", kOpts);
    return 1;
  }
  double sigma = 0;
  int nev = 0;
  bool ok = true;
  if (!v.count("sigma")) {
    std::printf("ERROR: --sigma not supplied.\n\n");
    ok = false;
  } else if (!mdsctk_cli::parse_double(v["sigma"], &sigma) || !(sigma > 0)) {
    std::printf("ERROR: bad --sigma value '%s'.\n\n", v["sigma"].c_str());
    ok = false;
  }
  if (!v.count("nevals")) {
    std::printf("ERROR: --nevals not supplied.\n\n");
    ok = false;
  } else if (!mdsctk_cli::parse_int(v["nevals"], &nev) || nev <= 0) {
    std::printf("ERROR: bad --nevals value '%s'.\n\n", v["nevals"].c_str());
    ok = false;
  }
  if (!ok) return -1;  // exit status 255, as decomp_sparse_nystrom.cpp:98-100
  const std::string ssm = v.count("ssm-file") ? v["ssm-file"] : "distances.ssm";
  const std::string gsm = v.count("gsm-file") ? v["gsm-file"] : "distances.gsm";
  const std::string fv = v.count("evals-file") ? v["evals-file"] : "eigenvalues.dat";
  const std::string fe = v.count("evecs-file") ? v["evecs-file"] : "eigenvectors.dat";
  const std::string fr = v.count("residuals-file") ? v["residuals-file"] : "residuals.dat";
  std::printf("Running with the following options:\n");
  std::printf("sigma =          %g\n", sigma);
  std::printf("nevals =         %d\n", nev);
  std::printf("ssm-file =       %s\n", ssm.c_str());
  std::printf("gsm-file =       %s\n", gsm.c_str());
  std::printf("evals-file =     %s\n", fv.c_str());
  std::printf("evecs-file =     %s\n", fe.c_str());
  std::printf("residuals-file = %s\n\n", fr.c_str());

  int32_t n = 0, m = 0;
  std::vector<int32_t> pa, ia, pb, ib;
  std::vector<double> va, vb;
  if (!read_csc(ssm, n, pa, ia, va) || !read_csc(gsm, m, pb, ib, vb)) {
    std::printf("ERROR: %s\n", mdsctk_io_last_error());
    return 3;
  }
  if (nev > n) {
    std::printf("ERROR: --nevals (%d) exceeds the matrix dimension (%d).\n", nev, n);
    return -1;
  }
  std::vector<double> ev(nev), evec((size_t)nev * n), ext((size_t)nev * m), res(nev);
  const int got = mdsctk_nystrom_decomp(n, pa.data(), ia.data(), va.data(), m, pb.data(),
                                        ib.data(), vb.data(), sigma, nev, ev.data(), evec.data(),
                                        ext.data(), res.data());
  if (got < 0) {
    std::printf("ERROR: %s\n", mdsctk_io_last_error());
    return 2;
  }
  std::printf("Number of converged eigenvalues/vectors found: %d\n", got);
  FILE* a = std::fopen(fv.c_str(), "w");
  FILE* b = std::fopen(fe.c_str(), "w");
  FILE* c = std::fopen(fr.c_str(), "w");
  if (!a || !b || !c) {
    std::printf("ERROR: cannot open an output file\n");
    return 3;
  }
  double max_res = 0.0;
  for (int x = 0; x < nev; ++x) {
    std::fprintf(a, "%.17g\n", ev[x]);
    for (int32_t y = 0; y < n; ++y) std::fprintf(b, "%.17g ", evec[(size_t)x * n + y]);
    for (int32_t y = 0; y < m; ++y) std::fprintf(b, "%.17g ", ext[(size_t)x * m + y]);
    std::fprintf(b, "\n");
    std::fprintf(c, "%.17g\n", res[x]);
    if (res[x] > max_res) max_res = res[x];
  }
  std::fclose(a);
  std::fclose(b);
  std::fclose(c);
  // eps as decomp_sparse_nystrom.cpp:145-147: the square root of the unit roundoff
  double eps = 1.0;
  do { eps /= 2.0; } while (1.0 + (eps / 2.0) != 1.0);
  eps = std::sqrt(eps);
  std::printf("Max residual: %g (eps: %g)\n", max_res, eps);
  if (max_res > eps) {
    std::printf("*** Sum of residuals too high (max_r > eps)!\n");
    std::printf("*** Please, check results manually...\n");
  }
  return got == nev ? 0 : 2;
}
