// make_gesparse: out-of-sample kNN rows (distances.dat / indices.dat from
// knn_rms -f) to the general CSC file (.gsm), one column per query row.
// Options, defaults, messages and exit codes of make_gesparse.cpp:60-105;
// the matrix is built by mdsctk_gsm_build (include/mdsctk_graph.h).
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
    {"knn", 'k', true, "Input:  K-nearest neighbors (int)"},
    {"output-knn", 'n', true, "Input:  K-nn to keep in output (int)"},
    {"symmetric", 's', false, "Input:  Enforce symmetry (bool)"},
    {"index-file", 'i', true, "Input:  Index file (string:filename)"},
    {"distance-file", 'd', true, "Input:  Distances file (string:filename)"},
    {"output-file", 'o', true, "Output: General sparse matrix file (string:filename)"},
};
}  // namespace

int main(int argc, char** argv) {
  std::printf("\n   make_gesparse (MDSCTK make_gesparse)\n");
  std::printf("   Converts the results from knn_* into general CSC format.\n\n");
  std::map<std::string, std::string> v;
  std::string err;
  if (!mdsctk_cli::parse(argc, argv, kOpts, v, err)) {
    std::fprintf(stderr, "ERROR: %s\n", err.c_str());
    return 1;
  }
  if (v.count("help")) {
    mdsctk_cli::usage("make_gesparse", kOpts);
    return 1;
  }
  int maxk = 0, k = 0;
  bool ok = true;
  if (!v.count("knn")) {
    std::printf("ERROR: --knn not supplied.\n\n");
    ok = false;
  } else if (!mdsctk_cli::parse_int(v["knn"], &maxk) || maxk <= 0) {
    std::printf("ERROR: bad --knn value '%s'.\n\n", v["knn"].c_str());
    ok = false;
  }
  k = maxk;
  if (v.count("output-knn") && (!mdsctk_cli::parse_int(v["output-knn"], &k) || k <= 0)) {
    std::printf("ERROR: bad --output-knn value '%s'.\n\n", v["output-knn"].c_str());
    ok = false;
  }
  if (ok && maxk < k) {
    std::printf("ERROR: Output k (%d) is not less than the input k (%d).\n\n", k, maxk);
    ok = false;
  }
  if (!ok) return -1;  // exit status 255, as make_gesparse.cpp:103-105
  const bool symmetric = v.count("symmetric") > 0;
  const std::string i_file = v.count("index-file") ? v["index-file"] : "indices.dat";
  const std::string d_file = v.count("distance-file") ? v["distance-file"] : "distances.dat";
  const std::string o_file = v.count("output-file") ? v["output-file"] : "distances.gsm";
  std::printf("Running with the following options:\n");
  std::printf("knn =            %d\n", maxk);
  std::printf("output-knn =     %d\n", k);
  std::printf("index-file =     %s\n", i_file.c_str());
  std::printf("distances-file = %s\n", d_file.c_str());
  std::printf("output-file =    %s\n\n", o_file.c_str());

  std::vector<char> dbytes, ibytes;
  if (!mdsctk_cli::read_all(d_file, dbytes)) {
    std::printf("***ERROR***\nCould not open file: %s\n\n", d_file.c_str());
    return -1;
  }
  if (!mdsctk_cli::read_all(i_file, ibytes)) {
    std::printf("***ERROR***\nCould not open file: %s\n\n", i_file.c_str());
    return -1;
  }
  // Whole rows only, like the reference's read loop (make_gesparse.cpp:226-279).
  const int64_t rows_i = (int64_t)(ibytes.size() / (sizeof(int32_t) * maxk));
  const int64_t rows_d = (int64_t)(dbytes.size() / (sizeof(double) * maxk));
  const int64_t n = rows_i < rows_d ? rows_i : rows_d;
  std::printf("Creating sparse matrix...\n");
  int64_t nnz = 0;
  const double* dist = (const double*)dbytes.data();
  const int32_t* idx = (const int32_t*)ibytes.data();
  const int sym = symmetric ? 1 : 0;
  if (mdsctk_gsm_build(dist, idx, n, maxk, k, sym, nullptr, nullptr, nullptr, &nnz) < 0) {
    std::printf("ERROR: %s\n", mdsctk_io_last_error());
    return 2;
  }
  std::vector<int32_t> pcol((size_t)n + 1), irow((size_t)nnz);
  std::vector<double> vals((size_t)nnz);
  if (mdsctk_gsm_build(dist, idx, n, maxk, k, sym, pcol.data(), irow.data(), vals.data(), &nnz) <
          0 ||
      mdsctk_ssm_write(o_file.c_str(), (int32_t)n, pcol.data(), irow.data(), vals.data()) < 0) {
    std::printf("ERROR: %s\n", mdsctk_io_last_error());
    return 2;
  }
  std::printf("%lld rows, %lld stored entries -> %s\n\n", (long long)n, (long long)nnz,
              o_file.c_str());
  return 0;
}
