// knn_data: drop-in replacement for MDSCTK's knn_data executable on MI355X.
//
// Same options, defaults, outputs and exit codes as knn_data.cpp:43-263:
//   -h/--help (exit 1), -t/--threads, -k/--knn (required when sorting),
//   -s/--sort, -v/--vector-size (required; either missing exits 255),
//   -b/--block-size, -c/--correlation (switch), -r/--reference-file
//   (default reference.pts), -f/--fit-file (default: the reference file),
//   -d/--distance-file, -i/--index-file.
// Inputs are headerless FP64 vectors of --vector-size values (whole vectors
// only, as the reference's read loop keeps them).  Outputs: distances
// (float64 [n_fit][k], Euclidean or correlation distance, rank 0 dropped)
// and indices (int32 [n_fit][k], 0-based); with --sort false the full
// unsorted rows [n_fit][n_ref] and identity indices.
// Options added here (no reference counterpart): --precision
// {bf16,fp32,exact} (contraction form, default bf16) and --device N.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "cli_options.h"
#include "mdsctk_knn.h"

namespace {

using mdsctk_cli::Opt;

const std::vector<Opt> kOpts = {
    {"help", 'h', false, "show this help message and exit"},
    {"threads", 't', true, "Input:  Number of threads to start (int)"},
    {"knn", 'k', true, "Input:  K-nearest neighbors (int)"},
    {"sort", 's', true, "Input:  Find K-nn,false=full distance matix (bool)"},
    {"vector-size", 'v', true, "Input:  Data vector length (int)"},
    {"block-size", 'b', true, "Input:  Workgroup block size in # frames (int)"},
    {"correlation", 'c', false, "Input:  Use correlation distance (bool)"},
    {"reference-file", 'r', true, "Input:  Reference data file (string:filename)"},
    {"fit-file", 'f', true, "Input:  Fitting data file (string:filename)"},
    {"distance-file", 'd', true, "Output: K-nn distances file (string:filename)"},
    {"index-file", 'i', true, "Output: K-nn indices file (string:filename)"},
    {"precision", 0, true, "Input:  Contraction form bf16|fp32|exact (string)"},
    {"device", 0, true, "Input:  HIP device index (int)"},
};

// Whole vectors of vdim doubles from a .pts file (knn_data.cpp:150-175).
bool read_vectors(const std::string& path, int vdim, std::vector<double>& out, int64_t* n) {
  std::vector<char> bytes;
  if (!mdsctk_cli::read_all(path, bytes)) return false;
  *n = (int64_t)(bytes.size() / (sizeof(double) * (size_t)vdim));
  out.resize((size_t)*n * vdim);
  if (*n > 0) std::copy_n((const double*)bytes.data(), out.size(), out.data());
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  std::printf("\n   knn_data (MDSCTK knn_data on MI355X)\n");
  std::printf("   Computes the k nearest neighbors of all pairs of\n");
  std::printf("   vectors in the given binary data files.\n\n");
  std::printf("   Use -h or --help to see the complete list of options.\n\n");

  std::map<std::string, std::string> v;
  std::string perr;
  if (!mdsctk_cli::parse(argc, argv, kOpts, v, perr)) {
    std::fprintf(stderr, "ERROR: %s\n", perr.c_str());
    return 1;
  }
  if (v.count("help")) {
    mdsctk_cli::usage("knn_data", kOpts);
    return 1;
  }
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  int nthreads = (int)hw, k = 0, vdim = 0, blksize = 128, device = 0;
  bool sort = true;
  std::string reff = "reference.pts", fitf, df = "distances.dat", inf = "indices.dat",
              prec = "bf16";
  auto bad = [](const char* o, const std::string& val) {
    std::fprintf(stderr, "ERROR: invalid value '%s' for option --%s\n", val.c_str(), o);
    return 1;
  };
  if (v.count("threads") && !mdsctk_cli::parse_int(v["threads"], &nthreads))
    return bad("threads", v["threads"]);
  if (v.count("knn") && !mdsctk_cli::parse_int(v["knn"], &k)) return bad("knn", v["knn"]);
  if (v.count("sort") && !mdsctk_cli::parse_bool(v["sort"], &sort)) return bad("sort", v["sort"]);
  if (v.count("vector-size") && !mdsctk_cli::parse_int(v["vector-size"], &vdim))
    return bad("vector-size", v["vector-size"]);
  if (v.count("block-size") && !mdsctk_cli::parse_int(v["block-size"], &blksize))
    return bad("block-size", v["block-size"]);
  if (v.count("device") && !mdsctk_cli::parse_int(v["device"], &device))
    return bad("device", v["device"]);
  const bool corr = v.count("correlation") > 0;
  if (v.count("reference-file")) reff = v["reference-file"];
  if (v.count("distance-file")) df = v["distance-file"];
  if (v.count("index-file")) inf = v["index-file"];
  if (v.count("precision")) prec = v["precision"];
  bool ok = true;
  if (!v.count("knn") && sort) {
    std::printf("ERROR: --knn not supplied.\n\n");
    ok = false;
  }
  if (!v.count("vector-size")) {
    std::printf("ERROR: --vector-size not supplied.\n\n");
    ok = false;
  }
  if (!ok) return -1;  // knn_data.cpp:94-111 (exit status 255)
  if (vdim <= 0) return bad("vector-size", v["vector-size"]);
  const bool self = !v.count("fit-file") || v["fit-file"] == reff;
  fitf = v.count("fit-file") ? v["fit-file"] : reff;
  int precision;
  if (prec == "bf16") precision = MDSCTK_PREC_SPLIT_BF16;
  else if (prec == "fp32") precision = MDSCTK_PREC_FP32;
  else if (prec == "exact") precision = MDSCTK_PREC_EXACT_FP64;
  else return bad("precision", prec);

  std::printf("Running with the following options:\n");
  std::printf("threads =        %d\n", nthreads);
  std::printf("knn =            %d\n", k);
  std::printf("sort =           %d\n", sort ? 1 : 0);
  std::printf("vector-size =    %d\n", vdim);
  std::printf("correlation =    %d\n", corr ? 1 : 0);
  std::printf("reference-file = %s\n", reff.c_str());
  std::printf("fit-file =       %s\n", fitf.c_str());
  std::printf("distance-file =  %s\n", df.c_str());
  std::printf("index-file =     %s\n", inf.c_str());
  std::printf("precision =      %s\n\n", prec.c_str());
  std::fflush(stdout);

  std::vector<double> ref, fit;
  int64_t n_ref = 0, n_fit = 0;
  std::printf("Reading reference coordinates from file: %s ... ", reff.c_str());
  if (!read_vectors(reff, vdim, ref, &n_ref)) {
    std::printf("\nERROR: cannot read %s\n", reff.c_str());
    return 3;
  }
  std::printf("done.\nNumber of reference coordinates: %lld\n", (long long)n_ref);
  if (!self) {
    std::printf("Reading fitting coordinates from file: %s ... ", fitf.c_str());
    if (!read_vectors(fitf, vdim, fit, &n_fit)) {
      std::printf("\nERROR: cannot read %s\n", fitf.c_str());
      return 3;
    }
    std::printf("done.\n");
  } else {
    n_fit = n_ref;
  }
  std::printf("Number of fitting coordinates: %lld\n", (long long)n_fit);
  if (n_ref == 0) {
    std::printf("ERROR: %s holds no vectors.\n", reff.c_str());
    return 3;
  }
  std::printf("Block size: %d\n", std::max(1, (int)std::min<int64_t>(blksize, n_fit)));

  mdsctk_knn_ctx* ctx = nullptr;
  int rc = mdsctk_knn_create(&ctx, device, precision);
  if (rc < 0) {
    std::printf("ERROR: cannot create the GPU context on device %d (code %d)\n", device, rc);
    return 2;
  }
  const auto t0 = std::chrono::steady_clock::now();
  rc = mdsctk_data_set_reference(ctx, ref.data(), n_ref, vdim,
                                 corr ? MDSCTK_METRIC_CORRELATION : MDSCTK_METRIC_EUCLIDEAN);
  if (rc < 0) {
    std::printf("ERROR: %s\n", mdsctk_knn_last_error(ctx));
    mdsctk_knn_destroy(ctx);
    return 2;
  }
  const uint32_t flags = sort ? 0u : MDSCTK_KNN_FULL_ROWS;
  const int width = mdsctk_knn_row_width(ctx, k, flags);
  std::vector<double> dist((size_t)n_fit * std::max(width, 0));
  std::vector<int32_t> idx((size_t)n_fit * std::max(width, 0));
  rc = mdsctk_data_query(ctx, self ? nullptr : fit.data(), n_fit, k, flags, dist.data(),
                         idx.data());
  if (rc < 0) {
    std::printf("ERROR: %s\n", mdsctk_knn_last_error(ctx));
    mdsctk_knn_destroy(ctx);
    return 2;
  }
  mdsctk_knn_stats st;
  mdsctk_knn_get_stats(ctx, &st);
  mdsctk_knn_destroy(ctx);
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (!mdsctk_cli::write_all(df, dist.data(), dist.size() * sizeof(double)) ||
      !mdsctk_cli::write_all(inf, idx.data(), idx.size() * sizeof(int32_t))) {
    std::printf("ERROR: cannot write %s / %s\n", df.c_str(), inf.c_str());
    return 3;
  }
  std::printf("\n%lld x %lld pairs, k=%d, %s: %.3f s (GPU %.3f ms, %.3g pairs/s), "
              "%lld candidates re-scored, %lld exact rows\n\n",
              (long long)n_fit, (long long)n_ref, width, prec.c_str(), secs, st.ms_total,
              st.ms_total > 0 ? (double)st.n_pairs / (st.ms_total * 1e-3) : 0.0,
              (long long)st.n_candidates, (long long)st.n_fallback_rows);
  return 0;
}
