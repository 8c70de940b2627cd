// knn_rms: drop-in replacement for MDSCTK's knn_rms executable on MI355X.
//
// Same options, defaults, outputs and exit codes as knn_rms.cpp:43-305:
//   -h/--help (exit 1), -t/--threads, -k/--knn (required when sorting,
//   exit 255 otherwise), -s/--sort, -n/--nofit, -b/--block-size,
//   -p/--topology-file, -r/--reference-file, -f/--fit-file,
//   -d/--distance-file, -i/--index-file; an atom-count mismatch between the
//   topology and an XTC file exits with 4 (knn_rms.cpp:158-178).
// Outputs: distances (float64 [n_fit][k], Angstrom, rank 0 dropped) and
// indices (int32 [n_fit][k], 0-based), headerless; with --sort false the
// full unsorted rows [n_fit][n_ref] and identity indices.
// Options added here (no reference counterpart): --precision
// {bf16,fp16,fp32,exact} (contraction form, default bf16) and --device N.
// The pair computation runs on the GPU through libmdsctk_knn; -t sets the
// XTC decode threads and -b is accepted and reported (the GPU tiles the
// work itself).
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "cli_options.h"
#include "mdsctk_io.h"
#include "mdsctk_knn.h"

namespace {

using mdsctk_cli::Opt;
using mdsctk_cli::parse_bool;
using mdsctk_cli::parse_int;
using mdsctk_cli::write_all;

const std::vector<Opt> kOpts = {
    {"help", 'h', false, "show this help message and exit"},
    {"threads", 't', true, "Input:  Number of threads to start (int)"},
    {"knn", 'k', true, "Input:  K-nearest neighbors (int)"},
    {"sort", 's', true, "Input:  Find K-nn,false=full distance matix (bool)"},
    {"nofit", 'n', true, "Input:  DO NOT rotate+fit before RMSD calculation (bool)"},
    {"block-size", 'b', true, "Input:  Workgroup block size in # frames (int)"},
    {"topology-file", 'p', true, "Input:  Topology file [.pdb,.gro] (string:filename)"},
    {"reference-file", 'r', true, "Input:  Reference [.xtc] file (string:filename)"},
    {"fit-file", 'f', true, "Input:  Fitting [.xtc] file (string:filename)"},
    {"distance-file", 'd', true, "Output: K-nn distances file (string:filename)"},
    {"index-file", 'i', true, "Output: K-nn indices file (string:filename)"},
    {"precision", 0, true, "Input:  Contraction form bf16|fp16|fp32|exact (string)"},
    {"device", 0, true, "Input:  HIP device index (int)"},
};

void usage() { mdsctk_cli::usage("knn_rms", kOpts); }

bool parse(int argc, char** argv, std::map<std::string, std::string>& vals, std::string& err) {
  return mdsctk_cli::parse(argc, argv, kOpts, vals, err);
}

}  // namespace

int main(int argc, char** argv) {
  std::printf("\n   knn_rms (MDSCTK knn_rms on MI355X)\n");
  std::printf("   Computes the k nearest neighbors of all reference\n");
  std::printf("   structures in the given xtc file for each structure\n");
  std::printf("   in the given fitting xtc file. (Uses the same file for\n");
  std::printf("   as reference by default to make a symmetric comparison.)\n");
  std::printf("   A topology PDB file should be provided for determining\n");
  std::printf("   the mass of each atom.\n\n");
  std::printf("   Use -h or --help to see the complete list of options.\n\n");

  std::map<std::string, std::string> v;
  std::string perr;
  if (!parse(argc, argv, v, perr)) {
    std::fprintf(stderr, "ERROR: %s\n", perr.c_str());
    return 1;
  }
  if (v.count("help")) {
    usage();
    return 1;
  }
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  int nthreads = (int)hw, k = 0, blksize = 128, device = 0;
  bool sort = true, nofit = false;
  std::string top = "topology.pdb", reff = "reference.xtc", fitf, df = "distances.dat",
              inf = "indices.dat", prec = "bf16";
  auto bad = [](const char* o, const std::string& val) {
    std::fprintf(stderr, "ERROR: invalid value '%s' for option --%s\n", val.c_str(), o);
    return 1;
  };
  if (v.count("threads") && !parse_int(v["threads"], &nthreads)) return bad("threads", v["threads"]);
  if (v.count("knn") && !parse_int(v["knn"], &k)) return bad("knn", v["knn"]);
  if (v.count("sort") && !parse_bool(v["sort"], &sort)) return bad("sort", v["sort"]);
  if (v.count("nofit") && !parse_bool(v["nofit"], &nofit)) return bad("nofit", v["nofit"]);
  if (v.count("block-size") && !parse_int(v["block-size"], &blksize)) return bad("block-size", v["block-size"]);
  if (v.count("device") && !parse_int(v["device"], &device)) return bad("device", v["device"]);
  if (v.count("topology-file")) top = v["topology-file"];
  if (v.count("reference-file")) reff = v["reference-file"];
  if (v.count("distance-file")) df = v["distance-file"];
  if (v.count("index-file")) inf = v["index-file"];
  if (v.count("precision")) prec = v["precision"];
  if (!v.count("knn") && sort) {
    std::printf("ERROR: --knn not supplied.\n\n");
    return -1;  // knn_rms.cpp:98-108 (exit status 255)
  }
  const bool self = !v.count("fit-file") || v["fit-file"] == reff;
  fitf = v.count("fit-file") ? v["fit-file"] : reff;
  int precision;
  if (prec == "bf16") precision = MDSCTK_PREC_SPLIT_BF16;
  else if (prec == "fp16") precision = MDSCTK_PREC_SPLIT_FP16;
  else if (prec == "fp32") precision = MDSCTK_PREC_FP32;
  else if (prec == "exact") precision = MDSCTK_PREC_EXACT_FP64;
  else return bad("precision", prec);

  std::printf("Running with the following options:\n");
  std::printf("threads =        %d\n", nthreads);
  std::printf("knn =            %d\n", k);
  std::printf("sort =           %d\n", sort ? 1 : 0);
  std::printf("nofit =          %d\n", nofit ? 1 : 0);
  std::printf("topology-file =  %s\n", top.c_str());
  std::printf("reference-file = %s\n", reff.c_str());
  std::printf("fit-file =       %s\n", fitf.c_str());
  std::printf("distance-file =  %s\n", df.c_str());
  std::printf("index-file =     %s\n", inf.c_str());
  std::printf("precision =      %s\n\n", prec.c_str());
  std::fflush(stdout);

  std::printf("Reading topology information from %s ... ", top.c_str());
  int32_t natoms = 0;
  if (mdsctk_topology_read(top.c_str(), &natoms, nullptr, 0) < 0) {
    std::printf("\nERROR: %s\n", mdsctk_io_last_error());
    return 3;
  }
  std::vector<float> masses(natoms);
  if (mdsctk_topology_read(top.c_str(), &natoms, masses.data(), natoms) < 0) {
    std::printf("\nERROR: %s\n", mdsctk_io_last_error());
    return 3;
  }
  std::printf("done.\n");

  auto load = [&](const std::string& path, std::vector<float>& xyz, int64_t* n) -> int {
    int64_t nf = 0;
    int32_t na = 0;
    float pr = 0;
    if (mdsctk_xtc_info(path.c_str(), &nf, &na, &pr) < 0) {
      std::printf("\nERROR: %s\n", mdsctk_io_last_error());
      return 3;
    }
    if (na != natoms) {
      std::printf("*** ERROR ***\nNumber of atoms in topology file (%d) does not match the number "
                  "of atoms in the XTC file (%s : %d).\n", natoms, path.c_str(), na);
      return 4;
    }
    xyz.resize((size_t)nf * natoms * 3);
    int64_t got = 0;
    if (mdsctk_xtc_read(path.c_str(), xyz.data(), nf, natoms, &got, nullptr, nullptr, nullptr,
                        nthreads) < 0) {
      std::printf("\nERROR: %s\n", mdsctk_io_last_error());
      return 3;
    }
    *n = got;
    return 0;
  };

  std::vector<float> ref_xyz, fit_xyz;
  int64_t n_ref = 0, n_fit = 0;
  std::printf("Reading reference coordinates from file: %s ... ", reff.c_str());
  std::fflush(stdout);
  if (int rc = load(reff, ref_xyz, &n_ref)) return rc;
  std::printf("done.\n");
  if (!self) {
    std::printf("Reading fitting coordinates from file: %s ... ", fitf.c_str());
    std::fflush(stdout);
    if (int rc = load(fitf, fit_xyz, &n_fit)) return rc;
    std::printf("done.\n");
  } else {
    n_fit = n_ref;
  }
  if (n_ref == 0) {
    std::printf("ERROR: %s holds no frames.\n", reff.c_str());
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
  rc = mdsctk_knn_set_reference(ctx, ref_xyz.data(), n_ref, natoms, masses.data());
  if (rc < 0) {
    std::printf("ERROR: %s\n", mdsctk_knn_last_error(ctx));
    mdsctk_knn_destroy(ctx);
    return 2;
  }
  uint32_t flags = 0;
  if (nofit) flags |= MDSCTK_KNN_NOFIT;
  if (!sort) flags |= MDSCTK_KNN_FULL_ROWS;
  const int width = mdsctk_knn_row_width(ctx, k, flags);
  std::vector<double> dist((size_t)n_fit * std::max(width, 0));
  std::vector<int32_t> idx((size_t)n_fit * std::max(width, 0));
  rc = mdsctk_knn_query(ctx, self ? nullptr : fit_xyz.data(), n_fit, k, flags, dist.data(),
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

  if (!write_all(df, dist.data(), dist.size() * sizeof(double)) ||
      !write_all(inf, idx.data(), idx.size() * sizeof(int32_t))) {
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
