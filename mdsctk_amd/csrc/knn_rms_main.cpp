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
// The pair computation runs on the GPU through libmdsctk_knn.  The XTC
// files are decoded on the GPU as they are read (mdsctk_xtc_read_device),
// so the trajectories never sit in host memory; outputs are produced and
// written in row blocks of at most 2^27 values (the reference's 2^30 / 8
// doubles per block, knn_rms.cpp:143, 213-221, 281-291), so host memory stays
// bounded for --sort false too.  -t and -b are accepted and reported (the
// GPU decodes and tiles the work itself).
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

#include <hip/hip_runtime.h>

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

  using clk = std::chrono::steady_clock;
  auto secs_since = [](clk::time_point t) {
    return std::chrono::duration<double>(clk::now() - t).count();
  };
  const auto t_start = clk::now();
  // device buffers and pinned host buffers of the run, released at exit
  struct Held {
    float* d_ref = nullptr;
    float* d_fit = nullptr;
    double* d_dist = nullptr;
    int32_t* d_idx = nullptr;
    double* h_dist = nullptr;
    int32_t* h_idx = nullptr;
    mdsctk_knn_ctx* ctx = nullptr;
    FILE* fd = nullptr;
    FILE* fi = nullptr;
    ~Held() {
      if (ctx) mdsctk_knn_destroy(ctx);
      mdsctk_xtc_free_device(d_ref);
      mdsctk_xtc_free_device(d_fit);
      (void)hipFree(d_dist);
      (void)hipFree(d_idx);
      if (h_dist) (void)hipHostFree(h_dist);
      if (h_idx) (void)hipHostFree(h_idx);
      if (fd) std::fclose(fd);
      if (fi) std::fclose(fi);
    }
  } h;

  // XTC -> device (knn_rms.cpp:155-206); a natoms mismatch exits with 4
  auto load = [&](const std::string& path, float** d, int64_t* n) -> int {
    const int rc = mdsctk_xtc_read_device(path.c_str(), natoms, device, d, n);
    if (rc == MDSCTK_ERR_NATOMS) {
      int32_t na = 0;
      mdsctk_xtc_natoms(path.c_str(), &na);
      std::printf("*** ERROR ***\nNumber of atoms in topology file (%d) does not match the number "
                  "of atoms in the XTC file (%s : %d).\n", natoms, path.c_str(), na);
      return 4;
    }
    if (rc < 0) {
      std::printf("\nERROR: %s\n", mdsctk_io_last_error());
      return 3;
    }
    return 0;
  };

  // Both files' atom counts are checked at open, before any frame is read
  // (knn_rms.cpp:155-178).
  for (const std::string* path : {&reff, &fitf}) {
    int32_t na = 0;
    if (mdsctk_xtc_natoms(path->c_str(), &na) < 0) {
      std::printf("\nERROR: %s\n", mdsctk_io_last_error());
      return 3;
    }
    if (na != natoms) {
      std::printf("*** ERROR ***\nNumber of atoms in topology file (%d) does not match the number "
                  "of atoms in the XTC file (%s : %d).\n", natoms, path->c_str(), na);
      return 4;
    }
  }

  int64_t n_ref = 0, n_fit = 0;
  std::printf("Reading reference coordinates from file: %s ... ", reff.c_str());
  std::fflush(stdout);
  if (int rc = load(reff, &h.d_ref, &n_ref)) return rc;
  std::printf("done.\n");
  if (!self) {
    std::printf("Reading fitting coordinates from file: %s ... ", fitf.c_str());
    std::fflush(stdout);
    if (int rc = load(fitf, &h.d_fit, &n_fit)) return rc;
    std::printf("done.\n");
  } else {
    n_fit = n_ref;
  }
  const double t_read = secs_since(t_start);
  if (n_ref == 0) {
    std::printf("ERROR: %s holds no frames.\n", reff.c_str());
    return 3;
  }
  std::printf("Block size: %d\n", std::max(1, (int)std::min<int64_t>(blksize, n_fit)));

  int rc = mdsctk_knn_create(&h.ctx, device, precision);
  if (rc < 0) {
    h.ctx = nullptr;
    std::printf("ERROR: cannot create the GPU context on device %d (code %d)\n", device, rc);
    return 2;
  }
  mdsctk_knn_ctx* ctx = h.ctx;
  const auto t0 = clk::now();
  rc = mdsctk_knn_set_reference_device(ctx, h.d_ref, n_ref, natoms, masses.data());
  if (rc < 0) {
    std::printf("ERROR: %s\n", mdsctk_knn_last_error(ctx));
    return 2;
  }
  uint32_t flags = 0;
  if (nofit) flags |= MDSCTK_KNN_NOFIT;
  if (!sort) flags |= MDSCTK_KNN_FULL_ROWS;
  const int width = mdsctk_knn_row_width(ctx, k, flags);
  if (width < 0) {
    std::printf("ERROR: %s\n", mdsctk_knn_last_error(ctx));
    return 2;
  }
  h.fd = std::fopen(df.c_str(), "wb");
  h.fi = std::fopen(inf.c_str(), "wb");
  if (!h.fd || !h.fi) {
    std::printf("ERROR: cannot write %s / %s\n", df.c_str(), inf.c_str());
    return 3;
  }
  // Row blocks of at most 2^27 output values (1 GiB of doubles).  A self
  // query that fits one block runs as one self query (each unordered pair
  // scored once); otherwise each block queries its frames' rows.
  const int64_t w = std::max(width, 1);
  const int64_t rows_per_block = std::max<int64_t>(1, std::min<int64_t>(n_fit, (int64_t(1) << 27) / w));
  const size_t blk_vals = (size_t)rows_per_block * w;
  if (hipMalloc((void**)&h.d_dist, blk_vals * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&h.d_idx, blk_vals * sizeof(int32_t)) != hipSuccess ||
      hipHostMalloc((void**)&h.h_dist, blk_vals * sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&h.h_idx, blk_vals * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
    std::printf("ERROR: cannot allocate the output blocks (%zu values)\n", blk_vals);
    return 2;
  }
  const float* fit_frames = self ? h.d_ref : h.d_fit;
  const bool one_self_query = self && rows_per_block >= n_fit;
  double gpu_ms = 0;
  int64_t n_cand = 0, n_exact = 0;
  double t_write = 0;
  for (int64_t r0 = 0; r0 < n_fit; r0 += rows_per_block) {
    const int64_t nr = std::min(rows_per_block, n_fit - r0);
    rc = mdsctk_knn_query_device(ctx, one_self_query ? nullptr : fit_frames + r0 * natoms * 3, nr,
                                 k, flags, h.d_dist, h.d_idx);
    if (rc < 0) {
      std::printf("ERROR: %s\n", mdsctk_knn_last_error(ctx));
      return 2;
    }
    mdsctk_knn_stats st;
    mdsctk_knn_get_stats(ctx, &st);
    gpu_ms += st.ms_total;
    n_cand += st.n_candidates;
    n_exact += st.n_fallback_rows;
    const size_t nv = (size_t)nr * width;
    if (hipMemcpy(h.h_dist, h.d_dist, nv * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(h.h_idx, h.d_idx, nv * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) {
      std::printf("ERROR: cannot copy the outputs back\n");
      return 2;
    }
    const auto tw = clk::now();
    if (std::fwrite(h.h_dist, sizeof(double), nv, h.fd) != nv ||
        std::fwrite(h.h_idx, sizeof(int32_t), nv, h.fi) != nv) {
      std::printf("ERROR: cannot write %s / %s\n", df.c_str(), inf.c_str());
      return 3;
    }
    t_write += secs_since(tw);
  }
  if (std::fclose(h.fd) != 0 || std::fclose(h.fi) != 0) {
    h.fd = h.fi = nullptr;
    std::printf("ERROR: cannot write %s / %s\n", df.c_str(), inf.c_str());
    return 3;
  }
  h.fd = h.fi = nullptr;
  const double secs = secs_since(t0);
  const double pairs = (double)n_fit * (double)n_ref;
  std::printf("\n%lld x %lld pairs, k=%d, %s: %.3f s (GPU %.3f ms, %.3g pairs/s), "
              "%lld candidates re-scored, %lld exact rows; XTC read and decode %.3f s, "
              "file writes %.3f s, wall %.3f s\n\n",
              (long long)n_fit, (long long)n_ref, width, prec.c_str(), secs, gpu_ms,
              gpu_ms > 0 ? pairs / (gpu_ms * 1e-3) : 0.0, (long long)n_cand, (long long)n_exact,
              t_read, t_write, secs_since(t_start));
  return 0;
}
