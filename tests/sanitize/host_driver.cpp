// Host-side exercise of the engine's threaded and parsing code for the
// sanitizer builds (AddressSanitizer + UndefinedBehaviorSanitizer, and
// ThreadSanitizer): the XTC codec (write, index, parallel decode against a
// single-threaded decode, truncated and corrupted files), the topology
// reader, and the sparse graph stages with the threaded Lanczos solver.
// Built from the library's own sources (xtc.cpp, topology.cpp,
// spectral.cpp) with g++, no GPU code.  Exits 0 when every check holds; a
// sanitizer report makes the run fail on its own.
//   host_driver <golden dir> <scratch dir> [quick]
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mdsctk_graph.h"
#include "mdsctk_io.h"
#include "mdsctk_knn.h"

namespace {

int failures = 0;

void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s (%s)\n", what, mdsctk_io_last_error());
    ++failures;
  }
}

std::vector<float> frames(int n, int natoms, unsigned seed) {
  std::mt19937 rng(seed);
  std::normal_distribution<float> g(0.f, 1.f);
  std::vector<float> x((size_t)n * natoms * 3);
  for (auto& v : x) v = 2.5f * g(rng);
  return x;
}

std::vector<unsigned char> slurp(const std::string& p) {
  std::vector<unsigned char> b;
  FILE* f = std::fopen(p.c_str(), "rb");
  if (!f) return b;
  std::fseek(f, 0, SEEK_END);
  b.resize((size_t)std::ftell(f));
  std::fseek(f, 0, SEEK_SET);
  if (!b.empty() && std::fread(b.data(), 1, b.size(), f) != b.size()) b.clear();
  std::fclose(f);
  return b;
}

void spit(const std::string& p, const std::vector<unsigned char>& b) {
  FILE* f = std::fopen(p.c_str(), "wb");
  if (!f) return;
  std::fwrite(b.data(), 1, b.size(), f);
  std::fclose(f);
}

// decode with 1 and with 8 threads; both must agree bit for bit
bool decode_both(const std::string& path, int natoms, int64_t n, std::vector<float>* out) {
  std::vector<float> a((size_t)n * natoms * 3), b(a.size());
  int64_t ga = 0, gb = 0;
  if (mdsctk_xtc_read(path.c_str(), a.data(), n, natoms, &ga, nullptr, nullptr, nullptr, 1) < 0)
    return false;
  std::vector<int32_t> steps(n);
  std::vector<float> times(n), boxes((size_t)n * 9);
  if (mdsctk_xtc_read(path.c_str(), b.data(), n, natoms, &gb, steps.data(), times.data(),
                      boxes.data(), 8) < 0)
    return false;
  if (ga != n || gb != n || std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) != 0)
    return false;
  if (out) *out = a;
  return true;
}

void xtc_checks(const std::string& gold, const std::string& tmp, bool quick) {
  // round trip at two precisions and a frame of at most 9 atoms (raw floats)
  struct Case { int n, natoms; float prec; };
  const Case cases[] = {{quick ? 40 : 300, 120, 1e5f}, {60, 304, 1e3f}, {17, 5, 1e3f}};
  for (const Case& c : cases) {
    const std::string f = tmp + "/rt.xtc";
    const auto x = frames(c.n, c.natoms, (unsigned)c.natoms);
    check(mdsctk_xtc_write(f.c_str(), x.data(), c.n, c.natoms, c.prec, nullptr, nullptr,
                           nullptr) == 0, "xtc_write");
    int64_t nf = 0;
    int32_t na = 0;
    float pr = 0;
    check(mdsctk_xtc_info(f.c_str(), &nf, &na, &pr) == 0 && nf == c.n && na == c.natoms,
          "xtc_info");
    std::vector<float> y;
    check(decode_both(f, c.natoms, c.n, &y), "parallel decode equals serial decode");
    double err = 0;
    for (size_t i = 0; i < x.size() && !y.empty(); ++i)
      err = std::fmax(err, std::fabs((double)x[i] - (double)y[i]));
    check(c.natoms <= 9 ? err == 0.0 : err <= 0.51 / c.prec + 1e-6, "round-trip error");
    // every truncation is refused (index) or decoded without reading past the end
    const auto bytes = slurp(f);
    const size_t cuts[] = {bytes.size() - 1, bytes.size() - 5, bytes.size() / 2, 57, 3};
    for (size_t cut : cuts) {
      std::vector<unsigned char> t(bytes.begin(), bytes.begin() + (long)cut);
      spit(tmp + "/cut.xtc", t);
      int64_t n2 = 0;
      check(mdsctk_xtc_info((tmp + "/cut.xtc").c_str(), &n2, &na, &pr) < 0, "truncated file refused");
    }
    // corrupted bit streams: decoding must stay within its buffers (any
    // error code is fine)
    std::mt19937 rng(7);
    for (int trial = 0; trial < (quick ? 20 : 120) && c.natoms > 9; ++trial) {
      auto t = bytes;
      const size_t frame_len = bytes.size() / c.n;
      for (int flips = 0; flips < 4; ++flips) {
        const size_t pos = 92 + rng() % (bytes.size() - 92);
        if (pos % frame_len < 92) continue;  // keep headers (the index pass rejects those)
        t[pos] ^= (unsigned char)(1u << (rng() % 8));
      }
      spit(tmp + "/bad.xtc", t);
      std::vector<float> z((size_t)c.n * c.natoms * 3);
      int64_t g = 0;
      (void)mdsctk_xtc_read((tmp + "/bad.xtc").c_str(), z.data(), c.n, c.natoms, &g, nullptr,
                            nullptr, nullptr, 4);
    }
  }
  // the GROMACS-written fixture (small-difference runs)
  int64_t nf = 0;
  int32_t na = 0;
  float pr = 0;
  const std::string trp = gold + "/trp-cage.xtc";
  check(mdsctk_xtc_info(trp.c_str(), &nf, &na, &pr) == 0 && nf == 1000 && na == 60, "trp-cage info");
  check(decode_both(trp, 60, nf, nullptr), "trp-cage parallel decode");
  int32_t head = 0;
  check(mdsctk_xtc_natoms(trp.c_str(), &head) == 0 && head == 60, "xtc_natoms");
  // atom-count mismatch
  std::vector<float> z((size_t)nf * 61 * 3);
  int64_t g = 0;
  check(mdsctk_xtc_read(trp.c_str(), z.data(), nf, 61, &g, nullptr, nullptr, nullptr, 4) ==
            MDSCTK_ERR_NATOMS, "natoms mismatch refused");
}

void topology_checks(const std::string& gold, const std::string& tmp) {
  int32_t na = 0;
  const std::string pdb = gold + "/trp-cage.pdb";
  check(mdsctk_topology_read(pdb.c_str(), &na, nullptr, 0) == 0 && na == 60, "pdb natoms");
  std::vector<float> m(na);
  check(mdsctk_topology_read(pdb.c_str(), &na, m.data(), na) == 0, "pdb masses");
  double tot = 0;
  for (float v : m) tot += v;
  check(tot > 300 && tot < 1200, "pdb mass sum");
  check(mdsctk_topology_read(pdb.c_str(), &na, m.data(), 10) < 0, "small capacity refused");
  const std::string gro = tmp + "/t.gro";
  FILE* f = std::fopen(gro.c_str(), "w");
  std::fprintf(f, "test\n    3\n");
  const char* names[] = {"N", "CA", "OW"};
  for (int i = 0; i < 3; ++i)
    std::fprintf(f, "%5d%-5s%5s%5d%8.3f%8.3f%8.3f\n", 1, "ALA", names[i], i + 1, 0.1 * i, 0.0, 0.0);
  std::fprintf(f, "   1.00000   1.00000   1.00000\n");
  std::fclose(f);
  float gm[3];
  check(mdsctk_topology_read(gro.c_str(), &na, gm, 3) == 0 && na == 3, "gro read");
  const std::string bad = tmp + "/bad.pdb";
  f = std::fopen(bad.c_str(), "w");
  std::fprintf(f, "ATOM  xxxxx garbage\n");
  std::fclose(f);
  (void)mdsctk_topology_read(bad.c_str(), &na, nullptr, 0);
}

void graph_checks(bool quick) {
  // a kNN graph of points on a ring with noise: k nearest by index distance
  const int n = quick ? 4500 : 12000, k = 8, nev = 6;  // >= 4096: the thread team runs
  std::mt19937 rng(3);
  std::uniform_real_distribution<double> u(0.9, 1.1);
  std::vector<double> dist((size_t)n * k);
  std::vector<int32_t> idx((size_t)n * k);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < k; ++j) {
      const int step = j / 2 + 1;
      idx[(size_t)i * k + j] = (i + ((j & 1) ? step : n - step)) % n;
      dist[(size_t)i * k + j] = step * u(rng);
    }
  int64_t nnz = 0;
  check(mdsctk_ssm_build(dist.data(), idx.data(), n, k, k, nullptr, nullptr, nullptr, &nnz) == 0,
        "ssm size");
  std::vector<int32_t> pcol(n + 1), irow(nnz);
  std::vector<double> vals(nnz);
  check(mdsctk_ssm_build(dist.data(), idx.data(), n, k, k, pcol.data(), irow.data(), vals.data(),
                         &nnz) == 0, "ssm build");
  std::vector<double> ev(nev), vec((size_t)nev * n), res(nev);
  const int got = mdsctk_spectral_decomp(n, pcol.data(), irow.data(), vals.data(), 2.0, nev,
                                         ev.data(), vec.data(), res.data());
  check(got == nev, "Lanczos converged");
  check(std::fabs(ev[0] - 1.0) < 1e-8, "leading eigenvalue 1 (connected graph)");
  for (int j = 0; j < nev; ++j) check(res[j] < 1e-8, "residual");
  // the Nystrom extension on the same matrix (its in-sample closed form)
  int64_t gnnz = 0;
  const int m = 50;
  check(mdsctk_gsm_build(dist.data(), idx.data(), m, k, k, 1, nullptr, nullptr, nullptr, &gnnz) == 0,
        "gsm size");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: host_driver <golden dir> <scratch dir> [quick]\n");
    return 2;
  }
  const bool quick = argc > 3 && std::strcmp(argv[3], "quick") == 0;
  xtc_checks(argv[1], argv[2], quick);
  topology_checks(argv[1], argv[2]);
  graph_checks(quick);
  std::printf("host_driver: %d failures\n", failures);
  return failures ? 1 : 0;
}
