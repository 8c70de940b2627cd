// Topology reader for the knn_rms path: atom count and per-atom masses from
// PDB or GRO files (the reference calls GROMACS read_tps_conf and uses
// top.atoms.atom[i].m as fit/RMSD weights, knn_rms.cpp:149-153, 181-182).
//
// Mass assignment: the PDB element column (77-78) when present, otherwise
// the element implied by the atom name: leading digits are skipped and the
// first letter is taken, except that an ion whose residue name equals its
// atom name (NA, CL, MG, ZN, CA, K, ...) takes the two-letter element.
// Masses are the standard atomic weights GROMACS tabulates for these
// elements (H 1.008, C 12.011, N 14.0067, O 15.9994, P 30.97376, S 32.06).
// GROMACS resolves masses through its residue/atom-name database; for
// protein and nucleic-acid names both rules give the same element.
#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mdsctk_io.h"
#include "../../include/mdsctk_knn.h"

extern "C" void mdsctk_io_set_error(const char* msg);  // defined in xtc.cpp

namespace {

struct ElementMass {
  const char* sym;
  float mass;
};

const ElementMass kMasses[] = {
    {"H", 1.00800f},   {"HE", 4.00260f},  {"LI", 6.94100f},  {"BE", 9.01218f},
    {"B", 10.81000f},  {"C", 12.01100f},  {"N", 14.00670f},  {"O", 15.99940f},
    {"F", 18.99840f},  {"NE", 20.17900f}, {"NA", 22.98977f}, {"MG", 24.30500f},
    {"AL", 26.98154f}, {"SI", 28.08550f}, {"P", 30.97376f},  {"S", 32.06000f},
    {"CL", 35.45300f}, {"AR", 39.94800f}, {"K", 39.09830f},  {"CA", 40.08000f},
    {"MN", 54.93800f}, {"FE", 55.84700f}, {"CO", 58.93320f}, {"NI", 58.69000f},
    {"CU", 63.54600f}, {"ZN", 65.39000f}, {"SE", 78.96000f}, {"BR", 79.90400f},
    {"I", 126.90450f}, {"CS", 132.90540f},
};

const char* kIons[] = {"NA", "CL", "MG", "ZN", "CA", "K", "FE", "MN", "CU", "CO",
                       "NI", "LI", "CS", "BR", "I", "F"};

std::string trim_upper(const std::string& s) {
  std::string o;
  for (char ch : s)
    if (!std::isspace((unsigned char)ch)) o.push_back((char)std::toupper((unsigned char)ch));
  return o;
}

bool mass_of(const std::string& sym, float* m) {
  for (const auto& e : kMasses)
    if (sym == e.sym) {
      *m = e.mass;
      return true;
    }
  return false;
}

bool element_from_name(const std::string& atom, const std::string& res, std::string* el) {
  const std::string a = trim_upper(atom), r = trim_upper(res);
  for (const char* ion : kIons)
    if (a == ion && r == ion) {
      *el = ion;
      return true;
    }
  for (char ch : a)
    if (std::isalpha((unsigned char)ch)) {
      *el = std::string(1, ch);
      return true;
    }
  return false;
}

int fail(int code, const std::string& msg) {
  mdsctk_io_set_error(msg.c_str());
  return code;
}

struct Atom {
  std::string name, res, element;
};

int read_pdb(FILE* f, std::vector<Atom>& atoms) {
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    std::string s(line);
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
    if (s.compare(0, 6, "ENDMDL") == 0) break;
    if (s.compare(0, 6, "ATOM  ") != 0 && s.compare(0, 6, "HETATM") != 0) continue;
    if (s.size() < 27) return fail(MDSCTK_ERR_FORMAT, "short PDB ATOM record: " + s);
    Atom a;
    a.name = s.substr(12, 4);
    a.res = s.substr(17, 3);
    if (s.size() >= 78) a.element = trim_upper(s.substr(76, 2));
    atoms.push_back(a);
  }
  return 0;
}

int read_gro(FILE* f, std::vector<Atom>& atoms) {
  char line[512];
  if (!std::fgets(line, sizeof line, f)) return fail(MDSCTK_ERR_FORMAT, "empty GRO file");
  if (!std::fgets(line, sizeof line, f)) return fail(MDSCTK_ERR_FORMAT, "GRO file has no atom count");
  long n = std::strtol(line, nullptr, 10);
  if (n <= 0) return fail(MDSCTK_ERR_FORMAT, "bad GRO atom count");
  for (long i = 0; i < n; ++i) {
    if (!std::fgets(line, sizeof line, f)) return fail(MDSCTK_ERR_FORMAT, "GRO file ends early");
    std::string s(line);
    if (s.size() < 20) return fail(MDSCTK_ERR_FORMAT, "short GRO atom line");
    Atom a;
    a.res = s.substr(5, 5);
    a.name = s.substr(10, 5);
    atoms.push_back(a);
  }
  return 0;
}

bool ends_with(const std::string& s, const char* suf) {
  const size_t n = std::strlen(suf);
  if (s.size() < n) return false;
  for (size_t i = 0; i < n; ++i)
    if (std::tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
  return true;
}

}  // namespace

extern "C" int mdsctk_topology_read(const char* path, int32_t* natoms, float* masses,
                                    int32_t capacity) {
  if (!path || !natoms) return fail(MDSCTK_ERR_ARG, "null argument");
  const std::string p(path);
  const bool gro = ends_with(p, ".gro");
  if (ends_with(p, ".tpr"))
    return fail(MDSCTK_ERR_FORMAT, p + ": .tpr topologies are not supported; give a .pdb or .gro");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(MDSCTK_ERR_IO, "cannot open " + p);
  std::vector<Atom> atoms;
  int rc = gro ? read_gro(f, atoms) : read_pdb(f, atoms);
  std::fclose(f);
  if (rc) return rc;
  if (atoms.empty()) return fail(MDSCTK_ERR_FORMAT, p + ": no atoms");
  *natoms = (int32_t)atoms.size();
  if (!masses) return 0;
  if (capacity < (int32_t)atoms.size()) return fail(MDSCTK_ERR_ARG, "mass buffer too small");
  for (size_t i = 0; i < atoms.size(); ++i) {
    std::string el = atoms[i].element;
    float m = 0.f;
    if (el.empty() || !mass_of(el, &m)) {
      if (!element_from_name(atoms[i].name, atoms[i].res, &el) || !mass_of(el, &m))
        return fail(MDSCTK_ERR_FORMAT, p + ": no mass for atom '" + atoms[i].name + "'");
    }
    masses[i] = m;
  }
  return 0;
}
