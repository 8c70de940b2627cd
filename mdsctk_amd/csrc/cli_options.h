// Command-line handling shared by the knn_rms, make_sysparse and
// decomp_sparse executables, in the style of boost::program_options'
// defaults used by the MDSCTK tools: "-k 30", "-k30", "--knn 30",
// "--knn=30", unambiguous prefixes of long names, and boolean values
// true/false, 1/0, yes/no, on/off.
#pragma once

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

namespace mdsctk_cli {

struct Opt {
  const char* lng;
  char shrt;
  bool takes_value;
  const char* help;
};

inline void usage(const char* program, const std::vector<Opt>& opts) {
  std::printf("usage: %s [options]\nProgram options:\n", program);
  for (const auto& o : opts) {
    char left[64];
    if (o.shrt)
      std::snprintf(left, sizeof left, "  -%c [ --%s ]%s", o.shrt, o.lng, o.takes_value ? " arg" : "");
    else
      std::snprintf(left, sizeof left, "  --%s%s", o.lng, o.takes_value ? " arg" : "");
    std::printf("%-36s %s\n", left, o.help);
  }
}

inline bool parse_bool(const std::string& v, bool* out) {
  std::string s;
  for (char c : v) s.push_back((char)std::tolower((unsigned char)c));
  if (s == "true" || s == "1" || s == "yes" || s == "on") { *out = true; return true; }
  if (s == "false" || s == "0" || s == "no" || s == "off") { *out = false; return true; }
  return false;
}

inline bool parse_int(const std::string& v, int* out) {
  char* end = nullptr;
  long x = std::strtol(v.c_str(), &end, 10);
  if (v.empty() || *end) return false;
  *out = (int)x;
  return true;
}

inline bool parse_double(const std::string& v, double* out) {
  char* end = nullptr;
  double x = std::strtod(v.c_str(), &end);
  if (v.empty() || *end) return false;
  *out = x;
  return true;
}

// Returns false (with a message) on an unknown or incomplete option.
inline bool parse(int argc, char** argv, const std::vector<Opt>& opts,
                  std::map<std::string, std::string>& vals, std::string& err) {
  for (int a = 1; a < argc; ++a) {
    std::string arg = argv[a];
    const Opt* opt = nullptr;
    std::string value;
    bool have_value = false;
    if (arg.rfind("--", 0) == 0) {
      std::string name = arg.substr(2);
      const size_t eq = name.find('=');
      if (eq != std::string::npos) {
        value = name.substr(eq + 1);
        name = name.substr(0, eq);
        have_value = true;
      }
      int matches = 0;
      for (const auto& o : opts) {
        if (name == o.lng) { opt = &o; matches = 1; break; }
        if (std::string(o.lng).rfind(name, 0) == 0) { opt = &o; ++matches; }
      }
      if (!opt || matches != 1) {
        err = "unrecognised option '" + arg + "'";
        return false;
      }
    } else if (arg.size() >= 2 && arg[0] == '-') {
      for (const auto& o : opts)
        if (o.shrt == arg[1]) opt = &o;
      if (!opt) {
        err = "unrecognised option '" + arg + "'";
        return false;
      }
      if (arg.size() > 2) {
        value = arg.substr(2);
        have_value = true;
      }
    } else {
      err = "unexpected argument '" + arg + "'";
      return false;
    }
    if (opt->takes_value && !have_value) {
      if (a + 1 >= argc) {
        err = std::string("the required argument for option '--") + opt->lng + "' is missing";
        return false;
      }
      value = argv[++a];
    }
    vals[opt->lng] = value;
  }
  return true;
}

inline bool write_all(const std::string& path, const void* data, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = bytes == 0 || std::fwrite(data, 1, bytes, f) == bytes;
  return std::fclose(f) == 0 && ok;
}

inline bool read_all(const std::string& path, std::vector<char>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(size > 0 ? (size_t)size : 0);
  const bool ok = out.empty() || std::fread(out.data(), 1, out.size(), f) == out.size();
  std::fclose(f);
  return ok;
}

}  // namespace mdsctk_cli
