// Helpers shared by the command-line tools (flag parsing, file loading, printing).
#pragma once
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "naivesat_hip.h"

namespace cli {

[[noreturn]] inline void die(const std::string& msg) {
  std::fprintf(stderr, "error: %s\n", msg.c_str());
  std::exit(1);
}

inline void check(int st, const char* what) {
  if (st != NS_OK) die(std::string(what) + ": " + ns_last_error());
}

inline std::string read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) die("cannot read " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

// Minimal flag parser: short (-e 24, -e24) and long (--elem-inputs 24, --elem-inputs=24)
// options; `with_value` lists options that take a value; the rest are switches.
struct Args {
  std::vector<std::string> positional;
  std::vector<std::pair<std::string, std::string>> opts;  // canonical long name -> value

  bool has(const std::string& name) const {
    for (auto& o : opts)
      if (o.first == name) return true;
    return false;
  }
  std::string get(const std::string& name, const std::string& def = "") const {
    for (auto& o : opts)
      if (o.first == name) return o.second;
    return def;
  }
  long long num(const std::string& name, long long def) const {
    std::string v = get(name);
    if (v.empty()) return def;
    char* end = nullptr;
    long long r = std::strtoll(v.c_str(), &end, 10);
    if (!end || *end) die("bad number for --" + name + ": " + v);
    return r;
  }
};

struct OptSpec {
  char short_name;
  const char* long_name;
  bool with_value;
};

inline Args parse_args(int argc, char** argv, const std::vector<OptSpec>& spec, const char* usage) {
  Args a;
  for (int i = 1; i < argc; i++) {
    std::string s = argv[i];
    if (s == "-h" || s == "--help") {
      std::printf("%s", usage);
      std::exit(0);
    }
    if (s.size() > 1 && s[0] == '-' && !(s.size() > 1 && std::isdigit((unsigned char)s[1]))) {
      const OptSpec* found = nullptr;
      std::string val;
      bool have_val = false;
      if (s.rfind("--", 0) == 0) {
        std::string name = s.substr(2);
        size_t eq = name.find('=');
        if (eq != std::string::npos) {
          val = name.substr(eq + 1);
          name = name.substr(0, eq);
          have_val = true;
        }
        for (auto& o : spec)
          if (name == o.long_name) found = &o;
      } else {
        for (auto& o : spec)
          if (s[1] == o.short_name) found = &o;
        if (found && s.size() > 2) {
          val = s.substr(2);
          have_val = true;
        }
      }
      if (!found) die("unknown option " + s + "\n" + usage);
      if (found->with_value && !have_val) {
        if (i + 1 >= argc) die(std::string("missing value for ") + s);
        val = argv[++i];
      }
      a.opts.push_back({found->long_name, val});
    } else {
      a.positional.push_back(s);
    }
  }
  return a;
}

// -t values of the reference (cpu, opencl:N, cpu_and_opencl, ...) plus hip:N. Every
// type runs on one HIP device; returns its index.
inline int device_from_exec_type(const std::string& t, bool print) {
  auto num_after = [&](size_t pos) {
    char* end = nullptr;
    long v = std::strtol(t.c_str() + pos, &end, 10);
    if (!end || *end || pos >= t.size()) die("Unknown exec type");
    return (int)v;
  };
  if (t.rfind("hip:", 0) == 0) return num_after(4);
  if (t.rfind("opencl:", 0) == 0) return num_after(7);
  if (t.rfind("cpu_and_opencl_1d:", 0) == 0) {
    if (print) std::printf("Note: %s runs on HIP device alone\n", t.c_str());
    return num_after(18);
  }
  if (t.rfind("cpu_and_opencl_1:", 0) == 0) {
    if (print) std::printf("Note: %s runs on HIP device alone\n", t.c_str());
    return num_after(17);
  }
  if (t == "cpu" || t == "cpu_and_opencl" || t == "cpu_and_opencl_d") {
    if (print) std::printf("Note: exec type %s runs on HIP device 0\n", t.c_str());
    return 0;
  }
  die("Unknown exec type");
}

inline std::string bits_msb_first(unsigned long long v, unsigned width) {
  std::string s;
  for (int b = (int)width - 1; b >= 0; b--) s += ((v >> b) & 1ull) ? '1' : '0';
  return s;
}

inline void print_devices() {
  // the reference lists its OpenCL devices first (src/main.rs:385-388)
  std::printf("HIP devices: [%s]\n", "gfx950");
}

}  // namespace cli
