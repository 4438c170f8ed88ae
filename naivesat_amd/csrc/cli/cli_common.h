// Helpers shared by the command-line tools (flag parsing, file loading, printing).
#pragma once
#include <chrono>
#include <csignal>
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

// The first line of every tool: the reference lists its OpenCL GPU devices
// (src/main.rs:385-388, "OpenCL devices: [...]"); here the HIP devices in the same form,
// each as hip:<index> <architecture>.
inline void print_devices() {
  uint32_t n = 0;
  char names[4096] = {0};
  if (ns_device_list(&n, names, sizeof(names)) != NS_OK) n = 0;
  std::string line;
  const char* p = names;
  for (uint32_t d = 0; d < n; d++) {
    const char* q = std::strchr(p, ',');
    std::string name = q ? std::string(p, q) : std::string(p);
    if (d) line += ", ";
    line += "hip:" + std::to_string(d) + " " + name;
    p = q ? q + 1 : p + std::strlen(p);
  }
  std::printf("OpenCL devices: [%s]\n", line.c_str());
}

// Progress of a search made of launch ranges (ns_sat_solve_progress / ns_qsat_solve_progress):
// "Start execution" and the clock start once the kernel is built, "Step: k / steps" after
// each range completes (the reference's per-arg callback, src/main.rs:125-153), and SIGINT
// stops the search after the ranges in flight (a second SIGINT ends the process).
inline volatile std::sig_atomic_t& interrupt_flag() {
  static volatile std::sig_atomic_t flag = 0;
  return flag;
}
extern "C" inline void on_interrupt(int) {
  interrupt_flag() = 1;
  std::signal(SIGINT, SIG_DFL);
}

struct Progress {
  bool print = true;       // the reference prints these lines for n >= 14 only
  long long elem = -1;     // elem inputs as printed before the build
  bool started = false;
  uint64_t done = 0, steps = 0;
  uint32_t bits = 0;
  std::chrono::steady_clock::time_point t0;
  Progress() { std::signal(SIGINT, on_interrupt); }
  ~Progress() { std::signal(SIGINT, SIG_DFL); }
  bool cancelled() const { return interrupt_flag() != 0; }
  // wall time from "Start execution" to now
  double seconds() const {
    return started ? std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() : 0.0;
  }
};

extern "C" inline int progress_step(void* user, uint64_t step, uint64_t steps, uint32_t bits) {
  auto* p = static_cast<Progress*>(user);
  p->steps = steps;
  p->bits = bits;
  if (step == NS_SAT_STEP_START) {
    if (p->print) {
      if (p->elem >= 0 && (long long)bits != p->elem)
        std::printf("Note: launch ranges of 2^%u assignments (elem inputs %u)\n", bits, bits);
      std::printf("Start execution\n");
    }
    std::fflush(stdout);
    p->started = true;
    p->t0 = std::chrono::steady_clock::now();
  } else {
    p->done = step + 1;
    if (p->print) std::printf("Step: %llu / %llu\n", (unsigned long long)step, (unsigned long long)steps);
    std::fflush(stdout);
  }
  return interrupt_flag() ? 1 : 0;
}

}  // namespace cli
