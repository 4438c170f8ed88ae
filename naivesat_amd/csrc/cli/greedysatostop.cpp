// greedysatostop: SATOSTOP by a device-resident next-state table and pointer jumping, or
// with -p by the reference's partitioned search over an on-disk image (the joins on the
// device). Command line and printed lines follow the reference's greedysatostop
// (src/greedysatostop.rs:28-49,692-837,1042-1138,1356-1369).
#include "cli_common.h"

static const char* kUsage =
    "Usage: greedysatostop [OPTIONS] <CIRCUIT> <UNKNOWNS>\n"
    "  -C, --opencl <N>                  HIP device index [default: 0]\n"
    "  -n, --no-optimize-negs            accepted for compatibility\n"
    "  -p, --partitions <P>              partitioned search over an on-disk image (power of two >= 2)\n"
    "  -x, --file-image-prefix <PREFIX>  where the image file goes (prefix of its path)\n"
    "  -v, --verify                      verify the solution by walking the circuit\n"
    "  -a, --all                         run every pass and report all stopping unknowns\n";

int main(int argc, char** argv) {
  auto a = cli::parse_args(argc, argv,
                           {{'C', "opencl", true},
                            {'n', "no-optimize-negs", false},
                            {'p', "partitions", true},
                            {'x', "file-image-prefix", true},
                            {'v', "verify", false},
                            {'a', "all", false}},
                           kUsage);
  if (a.positional.size() != 2) cli::die(std::string("expected <CIRCUIT> <UNKNOWNS>\n") + kUsage);
  cli::print_devices();
  std::string text = cli::read_file(a.positional[0]);
  ns_circuit* c = nullptr;
  cli::check(ns_circuit_parse(text.data(), text.size(), &c), "circuit");
  uint32_t n = 0, g = 0, no = 0;
  cli::check(ns_circuit_info(c, &n, &g, &no), "circuit");
  if (no != n + 1) cli::die("circuit must have one output more than inputs");
  long long U = std::atoll(a.positional[1].c_str());
  if (U < 0 || U > (long long)n) cli::die("unknowns must not exceed the inputs");
  ns_ctx* ctx = nullptr;
  ns_satostop_result r{};
  if (a.has("partitions")) {
    // do_solve_with_cpu_builder_with_partitions (src/greedysatostop.rs:692-837)
    if (a.has("opencl")) cli::die("--partitions runs without --opencl");
    const long long P = a.num("partitions", 0);
    if (P < 2 || (P & (P - 1))) cli::die("Number of partitions must be power of 2 (at least 2)");
    uint32_t pb = 0;
    while ((1ll << pb) < P) pb++;
    std::printf("Load partition bits: %u\n", std::min(pb + 2, n > 10 ? n - 10 : 0u));
    cli::check(ns_ctx_create(0, &ctx), "device");
    std::printf("Calculate first nexts\n");
    ns_partitioned_result pr{};
    const std::string prefix = a.has("file-image-prefix") ? a.get("file-image-prefix") : "";
    cli::check(ns_greedy_partitioned_solve(ctx, c, (uint32_t)U, (uint64_t)P, prefix.c_str(), &pr), "search");
    for (uint32_t st = 0; st < pr.stages; st++) std::printf("Stage %u / %u\n", st, n);
    std::printf("Time: %g\n", (pr.wall_ms - pr.compile_ms) / 1000.0);
    r.n_stopped = pr.found;
    r.start = pr.start;
    r.end = pr.end;
  } else {
    int dev = (int)a.num("opencl", 0);
    cli::check(ns_ctx_create(dev, &ctx), "device");
    std::printf("Calculate first nexts\n");
    ns_satostop_opts o{};
    // greedysatostop's own search runs over the next-state table (its rule: the first jump
    // pass in which a start stops; --all: every start, still over the table)
    o.flags = a.has("all") ? NS_STOP_TABLE_WALKS : NS_STOP_FIRST_PASS;
    cli::check(ns_satostop_solve(ctx, c, (uint32_t)U, &o, &r, nullptr, nullptr), "search");
    for (uint32_t i = 0; i < r.passes; i++) std::printf("Joining nexts: Stage: %u / %u\n", i, n);
    // wall time after the kernel build (src/greedysatostop.rs:655-688)
    std::printf("Time: %g\n", (r.wall_ms - r.compile_ms) / 1000.0);
    if (a.has("all")) std::printf("Stopping unknowns: %llu\n", (unsigned long long)r.n_stopped);
  }
  if (r.n_stopped) {
    std::printf("Solution: Solution { start: %llu, end: %llu }\n", (unsigned long long)r.start,
                (unsigned long long)r.end);
    if (a.has("verify")) {
      unsigned long long s = r.start;
      std::vector<uint8_t> in(n), out(n + 1);
      for (unsigned long long step = 0; step <= (1ull << n); step++) {
        for (uint32_t k = 0; k < n; k++) in[k] = (uint8_t)((s >> k) & 1ull);
        cli::check(ns_circuit_eval(c, in.data(), out.data()), "verify");
        s = 0;
        for (uint32_t k = 0; k < n; k++) s |= (unsigned long long)out[k] << k;
        if (out[n]) break;
      }
      std::printf("Verified: end=%llu\n", s);
      if (s != r.end) std::printf("INCORRECT\n");
    }
  } else {
    std::printf("Unsatisfiable!\n");
  }
  ns_ctx_destroy(ctx);
  ns_circuit_free(c);
  return 0;
}
