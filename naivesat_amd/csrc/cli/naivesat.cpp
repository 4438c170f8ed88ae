// naivesat: brute-force SAT of a one-output gate circuit on one HIP device.
// Command line and printed lines follow the reference's naivesat (src/main.rs:52-74,266-394).
#include "cli_common.h"

static const char* kUsage =
    "Usage: naivesat [OPTIONS] --exec-type <EXEC_TYPE> <CIRCUIT>\n"
    "  -e, --elem-inputs <N>      log2 of assignments per launch [default: 24]\n"
    "  -t, --exec-type <TYPE>     cpu|opencl:N|hip:N|cpu_and_opencl* (all run on one HIP device)\n"
    "  -n, --no-optimize-negs     accepted for compatibility\n"
    "  -G, --opencl-group-len <G> threads per block\n"
    "  -W, --words-per-lane <W>   32-assignment words per lane\n"
    "  -c, --count                search the whole space and print the satisfying count\n";

int main(int argc, char** argv) {
  auto a = cli::parse_args(argc, argv,
                           {{'e', "elem-inputs", true},
                            {'t', "exec-type", true},
                            {'n', "no-optimize-negs", false},
                            {'G', "opencl-group-len", true},
                            {'W', "words-per-lane", true},
                            {'c', "count", false}},
                           kUsage);
  if (a.positional.size() != 1) cli::die(std::string("expected one circuit file\n") + kUsage);
  if (!a.has("exec-type")) cli::die(std::string("--exec-type is required\n") + kUsage);
  cli::print_devices();
  std::string text = cli::read_file(a.positional[0]);
  ns_circuit* c = nullptr;
  cli::check(ns_circuit_parse(text.data(), text.size(), &c), "circuit");
  uint32_t n = 0, g = 0, no = 0;
  cli::check(ns_circuit_info(c, &n, &g, &no), "circuit");
  if (no != 1) cli::die("circuit must have exactly one output");
  long long e = a.num("elem-inputs", 24);
  if (n >= 14) {
    if (e >= (long long)n) e = n - 1;
    if (!(e > 12 && e <= 64)) cli::die("elem inputs must be in 13..64");
    std::printf("Elem inputs: %lld\n", e);
  }
  int dev = cli::device_from_exec_type(a.get("exec-type"), n >= 14);
  if (n >= 14) std::printf("Execute in HIP device=%d\n", dev);
  ns_ctx* ctx = nullptr;
  cli::check(ns_ctx_create(dev, &ctx), "device");
  ns_sat_opts o{};
  // one launch covers at least 2^32 assignments: many launches of 2^24 cannot fill the GPU
  o.elem_bits = (uint32_t)(e < 32 ? 32 : e);
  o.group_len = (uint32_t)a.num("opencl-group-len", 0);
  o.words_per_lane = (uint32_t)a.num("words-per-lane", 0);
  o.flags = a.has("count") ? 0u : NS_SAT_STOP_AT_FIRST;
  if (n >= 14) std::printf("Start execution\n");
  ns_sat_result r{};
  cli::check(ns_sat_solve(ctx, c, &o, &r), "search");
  if (n >= 14) std::printf("Time: %g\n", r.kernel_ms / 1000.0);
  if (a.has("count")) std::printf("Count: %llu\n", (unsigned long long)r.count);
  if (r.found) {
    std::vector<uint8_t> in(n), out(1);
    for (uint32_t k = 0; k < n; k++) in[k] = (uint8_t)((r.lowest >> k) & 1ull);
    cli::check(ns_circuit_eval(c, in.data(), out.data()), "verify");
    if (!out[0])
      std::printf("INCORRECT!! %s\n", cli::bits_msb_first(r.lowest, n).c_str());
    else
      std::printf("Found Input: %s\n", cli::bits_msb_first(r.lowest, n).c_str());
  } else {
    std::printf("Unsatisfiable!\n");
  }
  ns_ctx_destroy(ctx);
  ns_circuit_free(c);
  return 0;
}
