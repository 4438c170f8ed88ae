// naivesat: brute-force SAT of a one-output gate circuit on one HIP device.
// Command line and printed lines follow the reference's naivesat (src/main.rs:52-74,266-394).
#include "cli_common.h"

static const char* kUsage =
    "Usage: naivesat [OPTIONS] --exec-type <EXEC_TYPE> <CIRCUIT>\n"
    "  -e, --elem-inputs <N>      log2 of assignments per launch range, 13..64, kept exactly up to 40\n"
    "                             [default: 32; the reference's 24 leaves the GPU idle between launches]\n"
    "  -t, --exec-type <TYPE>     cpu|opencl:N|hip:N|cpu_and_opencl* (all run on one HIP device)\n"
    "  -n, --no-optimize-negs     accepted for compatibility\n"
    "  -G, --opencl-group-len <G> threads per block\n"
    "  -W, --words-per-lane <W>   32-assignment words per lane\n"
    "  -c, --count                search the whole space and print the satisfying count\n"
    "  -X, --no-group-exits       evaluate every gate of every assignment (by default a wave\n"
    "                             leaves a word once the output's groups so far are zero)\n"
    "Ctrl-C stops the search after the launch ranges in flight.\n";

int main(int argc, char** argv) {
  auto a = cli::parse_args(argc, argv,
                           {{'e', "elem-inputs", true},
                            {'t', "exec-type", true},
                            {'n', "no-optimize-negs", false},
                            {'G', "opencl-group-len", true},
                            {'W', "words-per-lane", true},
                            {'c', "count", false},
                            {'X', "no-group-exits", false}},
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
  long long e = a.num("elem-inputs", 32);
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
  o.elem_bits = n >= 14 ? (uint32_t)e : 0u;
  o.group_len = (uint32_t)a.num("opencl-group-len", 0);
  o.words_per_lane = (uint32_t)a.num("words-per-lane", 0);
  o.flags = (a.has("count") ? 0u : NS_SAT_STOP_AT_FIRST) | (a.has("no-group-exits") ? 0u : NS_SAT_GROUP_EXITS);
  cli::Progress prog;
  prog.print = n >= 14;
  prog.elem = n >= 14 ? e : -1;
  ns_sat_result r{};
  cli::check(ns_sat_solve_progress(ctx, c, &o, cli::progress_step, &prog, &r), "search");
  if (n >= 14) std::printf("Time: %g\n", prog.seconds());
  // interrupted: the answer over the ranges searched (a prefix of the space), exit status 130
  const bool cancelled = !r.complete && prog.cancelled();
  if (cancelled) {
    std::printf("Cancelled after %llu / %llu launch ranges (%llu assignments searched)\n",
                (unsigned long long)r.launches, (unsigned long long)prog.steps,
                (unsigned long long)((unsigned long long)r.launches << prog.bits));
    if (a.has("count")) std::printf("Count: %llu (searched ranges only)\n", (unsigned long long)r.count);
  } else if (a.has("count")) {
    std::printf("Count: %llu\n", (unsigned long long)r.count);
  }
  if (cancelled && !r.found) {
    std::printf("No solution in the searched ranges\n");
  } else if (r.found) {
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
  return cancelled ? 130 : 0;
}
