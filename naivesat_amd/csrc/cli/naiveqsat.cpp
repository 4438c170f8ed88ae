// naiveqsat: quantified SAT of a one-output gate circuit on one HIP device.
// Command line and printed lines follow the reference's naiveqsat
// (src/qsat.rs:74-96,156-170,1838-2000,6270-6288).
#include "cli_common.h"

static const char* kUsage =
    "Usage: naiveqsat [OPTIONS] --exec-type <EXEC_TYPE> <CIRCUIT>\n"
    "  -e, --elem-inputs <N>      log2 of assignments per launch range, 13..64, kept exactly up to 40\n"
    "                             [default: 32; the reference's 24 leaves the GPU idle between launches]\n"
    "  -t, --exec-type <TYPE>     cpu|opencl:N|hip:N|cpu_and_opencl* (all run on one HIP device)\n"
    "  -n, --no-optimize-negs     accepted for compatibility\n"
    "  -G, --opencl-group-len <G> threads per block\n"
    "  -W, --words-per-lane <W>   32-assignment words per lane\n"
    "Ctrl-C stops the search after the launch ranges in flight (no answer).\n";

int main(int argc, char** argv) {
  auto a = cli::parse_args(argc, argv,
                           {{'e', "elem-inputs", true},
                            {'t', "exec-type", true},
                            {'n', "no-optimize-negs", false},
                            {'G', "opencl-group-len", true},
                            {'W', "words-per-lane", true}},
                           kUsage);
  if (a.positional.size() != 1) cli::die(std::string("expected one circuit file\n") + kUsage);
  if (!a.has("exec-type")) cli::die(std::string("--exec-type is required\n") + kUsage);
  cli::print_devices();
  std::string text = cli::read_file(a.positional[0]);
  ns_circuit* c = nullptr;
  std::vector<uint8_t> q(text.size() + 1, 0);
  int is_quant = 0;
  cli::check(ns_quant_circuit_parse(text.data(), text.size(), &c, q.data(), q.size(), &is_quant), "circuit");
  if (!is_quant) std::printf("Load as circuit with quantifiers\n");
  uint32_t n = 0, g = 0, no = 0;
  cli::check(ns_circuit_info(c, &n, &g, &no), "circuit");
  if (no != 1) cli::die("circuit must have exactly one output");
  long long e = a.num("elem-inputs", 32);
  int dev = cli::device_from_exec_type(a.get("exec-type"), n >= 14);
  long long G = a.num("opencl-group-len", 256);
  if (n >= 14) {
    if (e >= (long long)n) e = n - 1;
    if (!(e > 12 && e <= 64)) cli::die("elem inputs must be in 13..64");
    std::printf("Elem inputs: %lld\n", e);
    std::printf("Execute in HIP device=%d\n", dev);
    std::printf("GroupLen: %lld\n", G);
  }
  ns_ctx* ctx = nullptr;
  cli::check(ns_ctx_create(dev, &ctx), "device");
  ns_qsat_opts o{};
  o.elem_bits = n >= 14 ? (uint32_t)e : 0u;
  o.group_len = (uint32_t)G;
  o.words_per_lane = (uint32_t)a.num("words-per-lane", 0);
  cli::Progress prog;
  prog.print = n >= 14;
  prog.elem = n >= 14 ? e : -1;
  ns_qsat_result r{};
  int st = ns_qsat_solve_progress(ctx, c, q.data(), &o, cli::progress_step, &prog, &r, nullptr, 0);
  if (st == NS_ERR_CANCELLED) {
    std::printf("Cancelled after %llu / %llu launch ranges: no result\n", (unsigned long long)prog.done,
                (unsigned long long)prog.steps);
    ns_ctx_destroy(ctx);
    ns_circuit_free(c);
    return 130;
  }
  cli::check(st, "search");
  if (n >= 14) std::printf("Time: %g\n", prog.seconds());
  // FinalResult display: witness printed with bit b = input b, most significant first
  std::string res;
  if (r.has_witness) {
    unsigned long long sol = 0;
    for (uint32_t b = 0; b < r.first_bits; b++)
      if ((r.witness >> (r.first_bits - 1 - b)) & 1ull) sol |= 1ull << b;
    res = std::string(r.first_all ? "(Unsatisfiable: " : "(Satisfiable: ") +
          cli::bits_msb_first(sol, r.first_bits) + ")";
  } else {
    res = r.first_all ? "(Satisfiable)" : "(Unsatisfiable)";
  }
  std::printf("Result: %s\n", res.c_str());
  ns_ctx_destroy(ctx);
  ns_circuit_free(c);
  return 0;
}
