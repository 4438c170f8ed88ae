// naivesatostop: SATOSTOP answer (start, end, steps) on one HIP device.
// Command line and printed lines follow the reference's naivesatostop
// (src/satostop.rs:62-99,1650-1782,4443-4467). The search finds every unknown's answer (and
// prints the lowest stopping one): on the device-resident next-state table up to 35 state
// bits, by walks that evaluate the circuit without a table up to 62 (at most 28 unknowns);
// otherwise, or with --hashmap-search, it runs the reference's hash-map search (any stopping
// unknown) with the -e/-H/-U/-P/-C options.
#include "cli_common.h"

static const char* kUsage =
    "Usage: naivesatostop [OPTIONS] --exec-type <EXEC_TYPE> <CIRCUIT> <UNKNOWNS>\n"
    "  -S, --simple                          lowest stopping unknown, printed as 'Stopped at'\n"
    "  -e, --elem-inputs <N>                 states per block of the hash-map search [default: 24]\n"
    "  -H, --hashmap-len-bits <N>            log2 of hash-map entries [default: 24]\n"
    "  -U, --unknown-fill-bits <N>           [default: unknowns - (H - 1), at least 0]\n"
    "  -P, --max-predecessors <N>            [default: 0]\n"
    "  -C, --iters-per-clear-predecessors <N> [default: 20]\n"
    "  -M, --hashmap-search                  hash-map search also where the table would do\n"
    "  -t, --exec-type <TYPE>                cpu|opencl:N|hip:N (all run on one HIP device)\n"
    "  -n, --no-optimize-negs                accepted for compatibility\n"
    "  -G, --opencl-group-len <G>            threads per block of the table build\n"
    "  -v, --verify                          verify the solution by walking the circuit\n";

int main(int argc, char** argv) {
  auto a = cli::parse_args(argc, argv,
                           {{'S', "simple", false},
                            {'e', "elem-inputs", true},
                            {'H', "hashmap-len-bits", true},
                            {'U', "unknown-fill-bits", true},
                            {'P', "max-predecessors", true},
                            {'C', "iters-per-clear-predecessors", true},
                            {'t', "exec-type", true},
                            {'n', "no-optimize-negs", false},
                            {'G', "opencl-group-len", true},
                            {'v', "verify", false},
                            {'M', "hashmap-search", false}},
                           kUsage);
  if (a.positional.size() != 2) cli::die(std::string("expected <CIRCUIT> <UNKNOWNS>\n") + kUsage);
  if (!a.has("exec-type")) cli::die(std::string("--exec-type is required\n") + kUsage);
  cli::print_devices();
  std::string text = cli::read_file(a.positional[0]);
  ns_circuit* c = nullptr;
  cli::check(ns_circuit_parse(text.data(), text.size(), &c), "circuit");
  uint32_t n = 0, g = 0, no = 0;
  cli::check(ns_circuit_info(c, &n, &g, &no), "circuit");
  if (no != n + 1) cli::die("circuit must have one output more than inputs");
  long long U = std::atoll(a.positional[1].c_str());
  if (U < 0 || U > (long long)n) cli::die("unknowns must not exceed the inputs");
  const bool simple = a.has("simple");
  // the per-start search: the device table up to 35 state bits (packed from 32 on, with at
  // most n - 10 unknowns), then walks without a table (at most 28 unknowns)
  const bool table_ok = n <= 33 || (n <= 35 && U <= (long long)n - 10) || (n <= 62 && U <= 28);
  const bool hashmap = !simple && (!table_ok || a.has("hashmap-search"));
  if (simple && !table_ok) cli::die("--simple needs the per-start search (at most 28 unknowns over 35 state bits)");
  if (hashmap && n < 10) cli::die("the hash-map search takes at least 10 state bits");
  int dev = cli::device_from_exec_type(a.get("exec-type"), !simple);
  const long long H = a.num("hashmap-len-bits", 24);
  long long e = a.num("elem-inputs", 24);
  if (e >= (long long)n) e = n - 1;
  // calculate_default_unknown_fill_bits, src/satostop.rs:101-107
  const long long ufb = a.num("unknown-fill-bits", H - 1 < U ? U - (H - 1) : 0);
  if (!simple) {
    std::printf("Elem inputs: %lld\n", e);
    if (hashmap) std::printf("HashMap length: %llu bytes, bits=%lld\n", (1ull << H) * 32ull, H);
    std::printf("Unknowns: %lld\n", U);
    if (hashmap) std::printf("Unknown fill bits: %lld\n", ufb);
    std::printf("Execute in HIP device=%d\n", dev);
  }
  ns_ctx* ctx = nullptr;
  cli::check(ns_ctx_create(dev, &ctx), "device");
  ns_satostop_result r{};
  if (hashmap) {
    ns_hashmap_opts ho{};
    ho.hashmap_len_bits = (uint32_t)H;
    ho.elem_bits = (uint32_t)e;
    ho.unknown_fill_bits = (int32_t)ufb;
    ho.max_predecessors = (uint32_t)a.num("max-predecessors", 0);
    ho.iters_per_clear_predecessors = (uint32_t)a.num("iters-per-clear-predecessors", 20);
    ns_hashmap_result hr{};
    // per block: the reference's "Step: arg / arg_steps" and resolved unknowns lines
    // (src/satostop.rs:1548-1558); Ctrl-C stops the search
    cli::Progress prog;
    prog.print = !simple;
    const int st = ns_satostop_hashmap_solve_progress(
        ctx, c, (uint32_t)U, &ho,
        [](void* u, uint64_t step, uint64_t steps, uint64_t resolved) -> int {
          auto* p = static_cast<cli::Progress*>(u);
          if (p->print) {
            std::printf("Step: %llu / %llu\n", (unsigned long long)step, (unsigned long long)steps);
            std::printf("Resolved unknowns entries: %llu\n", (unsigned long long)resolved);
            std::fflush(stdout);
          }
          return p->cancelled() ? 1 : 0;
        },
        &prog, &hr);
    if (st == NS_ERR_CANCELLED) {
      std::printf("Cancelled: no result\n");
      ns_ctx_destroy(ctx);
      ns_circuit_free(c);
      return 130;
    }
    cli::check(st, "search");
    r.n_stopped = hr.found == 1 ? 1 : 0;
    r.start = hr.start;
    r.end = hr.end;
    r.steps = hr.steps;
    r.wall_ms = hr.wall_ms;
    r.compile_ms = hr.compile_ms;
  } else {
    ns_satostop_opts o{};
    o.group_len = (uint32_t)a.num("opencl-group-len", 0);
    cli::check(ns_satostop_solve(ctx, c, (uint32_t)U, &o, &r, nullptr, nullptr), "search");
  }
  if (simple) {
    // src/satostop.rs:1776-1781: decimal, zero-padded to the unknown / state widths
    if (r.n_stopped)
      std::printf("Stopped at %0*llu %0*llu\n", (int)U, (unsigned long long)r.lowest_u, (int)n,
                  (unsigned long long)r.end);
    else
      std::printf("Unsatisfiable!\n");
  } else {
    // wall time after the kernel build (the reference starts its clock after build(),
    // src/satostop.rs:1544-1565)
    std::printf("Time: %g\n", (r.wall_ms - r.compile_ms) / 1000.0);
    if (r.n_stopped) {
      std::printf("Solution: Solution { start: %llu, end: %llu, steps: %llu }\n", (unsigned long long)r.start,
                  (unsigned long long)r.end, (unsigned long long)r.steps);
      if (a.has("verify")) {
        unsigned long long s = r.start, steps = 0;
        std::vector<uint8_t> in(n), out(n + 1);
        for (unsigned long long step = 0; step <= (1ull << n); step++) {
          for (uint32_t k = 0; k < n; k++) in[k] = (uint8_t)((s >> k) & 1ull);
          cli::check(ns_circuit_eval(c, in.data(), out.data()), "verify");
          s = 0;
          for (uint32_t k = 0; k < n; k++) s |= (unsigned long long)out[k] << k;
          steps++;
          if (out[n]) break;
        }
        std::printf("Verified: end=%llu steps=%llu\n", s, steps);
        if (s != r.end || steps != r.steps) std::printf("INCORRECT\n");
      }
    } else {
      std::printf("Unsatisfiable!\n");
    }
  }
  ns_ctx_destroy(ctx);
  ns_circuit_free(c);
  return 0;
}
