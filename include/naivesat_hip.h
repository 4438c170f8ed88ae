/*
 * naivesat_hip — C ABI of the MI355X (gfx950) brute-force SAT / QSAT / SATOSTOP engine.
 *
 * Every entry point returns an int status (NS_OK on success) and never throws or aborts;
 * the message of the last failure on the calling thread is ns_last_error(). Handles are
 * opaque. A context owns one HIP stream and its device buffers; use a context from one
 * thread at a time. Result structs are caller-allocated.
 *
 * The reference (matszpk/naivesat 0.1.3, Rust) reaches its hot path through the external
 * `gatenative` builder/executor API and `gatesim` circuits; each entry point below names
 * the reference interface it replaces.
 */
#ifndef NAIVESAT_HIP_H
#define NAIVESAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NS_OK 0
#define NS_ERR_INVALID 1     /* bad argument or malformed circuit */
#define NS_ERR_HIP 2         /* HIP runtime failure */
#define NS_ERR_COMPILE 3     /* run-time kernel compilation failed */
#define NS_ERR_UNSUPPORTED 4 /* size or device outside what the engine handles */
#define NS_ERR_NOMEM 5       /* device or host allocation failed */
#define NS_ERR_INTERNAL 6
#define NS_ERR_CANCELLED 7 /* a progress callback stopped a search that has no partial answer */

/* gate functions (gatesim's gate set) */
#define NS_FUNC_AND 0   /* a & b    */
#define NS_FUNC_NOR 1   /* ~(a | b) */
#define NS_FUNC_NIMPL 2 /* a & ~b   */
#define NS_FUNC_XOR 3   /* a ^ b    */

typedef struct ns_circuit ns_circuit;
typedef struct ns_ctx ns_ctx;

const char* ns_last_error(void);
const char* ns_version(void);

/* ---- circuits (replace gatesim `Circuit<usize>`: src/main.rs:6,391; src/qsat.rs:6277-6285) ---- */

/* Parses the circuit text form "{0 1 2 and(0,1) xor(2,3):0n}" (replaces
 * `Circuit::<usize>::from_str`, src/main.rs:391). */
int ns_circuit_parse(const char* text, size_t len, ns_circuit** out);
/* Parses "eaa {circuit}" (one quantifier letter per input, input 0 first) into the
 * circuit plus quantifiers (quants_all[k] = 1 for 'a'); *is_quant = 0 and a plain parse
 * when the text has no quantifier prefix (replaces `QuantCircuit::from_str` with the
 * all-exists fallback, src/qsat.rs:6277-6285). quants_all needs room for the input count. */
int ns_quant_circuit_parse(const char* text, size_t len, ns_circuit** out, uint8_t* quants_all,
                           size_t quants_cap, int* is_quant);
/* Builds a circuit from arrays: gate j drives wire n_inputs + j (replaces
 * `Circuit::new`, src/satostop.rs:1793). */
int ns_circuit_from_arrays(uint32_t n_inputs, uint32_t n_gates, const uint32_t* in0,
                           const uint32_t* in1, const uint8_t* func, uint32_t n_outputs,
                           const uint32_t* out_wire, const uint8_t* out_neg, ns_circuit** out);
int ns_circuit_info(const ns_circuit* c, uint32_t* n_inputs, uint32_t* n_gates, uint32_t* n_outputs);
/* Copies the circuit into caller arrays sized from ns_circuit_info (any may be NULL). */
int ns_circuit_to_arrays(const ns_circuit* c, uint32_t* in0, uint32_t* in1, uint8_t* func,
                         uint32_t* out_wire, uint8_t* out_neg);
/* Text form; *needed gets the length without the terminating NUL. */
int ns_circuit_format(const ns_circuit* c, char* buf, size_t cap, size_t* needed);
/* Scalar evaluation on the host (replaces `Circuit::eval`, src/main.rs:374), used to verify
 * answers the way the reference's CLIs do. inputs[k], outputs[i] are 0/1. */
int ns_circuit_eval(const ns_circuit* c, const uint8_t* inputs, uint8_t* outputs);
/* Gates in the output cone and the peak live values of the kernel's evaluation order. */
int ns_circuit_cone(const ns_circuit* c, uint32_t* cone_gates, uint32_t* peak_live);
void ns_circuit_free(ns_circuit* c);

/* ---- synthetic workloads (deterministic in the seed; SplitMix64 stream) ---- */
int ns_gen_sat(uint64_t seed, uint32_t n, uint32_t gates, uint32_t word_bits, ns_circuit** out,
               uint64_t* key);
int ns_gen_qsat(uint64_t seed, uint32_t n, uint32_t gates, uint32_t last, ns_circuit** out);
/* Factoring: inputs a = x0..x(bits-1), b = x(bits)..x(2 bits - 1); output = [a * b ==
 * product] and a != 1 and b != 1 (the kind of circuit the reference's benchmarks.txt runs
 * naivesat on, prime_69690421.circuit: 54 inputs; that circuit itself is not in the
 * reference). 2 <= bits <= 32, product < 2^(2 bits). */
int ns_gen_factor(uint64_t product, uint32_t bits, ns_circuit** out);
int ns_gen_satostop(uint64_t seed, uint32_t half, uint32_t rounds, uint32_t round_gates,
                    uint32_t unknowns, ns_circuit** out, uint64_t* target);

/* ---- device context (replaces OpenCLBuilder::new(&device, cfg) + its command queue,
 *      src/main.rs:292-301) ---- */
/* The HIP devices (replaces get_all_devices(CL_DEVICE_TYPE_GPU), src/main.rs:385-388, that
 * the tools list first): *count devices, their gfx architecture names comma-separated in
 * names (NUL-terminated, cut to cap bytes). No device: count 0. */
int ns_device_list(uint32_t* count, char* names, size_t cap);
int ns_ctx_create(int device, ns_ctx** out);
void ns_ctx_destroy(ns_ctx* ctx);
/* Releases the context's SATOSTOP tables and images (device memory kept between solves so a
 * repeated search reuses it); the next solve allocates again. No counterpart in the
 * reference, whose buffers live as long as its executors. */
int ns_ctx_trim(ns_ctx* ctx);

/* ---- SAT (replaces the naivesat mapper run: BasicMapperBuilder + OpenCLBuilder with
 *      AGGR_OUTPUT_OPENCL_CODE and the per-arg callback, src/main.rs:91-104,157-201) ----
 * Input k of the circuit is bit k of the assignment x; the circuit must have one output.
 * Finds the number of satisfying assignments and the lowest one. */
#define NS_SAT_STOP_AT_FIRST 1u /* stop after the first launch range holding a solution */
/* Group exits (no counterpart in the reference, which evaluates every gate of every
 * assignment): when the kernel evaluates the output's conjunction in groups of its literals
 * (large circuits, from 2^34 assignments), a wave leaves a 32-assignment word as soon as the
 * groups so far are zero in all of its lanes. The answers are the same; the gates skipped are
 * not evaluated, so the search's time no longer measures a gate-evaluation rate. */
#define NS_SAT_GROUP_EXITS 2u
typedef struct {
  uint32_t elem_bits;      /* log2 assignments per launch (0 = engine default) */
  uint32_t group_len;      /* threads per block, multiple of 64 (0 = default) */
  uint32_t words_per_lane; /* 32-assignment words per lane, power of two (0 = default) */
  uint32_t flags;
} ns_sat_opts;
typedef struct {
  uint64_t count;     /* satisfying assignments in the searched range */
  uint64_t lowest;    /* lowest satisfying assignment (valid when found) */
  uint32_t found;
  uint32_t complete;  /* 1 when the whole 2^n space was searched */
  uint32_t launches;
  uint32_t cone_gates;
  uint32_t valu_ops;   /* v_bitop3 LUTs per 32-assignment word */
  uint32_t scalar_ops; /* wave-uniform gates per word (scalar unit) */
  uint32_t backend;    /* 1 = native assembly path, 0 = compiler path */
  uint32_t vgpr_class; /* VGPR budget of the kernel (native path) */
  double kernel_ms;   /* device time of the search launches (HIP events) */
  double compile_ms;  /* run-time compilation (0 when cached) */
  double wall_ms;
} ns_sat_result;
int ns_sat_solve(ns_ctx* ctx, const ns_circuit* c, const ns_sat_opts* opts, ns_sat_result* res);
/* ns_sat_solve with progress and cancellation (replaces execute_direct's per-arg callback and
 * stop predicate, src/main.rs:178-196: the reference prints "Step: arg / arg_steps" from it).
 * The search runs as `steps` launch ranges of 2^range_bits assignments each, range k covering
 * assignments [k << range_bits, (k + 1) << range_bits); elem_bits in opts sets range_bits
 * exactly (from 11 to 40; 0 = the engine default, 2^32). fn is called on the calling thread
 * once with step = NS_SAT_STEP_START when the kernel is built and nothing is launched yet
 * (the reference's "Start execution"), then once per range, in order, when it has completed.
 * A nonzero return stops the search after the ranges already launched: the result covers
 * the first res->launches ranges and res->complete is 0. Answers equal ns_sat_solve's. */
#define NS_SAT_STEP_START 0xFFFFFFFFFFFFFFFFull
typedef int (*ns_sat_progress_fn)(void* user, uint64_t step, uint64_t steps, uint32_t range_bits);
int ns_sat_solve_progress(ns_ctx* ctx, const ns_circuit* c, const ns_sat_opts* opts, ns_sat_progress_fn fn,
                          void* user, ns_sat_result* res);
/* No counterpart in the reference: ns_sat_solve for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It makes no
 * device or host allocation, loads no module and builds no graph: the circuit's kernel must
 * already be built on this context (an earlier ns_sat_solve of the same circuit with the same
 * opts; otherwise NS_ERR_UNSUPPORTED), and the search counters live in a 12 KiB device array
 * the library holds per GPU for the life of the process. Every call on one GPU shares that
 * array and clears it first: calls that overlap on one GPU race. Same results as
 * ns_sat_solve (compile_ms 0; no graph replay, no words-per-lane tuning run). */
int ns_sat_solve_noalloc(ns_ctx* ctx, const ns_circuit* c, const ns_sat_opts* opts, ns_sat_result* res);

/* ---- QSAT (replaces the naiveqsat run: AGGR_OUTPUT_OPENCL_CODE tree + OpenCLQuantReducer +
 *      QuantReducer, src/qsat.rs:172-291,405-720,942-1321,1658-1720) ----
 * quants_all[k] = 1 when input k is universally quantified; input 0 is outermost. The
 * outermost block is the maximal prefix of equal quantifiers (first_bits inputs). For each
 * value u of that block (u's most significant bit = input 0) the formula below it is
 * reduced; truth = OR (exists) / AND (all) over u. The witness is the lowest u whose value
 * decides the truth (1 for exists, 0 for all) — the solution the reference prints. */
typedef struct {
  uint32_t elem_bits;
  uint32_t group_len;
  uint32_t words_per_lane;
  uint32_t flags;
} ns_qsat_opts;
typedef struct {
  uint32_t truth;
  uint32_t first_bits;  /* size of the outermost block */
  uint32_t first_all;   /* outermost quantifier is 'all' */
  uint32_t has_witness;
  uint64_t witness;       /* lowest witness u (index order, input 0 = MSB) */
  uint64_t witness_count; /* number of u values that are witnesses */
  uint32_t passes;        /* reduction passes after the circuit kernel */
  uint32_t cone_gates;
  uint32_t valu_ops;
  uint32_t scalar_ops;
  uint32_t backend;
  uint32_t vgpr_class;
  double kernel_ms;
  double compile_ms;
  double wall_ms;
} ns_qsat_result;
/* bitmap (optional): 2^first_bits bits, bit u = value for block value u; bitmap_words
 * u32 words of room (needs ceil(2^first_bits / 32)). */
int ns_qsat_solve(ns_ctx* ctx, const ns_circuit* c, const uint8_t* quants_all,
                  const ns_qsat_opts* opts, ns_qsat_result* res, uint32_t* bitmap,
                  uint64_t bitmap_words);
/* ns_qsat_solve with per-range progress (the reference's "Step: arg / arg_steps" callback of
 * the QSAT mapper, src/qsat.rs:1634-1654,1701-1718): fn as for ns_sat_solve_progress, over
 * the circuit kernel's launch ranges (the reduction passes follow them). A nonzero return
 * stops the search and the call fails with NS_ERR_CANCELLED: a QSAT answer needs every range. */
int ns_qsat_solve_progress(ns_ctx* ctx, const ns_circuit* c, const uint8_t* quants_all,
                           const ns_qsat_opts* opts, ns_sat_progress_fn fn, void* user, ns_qsat_result* res,
                           uint32_t* bitmap, uint64_t bitmap_words);

/* ---- SATOSTOP on a device-resident next-state table (replaces greedysatostop's
 *      do_solve_with_opencl_builder: transpose epilogue + check_stop / join_nexts /
 *      find_solution, src/greedysatostop.rs:841-1099; and the answer contract of
 *      naivesatostop, src/satostop.rs:322-327,1650-1782) ----
 * The circuit has n inputs (state) and n+1 outputs: outputs 0..n-1 are the next state,
 * output n is STOP. Unknowns are the top `unknowns` state bits; start state of unknown u
 * is u << (n - unknowns). For every u the engine finds whether iterating the circuit from
 * its start reaches STOP and after how many evaluations (the stopping one included). */
#define NS_STOP_FIRST_PASS 1u /* end after the first jump pass where some start stopped */
#define NS_STOP_PASSES 2u     /* full search by passes over the whole table, not waypoint walks */
#define NS_STOP_TABLE_WALKS 4u /* full search by waypoint walks over the table (up to 35 state bits) */
#define NS_STOP_BIT_WALKS 8u   /* full search by waypoint walks without a table (the default from 34 bits) */
typedef struct {
  uint32_t flags;
  uint32_t max_passes;     /* 0 = n (enough for every path) */
  uint32_t group_len;      /* table build block size (0 = default) */
  uint32_t words_per_lane; /* table build words per lane (0 = default) */
} ns_satostop_opts;
typedef struct {
  uint32_t any_stop;   /* some state's STOP output is 1 */
  uint32_t passes;     /* jump passes run over the whole table (P) */
  uint64_t n_stopped;  /* unknown starts that stop */
  uint64_t lowest_u;   /* lowest stopping unknown (valid when n_stopped > 0) */
  uint64_t start;      /* its start state */
  uint64_t end;        /* state produced by the stopping evaluation */
  uint64_t steps;      /* evaluations until STOP, the stopping one included */
  uint32_t cone_gates;
  uint32_t valu_ops;
  uint32_t backend;
  uint32_t vgpr_class;
  double build_ms;     /* table build kernel time */
  double jump_ms;      /* total jump-pass kernel time */
  double compile_ms;
  double wall_ms;
  uint64_t walk_jumps; /* per start, up to this many jumps of 2^P evaluations follow the passes */
  double walk_ms;      /* kernel time of those walks (waypoint method: of the waypoint walks) */
  /* The full search (no NS_STOP_FIRST_PASS, max_passes 0 or n) runs as waypoint walks
   * when the unknowns allow: one walk from every state whose low waypoint_bits bits are
   * zero (the starts among them) to the next such state or STOP, then pointer jumping
   * over the 2^(n - waypoint_bits) walk records (waypoint_passes passes, jump_ms). */
  uint32_t method;          /* 0 passes over the table + walks, 1 waypoint walks over the table,
                             2 waypoint walks without a table (bit-sliced, 32 walks per lane;
                             build_ms 0, any_stop = some start stops) */
  uint32_t waypoint_bits;
  uint32_t waypoint_passes;
  uint32_t reserved;
  uint64_t walk_evaluations; /* waypoint method: table entries read by the waypoint walks */
} ns_satostop_result;
/* steps_out / ends_out (optional): 2^unknowns entries each; steps 0 = never stops.
 * These 32-bit arrays need n <= 31; ns_satostop_solve_wide takes up to 33 state bits. */
int ns_satostop_solve(ns_ctx* ctx, const ns_circuit* c, uint32_t unknowns,
                      const ns_satostop_opts* opts, ns_satostop_result* res, uint32_t* steps_out,
                      uint32_t* ends_out);

/* Same contract for state spaces of up to 62 bits (the reference handles states wider
 * than 32 bits with u64 entries: join_nexts_64, src/greedysatostop.rs:169-194, and
 * HashEntry, src/satostop.rs:168-174), with 64-bit per-start results. From 32 state bits
 * on the full search reads a packed (n+1)-bit image (the reference's MemImage layout;
 * n = 35: 155 GiB); 34 and 35 state bits take only the full search, with at most n - 10
 * unknowns. The passes (first-pass mode, max_passes, fallback) use split u32 tables and
 * u64 jump tables, up to 33 state bits (n = 33: 192 GiB). From 34 state bits on (or with
 * NS_STOP_BIT_WALKS) the full search walks without a table: the circuit is evaluated for 32
 * walks per lane at each step, and only the waypoint records (at most 2^28 of 16 bytes)
 * stay in device memory; unknowns <= n - 8 there. */
int ns_satostop_solve_wide(ns_ctx* ctx, const ns_circuit* c, uint32_t unknowns,
                           const ns_satostop_opts* opts, ns_satostop_result* res, uint64_t* steps_out,
                           uint64_t* ends_out);

/* naivesatostop's hash-map search (src/satostop.rs:1152-1486, 1585-1648) for state spaces
 * beyond the device table (n up to 63): the circuit's outputs for one block of 2^e states
 * at a time are joined into a hash map of 2^H entries that jump along the paths, until some
 * unknown start reaches STOP (a solution) or every unknown start is known to stop or loop
 * (no solution). Which solution is reported depends on the hash map's contents, as in the
 * reference; it is verified by walking the circuit. */
typedef struct {
  uint32_t hashmap_len_bits;             /* H: 2^H entries of 32 bytes (0 = 24, -H) */
  uint32_t elem_bits;                    /* e: 2^e states per block (0 = 24, -e; at most n - 1) */
  int32_t unknown_fill_bits;             /* -U (-1 = max(0, unknowns - (H - 1))) */
  uint32_t max_predecessors;             /* -P (default 0) */
  uint32_t iters_per_clear_predecessors; /* -C (0 = 20) */
  uint32_t max_sweeps;                   /* sweeps over all blocks before giving up (0 = no limit) */
} ns_hashmap_opts;
typedef struct {
  uint32_t found;      /* 1 solution, 0 no solution, 2 max_sweeps reached without either */
  uint32_t sweeps;     /* sweeps over the blocks started */
  uint64_t start, end, steps;
  uint64_t resolved_unknowns; /* unknown-fill groups whose starts all stop or loop */
  uint64_t blocks;     /* blocks processed */
  double compile_ms, kernel_ms, wall_ms;
} ns_hashmap_result;
int ns_satostop_hashmap_solve(ns_ctx* ctx, const ns_circuit* c, uint32_t unknowns, const ns_hashmap_opts* opts,
                              ns_hashmap_result* res);
/* ns_satostop_hashmap_solve with per-block progress (the reference's per-arg callback,
 * src/satostop.rs:1548-1558, which prints "Step: arg / arg_steps" and "Resolved unknowns
 * entries: ru"): fn(user, step, steps, resolved_unknowns) after each block of 2^elem_bits
 * states, in order (steps = blocks per sweep; a new sweep starts again at 0). A nonzero
 * return stops the search: NS_ERR_CANCELLED. */
typedef int (*ns_hashmap_progress_fn)(void* user, uint64_t step, uint64_t steps, uint64_t resolved_unknowns);
int ns_satostop_hashmap_solve_progress(ns_ctx* ctx, const ns_circuit* c, uint32_t unknowns,
                                       const ns_hashmap_opts* opts, ns_hashmap_progress_fn fn, void* user,
                                       ns_hashmap_result* res);

/* Diagnostic (one device step of the hash-map search on caller arrays, for checking it
 * against the reference's data sets): which = 0 join_to_hashmap (src/satostop.rs:186-266),
 * 1 join_hashmap_itself_and_check_solution (391-594, from hm into hm_out), 2
 * add_to_hashmap_and_check_solution (713-942, `test` skips states whose next state is 0 as
 * the reference's test routine does), 3 join_to_hashmap then join_hashmap_itself as the
 * search runs them, in one pass (from hm into hm_out). hm / hm_out: hm_len entries of 32 bytes (HashEntry);
 * outputs: the block's output words, two per state when state_len + 1 > 32; fills: the
 * unknown-fill counters; sol: {resolved_unknowns, start, end, steps: u64; defined: u32}. */
int ns_hashmap_step(ns_ctx* ctx, uint32_t which, uint32_t state_len, uint32_t arg_bit_place, uint64_t arg,
                    uint32_t unknown_bits, uint32_t unknown_fill_bits, uint32_t max_predecessors, uint32_t test,
                    const uint32_t* outputs, uint64_t output_words, void* hm, void* hm_out, uint64_t hm_len,
                    uint32_t* fills, uint64_t fills_len, void* sol);

/* greedysatostop's partitioned search (do_solve_with_cpu_builder_with_partitions,
 * src/greedysatostop.rs:692-837) with the next-state image on disk (FileImage, 478-594, in
 * `<file_prefix>greedy_temp_<ns>`, removed afterwards): the packed image is written chunk by
 * chunk, then for stage = 0..n-1 and every partition i: partition i is joined in place with
 * every partition j in order (j = i: a copy of its current self; otherwise j as last saved),
 * searched for its lowest stopping unknown start, and saved. The joins run on the device, a
 * partition (2^n / partitions packed entries) at a time in HBM. */
typedef struct {
  uint32_t found;      /* 1: start/end hold the solution; 0: none */
  uint32_t stages;     /* stages begun (0 when no state's STOP output is 1) */
  uint64_t start, end;
  uint64_t partitions;
  double build_ms;     /* table build and image write */
  double join_ms;      /* device join kernels */
  double io_ms;        /* partition reads, writes and host-device copies */
  double compile_ms;
  double wall_ms;
} ns_partitioned_result;
int ns_greedy_partitioned_solve(ns_ctx* ctx, const ns_circuit* c, uint32_t unknowns, uint64_t partitions,
                                const char* file_prefix, ns_partitioned_result* res);

/* FileImage on its own (src/greedysatostop.rs:478-594; host only, no device): create,
 * append chunks of packed words, then read and rewrite partitions. */
typedef struct ns_file_image ns_file_image;
int ns_file_image_create(uint32_t state_len, uint64_t partitions, const char* prefix, ns_file_image** out);
int ns_file_image_save_chunk(ns_file_image* fi, const uint64_t* words, uint64_t nwords);
int ns_file_image_load_partition(ns_file_image* fi, uint64_t part, uint64_t* words, uint64_t words_cap);
int ns_file_image_save_partition(ns_file_image* fi, uint64_t part, const uint64_t* words, uint64_t nwords);
void ns_file_image_free(ns_file_image* fi);

/* Table build alone (replaces greedysatostop's first-nexts launch with the transpose
 * epilogue, src/greedysatostop.rs:1042-1072; identity layout pinned by
 * src/satostop.rs:1788-1915): table_out gets 2^n u32, entry s = sum_i output_i(s) << i.
 * At most 32 outputs and n <= 31. */
int ns_table_build(ns_ctx* ctx, const ns_circuit* c, uint32_t* table_out, uint64_t table_len,
                   double* kernel_ms);

/* ---- Serving (no counterpart in the reference, which solves one circuit per process):
 *      requests taken as they arrive and solved by `workers` lanes (a host thread, a HIP
 *      stream and buffers each) that run concurrently on the device. A free lane takes the
 *      oldest admissible request: SATOSTOP requests are admitted when their tables fit in
 *      the device arena (a request waiting for memory holds back only the memory-needing
 *      requests behind it); with batching on, a lane takes every waiting SAT / QSAT request
 *      (up to batch_max) and solves them together in one persistent kernel whose code
 *      object holds all their circuit bodies. Every answer is the one the request gets
 *      alone. ---- */
#define NS_REQ_SAT 0u
#define NS_REQ_QSAT 1u
#define NS_REQ_SATOSTOP 2u
typedef struct {
  uint32_t kind;
  uint32_t unknowns;          /* SATOSTOP: unknown state bits (at most 31 state bits served) */
  const ns_circuit* circuit;
  const uint8_t* quants_all;  /* QSAT: one quantifier per input */
  double arrival_s;           /* ns_serve: arrival time after the start of the run */
} ns_request;
typedef struct {
  int32_t status;             /* NS_OK or the error of this request */
  uint32_t worker;
  /* SAT: count, lowest, found; QSAT: truth, witness, witness_count, has_witness;
   * SATOSTOP: n_stopped, lowest_u, steps, end */
  uint64_t v0, v1, v2, v3;
  double start_s, finish_s;   /* service start and completion after the start of the server */
  double compile_ms, kernel_ms;
  double submit_s;            /* arrival (submission) after the start of the server */
} ns_request_result;
typedef struct {
  uint32_t workers;           /* concurrent lanes (0 = 8, at most 64) */
  uint32_t realtime;          /* ns_serve: submit each request at its arrival_s */
  uint64_t hbm_budget_bytes;  /* device arena for SATOSTOP tables; ns_serve: 0 = workers x the
                                 largest request's tables, at most 16 GiB; ns_serve_open: 0 =
                                 4 GiB */
  uint32_t batch_max;         /* > 1: batch SAT / QSAT requests, at most this many per kernel */
  uint32_t reserved;
} ns_serve_opts;
/* Replays a trace: requests submitted at their arrival times (realtime) or all at once,
 * every result waited for; *total_s = time to the last completion. */
int ns_serve(ns_ctx* ctx, const ns_request* reqs, uint32_t n, const ns_serve_opts* opts,
             ns_request_result* results, double* total_s);

/* Online serving: open a server on a context, submit requests as they arrive (each gets a
 * ticket; its arrival is the submit time), poll or wait for results, close. The server
 * copies each request's circuit and quantifiers, so the caller may free them after
 * ns_serve_submit returns. ns_serve_poll / ns_serve_wait set *done = 1 and fill *res once
 * the request has finished; a finished ticket's result is handed out once, after which the
 * ticket is unknown (NS_ERR_INVALID). ns_serve_wait with timeout_s < 0 waits without limit.
 * ns_serve_close waits for every submitted request, stops the lanes and frees the server.
 * A server may be used from several threads; its context's own solves may run meanwhile
 * from another thread. */
typedef struct ns_server ns_server;
int ns_serve_open(ns_ctx* ctx, const ns_serve_opts* opts, ns_server** out);
int ns_serve_submit(ns_server* srv, const ns_request* req, uint64_t* ticket);
int ns_serve_poll(ns_server* srv, uint64_t ticket, ns_request_result* res, int* done);
int ns_serve_wait(ns_server* srv, uint64_t ticket, double timeout_s, ns_request_result* res, int* done);
int ns_serve_close(ns_server* srv);

/* Many SAT / QSAT circuits solved together by the persistent batch kernels on the
 * context's lane (no counterpart in the reference: its mapper runs one circuit per
 * process). SAT results as ns_sat_solve's count / lowest / found, QSAT as ns_qsat_solve's
 * truth / witness / witness_count / has_witness, in v0..v3 of results[i]; requests of
 * kind SATOSTOP, and circuits the batch kernels cannot take, are solved one at a time.
 * The context keeps its last four batches: the same batch again (same circuits, kinds and
 * prefixes, in order) skips code generation and the module load and runs as a HIP graph,
 * captured on the first repeat and replayed afterwards.
 * info (optional, 10 doubles): requests, fallbacks, largest register class, code bytes, code
 * generation ms, module load ms, kernel ms, wall ms, graph launches of this batch so far,
 * batch kernel launches (one per kind and register class present). */
int ns_batch_solve(ns_ctx* ctx, const ns_request* reqs, uint32_t n, ns_request_result* results, double* info);

/* Diagnostic (no counterpart in the reference; no device needed): builds the native-path
 * kernel for a circuit, assembles and links it, and returns the circuit body's assembly.
 * kind: 0 SAT (output 0), 1 QSAT one-bit-per-thread, 2 QSAT outer-block-in-thread,
 * 3 next-state table (all outputs), 4 batch body with the SAT input order, 5 batch body with
 * the QSAT input order, 6 SATOSTOP walk body (two walk steps and their bookkeeping; lane_shift
 * = the waypoint bits k), 7 SAT with NS_SAT_GROUP_EXITS. Input k sits at index bit k (SAT,
 * table, 4, 7) or n-1-k (QSAT, 5);
 * lane_shift places the six lane bits at index bits 5+lane_shift..10+lane_shift.
 * Returns NS_ERR_UNSUPPORTED when the circuit exceeds the native register budget. */
int ns_native_body(const ns_circuit* c, uint32_t kind, uint32_t lane_shift, char* body, size_t cap,
                   size_t* needed, uint32_t* vgpr_class, uint64_t* code_bytes);

/* Diagnostic (no counterpart in the reference; no device needed): encodes native body text
 * (the instruction subset ns_native_body returns) into gfx950 machine-code words.
 * Returns NS_ERR_UNSUPPORTED for an instruction outside that subset. */
int ns_native_encode(const char* body, size_t len, uint32_t* words, size_t cap, size_t* n_words);

/* Diagnostic (no counterpart in the reference; no device needed): the code object the
 * engine loads for a circuit's native kernel (kind and lane_shift as for ns_native_body):
 * the template's cached code object with the encoded body copied into its body slot at
 * byte *slot_offset (*patched = 1), or the fully assembled module (*patched = 0) when the
 * body does not fit the largest slot. Kinds 4 and 5: the batch code object (the batch
 * kernels of every register class) with this one body in its region, *slot_offset = the
 * body's first byte. Kind 6: the walk template's code object with the walk body. */
int ns_native_code(const ns_circuit* c, uint32_t kind, uint32_t lane_shift, void* code, size_t cap,
                   size_t* needed, uint32_t* patched, uint64_t* slot_offset);

/* Per-pass timing of the last ns_satostop_solve on this context: pass_ms[i] for jump pass i
 * (HIP events on the context stream). Returns the number of passes written. */
int ns_satostop_pass_times(ns_ctx* ctx, double* pass_ms, uint32_t cap, uint32_t* written);

/* Diagnostic: the device's packing of next-state entries into the reference's MemImage
 * layout (src/greedysatostop.rs:339-400: entry i in bits [i*width, (i+1)*width) of
 * little-endian u64 words), as the packed SATOSTOP image is built: values[i] masked to width
 * bits, count a multiple of 64, count * width / 64 words out. */
int ns_pack_entries(ns_ctx* ctx, uint32_t width, const uint64_t* values, uint64_t count, uint64_t* words,
                    uint64_t words_cap);

/* Measurement probe, not part of the search path (bench.py uses it as the ceiling for the
 * waypoint walks; no counterpart in the reference): the rate of scattered 4-byte reads from a
 * 2^table_bits-entry device table, lanes x reads_per_lane reads at independent
 * pseudo-random indices (flags bit 0: with the non-temporal load hint); *ms = kernel time
 * of one run (HIP events), *reads = reads made. The ceiling against which the waypoint
 * walks' table reads are measured. */
int ns_probe_gather(ns_ctx* ctx, uint32_t table_bits, uint32_t reads_per_lane, uint32_t lanes, uint32_t flags,
                    double* ms, uint64_t* reads);

#ifdef __cplusplus
}
#endif
#endif /* NAIVESAT_HIP_H */
