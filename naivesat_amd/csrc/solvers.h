// Solver drivers: SAT, QSAT and SATOSTOP searches on one device context.
#pragma once
#include <string>
#include <vector>

#include "../../include/naivesat_hip.h"
#include "circuit.h"
#include "runtime.h"

namespace ns {

// 128-bit fingerprint of a circuit's content (hex): the kernel caches' key
std::string circuit_fingerprint(const Circuit& c);

// progress: optional per-range callback (ns_sat_solve_progress); a nonzero return stops the
// search after the ranges already launched
void sat_solve(Lane& ctx, const Circuit& c, const ns_sat_opts& o, ns_sat_result& r,
               ns_sat_progress_fn progress = nullptr, void* user = nullptr);
// sat_solve without any allocation: the kernel must be in the context's cache (built by an
// earlier sat_solve with the same options), counters in `scratch` (device address)
void sat_solve_noalloc(Lane& ctx, const Circuit& c, const ns_sat_opts& o, ns_sat_result& r,
                       unsigned long long* scratch);
// progress: as for sat_solve; a cancelled QSAT search fails with ST_CANCELLED (no answer)
void qsat_solve(Lane& ctx, const Circuit& c, const std::vector<uint8_t>& quants_all,
                const ns_qsat_opts& o, ns_qsat_result& r, uint32_t* bitmap, uint64_t bitmap_words,
                ns_sat_progress_fn progress = nullptr, void* user = nullptr);
// steps_out / ends_out (optional): 2^unknowns host entries each
void satostop_solve(Lane& ctx, const Circuit& c, uint32_t unknowns, const ns_satostop_opts& o,
                    ns_satostop_result& r, uint64_t* steps_out, uint64_t* ends_out);
double table_build(Lane& ctx, const Circuit& c, uint32_t* table_out, uint64_t table_len);
// Bytes satostop_solve's default full search (no flags) may need in ctx.tab_a .. tab_e for
// n <= 31 state bits: table, two record or jump-table buffers, per-start results.
void satostop_buffer_plan(uint32_t n, uint32_t unknowns, uint64_t bytes[5]);

// naivesatostop's hash-map search (state spaces beyond the device table)
// progress (optional): called after each block of the sweep, in order; nonzero cancels
void satostop_hashmap_solve(Lane& ctx, const Circuit& c, uint32_t unknowns, const ns_hashmap_opts& o,
                            ns_hashmap_result& r, ns_hashmap_progress_fn progress = nullptr, void* user = nullptr);
// one device step of it on host arrays (diagnostic; see ns_hashmap_step)
void hashmap_step(Lane& ctx, uint32_t which, uint32_t state_len, uint32_t arg_bit_place, uint64_t arg,
                  uint32_t unknown_bits, uint32_t unknown_fill_bits, uint32_t max_predecessors, bool test,
                  const uint32_t* outputs, uint64_t output_words, void* hm, void* hm_out, uint64_t hm_len,
                  uint32_t* fills, uint64_t fills_len, void* sol);

// greedysatostop's partitioned search with the image on disk
void greedy_partitioned_solve(Lane& ctx, const Circuit& c, uint32_t unknowns, uint64_t partitions,
                              const std::string& file_prefix, ns_partitioned_result& r);

// per-pass jump times of the last satostop_solve on a context
const std::vector<double>& last_pass_times(Lane& ctx);

}  // namespace ns
