// Solver drivers: SAT, QSAT and SATOSTOP searches on one device context.
#pragma once
#include <vector>

#include "../../include/naivesat_hip.h"
#include "circuit.h"
#include "runtime.h"

namespace ns {

void sat_solve(Lane& ctx, const Circuit& c, const ns_sat_opts& o, ns_sat_result& r);
void qsat_solve(Lane& ctx, const Circuit& c, const std::vector<uint8_t>& quants_all,
                const ns_qsat_opts& o, ns_qsat_result& r, uint32_t* bitmap, uint64_t bitmap_words);
// steps_out / ends_out (optional): 2^unknowns host entries each
void satostop_solve(Lane& ctx, const Circuit& c, uint32_t unknowns, const ns_satostop_opts& o,
                    ns_satostop_result& r, uint64_t* steps_out, uint64_t* ends_out);
double table_build(Lane& ctx, const Circuit& c, uint32_t* table_out, uint64_t table_len);

// per-pass jump times of the last satostop_solve on a context
const std::vector<double>& last_pass_times(Lane& ctx);

}  // namespace ns
