// HIP source generation for circuit-specialised kernels (compiled at run time).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "circuit.h"

namespace ns {

// Launch geometry shared by all generated kernels.
//   word w (32 assignments, bit b of the word <-> assignment index 32*w + b)
//   w = w0 + ((blockIdx.x * BLOCK/64 + wave) * 64 * WPL) + (lane << lane_shift) + (j << j_shift)
// lane_shift/j_shift are (0, 6) when consecutive lanes take consecutive words and
// (log2 WPL, 0) when the WPL words of one lane are consecutive.
struct Geometry {
  uint32_t block = 256;  // threads per block (multiple of 64)
  uint32_t wpl = 4;      // words per lane (power of two)
  uint32_t lane_shift = 0;
  uint32_t j_shift = 6;
  uint32_t words_per_block() const { return block * wpl; }
};

struct GenInfo {
  uint32_t cone_gates = 0;  // gates actually emitted
  uint32_t peak_live = 0;   // peak live values of the emitted order
  uint32_t valu_ops = 0;    // v_bitop3 LUT operations emitted per 32-assignment word
  uint32_t scalar_ops = 0;  // wave-uniform gates (scalar instructions, once per wave and word)
};

// SAT: count of satisfying assignments and lowest one.
// kernel ns_sat(u64 w0, u64 wend, unsigned long long* res): res[0] += count,
// res[1] = min(res[1], lowest).
std::string gen_sat_kernel(const Circuit& c, const Geometry& g, GenInfo* info);

// QSAT, first phase. `quants_all[k]` is 1 when input k is universally quantified.
// Enumeration index v has input 0 as its most significant bit (input n-1 is bit 0).
// L = number of levels below the outermost quantifier block.
//  * L >= 5 + log2(WPL): kernel ns_qsat(u64 w0, u64 wend, unsigned long long* bits):
//    one reduced bit per thread, bit t of `bits` for thread index t (w >> log2 WPL).
//  * otherwise kernel ns_qsat(u64 w0, u64 wend, unsigned long long* stats, u32* bitmap):
//    the outermost block is reached inside the thread; stats[0] += witness count,
//    stats[1] = min(lowest witness), bitmap (may be null) gets bit u for block value u.
std::string gen_qsat_kernel(const Circuit& c, const std::vector<uint8_t>& quants_all,
                            const Geometry& g, GenInfo* info);

// Table build: for every state s (input k = bit k of s) writes
// table[s] = sum_i output_i(s) << i (at most 32 outputs) and sets *anylast when the
// last output is 1 for any state.
// kernel ns_table(u64 w0, u64 wend, u32* table, u32* anylast)
std::string gen_table_kernel(const Circuit& c, const Geometry& g, GenInfo* info);

}  // namespace ns
