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

// One operation of a mapped circuit. lut = 1: dst = v_bitop3(src0, src1, src2) with
// truth table fn (bit (s0<<2 | s1<<1 | s2)); lut = 0: wave-uniform two-input gate
// dst = func fn (FUNC_*) of (src0, src1), evaluated with scalar instructions.
struct LutOp {
  uint8_t lut = 0;
  uint8_t fn = 0;
  uint32_t dst = 0;
  uint32_t src[3] = {0, 0, 0};
};

struct LutProgram {
  std::vector<LutOp> ops;        // evaluation order
  std::vector<uint8_t> uniform;  // per wire: value is the same in all lanes of a wave
  uint32_t valu_ops = 0;
  uint32_t scalar_ops = 0;
};

// Maps the gates of `order` onto 3-input LUTs (single-use gates folded into their consumer
// while the function keeps at most three leaves). uni_input[k] marks wave-uniform inputs;
// gates over uniform values only become scalar ops. out_wires are always materialised.
LutProgram map_luts(const Circuit& c, const std::vector<uint32_t>& order, const std::vector<uint32_t>& out_wires,
                    const std::vector<uint8_t>& uni_input);

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

// SATOSTOP waypoint walks without a table: kernel ns_walk(u64 count, u64 limit, u64 budget,
// ulonglong2* out, u32* long_flag, u64* evals). Each lane carries 32 walks in the bits of its
// n state registers (bit b of register i = bit i of walk b's state) and advances all of them
// with one evaluation of the circuit (n inputs, n + 1 outputs, STOP last). A walk starts at
// waypoint idx (state idx << k, idx < count; the lane's walks are idx = lane + q * lanes,
// q = 0, 1, ...) and ends at the first STOP, at a next state whose low k bits are zero, at
// a repeated state (Brent's test against its state saved after 1, 2, 4, ... evaluations) or
// after `limit` evaluations; then out[idx] = {target | kind << 62, evaluations} with the
// kinds and targets of the waypoint records (0 next waypoint, 1 STOP, 2 never, 3 too long,
// which also sets *long_flag), and the slot takes the lane's next walk. A lane stops after
// `budget` evaluations (raising the flag) and every lane once it sees the flag. evals[64
// slots, 16 words apart] += evaluations done.
std::string gen_walk_kernel(const Circuit& c, uint32_t k, GenInfo* info);

}  // namespace ns
