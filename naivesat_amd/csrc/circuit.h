// Gate-circuit IR used by every solver in this library.
//
// Semantics follow the reference's circuit model (gatesim `Circuit<usize>` as used at
// src/main.rs:365,374 and src/qsat.rs:1291-1304): wires 0..n-1 are the inputs, wire n+j
// is gate j, a gate has two operand wires that precede it and one of four functions,
// and every output is a (wire, negated) pair.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace ns {

// Gate functions (same four as the reference's gate set).
enum GateFunc : uint8_t {
  FUNC_AND = 0,    // a & b
  FUNC_NOR = 1,    // ~(a | b)
  FUNC_NIMPL = 2,  // a & ~b
  FUNC_XOR = 3,    // a ^ b
};

struct Gate {
  uint32_t i0, i1;
  uint8_t func;
};

struct Output {
  uint32_t wire;
  bool neg;
};

struct Circuit {
  uint32_t n = 0;              // input count
  std::vector<Gate> gates;     // gate j drives wire n + j
  std::vector<Output> outputs;

  uint32_t wires() const { return n + (uint32_t)gates.size(); }
  // Throws std::runtime_error when a gate refers to itself or to a later wire.
  void validate() const;
  // Scalar evaluation; inputs[k] is input k; returns one bool per output.
  std::vector<bool> eval(const std::vector<bool>& inputs) const;
  // 64-lane evaluation: bit b of inputs[k] is input k of lane b.
  std::vector<uint64_t> eval64(const std::vector<uint64_t>& inputs) const;
};

// Text form of a circuit: "{0 1 2 and(0,1) xor(2,3):0n}" — inputs listed first,
// then gates in wire order as func(a,b), output markers ":k" or ":kn" (negated)
// after the wire that drives output k. Throws std::runtime_error on bad input.
Circuit parse_circuit(const std::string& text);
std::string format_circuit(const Circuit& c);

// Quantified circuit text: quantifier letters ('e' exists, 'a' all, one per input,
// input 0 first) followed by the circuit text, e.g. "eea {0 1 2 ...}".
// Returns false when the text does not start with a quantifier prefix.
bool parse_quant_circuit(const std::string& text, std::vector<uint8_t>& quants_all,
                         Circuit& c);

// Output cone: wires reachable from the selected outputs (all outputs when `outs`
// is empty). mark[w] = 1 for every gate wire in the cone.
std::vector<uint8_t> output_cone(const Circuit& c, const std::vector<uint32_t>& outs = {});

// Evaluation order of the cone's gates chosen to keep few values live at once:
// depth-first post-order from the outputs, deeper operand first.
std::vector<uint32_t> schedule_gates(const Circuit& c, const std::vector<uint8_t>& cone,
                                     const std::vector<uint32_t>& outs = {});

// Peak number of simultaneously live gate values for a given order.
uint32_t peak_live(const Circuit& c, const std::vector<uint32_t>& order,
                   const std::vector<uint32_t>& outs = {});

// ---- synthetic workload generators (deterministic in the seed) ----

// SplitMix64 stream used by all generators.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  // uniform in [0, m) via the high half of a 64x64 product
  uint64_t below(uint64_t m) { return (uint64_t)(((unsigned __int128)next() * m) >> 64); }
  bool bit() { return (next() >> 63) != 0; }
};

// Builder that accepts AND/OR/XOR gates over literals (wire, negated) and lowers
// them to the four gate functions, tracking output polarity of each node.
struct LitBuilder {
  Circuit c;
  std::vector<uint8_t> pol;  // logical value of wire w = physical value ^ pol[w]
  explicit LitBuilder(uint32_t n) { c.n = n; pol.assign(n, 0); }
  // op: 0 AND, 1 OR, 2 XOR. Returns the new node id (a wire).
  uint32_t gate(int op, uint32_t a, bool na, uint32_t b, bool nb);
  void output(uint32_t node, bool neg) { c.outputs.push_back({node, (bool)(pol[node] ^ neg)}); }
};

// Random gate list: `gates` gates, AND/OR/XOR uniform, each operand uniform over the
// inputs and earlier gates of this block and negated with p=0.5. `in` maps the
// block's local inputs to builder nodes. Returns the node ids of the gates.
std::vector<uint32_t> random_block(LitBuilder& lb, Rng& rng, const std::vector<uint32_t>& in,
                                   uint32_t gates);

// SAT workload: n inputs, `gates` random gates, the last `wbits` gates form W(x),
// single output = [W(x) == W(k)] for a key k drawn from the seed.
Circuit gen_sat_circuit(uint64_t seed, uint32_t n, uint32_t gates, uint32_t wbits,
                        uint64_t* key);
// Factoring workload (the kind of circuit the reference's published naivesat runs search,
// benchmarks.txt): inputs a = x0..x(bits-1) and b = x(bits)..x(2 bits - 1), single output =
// [a * b == N] and a != 1 and b != 1, over the full 2 bits-bit product (carry-save rounds of
// full adders, then a ripple carry).
Circuit gen_factor_circuit(uint64_t N, uint32_t bits);
// QSAT workload: n inputs, `gates` random gates, output = XOR of 3 distinct gates
// drawn from the last `last` gates.
Circuit gen_qsat_circuit(uint64_t seed, uint32_t n, uint32_t gates, uint32_t last);
// SATOSTOP workload: state of 2*half bits, `rounds` Feistel rounds whose round
// functions are random `round_gates`-gate blocks on `half` inputs (last `half`
// gates are the outputs), STOP = [next == target]. The target is planted as the
// state reached after m (1..64) steps from a random start u0 << (2*half-unknowns).
Circuit gen_satostop_circuit(uint64_t seed, uint32_t half, uint32_t rounds,
                             uint32_t round_gates, uint32_t unknowns, uint64_t* target);

}  // namespace ns
