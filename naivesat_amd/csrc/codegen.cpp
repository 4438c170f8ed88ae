// Circuit -> straight-line HIP kernel source.
//
// Every generated kernel evaluates the circuit bit-sliced: one 32-bit word holds 32
// consecutive assignments. The five lowest index bits are the constant masks
// 0xAAAAAAAA .. 0xFFFF0000 (the in-word enumeration the reference's generated kernels
// use, src/qsat.rs:1299-1301), six index bits come from the lane id, the j-loop bits
// and everything above are wave-uniform and become scalar (SGPR) operands.
#include "codegen.h"

#include <sstream>
#include <stdexcept>

namespace ns {

namespace {

const uint32_t kWordMasks[5] = {0xAAAAAAAAu, 0xCCCCCCCCu, 0xF0F0F0F0u, 0xFF00FF00u, 0xFFFF0000u};

uint32_t log2u(uint32_t x) {
  uint32_t r = 0;
  while ((1u << r) < x) r++;
  if ((1u << r) != x) throw std::runtime_error("value must be a power of two");
  return r;
}

// name of the value on wire w
std::string wname(const Circuit& c, uint32_t w) {
  return (w < c.n ? "x" : "g") + std::to_string(w);
}

// Emits the common kernel head: lane/wave ids, the wave-uniform base U, lane masks.
void emit_head(std::ostringstream& os, const Geometry& g) {
  os << "  const unsigned int lane = threadIdx.x & 63u;\n"
        "  const unsigned int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);\n"
        "  const unsigned long long U = w0 + ((unsigned long long)blockIdx.x * "
     << (g.block / 64) << "ull + wave) * " << (64ull * g.wpl) << "ull;\n";
  for (int i = 0; i < 6; i++)
    os << "  const unsigned int LM" << i << " = 0u - ((lane >> " << i << ") & 1u);\n";
}

// Emits the definitions of the inputs for one word. vbit_of_input[k] is the position
// of input k inside the assignment index.
void emit_inputs(std::ostringstream& os, const Circuit& c, const std::vector<uint32_t>& vbit_of_input,
                 const std::vector<uint8_t>& used, const Geometry& g) {
  uint32_t jb = log2u(g.wpl);
  for (uint32_t k = 0; k < c.n; k++) {
    if (!used[k]) continue;
    uint32_t vb = vbit_of_input[k];
    os << "    const unsigned int x" << k << " = ";
    if (vb < 5) {
      os << "0x" << std::hex << kWordMasks[vb] << std::dec << "u;\n";
      continue;
    }
    uint32_t wb = vb - 5;
    if (wb >= g.lane_shift && wb < g.lane_shift + 6)
      os << "LM" << (wb - g.lane_shift) << ";\n";
    else if (wb >= g.j_shift && wb < g.j_shift + jb)
      os << "0u - ((j >> " << (wb - g.j_shift) << ") & 1u);\n";
    else
      os << "0u - (unsigned int)((U >> " << wb << ") & 1ull);\n";
  }
}

// ---------------------------------------------------------------------------
// 3-input LUT mapping.
//
// Every emitted operation is one v_bitop3_b32: an arbitrary boolean function of up to
// three 32-bit operands given by an 8-bit truth table (bit (s0<<2 | s1<<1 | s2)). The
// mapper walks the gates in evaluation order and folds a gate into its consumer when the
// gate has a single use and the combined function still has at most three distinct
// leaves; negations are absorbed into the truth tables. The LUTs are emitted as inline
// assembly so that each one is exactly one VALU instruction and the compiler does not
// re-combine them (gfx950's compiler-formed bitop3 combinations have been observed to
// compute wrong functions for some gate patterns).
// ---------------------------------------------------------------------------
struct Expr {
  uint8_t nl = 0;
  uint32_t leaf[3] = {0, 0, 0};
  uint8_t tt = 0;  // bit (v0<<2 | v1<<1 | v2) where vi is the value of leaf i
};

inline Expr leaf_expr(uint32_t w) {
  Expr e;
  e.nl = 1;
  e.leaf[0] = e.leaf[1] = e.leaf[2] = w;
  e.tt = 0xF0;  // value = bit of s0
  return e;
}

// value of expression e for leaf assignment given as (wire -> bit) lookup
inline uint32_t eval_expr(const Expr& e, const uint32_t* ul, uint32_t un, uint32_t idx) {
  uint32_t bits[3] = {0, 0, 0};
  for (uint32_t i = 0; i < e.nl; i++) {
    for (uint32_t k = 0; k < un; k++)
      if (ul[k] == e.leaf[i]) {
        bits[i] = (idx >> (2 - k)) & 1u;
        break;
      }
  }
  // leaves beyond nl repeat leaf 0 (their table bits are don't-care but consistent)
  for (uint32_t i = e.nl; i < 3; i++) bits[i] = bits[0];
  return (e.tt >> ((bits[0] << 2) | (bits[1] << 1) | bits[2])) & 1u;
}

inline uint32_t gate_fn(uint8_t f, uint32_t a, uint32_t b) {
  switch (f) {
    case FUNC_AND: return a & b;
    case FUNC_NOR: return (~(a | b)) & 1u;
    case FUNC_NIMPL: return a & (~b & 1u);
    default: return a ^ b;
  }
}

class LutMapper {
 public:
  // uni_input[k] = 1 when input k is the same in every lane of a wave (constant word
  // masks, loop and block bits); gates over uniform values only are uniform too and are
  // computed once per wave with scalar instructions.
  LutMapper(const Circuit& c, const std::vector<uint8_t>& uni_input)
      : c_(c), state_(c.wires(), 0), pend_(c.wires()) {
    prog_.uniform.assign(c.wires(), 0);
    for (uint32_t k = 0; k < c.n; k++) prog_.uniform[k] = uni_input[k];
  }

  // maps the gates in `order`; outputs in `out_wires` are materialised
  void run(const std::vector<uint32_t>& order, const std::vector<uint32_t>& out_wires) {
    std::vector<uint32_t> fan(c_.wires(), 0);
    for (uint32_t w : order) {
      const Gate& g = c_.gates[w - c_.n];
      fan[g.i0]++;
      if (g.i1 != g.i0) fan[g.i1]++;
    }
    for (uint32_t w : out_wires) fan[w] += 2;
    std::vector<uint8_t>& uni = prog_.uniform;
    for (uint32_t w : order) {
      const Gate& g = c_.gates[w - c_.n];
      if (uni[g.i0] && uni[g.i1]) {
        // wave-uniform gate: one scalar operation, always materialised
        uni[w] = 1;
        LutOp op;
        op.lut = 0;
        op.fn = g.func;
        op.dst = w;
        op.src[0] = g.i0;
        op.src[1] = g.i1;
        op.src[2] = g.i0;
        prog_.ops.push_back(op);
        state_[w] = 2;
        prog_.scalar_ops++;
        continue;
      }
      Expr ea = get(g.i0), eb = get(g.i1);
      uint32_t ul[6], un;
      while (true) {
        un = 0;
        for (const Expr* e : {&ea, &eb})
          for (uint32_t i = 0; i < e->nl; i++) {
            bool dup = false;
            for (uint32_t k = 0; k < un; k++) dup |= ul[k] == e->leaf[i];
            if (!dup) ul[un++] = e->leaf[i];
          }
        if (un <= 3) break;
        // materialise the pending operand with more leaves
        uint32_t x = (ea.nl >= eb.nl && state_[g.i0] == 1) ? g.i0 : (state_[g.i1] == 1 ? g.i1 : g.i0);
        materialize(x);
        ea = get(g.i0);
        eb = get(g.i1);
      }
      Expr r;
      r.nl = (uint8_t)un;
      for (uint32_t k = 0; k < 3; k++) r.leaf[k] = ul[k < un ? k : 0];
      uint8_t tt = 0;
      for (uint32_t idx = 0; idx < 8; idx++) {
        uint32_t a = eval_expr(ea, ul, un, idx), b = eval_expr(eb, ul, un, idx);
        if (gate_fn(g.func, a, b)) tt |= (uint8_t)(1u << idx);
      }
      // leaves beyond un repeat leaf 0: make the table consistent for those positions
      r.tt = normalize(tt, un);
      pend_[w] = r;
      state_[w] = 1;
      if (fan[w] != 1) materialize(w);
    }
  }

  LutProgram take() { return std::move(prog_); }

 private:
  // table over (s0,s1,s2) where s_k for k >= un equals s0
  static uint8_t normalize(uint8_t tt, uint32_t un) {
    uint8_t r = 0;
    for (uint32_t idx = 0; idx < 8; idx++) {
      uint32_t b0 = (idx >> 2) & 1, b1 = (idx >> 1) & 1, b2 = idx & 1;
      if (un < 2) b1 = b0;
      if (un < 3) b2 = b0;
      uint32_t src = (b0 << 2) | (b1 << 1) | b2;
      if ((tt >> src) & 1) r |= (uint8_t)(1u << idx);
    }
    return r;
  }

  Expr get(uint32_t w) {
    if (w < c_.n || state_[w] == 2) return leaf_expr(w);
    return pend_[w];
  }

  void materialize(uint32_t w) {
    if (w < c_.n || state_[w] == 2) return;
    const Expr& e = pend_[w];
    LutOp op;
    op.lut = 1;
    op.fn = e.tt;
    op.dst = w;
    for (int i = 0; i < 3; i++) op.src[i] = e.leaf[i];
    prog_.ops.push_back(op);
    state_[w] = 2;
    prog_.valu_ops++;
  }

  const Circuit& c_;
  std::vector<uint8_t> state_;  // 0 not reached, 1 pending expression, 2 materialised
  std::vector<Expr> pend_;
  LutProgram prog_;
};

}  // namespace

LutProgram map_luts(const Circuit& c, const std::vector<uint32_t>& order, const std::vector<uint32_t>& out_wires,
                    const std::vector<uint8_t>& uni_input) {
  LutMapper m(c, uni_input);
  m.run(order, out_wires);
  return m.take();
}

namespace {

// Emits the mapped program as HIP source: scalar gates as C++ expressions, LUTs as one
// inline-assembly v_bitop3_b32 each (at most one wave-uniform operand, read from an SGPR).
void emit_program(std::ostringstream& os, const Circuit& c, const LutProgram& p) {
  for (const LutOp& op : p.ops) {
    if (!op.lut) {
      std::string a = wname(c, op.src[0]), b = wname(c, op.src[1]);
      os << "    const unsigned int g" << op.dst << " = ";
      switch (op.fn) {
        case FUNC_AND: os << a << " & " << b; break;
        case FUNC_NOR: os << "~(" << a << " | " << b << ")"; break;
        case FUNC_NIMPL: os << a << " & ~" << b; break;
        default: os << a << " ^ " << b; break;
      }
      os << ";\n";
      continue;
    }
    bool sgpr_used = false;
    std::string cons[3];
    for (int i = 0; i < 3; i++) {
      bool first_use = true;
      for (int k = 0; k < i; k++) first_use &= op.src[k] != op.src[i];
      if (p.uniform[op.src[i]] && !sgpr_used && first_use) {
        cons[i] = "s";
        sgpr_used = true;
        for (int k = i + 1; k < 3; k++)
          if (op.src[k] == op.src[i]) cons[k] = "s";
      } else if (cons[i].empty()) {
        cons[i] = "v";
      }
    }
    os << "    unsigned int g" << op.dst << ";\n    asm(\"v_bitop3_b32 %0, %1, %2, %3 bitop3:0x" << std::hex
       << (unsigned)op.fn << std::dec << "\" : \"=v\"(g" << op.dst << ") : \"" << cons[0] << "\"("
       << wname(c, op.src[0]) << "), \"" << cons[1] << "\"(" << wname(c, op.src[1]) << "), \"" << cons[2]
       << "\"(" << wname(c, op.src[2]) << "));\n";
  }
}

// Emits the gates of the given order as mapped LUTs; returns the number of VALU ops per
// word and stores the number of scalar (per-wave) ops in *scalar_ops.
uint32_t emit_gates(std::ostringstream& os, const Circuit& c, const std::vector<uint32_t>& order,
                    const std::vector<uint32_t>& outs, const std::vector<uint8_t>& uni_input,
                    uint32_t* scalar_ops) {
  std::vector<uint32_t> out_wires;
  for (uint32_t k : outs) out_wires.push_back(c.outputs[k].wire);
  LutProgram p = map_luts(c, order, out_wires, uni_input);
  emit_program(os, c, p);
  if (scalar_ops) *scalar_ops = p.scalar_ops;
  return p.valu_ops;
}

// wave-uniform inputs for an index mapping: everything but the six lane bits
std::vector<uint8_t> uniform_inputs(const Circuit& c, const std::vector<uint32_t>& vbit_of_input,
                                    const Geometry& g) {
  std::vector<uint8_t> u(c.n, 1);
  for (uint32_t k = 0; k < c.n; k++) {
    uint32_t vb = vbit_of_input[k];
    if (vb >= 5 && vb - 5 >= g.lane_shift && vb - 5 < g.lane_shift + 6) u[k] = 0;
  }
  return u;
}

// Output k as an expression of materialised values (negation as one more VALU op).
std::string out_expr(const Circuit& c, uint32_t k) {
  const Output& o = c.outputs.at(k);
  std::string s = wname(c, o.wire);
  return o.neg ? "(~" + s + ")" : s;
}

// inputs referenced by the ordered gates or directly by the selected outputs
std::vector<uint8_t> used_inputs(const Circuit& c, const std::vector<uint32_t>& order,
                                 const std::vector<uint32_t>& outs) {
  std::vector<uint8_t> u(c.n, 0);
  for (uint32_t w : order) {
    const Gate& gt = c.gates[w - c.n];
    if (gt.i0 < c.n) u[gt.i0] = 1;
    if (gt.i1 < c.n) u[gt.i1] = 1;
  }
  for (uint32_t k : outs)
    if (c.outputs[k].wire < c.n) u[c.outputs[k].wire] = 1;
  return u;
}

void prelude(std::ostringstream& os) {
  os << "// generated: bit-sliced circuit kernel\n";
}

// valid bits of a word when the whole space has fewer than 32 assignments
uint32_t word_valid_bits(uint32_t n_bits) {
  return n_bits >= 5 ? 0xffffffffu : (uint32_t)((1ull << (1u << n_bits)) - 1);
}

}  // namespace

std::string gen_sat_kernel(const Circuit& c, const Geometry& g, GenInfo* info) {
  if (c.outputs.size() != 1) throw std::runtime_error("SAT needs exactly one output");
  std::vector<uint32_t> outs = {0};
  auto cone = output_cone(c, outs);
  auto order = schedule_gates(c, cone, outs);
  if (info) {
    info->cone_gates = (uint32_t)order.size();
    info->peak_live = peak_live(c, order, outs);
  }
  std::vector<uint32_t> vb(c.n);
  for (uint32_t k = 0; k < c.n; k++) vb[k] = k;
  std::ostringstream os;
  prelude(os);
  os << "extern \"C\" __global__ void __launch_bounds__(" << g.block
     << ") ns_sat(unsigned long long w0, unsigned long long wend, unsigned long long* res) {\n";
  emit_head(os, g);
  os << "  unsigned int cnt = 0u;\n  unsigned long long best = ~0ull;\n";
  os << "#pragma unroll 1\n  for (unsigned int j = 0; j < " << g.wpl << "u; j++) {\n";
  os << "    const unsigned long long w = U + ((unsigned long long)lane << " << g.lane_shift
     << ") + ((unsigned long long)j << " << g.j_shift << ");\n";
  emit_inputs(os, c, vb, used_inputs(c, order, outs), g);
  uint32_t sops = 0;
  uint32_t ops = emit_gates(os, c, order, outs, uniform_inputs(c, vb, g), &sops);
  if (info) {
    info->valu_ops = ops;
    info->scalar_ops = sops;
  }
  os << "    unsigned int o = " << out_expr(c, 0) << " & 0x" << std::hex << word_valid_bits(c.n)
     << std::dec << "u;\n";
  os << "    if (w >= wend) o = 0u;\n";
  os << "    cnt += __popc(o);\n"
        "    if (o != 0u && best == ~0ull) best = (w << 5) | (unsigned long long)__builtin_ctz(o);\n"
        "  }\n";
  os << "  if (__ballot(cnt != 0u) != 0ull) {\n"
        "    unsigned long long c64 = cnt;\n"
        "    for (int off = 32; off > 0; off >>= 1) {\n"
        "      unsigned long long ob = __shfl_xor(best, off);\n"
        "      best = ob < best ? ob : best;\n"
        "      c64 += __shfl_xor(c64, off);\n"
        "    }\n"
        "    if (lane == 0u) {\n"
        "      atomicAdd(&res[512u + ((blockIdx.x & 63u) << 4)], c64);\n"
        "      if (best < __hip_atomic_load(&res[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))\n"
        "        atomicMin(&res[1], best);\n"
        "    }\n"
        "  }\n}\n";
  return os.str();
}

std::string gen_qsat_kernel(const Circuit& c, const std::vector<uint8_t>& q, const Geometry& g,
                            GenInfo* info) {
  if (c.outputs.size() != 1) throw std::runtime_error("QSAT needs exactly one output");
  if (q.size() != c.n) throw std::runtime_error("one quantifier per input is required");
  const uint32_t n = c.n;
  uint32_t fb = 0;
  while (fb < n && q[fb] == q[0]) fb++;
  const uint32_t L = n - fb;
  const uint32_t jb = log2u(g.wpl);
  const bool q0all = q[0] != 0;
  // op of level l (level 0 = index bit 0 = input n-1): 1 = AND (all), 0 = OR (exists)
  auto op = [&](uint32_t l) { return q[n - 1 - l] ? "&" : "|"; };
  if (g.lane_shift != jb || g.j_shift != 0)
    throw std::runtime_error("QSAT kernels need the j bits below the lane bits");
  std::vector<uint32_t> outs = {0};
  auto cone = output_cone(c, outs);
  auto order = schedule_gates(c, cone, outs);
  if (info) {
    info->cone_gates = (uint32_t)order.size();
    info->peak_live = peak_live(c, order, outs);
  }
  std::vector<uint32_t> vb(n);
  for (uint32_t k = 0; k < n; k++) vb[k] = n - 1 - k;
  const uint32_t valid_bits = word_valid_bits(n);
  const bool case_a = L >= 5 + jb;
  std::ostringstream os;
  prelude(os);
  if (case_a)
    os << "extern \"C\" __global__ void __launch_bounds__(" << g.block
       << ") ns_qsat(unsigned long long w0, unsigned long long wend, unsigned long long* bits) {\n";
  else
    os << "extern \"C\" __global__ void __launch_bounds__(" << g.block
       << ") ns_qsat(unsigned long long w0, unsigned long long wend, unsigned long long* stats, "
          "unsigned int* bitmap) {\n";
  emit_head(os, g);
  os << "  const unsigned long long wl = U + ((unsigned long long)lane << " << jb << ");\n";
  if (!case_a && L <= 5)
    os << "  unsigned int cnt = 0u;\n  unsigned long long best = ~0ull;\n";
  else
    os << "  unsigned int R = 0u;\n";
  os << "#pragma unroll 1\n  for (unsigned int j = 0; j < " << g.wpl << "u; j++) {\n";
  os << "    const unsigned long long w = wl + j;\n";
  emit_inputs(os, c, vb, used_inputs(c, order, outs), g);
  uint32_t sops = 0;
  uint32_t ops = emit_gates(os, c, order, outs, uniform_inputs(c, vb, g), &sops);
  if (info) {
    info->valu_ops = ops;
    info->scalar_ops = sops;
  }
  os << "    unsigned int o = " << out_expr(c, 0) << ";\n";
  uint32_t inword = L < 5 ? L : 5;
  for (uint32_t l = 0; l < inword; l++)
    os << "    o = o " << op(l) << " (o >> " << (1u << l) << ");\n";
  if (!case_a && L <= 5) {
    // positions p * 2^L of the word hold outermost-block values u = (32w + p 2^L) >> L
    uint32_t pos = 0;
    for (uint32_t p = 0; p < (32u >> L); p++) pos |= 1u << (p << L);
    pos &= valid_bits;
    os << "    unsigned int wit = (" << (q0all ? "~o" : "o") << ") & 0x" << std::hex << pos << std::dec
       << "u;\n";
    os << "    if (w >= wend) wit = 0u;\n";
    os << "    cnt += __popc(wit);\n";
    os << "    if (wit != 0u && best == ~0ull) best = (w << " << (5 - L)
       << ") + (unsigned long long)(__builtin_ctz(wit) >> " << L << ");\n";
    os << "    if (bitmap != 0 && w < wend) {\n";
    if (L == 0) {
      os << "      bitmap[w] = o & 0x" << std::hex << valid_bits << std::dec << "u;\n";
    } else {
      os << "      unsigned int r = 0u;\n";
      for (uint32_t p = 0; p < (32u >> L); p++)
        if ((pos >> (p << L)) & 1u)
          os << "      r |= ((o >> " << (p << L) << ") & 1u) << " << p << ";\n";
      os << "      const unsigned long long ub = w << " << (5 - L) << ";\n";
      os << "      atomicOr(&bitmap[ub >> 5], r << (unsigned int)(ub & 31ull));\n";
    }
    os << "    }\n";
    os << "  }\n";
  } else {
    os << "    R |= (o & 1u) << j;\n  }\n";
    // j levels (levels 5 .. 5+jb-1)
    uint32_t jl = (L - 5) < jb ? (L - 5) : jb;
    for (uint32_t l = 0; l < jl; l++)
      os << "  R = R " << op(5 + l) << " (R >> " << (1u << l) << ");\n";
    os << "  const unsigned int nvalid = wend > wl ? (wend - wl >= " << g.wpl << "ull ? " << g.wpl
       << "u : (unsigned int)(wend - wl)) : 0u;\n";
    if (case_a) {
      os << "  const unsigned long long ball = __ballot((R & 1u) != 0u && nvalid != 0u);\n";
      os << "  if (lane == 0u) bits[U >> " << (jb + 6) << "] = ball;\n}\n";
    } else {
      // positions p * 2^(L-5) of R hold u = (wl >> (L-5)) + p
      uint32_t s = L - 5;
      uint32_t pos = 0;
      for (uint32_t p = 0; p < (g.wpl >> s); p++) pos |= 1u << (p << s);
      os << "  const unsigned int jmask = nvalid >= 32u ? 0xffffffffu : ((1u << nvalid) - 1u);\n";
      os << "  unsigned int wit = (" << (q0all ? "~R" : "R") << ") & 0x" << std::hex << pos << std::dec
         << "u & jmask;\n";
      os << "  unsigned int cnt = __popc(wit);\n";
      os << "  unsigned long long best = wit != 0u ? (wl >> " << s
         << ") + (unsigned long long)(__builtin_ctz(wit) >> " << s << ") : ~0ull;\n";
      os << "  if (bitmap != 0 && nvalid != 0u) {\n"
            "    unsigned int r = 0u;\n";
      for (uint32_t p = 0; p < (g.wpl >> s); p++)
        os << "    r |= ((R >> " << (p << s) << ") & 1u) << " << p << ";\n";
      os << "    r &= (nvalid >= " << g.wpl << "u) ? 0xffffffffu : ((1u << ((nvalid + "
         << ((1u << s) - 1) << "u) >> " << s << ")) - 1u);\n";
      os << "    const unsigned long long ub = wl >> " << s << ";\n";
      os << "    atomicOr(&bitmap[ub >> 5], r << (unsigned int)(ub & 31ull));\n";
      os << "  }\n";
    }
  }
  if (!case_a) {
    os << "  if (__ballot(cnt != 0u) != 0ull) {\n"
          "    unsigned long long c64 = cnt;\n"
          "    for (int off = 32; off > 0; off >>= 1) {\n"
          "      unsigned long long ob = __shfl_xor(best, off);\n"
          "      best = ob < best ? ob : best;\n"
          "      c64 += __shfl_xor(c64, off);\n"
          "    }\n"
          "    if (lane == 0u) {\n"
          "      atomicAdd(&stats[512u + ((blockIdx.x & 63u) << 4)], c64);\n"
          "      if (best < __hip_atomic_load(&stats[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))\n"
          "        atomicMin(&stats[1], best);\n"
          "    }\n"
          "  }\n}\n";
  }
  return os.str();
}

std::string gen_table_kernel(const Circuit& c, const Geometry& g, GenInfo* info) {
  const uint32_t m = (uint32_t)c.outputs.size();
  if (m == 0 || m > 32) throw std::runtime_error("table build supports 1..32 outputs");
  std::vector<uint32_t> outs;
  for (uint32_t k = 0; k < m; k++) outs.push_back(k);
  auto cone = output_cone(c, outs);
  auto order = schedule_gates(c, cone, outs);
  if (info) {
    info->cone_gates = (uint32_t)order.size();
    info->peak_live = peak_live(c, order, outs);
  }
  std::vector<uint32_t> vb(c.n);
  for (uint32_t k = 0; k < c.n; k++) vb[k] = k;
  std::ostringstream os;
  prelude(os);
  os << "extern \"C\" __global__ void __launch_bounds__(" << g.block
     << ") ns_table(unsigned long long w0, unsigned long long wend, unsigned int* table, "
        "unsigned int* anylast) {\n";
  emit_head(os, g);
  os << "  unsigned int last_any = 0u;\n";
  os << "#pragma unroll 1\n  for (unsigned int j = 0; j < " << g.wpl << "u; j++) {\n";
  os << "    const unsigned long long w = U + ((unsigned long long)lane << " << g.lane_shift
     << ") + ((unsigned long long)j << " << g.j_shift << ");\n";
  emit_inputs(os, c, vb, used_inputs(c, order, outs), g);
  uint32_t sops = 0;
  uint32_t ops = emit_gates(os, c, order, outs, uniform_inputs(c, vb, g), &sops);
  if (info) {
    info->valu_ops = ops;
    info->scalar_ops = sops;
  }
  os << "    unsigned int A[32];\n";
  for (uint32_t k = 0; k < 32; k++) {
    if (k < m) os << "    A[" << k << "] = " << out_expr(c, k) << ";\n";
    else os << "    A[" << k << "] = 0u;\n";
  }
  os << "    if (w < wend) last_any |= A[" << (m - 1) << "] & 0x" << std::hex << word_valid_bits(c.n)
     << std::dec << "u;\n";
  // 32x32 bit transpose: afterwards bit i of A[b] is bit b of output i
  os << "    {\n"
        "      unsigned int msk = 0x0000FFFFu;\n"
        "#pragma unroll\n"
        "      for (int s = 16; s != 0; s >>= 1, msk ^= (msk << s)) {\n"
        "#pragma unroll\n"
        "        for (int k = 0; k < 32; k = ((k | s) + 1) & ~s) {\n"
        "          const unsigned int t = ((A[k] >> s) ^ A[k | s]) & msk;\n"
        "          A[k] ^= t << s;\n"
        "          A[k | s] ^= t;\n"
        "        }\n"
        "      }\n"
        "    }\n";
  if (c.n >= 5) {
    os << "    if (w < wend) {\n"
          "      uint4* dst = (uint4*)(table + ((w - w0) << 5));\n"
          "#pragma unroll\n"
          "      for (int q = 0; q < 8; q++) dst[q] = make_uint4(A[4*q], A[4*q+1], A[4*q+2], A[4*q+3]);\n"
          "    }\n";
  } else {
    os << "    if (w < wend) {\n";
    for (uint32_t b = 0; b < (1u << c.n); b++) os << "      table[" << b << "] = A[" << b << "];\n";
    os << "    }\n";
  }
  os << "  }\n";
  // one flag write per wave, none once it is set (atomics on one word serialize)
  os << "  const unsigned long long any_b = __ballot(last_any != 0u);\n"
        "  if (any_b != 0ull && (threadIdx.x & 63u) == (unsigned int)(__ffsll((long long)any_b) - 1) &&\n"
        "      __hip_atomic_load(anylast, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)\n"
        "    atomicOr(anylast, 1u);\n}\n";
  return os.str();
}

std::string gen_walk_kernel(const Circuit& c, uint32_t k, GenInfo* info) {
  const uint32_t n = c.n;
  if (c.outputs.size() != (size_t)n + 1) throw std::runtime_error("SATOSTOP needs n inputs and n+1 outputs");
  if (n < 1 || n > 62 || k > n) throw std::runtime_error("walks take 1..62 state bits");
  std::vector<uint32_t> outs;
  for (uint32_t i = 0; i <= n; i++) outs.push_back(i);
  auto cone = output_cone(c, outs);
  auto order = schedule_gates(c, cone, outs);
  if (info) {
    info->cone_gates = (uint32_t)order.size();
    info->peak_live = peak_live(c, order, outs);
  }
  const std::vector<uint8_t> lane_inputs(n, 0);  // every input varies per walk
  std::ostringstream os;
  prelude(os);
  auto each = [&](const std::string& line) {
    // the line with @ replaced by 0..n-1, once per state bit
    for (uint32_t i = 0; i < n; i++) {
      std::string l = line;
      for (size_t p; (p = l.find('@')) != std::string::npos;) l.replace(p, 1, std::to_string(i));
      os << l;
    }
  };
  os << "extern \"C\" __global__ void __launch_bounds__(256) ns_walk(unsigned long long count, "
        "unsigned long long limit, unsigned long long budget, ulonglong2* __restrict__ out, unsigned int* long_flag, "
        "unsigned long long* evals, unsigned long long* queue) {\n";
  // per slot: start time (bits 0..31) and waypoint index (32..63) of its walk
  os << "  __shared__ unsigned long long meta[32 * 256];\n"
        "  const unsigned int tid = threadIdx.x, lane = tid & 63u;\n"
        "  const unsigned long long lim = limit < 0x7fffffffull ? limit : 0x7fffffffull;\n";
  each("  unsigned int s@ = 0u, v@ = 0u, f@ = 0u;\n");
  // fin: slots whose walk ended since the last check (their last state in f@), with the kind
  // of ending in fstop / fwp / floop (none of them: the limit) and, bit-sliced in a0..a3, the
  // evaluations since it ended
  os << "  unsigned int active = 0u, fin = 0u, fstop = 0u, fwp = 0u, floop = 0u, t = 0u;\n"
        "  unsigned int a0 = 0u, a1 = 0u, a2 = 0u, a3 = 0u;\n"
        "  unsigned long long used = 0ull, ev = 0ull;\n"
        "  bool drained = false;\n";
  // The slots in `need` take the next waypoints of the global queue at time t (one atomic
  // per wave: the lanes' shares are an exclusive prefix sum of their counts); a slot's bits
  // of the state and of the saved state become its waypoint's state.
  std::ostringstream refill;
  refill << "    if (!drained) {\n"
            "      const unsigned int r = __popc(need);\n"
            "      unsigned int incl = r;\n"
            "      for (unsigned int off = 1; off < 64u; off <<= 1) {\n"
            "        const unsigned int y = __shfl_up(incl, off);\n"
            "        if (lane >= off) incl += y;\n"
            "      }\n"
            "      const unsigned int total = __shfl(incl, 63);\n"
            "      unsigned long long base = 0ull;\n"
            "      if (total != 0u) {\n"
            "        if (lane == 0u) base = atomicAdd(queue, (unsigned long long)total);\n"
            "        base = __shfl(base, 0);\n"
            "      }\n"
            "      if (base + total >= count) drained = true;\n"
            "      unsigned long long nx = base + (incl - r);\n"
            "      for (unsigned int rem = need; rem; rem &= rem - 1u) {\n"
            "        const unsigned int b = __builtin_ctz(rem), mb = 1u << b;\n"
            "        const unsigned long long idx = nx++;\n"
            "        if (idx >= count) break;\n"
            "        const unsigned long long st = idx << "
         << k << ";\n";
  for (uint32_t i = 0; i < n; i++)
    refill << "        s" << i << " = (s" << i << " & ~mb) | ((unsigned int)((st >> " << i << ") & 1ull) * mb); v" << i
           << " = (v" << i << " & ~mb) | (s" << i << " & mb);\n";
  refill << "        meta[b * 256u + tid] = (unsigned long long)t | (idx << 32);\n"
            "        active |= mb;\n"
            "      }\n"
            "    }\n";
  os << "  {\n    const unsigned int need = 0xffffffffu;\n" << refill.str() << "  }\n";
  // the loop runs while any lane of the wave has a walk, so every lane takes part in the
  // queue's prefix sums
  os << "  while (__any((active | fin) != 0u)) {\n"
        "    used += 2ull;\n"
        "    if (used > budget) {\n"
        "      atomicOr(long_flag, 1u);\n"
        "      break;\n"
        "    }\n"
        "    if ((used & 255ull) == 0ull && __hip_atomic_load(long_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)\n"
        "      break;\n";
  // Two steps per iteration: A from s@ to p@, B from p@ back to s@. The saved-state test runs
  // in step B only: a walk that returns to its saved state does so again after its cycle
  // length, and saves happen at even times, so it is caught at most one cycle later.
  uint32_t sops = 0, ops = 0;
  each("    unsigned int p@;\n");
  for (int half = 0; half < 2; half++) {
    const char* src = half == 0 ? "s" : "p";
    const char* dst = half == 0 ? "p" : "s";
    os << "    {\n";
    for (uint32_t i = 0; i < n; i++) os << "      const unsigned int x" << i << " = " << src << i << ";\n";
    ops = emit_gates(os, c, order, outs, lane_inputs, &sops);
    os << "      const unsigned int on = " << out_expr(c, n) << ";\n";
    for (uint32_t i = 0; i < n; i++) os << "      " << dst << i << " = " << out_expr(c, i) << ";\n";
    os << "      const unsigned int stop = on & active;\n";
    os << "      const unsigned int nz = 0u";
    for (uint32_t i = 0; i < k; i++) os << " | " << dst << i;
    os << ";\n      const unsigned int wp = ~nz & active & ~stop;\n";
    if (half == 1) {
      os << "      const unsigned int eq = 0xffffffffu";
      for (uint32_t i = 0; i < n; i++) os << " & ~(" << dst << i << " ^ v" << i << ")";
      os << ";\n      const unsigned int loop = eq & active & ~stop & ~wp;\n"
            "      const unsigned int done = stop | wp | loop;\n"
            "      floop |= loop;\n";
    } else {
      os << "      const unsigned int done = stop | wp;\n";
    }
    // an ending slot keeps the state its last evaluation gave in f@
    os << "      if (__any(done != 0u)) {\n";
    for (uint32_t i = 0; i < n; i++)
      os << "        f" << i << " = (f" << i << " & ~done) | (" << dst << i << " & done);\n";
    os << "      }\n"
          "      unsigned int cy = fin;\n"
          "      a0 ^= cy;\n"
          "      cy &= ~a0;\n"
          "      a1 ^= cy;\n"
          "      cy &= ~a1;\n"
          "      a2 ^= cy;\n"
          "      cy &= ~a2;\n"
          "      a3 ^= cy;\n"
          "      fin |= done;\n"
          "      fstop |= stop;\n"
          "      fwp |= wp;\n"
          "      active &= ~done;\n"
          "      t++;\n"
          "    }\n";
  }
  if (info) {
    info->valu_ops = ops;
    info->scalar_ops = sops;
  }
  // every 16 evaluations: the evaluations of the window (16 per slot running at its start,
  // less the steps since each ended slot finished), Brent's saves (a slot saves its state at
  // the first check after each power of two of its evaluations), the limit, the records of
  // the ended walks and new walks in their slots
  os << "    if ((t & 15u) == 0u) {\n"
        "      ev += 16ull * (unsigned long long)__popc(active | fin) -\n"
        "            (unsigned long long)(__popc(a0) + 2u * __popc(a1) + 4u * __popc(a2) + 8u * __popc(a3));\n"
        "      unsigned int save = 0u, lng = 0u;\n"
        "      for (unsigned int rem = active; rem; rem &= rem - 1u) {\n"
        "        const unsigned int b = __builtin_ctz(rem);\n"
        "        const unsigned int start = (unsigned int)meta[b * 256u + tid];\n"
        "        const unsigned int steps = t - start, before = steps > 16u ? steps - 16u : 0u;\n"
        "        if ((unsigned long long)steps >= lim) lng |= 1u << b;\n"
        "        else if (before == 0u || __builtin_clz(steps) < __builtin_clz(before)) save |= 1u << b;\n"
        "      }\n"
        "      if (save) {\n";
  each("        v@ = (v@ & ~save) | (s@ & save);\n");
  os << "      }\n"
        "      if (lng) {\n"
        "        atomicOr(long_flag, 1u);\n"
        "        fin |= lng;\n"
        "        active &= ~lng;\n"
        "      }\n"
        "      for (unsigned int rem = fin; rem; rem &= rem - 1u) {\n"
        "        const unsigned int b = __builtin_ctz(rem), mb = 1u << b;\n"
        "        const unsigned long long m = meta[b * 256u + tid];\n"
        "        const unsigned int age = ((a0 >> b) & 1u) | (((a1 >> b) & 1u) << 1) | (((a2 >> b) & 1u) << 2) |\n"
        "                                 (((a3 >> b) & 1u) << 3);\n"
        "        const unsigned int steps = t - age - (unsigned int)m;\n"
        "        const unsigned int kind = (fstop & mb) ? 1u : (fwp & mb) ? 0u : (floop & mb) ? 2u : 3u;\n"
        "        unsigned long long st = 0ull;\n";
  each("        st |= (unsigned long long)((f@ >> b) & 1u) << @;\n");
  os << "        const unsigned long long target = kind == 1u ? st : kind == 0u ? (st >> " << k
     << ") : 0ull;\n"
        "        out[m >> 32] = make_ulonglong2(target | ((unsigned long long)kind << 62), steps);\n"
        "      }\n"
        "      {\n"
        "        const unsigned int need = fin;\n"
     << refill.str()
     << "      }\n"
        "      fin = 0u;\n"
        "      fstop = 0u;\n"
        "      fwp = 0u;\n"
        "      floop = 0u;\n"
        "      a0 = a1 = a2 = a3 = 0u;\n"
        "    }\n"
        "  }\n";
  // evaluations done by the wave, into one of 64 count slots 16 words apart
  os << "  for (int off = 32; off > 0; off >>= 1) ev += __shfl_xor(ev, off);\n"
        "  if ((tid & 63u) == 0u && ev != 0ull) atomicAdd(evals + ((blockIdx.x & 63u) << 4), ev);\n"
        "}\n";
  return os.str();
}

}  // namespace ns
