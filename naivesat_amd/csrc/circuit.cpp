// Circuit IR: validation, scalar evaluation, text format, cone/schedule analysis and
// the synthetic workload generators.
#include "circuit.h"

#include <algorithm>
#include <cctype>
#include <map>
#include <stdexcept>

namespace ns {

static inline uint64_t apply64(uint8_t f, uint64_t a, uint64_t b) {
  switch (f) {
    case FUNC_AND: return a & b;
    case FUNC_NOR: return ~(a | b);
    case FUNC_NIMPL: return a & ~b;
    default: return a ^ b;
  }
}

void Circuit::validate() const {
  for (size_t j = 0; j < gates.size(); j++) {
    uint32_t w = n + (uint32_t)j;
    if (gates[j].i0 >= w || gates[j].i1 >= w)
      throw std::runtime_error("gate " + std::to_string(j) + " refers to a later wire");
    if (gates[j].func > FUNC_XOR) throw std::runtime_error("bad gate function");
  }
  for (auto& o : outputs)
    if (o.wire >= wires()) throw std::runtime_error("output refers to a missing wire");
}

std::vector<uint64_t> Circuit::eval64(const std::vector<uint64_t>& inputs) const {
  std::vector<uint64_t> v(wires());
  for (uint32_t k = 0; k < n; k++) v[k] = inputs[k];
  for (size_t j = 0; j < gates.size(); j++)
    v[n + j] = apply64(gates[j].func, v[gates[j].i0], v[gates[j].i1]);
  std::vector<uint64_t> out(outputs.size());
  for (size_t i = 0; i < outputs.size(); i++)
    out[i] = v[outputs[i].wire] ^ (outputs[i].neg ? ~0ULL : 0ULL);
  return out;
}

std::vector<bool> Circuit::eval(const std::vector<bool>& inputs) const {
  std::vector<uint64_t> in(n);
  for (uint32_t k = 0; k < n; k++) in[k] = inputs[k] ? 1 : 0;
  auto o = eval64(in);
  std::vector<bool> r(o.size());
  for (size_t i = 0; i < o.size(); i++) r[i] = (o[i] & 1) != 0;
  return r;
}

// ---------------- text format ----------------

namespace {
struct Lexer {
  const std::string& s;
  size_t p = 0;
  explicit Lexer(const std::string& t) : s(t) {}
  void ws() {
    while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
  }
  bool eat(char c) {
    ws();
    if (p < s.size() && s[p] == c) {
      p++;
      return true;
    }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) throw std::runtime_error(std::string("expected '") + c + "' at offset " + std::to_string(p));
  }
  bool peek_digit() {
    ws();
    return p < s.size() && std::isdigit((unsigned char)s[p]);
  }
  uint64_t number() {
    ws();
    if (p >= s.size() || !std::isdigit((unsigned char)s[p]))
      throw std::runtime_error("expected a number at offset " + std::to_string(p));
    uint64_t v = 0;
    while (p < s.size() && std::isdigit((unsigned char)s[p])) {
      v = v * 10 + (uint64_t)(s[p] - '0');
      if (v > 0xffffffffULL) throw std::runtime_error("number too large");
      p++;
    }
    return v;
  }
  std::string ident() {
    ws();
    size_t b = p;
    while (p < s.size() && std::isalpha((unsigned char)s[p])) p++;
    return s.substr(b, p - b);
  }
};
}  // namespace

Circuit parse_circuit(const std::string& text) {
  Lexer lx(text);
  lx.expect('{');
  Circuit c;
  std::map<uint64_t, Output> outs;
  uint32_t wire = 0;
  bool in_gates = false;
  auto markers = [&](uint32_t w) {
    while (lx.p < text.size() && text[lx.p] == ':') {
      lx.p++;
      uint64_t k = lx.number();
      bool neg = false;
      if (lx.p < text.size() && text[lx.p] == 'n') {
        neg = true;
        lx.p++;
      }
      if (outs.count(k)) throw std::runtime_error("duplicate output index " + std::to_string(k));
      outs[k] = Output{w, neg};
    }
  };
  while (true) {
    if (lx.eat('}')) break;
    if (lx.p >= text.size()) throw std::runtime_error("unterminated circuit");
    if (lx.peek_digit()) {
      if (in_gates) throw std::runtime_error("input listed after a gate");
      uint64_t idx = lx.number();
      if (idx != wire) throw std::runtime_error("inputs must be listed in order 0..n-1");
      c.n++;
      markers(wire);
      wire++;
      continue;
    }
    std::string f = lx.ident();
    uint8_t func;
    if (f == "and") func = FUNC_AND;
    else if (f == "nor") func = FUNC_NOR;
    else if (f == "nimpl") func = FUNC_NIMPL;
    else if (f == "xor") func = FUNC_XOR;
    else throw std::runtime_error("unknown gate function '" + f + "'");
    in_gates = true;
    lx.expect('(');
    uint64_t a = lx.number();
    lx.expect(',');
    uint64_t b = lx.number();
    lx.expect(')');
    c.gates.push_back(Gate{(uint32_t)a, (uint32_t)b, func});
    markers(wire);
    wire++;
  }
  lx.ws();
  if (lx.p != text.size()) throw std::runtime_error("trailing text after circuit");
  for (uint64_t k = 0; k < outs.size(); k++) {
    auto it = outs.find(k);
    if (it == outs.end()) throw std::runtime_error("missing output " + std::to_string(k));
    c.outputs.push_back(it->second);
  }
  c.validate();
  return c;
}

std::string format_circuit(const Circuit& c) {
  std::vector<std::vector<std::pair<uint32_t, bool>>> marks(c.wires());
  for (size_t i = 0; i < c.outputs.size(); i++)
    marks[c.outputs[i].wire].push_back({(uint32_t)i, c.outputs[i].neg});
  static const char* names[4] = {"and", "nor", "nimpl", "xor"};
  std::string s = "{";
  for (uint32_t w = 0; w < c.wires(); w++) {
    if (w) s += ' ';
    if (w < c.n) {
      s += std::to_string(w);
    } else {
      const Gate& g = c.gates[w - c.n];
      s += names[g.func];
      s += '(' + std::to_string(g.i0) + ',' + std::to_string(g.i1) + ')';
    }
    for (auto& m : marks[w]) {
      s += ':' + std::to_string(m.first);
      if (m.second) s += 'n';
    }
  }
  s += '}';
  return s;
}

bool parse_quant_circuit(const std::string& text, std::vector<uint8_t>& quants_all, Circuit& c) {
  size_t p = 0;
  std::vector<uint8_t> q;
  while (p < text.size() && text[p] != '{') {
    char ch = (char)std::tolower((unsigned char)text[p]);
    if (ch == 'e') q.push_back(0);
    else if (ch == 'a') q.push_back(1);
    else if (!std::isspace((unsigned char)ch)) return false;
    p++;
  }
  if (q.empty()) return false;
  c = parse_circuit(text.substr(p));
  if (q.size() != c.n) throw std::runtime_error("quantifier count differs from input count");
  quants_all = q;
  return true;
}

// ---------------- analysis ----------------

std::vector<uint8_t> output_cone(const Circuit& c, const std::vector<uint32_t>& outs) {
  std::vector<uint8_t> mark(c.wires(), 0);
  std::vector<uint32_t> st;
  if (outs.empty())
    for (auto& o : c.outputs) st.push_back(o.wire);
  else
    for (auto k : outs) st.push_back(c.outputs.at(k).wire);
  while (!st.empty()) {
    uint32_t w = st.back();
    st.pop_back();
    if (w < c.n || mark[w]) continue;
    mark[w] = 1;
    const Gate& g = c.gates[w - c.n];
    st.push_back(g.i0);
    st.push_back(g.i1);
  }
  return mark;
}

std::vector<uint32_t> schedule_gates(const Circuit& c, const std::vector<uint8_t>& cone,
                                     const std::vector<uint32_t>& outs) {
  const uint32_t n = c.n;
  std::vector<uint32_t> depth(c.wires(), 0);
  for (size_t j = 0; j < c.gates.size(); j++)
    depth[n + j] = 1 + std::max(depth[c.gates[j].i0], depth[c.gates[j].i1]);
  std::vector<uint8_t> seen(c.wires(), 0);
  std::vector<uint32_t> order;
  std::vector<uint32_t> roots;
  if (outs.empty())
    for (auto& o : c.outputs) roots.push_back(o.wire);
  else
    for (auto k : outs) roots.push_back(c.outputs.at(k).wire);
  // iterative post-order DFS; frame = (wire, stage)
  std::vector<std::pair<uint32_t, int>> st;
  for (uint32_t r : roots) {
    st.push_back({r, 0});
    while (!st.empty()) {
      auto& fr = st.back();
      uint32_t w = fr.first;
      if (w < n || !cone[w] || (fr.second == 0 && seen[w])) {
        st.pop_back();
        continue;
      }
      const Gate& g = c.gates[w - n];
      uint32_t first = g.i0, second = g.i1;
      if (depth[second] > depth[first]) std::swap(first, second);
      if (fr.second == 0) {
        seen[w] = 1;
        fr.second = 1;
        st.push_back({first, 0});
      } else if (fr.second == 1) {
        fr.second = 2;
        st.push_back({second, 0});
      } else {
        order.push_back(w);
        st.pop_back();
      }
    }
  }
  return order;
}

uint32_t peak_live(const Circuit& c, const std::vector<uint32_t>& order,
                   const std::vector<uint32_t>& outs) {
  std::vector<int64_t> last(c.wires(), -1);
  for (size_t i = 0; i < order.size(); i++) {
    const Gate& g = c.gates[order[i] - c.n];
    last[g.i0] = (int64_t)i;
    last[g.i1] = (int64_t)i;
  }
  if (outs.empty())
    for (auto& o : c.outputs) last[o.wire] = (int64_t)order.size();
  else
    for (auto k : outs) last[c.outputs[k].wire] = (int64_t)order.size();
  std::vector<uint32_t> dying(order.size() + 1, 0);
  for (uint32_t w : order)
    if (last[w] >= 0) dying[last[w]]++;
  uint32_t live = 0, peak = 0;
  for (size_t i = 0; i < order.size(); i++) {
    live++;
    peak = std::max(peak, live);
    live -= dying[i];
  }
  return peak;
}

// ---------------- generators ----------------

uint32_t LitBuilder::gate(int op, uint32_t a, bool na, uint32_t b, bool nb) {
  bool ea = na ^ (pol[a] != 0), eb = nb ^ (pol[b] != 0);
  uint8_t func;
  uint8_t p = 0;
  uint32_t x = a, y = b;
  if (op == 2) {
    func = FUNC_XOR;
    p = (uint8_t)(ea ^ eb);
  } else {
    if (op == 1) {  // a|b = ~(~a & ~b)
      ea = !ea;
      eb = !eb;
      p = 1;
    }
    if (!ea && !eb) func = FUNC_AND;
    else if (!ea && eb) func = FUNC_NIMPL;
    else if (ea && !eb) {
      func = FUNC_NIMPL;
      std::swap(x, y);
    } else func = FUNC_NOR;
  }
  c.gates.push_back(Gate{x, y, func});
  pol.push_back(p);
  return c.wires() - 1;
}

std::vector<uint32_t> random_block(LitBuilder& lb, Rng& rng, const std::vector<uint32_t>& in,
                                   uint32_t gates) {
  std::vector<uint32_t> local(in);
  std::vector<uint32_t> made;
  made.reserve(gates);
  for (uint32_t j = 0; j < gates; j++) {
    int op = (int)rng.below(3);
    uint32_t a = local[rng.below(local.size())];
    bool na = rng.bit();
    uint32_t b = local[rng.below(local.size())];
    bool nb = rng.bit();
    uint32_t g = lb.gate(op, a, na, b, nb);
    local.push_back(g);
    made.push_back(g);
  }
  return made;
}

// logical value of builder nodes for one scalar input assignment
static std::vector<uint8_t> eval_nodes(const LitBuilder& lb, uint64_t x) {
  const Circuit& c = lb.c;
  std::vector<uint64_t> in(c.n);
  for (uint32_t k = 0; k < c.n; k++) in[k] = (x >> k) & 1;
  std::vector<uint64_t> v(c.wires());
  for (uint32_t k = 0; k < c.n; k++) v[k] = in[k];
  for (size_t j = 0; j < c.gates.size(); j++)
    v[c.n + j] = apply64(c.gates[j].func, v[c.gates[j].i0], v[c.gates[j].i1]) & 1;
  std::vector<uint8_t> r(c.wires());
  for (uint32_t w = 0; w < c.wires(); w++) r[w] = (uint8_t)((v[w] ^ lb.pol[w]) & 1);
  return r;
}

// AND of literals as a balanced tree; returns (node, neg) of the result literal
static std::pair<uint32_t, bool> and_tree(LitBuilder& lb, std::vector<std::pair<uint32_t, bool>> lits) {
  while (lits.size() > 1) {
    std::vector<std::pair<uint32_t, bool>> nl;
    for (size_t i = 0; i + 1 < lits.size(); i += 2)
      nl.push_back({lb.gate(0, lits[i].first, lits[i].second, lits[i + 1].first, lits[i + 1].second), false});
    if (lits.size() & 1) nl.push_back(lits.back());
    lits.swap(nl);
  }
  return lits[0];
}

Circuit gen_sat_circuit(uint64_t seed, uint32_t n, uint32_t gates, uint32_t wbits, uint64_t* key) {
  if (n == 0 || n > 64 || wbits == 0 || wbits > gates)
    throw std::runtime_error("bad SAT generator parameters");
  Rng rng(seed);
  uint64_t mask = n == 64 ? ~0ULL : ((1ULL << n) - 1);
  uint64_t k = rng.next() & mask;
  LitBuilder lb(n);
  std::vector<uint32_t> in(n);
  for (uint32_t i = 0; i < n; i++) in[i] = i;
  auto made = random_block(lb, rng, in, gates);
  auto val = eval_nodes(lb, k);
  std::vector<std::pair<uint32_t, bool>> lits;
  for (uint32_t i = 0; i < wbits; i++) {
    uint32_t w = made[gates - wbits + i];
    lits.push_back({w, val[w] == 0});
  }
  auto r = and_tree(lb, lits);
  lb.output(r.first, r.second);
  if (key) *key = k;
  return lb.c;
}

Circuit gen_factor_circuit(uint64_t N, uint32_t bits) {
  if (bits < 2 || 2 * bits > 64 || (2 * bits < 64 && (N >> (2 * bits)) != 0))
    throw std::runtime_error("bad factoring generator parameters");
  using Lit = std::pair<uint32_t, bool>;
  LitBuilder lb(2 * bits);
  auto op = [&](int o, Lit x, Lit y) -> Lit { return {lb.gate(o, x.first, x.second, y.first, y.second), false}; };
  // partial products a_i & b_j in the column of weight i + j
  std::vector<std::vector<Lit>> col(2 * bits + 1);
  for (uint32_t i = 0; i < bits; i++)
    for (uint32_t j = 0; j < bits; j++) col[i + j].push_back(op(0, {i, false}, {bits + j, false}));
  // carry-save rounds: every column of three or more takes full adders (sum in the column,
  // carry in the next) until each column holds at most two values
  for (bool more = true; more;) {
    more = false;
    std::vector<std::vector<Lit>> nc(2 * bits + 1);
    for (uint32_t w = 0; w < 2 * bits; w++) {
      auto& v = col[w];
      size_t k = 0;
      for (; k + 3 <= v.size(); k += 3) {
        const Lit t = op(2, v[k], v[k + 1]);
        nc[w].push_back(op(2, t, v[k + 2]));
        nc[w + 1].push_back(op(1, op(0, v[k], v[k + 1]), op(0, t, v[k + 2])));
      }
      for (; k < v.size(); k++) nc[w].push_back(v[k]);
    }
    col.swap(nc);
    for (uint32_t w = 0; w < 2 * bits; w++) more |= col[w].size() > 2;
  }
  // the two rows added with a ripple carry; the product has 2 * bits bits
  std::vector<Lit> prod(2 * bits);
  bool have_c = false;
  Lit carry{0, false};
  for (uint32_t w = 0; w < 2 * bits; w++) {
    std::vector<Lit> v = col[w];
    if (have_c) v.push_back(carry);
    if (v.empty()) {  // a column with nothing in it (weight 2 * bits - 1 for 2 bits): zero
      prod[w] = op(2, {0, false}, {0, false});
      have_c = false;
    } else if (v.size() == 1) {
      prod[w] = v[0];
      have_c = false;
    } else if (v.size() == 2) {
      prod[w] = op(2, v[0], v[1]);
      carry = op(0, v[0], v[1]);
      have_c = true;
    } else {
      const Lit t = op(2, v[0], v[1]);
      prod[w] = op(2, t, v[2]);
      carry = op(1, op(0, v[0], v[1]), op(0, t, v[2]));
      have_c = true;
    }
  }
  // [a * b == N] and a != 1 and b != 1 (an input of ones is not a factorisation)
  std::vector<Lit> lits;
  for (uint32_t w = 0; w < 2 * bits; w++) lits.push_back({prod[w].first, prod[w].second ^ (((N >> w) & 1) == 0)});
  for (uint32_t f = 0; f < 2; f++) {
    Lit any{f * bits, true};  // a != 1: ~a_0 or a_1 or ... or a_{bits-1}
    for (uint32_t i = 1; i < bits; i++) any = op(1, any, {f * bits + i, false});
    lits.push_back(any);
  }
  auto r = and_tree(lb, lits);
  lb.output(r.first, r.second);
  return lb.c;
}

Circuit gen_qsat_circuit(uint64_t seed, uint32_t n, uint32_t gates, uint32_t last) {
  if (n == 0 || n > 64 || last < 3 || last > gates) throw std::runtime_error("bad QSAT generator parameters");
  Rng rng(seed);
  LitBuilder lb(n);
  std::vector<uint32_t> in(n);
  for (uint32_t i = 0; i < n; i++) in[i] = i;
  auto made = random_block(lb, rng, in, gates);
  uint32_t pick[3];
  for (int i = 0; i < 3; i++) {
    while (true) {
      uint32_t t = (uint32_t)rng.below(last);
      bool dup = false;
      for (int j = 0; j < i; j++) dup |= pick[j] == t;
      if (!dup) {
        pick[i] = t;
        break;
      }
    }
  }
  uint32_t g0 = made[gates - last + pick[0]], g1 = made[gates - last + pick[1]],
           g2 = made[gates - last + pick[2]];
  uint32_t x = lb.gate(2, g0, false, g1, false);
  uint32_t y = lb.gate(2, x, false, g2, false);
  lb.output(y, false);
  return lb.c;
}

Circuit gen_satostop_circuit(uint64_t seed, uint32_t half, uint32_t rounds, uint32_t round_gates,
                             uint32_t unknowns, uint64_t* target) {
  uint32_t n = 2 * half;
  if (half == 0 || n > 62 || round_gates < half || unknowns > n)
    throw std::runtime_error("bad SATOSTOP generator parameters");
  Rng rng(seed);
  LitBuilder lb(n);
  std::vector<uint32_t> L(half), R(half);
  for (uint32_t i = 0; i < half; i++) {
    L[i] = i;
    R[i] = half + i;
  }
  for (uint32_t r = 0; r < rounds; r++) {
    auto made = random_block(lb, rng, R, round_gates);
    std::vector<uint32_t> nr(half);
    for (uint32_t i = 0; i < half; i++)
      nr[i] = lb.gate(2, L[i], false, made[round_gates - half + i], false);
    L = R;
    R = nr;
  }
  std::vector<uint32_t> next(n);
  for (uint32_t i = 0; i < half; i++) {
    next[i] = L[i];
    next[half + i] = R[i];
  }
  // plant the target: state reached after m steps from a random unknown start
  uint64_t u0 = unknowns ? rng.below(1ULL << unknowns) : 0;
  uint64_t s = u0 << (n - unknowns);
  uint32_t m = 1 + (uint32_t)rng.below(64);
  for (uint32_t step = 0; step < m; step++) {
    auto val = eval_nodes(lb, s);
    uint64_t ns = 0;
    for (uint32_t i = 0; i < n; i++) ns |= (uint64_t)val[next[i]] << i;
    s = ns;
  }
  std::vector<std::pair<uint32_t, bool>> lits;
  for (uint32_t i = 0; i < n; i++) lits.push_back({next[i], ((s >> i) & 1) == 0});
  auto stop = and_tree(lb, lits);
  for (uint32_t i = 0; i < n; i++) lb.output(next[i], false);
  lb.output(stop.first, stop.second);
  if (target) *target = s;
  return lb.c;
}

}  // namespace ns
