// Native code path: register allocation of a mapped LUT program and its assembly text.
//
// The body runs inside a template kernel (templates/gen_templates.py) with this contract:
//   v[lm .. lm+5]  lane masks (input), s20/s21 wave-uniform word index (input),
//   v0 .. v[m-1]   outputs, v[pool_lo .. pool_hi] and s40..s95 free for the body.
// Wave-uniform values (constant word masks, uniform index bits, uniform gates) live in
// SGPRs and are computed with scalar instructions; lane-varying values live in VGPRs and
// every LUT is one v_bitop3_b32 (at most one SGPR operand per instruction).
#include "asmgen.h"

#include <algorithm>
#include <sstream>

namespace ns {

namespace {

const uint32_t kWordMasks[5] = {0xAAAAAAAAu, 0xCCCCCCCCu, 0xF0F0F0F0u, 0xFF00FF00u, 0xFFFF0000u};
const uint32_t kSgprLo = 40, kSgprHi = 95;

struct Loc {
  char file = 0;  // 'v' or 's' (0 = not yet defined)
  uint32_t reg = 0;
};

std::string rname(const Loc& l) { return std::string(1, l.file) + std::to_string(l.reg); }

std::string hex(uint32_t v) {
  std::ostringstream o;
  o << "0x" << std::hex << v;
  return o.str();
}

class Allocator {
 public:
  Allocator(const Circuit& c, const LutProgram& p, const std::vector<uint32_t>& vbit, uint32_t lane_shift,
            uint32_t lm_base, uint32_t pool_lo, uint32_t pool_hi, const std::vector<uint32_t>& out_wires,
            const std::vector<uint8_t>& out_neg)
      : c_(c), p_(p), vbit_(vbit), ls_(lane_shift), lm_(lm_base), pool_lo_(pool_lo), pool_hi_(pool_hi),
        loc_(c.wires()), last_(c.wires(), -1), outw_(out_wires), outneg_(out_neg) {
    for (uint32_t r = pool_lo; r <= pool_hi; r++) vfree_.push_back(r);
    for (uint32_t r = kSgprLo; r <= kSgprHi; r++) sfree_.push_back(r);
    std::reverse(vfree_.begin(), vfree_.end());
    std::reverse(sfree_.begin(), sfree_.end());
    const int64_t end = (int64_t)p.ops.size();
    for (size_t i = 0; i < p.ops.size(); i++) {
      const LutOp& op = p.ops[i];
      for (int k = 0; k < (op.lut ? 3 : 2); k++) last_[op.src[k]] = (int64_t)i;
    }
    for (uint32_t w : out_wires) last_[w] = end;
    // lane inputs sit in their template registers
    for (uint32_t k = 0; k < c.n; k++) {
      uint32_t vb = vbit[k];
      if (vb >= 5 && vb - 5 >= ls_ && vb - 5 < ls_ + 6) {
        loc_[k].file = 'v';
        loc_[k].reg = lm_ + (vb - 5 - ls_);
      }
    }
    // first output position of each wire that is produced non-negated by a LUT
    for (size_t k = 0; k < out_wires.size(); k++)
      if (!out_neg[k] && out_wires[k] >= c.n && !p.uniform[out_wires[k]]) {
        bool seen = false;
        for (auto& pc : precolor_) seen |= pc.first == out_wires[k];
        if (!seen) precolor_.push_back({out_wires[k], (uint32_t)k});
      }
  }

  std::string run() {
    for (size_t i = 0; i < p_.ops.size(); i++) {
      const LutOp& op = p_.ops[i];
      if (!op.lut) {
        std::string a = use_scalar(op.src[0]), b = use_scalar(op.src[1]);
        release_dead(op.src, 2, (int64_t)i);
        Loc d = alloc('s');
        loc_[op.dst] = d;
        static const char* names[4] = {"s_and_b32", "s_nor_b32", "s_andn2_b32", "s_xor_b32"};
        out_ << "\t" << names[op.fn] << " " << rname(d) << ", " << a << ", " << b << "\n";
        // a value never used again (possible for dead uniform gates) is released at once
        if (last_[op.dst] <= (int64_t)i) release(op.dst);
        continue;
      }
      // sources: at most one distinct SGPR; further SGPR operands are copied to VGPRs
      std::string src[3];
      std::vector<Loc> temps;
      uint32_t sgpr_wire = UINT32_MAX;
      for (int k = 0; k < 3; k++) {
        uint32_t w = op.src[k];
        Loc l = define(w);
        if (l.file == 's') {
          if (sgpr_wire == UINT32_MAX || sgpr_wire == w) {
            sgpr_wire = w;
            src[k] = rname(l);
          } else {
            Loc t = alloc('v');
            temps.push_back(t);
            out_ << "\tv_mov_b32 " << rname(t) << ", " << rname(l) << "\n";
            src[k] = rname(t);
          }
        } else {
          src[k] = rname(l);
        }
      }
      release_dead(op.src, 3, (int64_t)i);
      for (auto& t : temps) free_reg(t);
      Loc d;
      bool pre = false;
      for (auto& pc : precolor_)
        if (pc.first == op.dst) {
          d.file = 'v';
          d.reg = pc.second;
          pre = true;
        }
      if (!pre) d = alloc('v');
      loc_[op.dst] = d;
      out_ << "\tv_bitop3_b32 " << rname(d) << ", " << src[0] << ", " << src[1] << ", " << src[2] << " bitop3:"
           << hex(op.fn) << "\n";
      if (last_[op.dst] <= (int64_t)i) release(op.dst);
    }
    // outputs into v0..
    for (size_t k = 0; k < outw_.size(); k++) {
      uint32_t w = outw_[k];
      Loc l = define(w);
      if (l.file == 'v' && l.reg == k && !outneg_[k]) continue;
      out_ << "\t" << (outneg_[k] ? "v_not_b32 v" : "v_mov_b32 v") << k << ", " << rname(l) << "\n";
    }
    return out_.str();
  }

  uint32_t max_vgpr() const { return vmax_; }

 private:
  // location of wire w, emitting the definition of a uniform input on first use
  Loc define(uint32_t w) {
    if (loc_[w].file) return loc_[w];
    if (w >= c_.n) throw std::runtime_error("internal: value used before definition");
    uint32_t vb = vbit_[w];
    Loc d = alloc('s');
    if (vb < 5) {
      out_ << "\ts_mov_b32 " << rname(d) << ", " << hex(kWordMasks[vb]) << "\n";
    } else {
      uint32_t wb = vb - 5;  // bit of the word index, taken from the uniform part s20/s21
      out_ << "\ts_bfe_i32 " << rname(d) << ", " << (wb < 32 ? "s20" : "s21") << ", "
           << hex((1u << 16) | (wb & 31u)) << "\n";
    }
    loc_[w] = d;
    return d;
  }

  std::string use_scalar(uint32_t w) {
    Loc l = define(w);
    if (l.file != 's') throw std::runtime_error("internal: scalar gate reads a lane value");
    return rname(l);
  }

  void release_dead(const uint32_t* src, int n, int64_t i) {
    for (int k = 0; k < n; k++) {
      uint32_t w = src[k];
      bool dup = false;
      for (int q = 0; q < k; q++) dup |= src[q] == w;
      if (!dup && last_[w] == i) release(w);
    }
  }

  void release(uint32_t w) {
    Loc& l = loc_[w];
    if (!l.file) return;
    // lane masks and output registers are not pool registers
    if (l.file == 's' || (l.file == 'v' && l.reg >= pool_lo_ && l.reg <= pool_hi_)) free_reg(l);
    l.file = 'x';  // released (never read again)
  }

  void free_reg(const Loc& l) {
    if (l.file == 'v') vfree_.push_back(l.reg);
    else if (l.file == 's') sfree_.push_back(l.reg);
  }

  Loc alloc(char file) {
    auto& fl = file == 'v' ? vfree_ : sfree_;
    if (fl.empty())
      throw NativeUnfit(std::string("out of ") + (file == 'v' ? "VGPRs" : "SGPRs") + " for the native body");
    Loc l;
    l.file = file;
    l.reg = fl.back();
    fl.pop_back();
    if (file == 'v') vmax_ = std::max(vmax_, l.reg);
    return l;
  }

  const Circuit& c_;
  const LutProgram& p_;
  const std::vector<uint32_t>& vbit_;
  uint32_t ls_, lm_, pool_lo_, pool_hi_;
  std::vector<Loc> loc_;
  std::vector<int64_t> last_;
  std::vector<uint32_t> vfree_, sfree_;
  std::vector<std::pair<uint32_t, uint32_t>> precolor_;
  const std::vector<uint32_t>& outw_;
  const std::vector<uint8_t>& outneg_;
  std::ostringstream out_;
  uint32_t vmax_ = 0;
};

const char* kind_name(NativeKind k) {
  switch (k) {
    case NativeKind::SAT: return "sat";
    case NativeKind::QSAT_A: return "qsata";
    case NativeKind::QSAT_B: return "qsatb";
    default: return "table";
  }
}

}  // namespace

NativeKernel build_native(const Circuit& c, NativeKind kind, const std::vector<uint32_t>& vbit,
                          uint32_t lane_shift, const std::vector<uint32_t>& outs) {
  NativeKernel nk;
  std::vector<uint32_t> out_wires;
  std::vector<uint8_t> out_neg;
  for (uint32_t k : outs) {
    out_wires.push_back(c.outputs.at(k).wire);
    out_neg.push_back(c.outputs.at(k).neg ? 1 : 0);
  }
  auto cone = output_cone(c, outs);
  auto order = schedule_gates(c, cone, outs);
  std::vector<uint8_t> uni(c.n, 1);
  for (uint32_t k = 0; k < c.n; k++)
    if (vbit[k] >= 5 && vbit[k] - 5 >= lane_shift && vbit[k] - 5 < lane_shift + 6) uni[k] = 0;
  LutProgram p = map_luts(c, order, out_wires, uni);
  nk.info.cone_gates = (uint32_t)order.size();
  nk.info.peak_live = peak_live(c, order, outs);
  nk.info.valu_ops = p.valu_ops;
  nk.info.scalar_ops = p.scalar_ops;
  const bool table = kind == NativeKind::TABLE;
  const uint32_t lm = table ? 32 : 1;
  const uint32_t pool_lo = table ? 40 : 8;
  struct Cls {
    uint32_t budget, hi;
  };
  const Cls classes[3] = {{64, 55}, {128, 119}, {256, 239}};
  std::string body;
  for (const Cls& cl : classes) {
    if (table && cl.budget < 128) continue;
    try {
      Allocator a(c, p, vbit, lane_shift, lm, pool_lo, cl.hi, out_wires, out_neg);
      body = a.run();
      nk.vgpr_class = cl.budget;
      break;
    } catch (const NativeUnfit&) {
      continue;
    }
  }
  if (!nk.vgpr_class) throw NativeUnfit("circuit needs more registers than the native templates provide");
  nk.kernel_name = std::string("ns_nat_") + kind_name(kind) + "_r" + std::to_string(nk.vgpr_class);
  const char* tpl = native_template(nk.kernel_name);
  if (!tpl) throw std::runtime_error("missing native template " + nk.kernel_name);
  std::string text(tpl);
  std::string marker = "; NS_BODY " + nk.kernel_name;
  size_t pos = text.find(marker);
  if (pos == std::string::npos) throw std::runtime_error("template placeholder not found");
  size_t eol = text.find('\n', pos);
  text.replace(pos, eol - pos, "; circuit body\n" + body);
  nk.asm_text = std::move(text);
  return nk;
}

}  // namespace ns
