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
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <set>
#include <sstream>
#include <thread>
#include <tuple>

namespace ns {

namespace {

const uint32_t kWordMasks[5] = {0xAAAAAAAAu, 0xCCCCCCCCu, 0xF0F0F0F0u, 0xFF00FF00u, 0xFFFF0000u};
// future uses of a new value whose other VGPR sources steer the bank of its register
constexpr size_t kBankLookahead = 8;
// labels of the group exits (one body per module or code object region: unique there)
const char* const kExitLabel = ".Lns_exit";
const char* const kDoneLabel = ".Lns_done";
// SGPRs of the body: s[40..101]; the SAT templates, whose own code keeps only s2 (its loop
// counter) across the body, also give it s0, s1, s[3..19] and s[22..39] (s20 and s21 carry
// the word index)
const uint32_t kSgprLoSat = 0, kSgprLo = 40, kSgprLoBatch = 48, kSgprHi = 101;
const uint32_t kSatLoopSgpr = 2;
// state tiers of the walk templates (templates/gen_templates.py WALK_TIERS)
const uint32_t kWalkTiers[5] = {16, 24, 32, 40, 48};

struct Loc {
  int v = -1;   // VGPR holding the value
  int a = -1;   // AGPR copy (spill slot of a lane value)
  int s = -1;   // SGPR (wave-uniform values)
  int sl = -1;  // lane slot of a wave-uniform value parked in a VGPR lane
  bool dead = false;
};

// wave-uniform values can be parked in lanes of the highest pool VGPRs (64 slots each)

std::string hex(uint32_t v) {
  char buf[12];
  char* e = buf + sizeof(buf);
  char* p = e;
  do {
    *--p = "0123456789abcdef"[v & 15u];
    v >>= 4;
  } while (v);
  *--p = 'x';
  *--p = '0';
  return std::string(p, e);
}

// Text of a body as it is built: appends without the stream machinery (a body is thousands
// of short lines and is generated for every circuit served)
class BodyText {
 public:
  BodyText() { s_.reserve(1 << 16); }
  BodyText& operator<<(const char* x) {
    s_ += x;
    return *this;
  }
  BodyText& operator<<(const std::string& x) {
    s_ += x;
    return *this;
  }
  BodyText& operator<<(char c) {
    s_ += c;
    return *this;
  }
  BodyText& operator<<(long long v) {
    char buf[24];
    char* e = buf + sizeof(buf);
    char* p = e;
    const bool neg = v < 0;
    unsigned long long u = neg ? 0ull - (unsigned long long)v : (unsigned long long)v;
    do {
      *--p = (char)('0' + u % 10);
      u /= 10;
    } while (u);
    if (neg) *--p = '-';
    s_.append(p, e);
    return *this;
  }
  BodyText& operator<<(int v) { return *this << (long long)v; }
  BodyText& operator<<(unsigned v) { return *this << (long long)v; }
  BodyText& operator<<(unsigned long v) { return *this << (long long)v; }
  const std::string& str() const { return s_; }

 private:
  std::string s_;
};

// Linear allocation in evaluation order. VGPRs come from the pool; when it is exhausted
// (and the template provides AGPRs) the resident value with the furthest next use is
// copied to an AGPR (v_accvgpr_write) and read back (v_accvgpr_read) before its next use.
class Allocator {
 public:
  // in_regs (optional): input k sits in VGPR in_regs[k] (the walk bodies: every input is a
  // template register); out_regs (optional): output k goes to VGPR out_regs[k] instead of vk,
  // which may be an input's register: an output is computed into it only once that input is
  // dead, otherwise copied there at the end
  // exits (optional): running conjunctions of the output groups, (wire, negated); after each
  // is computed the body leaves for its end with output 0 when it is zero in every lane
  // (s_cbranch_vccz), since the later groups cannot make the output nonzero
  Allocator(const Circuit& c, const LutProgram& p, const std::vector<uint32_t>& vbit, uint32_t lane_shift,
            uint32_t lm_base, uint32_t pool_lo, uint32_t pool_hi, uint32_t agprs, uint32_t slot_regs,
            const std::vector<uint32_t>& out_wires, const std::vector<uint8_t>& out_neg, uint32_t sgpr_lo,
            const std::vector<int>* in_regs = nullptr, const std::vector<uint32_t>* out_regs = nullptr,
            const std::vector<std::pair<uint32_t, bool>>* exits = nullptr, bool vcc_regs = false)
      : c_(c), p_(p), vbit_(vbit), ls_(lane_shift), lm_(lm_base), pool_lo_(pool_lo), pool_hi_(pool_hi),
        sgpr_lo_(sgpr_lo),
        loc_(c.wires()), uses_(c.wires()), ptr_(c.wires(), 0), outw_(out_wires), outneg_(out_neg) {
    for (size_t k = 0; k < out_wires.size(); k++) outreg_.push_back(out_regs ? (*out_regs)[k] : (uint32_t)k);
    if (exits) exits_ = *exits;
    slot_base_ = (int)pool_hi - (int)slot_regs + 1;
    pool_hi_ = pool_hi - slot_regs;
    for (int r = (int)pool_hi_; r >= (int)pool_lo; r--) vfree_[r & 3].push_back(r);
    for (int r = (int)(64 * slot_regs) - 1; r >= 0; r--) slotfree_.push_back(r);
    // vcc_lo / vcc_hi (SGPRs 106, 107) as two more, taken last, when the template gives vcc
    // and the body tests no exits
    if (vcc_regs) sregs_ = {107, 106};
    for (int r = (int)kSgprHi; r >= (int)sgpr_lo_; r--)
      if (r != 20 && r != 21 && !(sgpr_lo_ == kSgprLoSat && r == (int)kSatLoopSgpr)) sregs_.push_back(r);
    sfree_ = sregs_;
    for (int r = (int)agprs - 1; r >= 0; r--) afree_.push_back(r);
    vowner_.assign(512, UINT32_MAX);
    end_ = (uint32_t)p.ops.size();
    for (uint32_t i = 0; i < p.ops.size(); i++) {
      const LutOp& op = p.ops[i];
      const int ns = op.lut ? 3 : 2;
      for (int k = 0; k < ns; k++) {
        bool dup = false;
        for (int q = 0; q < k; q++) dup |= op.src[q] == op.src[k];
        if (!dup) uses_[op.src[k]].push_back(i);
      }
    }
    for (uint32_t w : out_wires) uses_[w].push_back(end_);
    // lane inputs sit in their template registers
    for (uint32_t k = 0; k < c.n; k++) {
      uint32_t vb = vbit[k];
      if (in_regs && (*in_regs)[k] >= 0) {
        loc_[k].v = (*in_regs)[k];
        vowner_[(size_t)(*in_regs)[k]] = k;
      } else if (vb >= 5 && vb - 5 >= ls_ && vb - 5 < ls_ + 6) {
        loc_[k].v = (int)(lm_ + (vb - 5 - ls_));
      }
    }
    // first output position of each wire that is produced non-negated by a LUT
    for (size_t k = 0; k < out_wires.size(); k++)
      if (!out_neg[k] && out_wires[k] >= c.n && !p.uniform[out_wires[k]]) {
        bool seen = false;
        for (auto& pc : precolor_) seen |= pc.first == out_wires[k];
        if (!seen) precolor_.push_back({out_wires[k], outreg_[k]});
      }
  }

  std::string run() {
    for (uint32_t k = 0; k < c_.n; k++)
      if (uses_[k].empty()) kill(k);  // unread inputs: their registers may take outputs
    for (uint32_t i = 0; i < p_.ops.size(); i++) {
      const LutOp& op = p_.ops[i];
      cur_ = i;
      prefetch_uniform(i + 2);
      if (!op.lut) {
        pinned_.assign(op.src, op.src + 2);
        std::string a = sreg(op.src[0]), b = sreg(op.src[1]);
        consume(op.src, 2);
        int d = alloc_s();
        pinned_.clear();
        set_s(op.dst, d);
        static const char* names[4] = {"s_and_b32", "s_nor_b32", "s_andn2_b32", "s_xor_b32"};
        out_ << "\t" << names[op.fn] << " " << sname(d) << ", " << a << ", " << b << "\n";
        if (uses_[op.dst].empty()) kill(op.dst);
        continue;
      }
      emit_lut(op);
    }
    return finish();
  }

  void emit_lut(const LutOp& op) {
    {
      // sources: at most one distinct SGPR; further SGPR operands are copied to VGPRs
      std::string src[3];
      std::vector<int> temps;
      uint32_t sgpr_wire = UINT32_MAX;
      pinned_.clear();
      for (int k = 0; k < 3; k++) pinned_.push_back(op.src[k]);
      for (int k = 0; k < 3; k++) {
        uint32_t w = op.src[k];
        if (p_.uniform[w]) {
          std::string r = sreg(w);
          if (sgpr_wire == UINT32_MAX || sgpr_wire == w) {
            sgpr_wire = w;
            src[k] = r;
          } else {
            int t = alloc_v(UINT32_MAX);
            temps.push_back(t);
            out_ << "\tv_mov_b32 v" << t << ", " << r << "\n";
            src[k] = "v" + std::to_string(t);
          }
        } else {
          src[k] = "v" + std::to_string(vreg(w));
        }
      }
      pinned_.clear();
      consume(op.src, 3);
      for (int t : temps) free_v(t);
      int d = -1;
      for (auto& pc : precolor_)
        if (pc.first == op.dst && output_reg_free((int)pc.second)) d = (int)pc.second;
      if (d < 0) d = alloc_v(op.dst);
      else vowner_[(size_t)d] = op.dst;
      loc_[op.dst].v = d;
      out_ << "\tv_bitop3_b32 v" << d << ", " << src[0] << ", " << src[1] << ", " << src[2] << " bitop3:"
           << hex(op.fn) << "\n";
      for (const auto& ex : exits_)
        if (ex.first == op.dst) {
          // a negated conjunction literal is zero where its wire is all ones
          out_ << "\tv_cmp_ne_u32 vcc, " << (ex.second ? "-1" : "0") << ", v" << d << "\n\ts_cbranch_vccz "
               << kExitLabel << "\n";
          exit_tests_++;
        }
      if (uses_[op.dst].empty()) kill(op.dst);
    }
  }

  std::string finish() {
    // outputs into their registers (v0.. or out_regs)
    cur_ = end_;
    // a value still to be copied from a register that another output is about to receive is
    // first moved to a free pool register (outputs may go to the inputs' registers)
    auto pending = [&](size_t k) {
      const uint32_t w = outw_[k];
      return p_.uniform[w] || outneg_[k] || loc_[w].v != (int)outreg_[k];
    };
    for (size_t k = 0; k < outw_.size(); k++) {
      const uint32_t w = outw_[k];
      if (!pending(k) || p_.uniform[w] || loc_[w].v < 0) continue;
      bool clash = false;
      for (size_t j = 0; j < outw_.size(); j++) clash |= j != k && pending(j) && (int)outreg_[j] == loc_[w].v;
      if (!clash) continue;
      const int t = alloc_v(UINT32_MAX);
      out_ << "\tv_mov_b32 v" << t << ", v" << loc_[w].v << "\n";
      loc_[w].v = t;
    }
    std::vector<uint8_t> done(outw_.size(), 0);
    for (size_t k = 0; k < outw_.size(); k++) done[k] = !pending(k);
    for (size_t k = 0; k < outw_.size(); k++) {
      if (done[k]) continue;
      uint32_t w = outw_[k];
      const uint32_t dk = outreg_[k];
      const char* mn = outneg_[k] ? "v_not_b32" : "v_mov_b32";
      if (p_.uniform[w]) {
        std::string r = sreg(w);  // may emit its own definition first
        out_ << "\t" << mn << " v" << dk << ", " << r << "\n";
        continue;
      }
      const Loc& l = loc_[w];
      if (l.v >= 0) {
        out_ << "\t" << mn << " v" << dk << ", v" << l.v << "\n";
      } else {
        out_ << "\tv_accvgpr_read_b32 v" << dk << ", a" << l.a << "\n";
        if (outneg_[k]) out_ << "\tv_not_b32 v" << dk << ", v" << dk << "\n";
      }
    }
    if (exit_tests_) {
      if (outw_.size() != 1 || outreg_[0] != 0) throw std::runtime_error("internal: group exits need one output in v0");
      out_ << "\ts_branch " << kDoneLabel << "\n" << kExitLabel << ":\n\tv_mov_b32 v0, 0\n" << kDoneLabel << ":\n";
    }
    return out_.str();
  }

  uint32_t exit_tests() const { return exit_tests_; }

  uint32_t spills() const { return spills_; }
  uint32_t reloads() const { return reloads_; }

 private:
  bool is_pool_v(int r) const { return r >= (int)pool_lo_ && r <= (int)pool_hi_; }

  static std::string sname(int r) {
    if (r == 106) return "vcc_lo";
    if (r == 107) return "vcc_hi";
    return "s" + std::to_string(r);
  }

  // an output may be computed straight into register r when r holds no value still needed
  bool output_reg_free(int r) const {
    const uint32_t w = vowner_[(size_t)r];
    return w == UINT32_MAX || loc_[w].dead || loc_[w].v != r;
  }

  // SGPR of a wave-uniform value: defines a uniform input on first use and reloads a value
  // parked in a VGPR lane
  std::string sreg(uint32_t w) {
    Loc& l = loc_[w];
    if (l.s >= 0) return sname(l.s);
    if (l.sl >= 0) {
      int d = alloc_s();
      reload_slot(w, d);
      out_ << "\ts_nop 1\n";  // two wait states between the SGPR write and its read
      return sname(d);
    }
    if (w >= c_.n) throw std::runtime_error("internal: uniform value used before definition");
    uint32_t vb = vbit_[w];
    int d = alloc_s();
    if (vb < 5) {
      out_ << "\ts_mov_b32 " << sname(d) << ", " << hex(kWordMasks[vb]) << "\n";
    } else {
      uint32_t wb = vb - 5;  // bit of the word index, taken from the uniform part s20/s21
      out_ << "\ts_bfe_i32 " << sname(d) << ", " << (wb < 32 ? "s20" : "s21") << ", " << hex((1u << 16) | (wb & 31u))
           << "\n";
    }
    set_s(w, d);
    return sname(d);
  }

  void reload_slot(uint32_t w, int d) {
    Loc& l = loc_[w];
    out_ << "\tv_readlane_b32 " << sname(d) << ", v" << (slot_base_ + l.sl / 64) << ", " << (l.sl % 64) << "\n";
    set_s(w, d);
    reloads_++;
  }

  // reads the parked uniform sources of op j into free SGPRs two ops ahead, so that the
  // v_readlane and the read of its result are separated without an s_nop
  void prefetch_uniform(uint32_t j) {
    if (j >= end_) return;
    const LutOp& op = p_.ops[j];
    for (int k = 0; k < (op.lut ? 3 : 2); k++) {
      uint32_t w = op.src[k];
      Loc& l = loc_[w];
      if (!p_.uniform[w] || l.s >= 0 || l.sl < 0 || sfree_.empty()) continue;
      int d = sfree_.back();
      sfree_.pop_back();
      sowner_[d] = UINT32_MAX;
      reload_slot(w, d);
    }
  }

  // VGPR of a lane value, reloading it from its AGPR copy when needed
  int vreg(uint32_t w) {
    Loc& l = loc_[w];
    if (l.v >= 0) return l.v;
    if (l.a < 0) throw std::runtime_error("internal: lane value used before definition");
    int r = alloc_v(w);
    out_ << "\tv_accvgpr_read_b32 v" << r << ", a" << l.a << "\n";
    l.v = r;
    reloads_++;
    return r;
  }

  // the current op has read its sources: advance their next-use pointers and free the dead
  void consume(const uint32_t* src, int n) {
    for (int k = 0; k < n; k++) {
      uint32_t w = src[k];
      bool dup = false;
      for (int q = 0; q < k; q++) dup |= src[q] == w;
      if (dup) continue;
      if (ptr_[w] < uses_[w].size() && uses_[w][ptr_[w]] == cur_) ptr_[w]++;
      if (ptr_[w] >= uses_[w].size()) kill(w);
    }
  }

  void kill(uint32_t w) {
    Loc& l = loc_[w];
    if (l.dead) return;
    if (l.v >= 0 && is_pool_v(l.v)) free_v(l.v);
    if (l.a >= 0) afree_.push_back(l.a);
    if (l.s >= 0) drop_s(w);
    if (l.sl >= 0) slotfree_.push_back(l.sl);
    l.v = l.a = l.s = l.sl = -1;
    l.dead = true;
  }

  uint32_t next_use(uint32_t w) const { return ptr_[w] < uses_[w].size() ? uses_[w][ptr_[w]] : UINT32_MAX; }

  void set_s(uint32_t w, int d) {
    loc_[w].s = d;
    sowner_[d] = w;
  }

  void drop_s(uint32_t w) {
    sowner_[loc_[w].s] = UINT32_MAX;
    sfree_.push_back(loc_[w].s);
    loc_[w].s = -1;
  }

  int alloc_s() {
    if (sfree_.empty()) evict_sgpr();
    int r = sfree_.back();
    sfree_.pop_back();
    sowner_[r] = UINT32_MAX;
    return r;
  }

  // parks the uniform value with the furthest next use in a VGPR lane
  void evict_sgpr() {
    uint32_t best = UINT32_MAX;
    uint64_t best_use = 0;
    for (int r : sregs_) {
      const uint32_t w = sowner_[r];
      if (w == UINT32_MAX || loc_[w].dead) continue;
      bool pin = false;
      for (uint32_t pw : pinned_) pin |= pw == w;
      if (pin) continue;
      // inputs are recomputed with one scalar instruction, so they go first
      uint64_t nu = next_use(w);
      if (w < c_.n) nu = nu * 4 + 4;
      if (best == UINT32_MAX || nu > best_use) {
        best = w;
        best_use = nu;
      }
    }
    if (best == UINT32_MAX) throw NativeUnfit("out of SGPRs for the native body");
    Loc& l = loc_[best];
    if (best < c_.n) {
      drop_s(best);
      return;
    }
    if (l.sl < 0) {
      if (slotfree_.empty()) throw NativeUnfit("out of SGPRs and lane slots for the native body");
      l.sl = slotfree_.back();
      slotfree_.pop_back();
      out_ << "\tv_writelane_b32 v" << (slot_base_ + l.sl / 64) << ", " << sname(l.s) << ", " << (l.sl % 64) << "\n";
      spills_++;
    }
    drop_s(best);
  }

  void free_v(int r) {
    vowner_[r] = UINT32_MAX;
    vfree_[r & 3].push_back(r);
  }

  // bank (register index mod 4) of a lane value's VGPR, or -1
  int bank_of(uint32_t w) const {
    if (p_.uniform[w]) return -1;
    return loc_[w].v >= 0 ? (loc_[w].v & 3) : -1;
  }

  // A VALU op that reads two different VGPRs of one bank stalls. The new register of
  // `owner` avoids the banks of the current op's other VGPR sources and of the sources
  // of owner's next uses that already sit in VGPRs.
  int alloc_v(uint32_t owner) {
    if (vfree_[0].empty() && vfree_[1].empty() && vfree_[2].empty() && vfree_[3].empty()) spill_one();
    int pen[4] = {0, 0, 0, 0};
    // a temporary copy (no owner) is a source of the current op; a destination is not
    if (owner == UINT32_MAX && cur_ < end_ && p_.ops[cur_].lut)
      for (uint32_t s : p_.ops[cur_].src)
        if (bank_of(s) >= 0) pen[bank_of(s)] += 8;
    if (owner != UINT32_MAX) {
      const auto& u = uses_[owner];
      for (size_t k = ptr_[owner], m = 0; k < u.size() && m < kBankLookahead; k++, m++) {
        if (u[k] >= end_ || u[k] == cur_) continue;
        const LutOp& op = p_.ops[u[k]];
        if (!op.lut) continue;
        for (uint32_t s : op.src)
          if (s != owner && bank_of(s) >= 0) pen[bank_of(s)] += m == 0 ? 4 : 2;
      }
    }
    int best = -1;
    for (int b = 0; b < 4; b++) {
      if (vfree_[b].empty()) continue;
      if (best < 0 || pen[b] < pen[best] || (pen[b] == pen[best] && vfree_[b].size() > vfree_[best].size()))
        best = b;
    }
    int r = vfree_[best].back();
    vfree_[best].pop_back();
    vowner_[r] = owner;
    return r;
  }

  // evicts the pool-resident value with the furthest next use into an AGPR
  void spill_one() {
    uint32_t best = UINT32_MAX, best_use = 0;
    for (int r = (int)pool_lo_; r <= (int)pool_hi_; r++) {
      uint32_t w = vowner_[r];
      if (w == UINT32_MAX) continue;
      bool pin = false;
      for (uint32_t pw : pinned_) pin |= pw == w;
      if (pin) continue;
      uint32_t nu = next_use(w);
      if (best == UINT32_MAX || nu > best_use) {
        best = w;
        best_use = nu;
      }
    }
    if (best == UINT32_MAX) throw NativeUnfit("no VGPR can be spilled");
    Loc& l = loc_[best];
    if (l.a < 0) {
      if (afree_.empty()) throw NativeUnfit("out of VGPRs and AGPRs for the native body");
      l.a = afree_.back();
      afree_.pop_back();
      out_ << "\tv_accvgpr_write_b32 a" << l.a << ", v" << l.v << "\n";
      spills_++;
    }
    free_v(l.v);
    l.v = -1;
  }

  const Circuit& c_;
  const LutProgram& p_;
  const std::vector<uint32_t>& vbit_;
  uint32_t ls_, lm_, pool_lo_, pool_hi_, sgpr_lo_;
  std::vector<Loc> loc_;
  std::vector<std::vector<uint32_t>> uses_;
  std::vector<uint32_t> ptr_;
  std::vector<int> vfree_[4], sfree_, afree_, slotfree_;
  std::vector<int> sregs_;  // the body's SGPRs
  int slot_base_ = 0;
  std::vector<uint32_t> vowner_;
  std::vector<uint32_t> sowner_ = std::vector<uint32_t>(128, UINT32_MAX);
  std::vector<uint32_t> pinned_;
  std::vector<std::pair<uint32_t, uint32_t>> precolor_;
  std::vector<uint32_t> outreg_;
  const std::vector<uint32_t>& outw_;
  const std::vector<uint8_t>& outneg_;
  BodyText out_;
  uint32_t cur_ = 0, end_ = 0;
  uint32_t spills_ = 0, reloads_ = 0;
  std::vector<std::pair<uint32_t, bool>> exits_;
  uint32_t exit_tests_ = 0;
};

// Reorders a LUT program to lower the number of simultaneously live lane values: greedy
// list scheduling over the lane-varying ops (an op is ready when its lane-varying sources
// exist), preferring the op that ends the most lifetimes, then program order. Wave-uniform
// ops are emitted just before their first consumer, so scalar values live briefly.
// With wgroup (a group number per wire), the ready ops of the lowest group go first, so a
// circuit built as a sequence of groups is scheduled one group after another.
LutProgram pressure_order(const LutProgram& p, const std::vector<uint32_t>& out_wires, uint32_t wires,
                          const std::vector<uint32_t>* wgroup = nullptr) {
  const uint32_t m = (uint32_t)p.ops.size();
  std::vector<int32_t> prod(wires, -1);
  for (uint32_t i = 0; i < m; i++) prod[p.ops[i].dst] = (int32_t)i;
  auto nsrc = [&](const LutOp& op) { return op.lut ? 3 : 2; };
  std::vector<uint32_t> rem(wires, 0), npred(m, 0);
  std::vector<std::vector<uint32_t>> succ(m);
  for (uint32_t i = 0; i < m; i++) {
    const LutOp& op = p.ops[i];
    for (int k = 0; k < nsrc(op); k++) {
      uint32_t s = op.src[k];
      bool dup = false;
      for (int q = 0; q < k; q++) dup |= op.src[q] == s;
      if (dup) continue;
      rem[s]++;
      if (prod[s] >= 0 && !p.uniform[s]) {
        npred[i]++;
        succ[(uint32_t)prod[s]].push_back(i);
      }
    }
  }
  for (uint32_t w : out_wires) rem[w]++;
  LutProgram q;
  q.uniform = p.uniform;
  q.valu_ops = p.valu_ops;
  q.scalar_ops = p.scalar_ops;
  q.ops.reserve(m);
  std::vector<uint8_t> done(m, 0);
  std::vector<uint32_t> stack;
  auto emit = [&](uint32_t root) {
    // uniform producers first (depth-first), then the op itself
    stack.assign(1, root);
    while (!stack.empty()) {
      uint32_t i = stack.back();
      const LutOp& op = p.ops[i];
      bool pending = false;
      for (int k = 0; k < nsrc(op); k++) {
        int32_t pr = prod[op.src[k]];
        if (pr >= 0 && p.uniform[op.src[k]] && !done[(uint32_t)pr]) {
          stack.push_back((uint32_t)pr);
          pending = true;
          break;
        }
      }
      if (pending) continue;
      stack.pop_back();
      if (done[i]) continue;
      done[i] = 1;
      q.ops.push_back(op);
      for (int k = 0; k < nsrc(op); k++) {
        bool dup = false;
        for (int r = 0; r < k; r++) dup |= op.src[r] == op.src[k];
        if (!dup) rem[op.src[k]]--;
      }
    }
  };
  // The ready lane ops in an ordered set keyed (group, 1 - freed, index), freed = the op's
  // distinct lane sources it is the last use of: the set's first element is the op the
  // greedy rule takes. A source's remaining uses dropping to one bumps that last consumer.
  std::vector<std::vector<uint32_t>> users(wires);
  for (uint32_t i = 0; i < m; i++) {
    const LutOp& op = p.ops[i];
    for (int k = 0; k < nsrc(op); k++) {
      const uint32_t s = op.src[k];
      bool dup = false;
      for (int t = 0; t < k; t++) dup |= op.src[t] == s;
      if (!dup && prod[s] >= 0 && !p.uniform[s]) users[s].push_back(i);
    }
  }
  auto group_of = [&](uint32_t i) { return wgroup ? (*wgroup)[p.ops[i].dst] : 0u; };
  std::vector<int> freed(m, 0);
  std::vector<uint8_t> in_ready(m, 0);
  std::set<std::tuple<uint32_t, int, uint32_t>> ready;
  auto make_ready = [&](uint32_t i) {
    const LutOp& op = p.ops[i];
    int f = 0;
    for (int k = 0; k < nsrc(op); k++) {
      const uint32_t s = op.src[k];
      bool dup = false;
      for (int t = 0; t < k; t++) dup |= op.src[t] == s;
      if (!dup && prod[s] >= 0 && !p.uniform[s] && rem[s] == 1) f++;
    }
    freed[i] = f;
    in_ready[i] = 1;
    ready.insert({group_of(i), 1 - f, i});
  };
  // after an emitted op: lane sources left with one use make that user's choice cheaper
  auto on_emitted = [&](uint32_t i) {
    const LutOp& op = p.ops[i];
    for (int k = 0; k < nsrc(op); k++) {
      const uint32_t s = op.src[k];
      bool dup = false;
      for (int t = 0; t < k; t++) dup |= op.src[t] == s;
      if (dup || prod[s] < 0 || p.uniform[s] || rem[s] != 1) continue;
      for (uint32_t u : users[s])
        if (!done[u] && in_ready[u]) {
          ready.erase({group_of(u), 1 - freed[u], u});
          freed[u]++;
          ready.insert({group_of(u), 1 - freed[u], u});
        }
    }
  };
  for (uint32_t i = 0; i < m; i++)
    if (!p.uniform[p.ops[i].dst] && npred[i] == 0) make_ready(i);
  while (!ready.empty()) {
    const uint32_t i = std::get<2>(*ready.begin());
    ready.erase(ready.begin());
    in_ready[i] = 0;
    emit(i);
    on_emitted(i);
    for (uint32_t j : succ[i])
      if (--npred[j] == 0) make_ready(j);
  }
  for (uint32_t i = 0; i < m; i++)
    if (!done[i]) emit(i);
  return q;
}

// A LUT over one lane-varying value x and two wave-uniform values u1, u2 has to copy one of
// the uniform operands to a VGPR first (a vector instruction reads at most one SGPR here).
// When its cofactors in (u1, u2) take at most two functions of x, the LUT equals g(x, U) with
// U = h(u1, u2): one scalar gate (appended to cx) and a LUT reading x and the SGPR of U.
void split_uniform_pairs(Circuit& cx, LutProgram& p) {
  std::vector<LutOp> ops;
  ops.reserve(p.ops.size() + 64);
  for (const LutOp& op : p.ops) {
    if (!op.lut || op.src[0] == op.src[1] || op.src[0] == op.src[2] || op.src[1] == op.src[2]) {
      ops.push_back(op);
      continue;
    }
    int xi = -1, ui = -1, uj = -1;
    for (int k = 0; k < 3; k++) {
      if (!p.uniform[op.src[k]]) {
        if (xi >= 0) xi = 3;  // two lane-varying sources
        else xi = k;
      } else if (ui < 0) {
        ui = k;
      } else {
        uj = k;
      }
    }
    if (xi < 0 || xi > 2 || uj < 0) {
      ops.push_back(op);
      continue;
    }
    // cofactor code of (a, b) = (u1, u2): bit 0 = value at x = 0, bit 1 = value at x = 1
    auto bit_at = [&](uint32_t x, uint32_t a, uint32_t b) {
      uint32_t v[3];
      v[xi] = x;
      v[ui] = a;
      v[uj] = b;
      return (op.fn >> ((v[0] << 2) | (v[1] << 1) | v[2])) & 1u;
    };
    uint32_t code[2][2];
    for (uint32_t a = 0; a < 2; a++)
      for (uint32_t b = 0; b < 2; b++) code[a][b] = bit_at(0, a, b) | (bit_at(1, a, b) << 1);
    const uint32_t f0 = code[0][0];
    uint32_t f1 = f0, distinct = 1;
    for (uint32_t a = 0; a < 2; a++)
      for (uint32_t b = 0; b < 2; b++)
        if (code[a][b] != f0) {
          if (distinct == 1) {
            f1 = code[a][b];
            distinct = 2;
          } else if (code[a][b] != f1) {
            distinct = 3;
          }
        }
    if (distinct != 2) {
      ops.push_back(op);
      continue;
    }
    uint32_t t = 0;  // bit (a << 1 | b): cofactor f1
    for (uint32_t a = 0; a < 2; a++)
      for (uint32_t b = 0; b < 2; b++)
        if (code[a][b] == f1) t |= 1u << ((a << 1) | b);
    const uint32_t u1 = op.src[ui], u2 = op.src[uj];
    // U as one gate of the circuit's four functions over (u1, u2) or (u2, u1), or its complement
    auto gate2 = [](int f, uint32_t a, uint32_t b) -> uint32_t {
      switch (f) {
        case FUNC_AND: return a & b;
        case FUNC_NOR: return (a | b) ^ 1u;
        case FUNC_NIMPL: return a & (b ^ 1u);
        default: return a ^ b;
      }
    };
    int func = -1;
    bool swap = false, pol = true;
    for (int f = 0; f < 4 && func < 0; f++)
      for (int sw = 0; sw < 2 && func < 0; sw++) {
        uint32_t h = 0;
        for (uint32_t a = 0; a < 2; a++)
          for (uint32_t b = 0; b < 2; b++)
            if (gate2(f, sw ? b : a, sw ? a : b)) h |= 1u << ((a << 1) | b);
        if (h == t || h == (~t & 15u)) {
          func = f;
          swap = sw != 0;
          pol = h == t;
        }
      }
    uint32_t uw;
    if (t == 0x3 || t == 0xc) {  // U = u1 or its complement
      uw = u1;
      pol = t == 0xc;
    } else if (t == 0x5 || t == 0xa) {  // U = u2 or its complement
      uw = u2;
      pol = t == 0xa;
    } else if (func >= 0) {
      const uint32_t a = swap ? u2 : u1, b = swap ? u1 : u2;
      cx.gates.push_back({a, b, (uint8_t)func});
      uw = cx.n + (uint32_t)cx.gates.size() - 1;
      p.uniform.push_back(1);
      LutOp sop;
      sop.lut = 0;
      sop.fn = (uint8_t)func;
      sop.dst = uw;
      sop.src[0] = a;
      sop.src[1] = b;
      sop.src[2] = a;
      ops.push_back(sop);
      p.scalar_ops++;
    } else {
      ops.push_back(op);
      continue;
    }
    // g(x, U): sources (x, U, x); the third position repeats x
    LutOp nop = op;
    nop.src[0] = op.src[xi];
    nop.src[1] = uw;
    nop.src[2] = op.src[xi];
    uint8_t tt = 0;
    for (uint32_t idx = 0; idx < 8; idx++) {
      const uint32_t xv = (idx >> 2) & 1u, uv = (idx >> 1) & 1u;
      const uint32_t sel = (uv != 0) == pol ? f1 : f0;
      if ((sel >> xv) & 1u) tt |= (uint8_t)(1u << idx);
    }
    nop.fn = tt;
    ops.push_back(nop);
  }
  p.ops = std::move(ops);
}

// The output as a conjunction of literals (wire, negated): descends through And, Nor and
// Nimpl gates from a non-negated output (a & b, ~a & ~b, a & ~b); Xor gates, inputs and
// negated literals are leaves.
std::vector<std::pair<uint32_t, bool>> and_literals(const Circuit& c, uint32_t wire) {
  std::vector<std::pair<uint32_t, bool>> lits, stack{{wire, false}};
  std::vector<uint8_t> seen(2 * (size_t)c.wires(), 0);
  while (!stack.empty()) {
    auto [w, neg] = stack.back();
    stack.pop_back();
    if (seen[2 * (size_t)w + (neg ? 1 : 0)]++) continue;
    if (!neg && w >= c.n && c.gates[w - c.n].func != FUNC_XOR) {
      const Gate& g = c.gates[w - c.n];
      const bool na = g.func == FUNC_NOR, nb = g.func != FUNC_AND;
      stack.push_back({g.i1, nb});
      stack.push_back({g.i0, na});
      continue;
    }
    lits.push_back({w, neg});
  }
  return lits;
}

// The local search of grouped_circuit: groups of literal indices, clist[i] = the cone gates
// of literal i (of G gates). One descent from the given groups, then `rounds` rounds of eight
// perturbed descents from the best so far (seeded, run on threads of their own); the lowest
// total is kept (the first on ties), so the result does not depend on timing.
std::vector<std::vector<size_t>> refine_groups(std::vector<std::vector<size_t>> groups,
                                               const std::vector<std::vector<uint32_t>>& clist, uint32_t G,
                                               uint32_t cap_pct, int rounds = 2) {
  const size_t ng = groups.size();
  uint32_t cap = 0;
  {
    std::vector<uint8_t> seen(G, 0);
    for (size_t a = 0; a < ng; a++) {
      std::fill(seen.begin(), seen.end(), 0);
      uint32_t sz = 0;
      for (size_t i : groups[a])
        for (uint32_t g : clist[i])
          if (!seen[g]++) sz++;
      cap = std::max(cap, sz);
    }
    cap = (uint32_t)((uint64_t)cap * cap_pct / 100);
  }
  // descent from g after `perturb` seeded random moves; returns the total cone gates
  auto descend = [&](std::vector<std::vector<size_t>>& gr, uint64_t seed, int perturb) -> uint64_t {
    std::vector<std::vector<uint16_t>> cnt(ng, std::vector<uint16_t>(G, 0));
    std::vector<uint32_t> gsize(ng, 0);
    for (size_t a = 0; a < ng; a++)
      for (size_t i : gr[a])
        for (uint32_t g : clist[i])
          if (cnt[a][g]++ == 0) gsize[a]++;
    auto move = [&](size_t i, size_t a, size_t b) {
      for (uint32_t g : clist[i]) {
        if (--cnt[a][g] == 0) gsize[a]--;
        if (cnt[b][g]++ == 0) gsize[b]++;
      }
      auto& ga = gr[a];
      ga.erase(std::find(ga.begin(), ga.end(), i));
      gr[b].push_back(i);
    };
    uint64_t rng = seed * 0x9e3779b97f4a7c15ull + 0x632be59bd9b4e019ull;
    auto next = [&]() {
      rng ^= rng << 13;
      rng ^= rng >> 7;
      rng ^= rng << 17;
      return rng;
    };
    for (int k = 0; k < perturb; k++) {
      const size_t a = next() % ng, b = next() % ng;
      if (a == b || gr[a].size() < 2) continue;
      const size_t i = gr[a][next() % gr[a].size()];
      int64_t added = 0;
      for (uint32_t g : clist[i]) added += cnt[b][g] == 0;
      if (gsize[b] + added > cap) continue;
      move(i, a, b);
    }
    std::vector<uint8_t> mark(G, 0);
    for (int pass = 0; pass < 256; pass++) {
      bool moved = false;
      // single moves, the best target for each literal
      for (size_t a = 0; a < ng; a++)
        for (size_t pos = 0; pos < gr[a].size();) {
          const size_t i = gr[a][pos];
          if (gr[a].size() == 1) break;
          int64_t freed = 0;
          for (uint32_t g : clist[i]) freed += cnt[a][g] == 1;
          size_t to = SIZE_MAX;
          int64_t gain = 0;
          for (size_t b = 0; b < ng; b++) {
            if (b == a) continue;
            int64_t added = 0;
            for (uint32_t g : clist[i]) added += cnt[b][g] == 0;
            if (gsize[b] + added > cap || freed - added <= gain) continue;
            gain = freed - added;
            to = b;
          }
          if (to == SIZE_MAX) {
            pos++;
            continue;
          }
          move(i, a, to);
          moved = true;
        }
      if (moved) continue;
      // swaps: i in a and j in b trade groups (the first that lowers the total)
      for (size_t a = 0; a < ng && !moved; a++)
        for (size_t b = a + 1; b < ng && !moved; b++)
          for (size_t pi = 0; pi < gr[a].size() && !moved; pi++) {
            const size_t i = gr[a][pi];
            for (uint32_t g : clist[i]) mark[g] = 1;
            int64_t freed_a = 0;
            for (uint32_t g : clist[i]) freed_a += cnt[a][g] == 1;
            for (size_t pj = 0; pj < gr[b].size(); pj++) {
              const size_t j = gr[b][pj];
              // a loses i and gains j; b loses j and gains i (mark bit 0: cone i, bit 1: cone j)
              int64_t added_a = 0, freed_b = 0, added_b = 0;
              for (uint32_t g : clist[j]) {
                added_a += cnt[a][g] - (mark[g] & 1) == 0;
                freed_b += cnt[b][g] == 1;
                mark[g] |= 2;
              }
              for (uint32_t g : clist[i]) added_b += cnt[b][g] - ((mark[g] >> 1) & 1) == 0;
              for (uint32_t g : clist[j]) mark[g] &= 1;
              const int64_t na = (int64_t)gsize[a] - freed_a + added_a, nb = (int64_t)gsize[b] - freed_b + added_b;
              if (na > cap || nb > cap || na + nb >= (int64_t)gsize[a] + gsize[b]) continue;
              for (uint32_t g : clist[i]) mark[g] = 0;
              move(i, a, b);
              move(j, b, a);
              moved = true;
              break;
            }
            if (!moved)
              for (uint32_t g : clist[i]) mark[g] = 0;
          }
      if (!moved) break;
    }
    uint64_t t = 0;
    for (uint32_t v : gsize) t += v;
    return t;
  };
  uint64_t best_total = descend(groups, 0, 0);
  constexpr int kRestarts = 8;
  for (int round = 0; round < rounds; round++) {
    std::vector<std::vector<std::vector<size_t>>> tries(kRestarts, groups);
    std::vector<uint64_t> totals(kRestarts, 0);
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> err(kRestarts);
    for (int r = 0; r < kRestarts; r++)
      th.emplace_back([&, r, round]() {
        try {
          totals[r] = descend(tries[r], (uint64_t)(round * kRestarts + r + 1), 6);
        } catch (...) {
          err[r] = std::current_exception();
        }
      });
    for (auto& t : th) t.join();
    for (auto& e : err)
      if (e) std::rethrow_exception(e);
    for (int r = 0; r < kRestarts; r++)
      if (totals[r] < best_total) {
        best_total = totals[r];
        groups = tries[r];
      }
  }
  return groups;
}

// Output k of c rebuilt as K groups of its conjunction's literals (grouped by shared cone
// gates; refined by refine_groups when refine_cap_pct is non-zero): each group's cone is
// copied (gates shared by groups are recomputed, except `pin` kept values that later groups
// reuse), its literals are ANDed, and each group's value is ANDed into an accumulator right
// after it. wgroup gets the group (1..K) of every new wire, so the schedule can keep groups
// apart and only the accumulator (and kept values) live from one group into the next.
// Returns false when the output has fewer than K literals.
// Literal order of an exit-first group: starting empty, the literal whose cone adds the fewest
// gates to the group's cone so far, `count` times. union_gates[i] = the cone after i + 1.
std::vector<size_t> cheapest_union_order(const std::vector<std::vector<uint8_t>>& cone, size_t count,
                                         std::vector<uint32_t>* union_gates = nullptr) {
  const size_t G = cone.empty() ? 0 : cone[0].size();
  std::vector<uint8_t> u(G, 0), used(cone.size(), 0);
  std::vector<size_t> order;
  uint32_t have = 0;
  while (order.size() < count && order.size() < cone.size()) {
    size_t pick = SIZE_MAX;
    uint32_t pbest = 0;
    for (size_t i = 0; i < cone.size(); i++) {
      if (used[i]) continue;
      uint32_t extra = 0;
      for (size_t g = 0; g < G; g++) extra += cone[i][g] & (u[g] ^ 1);
      if (pick == SIZE_MAX || extra < pbest) {
        pick = i;
        pbest = extra;
      }
    }
    used[pick] = 1;
    order.push_back(pick);
    for (size_t g = 0; g < G; g++) u[g] |= cone[pick][g];
    have += pbest;
    if (union_gates) union_gates->push_back(have);
  }
  return order;
}

std::vector<std::vector<uint8_t>> literal_cones(const Circuit& c, const std::vector<std::pair<uint32_t, bool>>& lits) {
  const uint32_t G = (uint32_t)c.gates.size();
  std::vector<std::vector<uint8_t>> cone(lits.size());
  for (size_t i = 0; i < lits.size(); i++) {
    cone[i].assign(G, 0);
    std::vector<uint32_t> st{lits[i].first};
    while (!st.empty()) {
      const uint32_t w = st.back();
      st.pop_back();
      if (w < c.n || cone[i][w - c.n]) continue;
      cone[i][w - c.n] = 1;
      st.push_back(c.gates[w - c.n].i0);
      st.push_back(c.gates[w - c.n].i1);
    }
  }
  return cone;
}

// The chance that a wave of the SAT search keeps an assignment standing after a first group
// of the first L literals of `order`, for every L (survive[L]), measured on `samples` waves at
// random places of the 2^n space: each wave's 2,048 assignments (words wu .. wu + 63, wu a
// multiple of 64) evaluated on the host 64 at a time, bit-sliced.
std::vector<double> sampled_survival(const Circuit& c, const std::vector<std::pair<uint32_t, bool>>& lits,
                                     const std::vector<size_t>& order, uint32_t samples, uint64_t seed) {
  const uint32_t n = c.n;
  std::vector<uint8_t> need(c.wires(), 0);
  for (const auto& l : lits) need[l.first] = 1;
  for (size_t j = c.gates.size(); j-- > 0;)
    if (need[n + j]) need[c.gates[j].i0] = need[c.gates[j].i1] = 1;
  std::vector<uint32_t> cone_gates;
  for (size_t j = 0; j < c.gates.size(); j++)
    if (need[n + j]) cone_gates.push_back((uint32_t)j);
  static const uint64_t kPat[6] = {0xAAAAAAAAAAAAAAAAull, 0xCCCCCCCCCCCCCCCCull, 0xF0F0F0F0F0F0F0F0ull,
                                   0xFF00FF00FF00FF00ull, 0xFFFF0000FFFF0000ull, 0xFFFFFFFF00000000ull};
  const uint64_t words = n >= 64 ? (1ull << 59) : (1ull << (n - 5));
  uint64_t st = seed;
  auto next = [&]() {  // SplitMix64
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  std::vector<uint64_t> v(c.wires());
  std::vector<uint32_t> alive(order.size() + 1, 0);
  for (uint32_t s = 0; s < samples; s++) {
    const uint64_t wu = (next() % words) & ~63ull;
    std::vector<uint8_t> hit(order.size() + 1, 0);
    for (uint32_t j = 0; j < 32; j++) {  // words wu + 2j, wu + 2j + 1: assignments base .. base + 63
      const uint64_t base = (wu + 2 * j) << 5;
      for (uint32_t k = 0; k < n; k++) v[k] = k < 6 ? kPat[k] : (((base >> k) & 1) ? ~0ull : 0ull);
      for (uint32_t g : cone_gates) {
        const uint64_t a = v[c.gates[g].i0], b = v[c.gates[g].i1];
        uint64_t r;
        switch (c.gates[g].func) {
          case FUNC_AND: r = a & b; break;
          case FUNC_NOR: r = ~(a | b); break;
          case FUNC_NIMPL: r = a & ~b; break;
          default: r = a ^ b; break;
        }
        v[n + g] = r;
      }
      uint64_t acc = ~0ull;
      hit[0] = 1;
      for (size_t L = 1; L <= order.size(); L++) {
        const auto& l = lits[order[L - 1]];
        acc &= l.second ? ~v[l.first] : v[l.first];
        if (acc) hit[L] = 1;
      }
    }
    for (size_t L = 0; L <= order.size(); L++) alive[L] += hit[L];
  }
  std::vector<double> out(order.size() + 1);
  for (size_t L = 0; L <= order.size(); L++) out[L] = (double)alive[L] / samples;
  return out;
}

bool grouped_circuit(const Circuit& c, uint32_t out_k, uint32_t K, Circuit& out, std::vector<uint32_t>& wgroup,
                     uint32_t refine_cap_pct, uint32_t pin = 0, const std::vector<uint8_t>* uni_in = nullptr,
                     uint32_t pin_u = 0, std::vector<std::pair<uint32_t, bool>>* accs = nullptr,
                     uint32_t first_lits = 0) {
  const Output o = c.outputs.at(out_k);
  const auto lits = and_literals(c, o.wire);
  if (K < 2 || lits.size() < K) return false;
  if (first_lits && lits.size() < first_lits + K - 1) return false;
  const uint32_t G = (uint32_t)c.gates.size();
  // the greedy grouping takes ~L^2 G steps and the cone marks L G bytes (L literals, G gates):
  // outputs past these bounds are not grouped
  const uint64_t L = lits.size();
  if (L * G > (1ull << 27) || L * L * G > (1ull << 32)) return false;
  // cone of each literal (gate marks)
  std::vector<std::vector<uint8_t>> cone(lits.size());
  for (size_t i = 0; i < lits.size(); i++) {
    cone[i].assign(G, 0);
    std::vector<uint32_t> st{lits[i].first};
    while (!st.empty()) {
      uint32_t w = st.back();
      st.pop_back();
      if (w < c.n || cone[i][w - c.n]) continue;
      cone[i][w - c.n] = 1;
      st.push_back(c.gates[w - c.n].i0);
      st.push_back(c.gates[w - c.n].i1);
    }
  }
  auto csize = [&](size_t i) {
    uint32_t k = 0;
    for (uint8_t b : cone[i]) k += b;
    return k;
  };
  std::vector<uint8_t> used(lits.size(), 0);
  std::vector<std::vector<size_t>> groups;
  size_t left = lits.size(), rest_groups = K;
  if (first_lits) {
    // Exit-first (for group exits): the first group takes the first_lits literals whose cones
    // add the fewest gates to it, one at a time; later groups run only in the waves where
    // the first one left some assignment standing, so their size matters little.
    std::vector<size_t> grp = cheapest_union_order(cone, first_lits);
    for (size_t i : grp) used[i] = 1;
    left -= grp.size();
    groups.push_back(std::move(grp));
    rest_groups = K - 1;
  }
  // greedy groups of ceil(m / K): the largest remaining cone, then the literals sharing most
  const size_t per = (left + rest_groups - 1) / rest_groups;
  for (; left;) {
    size_t first = SIZE_MAX;
    uint32_t best = 0;
    for (size_t i = 0; i < lits.size(); i++)
      if (!used[i] && (first == SIZE_MAX || csize(i) > best)) {
        first = i;
        best = csize(i);
      }
    std::vector<size_t> grp{first};
    used[first] = 1;
    left--;
    std::vector<uint8_t> u = cone[first];
    while (grp.size() < per && left) {
      size_t pick = SIZE_MAX;
      double pbest = 0;
      for (size_t i = 0; i < lits.size(); i++) {
        if (used[i]) continue;
        uint32_t shared = 0, extra = 0;
        for (uint32_t g = 0; g < G; g++) {
          shared += cone[i][g] & u[g];
          extra += cone[i][g] & (u[g] ^ 1);
        }
        const double sc = shared - 0.5 * extra;
        if (pick == SIZE_MAX || sc > pbest) {
          pick = i;
          pbest = sc;
        }
      }
      grp.push_back(pick);
      used[pick] = 1;
      left--;
      for (uint32_t g = 0; g < G; g++) u[g] |= cone[pick][g];
    }
    groups.push_back(std::move(grp));
  }
  if (refine_cap_pct) {
    // Local search from the greedy groups over the total of the groups' cone gates (the
    // gates evaluated), no group growing past refine_cap_pct percent of the largest greedy
    // group (a proxy for the registers it needs): single literals move to another group and
    // pairs of literals swap groups while the total drops; then a few seeded perturbations
    // (random moves) are searched again and the lowest total is kept. On the 36-input,
    // 20,000-gate search the 5 groups go from 3,966 cone gates to about 2,800.
    std::vector<std::vector<uint32_t>> clist(lits.size());
    for (size_t i = 0; i < lits.size(); i++)
      for (uint32_t g = 0; g < G; g++)
        if (cone[i][g]) clist[i].push_back(g);
    if (first_lits) {  // the first group stays as chosen
      std::vector<std::vector<size_t>> tail(groups.begin() + 1, groups.end());
      // the later groups run only where a wave survives the first: one descent, no perturbed
      // rounds (54-input factoring search: code generation 0.37 -> 0.20 s, the same body)
      if (tail.size() >= 2) tail = refine_groups(tail, clist, G, refine_cap_pct, 0);
      tail.insert(tail.begin(), groups[0]);
      groups = std::move(tail);
    } else {
      groups = refine_groups(groups, clist, G, refine_cap_pct);
    }
  }
  out = Circuit();
  out.n = c.n;
  wgroup.assign(c.n, 0);
  auto add = [&](uint32_t a, uint32_t b, uint8_t f, uint32_t grp) {
    out.gates.push_back({a, b, f});
    wgroup.push_back(grp);
    return c.n + (uint32_t)out.gates.size() - 1;
  };
  // a & b for literals (wire, negated), as one positive wire
  auto conj = [&](std::pair<uint32_t, bool> x, std::pair<uint32_t, bool> y, uint32_t grp) {
    if (!x.second && !y.second) return add(x.first, y.first, FUNC_AND, grp);
    if (!x.second) return add(x.first, y.first, FUNC_NIMPL, grp);
    if (!y.second) return add(y.first, x.first, FUNC_NIMPL, grp);
    return add(x.first, y.first, FUNC_NOR, grp);
  };
  // Kept values: the `pin` gates that the most groups share, weighted by their own cone, are
  // computed once (by the first group that needs them) and stay live into the later groups,
  // whose cones stop there instead of recomputing them.
  std::vector<uint8_t> pinned(G, 0), have(G, 0);
  if (pin) {
    std::vector<uint32_t> use(G, 0);
    for (size_t gi = 0; gi < groups.size(); gi++) {
      std::vector<uint8_t> u(G, 0);
      for (size_t i : groups[gi])
        for (uint32_t g = 0; g < G; g++) u[g] |= cone[i][g];
      for (uint32_t g = 0; g < G; g++) use[g] += u[g];
    }
    std::vector<std::pair<uint64_t, uint32_t>> cand;
    std::vector<uint32_t> mark(G, UINT32_MAX);
    for (uint32_t g = 0; g < G; g++) {
      if (use[g] < 2) continue;
      uint64_t sz = 0;
      std::vector<uint32_t> st{c.n + g};
      while (!st.empty()) {
        const uint32_t w = st.back();
        st.pop_back();
        if (w < c.n || mark[w - c.n] == g) continue;
        mark[w - c.n] = g;
        sz++;
        st.push_back(c.gates[w - c.n].i0);
        st.push_back(c.gates[w - c.n].i1);
      }
      cand.push_back({(uint64_t)(use[g] - 1) * sz, g});
    }
    std::sort(cand.begin(), cand.end(), [](const auto& a, const auto& b) { return a.first > b.first || (a.first == b.first && a.second < b.second); });
    for (size_t q = 0; q < cand.size() && q < pin; q++) pinned[cand[q].second] = 1;
  }
  if (pin_u && uni_in) {
    // Kept uniform values: wave-uniform gates (over wave-uniform inputs only) that several
    // groups share are computed once as well. They live in SGPRs (or lane slots), not in
    // the VGPRs that bound the groups, and each saves its scalar cone in every later group.
    std::vector<uint8_t> ug(G, 0);
    auto is_u = [&](uint32_t w) { return w < c.n ? (*uni_in)[w] != 0 : ug[w - c.n] != 0; };
    for (uint32_t g = 0; g < G; g++) ug[g] = is_u(c.gates[g].i0) && is_u(c.gates[g].i1);
    std::vector<uint32_t> use(G, 0);
    for (size_t gi = 0; gi < groups.size(); gi++) {
      std::vector<uint8_t> u(G, 0);
      for (size_t i : groups[gi])
        for (uint32_t g = 0; g < G; g++) u[g] |= cone[i][g];
      for (uint32_t g = 0; g < G; g++) use[g] += u[g];
    }
    std::vector<std::pair<uint64_t, uint32_t>> cand;
    std::vector<uint32_t> mark(G, UINT32_MAX);
    for (uint32_t g = 0; g < G; g++) {
      if (!ug[g] || use[g] < 2 || pinned[g]) continue;
      uint64_t sz = 0;
      std::vector<uint32_t> st{c.n + g};
      while (!st.empty()) {
        const uint32_t w = st.back();
        st.pop_back();
        if (w < c.n || mark[w - c.n] == g) continue;
        mark[w - c.n] = g;
        sz++;
        st.push_back(c.gates[w - c.n].i0);
        st.push_back(c.gates[w - c.n].i1);
      }
      if (sz >= 2) cand.push_back({(uint64_t)(use[g] - 1) * sz, g});
    }
    std::sort(cand.begin(), cand.end(), [](const auto& a, const auto& b) { return a.first > b.first || (a.first == b.first && a.second < b.second); });
    for (size_t q = 0; q < cand.size() && q < pin_u; q++) pinned[cand[q].second] = 1;
  }
  std::pair<uint32_t, bool> acc{0, false};
  bool have_acc = false;
  std::vector<uint32_t> remap(c.wires());
  for (uint32_t k = 0; k < c.n; k++) remap[k] = k;
  for (size_t gi = 0; gi < groups.size(); gi++) {
    const uint32_t grp = (uint32_t)gi + 1;
    // the group's gates: its literals' cones, stopping at kept values already computed
    std::vector<uint8_t> u(G, 0);
    std::vector<uint32_t> st;
    for (size_t i : groups[gi]) st.push_back(lits[i].first);
    while (!st.empty()) {
      const uint32_t w = st.back();
      st.pop_back();
      if (w < c.n) continue;
      const uint32_t g = w - c.n;
      if (u[g] || (pinned[g] && have[g])) continue;
      u[g] = 1;
      st.push_back(c.gates[g].i0);
      st.push_back(c.gates[g].i1);
    }
    for (uint32_t g = 0; g < G; g++)
      if (u[g]) {
        remap[c.n + g] = add(remap[c.gates[g].i0], remap[c.gates[g].i1], c.gates[g].func, grp);
        if (pinned[g]) have[g] = 1;
      }
    std::pair<uint32_t, bool> v{remap[lits[groups[gi][0]].first], lits[groups[gi][0]].second};
    for (size_t q = 1; q < groups[gi].size(); q++)
      v = {conj(v, {remap[lits[groups[gi][q]].first], lits[groups[gi][q]].second}, grp), false};
    acc = have_acc ? std::pair<uint32_t, bool>{conj(acc, v, grp), false} : v;
    have_acc = true;
    if (accs && gi + 1 < groups.size()) accs->push_back(acc);
  }
  out.outputs.push_back({acc.first, acc.second != o.neg});
  return true;
}

// wave-uniform gates that several output groups share and that are computed once (per word)
// in the groups tried from 2^36 assignments: on the 36-input, 20,000-gate search 16 of them
// take 2,296 -> 2,253 issue slots per word, 4.96 -> 4.83 ms per launch (more of them no
// longer fit the SGPRs and go to lane slots, whose v_readlane costs more than recomputing)
constexpr uint32_t kKeptUniform = 16;

// Weight of the issue slots after a group exit: a wave that keeps an assignment standing runs
// on while the waves around it leave, so at less than two waves per SIMD. Fitted on the
// 36-input, 20,000-gate search: 13 first literals took 22.2 ms, 16 took 19.7 ms, where equal
// weights ranked 13 first; weighted 2, the 54-input factoring search's best size (20 of 56:
// 164 ms, against 296 ms with 16 and 471 ms with 13) is ranked first too.
constexpr double kSurvivorWeight = 2.0;

const char* kind_name(NativeKind k) {
  switch (k) {
    case NativeKind::SAT: return "sat";
    case NativeKind::QSAT_A: return "qsata";
    case NativeKind::QSAT_B: return "qsatb";
    case NativeKind::BATCH: return "batch";
    default: return "table";
  }
}

}  // namespace

// One iteration of the walk templates (ns_nat_walk<T>) as a circuit: two walk steps with
// their per-slot bookkeeping. Inputs: s[0..n) the slots' states, f[0..n) their final states,
// v[0..n) their saved states, then active, fin, fstop, fwp, floop, a0..a3; outputs: the new
// s, f and masks in the same order. Per step, from state x to y = next(x):
//   stop = STOP(x) & active, wp = (y's low k bits all zero) & active & ~stop,
//   loop (second step only) = (y == v) & active & ~stop & ~wp, done = stop | wp | loop,
//   f = done ? y : f, the ages a3..a0 (4-bit counters) of the slots in fin grow by one,
//   fin |= done, fstop |= stop, fwp |= wp, floop |= loop, active &= ~done.
Circuit walk_iteration(const Circuit& c, uint32_t k) {
  const uint32_t n = c.n;
  Circuit w;
  w.n = 3 * n + 9;
  auto add = [&](uint8_t fn, uint32_t a, uint32_t b) {
    w.gates.push_back({a, b, fn});
    return w.n + (uint32_t)w.gates.size() - 1;
  };
  auto AND = [&](uint32_t a, uint32_t b) { return add(FUNC_AND, a, b); };
  auto NOR = [&](uint32_t a, uint32_t b) { return add(FUNC_NOR, a, b); };
  auto NIMPL = [&](uint32_t a, uint32_t b) { return add(FUNC_NIMPL, a, b); };
  auto XOR = [&](uint32_t a, uint32_t b) { return add(FUNC_XOR, a, b); };
  auto OR = [&](uint32_t a, uint32_t b) { return XOR(a, NIMPL(b, a)); };
  // AND over ~x for the wires xs (all zero), as nor pairs
  auto all_zero = [&](const std::vector<uint32_t>& xs) {
    uint32_t acc = UINT32_MAX;
    for (size_t i = 0; i < xs.size(); i += 2) {
      const uint32_t t = NOR(xs[i], i + 1 < xs.size() ? xs[i + 1] : xs[i]);
      acc = acc == UINT32_MAX ? t : AND(acc, t);
    }
    return acc;
  };
  std::vector<uint32_t> S(n), F(n), V(n);
  for (uint32_t i = 0; i < n; i++) {
    S[i] = i;
    F[i] = n + i;
    V[i] = 2 * n + i;
  }
  uint32_t active = 3 * n, fin = active + 1, fstop = active + 2, fwp = active + 3, floop = active + 4;
  uint32_t a[4] = {active + 5, active + 6, active + 7, active + 8};
  std::vector<uint32_t> map(c.wires());
  for (int half = 0; half < 2; half++) {
    for (uint32_t i = 0; i < n; i++) map[i] = S[i];
    for (size_t g = 0; g < c.gates.size(); g++)
      map[n + g] = add(c.gates[g].func, map[c.gates[g].i0], map[c.gates[g].i1]);
    std::vector<uint32_t> Y(n + 1);
    for (uint32_t i = 0; i <= n; i++) {
      const uint32_t x = map[c.outputs[i].wire];
      Y[i] = c.outputs[i].neg ? NOR(x, x) : x;
    }
    const uint32_t stop = AND(Y[n], active);
    const uint32_t ok = NIMPL(active, stop);  // active & ~stop
    const uint32_t wp = k ? AND(all_zero(std::vector<uint32_t>(Y.begin(), Y.begin() + k)), ok) : ok;
    uint32_t done = OR(stop, wp);
    if (half == 1) {
      std::vector<uint32_t> e(n);
      for (uint32_t i = 0; i < n; i++) e[i] = XOR(Y[i], V[i]);
      const uint32_t loop = AND(all_zero(e), NIMPL(ok, wp));
      done = OR(done, loop);
      floop = OR(floop, loop);
    }
    for (uint32_t i = 0; i < n; i++) F[i] = XOR(F[i], AND(done, XOR(Y[i], F[i])));
    a[0] = XOR(a[0], fin);
    uint32_t cy = NIMPL(fin, a[0]);
    a[1] = XOR(a[1], cy);
    cy = NIMPL(cy, a[1]);
    a[2] = XOR(a[2], cy);
    cy = NIMPL(cy, a[2]);
    a[3] = XOR(a[3], cy);
    fin = OR(fin, done);
    fstop = OR(fstop, stop);
    fwp = OR(fwp, wp);
    active = NIMPL(active, done);
    for (uint32_t i = 0; i < n; i++) S[i] = Y[i];
  }
  for (uint32_t i = 0; i < n; i++) w.outputs.push_back({S[i], false});
  for (uint32_t i = 0; i < n; i++) w.outputs.push_back({F[i], false});
  for (uint32_t x : {active, fin, fstop, fwp, floop, a[0], a[1], a[2], a[3]}) w.outputs.push_back({x, false});
  return w;
}

NativeKernel build_native(const Circuit& c, NativeKind kind, const std::vector<uint32_t>& vbit,
                          uint32_t lane_shift, const std::vector<uint32_t>& outs, bool group_exits) {
  NativeKernel nk;
  auto cone = output_cone(c, outs);
  auto order = schedule_gates(c, cone, outs);
  nk.info.cone_gates = (uint32_t)order.size();
  nk.info.peak_live = peak_live(c, order, outs);
  std::vector<uint8_t> uni(c.n, 1);
  for (uint32_t k = 0; k < c.n; k++)
    if (vbit[k] >= 5 && vbit[k] - 5 >= lane_shift && vbit[k] - 5 < lane_shift + 6) uni[k] = 0;
  const bool table = kind == NativeKind::TABLE;
  const uint32_t lm = table ? 32 : 1;
  const uint32_t pool_lo = table ? 40 : 8;
  struct Cls {
    uint32_t budget, hi, agprs, slots;
  };
  // each register class first without, then with two VGPRs of lane slots for uniform values
  // (the QSAT-B templates keep two more VGPRs for their own code: gen_templates.py QSATB_POOL_CUT)
  // (the SAT templates, whose bodies may write vcc, end the r64 / r128 pools one register lower:
  // gen_templates.py SAT_VCC_POOL_CUT)
  const uint32_t cut = kind == NativeKind::QSAT_B ? 2 : 0;
  const uint32_t cs = cut + (kind == NativeKind::SAT ? 1 : 0);  // r64 / r128 only
  const Cls classes[8] = {{64, 55 - cs, 0, 0},      {64, 55 - cs, 0, 2},     {128, 119 - cs, 0, 0},
                          {128, 119 - cs, 0, 2},    {256, 239 - cut, 0, 0},  {256, 239 - cut, 0, 2},
                          {512, 239 - cut, 256, 0}, {512, 239 - cut, 256, 2}};
  struct Attempt {
    std::string body, why = "no register class tried";
    uint32_t vgpr_class = 0, spills = 0, reloads = 0, valu = 0, scalar = 0, vector_lines = 0, scalar_lines = 0,
             lane_reads = 0;
    // issue slots per word: a wave issues one instruction at a time, scalar or vector, and a
    // v_readlane holds its wave for about two
    uint32_t slots() const { return vector_lines + scalar_lines + lane_reads; }
  };
  // LUT mapping, schedule and the smallest register class that holds the program of cc
  auto attempt = [&](const Circuit& cc, const std::vector<uint32_t>& outs_cc, const std::vector<uint32_t>* wgroup,
                     const std::vector<std::pair<uint32_t, bool>>* accs = nullptr,
                     const std::vector<uint32_t>* order_in = nullptr) {
    Attempt at;
    std::vector<uint32_t> out_wires;
    std::vector<uint8_t> out_neg;
    for (uint32_t k : outs_cc) {
      out_wires.push_back(cc.outputs.at(k).wire);
      out_neg.push_back(cc.outputs.at(k).neg ? 1 : 0);
    }
    const std::vector<uint32_t> order_cc = order_in ? *order_in : schedule_gates(cc, output_cone(cc, outs_cc), outs_cc);
    Circuit cx = cc;  // with the scalar gates of split_uniform_pairs appended
    // group exits test the groups' running conjunctions: those are materialised values. Only
    // for an output that is the conjunction itself: a negated one is 1 where it is 0.
    const bool exits = group_exits && kind == NativeKind::SAT && accs && !accs->empty() && !c.outputs.at(outs[0]).neg;
    std::vector<uint32_t> mapped = out_wires;
    if (exits)
      for (const auto& a : *accs) mapped.push_back(a.first);
    LutProgram p0 = map_luts(cc, order_cc, mapped, uni);
    split_uniform_pairs(cx, p0);
    LutProgram p = pressure_order(p0, out_wires, cx.wires(), wgroup);
    at.valu = p.valu_ops;
    at.scalar = p.scalar_ops;
    // lane values live at once in this order: classes whose pool cannot hold them and that
    // have no AGPRs to spill to are skipped
    uint32_t peak = 0;
    {
      std::vector<uint32_t> rem(cx.wires(), 0);
      for (const LutOp& op : p.ops)
        for (int k = 0; k < (op.lut ? 3 : 2); k++) {
          bool dup = false;
          for (int q = 0; q < k; q++) dup |= op.src[q] == op.src[k];
          if (!dup) rem[op.src[k]]++;
        }
      for (uint32_t w : out_wires) rem[w]++;
      uint32_t live = 0;
      for (const LutOp& op : p.ops) {
        for (int k = 0; k < (op.lut ? 3 : 2); k++) {
          uint32_t w = op.src[k];
          bool dup = false;
          for (int q = 0; q < k; q++) dup |= op.src[q] == w;
          if (!dup && --rem[w] == 0 && w >= cx.n && !p.uniform[w]) live--;
        }
        if (rem[op.dst] > 0 && !p.uniform[op.dst]) live++;
        peak = std::max(peak, live);
      }
    }
    for (const Cls& cl : classes) {
      if (table && cl.budget < 128) continue;
      if (!cl.agprs && peak + cl.slots > cl.hi - pool_lo + 1) continue;
      try {
        Allocator a(cx, p, vbit, lane_shift, lm, pool_lo, cl.hi, cl.agprs, cl.slots, out_wires, out_neg,
                    kind == NativeKind::SAT ? kSgprLoSat : kind == NativeKind::BATCH ? kSgprLoBatch : kSgprLo,
                    nullptr, nullptr, exits ? accs : nullptr, kind == NativeKind::SAT && !exits);
        at.body = a.run();
        at.vgpr_class = cl.budget;
        at.spills = a.spills();
        at.reloads = a.reloads();
        break;
      } catch (const NativeUnfit& e) {
        at.why = e.what();
        continue;
      }
    }
    // vector instructions of the body (LUTs, copies, AGPR and lane moves)
    for (size_t pos = 0; (pos = at.body.find("\tv_", pos)) != std::string::npos; pos += 3) at.vector_lines++;
    for (size_t pos = 0; (pos = at.body.find("\ts_", pos)) != std::string::npos; pos += 3)
      if (at.body.compare(pos, 7, "\ts_nop ") != 0) at.scalar_lines++;
    for (size_t pos = 0; (pos = at.body.find("\tv_readlane", pos)) != std::string::npos; pos += 3) at.lane_reads++;
    return at;
  };
  Attempt best = attempt(c, outs, nullptr, nullptr, &order);  // the schedule computed above
  uint32_t groups = 1;
  // One wave per SIMD (the AGPR class) issues fewer VALU ops per cycle than two or more. A
  // single output that is a conjunction can be evaluated as K groups of its literals, one
  // after the other, recomputing gates the groups share. Programs are compared by issue
  // slots (vector plus scalar instructions, a v_readlane counted twice): a wave issues one
  // instruction at a time, and at two waves per SIMD a scalar op costs about as much as a
  // LUT (the issue micro-benchmark). Measured on the 36-input, 20,000-gate search, two waves
  // per SIMD ran 1.45 times the instructions per cycle of one (not twice: lane moves and
  // scalar work grow with the groups), so the groups are taken when they reach a two-wave
  // class at under 1.45 times the slots.
  // Each try costs about one more code generation (tens of ms for large bodies), so they are
  // tried for spaces of 2^34 assignments and more only, where a search takes seconds.
  // NS_OUTPUT_GROUPS=K forces K (1: never).
  long forced = -1;
  if (const char* e = std::getenv("NS_OUTPUT_GROUPS")) forced = std::atol(e);
  // Groups are tried when the single-group program needs one wave per SIMD (from 2^34
  // assignments) or fits no class at all (any size: the alternative is the compiler path).
  const bool unfit = best.vgpr_class == 0;
  bool exit_first = false;
  if (group_exits && kind == NativeKind::SAT && outs.size() == 1 && !c.outputs.at(outs[0]).neg && forced < 0 &&
      (unfit || c.n >= 34)) {
    // Group exits: the first group takes the literals whose cones are cheapest together, so
    // that most waves leave after it. Expected issue slots per word = the slots before the
    // first exit + the chance that a wave keeps an assignment standing after the first group
    // x the slots after (weighted: kSurvivorWeight). That chance is measured on sampled
    // waves (sampled_survival): the literals of a real circuit are far from independent coins
    // (the low product bits of a multiplier fix the low bits of one factor given the other).
    // First-group sizes are ranked by first-group cone gates + weighted chance x all cone
    // gates, and the three sizes around the best are allocated with 2 to 6 groups.
    struct Try {
      uint32_t L, K;
    };
    std::vector<Try> tries;
    std::vector<double> survive;
    {
      const auto lits = and_literals(c, c.outputs.at(outs[0]).wire);
      // the same bounds as grouped_circuit's: the cone marks take L G bytes, the order ~L^2 G steps
      const uint64_t nl = lits.size(), ng = c.gates.size();
      if (nl >= 3 && nl * ng <= (1ull << 27) && nl * nl * ng <= (1ull << 32)) {
        const auto lc = literal_cones(c, lits);
        std::vector<uint32_t> ug;
        const auto order = cheapest_union_order(lc, std::min<size_t>(lits.size() - 1, 48), &ug);
        survive = sampled_survival(c, lits, order, 512, 0x5eedull ^ c.gates.size());
        std::vector<uint8_t> all(c.gates.size(), 0);
        for (const auto& m : lc)
          for (size_t g = 0; g < m.size(); g++) all[g] |= m[g];
        double total = 0;
        for (uint8_t b : all) total += b;
        size_t bestL = 1;
        double bp = 0;
        for (size_t L = 1; L <= order.size(); L++) {
          const double proxy = ug[L - 1] + kSurvivorWeight * survive[L] * total;
          if (L == 1 || proxy < bp) {
            bestL = L;
            bp = proxy;
          }
        }
        for (size_t L : {bestL > 2 ? bestL - 2 : 1, bestL, bestL + 2})
          if (L >= 1 && L <= order.size())
            for (uint32_t K : {2u, 3u, 4u, 5u, 6u})
              if (L + K - 1 <= lits.size()) tries.push_back({(uint32_t)L, K});
      }
    }
    std::vector<Attempt> ats(tries.size());
    std::vector<std::exception_ptr> err(tries.size());
    std::vector<std::thread> th;
    for (size_t q = 0; q < tries.size(); q++)
      th.emplace_back([&, q]() {
        try {
          Circuit g;
          std::vector<uint32_t> wg;
          std::vector<std::pair<uint32_t, bool>> accs;
          if (grouped_circuit(c, outs[0], tries[q].K, g, wg, 100, 16, &uni, kKeptUniform, &accs, tries[q].L))
            ats[q] = attempt(g, {0}, &wg, &accs);
        } catch (...) {
          err[q] = std::current_exception();
        }
      });
    for (auto& t : th) t.join();
    for (auto& e : err)
      if (e) std::rethrow_exception(e);
    // against the program without groups (one wave per SIMD in the AGPR class: 1.45 times
    // the slots, as below)
    double best_cost = best.vgpr_class ? (double)best.slots() * (best.vgpr_class == 512 ? 1.45 : 1.0) : 0.0;
    int pick = -1;
    for (size_t q = 0; q < tries.size(); q++) {
      const Attempt& at = ats[q];
      if (!at.vgpr_class || at.vgpr_class > 256) continue;
      const size_t cut = at.body.find("s_cbranch_vccz");
      if (cut == std::string::npos) continue;
      uint32_t first = 0;
      for (size_t pos = 0; (pos = at.body.find('\t', pos)) < cut; pos++)
        if (at.body.compare(pos, 7, "\ts_nop ") != 0) first++;
      const double cost = first + kSurvivorWeight * survive[tries[q].L] * ((double)at.slots() - first);
      if ((pick < 0 && !best.vgpr_class) || cost < best_cost) {
        pick = (int)q;
        best_cost = cost;
      }
    }
    if (pick >= 0) {
      best = std::move(ats[(size_t)pick]);
      groups = tries[(size_t)pick].K;
      exit_first = true;
    }
  }
  if (!exit_first && !table && outs.size() == 1 && forced != 1 &&
      (forced > 1 || unfit || (best.vgpr_class == 512 && c.n >= 34))) {
    const uint32_t ks[9] = {2, 3, 4, 5, 6, 8, 10, 12, 16};
    // For each K: the refined groups (fewer gates), then the greedy ones (smaller groups).
    // Once refined groups reach two waves, looser caps are tried too (from 2^36
    // assignments, where a search takes a tenth of a second or more): the first K that
    // reaches two waves keeps its cheapest fitting program. A program that fits no class
    // takes the first grouped one that does, and the search goes on for two waves.
    auto better = [&](const Attempt& at) {
      if (!at.vgpr_class) return false;
      if (!best.vgpr_class) return true;
      return at.vgpr_class <= 256 && (best.vgpr_class == 512 ? at.slots() * 100 < best.slots() * 145
                                                            : at.slots() < best.slots());
    };
    for (uint32_t K : ks) {
      if (forced > 1 && (uint32_t)forced != K) continue;
      bool fits = false;  // a two-wave program (or, forced, any program) at this K
      for (uint32_t cap_pct : {100u, 0u}) {
        Circuit g;
        std::vector<uint32_t> wg;
        std::vector<std::pair<uint32_t, bool>> accs;
        if (!grouped_circuit(c, outs[0], K, g, wg, cap_pct, 0, nullptr, 0, &accs)) break;
        Attempt at = attempt(g, {0}, &wg, &accs);
        if (forced > 1 && at.vgpr_class && (cap_pct || at.vgpr_class <= 256) && !fits) {
          best = std::move(at);
          groups = K;
          fits = true;
          continue;
        }
        if (better(at)) {
          fits = at.vgpr_class <= 256;
          best = std::move(at);
          groups = K;
        }
        if (fits) break;
      }
      if (fits && c.n >= 36 && forced <= 1) {
        // looser caps, each with and without 16 kept lane values (pin), all with 16 kept
        // uniform values, on threads of their own, compared in list order
        const uint32_t caps[6] = {102, 103, 104, 106, 108, 110};
        std::vector<std::pair<uint32_t, uint32_t>> tries{{100, 16}};
        for (uint32_t cp : caps) {
          tries.push_back({cp, 0});
          tries.push_back({cp, 16});
        }
        std::vector<Attempt> ats(tries.size());
        std::vector<std::exception_ptr> err(tries.size());
        std::vector<std::thread> th;
        for (size_t q = 0; q < tries.size(); q++)
          th.emplace_back([&, q]() {
            try {
              Circuit g;
              std::vector<uint32_t> wg;
              std::vector<std::pair<uint32_t, bool>> accs;
              grouped_circuit(c, outs[0], K, g, wg, tries[q].first, tries[q].second, &uni, kKeptUniform, &accs);
              ats[q] = attempt(g, {0}, &wg, &accs);
            } catch (...) {
              err[q] = std::current_exception();
            }
          });
        for (auto& t : th) t.join();
        for (auto& e : err)
          if (e) std::rethrow_exception(e);
        for (Attempt& at : ats)
          if (better(at)) best = std::move(at);
      }
      if (fits) break;  // more groups only add work
    }
  }
  nk.info.valu_ops = best.valu;
  nk.info.scalar_ops = best.scalar;
  nk.vgpr_class = best.vgpr_class;
  nk.spills = best.spills;
  nk.reloads = best.reloads;
  nk.output_groups = groups;
  std::string body = std::move(best.body), why = best.why;
  if (!nk.vgpr_class) throw NativeUnfit("circuit does not fit the native templates: " + why);
  // the table kernel packs all 32 output registers: those without an output are zero
  if (table)
    for (size_t k = outs.size(); k < 32; k++) body += "\tv_mov_b32 v" + std::to_string(k) + ", 0\n";
  nk.kernel_name = std::string("ns_nat_") + kind_name(kind) + "_r" + std::to_string(nk.vgpr_class);
  if (kind == NativeKind::BATCH) {  // called from a batch code object: no module text of its own
    nk.body = std::move(body);
    return nk;
  }
  const char* tpl = native_template(nk.kernel_name);
  if (!tpl) throw std::runtime_error("missing native template " + nk.kernel_name);
  std::string text(tpl);
  std::string marker = "; NS_BODY " + nk.kernel_name;
  size_t pos = text.find(marker);
  if (pos == std::string::npos) throw std::runtime_error("template placeholder not found");
  size_t eol = text.find('\n', pos);
  text.replace(pos, eol - pos, "; circuit body\n" + body);
  nk.asm_text = std::move(text);
  nk.body = std::move(body);
  return nk;
}

NativeKernel build_native_walk(const Circuit& c, uint32_t k) {
  const uint32_t n = c.n;
  if (c.outputs.size() != (size_t)n + 1) throw NativeUnfit("a walk circuit has n inputs and n+1 outputs");
  if (n == 0 || k > n) throw NativeUnfit("bad walk circuit or waypoint bits");
  uint32_t T = 0;
  for (uint32_t t : kWalkTiers)
    if (n <= t) {
      T = t;
      break;
    }
  if (!T) throw NativeUnfit("walk templates hold at most 48 state bits");
  NativeKernel nk;
  {
    auto cone = output_cone(c);
    nk.info.cone_gates = (uint32_t)schedule_gates(c, cone).size();
  }
  const Circuit w = walk_iteration(c, k);
  std::vector<uint32_t> outs(w.outputs.size()), out_wires, out_regs;
  std::vector<uint8_t> out_neg;
  for (uint32_t i = 0; i < outs.size(); i++) {
    outs[i] = i;
    out_wires.push_back(w.outputs[i].wire);
    out_neg.push_back(0);
  }
  // template registers (gen_templates.py walk_regs): s at v8, f at v8+T, v at v8+2T, masks at v8+3T
  std::vector<int> in_regs(w.n);
  for (uint32_t i = 0; i < n; i++) {
    in_regs[i] = (int)(8 + i);
    in_regs[n + i] = (int)(8 + T + i);
    in_regs[2 * n + i] = (int)(8 + 2 * T + i);
  }
  for (uint32_t j = 0; j < 9; j++) in_regs[3 * n + j] = (int)(8 + 3 * T + j);
  for (uint32_t i = 0; i < n; i++) out_regs.push_back(8 + i);
  for (uint32_t i = 0; i < n; i++) out_regs.push_back(8 + T + i);
  for (uint32_t j = 0; j < 9; j++) out_regs.push_back(8 + 3 * T + j);
  const uint32_t pool_lo = 8 + 3 * T + 9;
  auto cone = output_cone(w, outs);
  auto order = schedule_gates(w, cone, outs);
  const std::vector<uint8_t> uni(w.n, 0);
  const std::vector<uint32_t> vb(w.n, 0);
  Circuit cx = w;
  LutProgram p0 = map_luts(w, order, out_wires, uni);
  split_uniform_pairs(cx, p0);
  LutProgram p = pressure_order(p0, out_wires, cx.wires(), nullptr);
  std::string why = "no register class tried";
  struct Cls {
    uint32_t budget, hi, agprs;
  };
  for (const Cls& cl : {Cls{256, 239, 0}, Cls{512, 239, 256}}) {
    try {
      Allocator a(cx, p, vb, 0, 0, pool_lo, cl.hi, cl.agprs, 0, out_wires, out_neg, kSgprLo, &in_regs, &out_regs);
      nk.body = a.run();
      nk.vgpr_class = cl.budget;
      nk.spills = a.spills();
      nk.reloads = a.reloads();
      break;
    } catch (const NativeUnfit& e) {
      why = e.what();
    }
  }
  if (!nk.vgpr_class) throw NativeUnfit("walk body does not fit the native templates: " + why);
  // every input varies per slot, so the body has no scalar work: the walk templates leave
  // the SGPRs to their own code
  for (size_t pos = 0; (pos = nk.body.find("\ts_", pos)) != std::string::npos; pos += 3)
    if (nk.body.compare(pos, 7, "\ts_nop ") != 0) throw std::runtime_error("walk body with scalar instructions");
  if (nk.body.find("lane_b32 s") != std::string::npos) throw std::runtime_error("walk body writes SGPRs");
  nk.info.valu_ops = (p.valu_ops + 1) / 2;  // per step
  nk.info.scalar_ops = p.scalar_ops;
  nk.kernel_name = "ns_nat_walk" + std::to_string(T) + "_r" + std::to_string(nk.vgpr_class);
  const char* tpl = native_template(nk.kernel_name);
  if (!tpl) throw std::runtime_error("missing native template " + nk.kernel_name);
  std::string text(tpl);
  const std::string marker = "; NS_BODY " + nk.kernel_name;
  size_t pos = text.find(marker);
  if (pos == std::string::npos) throw std::runtime_error("template placeholder not found");
  size_t eol = text.find('\n', pos);
  text.replace(pos, eol - pos, "; circuit body\n" + nk.body);
  nk.asm_text = std::move(text);
  return nk;
}

}  // namespace ns
