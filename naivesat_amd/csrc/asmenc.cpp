// gfx950 machine-code encoding of native circuit bodies and their placement in templates
// (see asmenc.h). Encodings follow the GFX9 formats: VOP1, VOPC, VOP3 (with v_bitop3_b32's
// truth table spread over the ABS, OMOD and NEG fields), VOP3P (AGPR copies), SOP1, SOP2,
// SOPP (s_nop and branches to the body's own labels).
#include "asmenc.h"

#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>

#include "asmgen.h"
#include "runtime.h"

namespace ns {

namespace {

constexpr uint32_t kLiteral = 255;

struct Operand {
  enum Kind { SGPR, VGPR, AGPR, IMM } kind;
  int64_t value;
};

bool parse_int(const std::string& t, int64_t* v) {
  if (t.empty()) return false;
  char* end = nullptr;
  long long x = std::strtoll(t.c_str(), &end, 0);
  if (*end != '\0') return false;
  *v = x;
  return true;
}

bool parse_operand(const std::string& t, Operand* op) {
  if (t == "vcc_lo" || t == "vcc_hi") {  // SGPRs 106 and 107
    op->kind = Operand::SGPR;
    op->value = t == "vcc_lo" ? 106 : 107;
    return true;
  }
  if (t.size() >= 2 && (t[0] == 's' || t[0] == 'v' || t[0] == 'a') && t[1] >= '0' && t[1] <= '9') {
    int64_t r;
    if (!parse_int(t.substr(1), &r)) return false;
    op->kind = t[0] == 's' ? Operand::SGPR : (t[0] == 'v' ? Operand::VGPR : Operand::AGPR);
    op->value = r;
    if (op->kind == Operand::SGPR && r > 101) return false;
    if (op->kind != Operand::SGPR && r > 255) return false;
    return true;
  }
  op->kind = Operand::IMM;
  return parse_int(t, &op->value);
}

// inline constant code for an integer, or kLiteral
uint32_t inline_code(int64_t v) {
  if (v >= 0 && v <= 64) return 128 + (uint32_t)v;
  if (v >= -16 && v <= -1) return 192 + (uint32_t)(-v);
  return kLiteral;
}

// 9-bit source field of a VALU operand; *lit receives a literal when one is needed
bool vsrc(const Operand& o, uint32_t* code, bool* has_lit, uint32_t* lit) {
  switch (o.kind) {
    case Operand::SGPR: *code = (uint32_t)o.value; return true;
    case Operand::VGPR: *code = 256 + (uint32_t)o.value; return true;
    case Operand::IMM:
      *code = inline_code(o.value);
      if (*code == kLiteral) {
        *has_lit = true;
        *lit = (uint32_t)o.value;
      }
      return true;
    default: return false;
  }
}

// 8-bit source field of a SALU operand
bool ssrc(const Operand& o, uint32_t* code, bool* has_lit, uint32_t* lit) {
  if (o.kind == Operand::SGPR) {
    *code = (uint32_t)o.value;
    return true;
  }
  if (o.kind != Operand::IMM) return false;
  *code = inline_code(o.value);
  if (*code == kLiteral) {
    *has_lit = true;
    *lit = (uint32_t)o.value;
  }
  return true;
}

uint32_t vop3(uint32_t op, uint32_t vdst, uint32_t s0, uint32_t s1, uint32_t s2, uint32_t* w1, uint32_t tt = 0) {
  // v_bitop3 truth table: bits 2..0 -> NEG, 5..3 -> ABS, 7..6 -> OMOD
  const uint32_t neg = tt & 7u, abs = (tt >> 3) & 7u, omod = (tt >> 6) & 3u;
  *w1 = s0 | (s1 << 9) | (s2 << 18) | (omod << 27) | (neg << 29);
  return (0x34u << 26) | (op << 16) | (abs << 8) | vdst;
}

}  // namespace

namespace {

// "v<N>" or "s<N>" at p (after spaces), then spaces and `end_ch` (',' or the end): the 9-bit
// VALU source code (SGPR N: N, VGPR N: 256 + N); false when the text is anything else
bool fast_reg(const char*& p, const char* e, char end_ch, uint32_t* code, bool vgpr_only) {
  while (p < e && (*p == ' ' || *p == '\t')) p++;
  if (p >= e || (*p != 'v' && *p != 's') || (vgpr_only && *p != 'v')) return false;
  const bool v = *p == 'v';
  p++;
  uint32_t n = 0;
  const char* d = p;
  while (p < e && *p >= '0' && *p <= '9' && p - d < 4) n = n * 10 + (uint32_t)(*p++ - '0');
  if (p == d || (p < e && *p >= '0' && *p <= '9')) return false;
  if (v ? n > 255 : n > 101) return false;
  while (p < e && (*p == ' ' || *p == '\t')) p++;
  if (end_ch) {
    if (p >= e || *p != end_ch) return false;
    p++;
  }
  *code = v ? 256 + n : n;
  return true;
}

// "v_bitop3_b32 vD, a, b, c bitop3:T" with register operands, the bulk of every body, parsed
// without building strings; false sends the line to the general parser (and its errors)
bool fast_bitop3(const char* b, const char* e, std::vector<uint32_t>& w) {
  static const char kMn[] = "v_bitop3_b32";
  const size_t mn = sizeof(kMn) - 1;
  if ((size_t)(e - b) <= mn || std::memcmp(b, kMn, mn) != 0 || (b[mn] != ' ' && b[mn] != '\t')) return false;
  const char* p = b + mn;
  uint32_t d, s0, s1, s2;
  if (!fast_reg(p, e, ',', &d, true) || !fast_reg(p, e, ',', &s0, false) || !fast_reg(p, e, ',', &s1, false) ||
      !fast_reg(p, e, 0, &s2, false))
    return false;
  static const char kMod[] = "bitop3:";
  const size_t ml = sizeof(kMod) - 1;
  if ((size_t)(e - p) <= ml || std::memcmp(p, kMod, ml) != 0) return false;
  p += ml;
  uint32_t tt = 0;
  if (e - p > 2 && p[0] == '0' && (p[1] == 'x' || p[1] == 'X')) {
    p += 2;
    const char* h = p;
    for (; p < e; p++) {
      const char c = *p;
      const int dig = c >= '0' && c <= '9' ? c - '0' : c >= 'a' && c <= 'f' ? c - 'a' + 10 : c >= 'A' && c <= 'F' ? c - 'A' + 10 : -1;
      if (dig < 0 || tt > 0xFF) break;
      tt = tt * 16 + (uint32_t)dig;
    }
    if (p == h) return false;
  } else {
    return false;
  }
  while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) p++;
  if (p != e || tt > 255) return false;
  uint32_t w1;
  w.push_back(vop3(0x234, d - 256, s0, s1, s2, &w1, tt));
  w.push_back(w1);
  return true;
}

}  // namespace

bool encode_body(const std::string& body, std::vector<uint32_t>& w, std::string* err) {
  std::string line;
  auto fail = [&](const std::string& m) {
    if (err) *err = m + ": " + line;
    return false;
  };
  // local labels ("name:") and the branches to them, resolved at the end: a SOPP branch
  // holds the signed distance in words from the instruction after it
  std::map<std::string, size_t> labels;
  std::vector<std::pair<size_t, std::string>> fixups;
  w.reserve(w.size() + body.size() / 12);
  for (size_t pos = 0; pos < body.size();) {
    size_t eol = body.find('\n', pos);
    if (eol == std::string::npos) eol = body.size();
    const char* lb = body.data() + pos;
    const char* le = body.data() + eol;
    pos = eol + 1;
    while (lb < le && (*lb == ' ' || *lb == '\t')) lb++;
    if (lb < le && fast_bitop3(lb, le, w)) continue;
    line.assign(lb, le);
    size_t a = line.find_first_not_of(" \t");
    if (a == std::string::npos || line[a] == ';') continue;
    std::string t = line.substr(a);
    while (!t.empty() && (t.back() == ' ' || t.back() == '\t' || t.back() == '\r')) t.pop_back();
    if (!t.empty() && t.back() == ':' && t.find_first_of(" \t,") == std::string::npos) {
      if (!labels.emplace(t.substr(0, t.size() - 1), w.size()).second) return fail("label defined twice");
      continue;
    }
    size_t sp = t.find_first_of(" \t");
    std::string mn = t.substr(0, sp);
    if (mn == "s_branch" || mn == "s_cbranch_vccz") {
      if (sp == std::string::npos) return fail("branch without a label");
      std::string target = t.substr(sp + 1);
      target.erase(0, target.find_first_not_of(" \t"));
      if (target.empty() || target.find_first_of(" \t,") != std::string::npos) return fail("bad branch label");
      fixups.push_back({w.size(), target});
      w.push_back(mn == "s_branch" ? 0xBF820000u : 0xBF860000u);
      continue;
    }
    if (mn == "v_cmp_ne_u32") {
      // VOPC, result in vcc: "v_cmp_ne_u32 vcc, src0, vN"
      std::vector<std::string> args;
      std::string cur;
      for (char ch : t.substr(sp == std::string::npos ? t.size() : sp + 1)) {
        if (ch == ',') {
          args.push_back(cur);
          cur.clear();
        } else if (ch != ' ' && ch != '\t') {
          cur += ch;
        }
      }
      if (!cur.empty()) args.push_back(cur);
      Operand s0, s1;
      if (args.size() != 3 || args[0] != "vcc" || !parse_operand(args[1], &s0) || !parse_operand(args[2], &s1) ||
          s1.kind != Operand::VGPR)
        return fail("bad v_cmp_ne_u32");
      uint32_t c0, lv = 0;
      bool lit = false;
      if (!vsrc(s0, &c0, &lit, &lv)) return fail("bad source");
      w.push_back(0x7C000000u | (0xCDu << 17) | ((uint32_t)s1.value << 9) | c0);
      if (lit) w.push_back(lv);
      continue;
    }
    std::vector<std::string> args;
    std::string modifier;
    if (sp != std::string::npos) {
      std::string rest = t.substr(sp + 1);
      size_t bp = rest.find("bitop3:");
      if (bp != std::string::npos) {
        modifier = rest.substr(bp + 7);
        rest = rest.substr(0, bp);
      }
      std::string cur;
      for (char ch : rest) {
        if (ch == ',') {
          args.push_back(cur);
          cur.clear();
        } else if (ch != ' ' && ch != '\t') {
          cur += ch;
        }
      }
      if (!cur.empty()) args.push_back(cur);
    }
    std::vector<Operand> ops(args.size());
    for (size_t k = 0; k < args.size(); k++)
      if (!parse_operand(args[k], &ops[k])) return fail("bad operand '" + args[k] + "'");
    auto is = [&](size_t k, Operand::Kind kd) { return k < ops.size() && ops[k].kind == kd; };
    if (mn == "v_bitop3_b32") {
      int64_t tt;
      if (ops.size() != 4 || !is(0, Operand::VGPR) || !parse_int(modifier, &tt) || tt < 0 || tt > 255)
        return fail("bad v_bitop3_b32");
      uint32_t s[3];
      for (int k = 0; k < 3; k++) {
        bool lit = false;
        uint32_t lv = 0;
        if (ops[1 + k].kind == Operand::IMM || !vsrc(ops[1 + k], &s[k], &lit, &lv)) return fail("bad source");
      }
      uint32_t w1;
      w.push_back(vop3(0x234, (uint32_t)ops[0].value, s[0], s[1], s[2], &w1, (uint32_t)tt));
      w.push_back(w1);
    } else if (mn == "v_mov_b32" || mn == "v_not_b32") {
      if (ops.size() != 2 || !is(0, Operand::VGPR)) return fail("bad VOP1");
      uint32_t s0, lv = 0;
      bool lit = false;
      if (!vsrc(ops[1], &s0, &lit, &lv)) return fail("bad source");
      const uint32_t op = mn == "v_mov_b32" ? 0x01 : 0x2b;
      w.push_back(0x7E000000u | ((uint32_t)ops[0].value << 17) | (op << 9) | s0);
      if (lit) w.push_back(lv);
    } else if (mn == "v_accvgpr_write_b32") {
      if (ops.size() != 2 || !is(0, Operand::AGPR) || !is(1, Operand::VGPR)) return fail("bad accvgpr write");
      w.push_back(0xD3D94000u | (uint32_t)ops[0].value);
      w.push_back(0x18000100u | (uint32_t)ops[1].value);
    } else if (mn == "v_accvgpr_read_b32") {
      if (ops.size() != 2 || !is(0, Operand::VGPR) || !is(1, Operand::AGPR)) return fail("bad accvgpr read");
      w.push_back(0xD3D84000u | (uint32_t)ops[0].value);
      w.push_back(0x18000100u | (uint32_t)ops[1].value);
    } else if (mn == "v_writelane_b32") {
      if (ops.size() != 3 || !is(0, Operand::VGPR) || !is(1, Operand::SGPR) || !is(2, Operand::IMM) ||
          ops[2].value < 0 || ops[2].value > 63)
        return fail("bad v_writelane_b32");
      uint32_t w1;
      w.push_back(vop3(0x28a, (uint32_t)ops[0].value, (uint32_t)ops[1].value, 128 + (uint32_t)ops[2].value, 0, &w1));
      w.push_back(w1);
    } else if (mn == "v_readlane_b32") {
      if (ops.size() != 3 || !is(0, Operand::SGPR) || !is(1, Operand::VGPR) || !is(2, Operand::IMM) ||
          ops[2].value < 0 || ops[2].value > 63)
        return fail("bad v_readlane_b32");
      uint32_t w1;
      w.push_back(
          vop3(0x289, (uint32_t)ops[0].value, 256 + (uint32_t)ops[1].value, 128 + (uint32_t)ops[2].value, 0, &w1));
      w.push_back(w1);
    } else if (mn == "s_mov_b32") {
      if (ops.size() != 2 || !is(0, Operand::SGPR)) return fail("bad s_mov_b32");
      uint32_t s0, lv = 0;
      bool lit = false;
      if (!ssrc(ops[1], &s0, &lit, &lv)) return fail("bad source");
      w.push_back(0xBE800000u | ((uint32_t)ops[0].value << 16) | s0);
      if (lit) w.push_back(lv);
    } else if (mn == "s_bfe_i32" || mn == "s_and_b32" || mn == "s_nor_b32" || mn == "s_andn2_b32" ||
               mn == "s_xor_b32") {
      if (ops.size() != 3 || !is(0, Operand::SGPR)) return fail("bad SOP2");
      uint32_t op = mn == "s_bfe_i32" ? 0x26 : mn == "s_and_b32" ? 0x0c : mn == "s_nor_b32" ? 0x18
                                                : mn == "s_andn2_b32" ? 0x12 : 0x10;
      uint32_t s0, s1, l0 = 0, l1 = 0;
      bool lit0 = false, lit1 = false;
      if (!ssrc(ops[1], &s0, &lit0, &l0) || !ssrc(ops[2], &s1, &lit1, &l1)) return fail("bad source");
      if (lit0 && lit1 && l0 != l1) return fail("two different literals");
      w.push_back(0x80000000u | (op << 23) | ((uint32_t)ops[0].value << 16) | (s1 << 8) | s0);
      if (lit0 || lit1) w.push_back(lit0 ? l0 : l1);
    } else if (mn == "s_nop") {
      if (ops.size() != 1 || !is(0, Operand::IMM) || ops[0].value < 0 || ops[0].value > 15) return fail("bad s_nop");
      w.push_back(0xBF800000u | (uint32_t)ops[0].value);
    } else {
      return fail("instruction outside the encoded subset");
    }
  }
  for (const auto& f : fixups) {
    auto it = labels.find(f.second);
    if (it == labels.end()) {
      line = f.second;
      return fail("branch to an undefined label");
    }
    const long long d = (long long)it->second - (long long)(f.first + 1);
    if (d < -32768 || d > 32767) {
      line = f.second;
      return fail("branch out of range");
    }
    w[f.first] |= (uint32_t)(d & 0xFFFF);
  }
  return true;
}

namespace {

// slot marker: the first 8 bytes of every reserved body region
const unsigned char kMarker[8] = {0x4d, 0x42, 0x53, 0x4e, 0x42, 0x4f, 0x44, 0x59};
// slot sizes; the largest stays within reach of one s_branch (+-32767 dwords)
const size_t kSledBytes[3] = {16384, 65536, 131008};

struct SledTemplate {
  std::vector<char> elf;
  size_t offset = 0;  // byte offset of the slot in the code object
  size_t bytes = 0;
};

std::mutex g_sled_mu;
std::map<std::string, std::shared_ptr<SledTemplate>> g_sleds;

std::shared_ptr<SledTemplate> sled_template(const std::string& kernel_name, size_t bytes) {
  const std::string key = kernel_name + "/" + std::to_string(bytes);
  std::lock_guard<std::mutex> lk(g_sled_mu);
  auto it = g_sleds.find(key);
  if (it != g_sleds.end()) return it->second;
  const char* tpl = native_template(kernel_name);
  if (!tpl) throw std::runtime_error("missing native template " + kernel_name);
  std::string text(tpl);
  const std::string marker = "; NS_BODY " + kernel_name;
  size_t pos = text.find(marker);
  if (pos == std::string::npos) throw std::runtime_error("template placeholder not found");
  size_t eol = text.find('\n', pos);
  std::ostringstream sled;
  sled << "; body slot\n\t.long 0x4e53424d\n\t.long 0x59444f42\n\t.fill " << (bytes / 4 - 2) << ", 4, 0xbf800000\n";
  text.replace(pos, eol - pos, sled.str());
  auto st = std::make_shared<SledTemplate>();
  st->elf = assemble_native(text);
  st->bytes = bytes;
  size_t found = 0, at = 0;
  for (size_t i = 0; i + 8 <= st->elf.size(); i++)
    if (std::memcmp(&st->elf[i], kMarker, 8) == 0) {
      found++;
      at = i;
    }
  if (found != 1) throw std::runtime_error("body slot marker not unique in the template code object");
  st->offset = at;
  g_sleds[key] = st;
  return st;
}

}  // namespace

namespace {

// batch region sizes (the bodies are called, so no branch range limits them)
const size_t kRegionBytes[5] = {65536, 262144, 1048576, 4194304, 16777216};
// literals of the marker instructions "s_mov_b32 s36, <literal>" that follow each batch
// kernel's s_getpc_b64 (its pc base): SAT 0x534253c0 + class index, QSAT 0x514253c0 + index
const uint32_t kPcMarkSat = 0x534253c0u, kPcMarkQsat = 0x514253c0u;
const char* const kClassNames[4] = {"64", "128", "256", "512"};
const uint32_t kSMovS36Lit = 0xBEA400FFu;  // s_mov_b32 s36, <32-bit literal>
const uint32_t kSGetpcS = 0xBE801C00u;     // s_getpc_b64 s[d:d+1] with the sdst field cleared

std::mutex g_batch_mu;
std::map<std::string, std::shared_ptr<BatchTemplate>> g_batch;

// offset of the unique "s_getpc_b64; s_mov_b32 s36, lit" pair's s_mov (the pc base)
size_t find_pc_base(const std::vector<char>& elf, uint32_t lit) {
  size_t found = 0, at = 0;
  for (size_t i = 8; i + 8 <= elf.size(); i += 4) {
    uint32_t w0, w1, w2;
    std::memcpy(&w0, &elf[i - 4], 4);
    std::memcpy(&w1, &elf[i], 4);
    std::memcpy(&w2, &elf[i + 4], 4);
    if (w1 == kSMovS36Lit && w2 == lit && (w0 & 0xFF80FF00u) == kSGetpcS) {
      found++;
      at = i;
    }
  }
  if (found != 1) throw std::runtime_error("batch kernel pc base marker not unique in the template code object");
  return at;
}

}  // namespace

int batch_class_index(uint32_t vgpr_class) {
  switch (vgpr_class) {
    case 64: return 0;
    case 128: return 1;
    case 256: return 2;
    case 512: return 3;
    default: throw std::runtime_error("no batch kernels for register class " + std::to_string(vgpr_class));
  }
}

std::shared_ptr<const BatchTemplate> batch_template(size_t min_bytes) {
  size_t bytes = 0;
  for (size_t b : kRegionBytes)
    if (min_bytes <= b) {
      bytes = b;
      break;
    }
  if (!bytes) return nullptr;
  const std::string name = "ns_nat_batch";
  const std::string key = name + "/" + std::to_string(bytes);
  std::lock_guard<std::mutex> lk(g_batch_mu);
  auto it = g_batch.find(key);
  if (it != g_batch.end()) return it->second;
  const char* tpl = native_template(name);
  if (!tpl) throw std::runtime_error("missing native template " + name);
  // the region goes at the end of .text, after every kernel: never reached by falling
  // through, entered only by the kernels' computed calls
  std::ostringstream text;
  text << tpl << "\n\t.text\n\t.p2align 8\n; batch region\n\t.long 0x4e53424d\n\t.long 0x59444f42\n\t.fill "
       << (bytes / 4 - 2) << ", 4, 0xbf800000\n";
  auto bt = std::make_shared<BatchTemplate>();
  bt->elf = assemble_native(text.str());
  bt->region_bytes = bytes;
  size_t found = 0;
  for (size_t i = 0; i + 8 <= bt->elf.size(); i++)
    if (std::memcmp(&bt->elf[i], kMarker, 8) == 0) {
      found++;
      bt->region = i;
    }
  if (found != 1 || bt->region % 4) throw std::runtime_error("batch region marker not unique in the template code object");
  for (int c = 0; c < 4; c++) {
    bt->pc_base[0][c] = find_pc_base(bt->elf, kPcMarkSat + c);
    bt->pc_base[1][c] = find_pc_base(bt->elf, kPcMarkQsat + c);
    // every pc base precedes the region (entries are unsigned offsets)
    if (bt->pc_base[0][c] >= bt->region || bt->pc_base[1][c] >= bt->region)
      throw std::runtime_error("batch region placed before a kernel");
    bt->kernel[0][c] = std::string("ns_nat_bsat_r") + kClassNames[c];
    bt->kernel[1][c] = std::string("ns_nat_bqsat_r") + kClassNames[c];
  }
  g_batch[key] = bt;
  return bt;
}

std::vector<char> compose_batch(const BatchTemplate& t, const std::vector<std::vector<uint32_t>>& bodies,
                                std::vector<uint64_t>& body_offset) {
  std::vector<char> elf = t.elf;
  body_offset.assign(bodies.size(), 0);
  size_t at = 0;  // byte offset in the region
  for (size_t i = 0; i < bodies.size(); i++) {
    const size_t need = bodies[i].size() * 4 + 4;  // body + return
    if (at + need > t.region_bytes) throw std::runtime_error("batch bodies exceed the region");
    std::memcpy(&elf[t.region + at], bodies[i].data(), bodies[i].size() * 4);
    const uint32_t ret = kSSetpcRet;
    std::memcpy(&elf[t.region + at + bodies[i].size() * 4], &ret, 4);
    body_offset[i] = at;
    at = (at + need + 255) & ~(size_t)255;  // bodies start on 256-byte boundaries
  }
  return elf;
}

size_t max_batch_region_bytes() { return kRegionBytes[sizeof(kRegionBytes) / sizeof(kRegionBytes[0]) - 1]; }

size_t batch_bytes(const std::vector<std::vector<uint32_t>>& bodies) {
  size_t at = 0;
  for (const auto& b : bodies) at = (at + b.size() * 4 + 4 + 255) & ~(size_t)255;
  return at;
}

std::vector<char> patch_native(const std::string& kernel_name, const std::string& body, size_t* slot_offset) {
  std::vector<uint32_t> words;
  std::string err;
  if (!encode_body(body, words, &err)) return {};
  const size_t need = words.size() * 4 + 4;  // body + branch to the end of the slot
  size_t bytes = 0;
  for (size_t b : kSledBytes)
    if (need <= b) {
      bytes = b;
      break;
    }
  if (!bytes) return {};
  std::shared_ptr<SledTemplate> st = sled_template(kernel_name, bytes);
  std::vector<char> elf = st->elf;
  const size_t body_bytes = words.size() * 4;
  const uint32_t skip = (uint32_t)((st->bytes - body_bytes - 4) / 4);  // dwords after the branch
  words.push_back(0xBF820000u | skip);                                  // s_branch to the slot end
  std::memcpy(&elf[st->offset], words.data(), words.size() * 4);
  if (slot_offset) *slot_offset = st->offset;
  return elf;
}

}  // namespace ns
