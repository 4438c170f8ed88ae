// Native code path without an assembler per kernel: the circuit body (the instruction
// subset the register allocator emits) is encoded to gfx950 machine code here and copied
// into a code object of its template whose body slot is a reserved region ("sled"); the
// template with a sled is assembled once per (template, sled size) and cached.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace ns {

// Encodes body text (one instruction per line, ';' comments) into 32-bit words. Returns
// false and sets *err for an instruction outside the supported subset.
bool encode_body(const std::string& body, std::vector<uint32_t>& words, std::string* err);

// Code object of template `kernel_name` with `body` in its slot, or an empty vector when the
// body does not fit the largest slot or cannot be encoded (the caller then assembles the
// full text instead).
std::vector<char> patch_native(const std::string& kernel_name, const std::string& body,
                               size_t* slot_offset = nullptr);

// Batch code objects: the persistent batch kernels of every register class (SAT and QSAT:
// ns_nat_bsat_rC / ns_nat_bqsat_rC, C = 64, 128, 256, 512) with a region after them in .text
// that holds circuit bodies as subroutines. Kernel (kind k, class c) calls body i at its pc
// base + entry, where entry = region + body_offset[i] - pc_base[k][c] (byte offsets in the
// code object; the loader keeps .text contiguous, so the same distances hold on the device).
// Each body returns with s_setpc_b64 s[38:39].
struct BatchTemplate {
  std::vector<char> elf;
  size_t region = 0;  // byte offset of the region in the code object
  size_t region_bytes = 0;
  size_t pc_base[2][4] = {};  // [0 SAT / 1 QSAT][class index: 64, 128, 256, 512]
  std::string kernel[2][4];
};
constexpr uint32_t kSSetpcRet = 0xBE801D26u;  // s_setpc_b64 s[38:39]
// The cached template whose region holds at least min_bytes (null: too large).
std::shared_ptr<const BatchTemplate> batch_template(size_t min_bytes);
// index of a register class in BatchTemplate's tables (64 -> 0 ... 512 -> 3)
int batch_class_index(uint32_t vgpr_class);
// Region bytes the bodies take (each body + its return, on 256-byte boundaries).
size_t batch_bytes(const std::vector<std::vector<uint32_t>>& bodies);
// The largest batch region any template holds.
size_t max_batch_region_bytes();
// The template's code object with the bodies (encoded words) in its region, each followed
// by the return; body_offset[i] = byte offset of body i in the region.
std::vector<char> compose_batch(const BatchTemplate& t, const std::vector<std::vector<uint32_t>>& bodies,
                                std::vector<uint64_t>& body_offset);

}  // namespace ns
