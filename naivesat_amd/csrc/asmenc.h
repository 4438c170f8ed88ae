// Native code path without an assembler per kernel: the circuit body (the instruction
// subset the register allocator emits) is encoded to gfx950 machine code here and copied
// into a code object of its template whose body slot is a reserved region ("sled"); the
// template with a sled is assembled once per (template, sled size) and cached.
#pragma once
#include <cstdint>
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

}  // namespace ns
