// Native code path: circuit bodies written directly as gfx950 assembly into precompiled
// kernel templates (templates/gen_templates.py) and assembled in-process.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "circuit.h"
#include "codegen.h"

namespace ns {

// BATCH: a body for the persistent batch kernels (templates' ns_nat_bsat / ns_nat_bqsat):
// one output in v0, lane masks in v1..v6, SGPRs s48..s101; called as a subroutine.
enum class NativeKind { SAT, QSAT_A, QSAT_B, TABLE, BATCH };

// Thrown when a circuit does not fit the register budget of any template (the caller then
// uses the compiler path).
struct NativeUnfit : std::runtime_error {
  explicit NativeUnfit(const std::string& m) : std::runtime_error(m) {}
};

struct NativeKernel {
  std::string asm_text;     // complete assembly module
  std::string body;         // the circuit body alone (what fills the template's slot)
  std::string kernel_name;  // symbol to launch
  uint32_t vgpr_class = 0;  // 64 / 128 / 256 / 512 (VGPRs + AGPR spill slots)
  uint32_t spills = 0;      // values copied to AGPRs per word
  uint32_t reloads = 0;     // AGPR reads per word
  uint32_t output_groups = 1;  // the output's conjunction evaluated as this many groups
  GenInfo info;
};

// Builds the assembly module for `kind`. vbit_of_input[k] = position of input k in the
// assignment index; lane bits are index bits 5+lane_shift .. 10+lane_shift; outs = the
// outputs evaluated (their values land in v0.. in order). group_exits (SAT): when the output
// is evaluated in groups, the body leaves early with output 0 once the groups so far give 0
// in every lane of the wave (the answers do not change; the skipped gates are not evaluated).
NativeKernel build_native(const Circuit& c, NativeKind kind, const std::vector<uint32_t>& vbit_of_input,
                          uint32_t lane_shift, const std::vector<uint32_t>& outs, bool group_exits = false);

// Walk body of the table-free SATOSTOP walks (templates' ns_nat_walk<T>_r<class>, T = the
// smallest state tier of 16, 24, 32, 40, 48 that holds c.n): two walk steps and their
// bookkeeping, with waypoints every 2^k states. Throws NativeUnfit past 48 state bits or
// when the body fits no register class.
NativeKernel build_native_walk(const Circuit& c, uint32_t k);

// Number of embedded templates and lookup (templates_data.cpp, generated at build time).
const char* native_template(const std::string& kernel_name);

}  // namespace ns
