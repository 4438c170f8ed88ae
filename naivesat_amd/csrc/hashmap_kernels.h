// Hash-map search of naivesatostop (src/satostop.rs:127-1150): types and launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ns {

// HashEntry (src/satostop.rs:168-174), 32 bytes
struct HmEntry {
  unsigned long long current, next, steps;
  unsigned int predecessors, state;  // state: 0 unused, 1 used, 2 stopped, 3 looped (+4 reserved)
};
static_assert(sizeof(HmEntry) == 32, "hash entry layout");

// SolutionAndResUnknowns (src/satostop.rs:481-487)
struct HmSolution {
  unsigned long long resolved_unknowns, start, end, steps;
  unsigned int defined;
};

struct HmParams {
  unsigned int state_len, arg_bit_place, hbits, unknown_bits, unknown_fill_bits;
};

// outputs of the current argument block: element j = lo[j*stride] | hi[j*stride] << 32
// (hi null when the state and STOP fit 32 bits)
struct HmOutputs {
  const unsigned int* lo;
  const unsigned int* hi;
  unsigned long long stride;
};

hipError_t hm_join_to(hipStream_t s, const HmParams& p, unsigned long long arg, const HmOutputs& o, HmEntry* hm,
                      unsigned long long len);
hipError_t hm_join_itself(hipStream_t s, const HmParams& p, const HmEntry* in, HmEntry* out, unsigned long long len,
                          unsigned int* fills, HmSolution* sol);
hipError_t hm_add(hipStream_t s, const HmParams& p, unsigned long long arg, const HmOutputs& o, HmEntry* hm,
                  unsigned int* fills, HmSolution* sol, unsigned int max_predecessors, bool test);
hipError_t hm_clear_preds(hipStream_t s, HmEntry* hm, unsigned long long len);

}  // namespace ns
