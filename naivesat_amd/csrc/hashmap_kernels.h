// Hash-map search of naivesatostop (src/satostop.rs:127-1150): types and launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ns {

// HashEntry (src/satostop.rs:168-174), 32 bytes
struct alignas(16) HmEntry {
  unsigned long long current, next, steps;
  unsigned int predecessors, state;  // state: 0 unused, 1 used, 2 stopped, 3 looped
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

// join_to_hashmap alone, in place
hipError_t hm_join_to(hipStream_t s, const HmParams& p, unsigned long long arg, const HmOutputs& o, HmEntry* hm,
                      unsigned long long len);
// join_to_hashmap (when o is given) and join_hashmap_itself from `in` into `out`; new
// predecessor counts go to jumps (one u32 per slot, zero before) until hm_add folds them in
hipError_t hm_join(hipStream_t s, const HmParams& p, unsigned long long arg, const HmOutputs* o, const HmEntry* in,
                   HmEntry* out, unsigned long long len, unsigned int* jumps, unsigned int* fills, HmSolution* sol);
// add_to_hashmap: claims (one u64 per slot) and jumps are zero before and after; clear
// zeroes every predecessor count afterwards
hipError_t hm_add(hipStream_t s, const HmParams& p, unsigned long long arg, const HmOutputs& o, HmEntry* hm,
                  unsigned long long len, unsigned int* fills, HmSolution* sol, unsigned long long* claims,
                  unsigned int* jumps, unsigned int max_predecessors, bool clear, bool test);
// adds jumps into the entries' predecessor counts (claims must be zero), leaving jumps zero
hipError_t hm_fold(hipStream_t s, const HmParams& p, HmEntry* hm, unsigned long long len, const unsigned int* fills,
                   unsigned long long* claims, unsigned int* jumps);

}  // namespace ns
