// Device side of naivesatostop's hash-map search (src/satostop.rs:127-1150): one thread per
// hash entry (join passes) or per state of the current argument block (insertion). Entry
// layout, states and update rules follow the reference so that its data sets apply.
#include "hashmap_kernels.h"

namespace ns {

namespace {

constexpr unsigned int HS_UNUSED = 0, HS_USED = 1, HS_STOPPED = 2, HS_LOOPED = 3, HS_RESERVED = 4;

// hash_function_64 (src/satostop.rs:127-136), u64 wrapping arithmetic
__device__ inline unsigned long long hash64(unsigned int bits, unsigned long long value) {
  const unsigned long long mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
  const unsigned int half = bits >> 1;
  unsigned long long temp = value * 9615409803190489167ull;
  temp = (temp << half) | (bits - half >= 64 ? 0ull : (temp >> (bits - half)));
  return ((value * 6171710485021949031ull) ^ temp ^ 0xb89d2ecda078ca1full) & mask;
}

// output word(s) of element j of the current block: lo[j*stride] | hi[j*stride] << 32
__device__ inline unsigned long long output_at(const HmOutputs& o, unsigned long long j) {
  unsigned long long v = o.lo[j * o.stride];
  if (o.hi) v |= (unsigned long long)o.hi[j * o.stride] << 32;
  return v;
}

// resolve_unknowns (src/satostop.rs:330-381, device form 489-538)
__device__ inline void resolve_unknowns(const HmParams& p, unsigned long long current, unsigned long long next,
                                        unsigned long long steps, unsigned int state, unsigned int* fills,
                                        HmSolution* sol) {
  if ((current & ((1ull << (p.state_len - p.unknown_bits)) - 1ull)) != 0) return;
  if (state != HS_LOOPED && state != HS_STOPPED) return;
  if (state == HS_STOPPED && atomicOr(&sol->defined, 1u) == 0u) {
    sol->start = current;
    sol->end = next;
    sol->steps = steps;
  }
  const unsigned long long idx = current >> (p.state_len - p.unknown_bits + p.unknown_fill_bits);
  const unsigned int mask = (1u << p.unknown_fill_bits) - 1u;
  const unsigned int value = (unsigned int)((current >> (p.state_len - p.unknown_bits)) & mask);
  if (atomicCAS(fills + idx, value, value + 1u) == value && value == mask)
    atomicAdd(&sol->resolved_unknowns, 1ull);
}

}  // namespace

// join_to_hashmap (src/satostop.rs:186-266): entries whose next state lies in the block
// take the block's output for it
__global__ void __launch_bounds__(256) k_hm_join_to(HmParams p, unsigned long long arg, HmOutputs o,
                                                    HmEntry* __restrict__ hm, unsigned long long len) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long arg_start = arg << p.arg_bit_place, arg_end = arg_start + (1ull << p.arg_bit_place);
  const unsigned long long state_mask = (1ull << p.state_len) - 1ull;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
    HmEntry* he = hm + i;
    if (he->state != HS_USED || he->next < arg_start || he->next >= arg_end) continue;
    const unsigned long long output = output_at(o, he->next - arg_start);
    const unsigned long long old_next = he->next;
    he->next = output & state_mask;
    if ((output >> p.state_len) & 1ull)
      he->state = HS_STOPPED;
    else if (state_mask <= he->steps || he->next == he->current || he->next == old_next)
      he->state = HS_LOOPED;
    else
      he->state = HS_USED;
    he->steps += 1;
  }
}

__global__ void k_hm_zero_preds(HmEntry* __restrict__ hm, unsigned long long len) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride)
    hm[i].predecessors = 0u;
}

// join_hashmap_itself_and_check_solution (src/satostop.rs:540-594): an entry whose next
// state has an entry jumps over it; predecessor counts move to the entries jumped to
__global__ void __launch_bounds__(256) k_hm_join_itself(HmParams p, const HmEntry* __restrict__ in,
                                                        HmEntry* __restrict__ out, unsigned long long len,
                                                        unsigned int* fills, HmSolution* sol) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned int shift = p.state_len - p.hbits;
  const unsigned long long state_mask = p.state_len >= 64 ? ~0ull : (1ull << p.state_len) - 1ull;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
    const HmEntry inhe = in[i];
    HmEntry* outhe = out + i;
    bool copy = true;
    unsigned long long cur = inhe.current, nxt = inhe.next, steps = inhe.steps;
    unsigned int state = inhe.state;
    if (inhe.state == HS_USED) {
      const unsigned long long next_idx = hash64(p.state_len, inhe.next) >> shift;
      const HmEntry nexthe = in[next_idx];
      if (nexthe.state != HS_UNUSED && nexthe.current == inhe.next) {
        const unsigned long long sum = inhe.steps + nexthe.steps;
        nxt = nexthe.next;
        steps = sum;
        state = nexthe.state;
        if (state == HS_USED && (cur == nxt || sum < inhe.steps || sum > state_mask)) state = HS_LOOPED;
        atomicAdd(&out[next_idx].predecessors, 1u);
        copy = false;
      }
    }
    outhe->current = cur;
    outhe->next = copy ? inhe.next : nxt;
    outhe->steps = steps;
    outhe->state = state;
    atomicAdd(&outhe->predecessors, inhe.predecessors);
    resolve_unknowns(p, cur, copy ? inhe.next : nxt, steps, state, fills, sol);
  }
}

// add_to_hashmap_and_check_solution (src/satostop.rs:862-942): the block's states enter
// the map at their hash slot unless the slot holds a current-unknown start or a state with
// more than max_predecessors predecessors; a slot is claimed with the RESERVED flag
__global__ void __launch_bounds__(256) k_hm_add(HmParams p, unsigned long long arg, HmOutputs o,
                                                HmEntry* __restrict__ hm, unsigned int* fills, HmSolution* sol,
                                                unsigned int max_predecessors, unsigned int test) {
  const unsigned long long count = 1ull << p.arg_bit_place;
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long arg_start = arg << p.arg_bit_place;
  const unsigned long long state_mask = (1ull << p.state_len) - 1ull;
  const unsigned int shift = p.state_len - p.hbits;
  const unsigned long long fill_mask = (1ull << p.unknown_fill_bits) - 1ull;
  const unsigned long long before_unknowns = (1ull << (p.state_len - p.unknown_bits)) - 1ull;
  const unsigned int fill_shift = p.state_len - p.unknown_bits + p.unknown_fill_bits;
  for (unsigned long long idx = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; idx < count;
       idx += stride) {
    const unsigned long long output = output_at(o, idx);
    const unsigned long long current = arg_start + idx;
    const unsigned long long next = output & state_mask;
    const unsigned int state =
        ((output >> p.state_len) & 1ull) ? HS_STOPPED : (next == current ? HS_LOOPED : HS_USED);
    if (test && next == 0) continue;
    const bool cur_solved =
        (current & before_unknowns) == 0 &&
        atomicAdd(fills + (current >> fill_shift), 0u) == (unsigned int)((current >> (p.state_len - p.unknown_bits)) & fill_mask);
    HmEntry* curhe = hm + (hash64(p.state_len, current) >> shift);
    bool again = true;
    for (unsigned int trial = 0; trial < 10 && again; trial++) {
      const unsigned int old_state = atomicOr(&curhe->state, HS_RESERVED);
      __threadfence();
      bool old_solved = false;
      const unsigned long long old_current = curhe->current;
      if (old_state != HS_UNUSED && (old_state & HS_RESERVED) == 0)
        old_solved = (old_current & before_unknowns) == 0 &&
                     atomicAdd(fills + (old_current >> fill_shift), 0u) ==
                         (unsigned int)((old_current >> (p.state_len - p.unknown_bits)) & fill_mask);
      if ((old_state & HS_RESERVED) == 0 &&
          ((cur_solved && !old_solved) || (!old_solved && curhe->predecessors <= max_predecessors))) {
        curhe->current = current;
        curhe->next = next;
        curhe->steps = 1;
        curhe->predecessors = 0;
        __threadfence();
        atomicExch(&curhe->state, state);
        again = false;
      } else {
        again = (old_state & HS_RESERVED) != 0 && cur_solved && !old_solved;
        __threadfence();
        if ((old_state & HS_RESERVED) == 0) atomicAnd(&curhe->state, ~HS_RESERVED);
      }
    }
    resolve_unknowns(p, current, next, 1, state, fills, sol);
  }
}

// hashmap_clear_predecessors (src/satostop.rs:1132-1150)
__global__ void k_hm_clear_preds(HmEntry* __restrict__ hm, unsigned long long len) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride)
    hm[i].predecessors = 0u;
}

static unsigned int grid_for(unsigned long long items) {
  const unsigned long long b = (items + 255ull) / 256ull;
  return (unsigned int)(b < (1ull << 22) ? (b ? b : 1) : (1ull << 22));
}

hipError_t hm_join_to(hipStream_t s, const HmParams& p, unsigned long long arg, const HmOutputs& o, HmEntry* hm,
                      unsigned long long len) {
  hipLaunchKernelGGL(k_hm_join_to, dim3(grid_for(len)), dim3(256), 0, s, p, arg, o, hm, len);
  return hipGetLastError();
}

hipError_t hm_join_itself(hipStream_t s, const HmParams& p, const HmEntry* in, HmEntry* out, unsigned long long len,
                          unsigned int* fills, HmSolution* sol) {
  hipLaunchKernelGGL(k_hm_zero_preds, dim3(grid_for(len)), dim3(256), 0, s, out, len);
  hipLaunchKernelGGL(k_hm_join_itself, dim3(grid_for(len)), dim3(256), 0, s, p, in, out, len, fills, sol);
  return hipGetLastError();
}

hipError_t hm_add(hipStream_t s, const HmParams& p, unsigned long long arg, const HmOutputs& o, HmEntry* hm,
                  unsigned int* fills, HmSolution* sol, unsigned int max_predecessors, bool test) {
  hipLaunchKernelGGL(k_hm_add, dim3(grid_for(1ull << p.arg_bit_place)), dim3(256), 0, s, p, arg, o, hm, fills, sol,
                     max_predecessors, test ? 1u : 0u);
  return hipGetLastError();
}

hipError_t hm_clear_preds(hipStream_t s, HmEntry* hm, unsigned long long len) {
  hipLaunchKernelGGL(k_hm_clear_preds, dim3(grid_for(len)), dim3(256), 0, s, hm, len);
  return hipGetLastError();
}

}  // namespace ns
