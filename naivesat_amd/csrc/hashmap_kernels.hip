// Device side of naivesatostop's hash-map search (src/satostop.rs:127-1150): one thread per
// hash entry (join passes) or per state of the current argument block (insertion). Entry
// layout, states and update rules follow the reference so that its data sets apply.
#include "hashmap_kernels.h"

namespace ns {

namespace {

constexpr unsigned int HS_UNUSED = 0, HS_USED = 1, HS_STOPPED = 2, HS_LOOPED = 3;

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

// a start state of an unknown whose fill group has resolved every value below its own
// (the "currently solved" test of src/satostop.rs:889-894)
__device__ inline bool current_unknown(const HmParams& p, unsigned long long current, const unsigned int* fills) {
  const unsigned int low = p.state_len - p.unknown_bits;
  if ((current & ((1ull << low) - 1ull)) != 0) return false;
  const unsigned int mask = (1u << p.unknown_fill_bits) - 1u;
  const unsigned int v = __hip_atomic_load(fills + (current >> (low + p.unknown_fill_bits)), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
  return v == (unsigned int)((current >> low) & mask);
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

// join_to_hashmap's update of one entry (src/satostop.rs:186-266): a used entry whose next
// state lies in the block [arg_start, arg_start + 2^arg_bit_place) takes the block's output
// for it
__device__ inline void join_to(const HmParams& p, unsigned long long arg_start, const HmOutputs& o, HmEntry& he) {
  if (he.state != HS_USED || he.next < arg_start || he.next - arg_start >= (1ull << p.arg_bit_place)) return;
  const unsigned long long state_mask = (1ull << p.state_len) - 1ull;
  const unsigned long long output = output_at(o, he.next - arg_start);
  const unsigned long long old_next = he.next;
  he.next = output & state_mask;
  if ((output >> p.state_len) & 1ull)
    he.state = HS_STOPPED;
  else if (state_mask <= he.steps || he.next == he.current || he.next == old_next)
    he.state = HS_LOOPED;
  else
    he.state = HS_USED;
  he.steps += 1;
}

__device__ inline HmEntry load_entry(const HmEntry* he) {
  const ulonglong2* q = reinterpret_cast<const ulonglong2*>(he);
  const ulonglong2 a = q[0], b = q[1];
  HmEntry e;
  e.current = a.x;
  e.next = a.y;
  e.steps = b.x;
  e.predecessors = (unsigned int)b.y;
  e.state = (unsigned int)(b.y >> 32);
  return e;
}

__device__ inline void store_entry(HmEntry* he, const HmEntry& e) {
  ulonglong2* q = reinterpret_cast<ulonglong2*>(he);
  q[0] = make_ulonglong2(e.current, e.next);
  q[1] = make_ulonglong2(e.steps, (unsigned long long)e.predecessors | ((unsigned long long)e.state << 32));
}

}  // namespace

// join_to_hashmap (src/satostop.rs:186-266) on its own, in place
__global__ void __launch_bounds__(256) k_hm_join_to(HmParams p, unsigned long long arg, HmOutputs o,
                                                    HmEntry* __restrict__ hm, unsigned long long len) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long arg_start = arg << p.arg_bit_place;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
    HmEntry he = load_entry(hm + i);
    const unsigned long long old_next = he.next;
    const unsigned int old_state = he.state;
    join_to(p, arg_start, o, he);
    if (he.next != old_next || he.state != old_state) store_entry(hm + i, he);
  }
}

// join_to_hashmap followed by join_hashmap_itself_and_check_solution (src/satostop.rs:540-594)
// in one pass: every entry read from `in` is first brought up to date with the block (when
// has_block is set), then a used entry whose next state has an entry jumps over it. The
// entries jumped to count their new predecessors in `jumps` (one u32 per slot, zero on
// entry); out[i].predecessors holds in[i]'s count, and the insertion pass that follows adds
// jumps[i] to it and clears jumps[i].
__global__ void __launch_bounds__(256) k_hm_join(HmParams p, unsigned long long arg, HmOutputs o,
                                                 unsigned int has_block, const HmEntry* __restrict__ in,
                                                 HmEntry* __restrict__ out, unsigned long long len,
                                                 unsigned int* __restrict__ jumps, unsigned int* fills,
                                                 HmSolution* sol) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long arg_start = arg << p.arg_bit_place;
  const unsigned int shift = p.state_len - p.hbits;
  const unsigned long long state_mask = (1ull << p.state_len) - 1ull;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
    HmEntry e = load_entry(in + i);
    if (has_block) join_to(p, arg_start, o, e);
    if (e.state == HS_USED) {
      const unsigned long long next_idx = hash64(p.state_len, e.next) >> shift;
      HmEntry ne = load_entry(in + next_idx);
      if (ne.state != HS_UNUSED && ne.current == e.next) {
        if (has_block) join_to(p, arg_start, o, ne);
        const unsigned long long sum = e.steps + ne.steps;
        e.next = ne.next;
        e.state = ne.state;
        if (e.state == HS_USED && (e.current == e.next || sum < e.steps || sum > state_mask)) e.state = HS_LOOPED;
        e.steps = sum;
        atomicAdd(jumps + next_idx, 1u);
      }
    }
    store_entry(out + i, e);
    resolve_unknowns(p, e.current, e.next, e.steps, e.state, fills, sol);
  }
}

// add_to_hashmap_and_check_solution (src/satostop.rs:862-942). The block's states enter the
// map at their hash slot unless the slot holds a current-unknown start (the next unknown of
// its fill group to resolve) or a state with more than max_predecessors predecessors. The
// reference lets every state race for its slot behind a RESERVED flag, so which state keeps
// a contested slot depends on arrival order. Here one writer per slot is decided first, in
// claims (one u64 per slot: bit 40 marks a current-unknown start, the low bits hold the
// state's index + 1, or the inverted index for a current-unknown start): the block's other
// states store their key with a plain store (the last store to land wins), then the block's
// unknown starts apply theirs with atomicMax, so a current-unknown start wins over every
// other state and the lowest one wins among them. That is the reference's rule for one
// arrival order. A pass over the slots then writes the winners' entries with plain,
// coalesced stores (no fence is needed), and a last one resolves the block's unknown starts,
// as if every insertion thread reached resolve_unknowns after all of them had tried their
// slot: one of the executions the reference's kernel allows.
__global__ void __launch_bounds__(256) k_hm_claim(HmParams p, unsigned long long arg, HmOutputs o,
                                                  unsigned long long* __restrict__ claims, unsigned int test) {
  const unsigned long long count = 1ull << p.arg_bit_place;
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long arg_start = arg << p.arg_bit_place;
  const unsigned long long state_mask = (1ull << p.state_len) - 1ull;
  const unsigned long long start_mask = (1ull << (p.state_len - p.unknown_bits)) - 1ull;
  for (unsigned long long idx = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; idx < count;
       idx += stride) {
    const unsigned long long current = arg_start + idx;
    if ((current & start_mask) == 0) continue;  // unknown starts claim in k_hm_claim_starts
    if (test && (output_at(o, idx) & state_mask) == 0) continue;
    claims[hash64(p.state_len, current) >> (p.state_len - p.hbits)] = idx + 1ull;
  }
}

// the block's unknown start states: every 2^(state_len - unknown_bits)-th state from the
// first one at or after the block's start
struct BlockStarts {
  unsigned long long first, count;
  unsigned int low;
};

__device__ inline BlockStarts block_starts(const HmParams& p, unsigned long long arg) {
  BlockStarts b;
  b.low = p.state_len - p.unknown_bits;
  const unsigned long long arg_start = arg << p.arg_bit_place, step = 1ull << b.low;
  const unsigned long long block_end = arg_start + (1ull << p.arg_bit_place);
  b.first = (arg_start + step - 1ull) & ~(step - 1ull);
  b.count = b.first < block_end ? (block_end - b.first + step - 1ull) >> b.low : 0ull;
  return b;
}

__global__ void __launch_bounds__(256) k_hm_claim_starts(HmParams p, unsigned long long arg, HmOutputs o,
                                                         const unsigned int* fills,
                                                         unsigned long long* __restrict__ claims, unsigned int test) {
  const BlockStarts b = block_starts(p, arg);
  const unsigned long long arg_start = arg << p.arg_bit_place;
  const unsigned long long state_mask = (1ull << p.state_len) - 1ull;
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long k = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; k < b.count; k += stride) {
    const unsigned long long current = b.first + (k << b.low), idx = current - arg_start;
    if (test && (output_at(o, idx) & state_mask) == 0) continue;
    const unsigned long long key =
        current_unknown(p, current, fills) ? (1ull << 40) | ((1ull << 40) - 1ull - idx) : idx + 1ull;
    atomicMax(claims + (hash64(p.state_len, current) >> (p.state_len - p.hbits)), key);
  }
}

// Per slot: adds the join pass's new predecessors, writes the claim's winner when the
// reference's rule lets it replace the slot's entry, and (clear set) clears the predecessor
// counts as hashmap_clear_predecessors does (src/satostop.rs:1132-1150). Leaves claims and
// jumps zero.
__global__ void __launch_bounds__(256) k_hm_place(HmParams p, unsigned long long arg, HmOutputs o,
                                                  HmEntry* __restrict__ hm, unsigned long long len,
                                                  const unsigned int* fills, unsigned long long* __restrict__ claims,
                                                  unsigned int* __restrict__ jumps, unsigned int max_predecessors,
                                                  unsigned int clear) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long arg_start = arg << p.arg_bit_place;
  const unsigned long long state_mask = (1ull << p.state_len) - 1ull;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
    const unsigned long long key = claims[i];
    const unsigned int extra = jumps[i];
    if (key == 0 && extra == 0 && !clear) continue;
    HmEntry he = load_entry(hm + i);
    he.predecessors += extra;
    if (key != 0) {
      const bool cur_solved = (key >> 40) != 0;
      const unsigned long long idx = cur_solved ? (1ull << 40) - 1ull - (key & ((1ull << 40) - 1ull)) : key - 1ull;
      const bool old_solved = he.state != HS_UNUSED && current_unknown(p, he.current, fills);
      if ((cur_solved && !old_solved) || (!old_solved && he.predecessors <= max_predecessors)) {
        const unsigned long long output = output_at(o, idx);
        he.current = arg_start + idx;
        he.next = output & state_mask;
        he.state = ((output >> p.state_len) & 1ull) ? HS_STOPPED : (he.next == he.current ? HS_LOOPED : HS_USED);
        he.steps = 1;
        he.predecessors = 0;
      }
      claims[i] = 0ull;
    }
    if (clear) he.predecessors = 0;
    if (extra) jumps[i] = 0u;
    store_entry(hm + i, he);
  }
}

// The block's unknown start states that loop or stop resolve after every insertion attempt
// of the block, as when each of the reference's insertion threads reaches resolve_unknowns
// after all of them have tried their slot.
__global__ void __launch_bounds__(256) k_hm_resolve_block(HmParams p, unsigned long long arg, HmOutputs o,
                                                          unsigned int* fills, HmSolution* sol, unsigned int test) {
  const BlockStarts b = block_starts(p, arg);
  const unsigned long long arg_start = arg << p.arg_bit_place;
  const unsigned long long state_mask = (1ull << p.state_len) - 1ull;
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long k = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; k < b.count; k += stride) {
    const unsigned long long current = b.first + (k << b.low);
    const unsigned long long output = output_at(o, current - arg_start);
    const unsigned long long next = output & state_mask;
    if (test && next == 0) continue;
    const unsigned int state = ((output >> p.state_len) & 1ull) ? HS_STOPPED : (next == current ? HS_LOOPED : HS_USED);
    resolve_unknowns(p, current, next, 1, state, fills, sol);
  }
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

hipError_t hm_join(hipStream_t s, const HmParams& p, unsigned long long arg, const HmOutputs* o, const HmEntry* in,
                   HmEntry* out, unsigned long long len, unsigned int* jumps, unsigned int* fills, HmSolution* sol) {
  const HmOutputs none{nullptr, nullptr, 1};
  hipLaunchKernelGGL(k_hm_join, dim3(grid_for(len)), dim3(256), 0, s, p, arg, o ? *o : none, o ? 1u : 0u, in, out,
                     len, jumps, fills, sol);
  return hipGetLastError();
}

hipError_t hm_add(hipStream_t s, const HmParams& p, unsigned long long arg, const HmOutputs& o, HmEntry* hm,
                  unsigned long long len, unsigned int* fills, HmSolution* sol, unsigned long long* claims,
                  unsigned int* jumps, unsigned int max_predecessors, bool clear, bool test) {
  const unsigned int low = p.state_len - p.unknown_bits;
  const unsigned long long starts = p.arg_bit_place > low ? 1ull << (p.arg_bit_place - low) : 1ull;
  hipLaunchKernelGGL(k_hm_claim, dim3(grid_for(1ull << p.arg_bit_place)), dim3(256), 0, s, p, arg, o, claims,
                     test ? 1u : 0u);
  hipLaunchKernelGGL(k_hm_claim_starts, dim3(grid_for(starts)), dim3(256), 0, s, p, arg, o, fills, claims,
                     test ? 1u : 0u);
  hipLaunchKernelGGL(k_hm_place, dim3(grid_for(len)), dim3(256), 0, s, p, arg, o, hm, len, fills, claims, jumps,
                     max_predecessors, clear ? 1u : 0u);
  hipLaunchKernelGGL(k_hm_resolve_block, dim3(grid_for(starts)), dim3(256), 0, s, p, arg, o, fills, sol,
                     test ? 1u : 0u);
  return hipGetLastError();
}

hipError_t hm_fold(hipStream_t s, const HmParams& p, HmEntry* hm, unsigned long long len, const unsigned int* fills,
                   unsigned long long* claims, unsigned int* jumps) {
  const HmOutputs none{nullptr, nullptr, 1};
  hipLaunchKernelGGL(k_hm_place, dim3(grid_for(len)), dim3(256), 0, s, p, 0ull, none, hm, len, fills, claims, jumps,
                     0u, 0u);
  return hipGetLastError();
}

}  // namespace ns
