// Circuit-independent kernels: QSAT reduction passes over bit arrays, the
// next-state table's pointer-jumping passes and the unknown-start query.
#include <hip/hip_runtime.h>

#include <mutex>

#include "static_kernels.h"

namespace ns {

// ---------------------------------------------------------------------------
// QSAT reduction pass.
//
// `in` holds 2^m bits; bit i is the value of the quantified formula below the
// already reduced levels for index i. The pass reduces the next k levels (k <= 11:
// five inside a 32-bit word, then up to six across lanes with shuffles): level l
// combines the pairs of values whose indices differ in bit l, with AND when bit
// (opmask >> l) is set (universal quantifier) and OR otherwise. Output: 2^(m-k)
// bits in `out` (must be zeroed). When `final` is set the output values are the
// outermost block's values u: stats[0] += number of witnesses (value != q0all),
// stats[1] = min(lowest witness u).
// ---------------------------------------------------------------------------
__device__ inline unsigned int red_level(unsigned int x, unsigned int sh, bool all) {
  return all ? (x & (x >> sh)) : (x | (x >> sh));
}

__device__ inline void wave_stats(unsigned long long cnt, unsigned long long best,
                                  unsigned long long* stats) {
  if (__ballot(cnt != 0ull) == 0ull) return;
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long ob = __shfl_xor(best, off);
    best = ob < best ? ob : best;
    cnt += __shfl_xor(cnt, off);
  }
  if ((threadIdx.x & 63u) == 0u) {
    atomicAdd(&stats[0], cnt);
    atomicMin(&stats[1], best);
  }
}

__global__ void __launch_bounds__(256) k_qsat_reduce(const unsigned int* __restrict__ in,
                                                     unsigned int m, unsigned int k,
                                                     unsigned int opmask, unsigned int* out,
                                                     unsigned int final_pass, unsigned int q0all,
                                                     unsigned long long* stats) {
  const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long nwords = m >= 5 ? (1ull << (m - 5)) : 1ull;
  const bool valid = t < nwords;
  const unsigned int lane = threadIdx.x & 63u;
  unsigned int x = valid ? in[t] : 0u;
  const unsigned int kin = k < 5 ? k : 5;
  for (unsigned int l = 0; l < kin; l++) x = red_level(x, 1u << l, (opmask >> l) & 1u);
  unsigned long long cnt = 0, best = ~0ull;
  if (k <= 5) {
    const unsigned int per_word = (m >= 5 ? 32u : (1u << m)) >> k;
    unsigned int r = 0;
    for (unsigned int p = 0; p < per_word; p++) r |= ((x >> (p << k)) & 1u) << p;
    const unsigned int vmask = per_word >= 32 ? 0xffffffffu : ((1u << per_word) - 1u);
    if (!valid) r = 0;
    const unsigned long long ob = t * (unsigned long long)(32u >> k);
    if (final_pass && valid) {
      unsigned int wit = (q0all ? ~r : r) & vmask;
      cnt = __popc(wit);
      if (wit) best = ob + (unsigned long long)__builtin_ctz(wit);
    }
    if (r) atomicOr(&out[ob >> 5], r << (unsigned int)(ob & 31ull));
  } else {
    unsigned int b = x & 1u;
    for (unsigned int l = 5; l < k; l++) {
      unsigned int y = __shfl_xor(b, 1 << (l - 5));
      b = ((opmask >> l) & 1u) ? (b & y) : (b | y);
    }
    const unsigned long long ball = __ballot(valid && b != 0u);
    const unsigned long long t0 = t - lane;
    if (lane == 0 && t0 < nwords) {
      const unsigned int s = k - 5;
      const unsigned long long lanes_valid = (nwords - t0) < 64ull ? (nwords - t0) : 64ull;
      const unsigned int nres = (unsigned int)(lanes_valid >> s);
      unsigned long long r = 0;
      for (unsigned int p = 0; p < nres; p++) r |= ((ball >> (p << s)) & 1ull) << p;
      const unsigned long long vmask = nres >= 64 ? ~0ull : ((1ull << nres) - 1ull);
      const unsigned long long ob = t0 >> s;
      if (final_pass) {
        unsigned long long wit = (q0all ? ~r : r) & vmask;
        cnt = __popcll(wit);
        if (wit) best = ob + (unsigned long long)__builtin_ctzll(wit);
      }
      if (r & 0xffffffffull) atomicOr(&out[ob >> 5], (unsigned int)r << (unsigned int)(ob & 31ull));
      if (r >> 32) atomicOr(&out[(ob >> 5) + 1], (unsigned int)(r >> 32));
    }
  }
  if (final_pass) wave_stats(cnt, best, stats);
}

// ---------------------------------------------------------------------------
// Pointer jumping over the next-state table.
//
// Entry layouts (u64). Narrow (n <= 31): bits 0..n-1 next state, bit n STOP, bits
// 32..63 number of circuit evaluations the entry stands for. Wide (n <= 33): the same
// with the evaluation count in bits 34..63 (at most 2^30 - 1; the wide path runs at most
// 29 passes). A stopped entry is final; any other entry i is replaced by T[next(i)] with
// the counts added, so after pass p every entry stands for exactly 2^p evaluations or for
// the path up to its STOP.
// ---------------------------------------------------------------------------
template <bool W>
struct Layout {
  static constexpr unsigned int shift = W ? 34u : 32u;  // first bit of the count
  static constexpr unsigned long long low = (1ull << shift) - 1ull;
};

// (the table kernels loop over their entries: a launch holds fewer than 2^32 threads)
__global__ void __launch_bounds__(256) k_jump_first(const unsigned int* __restrict__ t0,
                                                    unsigned long long* __restrict__ dst,
                                                    unsigned long long count, unsigned int n) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned int stop_bit = 1u << n;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const unsigned int e = t0[i];
    unsigned long long r;
    if (e & stop_bit) {
      r = (unsigned long long)e | (1ull << 32);
    } else {
      const unsigned int f = t0[e & (stop_bit - 1u)];
      r = (unsigned long long)f | (2ull << 32);
    }
    dst[i] = r;
  }
}

// wide first pass: entry s of T0 = lo[s] | hi[s] << split (outputs 0..split-1 and split..n)
__global__ void __launch_bounds__(256) k_jump_first_wide(const unsigned int* __restrict__ lo,
                                                         const unsigned int* __restrict__ hi,
                                                         unsigned long long* __restrict__ dst,
                                                         unsigned long long count, unsigned int n,
                                                         unsigned int split) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long mask = (1ull << n) - 1ull;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const unsigned long long e = (unsigned long long)lo[i] | ((unsigned long long)hi[i] << split);
    unsigned long long r;
    if ((e >> n) & 1ull) {
      r = (e & Layout<true>::low) | (1ull << 34);
    } else {
      const unsigned long long x = e & mask;
      const unsigned long long f = (unsigned long long)lo[x] | ((unsigned long long)hi[x] << split);
      r = (f & Layout<true>::low) | (2ull << 34);
    }
    dst[i] = r;
  }
}

template <bool W>
__global__ void __launch_bounds__(256) k_jump_pass(const unsigned long long* __restrict__ src,
                                                   unsigned long long* __restrict__ dst,
                                                   unsigned long long count, unsigned int n) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x * 2ull;
  const unsigned long long stop_bit = 1ull << n;
  constexpr unsigned int sh = Layout<W>::shift;
  for (unsigned long long base = ((unsigned long long)blockIdx.x * blockDim.x + threadIdx.x) * 2ull; base < count;
       base += stride) {
    // two entries per thread: two independent gathers in flight
    const ulonglong2 e2 = *reinterpret_cast<const ulonglong2*>(src + base);
    unsigned long long e[2] = {e2.x, e2.y};
    unsigned long long f[2];
#pragma unroll
    for (int q = 0; q < 2; q++) f[q] = (e[q] & stop_bit) ? 0ull : src[e[q] & (stop_bit - 1ull)];
#pragma unroll
    for (int q = 0; q < 2; q++)
      if (!(e[q] & stop_bit)) e[q] = (f[q] & Layout<W>::low) | (((e[q] >> sh) + (f[q] >> sh)) << sh);
    *reinterpret_cast<ulonglong2*>(dst + base) = make_ulonglong2(e[0], e[1]);
  }
}

// Per unknown start u (state u << (n - unknowns)), with tab = the table after P passes
// (a non-stopped entry covers 2^P evaluations; a stopped one holds the evaluations until
// STOP and the end state): follows up to `jumps` entries from the start, so the walk covers
// jumps x 2^P evaluations. steps_out[u] = evaluations until STOP (0 when it does not stop
// within them), ends_out[u] = state produced by the stopping evaluation; stats[0] +=
// stopped starts, stats[1] = min(lowest stopped u).
template <bool W>
__global__ void __launch_bounds__(256) k_starts(const unsigned long long* __restrict__ tab,
                                                unsigned int n, unsigned int unknowns,
                                                unsigned long long jumps, unsigned long long* steps_out,
                                                unsigned long long* ends_out, unsigned long long* stats) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long limit = ((1ull << unknowns) + 63ull) & ~63ull;  // whole waves take part
  unsigned long long cnt = 0, best = ~0ull;
  for (unsigned long long u = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; u < limit; u += stride) {
    if (u >= (1ull << unknowns)) continue;
    constexpr unsigned int sh = Layout<W>::shift;
    const unsigned long long mask = (1ull << n) - 1ull;
    unsigned long long x = u << (n - unknowns);
    unsigned long long total = 0, end = 0;
    bool stop = false;
    for (unsigned long long j = 0; j < jumps; j++) {
      const unsigned long long e = tab[x];
      total += e >> sh;
      if ((e >> n) & 1ull) {
        stop = true;
        end = e & mask;
        break;
      }
      x = e & mask;
    }
    if (steps_out) steps_out[u] = stop ? total : 0ull;
    if (ends_out) ends_out[u] = stop ? end : 0ull;
    if (stop) {
      cnt++;
      best = u < best ? u : best;
    }
  }
  wave_stats(cnt, best, stats);
}

// ---------------------------------------------------------------------------
// Waypoint walks: the full search without passes over the whole table.
//
// Waypoints are the states whose low k bits are zero; k <= n - unknowns, so every
// unknown's start state is one. One walk per waypoint follows the next-state table from
// it, evaluation by evaluation, until the first STOP, the next waypoint or a cycle of
// non-waypoint states (Brent's test against a state saved at power-of-two distances). Its
// outcome is a 16-byte record of the waypoint graph:
//   x = target | kind << 62, y = evaluations
//   WP_NEXT: target = index of the waypoint reached (its state >> k)
//   WP_STOP: target = state produced by the stopping evaluation
//   WP_LOOP: the walk entered a cycle without STOP or waypoint: it never stops
//   WP_LONG: more than `limit` evaluations (the caller falls back to the passes)
// Pointer jumping over these records (at most n - k passes, since a path visits each of
// the 2^(n-k) waypoints at most once before it repeats) resolves every waypoint to STOP
// or never; a start's answer is its own waypoint's record. Every state's path is followed
// exactly as in the passes, so the answers are the same; only the order of work differs.
// ---------------------------------------------------------------------------
constexpr unsigned int WP_NEXT = 0u, WP_STOP = 1u, WP_LOOP = 2u, WP_LONG = 3u;
constexpr unsigned long long WP_TARGET = (1ull << 62) - 1ull;

// Entry s of the next-state table (next state in bits 0..n-1, STOP at bit n; bits above n
// undefined): packed (n+1)-bit entries, entry s at bit s*(n+1) of little-endian u64 words
// (the reference's MemImage layout, src/greedysatostop.rs:339-400), or lo[s] | hi[s] << split.
__device__ inline unsigned long long table_entry(const TableView& t, unsigned long long s) {
  if (t.packed) {
    const unsigned long long bit = s * t.width;
    const unsigned long long idx = bit >> 6;
    const unsigned int b = (unsigned int)(bit & 63ull);
    unsigned long long v = t.packed[idx] >> b;
    if (b + t.width > 64u) v |= t.packed[idx + 1] << (64u - b);
    return v;
  }
  unsigned long long e = t.lo[s];
  if (t.hi) e |= (unsigned long long)t.hi[s] << t.split;
  return e;
}

// A lane that finishes a walk takes its next waypoint (grid stride) at once, so lanes
// stay busy while walks of different lengths end. A lane also stops, raising the flag,
// after `budget` evaluations in all, and every lane stops once it sees the flag (checked
// every 256 evaluations): maps whose walks are far longer than the waypoint spacing end
// the kernel early and go to the passes.
__global__ void __launch_bounds__(256) k_wp_walk(TableView tab, unsigned int n, unsigned int k,
                                                 unsigned long long count,
                                                 unsigned long long limit, unsigned long long budget,
                                                 ulonglong2* __restrict__ out, unsigned int* long_flag,
                                                 unsigned long long* evals) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long mask = (1ull << n) - 1ull, wmask = (1ull << k) - 1ull;
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long s = i << k, saved = s, steps = 0, power = 1, lam = 0, used = 0;
  for (bool active = i < count; active;) {
    if (++used > budget) {
      atomicOr(long_flag, 1u);
      break;
    }
    if ((used & 255ull) == 0ull &&
        __hip_atomic_load(long_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
      break;
    const unsigned long long e = table_entry(tab, s);
    steps++;
    const unsigned long long nx = e & mask;
    unsigned int kind = 4u;
    unsigned long long target = 0;
    if ((e >> n) & 1ull) {
      kind = WP_STOP;
      target = nx;
    } else if ((nx & wmask) == 0ull) {
      kind = WP_NEXT;
      target = nx >> k;
    } else if (nx == saved) {
      kind = WP_LOOP;
    } else if (steps >= limit) {
      kind = WP_LONG;
      atomicOr(long_flag, 1u);
    } else {
      s = nx;
      if (++lam == power) {
        saved = s;
        power <<= 1;
        lam = 0;
      }
    }
    if (kind != 4u) {
      out[i] = make_ulonglong2(target | ((unsigned long long)kind << 62), steps);
      i += stride;
      if (i >= count) break;
      s = i << k;
      saved = s;
      steps = 0;
      power = 1;
      lam = 0;
    }
  }
  // evaluations done by the wave, into one of 64 count slots 16 words apart
  unsigned long long total = used;
  for (int off = 32; off > 0; off >>= 1) total += __shfl_xor(total, off);
  if ((threadIdx.x & 63u) == 0u && total != 0ull) atomicAdd(evals + ((blockIdx.x & 63u) << 4), total);
}

// One pointer-jumping pass over the waypoint records; *pending |= some record still
// points at a waypoint.
__global__ void __launch_bounds__(256) k_wp_jump(const ulonglong2* __restrict__ in, ulonglong2* __restrict__ out,
                                                 unsigned long long count, unsigned int* pending) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  unsigned int any = 0u;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    ulonglong2 e = in[i];
    if ((unsigned int)(e.x >> 62) == WP_NEXT) {
      const ulonglong2 t = in[e.x & WP_TARGET];
      const unsigned int tk = (unsigned int)(t.x >> 62);
      e.x = t.x;
      e.y = tk == WP_NEXT || tk == WP_STOP ? e.y + t.y : 0ull;
      any |= tk == WP_NEXT ? 1u : 0u;
    }
    out[i] = e;
  }
  // one flag write per block at most, and none once some block has set it: atomics from
  // every wave to this one word would serialize
  if (__syncthreads_or((int)any) && threadIdx.x == 0 &&
      __hip_atomic_load(pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
    atomicOr(pending, 1u);
}

// Coarser waypoints from finer ones: the waypoints of spacing k + d are the spacing-k ones
// whose index has its low d bits zero. Coarse record j follows the fine records from fine
// index j << d until a fine WP_NEXT points at a coarse waypoint (WP_NEXT to its coarse
// index), a WP_STOP (its stopping state) or a WP_LOOP / WP_LONG record; evaluations add up
// along the way. A chain that comes back to a fine record it passed (Brent's test against
// the index saved at power-of-two hop counts) is a cycle without a coarse waypoint: WP_LOOP.
// A chain of more than max_hops records sets *overflow (the caller then keeps the fine
// level). Pointer jumping over the coarse records gives the answers pointer jumping over the
// fine ones would: a path between two coarse waypoints is the same sequence of walks.
__global__ void __launch_bounds__(256) k_wp_chase(const ulonglong2* __restrict__ in, unsigned int d,
                                                  ulonglong2* __restrict__ out, unsigned long long count_out,
                                                  unsigned int max_hops, unsigned int* overflow) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long dmask = (1ull << d) - 1ull;
  unsigned int over = 0u;
  for (unsigned long long j = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; j < count_out; j += stride) {
    unsigned long long idx = j << d, saved = idx, steps = 0ull;
    ulonglong2 res = make_ulonglong2((unsigned long long)WP_LOOP << 62, 0ull);
    for (unsigned int h = 1;; h++) {
      const ulonglong2 e = in[idx];
      const unsigned int kind = (unsigned int)(e.x >> 62);
      if (kind == WP_STOP) {
        res = make_ulonglong2(e.x, steps + e.y);
        break;
      }
      if (kind != WP_NEXT) {  // WP_LOOP or WP_LONG: never stops
        res = make_ulonglong2(e.x & ~WP_TARGET, 0ull);
        break;
      }
      steps += e.y;
      const unsigned long long t = e.x & WP_TARGET;
      if ((t & dmask) == 0ull) {
        res = make_ulonglong2((t >> d) | ((unsigned long long)WP_NEXT << 62), steps);
        break;
      }
      if (t == saved) break;  // res stays WP_LOOP
      if (h >= max_hops) {
        over = 1u;
        break;
      }
      if ((h & (h - 1u)) == 0u) saved = t;
      idx = t;
    }
    out[j] = res;
  }
  if (__syncthreads_or((int)over) && threadIdx.x == 0) atomicOr(overflow, 1u);
}

// Per unknown u: the record of its start's waypoint (u << (n - unknowns - k)).
__global__ void __launch_bounds__(256) k_wp_starts(const ulonglong2* __restrict__ rec, unsigned int n,
                                                   unsigned int unknowns, unsigned int k,
                                                   unsigned long long* steps_out, unsigned long long* ends_out,
                                                   unsigned long long* stats) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long limit = ((1ull << unknowns) + 63ull) & ~63ull;  // whole waves take part
  unsigned long long cnt = 0, best = ~0ull;
  for (unsigned long long u = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; u < limit; u += stride) {
    if (u >= (1ull << unknowns)) continue;
    const ulonglong2 e = rec[u << (n - unknowns - k)];
    const bool stop = (unsigned int)(e.x >> 62) == WP_STOP;
    if (steps_out) steps_out[u] = stop ? e.y : 0ull;
    if (ends_out) ends_out[u] = stop ? e.x & WP_TARGET : 0ull;
    if (stop) {
      cnt++;
      best = u < best ? u : best;
    }
  }
  wave_stats(cnt, best, stats);
}

// Random-read probe: each lane reads `reads` entries of t (2^bits u32) at indices drawn from
// its own linear congruential sequence (independent of the values read, so the reads can
// overlap) and folds them into one output word: the rate of scattered 4-byte reads that
// the memory system sustains, the ceiling of the waypoint walks' table reads.
// nontemporal: the loads carry the non-temporal hint (__builtin_nontemporal_load).
template <bool NT>
__global__ void __launch_bounds__(256) k_gather_probe(const unsigned int* __restrict__ t, unsigned int bits,
                                                      unsigned int reads, unsigned int* out) {
  const unsigned long long tid = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long mask = (1ull << bits) - 1ull;
  unsigned long long x = (tid + 1ull) * 0x9e3779b97f4a7c15ull;
  unsigned int acc = 0u;
#pragma unroll 8
  for (unsigned int r = 0; r < reads; r++) {
    x = x * 6364136223846793005ull + 1442695040888963407ull;
    const unsigned int* p = t + ((x >> 24) & mask);
    acc ^= NT ? __builtin_nontemporal_load(p) : *p;
  }
  out[tid] = acc;
}

// Packs entries lo[e] | hi[e] << split (e < entries, entries a multiple of 64) into
// width-bit fields: output word w holds bits [64w, 64w + 64) of the packed stream, so every
// lane writes one whole word and reads the two or three entries it overlaps.
__global__ void __launch_bounds__(256) k_pack(const unsigned int* __restrict__ lo, const unsigned int* __restrict__ hi,
                                              unsigned int split, unsigned int width, unsigned long long entries,
                                              unsigned long long* __restrict__ out) {
  const unsigned long long words = (entries * width) >> 6;
  const unsigned long long fmask = width >= 64 ? ~0ull : (1ull << width) - 1ull;
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long w = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
    const unsigned long long bit = w << 6;
    unsigned long long e = bit / width;
    unsigned int off = (unsigned int)(bit - e * width), pos = 0;
    unsigned long long word = 0;
    while (pos < 64u && e < entries) {
      unsigned long long v = lo[e];
      if (hi) v |= (unsigned long long)hi[e] << split;
      word |= ((v & fmask) >> off) << pos;
      pos += width - off;
      off = 0;
      e++;
    }
    out[w] = word;
  }
}

// ---------------------------------------------------------------------------
// greedysatostop's partitioned search over packed images (src/greedysatostop.rs:419-475).
// ---------------------------------------------------------------------------
__device__ inline unsigned long long mem_get(const unsigned long long* __restrict__ w, unsigned int width,
                                             unsigned long long i) {
  const unsigned long long bit = i * width, idx = bit >> 6;
  const unsigned int b = (unsigned int)(bit & 63ull);
  unsigned long long v = w[idx] >> b;
  if (b + width > 64u) v |= w[idx + 1] << (64u - b);
  return width >= 64 ? v : v & ((1ull << width) - 1ull);
}

// join_nexts: entry e of dest (len entries) that has not stopped and whose next state lies in
// [s_start, s_start + s_len) takes second's entry for it. Out of place (every output word is
// written by one lane from the entries it overlaps, read from the unchanged dest), so each
// entry's update reads the old dest and second only, as in the reference.
__global__ void __launch_bounds__(256) k_mem_join(const unsigned long long* __restrict__ dest, unsigned long long len,
                                                  const unsigned long long* __restrict__ second,
                                                  unsigned long long s_start, unsigned long long s_len,
                                                  unsigned int state_len, unsigned long long* __restrict__ out) {
  const unsigned int width = state_len + 1;
  const unsigned long long state_mask = (1ull << state_len) - 1ull, stop = 1ull << state_len;
  const unsigned long long words = (len * width) >> 6;
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long w = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
    const unsigned long long bit = w << 6;
    unsigned long long e = bit / width;
    unsigned int off = (unsigned int)(bit - e * width), pos = 0;
    unsigned long long word = 0;
    while (pos < 64u && e < len) {
      unsigned long long v = mem_get(dest, width, e);
      const unsigned long long nx = v & state_mask;
      if ((v & stop) == 0ull && nx >= s_start && nx - s_start < s_len) v = mem_get(second, width, nx - s_start);
      word |= (v >> off) << pos;
      pos += width - off;
      off = 0;
      e++;
    }
    out[w] = word;
  }
}

// find_solution: *best = min(*best, lowest unknown start in [start, start + len) whose entry
// has stopped); unknown starts are the multiples of 2^(state_len - unknowns).
__global__ void __launch_bounds__(256) k_mem_find(const unsigned long long* __restrict__ words, unsigned long long start,
                                                  unsigned long long len, unsigned int state_len, unsigned int unknowns,
                                                  unsigned long long* best) {
  const unsigned int width = state_len + 1, low = state_len - unknowns;
  const unsigned long long step = 1ull << low;
  const unsigned long long first = (start + step - 1ull) & ~(step - 1ull), end = start + len;
  const unsigned long long count = first < end ? (end - first + step - 1ull) >> low : 0ull;
  const unsigned long long limit = (count + 63ull) & ~63ull;  // whole waves take part
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  unsigned long long mine = ~0ull;
  for (unsigned long long k = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; k < limit; k += stride) {
    if (k >= count) continue;
    const unsigned long long u = first + (k << low);
    if ((mem_get(words, width, u - start) >> state_len) & 1ull) mine = u < mine ? u : mine;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(mine, off);
    mine = o < mine ? o : mine;
  }
  if ((threadIdx.x & 63u) == 0u && mine != ~0ull) atomicMin(best, mine);
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, unsigned long long count) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
    p[i] = v;
}

// Per request of a SAT batch (one block each): the sum of its chunks' counts and the lowest
// of their lowest hits (chunk_out[2c], [2c + 1]) into its result {count, lowest}.
__global__ void __launch_bounds__(256) k_batch_fold(const BatchFoldDesc* descs, const unsigned long long* chunk_out) {
  const BatchFoldDesc d = descs[blockIdx.x];
  const unsigned int n_ch = (d.wend + d.words - 1u) / d.words;
  unsigned long long cnt = 0, best = ~0ull;
  for (unsigned int j = threadIdx.x; j < n_ch; j += blockDim.x) {
    const unsigned long long c = d.chunk0 + j;
    cnt += chunk_out[2 * c];
    const unsigned long long b = chunk_out[2 * c + 1];
    best = b < best ? b : best;
  }
  __shared__ unsigned long long s_cnt[256], s_best[256];
  s_cnt[threadIdx.x] = cnt;
  s_best[threadIdx.x] = best;
  __syncthreads();
  for (unsigned int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      s_cnt[threadIdx.x] += s_cnt[threadIdx.x + h];
      const unsigned long long o = s_best[threadIdx.x + h];
      if (o < s_best[threadIdx.x]) s_best[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    d.out[0] = s_cnt[0];
    d.out[1] = s_best[0];
  }
}

__global__ void k_fill_strided_u64(unsigned long long* p, unsigned long long v, unsigned long long count,
                                   unsigned long long stride) {
  const unsigned long long step = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += step)
    p[i * stride] = v;
}

__global__ void k_fill_u32(unsigned int* p, unsigned int v, unsigned long long count) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// Counters of the allocation-free SAT entry point (ns_sat_solve_noalloc): [0] count, [1]
// lowest hit, 64 count slots 16 words apart from [512], as the SAT kernels write them. One
// array per device for the life of the process; calls that overlap on one device race.
// In host code `scratch` names a host-side stand-in: only the address that
// hipGetSymbolAddress returns for the current device may reach a kernel or a copy.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned long long scratch[kNoallocScratchWords];

namespace {
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kNoallocMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_setup(int device) {
  if (device < 0 || device >= (int)kNoallocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  int cur = 0;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != device) return hipErrorInvalidDevice;  // the caller makes the device current
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = static_cast<unsigned long long*>(p);
  return hipSuccess;
}

unsigned long long* noalloc_scratch(int device) {
  if (device < 0 || device >= (int)kNoallocMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline unsigned int blocks_for(unsigned long long threads, unsigned int bs) {
  return (unsigned int)((threads + bs - 1) / bs);
}

// grid of a kernel that loops over its items: one item per thread up to 2^31 threads
static inline unsigned int looped_blocks(unsigned long long items, unsigned int bs) {
  const unsigned long long cap = (1ull << 31) / bs;
  const unsigned long long b = (items + bs - 1) / bs;
  return (unsigned int)(b < cap ? b : cap);
}

hipError_t launch_qsat_reduce(hipStream_t s, const unsigned int* in, unsigned int m, unsigned int k,
                              unsigned int opmask, unsigned int* out, bool final_pass, bool q0all,
                              unsigned long long* stats) {
  unsigned long long nwords = m >= 5 ? (1ull << (m - 5)) : 1ull;
  unsigned long long threads = (nwords + 63ull) & ~63ull;
  hipLaunchKernelGGL(k_qsat_reduce, dim3(blocks_for(threads, 256)), dim3(256), 0, s, in, m, k, opmask,
                     out, final_pass ? 1u : 0u, q0all ? 1u : 0u, stats);
  return hipGetLastError();
}

hipError_t launch_jump_first(hipStream_t s, const unsigned int* t0, unsigned long long* dst,
                             unsigned long long count, unsigned int n) {
  hipLaunchKernelGGL(k_jump_first, dim3(looped_blocks(count, 256)), dim3(256), 0, s, t0, dst, count, n);
  return hipGetLastError();
}

hipError_t launch_jump_first_wide(hipStream_t s, const unsigned int* lo, const unsigned int* hi,
                                  unsigned long long* dst, unsigned long long count, unsigned int n,
                                  unsigned int split) {
  hipLaunchKernelGGL(k_jump_first_wide, dim3(looped_blocks(count, 256)), dim3(256), 0, s, lo, hi, dst, count, n,
                     split);
  return hipGetLastError();
}

hipError_t launch_jump_pass(hipStream_t s, const unsigned long long* src, unsigned long long* dst,
                            unsigned long long count, unsigned int n, bool wide) {
  if (wide)
    hipLaunchKernelGGL(k_jump_pass<true>, dim3(looped_blocks(count / 2, 256)), dim3(256), 0, s, src, dst, count, n);
  else
    hipLaunchKernelGGL(k_jump_pass<false>, dim3(looped_blocks(count / 2, 256)), dim3(256), 0, s, src, dst, count, n);
  return hipGetLastError();
}

hipError_t launch_starts(hipStream_t s, const unsigned long long* tab, unsigned int n, unsigned int unknowns,
                         unsigned long long jumps, bool wide, unsigned long long* steps_out,
                         unsigned long long* ends_out, unsigned long long* stats) {
  unsigned long long cnt = 1ull << unknowns;
  unsigned long long threads = (cnt + 63ull) & ~63ull;
  if (wide)
    hipLaunchKernelGGL(k_starts<true>, dim3(looped_blocks(threads, 256)), dim3(256), 0, s, tab, n, unknowns, jumps,
                       steps_out, ends_out, stats);
  else
    hipLaunchKernelGGL(k_starts<false>, dim3(looped_blocks(threads, 256)), dim3(256), 0, s, tab, n, unknowns, jumps,
                       steps_out, ends_out, stats);
  return hipGetLastError();
}

hipError_t launch_pack(hipStream_t s, const unsigned int* lo, const unsigned int* hi, unsigned int split,
                       unsigned int width, unsigned long long entries, unsigned long long* out) {
  hipLaunchKernelGGL(k_pack, dim3(looped_blocks((entries * width) >> 6, 256)), dim3(256), 0, s, lo, hi, split, width,
                     entries, out);
  return hipGetLastError();
}

hipError_t launch_wp_walk(hipStream_t s, const TableView& tab, unsigned int n, unsigned int k,
                          unsigned long long count, unsigned long long limit, void* out, unsigned int* long_flag,
                          unsigned long long* evals) {
  // enough lanes to keep the memory system's queues full; each takes several walks
  const unsigned long long lanes = count < (1ull << 19) ? count : (1ull << 19);
  // a lane's walks average 2^k evaluations each on a random map; allow 16 times that
  // over its walks plus one walk of the full limit
  const unsigned long long per_lane = (count + lanes - 1) / lanes;
  const unsigned long long budget = 16ull * per_lane * (1ull << k) + limit;
  hipLaunchKernelGGL(k_wp_walk, dim3(blocks_for(lanes, 256)), dim3(256), 0, s, tab, n, k, count, limit, budget,
                     static_cast<ulonglong2*>(out), long_flag, evals);
  return hipGetLastError();
}

hipError_t launch_wp_jump(hipStream_t s, const void* in, void* out, unsigned long long count, unsigned int* pending) {
  hipLaunchKernelGGL(k_wp_jump, dim3(looped_blocks(count, 256)), dim3(256), 0, s,
                     static_cast<const ulonglong2*>(in), static_cast<ulonglong2*>(out), count, pending);
  return hipGetLastError();
}

hipError_t launch_wp_chase(hipStream_t s, const void* in, unsigned int d, void* out, unsigned long long count_out,
                           unsigned int max_hops, unsigned int* overflow) {
  hipLaunchKernelGGL(k_wp_chase, dim3(looped_blocks(count_out, 256)), dim3(256), 0, s,
                     static_cast<const ulonglong2*>(in), d, static_cast<ulonglong2*>(out), count_out, max_hops,
                     overflow);
  return hipGetLastError();
}

hipError_t launch_wp_starts(hipStream_t s, const void* rec, unsigned int n, unsigned int unknowns, unsigned int k,
                            unsigned long long* steps_out, unsigned long long* ends_out, unsigned long long* stats) {
  const unsigned long long threads = ((1ull << unknowns) + 63ull) & ~63ull;
  hipLaunchKernelGGL(k_wp_starts, dim3(looped_blocks(threads, 256)), dim3(256), 0, s,
                     static_cast<const ulonglong2*>(rec), n, unknowns, k, steps_out, ends_out, stats);
  return hipGetLastError();
}

hipError_t launch_gather_probe(hipStream_t s, const unsigned int* t, unsigned int bits, unsigned int reads,
                               unsigned long long lanes, unsigned int* out, bool nontemporal) {
  if (nontemporal)
    hipLaunchKernelGGL(k_gather_probe<true>, dim3(blocks_for(lanes, 256)), dim3(256), 0, s, t, bits, reads, out);
  else
    hipLaunchKernelGGL(k_gather_probe<false>, dim3(blocks_for(lanes, 256)), dim3(256), 0, s, t, bits, reads, out);
  return hipGetLastError();
}

hipError_t launch_mem_join(hipStream_t s, const unsigned long long* dest, unsigned long long len,
                           const unsigned long long* second, unsigned long long s_start, unsigned long long s_len,
                           unsigned int state_len, unsigned long long* out) {
  hipLaunchKernelGGL(k_mem_join, dim3(looped_blocks((len * (state_len + 1)) >> 6, 256)), dim3(256), 0, s, dest, len,
                     second, s_start, s_len, state_len, out);
  return hipGetLastError();
}

hipError_t launch_mem_find(hipStream_t s, const unsigned long long* words, unsigned long long start,
                           unsigned long long len, unsigned int state_len, unsigned int unknowns,
                           unsigned long long* best) {
  const unsigned int low = state_len - unknowns;
  const unsigned long long starts = (len >> low) + 1ull;
  hipLaunchKernelGGL(k_mem_find, dim3(looped_blocks((starts + 63ull) & ~63ull, 256)), dim3(256), 0, s, words, start,
                     len, state_len, unknowns, best);
  return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, unsigned long long v,
                           unsigned long long count) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(looped_blocks(count, 256)), dim3(256), 0, s, p, v, count);
  return hipGetLastError();
}

hipError_t launch_batch_fold(hipStream_t s, const BatchFoldDesc* descs, unsigned int n,
                             const unsigned long long* chunk_out) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_batch_fold, dim3(n), dim3(256), 0, s, descs, chunk_out);
  return hipGetLastError();
}

hipError_t launch_fill_strided_u64(hipStream_t s, unsigned long long* p, unsigned long long v,
                                   unsigned long long count, unsigned long long stride) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_fill_strided_u64, dim3(looped_blocks(count, 256)), dim3(256), 0, s, p, v, count, stride);
  return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t s, unsigned int* p, unsigned int v, unsigned long long count) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(looped_blocks(count, 256)), dim3(256), 0, s, p, v, count);
  return hipGetLastError();
}

}  // namespace ns
