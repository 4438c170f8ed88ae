// Circuit-independent kernels: QSAT reduction passes over bit arrays, the
// next-state table's pointer-jumping passes and the unknown-start query.
#include <hip/hip_runtime.h>

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
// Entry layout (u64): bits 0..n-1 next state, bit n STOP, bits 32..63 number of
// circuit evaluations the entry stands for. A stopped entry is final; any other
// entry i is replaced by T[next(i)] with the step counts added, so after pass p
// every entry stands for exactly 2^p evaluations or for the path up to its STOP.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_jump_first(const unsigned int* __restrict__ t0,
                                                    unsigned long long* __restrict__ dst,
                                                    unsigned long long count, unsigned int n) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const unsigned int e = t0[i];
  const unsigned int stop_bit = 1u << n;
  unsigned long long r;
  if (e & stop_bit) {
    r = (unsigned long long)e | (1ull << 32);
  } else {
    const unsigned int f = t0[e & (stop_bit - 1u)];
    r = (unsigned long long)f | (2ull << 32);
  }
  dst[i] = r;
}

__global__ void __launch_bounds__(256) k_jump_pass(const unsigned long long* __restrict__ src,
                                                   unsigned long long* __restrict__ dst,
                                                   unsigned long long count, unsigned int n) {
  const unsigned long long base = ((unsigned long long)blockIdx.x * blockDim.x + threadIdx.x) * 2ull;
  if (base >= count) return;
  const unsigned int stop_bit = 1u << n;
  // two entries per thread: two independent gathers in flight
  const ulonglong2 e2 = *reinterpret_cast<const ulonglong2*>(src + base);
  unsigned long long e[2] = {e2.x, e2.y};
  unsigned long long f[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const unsigned int lo = (unsigned int)e[q];
    f[q] = (lo & stop_bit) ? 0ull : src[lo & (stop_bit - 1u)];
  }
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const unsigned int lo = (unsigned int)e[q];
    if (!(lo & stop_bit))
      e[q] = (f[q] & 0xffffffffull) | (((e[q] >> 32) + (f[q] >> 32)) << 32);
  }
  *reinterpret_cast<ulonglong2*>(dst + base) = make_ulonglong2(e[0], e[1]);
}

// Per unknown start u (state u << (n - unknowns)), with tab = the table after P passes
// (a non-stopped entry covers 2^P evaluations; a stopped one holds the evaluations until
// STOP and the end state): follows up to `jumps` entries from the start, so the walk covers
// jumps x 2^P evaluations. steps_out[u] = evaluations until STOP (0 when it does not stop
// within them), ends_out[u] = state produced by the stopping evaluation; stats[0] +=
// stopped starts, stats[1] = min(lowest stopped u).
__global__ void __launch_bounds__(256) k_starts(const unsigned long long* __restrict__ tab,
                                                unsigned int n, unsigned int unknowns,
                                                unsigned long long jumps, unsigned int* steps_out,
                                                unsigned int* ends_out, unsigned long long* stats) {
  const unsigned long long u = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cnt = 0, best = ~0ull;
  if (u < (1ull << unknowns)) {
    const unsigned int mask = (1u << n) - 1u;
    unsigned int x = (unsigned int)(u << (n - unknowns));
    unsigned long long total = 0;
    bool stop = false;
    unsigned int end = 0;
    for (unsigned long long j = 0; j < jumps; j++) {
      const unsigned long long e = tab[x];
      const unsigned int lo = (unsigned int)e;
      total += e >> 32;
      if ((lo >> n) & 1u) {
        stop = true;
        end = lo & mask;
        break;
      }
      x = lo & mask;
    }
    if (steps_out) steps_out[u] = stop ? (unsigned int)total : 0u;
    if (ends_out) ends_out[u] = stop ? end : 0u;
    if (stop) {
      cnt = 1;
      best = u;
    }
  }
  wave_stats(cnt, best, stats);
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, unsigned long long count) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

__global__ void k_fill_u32(unsigned int* p, unsigned int v, unsigned long long count) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline unsigned int blocks_for(unsigned long long threads, unsigned int bs) {
  return (unsigned int)((threads + bs - 1) / bs);
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
  hipLaunchKernelGGL(k_jump_first, dim3(blocks_for(count, 256)), dim3(256), 0, s, t0, dst, count, n);
  return hipGetLastError();
}

hipError_t launch_jump_pass(hipStream_t s, const unsigned long long* src, unsigned long long* dst,
                            unsigned long long count, unsigned int n) {
  hipLaunchKernelGGL(k_jump_pass, dim3(blocks_for(count / 2, 256)), dim3(256), 0, s, src, dst, count, n);
  return hipGetLastError();
}

hipError_t launch_starts(hipStream_t s, const unsigned long long* tab, unsigned int n,
                         unsigned int unknowns, unsigned long long jumps, unsigned int* steps_out,
                         unsigned int* ends_out, unsigned long long* stats) {
  unsigned long long cnt = 1ull << unknowns;
  unsigned long long threads = (cnt + 63ull) & ~63ull;
  hipLaunchKernelGGL(k_starts, dim3(blocks_for(threads, 256)), dim3(256), 0, s, tab, n, unknowns, jumps,
                     steps_out, ends_out, stats);
  return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, unsigned long long v,
                           unsigned long long count) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(count, 256)), dim3(256), 0, s, p, v, count);
  return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t s, unsigned int* p, unsigned int v, unsigned long long count) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(count, 256)), dim3(256), 0, s, p, v, count);
  return hipGetLastError();
}

}  // namespace ns
