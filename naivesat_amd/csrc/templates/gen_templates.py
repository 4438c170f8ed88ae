"""Writes templates.hip: the circuit-independent kernels of the native code path.

Each kernel evaluates a circuit through an inline-assembly placeholder whose register
contract is fixed; at run time the engine writes the circuit's register-allocated
instructions into the placeholder of the compiled assembly and assembles it.

Register contract of every placeholder ("; NS_BODY <name>"):
  inputs   v[LM..LM+5]   lane masks: v[LM+i] = 0 - ((lane >> i) & 1)
           s20, s21      wave-uniform part of the word index (lo, hi): every index bit except
                         the six lane bits
  outputs  v0 .. v[OUT-1] circuit outputs for the current word
  scratch  v[POOL_LO..POOL_HI], s[40..101], scc (clobbered); the SAT kernels, whose own code
           keeps to s2 across the body, also give it s0, s1, s[3..19] and s[22..39], and vcc (their
           r64 / r128 pools end one register lower: with vcc clobbered their own code takes one
           more)
Register classes bound the VGPR budget (and with it the occupancy).

Usage: python gen_templates.py OUT.hip
"""
import sys

CLASSES = {  # name: (total VGPR budget, pool high register); r512 adds a0..a255 as spill slots
    "r64": (64, 55),
    "r128": (128, 119),
    "r256": (256, 239),
    "r512": (512, 239),
}
SGPR_POOL = list(range(40, 102))
# the QSAT-B templates' own code needs two VGPRs more than the others: their bodies' pool ends
# two registers lower, so that the kernels stay within their class (64 / 128 VGPRs)
QSATB_POOL_CUT = 2
SAT_SGPR_POOL = [0, 1] + list(range(3, 20)) + list(range(22, 102))  # s20, s21: the word index
# the SAT bodies may write vcc (group exits); with vcc clobbered the r64 / r128 SAT kernels'
# own code takes one VGPR more, so their bodies' pool ends one register lower
SAT_VCC_POOL_CUT = 1


def clobbers(lo, hi, agprs=False, sgprs=SGPR_POOL, vcc=False):
    regs = [f'"v{i}"' for i in range(lo, hi + 1)]
    if agprs:
        regs += [f'"a{i}"' for i in range(256)]
    regs += [f'"s{i}"' for i in sgprs]
    regs.append('"scc"')
    if vcc:  # the group exits of the SAT bodies compare into vcc
        regs.append('"vcc"')
    return ", ".join(regs)


HEAD = r'''// generated by gen_templates.py - circuit-independent kernels of the native code path
#include <hip/hip_runtime.h>
typedef unsigned int u32;
typedef unsigned long long u64;

// lane masks and wave-uniform word base, as the circuit bodies expect them
#define NS_PROLOGUE                                                              \
  const u32 lane = threadIdx.x & 63u;                                             \
  const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);              \
  const u64 U = w0 + ((u64)blockIdx.x * (blockDim.x >> 6) + wave) * (64ull * wpl); \
  const u32 LM0 = 0u - (lane & 1u), LM1 = 0u - ((lane >> 1) & 1u);                \
  const u32 LM2 = 0u - ((lane >> 2) & 1u), LM3 = 0u - ((lane >> 3) & 1u);         \
  const u32 LM4 = 0u - ((lane >> 4) & 1u), LM5 = 0u - ((lane >> 5) & 1u);

// Wave totals into stats: the count goes to one of 64 slots 128 bytes apart from stats[512]
// on (by block; atomics on one word serialize, and a search of many hits would queue on
// it), the lowest hit to stats[1], skipped when that is already lower.
__device__ inline void ns_wave_stats(u64 cnt, u64 best, u64* stats) {
  if (__ballot(cnt != 0ull) == 0ull) return;
  for (int off = 32; off > 0; off >>= 1) {
    u64 ob = __shfl_xor(best, off);
    best = ob < best ? ob : best;
    cnt += __shfl_xor(cnt, off);
  }
  if ((threadIdx.x & 63u) == 0u) {
    atomicAdd(&stats[512u + ((blockIdx.x & 63u) << 4)], cnt);
    if (best < __hip_atomic_load(&stats[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(&stats[1], best);
  }
}

// a wave-uniform value copied into a VGPR (the compiler would keep it in an SGPR)
__device__ inline u32 ns_to_vgpr(u32 x) {
  u32 r;
  asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
  return r;
}

__device__ inline u32 ns_red(u32 x, u32 sh, u32 all) { return all ? (x & (x >> sh)) : (x | (x >> sh)); }

// Exclusive prefix sum over the wave's lanes of r (0..63) and, in *total, the wave's sum: one
// ballot per bit of r, each counted below the lane (mbcnt) and in all (no lane shuffles, no
// per-lane address registers)
__device__ inline u32 ns_wave_excl(u32 r, u32* total) {
  u32 ex = 0u, tot = 0u;
#pragma unroll
  for (u32 j = 0; j < 6u; j++) {
    const u64 m = __ballot((r >> j) & 1u);
    ex += __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u)) << j;
    tot += (u32)__popcll(m) << j;
  }
  *total = tot;
  return ex;
}
'''


def body1(name, lm, lo, hi, sgprs=SGPR_POOL, vcc=False):
    """placeholder with one output in v0"""
    ins = ", ".join([f'"{{v{lm + i}}}"(LM{i})' for i in range(6)] + ['"{s20}"(wlo)', '"{s21}"(whi)'])
    agprs = name.endswith("r512")
    return (f'    u32 o;\n'
            f'    asm volatile("; NS_BODY {name}" : "={{v0}}"(o) : {ins} : {clobbers(lo, hi, agprs, sgprs, vcc)});\n')


def sat_kernel(cls, vb, pool_hi):
    name = f"ns_nat_sat_{cls}"
    return f'''
// SAT: res[0] += satisfying assignments, res[1] = min(lowest); w = U + lane + (j << 6).
// Across the body the loop keeps one scalar (its counter, in s2) besides the word index in
// s20/s21, so that the body has the other SGPRs: the word limit is a per-lane bound on j, the
// valid mask a VGPR, and the values used after the loop wait in lanes of one VGPR.
extern "C" __global__ void __launch_bounds__(256) {name}(u64 w0, u64 wend, u64* res, u32 wpl, u32 valid) {{
  NS_PROLOGUE
  u32 cnt = 0u;
  u32 brel = ~0u;  // lowest hit as ((word - U) << 5) | bit: one VGPR across the body
  // words U + (j << 6) + lane at or past wend are not counted: j < jlim for this lane
  const u64 first = U + lane;
  const u64 left = wend > first ? wend - first : 0ull;
  const u32 jlim = left >= ((u64)wpl << 6) ? wpl : (u32)((left + 63ull) >> 6);
  u32 jlimrel = jlim << 11, jrel = lane << 5;  // (j << 11) + (lane << 5) against jlim << 11
  // the valid mask in a VGPR, and the loop counter as a scalar of its own (a copy through a
  // VGPR: a part of the kernel-argument registers would keep all of them live across the body)
  const u32 vvalid = ns_to_vgpr(valid);
  // U and res in lanes 0..3 of one VGPR across the loop, read back after it
  u32 keep = 0u;
  asm volatile("v_writelane_b32 %0, %1, 0\\n\\tv_writelane_b32 %0, %2, 1\\n\\t"
               "v_writelane_b32 %0, %3, 2\\n\\tv_writelane_b32 %0, %4, 3"
               : "+v"(keep) : "s"((u32)U), "s"((u32)(U >> 32)), "s"((u32)(u64)res), "s"((u32)((u64)res >> 32)));
  u32 wlo = (u32)U, whi = (u32)(U >> 32);
#pragma unroll 1
  for (u32 jj = __builtin_amdgcn_readfirstlane(ns_to_vgpr(wpl)); jj != 0u; jj--) {{
{body1(name, 1, 8, pool_hi if vb >= 256 else pool_hi - SAT_VCC_POOL_CUT, SAT_SGPR_POOL, vcc=True)}    o &= vvalid;
    if (jrel >= jlimrel) o = 0u;
    cnt += __popc(o);
    if (o != 0u && brel == ~0u) brel = jrel | (u32)__builtin_ctz(o);
    jrel += 2048u;
    const u64 wn = (((u64)whi << 32) | wlo) + 64ull;
    wlo = (u32)wn;
    whi = (u32)(wn >> 32);
  }}
  // (readlane returns int: each half goes through u32, so that the low half is not sign-extended)
  const u64 Ub = ((u64)(u32)__builtin_amdgcn_readlane(keep, 1u) << 32) | (u32)__builtin_amdgcn_readlane(keep, 0u);
  u64* resb = (u64*)(((u64)(u32)__builtin_amdgcn_readlane(keep, 3u) << 32) | (u32)__builtin_amdgcn_readlane(keep, 2u));
  ns_wave_stats(cnt, brel == ~0u ? ~0ull : (Ub << 5) + brel, resb);
}}
'''


def qsat_a_kernel(cls, pool_hi):
    name = f"ns_nat_qsata_{cls}"
    return f'''
// QSAT, one reduced bit per thread (all in-thread levels below the outermost block):
// w = U + (lane << jb) + j; opm bit l = AND for level l (level 0 = index bit 0).
extern "C" __global__ void __launch_bounds__(256) {name}(u64 w0, u64 wend, u64* bits, u32 wpl, u32 jb, u32 opm) {{
  NS_PROLOGUE
  const u64 wl = U + ((u64)lane << jb);
  u32 R = 0u;
#pragma unroll 1
  for (u32 j = 0; j < wpl; j++) {{
    const u64 wu = U + j;
    const u32 wlo = (u32)wu, whi = (u32)(wu >> 32);
{body1(name, 1, 8, pool_hi)}#pragma unroll
    for (u32 l = 0; l < 5; l++) o = ns_red(o, 1u << l, (opm >> l) & 1u);
    R |= (o & 1u) << j;
  }}
  for (u32 l = 0; l < jb; l++) R = ns_red(R, 1u << l, (opm >> (5 + l)) & 1u);
  const u32 nvalid = wend > wl ? 1u : 0u;
  const u64 ball = __ballot((R & 1u) != 0u && nvalid != 0u);
  if (lane == 0u) bits[U >> (jb + 6)] = ball;
}}
'''


def qsat_b_kernel(cls, pool_hi):
    name = f"ns_nat_qsatb_{cls}"
    return f'''
// QSAT, outermost block reached inside the thread (L < 5 + jb levels below it):
// stats[0] += witnesses, stats[1] = min(lowest witness u), bitmap bit u = block value.
extern "C" __global__ void __launch_bounds__(256) {name}(u64 w0, u64 wend, u64* stats, u32* bitmap, u32 wpl,
                                                      u32 jb, u32 opm, u32 L, u32 q0all, u32 valid) {{
  NS_PROLOGUE
  const u64 wl = U + ((u64)lane << jb);
  u32 cnt = 0u;
  u64 best = ~0ull;
  u32 R = 0u;
  const u32 inword = L < 5u ? L : 5u;
#pragma unroll 1
  for (u32 j = 0; j < wpl; j++) {{
    const u64 wu = U + j;
    const u64 w = wl + j;
    const u32 wlo = (u32)wu, whi = (u32)(wu >> 32);
{body1(name, 1, 8, pool_hi)}    for (u32 l = 0; l < inword; l++) o = ns_red(o, 1u << l, (opm >> l) & 1u);
    if (L <= 5u) {{
      u32 pos = 0u;
      for (u32 p = 0; p < (32u >> L); p++) pos |= 1u << (p << L);
      pos &= valid;
      u32 wit = (q0all ? ~o : o) & pos;
      if (w >= wend) wit = 0u;
      cnt += __popc(wit);
      if (wit != 0u && best == ~0ull) best = (w << (5u - L)) + (u64)(__builtin_ctz(wit) >> L);
      if (bitmap != 0 && w < wend) {{
        u32 r = 0u;
        for (u32 p = 0; p < (32u >> L); p++) r |= ((o >> (p << L)) & 1u) << p;
        r &= (L == 0u) ? valid : 0xffffffffu;
        const u64 ub = w << (5u - L);
        atomicOr(&bitmap[ub >> 5], r << (u32)(ub & 31ull));
      }}
    }} else {{
      R |= (o & 1u) << j;
    }}
  }}
  if (L > 5u) {{
    const u32 s = L - 5u;
    for (u32 l = 0; l < s; l++) R = ns_red(R, 1u << l, (opm >> (5u + l)) & 1u);
    const u32 nvalid = wend > wl ? (wend - wl >= wpl ? wpl : (u32)(wend - wl)) : 0u;
    u32 pos = 0u;
    for (u32 p = 0; p < (wpl >> s); p++) pos |= 1u << (p << s);
    const u32 jmask = nvalid >= 32u ? 0xffffffffu : ((1u << nvalid) - 1u);
    const u32 wit = (q0all ? ~R : R) & pos & jmask;
    cnt = __popc(wit);
    best = wit != 0u ? (wl >> s) + (u64)(__builtin_ctz(wit) >> s) : ~0ull;
    if (bitmap != 0 && nvalid != 0u) {{
      u32 r = 0u;
      for (u32 p = 0; p < (wpl >> s); p++) r |= ((R >> (p << s)) & 1u) << p;
      r &= (nvalid >= wpl) ? 0xffffffffu : ((1u << ((nvalid + (1u << s) - 1u) >> s)) - 1u);
      const u64 ub = wl >> s;
      atomicOr(&bitmap[ub >> 5], r << (u32)(ub & 31ull));
    }}
  }}
  ns_wave_stats(cnt, best, stats);
}}
'''


def table_kernel(cls, pool_hi):
    name = f"ns_nat_table_{cls}"
    outs = ", ".join(f'"={{v{k}}}"(A[{k}])' for k in range(32))
    ins = ", ".join([f'"{{v{32 + i}}}"(LM{i})' for i in range(6)] + ['"{s20}"(wlo)', '"{s21}"(whi)'])
    return f'''
// Next-state table: table[32(w - w0) + b] = outputs of state 32w + b packed (output i at
// bit i), so a launch over words [w0, wend) fills its own slice;
// *anylast |= any state whose last output (STOP) is 1. w = U + lane + (j << 6).
extern "C" __global__ void __launch_bounds__(256) {name}(u64 w0, u64 wend, u32* table, u32* anylast, u32 wpl,
                                                      u32 last, u32 valid, u32 small) {{
  NS_PROLOGUE
  u32 last_any = 0u;
#pragma unroll 1
  for (u32 j = 0; j < wpl; j++) {{
    const u64 wu = U + ((u64)j << 6);
    const u64 w = wu + lane;
    const u32 wlo = (u32)wu, whi = (u32)(wu >> 32);
    u32 A[32];
    asm volatile("; NS_BODY {name}" : {outs} : {ins} : {clobbers(40, pool_hi, cls == "r512")});
    if (w < wend) last_any |= A[last] & valid;
    {{
      u32 msk = 0x0000FFFFu;
#pragma unroll
      for (int s = 16; s != 0; s >>= 1, msk ^= (msk << s)) {{
#pragma unroll
        for (int k = 0; k < 32; k = ((k | s) + 1) & ~s) {{
          const u32 t = ((A[k] >> s) ^ A[k | s]) & msk;
          A[k] ^= t << s;
          A[k | s] ^= t;
        }}
      }}
    }}
    if (w < wend) {{
      if (small == 0u) {{
        uint4* dst = (uint4*)(table + ((w - w0) << 5));
#pragma unroll
        for (int q = 0; q < 8; q++) dst[q] = make_uint4(A[4 * q], A[4 * q + 1], A[4 * q + 2], A[4 * q + 3]);
      }} else {{
        for (u32 b = 0; b < small; b++) table[b] = A[b];
      }}
    }}
  }}
  // one flag write per wave, none once it is set (atomics on one word serialize)
  const u64 any_b = __ballot(last_any != 0u);
  if (any_b != 0ull && lane == (u32)(__ffsll((long long)any_b) - 1) &&
      __hip_atomic_load(anylast, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
    atomicOr(anylast, 1u);
}}
'''


BATCH_HEAD = r'''
// Work descriptor of one request in a batch (64 bytes): its circuit body is a subroutine in
// the batch's code object, entered at pc base + entry with the return address in s[38:39];
// the body may use s48..s101 (and the class's VGPR pool), the caller keeps s0..s35, s40..s47.
struct NsBatchDesc {
  u32 wend;      // words of the request's space; words at or past it are not counted
  u32 chunk0;    // first chunk of the request in the batch's chunk numbering
  u64* out;      // SAT: its result {count, lowest}; QSAT: its bit array
  u32 entry;     // byte offset of the body from its kernel's pc base
  u32 words;     // words per chunk (a multiple of 64 << jb)
  u32 jb;        // QSAT: in-thread levels above the word
  u32 opm;       // QSAT: bit l set = AND at level l (level 0 = index bit 0)
  u32 valid;     // valid assignments of a word (spaces under 32 assignments)
  u32 pad[7];
};

// Wave totals of a batch chunk: sum of cnt and min of rel over the wave, wave-uniform. The
// reduction swizzles within each half-wave (XOR pattern as an immediate: no address
// registers) and joins the halves by readlane; no branch, so the batch loops stay uniform.
#define NS_SWZ(v, m) ((u32)__builtin_amdgcn_ds_swizzle((int)(v), ((m) << 10) | 0x1f))
__device__ inline void ns_chunk_totals(u32 cnt, u32 rel, u32* c_out, u32* r_out) {
  cnt += NS_SWZ(cnt, 16);
  { const u32 o = NS_SWZ(rel, 16); rel = o < rel ? o : rel; }
  cnt += NS_SWZ(cnt, 8);
  { const u32 o = NS_SWZ(rel, 8); rel = o < rel ? o : rel; }
  cnt += NS_SWZ(cnt, 4);
  { const u32 o = NS_SWZ(rel, 4); rel = o < rel ? o : rel; }
  cnt += NS_SWZ(cnt, 2);
  { const u32 o = NS_SWZ(rel, 2); rel = o < rel ? o : rel; }
  cnt += NS_SWZ(cnt, 1);
  { const u32 o = NS_SWZ(rel, 1); rel = o < rel ? o : rel; }
  *c_out = __builtin_amdgcn_readlane(cnt, 0) + __builtin_amdgcn_readlane(cnt, 32);
  const u32 r0 = __builtin_amdgcn_readlane(rel, 0), r1 = __builtin_amdgcn_readlane(rel, 32);
  *r_out = r0 < r1 ? r0 : r1;
}

// a wave-uniform 64-bit value in SGPRs (readfirstlane returns int: each half goes through
// u32 so that the low half is not sign-extended into the high one)
__device__ inline u64 ns_uni64(u64 x) {
  const u32 hi = (u32)__builtin_amdgcn_readfirstlane((u32)(x >> 32));
  const u32 lo = (u32)__builtin_amdgcn_readfirstlane((u32)x);
  return ((u64)hi << 32) | (u64)lo;
}
'''


BATCH_SGPRS = [36, 37, 38, 39] + list(range(48, 102))  # call target, return address, body pool


def batch_call(cls, lo, hi):
    """computed call of the descriptor's body: s[36:37] = pc base + entry, return address in s[38:39]"""
    ins = ", ".join([f'"{{v{1 + i}}}"(LM{i})' for i in range(6)] +
                    ['"{s20}"(wlo)', '"{s21}"(whi)', '[pl] "s"(pclo)', '[ph] "s"(pchi)', '[en] "s"(entry)'])
    clob = clobbers(lo, hi, cls == "r512", BATCH_SGPRS)
    return (f'          u32 o;\n'
            f'          asm volatile("s_add_u32 s36, %[pl], %[en]\\n\\ts_addc_u32 s37, %[ph], 0\\n\\t'
            f's_swappc_b64 s[38:39], s[36:37]" : "={{v0}}"(o) : {ins} : {clob});\n')


def batch_kernels(cls, pool_hi, ci):
    common = """  const u32 lane = threadIdx.x & 63u;
  const u32 LM0 = 0u - (lane & 1u), LM1 = 0u - ((lane >> 1) & 1u);
  const u32 LM2 = 0u - ((lane >> 2) & 1u), LM3 = 0u - ((lane >> 3) & 1u);
  const u32 LM4 = 0u - ((lane >> 4) & 1u), LM5 = 0u - ((lane >> 5) & 1u);
  const u32 whi = 0u;
  const u32 waves = gridDim.x * (blockDim.x >> 6);
  const u32 wave0 = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
#pragma unroll 1
  for (u32 c = wave0; c < chunks; c += waves) {
    u32 lo = 0u, hi = nd;
#pragma unroll 1
    while (hi - lo > 1u) {
      const u32 mid = (lo + hi) >> 1;
      if ((u32)__builtin_amdgcn_readfirstlane(descs[mid].chunk0) <= c) lo = mid; else hi = mid;
    }
    // every descriptor field is wave-uniform: read into SGPRs
    const NsBatchDesc* d = descs + lo;
    const u32 wend = __builtin_amdgcn_readfirstlane(d->wend);
    const u32 first = (c - __builtin_amdgcn_readfirstlane(d->chunk0)) * __builtin_amdgcn_readfirstlane(d->words);
    const u32 last = min(first + __builtin_amdgcn_readfirstlane(d->words), wend);
    u64* out = (u64*)ns_uni64((u64)d->out);
    const u32 entry = __builtin_amdgcn_readfirstlane(d->entry);
"""
    sat = f"""
// Persistent SAT batch kernel. Wave w of the grid takes chunks w, w + W, w + 2W, ... (W =
// waves in the grid); a chunk is `words` consecutive words of one request (found by binary
// search over chunk0), each 64-word step running that request's body with the lane bits at
// index bits 5..10. The chunk's count and lowest hit go to chunk_out[2c], [2c + 1] (stored
// by every lane: the same value at the same address), folded per request afterwards. The
// loops hold no divergent branch: every condition is wave-uniform. The bodies are
// subroutines in the code object's batch region; the pc base is the address of the marker
// instruction that follows s_getpc_b64. Word indices are 32-bit (at most 2^36 assignments).
extern "C" __global__ void __launch_bounds__(256) ns_nat_bsat_{cls}(const NsBatchDesc* descs, u32 nd, u32 chunks,
                                                      u64* chunk_out) {{
  u64 pcb;
  asm volatile("s_getpc_b64 %0\\n\\ts_mov_b32 s36, 0x{0x534253c0 + ci:08x}" : "=s"(pcb) : : "s36");
  const u32 pclo = __builtin_amdgcn_readfirstlane((u32)pcb), pchi = __builtin_amdgcn_readfirstlane((u32)(pcb >> 32));
{common}    const u32 valid = __builtin_amdgcn_readfirstlane(d->valid);
    u32 cnt = 0u;
    u32 brel = ~0u;  // lowest hit of the chunk as ((word - first) << 5) | bit
#pragma unroll 1
    for (u32 w = first; w < last; w += 64u) {{
      const u32 wlo = __builtin_amdgcn_readfirstlane(w);
      const u32 rem = last - wlo;  // words wlo + lane with lane >= rem are not counted
{batch_call(cls, 8, pool_hi)}      o &= valid;
      o = lane < rem ? o : 0u;
      cnt += __popc(o);
      const u32 hit = ((wlo - first + lane) << 5) | (u32)__builtin_ctz(o | 0x80000000u);
      brel = (o != 0u && brel == ~0u) ? hit : brel;
    }}
    u32 tc, tr;
    ns_chunk_totals(cnt, brel, &tc, &tr);
    chunk_out[2ull * c] = (u64)tc;
    chunk_out[2ull * c + 1ull] = tr == ~0u ? ~0ull : ((u64)first << 5) + tr;
  }}
}}
"""
    qsat = f"""
// Persistent QSAT batch kernel (one reduced bit per thread): chunks as in the SAT one, with
// the lane bits at index bits 5+jb..10+jb; a 64 x 2^jb-word tile gives one ballot word of
// the request's bit array (bit array index = tile index), stored by every lane.
extern "C" __global__ void __launch_bounds__(256) ns_nat_bqsat_{cls}(const NsBatchDesc* descs, u32 nd, u32 chunks) {{
  u64 pcb;
  asm volatile("s_getpc_b64 %0\\n\\ts_mov_b32 s36, 0x{0x514253c0 + ci:08x}" : "=s"(pcb) : : "s36");
  const u32 pclo = __builtin_amdgcn_readfirstlane((u32)pcb), pchi = __builtin_amdgcn_readfirstlane((u32)(pcb >> 32));
{common}    const u32 jb = __builtin_amdgcn_readfirstlane(d->jb), opm = __builtin_amdgcn_readfirstlane(d->opm);
    const u32 tw = 64u << jb;
#pragma unroll 1
    for (u32 U = first; U < last; U += tw) {{
      u32 R = 0u;
#pragma unroll 1
      for (u32 j = 0; j < (1u << jb); j++) {{
        const u32 wlo = __builtin_amdgcn_readfirstlane(U + j);
{batch_call(cls, 8, pool_hi)}#pragma unroll
        for (u32 l = 0; l < 5; l++) o = ns_red(o, 1u << l, (opm >> l) & 1u);
        R |= (o & 1u) << j;
      }}
      for (u32 l = 0; l < jb; l++) R = ns_red(R, 1u << l, (opm >> (5 + l)) & 1u);
      out[U >> (jb + 6)] = __ballot((R & 1u) != 0u && wend - U > (lane << jb));
    }}
  }}
}}
"""
    return sat + qsat


# Walks without a table (SATOSTOP): the walk bodies hold no wave-uniform value, so they write
# no SGPR and the template keeps its own scalars anywhere across the body. State tiers (the template holds T state bits per walk
# slot; circuits of n <= T state bits leave the bits above n at zero) and the register classes
# each tier is built for. The 64 KiB of slot records in LDS per block already hold a CU to two
# blocks (two waves per SIMD), so no class below r256 is built.
WALK_TIERS = {16: ("r256", "r512"), 24: ("r256", "r512"), 32: ("r256", "r512"), 40: ("r256", "r512"),
              48: ("r256", "r512")}
WALK_MASKS = ("active", "fin", "fstop", "fwp", "floop", "a0", "a1", "a2", "a3")


def walk_regs(T):
    """fixed registers of the walk body: s (state), f (final state), v (saved state), masks"""
    base = 8
    return {"s": base, "f": base + T, "v": base + 2 * T, "m": base + 3 * T, "pool_lo": base + 3 * T + len(WALK_MASKS)}


WALK_CHUNK = 2048  # waypoints a wave takes from the queue at once (one full refill of its slots)
WALK_BULK = 384  # refills or ended walks per wave from which the transposed path is taken


def t32(arr):
    """32 x 32 bit transpose of arr[0..31] (word k bit b <-> word b bit k) in registers"""
    return (f"    {{\n      u32 msk = 0x0000FFFFu;\n#pragma unroll\n      for (int sh = 16; sh != 0; sh >>= 1, msk ^= (msk << sh)) {{\n"
            f"#pragma unroll\n        for (int q = 0; q < 32; q = ((q | sh) + 1) & ~sh) {{\n"
            f"          const u32 tt = (({arr}[q] >> sh) ^ {arr}[q | sh]) & msk;\n"
            f"          {arr}[q] ^= tt << sh;\n          {arr}[q | sh] ^= tt;\n        }}\n      }}\n    }}\n")


def walk_kernel(T, cls, pool_hi):
    name = f"ns_nat_walk{T}_{cls}"
    R = walk_regs(T)

    def each(fmt, lo=0, hi=None):
        return "".join(fmt.replace("@", str(i)) for i in range(lo, T if hi is None else hi))

    ops = ([f'"+{{v{R["s"] + i}}}"(s{i})' for i in range(T)] + [f'"+{{v{R["f"] + i}}}"(f{i})' for i in range(T)] +
           [f'"+{{v{R["v"] + i}}}"(v{i})' for i in range(T)] +
           [f'"+{{v{R["m"] + j}}}"({m})' for j, m in enumerate(WALK_MASKS)])
    wide = T > 32
    # Refill of the slots in `need` from the wave's waypoint range [lb, le), then its next
    # chunk nb (taken from the prefetch or the queue): rank q of the wave's refills (lanes in
    # order, a lane's slots in bit order) gets waypoint q < avail ? lb + q : nb + q - avail.
    refill = (
        "    if (!exhausted) {\n"
        "      u32 total;\n"
        "      const u32 q0 = ns_wave_excl(__popc(need), &total);\n"
        "      if (total != 0u) {\n"
        "        const u32 avail = le - lb;\n"
        "        u32 nb = 0u;\n"
        "        if (total > avail) {\n"
        "          if (have_pf) {\n"
        "            // the prefetched chunk is read here, not at its atomic: the wait for its\n"
        "            // result stays here too\n"
        "            asm volatile(\"\" : \"+v\"(pf));\n"
        "            nb = __builtin_amdgcn_readfirstlane(pf);\n"
        "            have_pf = false;\n"
        "          } else {\n"
        "            u32 b0 = 0u;\n"
        f"            if (lane == 0u) b0 = (u32)atomicAdd(queue, {WALK_CHUNK}ull);\n"
        "            nb = __builtin_amdgcn_readfirstlane(b0);\n"
        "          }\n"
        "        }\n"
        f"        if (total < {WALK_BULK}u) {{\n"
        "          u32 q = q0;\n"
        "          for (u32 rem = need; rem; rem &= rem - 1u) {\n"
        "            const u32 b = __builtin_ctz(rem), mb = 1u << b;\n"
        "            const u64 idx = q < avail ? (u64)(lb + q) : (u64)nb + (q - avail);\n"
        "            q++;\n"
        "            if (idx >= count) break;\n"
        "            const u64 st = idx << k;\n"
        + each("            s@ = (s@ & ~mb) | ((u32)((st >> @) & 1ull) * mb);\n            v@ = (v@ & ~mb) | (s@ & mb);\n") +
        "            meta[b * 256u + tid] = (u64)t | (idx << 32);\n"
        "            active |= mb;\n"
        "          }\n"
        "        } else {\n"
        "          // many refills: the new states of all 32 slots at once, by bit transposes\n"
        "          u32 X[32];\n"
        + ("          u32 XH[32];\n" if wide else "") +
        "          u32 vm = 0u, j = 0u;\n"
        "#pragma unroll\n"
        "          for (u32 b = 0; b < 32u; b++) {\n"
        "            const u32 hb = (need >> b) & 1u;\n"
        "            const u32 q = q0 + j;\n"
        "            const u64 idx = q < avail ? (u64)(lb + q) : (u64)nb + (q - avail);\n"
        "            const bool ok = hb != 0u && idx < count;\n"
        "            const u64 st = ok ? idx << k : 0ull;\n"
        "            X[b] = (u32)st;\n"
        + ("            XH[b] = (u32)(st >> 32);\n" if wide else "") +
        "            if (ok) meta[b * 256u + tid] = (u64)t | (idx << 32);\n"
        "            vm |= (ok ? 1u : 0u) << b;\n"
        "            j += hb;\n"
        "          }\n"
        + t32("X") + (t32("XH") if wide else "")
        + each("          s@ = (s@ & ~vm) | (X[@] & vm);\n          v@ = (v@ & ~vm) | (s@ & vm);\n", 0, min(T, 32))
        + (each("          s@ = (s@ & ~vm) | (XH[@ - 32] & vm);\n          v@ = (v@ & ~vm) | (s@ & vm);\n", 32, T)
           if wide else "") +
        "          active |= vm;\n"
        "        }\n"
        "        if (total > avail) {\n"
        "          lb = nb + (total - avail);\n"
        f"          le = nb + {WALK_CHUNK}u;\n"
        "        } else {\n"
        "          lb += total;\n"
        "        }\n"
        "        if ((u64)le > count) le = (u32)count;\n"
        "        if ((u64)lb >= count) exhausted = true;\n"
        "        // the next chunk is asked for one refill ahead: its atomic's latency overlaps\n"
        "        // the steps until it is used\n"
        f"        if (!exhausted && !have_pf && le - lb < {WALK_CHUNK}u) {{\n"
        f"          if (lane == 0u) pf = (u32)atomicAdd(queue, {WALK_CHUNK}ull);\n"
        "          have_pf = true;\n"
        "        }\n"
        "      }\n"
        "    }\n")
    # the records of the slots that ended in the window
    rec_one = (
        "        const u64 m = meta[b * 256u + tid];\n"
        "        const u32 age = ((a0 >> b) & 1u) | (((a1 >> b) & 1u) << 1) | (((a2 >> b) & 1u) << 2) | (((a3 >> b) & 1u) << 3);\n"
        "        const u32 steps = t - age - (u32)m;\n"
        "        const u32 kind = (fstop & mb) ? 1u : (fwp & mb) ? 0u : (floop & mb) ? 2u : 3u;\n"
        "        const u64 target = kind == 1u ? st : kind == 0u ? (st >> k) : 0ull;\n"
        "        out[m >> 32] = make_ulonglong2(target | ((u64)kind << 62), steps);\n"
        "        nrec++;\n")
    return f"""
// SATOSTOP walks without a table: 32 walk slots per lane, bit-sliced (bit b of s@ = state
// bit @ of slot b; up to {T} state bits, the bits above the circuit's n stay zero). A slot
// walks from its waypoint (a state whose low k bits are zero) until the next waypoint, STOP,
// a return to its saved state (Brent: saved after each power of two of its evaluations) or
// the limit, and writes out[waypoint] = (target | kind << 62, evaluations), kind 0 waypoint
// (target = its index), 1 STOP (target = the stopping state), 2 loop, 3 limit. Waypoints come
// from one counter (queue) in chunks of {WALK_CHUNK} per wave, the next chunk asked for one
// refill ahead, so every slot keeps taking walks while any are left. The body runs two walk
// steps with their bookkeeping (per slot: STOP, waypoint and loop tests, the final state,
// the evaluation ages of ended slots); every 16 steps this code records the ended walks,
// saves states and refills the slots (by bit transposes when a wave has many of them). The
// walks' evaluations go to evals[16 s] and their records to evals[16 s + 1] (64 slots s).
extern "C" __global__ void __launch_bounds__(256){" __attribute__((amdgpu_waves_per_eu(2, 2)))" if cls == "r256" else ""} {name}(u64 count, u64 limit, u64 budget, ulonglong2* __restrict__ out,
                                                      u32* long_flag, u64* evals, u64* queue, u32 k) {{
  __shared__ u64 meta[32 * 256];  // per slot: start time (bits 0..31), waypoint index (32..63)
  const u32 tid = threadIdx.x, lane = tid & 63u;
  const u64 lim = limit < 0x7fffffffull ? limit : 0x7fffffffull;
{each("  u32 s@ = 0u, f@ = 0u, v@ = 0u;" + chr(10))}  u32 active = 0u, fin = 0u, fstop = 0u, fwp = 0u, floop = 0u, t = 0u;
  u32 a0 = 0u, a1 = 0u, a2 = 0u, a3 = 0u;
  u64 used = 0ull, ev = 0ull;
  u32 nrec = 0u;
  // the wave's waypoint range [lb, le) (indices stay below 2^32: at most 2^28 waypoints, and the
  // queue passes them by at most one chunk per wave); pf: the prefetched chunk's base (lane 0)
  u32 lb = 0u, le = 0u, pf = 0u;
  bool have_pf = false, exhausted = false;
  {{
    const u32 need = 0xffffffffu;
{refill}  }}
  // every lane of the wave runs every iteration (the loop test is wave-uniform), so all take
  // part in the refills' prefix sums; a wave with waypoints left keeps stepping to its next
  // refill even with no walk running
  while (__any((active | fin) != 0u) || !exhausted) {{
    used += 2ull;
    if (used > budget) {{
      atomicOr(long_flag, 1u);
      break;
    }}
    if ((used & 255ull) == 0ull && __hip_atomic_load(long_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
      break;
    asm volatile("; NS_BODY {name}" : {", ".join(ops)} : : {clobbers(R["pool_lo"], pool_hi, cls == "r512", [])});
    t += 2u;
    if ((t & 15u) == 0u) {{
      // evaluations of the window: 16 per slot running at its start, less the steps since
      // each ended slot finished (its age)
      ev += 16ull * (u64)__popc(active | fin) -
            (u64)(__popc(a0) + 2u * __popc(a1) + 4u * __popc(a2) + 8u * __popc(a3));
      // saves and limits of the running slots: with most slots running, all 32 start times
      // are read at once (independent LDS reads, one wait) and the masks built without
      // branches; with few (the end of the search), slot by slot
      u32 save = 0u, lng = 0u;
      u32 nact;
      (void)ns_wave_excl(__popc(active), &nact);
      if (nact >= 1024u) {{
        const u32* mlo = reinterpret_cast<const u32*>(meta);
#pragma unroll
        for (u32 b = 0; b < 32u; b++) {{
          const u32 steps = t - mlo[2u * (b * 256u + tid)], before = steps > 16u ? steps - 16u : 0u;
          const u32 act = (active >> b) & 1u;
          const u32 over = (u64)steps >= lim ? 1u : 0u;
          const u32 pow2 = (before == 0u || __builtin_clz(steps) < __builtin_clz(before)) ? 1u : 0u;
          lng |= (act & over) << b;
          save |= (act & (over ^ 1u) & pow2) << b;
        }}
      }} else {{
        for (u32 rem = active; rem; rem &= rem - 1u) {{
          const u32 b = __builtin_ctz(rem);
          const u32 start = (u32)meta[b * 256u + tid];
          const u32 steps = t - start, before = steps > 16u ? steps - 16u : 0u;
          if ((u64)steps >= lim) lng |= 1u << b;
          else if (before == 0u || __builtin_clz(steps) < __builtin_clz(before)) save |= 1u << b;
        }}
      }}
      if (save) {{
{each("        v@ = (v@ & ~save) | (s@ & save);" + chr(10))}      }}
      if (lng) {{
        atomicOr(long_flag, 1u);
        fin |= lng;
        active &= ~lng;
      }}
      u32 nfin;
      (void)ns_wave_excl(__popc(fin), &nfin);
      if (nfin < {WALK_BULK}u) {{
        for (u32 rem = fin; rem; rem &= rem - 1u) {{
          const u32 b = __builtin_ctz(rem), mb = 1u << b;
          u64 st = 0ull;
{each("          st |= (u64)((f@ >> b) & 1u) << @;" + chr(10))}{rec_one.replace(chr(10) + "        ", chr(10) + "          ").replace("        const u64 m", "          const u64 m", 1)}        }}
      }} else {{
        // many ended walks: their final states by bit transposes of f
        u32 Z[32];
{each("        Z[@] = f@;" + chr(10), 0, min(T, 32))}{each("        Z[@] = 0u;" + chr(10), T, 32) if T < 32 else ""}{t32("Z")}{("        u32 ZH[32];" + chr(10) + each("        ZH[@ - 32] = f@;" + chr(10), 32, T) + "".join(f"        ZH[{i}] = 0u;" + chr(10) for i in range(T - 32, 32)) + t32("ZH")) if wide else ""}#pragma unroll
        for (u32 b = 0; b < 32u; b++) {{
          const u32 mb = 1u << b;
          if (fin & mb) {{
            const u64 st = (u64)Z[b]{" | ((u64)ZH[b] << 32)" if wide else ""};
{rec_one.replace(chr(10) + "        ", chr(10) + "            ").replace("        const u64 m", "            const u64 m", 1)}          }}
        }}
      }}
      {{
        const u32 need = ~active;
{refill}      }}
      fin = 0u;
      fstop = 0u;
      fwp = 0u;
      floop = 0u;
      a0 = a1 = a2 = a3 = 0u;
    }}
  }}
  // evaluations and records of the wave, into one of 64 count slots 16 words apart
  for (int off = 32; off > 0; off >>= 1) {{
    ev += __shfl_xor(ev, off);
    nrec += __shfl_xor(nrec, off);
  }}
  if (lane == 0u && ev != 0ull) atomicAdd(evals + ((blockIdx.x & 63u) << 4), ev);
  if (lane == 0u && nrec != 0u) atomicAdd(evals + ((blockIdx.x & 63u) << 4) + 1, (u64)nrec);
}}
"""


def main():
    out = [HEAD, BATCH_HEAD]
    batch = []
    for ci, (cls, (budget, hi)) in enumerate(CLASSES.items()):
        kernels = [("sat", sat_kernel(cls, budget, hi)), ("qsata", qsat_a_kernel(cls, hi)),
                   ("qsatb", qsat_b_kernel(cls, hi - QSATB_POOL_CUT))]
        if budget >= 128:
            kernels.append(("table", table_kernel(cls, hi)))
        for kind, src in kernels:
            # one kernel per compilation: NS_T_<name> selects it
            out.append(f"#if defined(NS_T_ns_nat_{kind}_{cls})\n{src}#endif\n")
        batch.append(batch_kernels(cls, hi, ci))
    for T, classes in WALK_TIERS.items():
        for cls in classes:
            out.append(f"#if defined(NS_T_ns_nat_walk{T}_{cls})\n{walk_kernel(T, cls, CLASSES[cls][1])}#endif\n")
    # the batch kernels of every class in one compilation (NS_T_ns_nat_batch): a batch's code
    # object holds them all over one body region, so each class runs at its own occupancy
    out.append("#if defined(NS_T_ns_nat_batch)\n" + "".join(batch) + "#endif\n")
    open(sys.argv[1], "w").write("".join(out))


if __name__ == "__main__":
    main()
