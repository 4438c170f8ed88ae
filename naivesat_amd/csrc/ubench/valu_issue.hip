// VALU issue micro-benchmark for gfx950: time per instruction of the patterns the native
// circuit bodies use (v_bitop3_b32 chains, SGPR operands, AGPR copies, v_readlane /
// v_writelane, scalar ops, LDS spills, operand banks), at 1, 2 and 4 waves per SIMD (256,
// 512, 1024 blocks of 4 waves on 256 CUs).
// Prints per-wave clock64 cycles per iteration and the wall-clock lane-op rate.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
typedef unsigned int u32;
typedef unsigned long long u64;

#define R8(x) x x x x x x x x
#define R64(x) R8(R8(x))

__global__ void __launch_bounds__(256) k_dep(u32* out, int iters, u64* cyc) {
  u32 a = threadIdx.x, b = a * 3, c = a * 7;
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R64("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96\n") : "+v"(a) : "v"(b), "v"(c));
  }
  u64 t1 = clock64();
  out[blockIdx.x * 256 + threadIdx.x] = a;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// 64 rotating-independent VALU ops per iteration (baseline for the next kernels)
__global__ void __launch_bounds__(256) k_rot(u32* out, int iters, u64* cyc) {
  u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, b = a0 * 3, c = a0 * 7;
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R8(R8("v_bitop3_b32 %0, %0, %4, %5 bitop3:0x96\nv_bitop3_b32 %1, %1, %4, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %2, %2, %4, %5 bitop3:0x96\nv_bitop3_b32 %3, %3, %4, %5 bitop3:0x96\n"))
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3) : "v"(b), "v"(c));
  }
  u64 t1 = clock64();
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// 256 rotating VALU ops + 64 SALU ops per iteration (one SALU after every 4 VALU)
__global__ void __launch_bounds__(256) k_salu(u32* out, int iters, u64* cyc) {
  u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, b = a0 * 3, c = a0 * 7;
  u32 s0 = __builtin_amdgcn_readfirstlane(blockIdx.x), s1 = s0 * 5;
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R8(R8("v_bitop3_b32 %0, %0, %4, %5 bitop3:0x96\nv_bitop3_b32 %1, %1, %4, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %2, %2, %4, %5 bitop3:0x96\nv_bitop3_b32 %3, %3, %4, %5 bitop3:0x96\n"
                       "s_xor_b32 %6, %6, %7\n"))
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(b), "+v"(c), "+s"(s0) : "s"(s1) : "scc");
  }
  u64 t1 = clock64();
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ s0;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// 256 rotating VALU ops + 64 AGPR write/read pairs per iteration
__global__ void __launch_bounds__(256) k_acc(u32* out, int iters, u64* cyc) {
  u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, b = a0 * 3, c = a0 * 7, t = 0;
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R8(R8("v_bitop3_b32 %0, %0, %4, %5 bitop3:0x96\nv_accvgpr_write_b32 a7, %0\n"
                       "v_bitop3_b32 %1, %1, %4, %5 bitop3:0x96\nv_bitop3_b32 %2, %2, %4, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %3, %3, %4, %5 bitop3:0x96\nv_accvgpr_read_b32 %6, a7\n"))
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(b), "+v"(c), "=v"(t) : : "a7");
    a0 ^= t;
  }
  u64 t1 = clock64();
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// 256 rotating VALU ops + 64 (readlane, s_nop 1, VALU use of the SGPR) per iteration
__global__ void __launch_bounds__(256) k_rl(u32* out, int iters, u64* cyc) {
  u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, b = a0 * 3, c = a0 * 7;
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R8(R8("v_readlane_b32 s60, %4, 5\ns_nop 1\nv_bitop3_b32 %0, %0, s60, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %1, %1, %4, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %2, %2, %4, %5 bitop3:0x96\nv_bitop3_b32 %3, %3, %4, %5 bitop3:0x96\n"))
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(b), "+v"(c) : : "s60");
  }
  u64 t1 = clock64();
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// like k_rl but the readlane is issued 3 VALU ops before its use and no s_nop
__global__ void __launch_bounds__(256) k_rl_early(u32* out, int iters, u64* cyc) {
  u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, b = a0 * 3, c = a0 * 7;
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R8(R8("v_readlane_b32 s60, %4, 5\n"
                       "v_bitop3_b32 %1, %1, %4, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %2, %2, %4, %5 bitop3:0x96\nv_bitop3_b32 %3, %3, %4, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %0, %0, s60, %5 bitop3:0x96\n"))
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(b), "+v"(c) : : "s60");
  }
  u64 t1 = clock64();
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// 256 VALU + 64 writelane per iteration
__global__ void __launch_bounds__(256) k_wl(u32* out, int iters, u64* cyc) {
  u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, b = a0 * 3, c = a0 * 7, t = 0;
  u32 s0 = __builtin_amdgcn_readfirstlane(blockIdx.x);
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R8(R8("v_writelane_b32 %6, %7, 3\nv_bitop3_b32 %0, %0, %4, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %1, %1, %4, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %2, %2, %4, %5 bitop3:0x96\nv_bitop3_b32 %3, %3, %4, %5 bitop3:0x96\n"))
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(b), "+v"(c), "+v"(t) : "s"(s0));
  }
  u64 t1 = clock64();
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ t;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// dependent chain through an SGPR operand: SALU result feeds the next VALU at once
__global__ void __launch_bounds__(256) k_salu_dep(u32* out, int iters, u64* cyc) {
  u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, b = a0 * 3, c = a0 * 7;
  u32 s0 = __builtin_amdgcn_readfirstlane(blockIdx.x), s1 = s0 * 5;
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R8(R8("s_xor_b32 %6, %6, %7\nv_bitop3_b32 %0, %0, %6, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %1, %1, %4, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %2, %2, %4, %5 bitop3:0x96\nv_bitop3_b32 %3, %3, %4, %5 bitop3:0x96\n"))
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(b), "+v"(c), "+s"(s0) : "s"(s1) : "scc");
  }
  u64 t1 = clock64();
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ s0;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// 256 VALU + 64 LDS spill/reload pairs per iteration: per group of 4 VALU ops a value is
// written to LDS and read back; the read is consumed after 3 VALU ops (short) or the
// reads of 4 groups are issued 16 VALU ops ahead of their uses (long)
__global__ void __launch_bounds__(256) k_lds_short(u32* out, int iters, u64* cyc) {
  __shared__ u32 slots[4 * 256];
  u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, b = a0 * 3, c = a0 * 7, t = 0;
  const u32 addr = (u32)(size_t)&slots[threadIdx.x];
  for (int k = 0; k < 4; k++) slots[threadIdx.x + 256 * k] = 0u;  // keeps the allocation
  __syncthreads();
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R8(R8("ds_write_b32 %7, %0\nds_read_b32 %6, %7 offset:1024\n"
                       "v_bitop3_b32 %1, %1, %4, %5 bitop3:0x96\nv_bitop3_b32 %2, %2, %4, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %3, %3, %4, %5 bitop3:0x96\ns_waitcnt lgkmcnt(0)\n"
                       "v_bitop3_b32 %0, %0, %6, %5 bitop3:0x96\n"))
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(b), "+v"(c), "=&v"(t) : "v"(addr) : "memory");
  }
  u64 t1 = clock64();
  __syncthreads();
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ t ^ slots[(threadIdx.x + 1) & 255];
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

__global__ void __launch_bounds__(256) k_lds_long(u32* out, int iters, u64* cyc) {
  __shared__ u32 slots[4 * 256];
  u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, b = a0 * 3, c = a0 * 7;
  u32 t0r = 0, t1r = 0, t2r = 0, t3r = 0;
  const u32 addr = (u32)(size_t)&slots[threadIdx.x];
  for (int k = 0; k < 4; k++) slots[threadIdx.x + 256 * k] = 0u;  // keeps the allocation
  __syncthreads();
#define ROT4 "v_bitop3_b32 %1, %1, %4, %5 bitop3:0x96\nv_bitop3_b32 %2, %2, %4, %5 bitop3:0x96\n" \
             "v_bitop3_b32 %3, %3, %4, %5 bitop3:0x96\nv_bitop3_b32 %0, %0, %4, %5 bitop3:0x96\n"
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R8(R8("ds_write_b32 %10, %0\nds_write_b32 %10, %1 offset:1024\n"
                       "ds_write_b32 %10, %2 offset:2048\nds_write_b32 %10, %3 offset:3072\n"
                       "ds_read_b32 %6, %10\nds_read_b32 %7, %10 offset:1024\n"
                       "ds_read_b32 %8, %10 offset:2048\nds_read_b32 %9, %10 offset:3072\n"
                       ROT4 ROT4 ROT4 "s_waitcnt lgkmcnt(0)\n"
                       "v_bitop3_b32 %0, %0, %6, %5 bitop3:0x96\nv_bitop3_b32 %1, %1, %7, %5 bitop3:0x96\n"
                       "v_bitop3_b32 %2, %2, %8, %5 bitop3:0x96\nv_bitop3_b32 %3, %3, %9, %5 bitop3:0x96\n"))
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(b), "+v"(c), "=&v"(t0r), "=&v"(t1r), "=&v"(t2r),
                   "=&v"(t3r)
                 : "v"(addr) : "memory");
  }
#undef ROT4
  u64 t1 = clock64();
  __syncthreads();
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ t0r ^ t1r ^ t2r ^ t3r ^ slots[(threadIdx.x + 1) & 255];
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// 256 VALU-heavy blocks with one LDS spill and one reload per 9 VALU ops, each reload read
// 9 ops after it was issued (s_waitcnt lgkmcnt(2): the block's own two LDS ops may still be
// in flight), the pattern of a body that keeps spilled values in LDS
__global__ void __launch_bounds__(256) k_lds_pipe(u32* out, int iters, u64* cyc) {
  __shared__ u32 slots[4 * 256];
  u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, b = a0 * 3, c = a0 * 7;
  u32 t0r = 0, t1r = 0;
  const u32 addr = (u32)(size_t)&slots[threadIdx.x];
  for (int k = 0; k < 4; k++) slots[threadIdx.x + 256 * k] = 0u;  // keeps the allocation
  __syncthreads();
#define ROT4 "v_bitop3_b32 %1, %1, %4, %5 bitop3:0x96\nv_bitop3_b32 %2, %2, %4, %5 bitop3:0x96\n" \
             "v_bitop3_b32 %3, %3, %4, %5 bitop3:0x96\nv_bitop3_b32 %0, %0, %4, %5 bitop3:0x96\n"
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R8(R8("ds_write_b32 %8, %0\nds_read_b32 %6, %8 offset:1024\n" ROT4 ROT4
                       "s_waitcnt lgkmcnt(2)\nv_bitop3_b32 %1, %1, %7, %5 bitop3:0x96\n"
                       "ds_write_b32 %8, %1 offset:2048\nds_read_b32 %7, %8 offset:3072\n" ROT4 ROT4
                       "s_waitcnt lgkmcnt(2)\nv_bitop3_b32 %0, %0, %6, %5 bitop3:0x96\n"))
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(b), "+v"(c), "+v"(t0r), "+v"(t1r)
                 : "v"(addr) : "memory");
  }
#undef ROT4
  u64 t1 = clock64();
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __syncthreads();
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ t0r ^ t1r ^ slots[(threadIdx.x + 1) & 255];
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// Instruction-stream size: 8 bank-disjoint chains of v_bitop3_b32 (v40, v43, v44, v47, v48,
// v51, v52, v55 reading v41 and v42) unrolled to 2,048 ops (16 KiB of code per iteration)
// or 256 ops (2 KiB), alone or with one scalar op after every 4 (s_xor_b32: 4 bytes;
// s_bfe_i32 with a 32-bit literal: 8 bytes)
#define B8 "v_bitop3_b32 v40, v40, v41, v42 bitop3:0x96\nv_bitop3_b32 v43, v43, v41, v42 bitop3:0x96\n" \
           "v_bitop3_b32 v44, v44, v41, v42 bitop3:0x96\nv_bitop3_b32 v47, v47, v41, v42 bitop3:0x96\n" \
           "v_bitop3_b32 v48, v48, v41, v42 bitop3:0x96\nv_bitop3_b32 v51, v51, v41, v42 bitop3:0x96\n" \
           "v_bitop3_b32 v52, v52, v41, v42 bitop3:0x96\nv_bitop3_b32 v55, v55, v41, v42 bitop3:0x96\n"
#define SX "s_xor_b32 s60, s60, s61\n"
#define SB "s_bfe_i32 s60, s61, 0x10016\n"
#define BIGK(name, unit, reps)                                                                          \
  __global__ void __launch_bounds__(256) name(u32* out, int iters, u64* cyc) {                          \
    u32 r;                                                                                              \
    asm volatile("v_mov_b32 v40, %1\nv_mov_b32 v41, %1\nv_mov_b32 v42, %1\nv_mov_b32 v43, %1\n"        \
                 "v_mov_b32 v44, %1\nv_mov_b32 v47, %1\nv_mov_b32 v48, %1\nv_mov_b32 v51, %1\n"         \
                 "v_mov_b32 v52, %1\nv_mov_b32 v55, %1\ns_mov_b32 s60, 0\ns_mov_b32 s61, 0\n"           \
                 "v_mov_b32 %0, v40" : "=v"(r) : "v"(threadIdx.x)                                      \
                 : "v40", "v41", "v42", "v43", "v44", "v47", "v48", "v51", "v52", "v55", "s60", "s61");    \
    u64 t0 = clock64();                                                                                 \
    for (int i = 0; i < iters; i++) {                                                                   \
      asm volatile(reps(unit) : : : "v40", "v41", "v42", "v43", "v44", "v47", "v48", "v51", "v52", "v55", \
                   "s60", "s61", "scc");                                                                \
    }                                                                                                   \
    u64 t1 = clock64();                                                                                 \
    asm volatile("v_xor_b32 %0, %0, v40\nv_xor_b32 %0, %0, v55" : "+v"(r) : : "v40", "v55");             \
    out[blockIdx.x * 256 + threadIdx.x] = r;                                                            \
    if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;                                                   \
  }
#define R4(x) x x x x
#define R256(x) R4(R8(R8(x)))
#define R32(x) R4(R8(x))
BIGK(k_big_v, B8, R256)
BIGK(k_big_xor, B8 SX SX, R256)
BIGK(k_big_bfe, B8 SB SB, R256)
BIGK(k_small_v, B8, R32)
BIGK(k_small_bfe, B8 SB SB, R32)

// Operand patterns of real bodies, 2,048 ops per iteration with no dependences between them:
// destinations rotate over v40..v47, sources are read-only registers v48..v63 (banks 0, 1, 2
// for the three operands, register index mod 4 being the bank); variants: every op with an
// SGPR operand, one op in four with one, one op in four a v_bfe_i32 (inline-constant field),
// every op with two sources of one bank
#define P3(d, a, b, c) "v_bitop3_b32 v" #d ", v" #a ", v" #b ", v" #c " bitop3:0x96\n"
#define PS(d, a, c) "v_bitop3_b32 v" #d ", v" #a ", s60, v" #c " bitop3:0x96\n"
#define PB(d, a) "v_bfe_i32 v" #d ", v" #a ", 5, 1\n"
#define D8 P3(40, 48, 49, 50) P3(41, 52, 53, 54) P3(42, 56, 57, 58) P3(43, 60, 61, 62) \
           P3(44, 48, 53, 58) P3(45, 52, 57, 62) P3(46, 56, 61, 50) P3(47, 60, 49, 54)
#define S8 PS(40, 48, 50) PS(41, 52, 54) PS(42, 56, 58) PS(43, 60, 62) \
           PS(44, 48, 58) PS(45, 52, 62) PS(46, 56, 50) PS(47, 60, 54)
#define S48 P3(40, 48, 49, 50) P3(41, 52, 53, 54) P3(42, 56, 57, 58) PS(43, 60, 62) \
            P3(44, 48, 53, 58) P3(45, 52, 57, 62) P3(46, 56, 61, 50) PS(47, 60, 54)
#define B48 P3(40, 48, 49, 50) P3(41, 52, 53, 54) P3(42, 56, 57, 58) PB(43, 60) \
            P3(44, 48, 53, 58) P3(45, 52, 57, 62) P3(46, 56, 61, 50) PB(47, 60)
#define C8 P3(40, 48, 52, 50) P3(41, 52, 56, 54) P3(42, 56, 60, 58) P3(43, 60, 48, 62) \
           P3(44, 48, 56, 58) P3(45, 52, 60, 62) P3(46, 56, 48, 50) P3(47, 60, 52, 54)
#define OPK(name, unit)                                                                              \
  __global__ void __launch_bounds__(256) name(u32* out, int iters, u64* cyc) {                       \
    u32 r;                                                                                           \
    asm volatile(R8(R8("v_mov_b32 v40, %1\n")) "s_mov_b32 s60, 0x5a5a5a5a\nv_mov_b32 %0, v40"       \
                 : "=v"(r) : "v"(threadIdx.x) : "v40", "s60");                                       \
    asm volatile("v_mov_b32 v48, %0\nv_mov_b32 v49, %0\nv_mov_b32 v50, %0\nv_mov_b32 v52, %0\n"      \
                 "v_mov_b32 v53, %0\nv_mov_b32 v54, %0\nv_mov_b32 v56, %0\nv_mov_b32 v57, %0\n"       \
                 "v_mov_b32 v58, %0\nv_mov_b32 v60, %0\nv_mov_b32 v61, %0\nv_mov_b32 v62, %0\n"       \
                 : : "v"(r) : "v48", "v49", "v50", "v52", "v53", "v54", "v56", "v57", "v58", "v60",    \
                   "v61", "v62");                                                                    \
    u64 t0 = clock64();                                                                              \
    for (int i = 0; i < iters; i++) {                                                                \
      asm volatile(R256(unit) : : : "v40", "v41", "v42", "v43", "v44", "v45", "v46", "v47", "v48",    \
                   "v49", "v50", "v52", "v53", "v54", "v56", "v57", "v58", "v60", "v61", "v62", "s60"); \
    }                                                                                                \
    u64 t1 = clock64();                                                                              \
    asm volatile("v_xor_b32 %0, %0, v47" : "+v"(r) : : "v47");                                        \
    out[blockIdx.x * 256 + threadIdx.x] = r;                                                         \
    if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;                                                \
  }
OPK(k_op_vgpr, D8)
OPK(k_op_sgpr, S8)
OPK(k_op_sgpr4, S48)
OPK(k_op_bfe4, B48)
OPK(k_op_bank, C8)

// 64 dependent 4-byte VOP2 ops (v_xor_b32) per iteration
__global__ void __launch_bounds__(256) k_vop2(u32* out, int iters, u64* cyc) {
  u32 a = threadIdx.x, b = a * 3;
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R64("v_xor_b32 %0, %1, %0\n") : "+v"(a) : "v"(b));
  }
  u64 t1 = clock64();
  out[blockIdx.x * 256 + threadIdx.x] = a;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// 64 independent 4-byte VOP2 ops over 4 chains
__global__ void __launch_bounds__(256) k_vop2_rot(u32* out, int iters, u64* cyc) {
  u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, b = a0 * 3;
  u64 t0 = clock64();
  for (int i = 0; i < iters; i++) {
    asm volatile(R8(R8("v_xor_b32 %0, %4, %0\nv_xor_b32 %1, %4, %1\nv_xor_b32 %2, %4, %2\nv_xor_b32 %3, %4, %3\n"))
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3) : "v"(b));
  }
  u64 t1 = clock64();
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

typedef void (*kfn)(u32*, int, u64*);
int main() {
  struct K { const char* name; kfn f; int valu; int other; } ks[] = {
      {"dep64", k_dep, 64, 0},         {"rot256", k_rot, 256, 0},       {"salu256+64", k_salu, 256, 64},
      {"acc256+128", k_acc, 256, 128}, {"rl256+64+nop", k_rl, 256, 64}, {"rlearly256+64", k_rl_early, 256, 64},
      {"wl256+64", k_wl, 256, 64},     {"saludep256+64", k_salu_dep, 256, 64},
      {"ldsshort256+128", k_lds_short, 256, 128}, {"ldslong256+512", k_lds_long, 256 + 256, 512},
      {"ldspipe1152+256", k_lds_pipe, 64 * 18, 256},
      {"big2048", k_big_v, 2048, 0}, {"big2048+xor512", k_big_xor, 2048, 512},
      {"big2048+bfe512", k_big_bfe, 2048, 512}, {"small256", k_small_v, 256, 0},
      {"small256+bfe64", k_small_bfe, 256, 64},
      {"ops-vgpr2048", k_op_vgpr, 2048, 0}, {"ops-sgpr2048", k_op_sgpr, 2048, 0},
      {"ops-sgpr1in4", k_op_sgpr4, 2048, 0}, {"ops-bfe1in4", k_op_bfe4, 2048, 0},
      {"ops-bank2048", k_op_bank, 2048, 0},
      {"vop2dep64", k_vop2, 64, 0},    {"vop2rot256", k_vop2_rot, 256, 0}};
  const int maxb = 256 * 8;
  u32* out;
  u64* cyc;
  if (hipMalloc(&out, sizeof(u32) * 256 * maxb) != hipSuccess) return 1;
  if (hipMalloc(&cyc, sizeof(u64) * maxb) != hipSuccess) return 1;
  u64* hc = (u64*)malloc(sizeof(u64) * maxb);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int iters = 2000;
  for (auto& k : ks) {
    for (int wps = 1; wps <= 4; wps *= 2) {
      int blocks = 256 * wps;  // 4 waves per block: wps waves per SIMD if spread evenly
      hipLaunchKernelGGL(k.f, dim3(blocks), dim3(256), 0, 0, out, 10, cyc);
      (void)hipEventRecord(e0, 0);
      hipLaunchKernelGGL(k.f, dim3(blocks), dim3(256), 0, 0, out, iters, cyc);
      (void)hipEventRecord(e1, 0);
      if (hipEventSynchronize(e1) != hipSuccess) return 2;
      float ms;
      (void)hipEventElapsedTime(&ms, e0, e1);
      (void)hipMemcpy(hc, cyc, sizeof(u64) * blocks, hipMemcpyDeviceToHost);
      double avg = 0;
      for (int i = 0; i < blocks; i++) avg += hc[i];
      avg /= blocks;
      double per_iter = avg / iters;
      const double lane_ops = (double)blocks * 256.0 * iters * (k.valu + k.other);
      printf("%-16s waves/SIMD=%d  ms=%.3f  clk/iter=%.1f  clk per VALU=%.2f  lane-ops/s=%.3e\n", k.name, wps, ms,
             per_iter, per_iter / k.valu, lane_ops / (ms * 1e-3));
    }
  }
  return 0;
}
