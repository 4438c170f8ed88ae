/*
 * CPU oracle for the naivesat hot path — TEST INFRASTRUCTURE ONLY (the checker used by
 * tests/ and by the benchmark's CPU-baseline leg; the product never links it).
 *
 * How each function is pinned (beyond agreeing with the GPU engine):
 *   or_eval_u64        gate truth tables and negated outputs, checked against
 *                      closed-form functions (tests/test_oracle.py).
 *   or_sat_count*      brute force with or_eval_u64 on n <= 16, planted-key property
 *                      (the key always satisfies, so count >= 1 and lowest <= key), the
 *                      reference's lowest-first scalar scan (src/main.rs:362-371).
 *   or_qr_*            the 30 cases of the reference's test_quant_reducer
 *                      (src/qsat.rs:2035-2466), transcribed to tests/golden.
 *   or_qsat            QuantReducer over the full truth table on n <= 14 (the reference's
 *                      own small-n path, src/qsat.rs:1991-1998) and recursive evaluation of
 *                      the quantified formula on n <= 12.
 *   or_satostop_table  scalar walks (src/satostop.rs:1721-1782) on n <= 12 for every
 *                      unknown; start-state rule of test_mem_image_find_solution
 *                      (src/greedysatostop.rs:1203-1268, starts are u << (n-U)).
 *   or_table_u32       identity circuit layout of src/satostop.rs:1788-1915.
 *   or_hash64, or_hm_* the hash-map data sets of src/satostop.rs:1917-4144 (join_to_hashmap
 *                      x3, join_hashmap_itself x1, add_to_hashmap x2), transcribed to
 *                      tests/golden/hashmap_cases.json; entries are placed and looked up with
 *                      or_hash64, so the add_to_hashmap collisions pin the hash too.
 *   or_mem_*           the reference's test_mem_image (src/greedysatostop.rs:1145-1201: the
 *                      packed layout checked bit by bit against its closed form) and
 *                      test_mem_image_find_solution (1203-1268); or_mem_join_nexts by
 *                      tables whose joins have closed forms (tests/test_oracle_memimage.py).
 */
#ifndef NS_ORACLE_H
#define NS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t n, ngates, nout;
  const uint32_t* in0;
  const uint32_t* in1;
  const uint8_t* func; /* 0 And, 1 Nor, 2 Nimpl, 3 Xor */
  const uint32_t* out_wire;
  const uint8_t* out_neg;
} or_circuit;

/* input k = bit k of x; returns output i in bit i (first 64 outputs) */
uint64_t or_eval_u64(const or_circuit* c, uint64_t x);

/* satisfying assignments of output 0 over all 2^n x, and the lowest one (~0 if none) */
int or_sat_count(const or_circuit* c, int threads, uint64_t* count, uint64_t* lowest);
/* the vector ISA the bit-sliced SAT evaluation runs with on this CPU ("avx512f" or "avx2") */
const char* or_simd_isa(void);
/* same over 64-assignment words [w_lo, w_hi) (assignment x = 64*w + b) */
int or_sat_count_range(const or_circuit* c, uint64_t w_lo, uint64_t w_hi, int threads, uint64_t* count,
                       uint64_t* lowest);

/* streaming QuantReducer; quants_all[k] = 1 for 'all' on input k (input 0 outermost) */
typedef struct or_qr or_qr;
or_qr* or_qr_new(const uint8_t* quants_all, uint32_t len);
int or_qr_push(or_qr* q, uint64_t index, int item);
/* returns 1 when a final result is available */
int or_qr_final(const or_qr* q, int* reversed, uint32_t* solution_bits, int* has_solution, uint64_t* solution);
void or_qr_free(or_qr* q);

/* QSAT by level reduction; index v has input 0 as MSB. Outputs truth, number of witness
 * values of the outermost block, the lowest witness (~0 if none) and, when bitmap_bytes is
 * not NULL, the block's value bitmap (2^first_bits bits). */
int or_qsat(const or_circuit* c, const uint8_t* quants_all, int threads, uint32_t* truth, uint64_t* wcount,
            uint64_t* lowest, uint8_t* bitmap_bytes);

/* scalar SATOSTOP walk from `start`: evaluations until STOP (0 when not within max_steps) */
uint64_t or_satostop_walk(const or_circuit* c, uint64_t start, uint64_t max_steps, uint64_t* end);
/* full table + n doubling passes; per unknown u: steps (0 = never stops) and end state */
int or_satostop_table(const or_circuit* c, uint32_t unknowns, int threads, uint32_t* steps_out,
                      uint32_t* ends_out, uint64_t* n_stopped, uint64_t* lowest_u);
/* table[s] = outputs of state s packed (at most 32 outputs, n <= 26) */
int or_table_u32(const or_circuit* c, uint32_t* table);

/* naivesatostop hash map (src/satostop.rs:127-1150) */
typedef struct {
  uint64_t current, next, steps;
  uint32_t predecessors, state; /* 0 unused, 1 used, 2 stopped, 3 looped */
} or_hash_entry;
typedef struct {
  uint64_t start, end, steps;
  uint32_t defined;
} or_hm_solution;
uint64_t or_hash64(uint32_t bits, uint64_t value);

/* MemImage (src/greedysatostop.rs:339-458): entry i of `width` = state_len + 1 bits at bit
 * i*width of little-endian u64 words. */
uint64_t or_mem_get(const uint64_t* data, uint32_t width, uint64_t i);
void or_mem_set(uint64_t* data, uint32_t width, uint64_t i, uint64_t value);
/* every entry of values[0..count) through or_mem_set / or_mem_get */
void or_mem_pack(const uint64_t* values, uint64_t count, uint32_t width, uint64_t* words);
void or_mem_unpack(const uint64_t* words, uint64_t count, uint32_t width, uint64_t* values);
/* join_nexts (src/greedysatostop.rs:419-444): entries of dest (states dest_start..+len) that
 * have not stopped and whose next state lies in second (second_start..+second_len) take
 * second's entry for it; in place, in index order. */
void or_mem_join_nexts(uint64_t* dest, uint64_t dest_start, uint64_t len, const uint64_t* second,
                       uint64_t second_start, uint64_t second_len, uint32_t state_len);
/* find_solution (src/greedysatostop.rs:446-475): the lowest unknown start in the image's
 * range whose entry has stopped; returns 1 and (start, end) or 0. */
int or_mem_find_solution(const uint64_t* data, uint64_t start, uint64_t len, uint32_t state_len,
                         uint32_t unknowns, uint64_t* sol_start, uint64_t* sol_end);
void or_hm_join_to(uint32_t state_len, uint32_t arg_bit_place, uint64_t arg, const uint32_t* outputs,
                   or_hash_entry* hm, uint64_t len);
void or_hm_join_itself(uint32_t state_len, const or_hash_entry* in, or_hash_entry* out, uint64_t len,
                       uint32_t unknown_bits, uint32_t unknown_fill_bits, uint32_t* unknown_fills,
                       uint64_t* resolved, or_hm_solution* sol);
void or_hm_add(uint32_t state_len, uint32_t arg_bit_place, uint64_t arg, const uint32_t* outputs,
               or_hash_entry* hm, uint64_t len, uint32_t unknown_bits, uint32_t unknown_fill_bits,
               uint32_t* unknown_fills, uint64_t* resolved, or_hm_solution* sol, uint32_t max_predecessors,
               int test);

#ifdef __cplusplus
}
#endif
#endif
