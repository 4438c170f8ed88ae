// Launchers for the circuit-independent kernels (static_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace ns {

// The allocation-free SAT entry point's counters: a device array of this many u64 words per
// device (12 KiB), held for the life of the process.
constexpr unsigned long long kNoallocScratchWords = 512 + 64 * 16;
constexpr unsigned int kNoallocMaxDevices = 64;
// Resolves the array's address on `device`, which must be the current device (once per
// device, at context set-up); noalloc_scratch returns it afterwards (null before).
hipError_t noalloc_scratch_setup(int device);
unsigned long long* noalloc_scratch(int device);

hipError_t launch_qsat_reduce(hipStream_t s, const unsigned int* in, unsigned int m, unsigned int k,
                              unsigned int opmask, unsigned int* out, bool final_pass, bool q0all,
                              unsigned long long* stats);
hipError_t launch_jump_first(hipStream_t s, const unsigned int* t0, unsigned long long* dst,
                             unsigned long long count, unsigned int n);
hipError_t launch_jump_first_wide(hipStream_t s, const unsigned int* lo, const unsigned int* hi,
                                  unsigned long long* dst, unsigned long long count, unsigned int n,
                                  unsigned int split);
hipError_t launch_jump_pass(hipStream_t s, const unsigned long long* src, unsigned long long* dst,
                            unsigned long long count, unsigned int n, bool wide);
hipError_t launch_starts(hipStream_t s, const unsigned long long* tab, unsigned int n, unsigned int unknowns,
                         unsigned long long jumps, bool wide, unsigned long long* steps_out,
                         unsigned long long* ends_out, unsigned long long* stats);
// A next-state table as the walks read it: packed (n+1)-bit entries (packed non-null,
// width = n + 1) or lo[s] | hi[s] << split (hi null: lo alone).
struct TableView {
  const unsigned int* lo;
  const unsigned int* hi;
  unsigned int split;
  const unsigned long long* packed;
  unsigned int width;
};

// Packs lo[e] | hi[e] << split, e < entries (a multiple of 64), into width-bit fields.
hipError_t launch_pack(hipStream_t s, const unsigned int* lo, const unsigned int* hi, unsigned int split,
                       unsigned int width, unsigned long long entries, unsigned long long* out);

// Waypoint walks (records of 16 bytes per waypoint; see static_kernels.hip)
hipError_t launch_wp_walk(hipStream_t s, const TableView& tab, unsigned int n, unsigned int k,
                          unsigned long long count, unsigned long long limit, void* out, unsigned int* long_flag,
                          unsigned long long* evals);
hipError_t launch_wp_jump(hipStream_t s, const void* in, void* out, unsigned long long count, unsigned int* pending);
// coarse waypoint records (spacing k + d, count_out of them) from the spacing-k ones
hipError_t launch_wp_chase(hipStream_t s, const void* in, unsigned int d, void* out, unsigned long long count_out,
                           unsigned int max_hops, unsigned int* overflow);
hipError_t launch_wp_starts(hipStream_t s, const void* rec, unsigned int n, unsigned int unknowns, unsigned int k,
                            unsigned long long* steps_out, unsigned long long* ends_out, unsigned long long* stats);
// lanes must be a multiple of 256; out holds one word per lane
hipError_t launch_gather_probe(hipStream_t s, const unsigned int* t, unsigned int bits, unsigned int reads,
                               unsigned long long lanes, unsigned int* out, bool nontemporal);
// greedysatostop's partitioned search over packed (state_len+1)-bit images
hipError_t launch_mem_join(hipStream_t s, const unsigned long long* dest, unsigned long long len,
                           const unsigned long long* second, unsigned long long s_start, unsigned long long s_len,
                           unsigned int state_len, unsigned long long* out);
hipError_t launch_mem_find(hipStream_t s, const unsigned long long* words, unsigned long long start,
                           unsigned long long len, unsigned int state_len, unsigned int unknowns,
                           unsigned long long* best);
hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, unsigned long long v,
                           unsigned long long count);
hipError_t launch_fill_u32(hipStream_t s, unsigned int* p, unsigned int v, unsigned long long count);
// A SAT batch's work descriptor as the fold reads it (the batch kernels' NsBatchDesc:
// 64 bytes; out = the request's {count, lowest})
struct BatchFoldDesc {
  unsigned int wend, chunk0;
  unsigned long long* out;
  unsigned int entry, words, jb, opm, valid;
  unsigned int pad[7];
};
// per descriptor (one block each): its chunks' counts summed and lowest hits' minimum
hipError_t launch_batch_fold(hipStream_t s, const BatchFoldDesc* descs, unsigned int n,
                             const unsigned long long* chunk_out);
// p[i * stride] = v for i < count
hipError_t launch_fill_strided_u64(hipStream_t s, unsigned long long* p, unsigned long long v,
                                   unsigned long long count, unsigned long long stride);

}  // namespace ns
