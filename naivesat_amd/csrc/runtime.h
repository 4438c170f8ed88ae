// HIP runtime plumbing: error checks, run-time compilation of generated kernels
// (hiprtc, gfx950 code objects) and the per-context stream/buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "codegen.h"

namespace ns {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// status codes shared with the C ABI
enum : int {
  ST_OK = 0,
  ST_INVALID = 1,
  ST_HIP = 2,
  ST_COMPILE = 3,
  ST_UNSUPPORTED = 4,
  ST_NOMEM = 5,
  ST_INTERNAL = 6,
  ST_CANCELLED = 7,
};

void check_hip(hipError_t e, const char* what);
#define NS_HIP(x) ::ns::check_hip((x), #x)

struct JitModule {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  double compile_ms = 0.0;  // 0 when served from the cache
  uint64_t last_use = 0;    // cache clock of the last lookup
  ~JitModule();
};

// A kernel built for one circuit and its code-generation parameters.
struct BuiltKernel {
  std::shared_ptr<JitModule> mod;
  GenInfo info;
  bool native = false;      // register-allocated assembly in a precompiled template
  uint32_t vgpr_class = 0;  // template register class (native only)
  uint32_t tuned_wpl = 0;   // words per lane chosen by timing (0: not tuned yet)
  uint32_t tuned_block = 256;  // threads per block chosen with it
};

// Compiles generated sources once per context; keyed by a 128-bit digest of the source
// text. Holds at most kCapacity modules: the least recently used is dropped (a solve still
// running it keeps it alive through its reference). Module loads and unloads on this runtime
// set the sizes: a load costs ~0.2 ms plus ~0.55 us per module loaded in the process (4,000
// distinct SAT solves in a row: 2.84 s with 256 modules kept, 3.51 s with 512, 6.90 s with
// 4,096), and an unload while other threads load and launch slows them (8 serving lanes on
// 4,096 distinct requests: ~1,500 solves/s with unloads during the run, ~3,100 without).
// So the cache is small, and while a server is open (park()) the modules it drops are
// parked instead of unloaded: its lanes unload them when they find no request waiting
// (reap), and past kParkedMax parked modules a drop unloads at once.
// A module holds tens of KB of device memory.
class Jit {
 public:
  // HIP source compiled with hiprtc
  std::shared_ptr<JitModule> get(const std::string& src, const char* kernel, double* compile_ms);
  // gfx950 assembly module assembled and linked in-process (amd_comgr)
  std::shared_ptr<JitModule> get_asm(const std::string& asm_text, const std::string& kernel, double* compile_ms);
  // native kernel: the body's machine code copied into the template's cached code object;
  // bodies outside the encoded subset or too large for the slot are assembled (asm_text)
  std::shared_ptr<JitModule> get_native(const std::string& kernel, const std::string& body,
                                        const std::string& asm_text, double* compile_ms);
  // kernels by fingerprint (circuit content, kernel kind, parameters): a repeated solve of
  // the same circuit skips code generation as well as compilation
  bool find_built(const std::string& fp, BuiltKernel* out);
  void put_built(const std::string& fp, const BuiltKernel& k);
  // an open server parks dropped modules (calls nest: one per open server)
  void park(bool on);
  // unloads up to n parked modules (all when n is 0); returns how many it unloaded
  size_t reap(size_t n);
  ~Jit();

 private:
  std::shared_ptr<JitModule> lookup(const std::string& key, double* compile_ms);
  std::shared_ptr<JitModule> insert(const std::string& key, std::shared_ptr<JitModule> m);
  // under mu_: a dropped module reference, parked or handed back to be released unlocked
  void drop(std::shared_ptr<JitModule> m, std::vector<std::shared_ptr<JitModule>>& release);
  static constexpr size_t kCapacity = 256;
  static constexpr size_t kParkedMax = 4096;
  std::mutex mu_;
  uint64_t clock_ = 0;
  int parking_ = 0;
  std::unordered_map<std::string, std::shared_ptr<JitModule>> cache_;
  std::unordered_map<std::string, std::pair<BuiltKernel, uint64_t>> built_;  // value, last use
  std::vector<std::shared_ptr<JitModule>> parked_;
};

// Assembles and links a gfx950 assembly module in-process; returns the code object.
std::vector<char> assemble_native(const std::string& asm_text);

// Growable device allocation (contents are not preserved when it grows). A buffer can
// also be lent memory it does not own (a slice of the serving arena): it then never grows
// and release() only forgets it.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  bool owned = true;
  void ensure(size_t bytes);
  void release();
  void borrow(void* q, size_t bytes) {
    release();
    p = q;
    cap = bytes;
    owned = false;
  }
  ~DevBuf() { release(); }
  template <class T>
  T* as() { return static_cast<T*>(p); }
};

// One in-flight solve: a HIP stream with its own result and scratch buffers. Several lanes
// run concurrently on the device (the serving path); the plain entry points use the
// context's main lane.
struct Lane {
  int device = 0;
  Jit& jit;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf stats;             // small result block
  // pinned host block the allocation-free SAT path reads its counters back into
  static constexpr size_t kPinnedBytes = 16384;
  unsigned long long* pinned = nullptr;
  DevBuf buf_a, buf_b, buf_c;         // QSAT scratch (bit arrays)
  DevBuf tab_a, tab_b, tab_c, tab_d;  // SATOSTOP: next-state table, jump tables, starts
  DevBuf tab_e;                       // SATOSTOP over 31 state bits: outputs 32..n of the table
  DevBuf tab_img;                     // SATOSTOP: packed (n+1)-bit next-state image
  DevBuf bstats, bdesc, bchunk;       // batches: per-request results, work descriptors, chunk results
  std::vector<double> pass_ms;  // jump-pass kernel times of the last SATOSTOP solve
  // launch sequences of repeated searches, captured once and replayed as HIP graphs; each
  // holds its kernel's module so the module outlives the graph
  struct CapturedGraph {
    hipGraphExec_t exec = nullptr;
    std::shared_ptr<JitModule> mod;
  };
  std::unordered_map<std::string, CapturedGraph> graphs;
  hipGraphExec_t find_graph(const std::string& key) const {
    auto it = graphs.find(key);
    return it == graphs.end() ? nullptr : it->second.exec;
  }
  void put_graph(const std::string& key, hipGraphExec_t exec, std::shared_ptr<JitModule> mod);
  void drop_graphs();
  Lane(int dev, Jit& j);
  ~Lane();
  Lane(const Lane&) = delete;
  Lane& operator=(const Lane&) = delete;
  // elapsed ms between ev0 and ev1 (both recorded on `stream`)
  float elapsed_ms();
  // device bytes held by the scratch buffers
  size_t held_bytes() const {
    return buf_a.cap + buf_b.cap + buf_c.cap + tab_a.cap + tab_b.cap + tab_c.cap + tab_d.cap + tab_e.cap +
           tab_img.cap + bstats.cap + bdesc.cap + bchunk.cap;
  }
  void trim_tables() {
    tab_a.release();
    tab_b.release();
    tab_c.release();
    tab_d.release();
    tab_e.release();
    tab_img.release();
  }
};

class Batcher;

struct Context {
  int device = 0;
  Jit jit;
  std::unique_ptr<Lane> main;
  explicit Context(int dev);
  ~Context();
  // the context's own batcher (ns_batch_solve): keeps recent batches, made on first use
  Batcher& batcher();

 private:
  std::mutex batcher_mu_;
  std::unique_ptr<Batcher> batcher_;
};

}  // namespace ns
