// Solver drivers. The search space is covered by launches over contiguous ranges of
// 32-assignment words (the reference's per-`arg` launches, src/main.rs:178-196, but
// with results accumulated on the device and read back once).
#include "solvers.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <functional>
#include <cstdlib>
#include <cstring>

#include "asmgen.h"
#include "codegen.h"
#include "file_image.h"
#include "hashmap_kernels.h"
#include "static_kernels.h"

namespace ns {

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t0) {
  return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

uint32_t ilog2(uint64_t x) {
  uint32_t r = 0;
  while ((1ull << r) < x) r++;
  return r;
}

bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

// Geometry for a search over `total_words` words.
Geometry pick_geometry(uint64_t total_words, uint32_t group_len, uint32_t wpl, bool j_low) {
  Geometry g;
  g.block = group_len ? group_len : 256;
  g.wpl = wpl ? wpl : 4;
  if (g.block % 64 || g.block > 1024) throw Error(ST_INVALID, "group_len must be a multiple of 64, at most 1024");
  if (!is_pow2(g.wpl) || g.wpl > 32) throw Error(ST_INVALID, "words_per_lane must be a power of two <= 32");
  // small spaces: one wave, one word per lane
  while ((uint64_t)g.block * g.wpl > total_words && (g.block > 64 || g.wpl > 1)) {
    if (g.wpl > 1) g.wpl >>= 1;
    else g.block >>= 1;
    if (g.block % 64) g.block = 64;
  }
  uint32_t jb = ilog2(g.wpl);
  if (j_low) {
    g.lane_shift = jb;
    g.j_shift = 0;
  } else {
    g.lane_shift = 0;
    g.j_shift = 6;
  }
  return g;
}

// SAT launch shape: words per lane, then threads per block, shrink until one block fits in a
// launch range of 2^(elem_bits-5) words, so a requested range size is kept exactly (down to
// 64 words = 2^11 assignments per range).
Geometry sat_geometry(uint64_t total_words, const ns_sat_opts& o) {
  uint64_t cap = total_words;
  if (o.elem_bits >= 5 && o.elem_bits <= 40 && (1ull << (o.elem_bits - 5)) < cap) cap = 1ull << (o.elem_bits - 5);
  return pick_geometry(cap, o.group_len, o.words_per_lane, false);
}

// Progress reports of a search made of launch ranges (ns_sat_solve_progress): one call before
// the first launch, then one per range in order once it has completed. With a callback, an
// event follows each launch and the host waits for range k only after range k+1 is queued,
// so the device stays busy while the host reports. A nonzero return from the callback
// cancels: the caller launches no further range.
class RangeProgress {
 public:
  RangeProgress(Lane& lane, ns_sat_progress_fn fn, void* user, uint64_t steps, uint32_t bits)
      : lane_(lane), fn_(fn), user_(user), steps_(steps), bits_(bits) {}
  ~RangeProgress() {
    for (uint32_t k = 0; k < npend_; k++) (void)hipEventDestroy(pend_[k]);
  }
  RangeProgress(const RangeProgress&) = delete;
  RangeProgress& operator=(const RangeProgress&) = delete;
  bool active() const { return fn_ != nullptr; }
  bool cancelled() const { return cancel_; }
  void start() { call(NS_SAT_STEP_START); }
  // range `step` is complete (the stream was synchronized after it)
  void completed(uint64_t step) { call(step); }
  // range `step` was just launched on the lane's stream
  void launched(uint64_t step) {
    if (!fn_) return;
    hipEvent_t e = nullptr;
    NS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    pend_[npend_] = e;
    step_[npend_] = step;
    npend_++;
    NS_HIP(hipEventRecord(e, lane_.stream));
    if (npend_ == 2) wait_oldest();
  }
  void drain() {
    while (npend_) wait_oldest();
  }

 private:
  void call(uint64_t step) {
    if (fn_ && !cancel_ && fn_(user_, step, steps_, bits_) != 0) cancel_ = true;
  }
  void wait_oldest() {
    hipEvent_t e = pend_[0];
    uint64_t st = step_[0];
    pend_[0] = pend_[1];
    step_[0] = step_[1];
    npend_--;
    hipError_t err = hipEventSynchronize(e);
    (void)hipEventDestroy(e);
    NS_HIP(err);
    call(st);
  }
  Lane& lane_;
  ns_sat_progress_fn fn_;
  void* user_;
  uint64_t steps_;
  uint32_t bits_;
  bool cancel_ = false;
  hipEvent_t pend_[2] = {nullptr, nullptr};
  uint64_t step_[2] = {0, 0};
  uint32_t npend_ = 0;
};

// Words per launch: 2^(elem_bits-5) rounded up to whole blocks.
uint64_t chunk_words(uint32_t elem_bits, const Geometry& g, uint64_t total_words) {
  uint32_t eb = elem_bits ? elem_bits : 32;
  if (eb < 5) eb = 5;
  if (eb > 40) eb = 40;
  uint64_t cw = 1ull << (eb - 5);
  uint64_t wpb = g.words_per_block();
  cw = (cw + wpb - 1) / wpb * wpb;
  uint64_t tot = (total_words + wpb - 1) / wpb * wpb;
  return cw < tot ? cw : tot;
}

void launch_range(Lane& ctx, hipFunction_t fn, const Geometry& g, uint64_t w0, uint64_t words,
                  void** extra_args, int n_extra, uint64_t wend) {
  uint64_t blocks = (words + g.words_per_block() - 1) / g.words_per_block();
  if (blocks > 0x7fffffffull) throw Error(ST_UNSUPPORTED, "launch range too large");
  if (n_extra > 10) throw Error(ST_INTERNAL, "too many kernel arguments");
  void* args[12];
  args[0] = &w0;
  args[1] = &wend;
  for (int i = 0; i < n_extra; i++) args[2 + i] = extra_args[i];
  NS_HIP(hipModuleLaunchKernel(fn, (unsigned)blocks, 1, 1, g.block, 1, 1, 0, ctx.stream, args, nullptr));
}

// The native path (assembly bodies in precompiled templates) is used unless the
// environment selects the compiler path (NS_BACKEND=hiprtc) or the launch needs more than
// the templates' 256 threads per block.
bool native_wanted(const Geometry& g) {
  const char* b = std::getenv("NS_BACKEND");
  if (b && std::strcmp(b, "hiprtc") == 0) return false;
  return g.block <= 256;
}

uint32_t word_valid_mask(uint32_t n) { return n >= 5 ? 0xffffffffu : (uint32_t)((1ull << (1u << n)) - 1); }

using Chosen = BuiltKernel;

constexpr uint32_t kTuneCandidates = 10;  // launch geometries the SAT autotuner times

// Search counters in ctx.stats: [0] count, [1] lowest hit, and from u64 index 512 on 64
// count slots 16 words apart that the search kernels' waves add to (by block), so that
// many hits do not queue on one word. The count is stats[0] plus the slots.
constexpr uint64_t kSlot0 = 512, kSlots = 64, kSlotStride = 16;

void reset_counts(Lane& ctx) {
  unsigned long long* st = ctx.stats.as<unsigned long long>();
  NS_HIP(launch_fill_u64(ctx.stream, st, 0ull, 1));
  NS_HIP(launch_fill_u64(ctx.stream, st + 1, ~0ull, 1));
  NS_HIP(launch_fill_u64(ctx.stream, st + kSlot0, 0ull, kSlots * kSlotStride));
}

// enqueues the copy; host[0..1] hold (count, lowest) after read_counts_done
struct CountsCopy {
  unsigned long long head[2];
  unsigned long long slots[kSlots * kSlotStride];
};

void read_counts_async(Lane& ctx, CountsCopy& cc, const unsigned long long* counters = nullptr) {
  const unsigned long long* st = counters ? counters : ctx.stats.as<unsigned long long>();
  NS_HIP(hipMemcpyAsync(cc.head, st, sizeof(cc.head), hipMemcpyDeviceToHost, ctx.stream));
  NS_HIP(hipMemcpyAsync(cc.slots, st + kSlot0, sizeof(cc.slots), hipMemcpyDeviceToHost, ctx.stream));
}

void read_counts_done(const CountsCopy& cc, unsigned long long host[2]) {
  unsigned long long total = cc.head[0];
  for (uint64_t s = 0; s < kSlots; s++) total += cc.slots[s * kSlotStride];
  host[0] = total;
  host[1] = cc.head[1];
}


// Kernel for a circuit: register-allocated assembly in a precompiled template when the body
// fits one, else generated HIP compiled by hiprtc. Cached per context under the circuit's
// fingerprint and every parameter the code depends on; *compile_ms counts code generation
// and compilation (0 on a cache hit).
Chosen choose_kernel(Lane& ctx, const Circuit& c, NativeKind kind, const std::vector<uint32_t>& vb,
                     uint32_t lane_shift, const std::vector<uint32_t>& outs, const Geometry& g,
                     const std::string& extra_key, const std::function<std::string(GenInfo*)>& hiprtc_src,
                     const char* hiprtc_kernel, double* compile_ms, std::string* fp_out = nullptr,
                     bool lookup_only = false, bool group_exits = false) {
  const bool native_ok = native_wanted(g);
  const char* og = std::getenv("NS_OUTPUT_GROUPS");  // changes the native body
  std::string fp = circuit_fingerprint(c) + "/" + (og ? og : "") + "/" + std::to_string((int)kind) + "/" +
                   std::to_string(lane_shift) +
                   "/" + std::to_string(g.block) + "/" + std::to_string(g.wpl) + "/" + std::to_string(g.j_shift) +
                   (native_ok ? "/n/" : "/r/") + extra_key + (group_exits ? "/x" : "");
  for (uint32_t k : outs) fp += "," + std::to_string(k);
  if (fp_out) *fp_out = fp;
  Chosen ch;
  if (ctx.jit.find_built(fp, &ch)) {
    if (compile_ms) *compile_ms = 0.0;
    return ch;
  }
  if (lookup_only)
    throw Error(ST_UNSUPPORTED, "the circuit's kernel is not built on this context: solve it once with the "
                                "same options first (building and loading a kernel allocates)");
  auto t0 = clk::now();
  double cms = 0.0;
  if (native_ok) {
    try {
      NativeKernel nk = build_native(c, kind, vb, lane_shift, outs, group_exits);
      ch.mod = ctx.jit.get_native(nk.kernel_name, nk.body, nk.asm_text, &cms);
      ch.info = nk.info;
      ch.native = true;
      ch.vgpr_class = nk.vgpr_class;
    } catch (const NativeUnfit&) {
    }
  }
  if (!ch.native) {
    std::string src = hiprtc_src(&ch.info);
    ch.mod = ctx.jit.get(src, hiprtc_kernel, &cms);
  }
  if (compile_ms) *compile_ms = ms_since(t0);
  ctx.jit.put_built(fp, ch);
  return ch;
}

}  // namespace

// 128-bit fingerprint of a circuit's content (two independent 64-bit hashes)
std::string circuit_fingerprint(const Circuit& c) {
  uint64_t h1 = 0xcbf29ce484222325ull, h2 = 0x9e3779b97f4a7c15ull;
  auto mix = [&](uint64_t v) {
    h1 = (h1 ^ v) * 0x100000001b3ull;
    h2 ^= v + 0x9e3779b97f4a7c15ull + (h2 << 6) + (h2 >> 2);
    h2 *= 0xff51afd7ed558ccdull;
  };
  mix(c.n);
  mix(c.gates.size());
  for (const Gate& g : c.gates) mix(((uint64_t)g.i0 << 34) | ((uint64_t)g.i1 << 2) | g.func);
  mix(c.outputs.size());
  for (const Output& o : c.outputs) mix(((uint64_t)o.wire << 1) | (o.neg ? 1u : 0u));
  char buf[40];
  std::snprintf(buf, sizeof(buf), "%016llx%016llx", (unsigned long long)h1, (unsigned long long)h2);
  return buf;
}

// ---------------------------------------------------------------------------
// SAT
// ---------------------------------------------------------------------------
void sat_solve(Lane& ctx, const Circuit& c, const ns_sat_opts& o, ns_sat_result& r, ns_sat_progress_fn progress,
               void* user) {
  auto t0 = clk::now();
  std::memset(&r, 0, sizeof(r));
  if (c.outputs.size() != 1) throw Error(ST_INVALID, "SAT needs a circuit with exactly one output");
  if (c.n == 0 || c.n > 64) throw Error(ST_UNSUPPORTED, "SAT supports 1..64 inputs");
  NS_HIP(hipSetDevice(ctx.device));
  const uint64_t total_words = c.n >= 5 ? (c.n == 64 ? (1ull << 59) : (1ull << (c.n - 5))) : 1ull;
  Geometry g = sat_geometry(total_words, o);
  std::vector<uint32_t> vb(c.n);
  for (uint32_t k = 0; k < c.n; k++) vb[k] = k;
  std::string fp;
  Chosen ch = choose_kernel(
      ctx, c, NativeKind::SAT, vb, 0, {0}, g, "",
      [&](GenInfo* info) { return gen_sat_kernel(c, g, info); }, "ns_sat", &r.compile_ms, &fp, false,
      (o.flags & NS_SAT_GROUP_EXITS) != 0);
  // Words per lane and threads per block are run-time parameters of the native kernel: a long
  // search times its first launches with each candidate pair and keeps the fastest (cached
  // with the kernel). Every word is still evaluated exactly once, so the answers do not
  // depend on them.
  const uint64_t cw_tune = chunk_words(o.elem_bits, g, total_words);
  const bool tune = ch.native && o.words_per_lane == 0 && o.group_len == 0 &&
                    (o.flags & NS_SAT_STOP_AT_FIRST) == 0 && total_words >= 16 * cw_tune &&
                    cw_tune % ((uint64_t)g.block * 16) == 0;
  if (ch.native && ch.tuned_wpl && o.words_per_lane == 0 && o.group_len == 0 &&
      cw_tune % ((uint64_t)ch.tuned_block * ch.tuned_wpl) == 0) {
    g.wpl = ch.tuned_wpl;
    g.block = ch.tuned_block;
  }
  auto mod = ch.mod;
  r.cone_gates = ch.info.cone_gates;
  r.valu_ops = ch.info.valu_ops;
  r.scalar_ops = ch.info.scalar_ops;
  r.backend = ch.native ? 1 : 0;
  r.vgpr_class = ch.vgpr_class;
  uint32_t wpl = g.wpl, valid = word_valid_mask(c.n);
  unsigned long long* res = ctx.stats.as<unsigned long long>();
  const uint64_t cw = chunk_words(o.elem_bits, g, total_words);
  const bool stop_first = (o.flags & NS_SAT_STOP_AT_FIRST) != 0;
  unsigned long long host[2] = {0, ~0ull};
  double kms = 0;
  uint64_t w0 = 0;
  r.complete = 1;
  // A repeated complete search (kernel from the cache, geometry settled) is captured once as
  // a HIP graph (result reset + every word-range launch) and replayed afterwards.
  const bool repeated = r.compile_ms == 0.0;
  RangeProgress prog(ctx, progress, user, (total_words + cw - 1) / cw, ilog2(cw) + 5);
  prog.start();  // the kernel is built, nothing is launched yet
  if (prog.cancelled()) {
    r.complete = 0;
    r.wall_ms = ms_since(t0);
    return;
  }
  if (repeated && !stop_first && !prog.active() && !(tune && !ch.tuned_wpl)) {
    const std::string key = fp + "/graph/" + std::to_string(g.block) + "/" + std::to_string(g.wpl) + "/" +
                            std::to_string(cw) + "/" + std::to_string(total_words);
    hipGraphExec_t exec = ctx.find_graph(key);
    if (!exec) {
      NS_HIP(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
      hipGraph_t graph = nullptr;
      try {
        reset_counts(ctx);
        for (uint64_t a = 0; a < total_words; a += cw) {
          uint64_t words = total_words - a < cw ? total_words - a : cw;
          void* extra[3] = {&res, &wpl, &valid};
          launch_range(ctx, mod->fn, g, a, words, extra, ch.native ? 3 : 1, total_words);
        }
      } catch (...) {
        (void)hipStreamEndCapture(ctx.stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      NS_HIP(hipStreamEndCapture(ctx.stream, &graph));
      hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      NS_HIP(ie);
      ctx.put_graph(key, exec, mod);
    }
    NS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
    NS_HIP(hipGraphLaunch(exec, ctx.stream));
    NS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
    CountsCopy cc;
    read_counts_async(ctx, cc);
    NS_HIP(hipStreamSynchronize(ctx.stream));
    read_counts_done(cc, host);
    r.launches = (uint32_t)((total_words + cw - 1) / cw);
    r.count = host[0];
    r.found = host[0] != 0;
    r.lowest = r.found ? host[1] : 0;
    r.kernel_ms = ctx.elapsed_ms();
    r.wall_ms = ms_since(t0);
    return;
  }
  reset_counts(ctx);
  NS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  if (tune && !ch.tuned_wpl) {
    // the autotuner: each (threads per block, words per lane) candidate runs one launch of
    // the search's own word ranges, timed with events; the fastest is kept with the kernel
    static const uint32_t cand[kTuneCandidates][2] = {{256, 1}, {256, 2}, {256, 4}, {256, 8}, {256, 16},
                                                      {128, 2}, {128, 4}, {128, 8}, {64, 4},  {64, 8}};
    hipEvent_t ev[kTuneCandidates + 1];
    for (auto& e : ev) NS_HIP(hipEventCreate(&e));
    NS_HIP(hipEventRecord(ev[0], ctx.stream));
    for (uint32_t k = 0; k < kTuneCandidates; k++) {
      Geometry gk = g;
      gk.block = cand[k][0];
      gk.wpl = cand[k][1];
      uint32_t wk = gk.wpl;
      void* extra[3] = {&res, &wk, &valid};
      launch_range(ctx, mod->fn, gk, w0, cw, extra, 3, total_words);
      NS_HIP(hipEventRecord(ev[k + 1], ctx.stream));
      r.launches++;
      w0 += cw;
    }
    NS_HIP(hipEventSynchronize(ev[kTuneCandidates]));
    float best = 0.f;
    for (uint32_t k = 0; k < kTuneCandidates; k++) {
      float ms = 0.f;
      NS_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      if (k == 0 || ms < best) {
        best = ms;
        g.block = cand[k][0];
        g.wpl = cand[k][1];
      }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    wpl = g.wpl;
    ch.tuned_wpl = g.wpl;
    ch.tuned_block = g.block;
    ctx.jit.put_built(fp, ch);
    for (uint32_t k = 0; k < kTuneCandidates; k++) prog.completed(k);
  }
  while (w0 < total_words && !prog.cancelled()) {
    uint64_t words = total_words - w0 < cw ? total_words - w0 : cw;
    void* extra[3] = {&res, &wpl, &valid};
    launch_range(ctx, mod->fn, g, w0, words, extra, ch.native ? 3 : 1, total_words);
    r.launches++;
    w0 += cw;
    if (stop_first) {
      // some hit so far: the lowest one is set
      NS_HIP(hipMemcpyAsync(host, res, sizeof(host), hipMemcpyDeviceToHost, ctx.stream));
      NS_HIP(hipStreamSynchronize(ctx.stream));
      prog.completed(r.launches - 1);
      if (host[1] != ~0ull) {
        r.complete = w0 >= total_words;
        break;
      }
    } else {
      prog.launched(r.launches - 1);
    }
  }
  prog.drain();
  if (prog.cancelled() && w0 < total_words) r.complete = 0;
  NS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  CountsCopy cc;
  read_counts_async(ctx, cc);
  NS_HIP(hipStreamSynchronize(ctx.stream));
  read_counts_done(cc, host);
  kms = ctx.elapsed_ms();
  r.count = host[0];
  r.found = host[0] != 0;
  r.lowest = r.found ? host[1] : 0;
  r.kernel_ms = kms;
  r.wall_ms = ms_since(t0);
}

// The same search with the counters in `scratch` (the file-scope device array of the
// allocation-free entry point) and its kernel taken from the context's cache: no device
// or host allocation, no module load, no graph. Counters are cleared by fill kernels and
// read back into the lane's pinned block.
void sat_solve_noalloc(Lane& ctx, const Circuit& c, const ns_sat_opts& o, ns_sat_result& r,
                       unsigned long long* scratch) {
  auto t0 = clk::now();
  std::memset(&r, 0, sizeof(r));
  if (!scratch) throw Error(ST_INTERNAL, "the allocation-free counters are not set up on this device");
  if (c.outputs.size() != 1) throw Error(ST_INVALID, "SAT needs a circuit with exactly one output");
  if (c.n == 0 || c.n > 64) throw Error(ST_UNSUPPORTED, "SAT supports 1..64 inputs");
  static_assert(kSlot0 + kSlots * kSlotStride <= kNoallocScratchWords, "scratch holds the SAT counters");
  static_assert(sizeof(CountsCopy) <= Lane::kPinnedBytes, "pinned block holds the counters");
  NS_HIP(hipSetDevice(ctx.device));
  const uint64_t total_words = c.n >= 5 ? (c.n == 64 ? (1ull << 59) : (1ull << (c.n - 5))) : 1ull;
  Geometry g = sat_geometry(total_words, o);
  std::vector<uint32_t> vb(c.n);
  for (uint32_t k = 0; k < c.n; k++) vb[k] = k;
  Chosen ch = choose_kernel(
      ctx, c, NativeKind::SAT, vb, 0, {0}, g, "", [&](GenInfo* info) { return gen_sat_kernel(c, g, info); },
      "ns_sat", &r.compile_ms, nullptr, true, (o.flags & NS_SAT_GROUP_EXITS) != 0);
  const uint64_t cw = chunk_words(o.elem_bits, g, total_words);
  if (ch.native && ch.tuned_wpl && o.words_per_lane == 0 && o.group_len == 0 &&
      cw % ((uint64_t)ch.tuned_block * ch.tuned_wpl) == 0) {
    g.wpl = ch.tuned_wpl;
    g.block = ch.tuned_block;
  }
  r.cone_gates = ch.info.cone_gates;
  r.valu_ops = ch.info.valu_ops;
  r.scalar_ops = ch.info.scalar_ops;
  r.backend = ch.native ? 1 : 0;
  r.vgpr_class = ch.vgpr_class;
  uint32_t wpl = g.wpl, valid = word_valid_mask(c.n);
  unsigned long long* res = scratch;
  const bool stop_first = (o.flags & NS_SAT_STOP_AT_FIRST) != 0;
  CountsCopy* cc = reinterpret_cast<CountsCopy*>(ctx.pinned);
  NS_HIP(launch_fill_u64(ctx.stream, res, 0ull, 1));
  NS_HIP(launch_fill_u64(ctx.stream, res + 1, ~0ull, 1));
  NS_HIP(launch_fill_u64(ctx.stream, res + kSlot0, 0ull, kSlots * kSlotStride));
  NS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  r.complete = 1;
  for (uint64_t w0 = 0; w0 < total_words; w0 += cw) {
    const uint64_t words = total_words - w0 < cw ? total_words - w0 : cw;
    void* extra[3] = {&res, &wpl, &valid};
    launch_range(ctx, ch.mod->fn, g, w0, words, extra, ch.native ? 3 : 1, total_words);
    r.launches++;
    if (stop_first) {
      NS_HIP(hipMemcpyAsync(cc->head, res, sizeof(cc->head), hipMemcpyDeviceToHost, ctx.stream));
      NS_HIP(hipStreamSynchronize(ctx.stream));
      if (cc->head[1] != ~0ull) {
        r.complete = w0 + words >= total_words;
        break;
      }
    }
  }
  NS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  read_counts_async(ctx, *cc, res);
  NS_HIP(hipStreamSynchronize(ctx.stream));
  unsigned long long host[2];
  read_counts_done(*cc, host);
  r.count = host[0];
  r.found = host[0] != 0;
  r.lowest = r.found ? host[1] : 0;
  r.kernel_ms = ctx.elapsed_ms();
  r.wall_ms = ms_since(t0);
}

// ---------------------------------------------------------------------------
// QSAT
// ---------------------------------------------------------------------------
void qsat_solve(Lane& ctx, const Circuit& c, const std::vector<uint8_t>& q, const ns_qsat_opts& o,
                ns_qsat_result& r, uint32_t* bitmap, uint64_t bitmap_words, ns_sat_progress_fn progress,
                void* user) {
  auto t0 = clk::now();
  std::memset(&r, 0, sizeof(r));
  if (c.outputs.size() != 1) throw Error(ST_INVALID, "QSAT needs a circuit with exactly one output");
  if (c.n == 0 || c.n > 60) throw Error(ST_UNSUPPORTED, "QSAT supports 1..60 inputs");
  if (q.size() != c.n) throw Error(ST_INVALID, "one quantifier per input is required");
  NS_HIP(hipSetDevice(ctx.device));
  const uint32_t n = c.n;
  uint32_t fb = 0;
  while (fb < n && q[fb] == q[0]) fb++;
  const uint32_t L = n - fb;
  const bool q0all = q[0] != 0;
  r.first_bits = fb;
  r.first_all = q0all;
  if (bitmap) {
    if (fb > 36) throw Error(ST_UNSUPPORTED, "bitmap of the outermost block limited to 2^36 bits");
    uint64_t need = fb >= 5 ? (1ull << (fb - 5)) : 1ull;
    if (bitmap_words < need) throw Error(ST_INVALID, "bitmap buffer too small");
  }
  const uint64_t total_words = n >= 5 ? (1ull << (n - 5)) : 1ull;
  // as for SAT, a requested range size is kept exactly (one block fits in a range)
  uint64_t gcap = total_words;
  if (o.elem_bits >= 5 && o.elem_bits <= 40 && (1ull << (o.elem_bits - 5)) < gcap) gcap = 1ull << (o.elem_bits - 5);
  Geometry g = pick_geometry(gcap, o.group_len, o.words_per_lane, true);
  const uint32_t jb = ilog2(g.wpl);
  const bool case_a = L >= 5 + jb;
  std::vector<uint32_t> vb(n);
  for (uint32_t k = 0; k < n; k++) vb[k] = n - 1 - k;
  std::string qkey(q.size(), '0');  // the generated HIP source depends on the prefix
  for (size_t k = 0; k < q.size(); k++) qkey[k] = q[k] ? '1' : '0';
  Chosen ch = choose_kernel(
      ctx, c, case_a ? NativeKind::QSAT_A : NativeKind::QSAT_B, vb, jb, {0}, g, qkey,
      [&](GenInfo* info) { return gen_qsat_kernel(c, q, g, info); }, "ns_qsat", &r.compile_ms);
  auto mod = ch.mod;
  r.cone_gates = ch.info.cone_gates;
  r.valu_ops = ch.info.valu_ops;
  r.scalar_ops = ch.info.scalar_ops;
  r.backend = ch.native ? 1 : 0;
  r.vgpr_class = ch.vgpr_class;
  // runtime parameters of the native templates: level ops (bit l = AND for level l)
  uint32_t wpl = g.wpl, jbv = jb, opm = 0, Lv = L, q0v = q0all ? 1u : 0u, valid = word_valid_mask(n);
  for (uint32_t l = 0; l < 32 && l < n; l++)
    if (q[n - 1 - l]) opm |= 1u << l;
  unsigned long long* stats = ctx.stats.as<unsigned long long>();
  reset_counts(ctx);
  const uint64_t cw = chunk_words(o.elem_bits, g, total_words);
  unsigned int* dev_bitmap = nullptr;  // final 2^fb-bit array on the device
  RangeProgress prog(ctx, progress, user, (total_words + cw - 1) / cw, ilog2(cw) + 5);
  prog.start();
  auto check_cancel = [&]() {
    if (prog.cancelled()) {
      prog.drain();
      throw Error(ST_CANCELLED, "QSAT search cancelled by its progress callback");
    }
  };
  check_cancel();
  uint64_t step = 0;
  NS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  if (case_a) {
    // phase 1: one bit per thread
    const uint32_t m0 = n - 5 - jb;
    const uint64_t xbytes = ((1ull << m0) + 63) / 64 * 8;
    ctx.buf_a.ensure(xbytes);
    unsigned long long* X = ctx.buf_a.as<unsigned long long>();
    for (uint64_t w0 = 0; w0 < total_words; w0 += cw) {
      uint64_t words = total_words - w0 < cw ? total_words - w0 : cw;
      void* extra[4] = {&X, &wpl, &jbv, &opm};
      launch_range(ctx, mod->fn, g, w0, words, extra, ch.native ? 4 : 1, total_words);
      prog.launched(step++);
      check_cancel();
    }
    prog.drain();
    check_cancel();
    // phase 2: reduction passes over the bit array
    uint32_t m = m0, remaining = L - 5 - jb, level = 5 + jb;
    const unsigned int* in = reinterpret_cast<unsigned int*>(X);
    DevBuf* outs[2] = {&ctx.buf_b, &ctx.buf_c};
    int which = 0;
    while (true) {
      uint32_t k = remaining < 11 ? remaining : 11;
      bool final_pass = k == remaining;
      uint32_t mo = m - k;
      uint64_t owords = mo >= 5 ? (1ull << (mo - 5)) : 1ull;
      outs[which]->ensure((owords + 1) * 4);
      unsigned int* out = outs[which]->as<unsigned int>();
      NS_HIP(launch_fill_u32(ctx.stream, out, 0u, owords + 1));
      uint32_t opmask = 0;
      for (uint32_t l = 0; l < k; l++)
        if (q[n - 1 - (level + l)]) opmask |= 1u << l;
      NS_HIP(launch_qsat_reduce(ctx.stream, in, m, k, opmask, out, final_pass, q0all, stats));
      r.passes++;
      if (final_pass) {
        dev_bitmap = out;
        break;
      }
      in = out;
      m = mo;
      remaining -= k;
      level += k;
      which ^= 1;
    }
  } else {
    if (bitmap) {
      uint64_t bw = fb >= 5 ? (1ull << (fb - 5)) : 1ull;
      ctx.buf_b.ensure(bw * 4);
      dev_bitmap = ctx.buf_b.as<unsigned int>();
      NS_HIP(launch_fill_u32(ctx.stream, dev_bitmap, 0u, bw));
    }
    for (uint64_t w0 = 0; w0 < total_words; w0 += cw) {
      uint64_t words = total_words - w0 < cw ? total_words - w0 : cw;
      void* extra[8] = {&stats, &dev_bitmap, &wpl, &jbv, &opm, &Lv, &q0v, &valid};
      launch_range(ctx, mod->fn, g, w0, words, extra, ch.native ? 8 : 2, total_words);
      prog.launched(step++);
      check_cancel();
    }
    prog.drain();
    check_cancel();
  }
  NS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  unsigned long long host[2];
  CountsCopy cc;
  read_counts_async(ctx, cc);
  if (bitmap && dev_bitmap) {
    uint64_t need = fb >= 5 ? (1ull << (fb - 5)) : 1ull;
    NS_HIP(hipMemcpyAsync(bitmap, dev_bitmap, need * 4, hipMemcpyDeviceToHost, ctx.stream));
  }
  NS_HIP(hipStreamSynchronize(ctx.stream));
  read_counts_done(cc, host);
  if (bitmap && fb < 5) bitmap[0] &= (uint32_t)((1ull << (1u << fb)) - 1);
  r.kernel_ms = ctx.elapsed_ms();
  r.witness_count = host[0];
  r.has_witness = host[0] != 0;
  r.witness = r.has_witness ? host[1] : 0;
  r.truth = (r.has_witness ? 1u : 0u) ^ (q0all ? 1u : 0u);
  r.wall_ms = ms_since(t0);
}

// ---------------------------------------------------------------------------
// Next-state table and SATOSTOP
// ---------------------------------------------------------------------------
namespace {

struct TableKernelInfo {
  GenInfo info;
  bool native = false;
  uint32_t vgpr_class = 0;
};

// Builds the u32 table of c's outputs (at most 32; output i at bit i) into dst; returns
// kernel ms; sets *anylast when the last output is 1 for some state.
double build_table_dev(Lane& ctx, const Circuit& c, uint32_t group_len, uint32_t wpl, DevBuf& dst,
                       bool* anylast, double* compile_ms, uint32_t* cone, TableKernelInfo* info_out = nullptr) {
  const uint32_t n = c.n;
  if (n == 0 || n > 33) throw Error(ST_UNSUPPORTED, "next-state tables support 1..33 state bits");
  if (c.outputs.empty() || c.outputs.size() > 32) throw Error(ST_UNSUPPORTED, "table entries hold at most 32 outputs");
  const uint64_t states = 1ull << n;
  const uint64_t total_words = n >= 5 ? (1ull << (n - 5)) : 1ull;
  Geometry g = pick_geometry(total_words, group_len, wpl ? wpl : 2, false);
  std::vector<uint32_t> vb(n), outs;
  for (uint32_t k = 0; k < n; k++) vb[k] = k;
  for (uint32_t k = 0; k < c.outputs.size(); k++) outs.push_back(k);
  Chosen ch = choose_kernel(
      ctx, c, NativeKind::TABLE, vb, 0, outs, g, "",
      [&](GenInfo* info) { return gen_table_kernel(c, g, info); }, "ns_table", compile_ms);
  auto mod = ch.mod;
  if (cone) *cone = ch.info.cone_gates;
  if (info_out) {
    info_out->info = ch.info;
    info_out->native = ch.native;
    info_out->vgpr_class = ch.vgpr_class;
  }
  uint32_t wpl_v = g.wpl, last = (uint32_t)c.outputs.size() - 1, valid = word_valid_mask(n),
           small = n < 5 ? (1u << n) : 0u;
  dst.ensure((states < 32 ? 32 : states) * 4);
  unsigned int* T = dst.as<unsigned int>();
  unsigned int* flag = reinterpret_cast<unsigned int*>(ctx.stats.as<char>() + 64);
  NS_HIP(launch_fill_u32(ctx.stream, flag, 0u, 1));
  const uint64_t cw = chunk_words(30, g, total_words);
  NS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  for (uint64_t w0 = 0; w0 < total_words; w0 += cw) {
    uint64_t words = total_words - w0 < cw ? total_words - w0 : cw;
    unsigned int* Tc = T + (w0 << 5);  // the kernel writes relative to its first word
    void* extra[6] = {&Tc, &flag, &wpl_v, &last, &valid, &small};
    launch_range(ctx, mod->fn, g, w0, words, extra, ch.native ? 6 : 2, total_words);
  }
  NS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  unsigned int hf = 0;
  NS_HIP(hipMemcpyAsync(&hf, flag, 4, hipMemcpyDeviceToHost, ctx.stream));
  NS_HIP(hipStreamSynchronize(ctx.stream));
  *anylast = hf != 0;
  return ctx.elapsed_ms();
}

}  // namespace

double table_build(Lane& ctx, const Circuit& c, uint32_t* table_out, uint64_t table_len) {
  NS_HIP(hipSetDevice(ctx.device));
  bool any = false;
  double cms = 0;
  if (c.n > 31) throw Error(ST_UNSUPPORTED, "u32 next-state tables support 1..31 state bits");
  double kms = build_table_dev(ctx, c, 0, 0, ctx.tab_a, &any, &cms, nullptr);
  const uint64_t states = 1ull << c.n;
  if (table_out) {
    if (table_len < states) throw Error(ST_INVALID, "table buffer too small");
    NS_HIP(hipMemcpyAsync(table_out, ctx.tab_a.p, states * 4, hipMemcpyDeviceToHost, ctx.stream));
    NS_HIP(hipStreamSynchronize(ctx.stream));
  }
  return kms;
}

// Passes over the whole table, then walks from the starts. After P passes every entry that
// has not stopped jumps 2^P evaluations, so 2^(C-P) jumps per start cover the same 2^C
// evaluations as C passes, with the same answers. A pass costs 2^n random gathers
// (throughput-bound, ~27 ps each); a walk costs 2^(C-P) dependent gathers per start
// (latency-bound, ~0.7 us each, while the starts are few). P minimises the modelled time;
// NS_JUMP_PASSES overrides it.
uint32_t choose_passes(uint32_t n, uint32_t unknowns, uint32_t cover) {
  if (cover == 0) return 0;
  if (const char* e = std::getenv("NS_JUMP_PASSES")) {
    long v = std::atol(e);
    if (v > 0) return v > (long)cover ? cover : (uint32_t)v;
  }
  const double per_gather = 27e-12, per_jump = 0.7e-6;
  uint32_t best = cover;
  double best_t = 0;
  for (uint32_t p = 1; p <= cover; p++) {
    const double jumps = std::ldexp(1.0, (int)(cover - p));
    const double t = p * std::ldexp(per_gather, (int)n) +
                     std::max(jumps * per_jump, std::ldexp(jumps * per_gather, (int)unknowns));
    if (p == 1 || t < best_t) {
      best = p;
      best_t = t;
    }
  }
  return best;
}

namespace {

// the circuit with only the outputs [from, to)
Circuit with_outputs(const Circuit& c, uint32_t from, uint32_t to) {
  Circuit s = c;
  s.outputs.assign(c.outputs.begin() + from, c.outputs.begin() + to);
  return s;
}

}  // namespace

// Waypoint spacing 2^k: at most 2^22 waypoints (their records stay cache-resident), states
// 2^8 apart where the state space allows, and k <= n - unknowns so that every start is a
// waypoint. NS_WAYPOINT_BITS overrides it.
uint32_t waypoint_bits(uint32_t n, uint32_t unknowns) {
  if (const char* e = std::getenv("NS_WAYPOINT_BITS")) {
    const long v = std::atol(e);
    if (v >= 0 && v <= (long)n) return (uint32_t)v;
  }
  const uint32_t want = std::max(n > 22 ? n - 22 : 0u, std::min(8u, n));
  return std::min(n - unknowns, want);
}

namespace {

// Packed next-state image of c: 2^n entries of n + 1 bits (next state, STOP at bit n) in the
// reference's MemImage layout, built 2^cb states at a time: the table kernels write outputs
// 0..31 (and 32..n) of the chunk to tab_a (tab_e) and the pack kernel writes the chunk's
// words to chunk_out(ci); after(ci) runs once they are enqueued. Returns kernel ms of the
// whole build (the stream is synchronised at the end); *any = some state's STOP output is 1.
double build_packed_chunks(Lane& ctx, const Circuit& c, const ns_satostop_opts& o, uint32_t cb,
                           const std::function<unsigned long long*(uint64_t)>& chunk_out,
                           const std::function<void(uint64_t)>& after, bool* any, double* compile_ms,
                           uint32_t* cone, TableKernelInfo* info) {
  const uint32_t n = c.n, width = n + 1;
  if (n < 6 || n > 62 || cb < 6 || cb > n) throw Error(ST_INTERNAL, "packed images take 6..62 state bits");
  const bool two = width > 32;
  const uint32_t split = two ? 32 : width;
  const uint64_t chunk = 1ull << cb, chunks = 1ull << (n - cb), words_chunk = chunk >> 5;
  Geometry g = pick_geometry(words_chunk, o.group_len, o.words_per_lane ? o.words_per_lane : 2, false);
  std::vector<uint32_t> vb(n);
  for (uint32_t k = 0; k < n; k++) vb[k] = k;
  const Circuit parts[2] = {with_outputs(c, 0, split), two ? with_outputs(c, split, width) : Circuit()};
  Chosen kern[2];
  *compile_ms = 0;
  *cone = 0;
  info->info = GenInfo();
  info->native = true;
  info->vgpr_class = 0;
  for (int h = 0; h < (two ? 2 : 1); h++) {
    std::vector<uint32_t> outs;
    for (uint32_t k = 0; k < parts[h].outputs.size(); k++) outs.push_back(k);
    double cms = 0;
    kern[h] = choose_kernel(
        ctx, parts[h], NativeKind::TABLE, vb, 0, outs, g, "",
        [&](GenInfo* gi) { return gen_table_kernel(parts[h], g, gi); }, "ns_table", &cms);
    *compile_ms += cms;
    *cone += kern[h].info.cone_gates;
    info->info.valu_ops += kern[h].info.valu_ops;
    if (!kern[h].native) info->native = false;
    info->vgpr_class = std::max(info->vgpr_class, kern[h].vgpr_class);
  }
  const uint64_t wpb = g.words_per_block();
  const uint64_t tab_bytes = (words_chunk + wpb - 1) / wpb * wpb * 32 * 4;
  ctx.tab_a.ensure(tab_bytes);
  if (two) ctx.tab_e.ensure(tab_bytes);
  unsigned int* flags = reinterpret_cast<unsigned int*>(ctx.stats.as<char>() + 64);  // [0] STOP seen, [1] unused
  NS_HIP(launch_fill_u32(ctx.stream, flags, 0u, 2));
  hipEvent_t e0, e1;
  NS_HIP(hipEventCreate(&e0));
  NS_HIP(hipEventCreate(&e1));
  NS_HIP(hipEventRecord(e0, ctx.stream));
  for (uint64_t ci = 0; ci < chunks; ci++) {
    const uint64_t w0 = ci * words_chunk;
    for (int h = 0; h < (two ? 2 : 1); h++) {
      unsigned int* T = h ? ctx.tab_e.as<unsigned int>() : ctx.tab_a.as<unsigned int>();
      // the STOP output is the last output of the last part
      unsigned int* flag = h == (two ? 1 : 0) ? flags : flags + 1;
      uint32_t wpl_v = g.wpl, last = (uint32_t)parts[h].outputs.size() - 1, valid = 0xffffffffu, small = 0;
      void* extra[6] = {&T, &flag, &wpl_v, &last, &valid, &small};
      launch_range(ctx, kern[h].mod->fn, g, w0, words_chunk, extra, kern[h].native ? 6 : 2, w0 + words_chunk);
    }
    NS_HIP(launch_pack(ctx.stream, ctx.tab_a.as<unsigned int>(), two ? ctx.tab_e.as<unsigned int>() : nullptr, split,
                       width, chunk, chunk_out(ci)));
    if (after) after(ci);
  }
  NS_HIP(hipEventRecord(e1, ctx.stream));
  unsigned int hf = 0;
  NS_HIP(hipMemcpyAsync(&hf, flags, 4, hipMemcpyDeviceToHost, ctx.stream));
  NS_HIP(hipStreamSynchronize(ctx.stream));
  float ms = 0.f;
  NS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *any = hf != 0;
  return ms;
}

// The whole packed image in ctx.tab_img (one word of padding after the last entry).
double build_packed_table(Lane& ctx, const Circuit& c, const ns_satostop_opts& o, bool* any, double* compile_ms,
                          uint32_t* cone, TableKernelInfo* info) {
  const uint32_t n = c.n, width = n + 1;
  if (n > 35) throw Error(ST_INTERNAL, "device images take at most 35 state bits");
  const uint32_t cb = std::min(n, 30u);
  const uint64_t total_words = ((1ull << n) * width) >> 6;
  ctx.tab_img.ensure((total_words + 1) * 8);
  unsigned long long* img = ctx.tab_img.as<unsigned long long>();
  NS_HIP(launch_fill_u64(ctx.stream, img + total_words, 0ull, 1));
  return build_packed_chunks(
      ctx, c, o, cb, [&](uint64_t ci) { return img + (((ci << cb) * width) >> 6); }, nullptr, any, compile_ms, cone,
      info);
}

// The full search as waypoint walks (see k_wp_walk) over the table tv: fills the per-start
// arrays and host[0..1] = (stopping starts, lowest one). Returns false, with nothing
// changed, when a walk exceeds its evaluation limit (a long path or cycle without
// waypoints); the passes then answer instead.
// A walk kernel enqueued on ctx.stream: (records, flag, evaluation slots, waypoints, limit).
using WalkLaunch = std::function<void(void*, unsigned int*, unsigned long long*, uint64_t, uint64_t)>;

// check_records: the walk kernel also counts the records it writes (word 1 of each count
// slot); every waypoint must have exactly one.
bool waypoint_search(Lane& ctx, uint32_t n, uint32_t unknowns, uint32_t k, const WalkLaunch& walk,
                     ns_satostop_result& r, unsigned long long* d_steps, unsigned long long* d_ends,
                     unsigned long long* stats, unsigned long long host[2], bool check_records = false) {
  const uint64_t count = 1ull << (n - k);
  // a walk ends after this many evaluations (far above the 2^k a walk takes on average)
  uint64_t limit = std::max<uint64_t>(1ull << 18, 1ull << std::min(k + 6, 31u));
  if (const char* e = std::getenv("NS_WAYPOINT_LIMIT")) {
    const long long v = std::atoll(e);
    if (v > 0) limit = (uint64_t)v;
  }
  ctx.tab_b.ensure(count * 16);
  ctx.tab_c.ensure(count * 16);
  void* cur = ctx.tab_b.p;
  void* nxt = ctx.tab_c.p;
  unsigned int* flag = reinterpret_cast<unsigned int*>(ctx.stats.as<char>() + 192);
  reset_counts(ctx);  // the walks count their evaluations in the count slots
  NS_HIP(launch_fill_u32(ctx.stream, flag, 0u, 1));
  NS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  walk(cur, flag, stats + kSlot0, count, limit);
  NS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  unsigned int hf = 0;
  NS_HIP(hipMemcpyAsync(&hf, flag, 4, hipMemcpyDeviceToHost, ctx.stream));
  CountsCopy cc;
  read_counts_async(ctx, cc);
  NS_HIP(hipStreamSynchronize(ctx.stream));
  const double walk_ms = ctx.elapsed_ms();
  if (hf != 0) return false;
  unsigned long long ev[2];
  read_counts_done(cc, ev);
  if (check_records) {
    uint64_t recs = 0;
    for (uint64_t s = 0; s < kSlots; s++) recs += cc.slots[s * kSlotStride + 1];
    if (recs != count)
      throw Error(ST_INTERNAL, "waypoint walks wrote " + std::to_string(recs) + " records for " +
                                   std::to_string(count) + " waypoints");
  }
  r.waypoint_bits = k;
  r.walk_ms = walk_ms;
  r.walk_evaluations = ev[0];
  // Coarser levels while the records are many: each chase pass keeps the waypoints whose
  // index has its low 4 bits zero (no coarser than the starts' spacing n - unknowns) and
  // follows the finer records between them (one dependent read per record); pointer jumping
  // then runs over 16 times fewer records per level. NS_WAYPOINT_CHASE=0: no chase.
  uint32_t kc = k;
  uint64_t cnt = count;
  const char* ce = std::getenv("NS_WAYPOINT_CHASE");
  const bool chase = !(ce && *ce == '0');
  constexpr uint32_t kChaseBits = 4, kChaseMaxHops = 1u << 16;
  while (chase && cnt > (1ull << 20) && kc + kChaseBits <= n - unknowns) {
    NS_HIP(launch_fill_u32(ctx.stream, flag, 0u, 1));
    NS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
    NS_HIP(launch_wp_chase(ctx.stream, cur, kChaseBits, nxt, cnt >> kChaseBits, kChaseMaxHops, flag));
    NS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
    NS_HIP(hipMemcpyAsync(&hf, flag, 4, hipMemcpyDeviceToHost, ctx.stream));
    NS_HIP(hipStreamSynchronize(ctx.stream));
    const double ms = ctx.elapsed_ms();
    ctx.pass_ms.push_back(ms);
    r.jump_ms += ms;
    r.waypoint_passes++;
    if (hf != 0) break;  // a chain of more than kChaseMaxHops records: jump over this level
    std::swap(cur, nxt);
    cnt >>= kChaseBits;
    kc += kChaseBits;
  }
  // a path visits at most 2^(n-kc) waypoints before it repeats one
  for (uint32_t p = 0; p < n - kc; p++) {
    NS_HIP(launch_fill_u32(ctx.stream, flag, 0u, 1));
    NS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
    NS_HIP(launch_wp_jump(ctx.stream, cur, nxt, cnt, flag));
    NS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
    NS_HIP(hipMemcpyAsync(&hf, flag, 4, hipMemcpyDeviceToHost, ctx.stream));
    NS_HIP(hipStreamSynchronize(ctx.stream));
    const double ms = ctx.elapsed_ms();
    ctx.pass_ms.push_back(ms);
    r.jump_ms += ms;
    r.waypoint_passes++;
    std::swap(cur, nxt);
    if (hf == 0) break;
  }
  NS_HIP(launch_fill_u64(ctx.stream, stats, 0ull, 1));
  NS_HIP(launch_fill_u64(ctx.stream, stats + 1, ~0ull, 1));
  NS_HIP(launch_wp_starts(ctx.stream, cur, n, unknowns, kc, d_steps, d_ends, stats));
  NS_HIP(hipMemcpyAsync(host, stats, 16, hipMemcpyDeviceToHost, ctx.stream));
  NS_HIP(hipStreamSynchronize(ctx.stream));
  return true;
}

// the waypoint walks over a table in device memory
WalkLaunch table_walk(Lane& ctx, uint32_t n, uint32_t k, const TableView& tv) {
  return [&ctx, n, k, tv](void* out, unsigned int* flag, unsigned long long* evals, uint64_t count, uint64_t limit) {
    NS_HIP(launch_wp_walk(ctx.stream, tv, n, k, count, limit, out, flag, evals));
  };
}

// Waypoint spacing of the walks without a table (k <= n - unknowns). The walks still running
// when the waypoints run out last ~2^k evaluations more (the tail: each wave waits for the
// longest of its 2,048 slots, ~8 x 2^k), so the waypoints should be many: 2^(n-k) >= 2^25,
// at least 8 walks for each of the 2^22 slots, and 2^5 states apart or more (the records and
// refills of shorter walks cost more than the tail). The chase passes keep the record jumps
// cheap (without them, k = 8 against 5 was best at 30 state bits). Measured on the
// Feistel maps: 1.90 ms of walks at k = 5 against 6.7 ms at k = 8 for 30 state bits, 41 ms
// at k = 8 against 72 ms at k = 11 for 36. At most 2^28 records.
uint32_t bitwalk_bits(uint32_t n, uint32_t unknowns) {
  if (const char* e = std::getenv("NS_WAYPOINT_BITS")) {
    const long long v = std::atoll(e);
    if (v >= 0 && v <= (long long)n) return (uint32_t)v;
  }
  const uint32_t k = n >= 27 ? std::max(5u, n - 28) : (n > 22 ? n - 22 : 0u);
  return std::min(n - unknowns, k);
}

// The waypoint walks without a table (bit-sliced walks, 32 per lane) for this circuit: the
// native walk template with the circuit's two-step body (register-allocated gfx950 code,
// ns_nat_walk<T>_r<class>) up to 48 state bits, the compiled ns_walk kernel otherwise
// (NS_BACKEND=hiprtc: always). Both take the same arguments (the native one also k) and
// write the same records.
WalkLaunch bit_walk(Lane& ctx, const Circuit& c, uint32_t k, double* compile_ms, uint32_t* cone, uint32_t* valu,
                    uint32_t* backend, uint32_t* vgpr_class) {
  const char* be = std::getenv("NS_BACKEND");
  const bool native_ok = !(be && std::strcmp(be, "hiprtc") == 0);
  const std::string fp = circuit_fingerprint(c) + "/walk/" + std::to_string(k) + (native_ok ? "/n" : "/r");
  Chosen ch;
  if (!ctx.jit.find_built(fp, &ch)) {
    auto t0 = clk::now();
    double cms = 0;
    if (native_ok) {
      try {
        NativeKernel nk = build_native_walk(c, k);
        ch.mod = ctx.jit.get_native(nk.kernel_name, nk.body, nk.asm_text, &cms);
        ch.info = nk.info;
        ch.native = true;
        ch.vgpr_class = nk.vgpr_class;
      } catch (const NativeUnfit&) {
      }
    }
    if (!ch.native) {
      const std::string src = gen_walk_kernel(c, k, &ch.info);
      ch.mod = ctx.jit.get(src, "ns_walk", &cms);
    }
    *compile_ms += ms_since(t0);
    ctx.jit.put_built(fp, ch);
  }
  *cone = ch.info.cone_gates;
  *valu = ch.info.valu_ops;
  *backend = ch.native ? 1 : 0;
  *vgpr_class = ch.vgpr_class;
  auto mod = ch.mod;
  return [&ctx, mod, k](void* out, unsigned int* flag, unsigned long long* evals, uint64_t count, uint64_t limit) {
    // enough lanes for two waves per SIMD, and at least 32 walks per lane where there are
    uint64_t lanes = std::min<uint64_t>(1ull << 17, std::max<uint64_t>(256, count / 32));
    lanes = (lanes + 255) / 256 * 256;
    const uint64_t per_lane = (count + lanes - 1) / lanes;
    // a lane's 32 slots advance together; its walks average 2^k evaluations each on a
    // random map: allow 16 times that, plus one walk of the full limit
    unsigned long long budget = 16ull * ((per_lane + 31) / 32) * (1ull << k) + limit;
    unsigned long long cnt = count, lim = limit;
    unsigned int kk = k;
    // the waypoints are handed out from one counter, so that every slot keeps taking walks
    // while any are left
    unsigned long long* queue = reinterpret_cast<unsigned long long*>(ctx.stats.as<char>() + 256);
    NS_HIP(launch_fill_u64(ctx.stream, queue, 0ull, 1));
    // the launch copies as many arguments as the kernel declares: the compiled ns_walk takes
    // the first seven, the native template all eight
    void* args[8] = {&cnt, &lim, &budget, &out, &flag, &evals, &queue, &kk};
    NS_HIP(hipModuleLaunchKernel(mod->fn, (unsigned)(lanes / 256), 1, 1, 256, 1, 1, 0, ctx.stream, args, nullptr));
  };
}

}  // namespace

void satostop_buffer_plan(uint32_t n, uint32_t unknowns, uint64_t bytes[5]) {
  if (n == 0 || n > 31 || unknowns > n)
    throw Error(ST_UNSUPPORTED, "the buffer plan covers the default search of 1..31 state bits");
  const uint64_t states = 1ull << n;
  // the waypoint walks over the table (two record buffers of 2^(n-k) x 16 bytes) where
  // satostop_solve takes them, and the passes (two u64 jump tables) they fall back on
  const uint32_t k = waypoint_bits(n, unknowns);
  const uint64_t records = (k <= n - unknowns && n - k <= 24) ? (1ull << (n - k)) * 16 : 0;
  // the walks without a table (the default full search) keep their records in the same two
  // buffers, and fall back on the table when a walk runs too long
  const uint32_t kb = bitwalk_bits(n, unknowns);
  const uint64_t bit_records = (kb <= n - unknowns && n - kb <= 28) ? (1ull << (n - kb)) * 16 : 0;
  bytes[0] = (states < 32 ? 32 : states) * 4;
  bytes[1] = std::max<uint64_t>(std::max<uint64_t>((states < 2 ? 2 : states) * 8, records), bit_records);
  bytes[2] = bytes[1];
  bytes[3] = (1ull << unknowns) * 16 + 64;
  bytes[4] = 0;  // outputs 32..n: only past 31 state bits
}

void satostop_solve(Lane& ctx, const Circuit& c, uint32_t unknowns, const ns_satostop_opts& o,
                    ns_satostop_result& r, uint64_t* steps_out, uint64_t* ends_out) {
  auto t0 = clk::now();
  std::memset(&r, 0, sizeof(r));
  const uint32_t n = c.n;
  if (c.outputs.size() != (size_t)n + 1) throw Error(ST_INVALID, "SATOSTOP needs n inputs and n+1 outputs");
  if (n == 0 || n > 62) throw Error(ST_UNSUPPORTED, "SATOSTOP supports 1..62 state bits");
  if (unknowns > n) throw Error(ST_INVALID, "more unknowns than state bits");
  NS_HIP(hipSetDevice(ctx.device));
  ctx.pass_ms.clear();
  // C = log2 of the evaluations every walk covers; the wide layout counts up to 2^30 - 1
  uint32_t cover = o.max_passes ? o.max_passes : n;
  if (cover > n) cover = n;
  const bool first_pass_mode = (o.flags & NS_STOP_FIRST_PASS) != 0;
  const bool full = !first_pass_mode && cover == n && (o.flags & NS_STOP_PASSES) == 0;
  // The full search without a table: walks of 32 states per lane that evaluate the circuit
  // as they go (the native walk template, ~5 ms of code generation and load), the default
  // wherever its waypoints fit: the 30-bit Feistel search takes 3.1 ms of device time this way
  // against 27 ms walking the table, and the packed image of 34 and 35 state bits would take
  // 75 and 155 GiB. NS_STOP_TABLE_WALKS keeps the table.
  const uint32_t kb = bitwalk_bits(n, unknowns);
  const bool bit_ok = full && kb <= n - unknowns && n - kb <= 28;
  const bool use_bit = bit_ok && (o.flags & NS_STOP_TABLE_WALKS) == 0;
  if (n > 35 && !use_bit)
    throw Error(ST_UNSUPPORTED, "over 35 state bits only the full search without a table, with at most 28 unknowns");
  const uint64_t starts_bw = 1ull << unknowns;
  if (use_bit) {
    ctx.tab_d.ensure(starts_bw * 16 + 64);
    unsigned long long* d_steps = ctx.tab_d.as<unsigned long long>();
    unsigned long long* d_ends = d_steps + starts_bw;
    unsigned long long* stats = ctx.stats.as<unsigned long long>();
    unsigned long long host[2] = {0, ~0ull};
    const WalkLaunch wl = bit_walk(ctx, c, kb, &r.compile_ms, &r.cone_gates, &r.valu_ops, &r.backend, &r.vgpr_class);
    if (waypoint_search(ctx, n, unknowns, kb, wl, r, d_steps, d_ends, stats, host, r.backend == 1)) {
      r.method = 2;
      r.n_stopped = host[0];
      r.any_stop = r.n_stopped != 0;
      if (r.n_stopped) {
        r.lowest_u = host[1];
        r.start = r.lowest_u << (n - unknowns);
      }
      unsigned long long se[2] = {0, 0};
      if (r.n_stopped) {
        NS_HIP(hipMemcpyAsync(&se[0], d_steps + r.lowest_u, 8, hipMemcpyDeviceToHost, ctx.stream));
        NS_HIP(hipMemcpyAsync(&se[1], d_ends + r.lowest_u, 8, hipMemcpyDeviceToHost, ctx.stream));
      }
      if (steps_out) NS_HIP(hipMemcpyAsync(steps_out, d_steps, starts_bw * 8, hipMemcpyDeviceToHost, ctx.stream));
      if (ends_out) NS_HIP(hipMemcpyAsync(ends_out, d_ends, starts_bw * 8, hipMemcpyDeviceToHost, ctx.stream));
      NS_HIP(hipStreamSynchronize(ctx.stream));
      if (r.n_stopped) {
        r.steps = se[0];
        r.end = se[1];
      }
      r.wall_ms = ms_since(t0);
      return;
    }
    if (n > 35)
      throw Error(ST_UNSUPPORTED, "a waypoint walk exceeded its limit (a long path or cycle without waypoints); "
                                  "over 35 state bits there is no table to fall back on: use the hash-map search");
    const double cms = r.compile_ms;
    std::memset(&r, 0, sizeof(r));
    r.compile_ms = cms;
    ctx.pass_ms.clear();
  }
  // The full search runs as waypoint walks when every start is a waypoint and the records
  // stay small; from 32 state bits on (NS_PACKED_TABLE=1: from 6) the walks read the packed
  // (n+1)-bit image, which holds 34 and 35 state bits in HBM too.
  const uint32_t k = waypoint_bits(n, unknowns);
  const bool walks = !first_pass_mode && cover == n && (o.flags & NS_STOP_PASSES) == 0 && k <= n - unknowns &&
                     n - k <= 24;
  const char* pe = std::getenv("NS_PACKED_TABLE");
  const bool packed = walks && n >= 6 && (n >= 32 || (pe && *pe == '1'));
  if (n > 33 && !packed)
    throw Error(ST_UNSUPPORTED, "34 and 35 state bits: only the full search (waypoint walks over the packed image, "
                                "at most n - 10 unknowns)");
  const uint64_t states = 1ull << n;
  const uint64_t starts = 1ull << unknowns;
  ctx.tab_d.ensure(starts * 16 + 64);
  unsigned long long* d_steps = ctx.tab_d.as<unsigned long long>();
  unsigned long long* d_ends = d_steps + starts;
  unsigned long long* stats = ctx.stats.as<unsigned long long>();
  unsigned long long host[2] = {0, ~0ull};
  bool any = false, searched = false, walks_failed = false;
  TableKernelInfo tki;
  if (packed) {
    r.build_ms = build_packed_table(ctx, c, o, &any, &r.compile_ms, &r.cone_gates, &tki);
    r.any_stop = any;
    if (!any) {
      NS_HIP(launch_fill_u64(ctx.stream, d_steps, 0ull, 2 * starts));
      host[0] = 0;
      searched = true;
    } else {
      const TableView tv{nullptr, nullptr, 0, ctx.tab_img.as<unsigned long long>(), n + 1};
      searched = waypoint_search(ctx, n, unknowns, k, table_walk(ctx, n, k, tv), r, d_steps, d_ends, stats, host);
      if (searched) r.method = 1;
      if (!searched) {
        if (n > 33)
          throw Error(ST_UNSUPPORTED, "a waypoint walk exceeded its limit (a long path or cycle without waypoints); "
                                      "over 33 state bits there are no passes to fall back on: use the hash-map search");
        walks_failed = true;
        ctx.tab_img.release();
      }
    }
  }
  // More than 31 state bits: the next-state table is split into outputs 0..31 and 32..n
  // and the jump entries use the wide layout. NS_WIDE_TABLE=1 takes that path for any n.
  const char* we = std::getenv("NS_WIDE_TABLE");
  const bool wide = n > 31 || (we && *we == '1');
  uint32_t split = 0;
  if (!searched) {
    double build_ms = 0, cms = 0;
    if (!wide) {
      build_ms = build_table_dev(ctx, c, o.group_len, o.words_per_lane, ctx.tab_a, &any, &cms, &r.cone_gates, &tki);
    } else {
      split = n + 1 > 32 ? 32 : n;  // outputs 0..split-1 in the low table
      bool any_lo = false;
      double cms_lo = 0, cms_hi = 0;
      uint32_t cone_lo = 0, cone_hi = 0;
      TableKernelInfo lo_info;
      build_ms = build_table_dev(ctx, with_outputs(c, 0, split), o.group_len, o.words_per_lane, ctx.tab_a, &any_lo,
                                 &cms_lo, &cone_lo, &lo_info);
      build_ms += build_table_dev(ctx, with_outputs(c, split, n + 1), o.group_len, o.words_per_lane, ctx.tab_e, &any,
                                  &cms_hi, &cone_hi, &tki);
      cms = cms_lo + cms_hi;
      r.cone_gates = cone_lo + cone_hi;
      tki.info.valu_ops += lo_info.info.valu_ops;
      if (!lo_info.native) tki.native = false;
      tki.vgpr_class = std::max(tki.vgpr_class, lo_info.vgpr_class);
    }
    r.build_ms += build_ms;
    r.compile_ms += cms;
    r.any_stop = any;
  }
  r.valu_ops = tki.info.valu_ops;
  r.backend = tki.native ? 1 : 0;
  r.vgpr_class = tki.vgpr_class;
  uint32_t max_passes = first_pass_mode ? cover : choose_passes(n, unknowns, cover);
  if (wide && max_passes > 29) max_passes = 29;
  if (first_pass_mode && wide && cover > 29) throw Error(ST_UNSUPPORTED, "first-pass mode covers at most 2^29 evaluations for more than 31 state bits");
  const uint64_t walk_jumps = 1ull << (cover - max_passes);
  const unsigned long long* final_tab = nullptr;
  if (any && !searched && walks && !walks_failed) {
    const TableView tv{ctx.tab_a.as<unsigned int>(), wide ? ctx.tab_e.as<unsigned int>() : nullptr, split, nullptr, 0};
    searched = waypoint_search(ctx, n, unknowns, k, table_walk(ctx, n, k, tv), r, d_steps, d_ends, stats, host);
    if (searched) r.method = 1;
  }
  if (any && !searched) {
    // even entry count for the two-entries-per-thread pass
    const uint64_t cnt = states < 2 ? 2 : states;
    ctx.tab_b.ensure(cnt * 8);
    ctx.tab_c.ensure(cnt * 8);
    unsigned long long* cur = ctx.tab_b.as<unsigned long long>();
    unsigned long long* nxt = ctx.tab_c.as<unsigned long long>();
    std::vector<hipEvent_t> evs(2 * max_passes);
    for (auto& e : evs) NS_HIP(hipEventCreate(&e));
    uint32_t p = 0;
    for (; p < max_passes; p++) {
      NS_HIP(hipEventRecord(evs[2 * p], ctx.stream));
      if (p == 0) {
        if (wide)
          NS_HIP(launch_jump_first_wide(ctx.stream, ctx.tab_a.as<unsigned int>(), ctx.tab_e.as<unsigned int>(), cur,
                                        states, n, split));
        else
          NS_HIP(launch_jump_first(ctx.stream, ctx.tab_a.as<unsigned int>(), cur, states, n));
      } else {
        NS_HIP(launch_jump_pass(ctx.stream, cur, nxt, states < 2 ? 2 : states, n, wide));
        std::swap(cur, nxt);
      }
      NS_HIP(hipEventRecord(evs[2 * p + 1], ctx.stream));
      if (first_pass_mode) {
        NS_HIP(launch_fill_u64(ctx.stream, stats, 0ull, 1));
        NS_HIP(launch_fill_u64(ctx.stream, stats + 1, ~0ull, 1));
        NS_HIP(launch_starts(ctx.stream, cur, n, unknowns, 1, wide, nullptr, nullptr, stats));
        NS_HIP(hipMemcpyAsync(host, stats, sizeof(host), hipMemcpyDeviceToHost, ctx.stream));
        NS_HIP(hipStreamSynchronize(ctx.stream));
        if (host[0] != 0) {
          p++;
          break;
        }
      }
    }
    NS_HIP(hipStreamSynchronize(ctx.stream));
    r.passes = p;
    for (uint32_t i = 0; i < p; i++) {
      float ms = 0;
      NS_HIP(hipEventElapsedTime(&ms, evs[2 * i], evs[2 * i + 1]));
      ctx.pass_ms.push_back(ms);
      r.jump_ms += ms;
    }
    for (auto& e : evs) (void)hipEventDestroy(e);
    final_tab = cur;
    const uint64_t jumps = p == max_passes ? walk_jumps : 1;  // first-pass mode may end early
    r.walk_jumps = jumps;
    NS_HIP(launch_fill_u64(ctx.stream, stats, 0ull, 1));
    NS_HIP(launch_fill_u64(ctx.stream, stats + 1, ~0ull, 1));
    hipEvent_t w0, w1;
    NS_HIP(hipEventCreate(&w0));
    NS_HIP(hipEventCreate(&w1));
    NS_HIP(hipEventRecord(w0, ctx.stream));
    NS_HIP(launch_starts(ctx.stream, final_tab, n, unknowns, jumps, wide, d_steps, d_ends, stats));
    NS_HIP(hipEventRecord(w1, ctx.stream));
    NS_HIP(hipMemcpyAsync(host, stats, sizeof(host), hipMemcpyDeviceToHost, ctx.stream));
    NS_HIP(hipStreamSynchronize(ctx.stream));
    float wms = 0.f;
    NS_HIP(hipEventElapsedTime(&wms, w0, w1));
    r.walk_ms = wms;
    (void)hipEventDestroy(w0);
    (void)hipEventDestroy(w1);
  } else if (!searched) {
    NS_HIP(launch_fill_u64(ctx.stream, d_steps, 0ull, 2 * starts));
    host[0] = 0;
  }
  NS_HIP(hipStreamSynchronize(ctx.stream));
  r.n_stopped = host[0];
  if (r.n_stopped) {
    r.lowest_u = host[1];
    r.start = r.lowest_u << (n - unknowns);
  }
  unsigned long long se[2] = {0, 0};
  if (r.n_stopped) {
    NS_HIP(hipMemcpyAsync(&se[0], d_steps + r.lowest_u, 8, hipMemcpyDeviceToHost, ctx.stream));
    NS_HIP(hipMemcpyAsync(&se[1], d_ends + r.lowest_u, 8, hipMemcpyDeviceToHost, ctx.stream));
  }
  if (steps_out) NS_HIP(hipMemcpyAsync(steps_out, d_steps, starts * 8, hipMemcpyDeviceToHost, ctx.stream));
  if (ends_out) NS_HIP(hipMemcpyAsync(ends_out, d_ends, starts * 8, hipMemcpyDeviceToHost, ctx.stream));
  NS_HIP(hipStreamSynchronize(ctx.stream));
  if (r.n_stopped) {
    r.steps = se[0];
    r.end = se[1];
  }
  r.wall_ms = ms_since(t0);
}

// ---------------------------------------------------------------------------
// SATOSTOP through naivesatostop's hash map
// ---------------------------------------------------------------------------
void hashmap_step(Lane& ctx, uint32_t which, uint32_t state_len, uint32_t arg_bit_place, uint64_t arg,
                  uint32_t unknown_bits, uint32_t unknown_fill_bits, uint32_t max_predecessors, bool test,
                  const uint32_t* outputs, uint64_t output_words, void* hm, void* hm_out, uint64_t hm_len,
                  uint32_t* fills, uint64_t fills_len, void* sol) {
  if (which > 3 || state_len == 0 || state_len > 63 || hm_len == 0 || (hm_len & (hm_len - 1)))
    throw Error(ST_INVALID, "bad hash-map step arguments");
  if (unknown_fill_bits > unknown_bits || unknown_bits > state_len || arg_bit_place >= state_len)
    throw Error(ST_INVALID, "bad unknown / block widths");
  uint32_t hbits = 0;
  while ((1ull << hbits) < hm_len) hbits++;
  if (hbits > state_len) throw Error(ST_INVALID, "hash map larger than the state space");
  const bool uses_block = which != 1;
  const bool to_out = which == 1 || which == 3;
  const uint64_t wpe = (state_len + 1 + 31) >> 5;
  if (uses_block && output_words < (wpe << arg_bit_place)) throw Error(ST_INVALID, "outputs array too small");
  if (fills_len < (1ull << (unknown_bits - unknown_fill_bits))) throw Error(ST_INVALID, "fills array too small");
  if (uses_block && (arg >> (state_len - arg_bit_place)) != 0) throw Error(ST_INVALID, "block index out of range");
  if (to_out && !hm_out) throw Error(ST_INVALID, "the join steps write a second hash map");
  // the kernels index the map and the fill counters by the entries' states
  for (uint64_t i = 0; i < hm_len; i++) {
    const HmEntry& he = static_cast<const HmEntry*>(hm)[i];
    if (he.state > 3) throw Error(ST_INVALID, "hash entry with an unknown state");
    if (he.state != 0 && ((he.current >> state_len) != 0 || (he.next >> state_len) != 0))
      throw Error(ST_INVALID, "hash entry state wider than state_len");
  }
  NS_HIP(hipSetDevice(ctx.device));
  DevBuf d_out, d_hm, d_hm2, d_fills, d_sol, d_side;
  const uint64_t hm_bytes = hm_len * sizeof(HmEntry);
  d_hm.ensure(hm_bytes);
  // claims (u64 per slot) and jumps (u32 per slot), zero
  d_side.ensure(hm_len * 12);
  unsigned long long* claims = d_side.as<unsigned long long>();
  unsigned int* jumps = reinterpret_cast<unsigned int*>(claims + hm_len);
  NS_HIP(launch_fill_u32(ctx.stream, d_side.as<unsigned int>(), 0u, hm_len * 3));
  d_fills.ensure(fills_len * 4);
  d_sol.ensure(sizeof(HmSolution));
  NS_HIP(hipMemcpyAsync(d_hm.p, hm, hm_bytes, hipMemcpyHostToDevice, ctx.stream));
  NS_HIP(hipMemcpyAsync(d_fills.p, fills, fills_len * 4, hipMemcpyHostToDevice, ctx.stream));
  NS_HIP(hipMemcpyAsync(d_sol.p, sol, sizeof(HmSolution), hipMemcpyHostToDevice, ctx.stream));
  HmParams hp{state_len, arg_bit_place, hbits, unknown_bits, unknown_fill_bits};
  HmOutputs ho{nullptr, nullptr, wpe};
  if (uses_block) {
    d_out.ensure(output_words * 4);
    NS_HIP(hipMemcpyAsync(d_out.p, outputs, output_words * 4, hipMemcpyHostToDevice, ctx.stream));
    ho.lo = d_out.as<unsigned int>();
    ho.hi = wpe == 2 ? ho.lo + 1 : nullptr;
  }
  unsigned int* dfills = d_fills.as<unsigned int>();
  HmSolution* dsol = d_sol.as<HmSolution>();
  if (which == 0) {
    NS_HIP(hm_join_to(ctx.stream, hp, arg, ho, d_hm.as<HmEntry>(), hm_len));
  } else if (which == 2) {
    NS_HIP(hm_add(ctx.stream, hp, arg, ho, d_hm.as<HmEntry>(), hm_len, dfills, dsol, claims, jumps, max_predecessors,
                  false, test));
  } else {
    d_hm2.ensure(hm_bytes);
    NS_HIP(hm_join(ctx.stream, hp, arg, which == 3 ? &ho : nullptr, d_hm.as<HmEntry>(), d_hm2.as<HmEntry>(), hm_len,
                   jumps, dfills, dsol));
    NS_HIP(hm_fold(ctx.stream, hp, d_hm2.as<HmEntry>(), hm_len, dfills, claims, jumps));
  }
  NS_HIP(hipMemcpyAsync(to_out ? hm_out : hm, to_out ? d_hm2.p : d_hm.p, hm_bytes, hipMemcpyDeviceToHost, ctx.stream));
  NS_HIP(hipMemcpyAsync(fills, d_fills.p, fills_len * 4, hipMemcpyDeviceToHost, ctx.stream));
  NS_HIP(hipMemcpyAsync(sol, d_sol.p, sizeof(HmSolution), hipMemcpyDeviceToHost, ctx.stream));
  NS_HIP(hipStreamSynchronize(ctx.stream));
}

void satostop_hashmap_solve(Lane& ctx, const Circuit& c, uint32_t unknowns, const ns_hashmap_opts& o,
                            ns_hashmap_result& r, ns_hashmap_progress_fn progress, void* user) {
  auto t0 = clk::now();
  std::memset(&r, 0, sizeof(r));
  const uint32_t n = c.n;
  if (c.outputs.size() != (size_t)n + 1) throw Error(ST_INVALID, "SATOSTOP needs n inputs and n+1 outputs");
  if (n < 10 || n > 63) throw Error(ST_UNSUPPORTED, "the hash-map search takes 10..63 state bits");
  if (unknowns == 0 || unknowns > n) throw Error(ST_INVALID, "unknowns must be in 1..n");
  const uint32_t H = o.hashmap_len_bits ? o.hashmap_len_bits : 24;
  if (H > n || H > 34) throw Error(ST_INVALID, "hash map bits must not exceed the state bits (and 34)");
  uint32_t e = o.elem_bits ? o.elem_bits : 24;
  if (e >= n) e = n - 1;
  if (e < 5) throw Error(ST_INVALID, "blocks need at least 32 states");
  // calculate_default_unknown_fill_bits, src/satostop.rs:101-107
  const uint32_t ufb = o.unknown_fill_bits >= 0 ? (uint32_t)o.unknown_fill_bits
                                                : (H - 1 < unknowns ? unknowns - (H - 1) : 0);
  if (ufb > unknowns) throw Error(ST_INVALID, "unknown fill bits exceed the unknowns");
  const uint32_t clear_every = o.iters_per_clear_predecessors ? o.iters_per_clear_predecessors : 20;
  NS_HIP(hipSetDevice(ctx.device));
  // the block's outputs: outputs 0..31 (and 32..n) as u32 tables of 2^e entries
  const bool two = n + 1 > 32;
  const uint32_t split = two ? 32 : n + 1;
  const uint64_t block = 1ull << e, block_words = block >> 5;
  Geometry g = pick_geometry(block_words, 0, 2, false);
  std::vector<uint32_t> vb(n);
  for (uint32_t k = 0; k < n; k++) vb[k] = k;
  Chosen kern[2];
  const Circuit parts[2] = {with_outputs(c, 0, split), two ? with_outputs(c, split, n + 1) : Circuit()};
  for (int h = 0; h < (two ? 2 : 1); h++) {
    std::vector<uint32_t> outs;
    for (uint32_t k = 0; k < parts[h].outputs.size(); k++) outs.push_back(k);
    double cms = 0;
    kern[h] = choose_kernel(
        ctx, parts[h], NativeKind::TABLE, vb, 0, outs, g, "",
        [&](GenInfo* info) { return gen_table_kernel(parts[h], g, info); }, "ns_table", &cms);
    r.compile_ms += cms;
  }
  // a launch covers whole thread blocks: small blocks leave slack at the end of the tables
  const uint64_t wpb = g.words_per_block();
  const uint64_t tab_bytes = (block_words + wpb - 1) / wpb * wpb * 32 * 4;
  ctx.tab_a.ensure(tab_bytes);
  if (two) ctx.tab_e.ensure(tab_bytes);
  const uint64_t hm_len = 1ull << H;
  ctx.tab_b.ensure(hm_len * sizeof(HmEntry));
  ctx.tab_c.ensure(hm_len * sizeof(HmEntry));
  const uint64_t nfills = 1ull << (unknowns - ufb);
  // tab_d: the fill counters, then per hash slot a claim word (u64) and a jump count (u32)
  const uint64_t fills_bytes = (nfills * 4 + 255) / 256 * 256;
  ctx.tab_d.ensure(fills_bytes + hm_len * 12);
  HmEntry* hm1 = ctx.tab_b.as<HmEntry>();
  HmEntry* hm2 = ctx.tab_c.as<HmEntry>();
  unsigned int* fills = ctx.tab_d.as<unsigned int>();
  unsigned long long* claims = reinterpret_cast<unsigned long long*>(ctx.tab_d.as<char>() + fills_bytes);
  unsigned int* jumps = reinterpret_cast<unsigned int*>(claims + hm_len);
  NS_HIP(launch_fill_u32(ctx.stream, reinterpret_cast<unsigned int*>(claims), 0u, hm_len * 3));
  HmSolution* sol = reinterpret_cast<HmSolution*>(ctx.stats.as<char>() + 128);
  NS_HIP(launch_fill_u64(ctx.stream, reinterpret_cast<unsigned long long*>(hm1), 0ull, hm_len * 4));
  NS_HIP(launch_fill_u64(ctx.stream, reinterpret_cast<unsigned long long*>(hm2), 0ull, hm_len * 4));
  NS_HIP(launch_fill_u32(ctx.stream, fills, 0u, nfills));
  NS_HIP(launch_fill_u64(ctx.stream, reinterpret_cast<unsigned long long*>(sol), 0ull, 5));
  HmParams hp{n, e, H, unknowns, ufb};
  HmOutputs ho{ctx.tab_a.as<unsigned int>(), two ? ctx.tab_e.as<unsigned int>() : nullptr, 1};
  unsigned int* flag = reinterpret_cast<unsigned int*>(ctx.stats.as<char>() + 64);
  const uint64_t args = 1ull << (n - e);
  uint32_t iter = 0;
  HmSolution host{};
  float kms_total = 0;
  NS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  for (r.sweeps = 1;; r.sweeps++) {
    bool done = false;
    for (uint64_t arg = 0; arg < args && !done; arg++) {
      // outputs of states [arg << e, (arg + 1) << e)
      const uint64_t w0 = arg * block_words;
      for (int h = 0; h < (two ? 2 : 1); h++) {
        unsigned int* T = h ? ctx.tab_e.as<unsigned int>() : ctx.tab_a.as<unsigned int>();
        uint32_t wpl_v = g.wpl, last = (uint32_t)parts[h].outputs.size() - 1, valid = 0xffffffffu, small = 0;
        void* extra[6] = {&T, &flag, &wpl_v, &last, &valid, &small};
        launch_range(ctx, kern[h].mod->fn, g, w0, block_words, extra, kern[h].native ? 6 : 2, w0 + block_words);
      }
      // the reference's process() (src/satostop.rs:1211-1257): join_to and join_itself in
      // one pass, add (which also clears the predecessor counts every clear_every blocks),
      // swap
      NS_HIP(hm_join(ctx.stream, hp, arg, &ho, hm1, hm2, hm_len, jumps, fills, sol));
      const bool clear = ++iter >= clear_every;
      if (clear) iter = 0;
      NS_HIP(hm_add(ctx.stream, hp, arg, ho, hm2, hm_len, fills, sol, claims, jumps, o.max_predecessors, clear, false));
      std::swap(hm1, hm2);
      r.blocks++;
      NS_HIP(hipMemcpyAsync(&host, sol, sizeof(host), hipMemcpyDeviceToHost, ctx.stream));
      NS_HIP(hipStreamSynchronize(ctx.stream));
      // the reference's per-arg callback prints "Step: arg / arg_steps" and the resolved
      // unknowns here (src/satostop.rs:1548-1558)
      if (progress && progress(user, arg, args, host.resolved_unknowns) != 0) {
        NS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
        NS_HIP(hipEventSynchronize(ctx.ev1));
        throw Error(ST_CANCELLED, "hash-map search cancelled by its progress callback");
      }
      if (host.defined) {
        r.found = 1;
        r.start = host.start;
        r.end = host.end;
        r.steps = host.steps;
        done = true;
      } else if (host.resolved_unknowns == nfills) {
        r.found = 0;
        done = true;
      }
    }
    if (done) break;
    if (o.max_sweeps && r.sweeps >= o.max_sweeps) {
      r.found = 2;
      break;
    }
  }
  NS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  NS_HIP(hipEventSynchronize(ctx.ev1));
  kms_total = ctx.elapsed_ms();
  r.resolved_unknowns = host.resolved_unknowns;
  r.kernel_ms = kms_total;
  r.wall_ms = ms_since(t0);
}

// ---------------------------------------------------------------------------
// greedysatostop's partitioned search with the image on disk
// ---------------------------------------------------------------------------
void greedy_partitioned_solve(Lane& ctx, const Circuit& c, uint32_t unknowns, uint64_t partitions,
                              const std::string& file_prefix, ns_partitioned_result& r) {
  auto t0 = clk::now();
  std::memset(&r, 0, sizeof(r));
  const uint32_t n = c.n, width = n + 1;
  if (c.outputs.size() != (size_t)n + 1) throw Error(ST_INVALID, "SATOSTOP needs n inputs and n+1 outputs");
  if (n < 10 || n > 48) throw Error(ST_UNSUPPORTED, "the partitioned search takes 10..48 state bits");
  if (unknowns > n) throw Error(ST_INVALID, "more unknowns than state bits");
  // do_solve_with_cpu_builder_with_partitions: at least two partitions, a power of two
  if (partitions < 2 || (partitions & (partitions - 1))) throw Error(ST_INVALID, "partitions: a power of two >= 2");
  FileImage fi(n, partitions, file_prefix);
  r.partitions = partitions;
  NS_HIP(hipSetDevice(ctx.device));
  // the image, chunk by chunk, through a device chunk buffer and a pinned host buffer
  const uint32_t cb = std::min(n, 30u);
  const uint64_t chunk_words = ((1ull << cb) * width) >> 6;
  const uint64_t part_words = fi.partition_words(), part_states = fi.partition_states();
  const uint64_t host_words = std::max(chunk_words, part_words);
  unsigned long long* host = nullptr;
  NS_HIP(hipHostMalloc(reinterpret_cast<void**>(&host), host_words * 8, hipHostMallocDefault));
  struct HostFree {
    unsigned long long* p;
    ~HostFree() { (void)hipHostFree(p); }
  } host_free{host};
  ctx.tab_img.ensure(chunk_words * 8);
  unsigned long long* chunk_dev = ctx.tab_img.as<unsigned long long>();
  bool any = false;
  TableKernelInfo tki;
  ns_satostop_opts o{};
  uint32_t cone = 0;
  auto tb = clk::now();
  build_packed_chunks(
      ctx, c, o, cb, [&](uint64_t) { return chunk_dev; },
      [&](uint64_t) {
        NS_HIP(hipMemcpyAsync(host, chunk_dev, chunk_words * 8, hipMemcpyDeviceToHost, ctx.stream));
        NS_HIP(hipStreamSynchronize(ctx.stream));
        fi.save_chunk(reinterpret_cast<const uint64_t*>(host), chunk_words);
      },
      &any, &r.compile_ms, &cone, &tki);
  r.build_ms = ms_since(tb);
  ctx.tab_img.release();
  if (any) {
    // partitions in HBM: dest (ping-pong) and second
    DevBuf d_dest, d_next, d_second;
    d_dest.ensure((part_words + 1) * 8);
    d_next.ensure((part_words + 1) * 8);
    d_second.ensure((part_words + 1) * 8);
    unsigned long long* dest = d_dest.as<unsigned long long>();
    unsigned long long* nxt = d_next.as<unsigned long long>();
    unsigned long long* second = d_second.as<unsigned long long>();
    unsigned long long* best = ctx.stats.as<unsigned long long>() + 2;
    hipEvent_t j0, j1;
    NS_HIP(hipEventCreate(&j0));
    NS_HIP(hipEventCreate(&j1));
    struct EvFree {
      hipEvent_t a, b;
      ~EvFree() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
      }
    } ev_free{j0, j1};
    auto load = [&](uint64_t part, unsigned long long* dst) {
      auto ti = clk::now();
      fi.load_partition(part, reinterpret_cast<uint64_t*>(host));
      NS_HIP(hipMemcpyAsync(dst, host, part_words * 8, hipMemcpyHostToDevice, ctx.stream));
      NS_HIP(hipStreamSynchronize(ctx.stream));
      r.io_ms += ms_since(ti);
    };
    bool done = false;
    for (uint32_t stage = 0; stage < n && !done; stage++) {
      r.stages = stage + 1;
      for (uint64_t i = 0; i < partitions && !done; i++) {
        load(i, dest);
        for (uint64_t j = 0; j < partitions; j++) {
          if (j != i) {
            load(j, second);
          } else {
            NS_HIP(hipMemcpyAsync(second, dest, part_words * 8, hipMemcpyDeviceToDevice, ctx.stream));
          }
          NS_HIP(hipEventRecord(j0, ctx.stream));
          NS_HIP(launch_mem_join(ctx.stream, dest, part_states, second, j * part_states, part_states, n, nxt));
          NS_HIP(hipEventRecord(j1, ctx.stream));
          NS_HIP(hipEventSynchronize(j1));
          float ms = 0.f;
          NS_HIP(hipEventElapsedTime(&ms, j0, j1));
          r.join_ms += ms;
          std::swap(dest, nxt);
        }
        // the lowest stopping unknown start of partition i
        NS_HIP(launch_fill_u64(ctx.stream, best, ~0ull, 1));
        NS_HIP(launch_mem_find(ctx.stream, dest, i * part_states, part_states, n, unknowns, best));
        unsigned long long b = ~0ull;
        NS_HIP(hipMemcpyAsync(&b, best, 8, hipMemcpyDeviceToHost, ctx.stream));
        NS_HIP(hipStreamSynchronize(ctx.stream));
        if (b != ~0ull) {
          // its entry: the two words it may span
          const uint64_t bit = (b - i * part_states) * width, idx = bit >> 6;
          unsigned long long w2[2] = {0, 0};
          NS_HIP(hipMemcpyAsync(w2, dest + idx, 16, hipMemcpyDeviceToHost, ctx.stream));
          NS_HIP(hipStreamSynchronize(ctx.stream));
          const uint32_t sh = (uint32_t)(bit & 63);
          unsigned long long v = w2[0] >> sh;
          if (sh + width > 64) v |= w2[1] << (64 - sh);
          r.found = 1;
          r.start = b;
          r.end = v & ((1ull << n) - 1ull);
          done = true;
          break;
        }
        auto ts = clk::now();
        NS_HIP(hipMemcpyAsync(host, dest, part_words * 8, hipMemcpyDeviceToHost, ctx.stream));
        NS_HIP(hipStreamSynchronize(ctx.stream));
        fi.save_partition(i, reinterpret_cast<const uint64_t*>(host));
        r.io_ms += ms_since(ts);
      }
    }
  }
  r.wall_ms = ms_since(t0);
}

const std::vector<double>& last_pass_times(Lane& ctx) { return ctx.pass_ms; }

}  // namespace ns
