#include "runtime.h"

#include "asmenc.h"
#include "batch.h"
#include "static_kernels.h"

#include <cstdio>

#include <hip/hiprtc.h>

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace ns {

void check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  int code = (e == hipErrorOutOfMemory) ? ST_NOMEM : ST_HIP;
  throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}

static void check_rtc(hiprtcResult r, const char* what) {
  if (r != HIPRTC_SUCCESS) throw Error(ST_COMPILE, std::string(what) + ": " + hiprtcGetErrorString(r));
}

// NS_TRACE_JIT=1 prints the assembly and module-load time of every native kernel
static bool trace_jit() {
  static const bool on = [] {
    const char* e = std::getenv("NS_TRACE_JIT");
    return e && *e && *e != '0';
  }();
  return on;
}

JitModule::~JitModule() {
  if (mod) (void)hipModuleUnload(mod);
}

// Cache key of a kernel: 128 bits over the text the code depends on (kernel name and source,
// assembly or body), as 32 hex digits, so that a cache entry does not keep its text (a body
// runs to hundreds of KB). Two independent 64-bit lanes, 8 bytes a step.
static std::string text_digest(std::initializer_list<const std::string*> parts) {
  uint64_t h1 = 0x243f6a8885a308d3ull, h2 = 0x13198a2e03707344ull, len = 0;
  auto rotl = [](uint64_t x, int r) { return (x << r) | (x >> (64 - r)); };
  auto step = [&](uint64_t w) {
    h1 = rotl(h1 ^ (w * 0x87c37b91114253d5ull), 31) * 0x4cf5ad432745937full;
    h2 = rotl(h2 + (w * 0x9e3779b97f4a7c15ull), 29) * 0xc2b2ae3d27d4eb4full + 0x52dce729u;
  };
  for (const std::string* p : parts) {
    const char* s = p->data();
    size_t n = p->size(), i = 0;
    for (; i + 8 <= n; i += 8) {
      uint64_t w;
      std::memcpy(&w, s + i, 8);
      step(w);
    }
    uint64_t tail = 0;
    std::memcpy(&tail, s + i, n - i);
    step(tail ^ ((uint64_t)(n - i) << 56));
    len += n;
    step(len);  // part boundary
  }
  auto fin = [](uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    return x ^ (x >> 33);
  };
  char buf[33];
  std::snprintf(buf, sizeof buf, "%016llx%016llx", (unsigned long long)fin(h1 ^ len),
                (unsigned long long)fin(h2 + len));
  return buf;
}

std::shared_ptr<JitModule> Jit::lookup(const std::string& key, double* compile_ms) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = cache_.find(key);
  if (it == cache_.end()) return nullptr;
  it->second->last_use = ++clock_;
  if (compile_ms) *compile_ms = 0.0;
  return it->second;
}

std::shared_ptr<JitModule> Jit::get(const std::string& src, const char* kernel, double* compile_ms) {
  const std::string kname(kernel), tag("rtc");
  std::string key = text_digest({&tag, &kname, &src});
  if (auto hit = lookup(key, compile_ms)) return hit;
  auto t0 = std::chrono::steady_clock::now();
  hiprtcProgram prog;
  check_rtc(hiprtcCreateProgram(&prog, src.c_str(), "ns_kernel.hip", 0, nullptr, nullptr),
            "hiprtcCreateProgram");
  // NS_HIPRTC_OPT overrides the optimisation level (e.g. "-O1") for debugging
  const char* opt_env = std::getenv("NS_HIPRTC_OPT");
  std::vector<const char*> opts = {"--offload-arch=gfx950", opt_env && *opt_env ? opt_env : "-O3",
                                   "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const char* extra_env = std::getenv("NS_HIPRTC_EXTRA");
  std::vector<std::string> extra;
  if (extra_env && *extra_env) {
    std::string s(extra_env);
    size_t p = 0;
    while (p < s.size()) {
      size_t q = s.find(' ', p);
      if (q == std::string::npos) q = s.size();
      if (q > p) extra.push_back(s.substr(p, q - p));
      p = q + 1;
    }
    for (auto& e : extra) opts.push_back(e.c_str());
  }
  hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  if (r != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw Error(ST_COMPILE, std::string("kernel compilation failed: ") + log.substr(0, 4000));
  }
  size_t cs = 0;
  check_rtc(hiprtcGetCodeSize(prog, &cs), "hiprtcGetCodeSize");
  std::vector<char> code(cs);
  check_rtc(hiprtcGetCode(prog, code.data()), "hiprtcGetCode");
  hiprtcDestroyProgram(&prog);
  auto m = std::make_shared<JitModule>();
  NS_HIP(hipModuleLoadData(&m->mod, code.data()));
  NS_HIP(hipModuleGetFunction(&m->fn, m->mod, kernel));
  auto t1 = std::chrono::steady_clock::now();
  m->compile_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (compile_ms) *compile_ms = m->compile_ms;
  return insert(key, m);
}

void Jit::drop(std::shared_ptr<JitModule> m, std::vector<std::shared_ptr<JitModule>>& release) {
  if (!m) return;
  // only the last reference is parked (the module cache and the built kernels both hold one;
  // dropping another holder's does not unload)
  if (parking_ > 0 && m.use_count() == 1 && parked_.size() < kParkedMax) parked_.push_back(std::move(m));
  else release.push_back(std::move(m));
}

void Jit::park(bool on) {
  std::lock_guard<std::mutex> lk(mu_);
  parking_ += on ? 1 : -1;
}

size_t Jit::reap(size_t n) {
  std::vector<std::shared_ptr<JitModule>> release;
  {
    std::lock_guard<std::mutex> lk(mu_);
    const size_t k = n == 0 || n > parked_.size() ? parked_.size() : n;
    release.assign(std::make_move_iterator(parked_.end() - (long)k), std::make_move_iterator(parked_.end()));
    parked_.resize(parked_.size() - k);
  }
  return release.size();  // released here, unlocked (a module another holder still uses stays loaded)
}

std::shared_ptr<JitModule> Jit::insert(const std::string& key, std::shared_ptr<JitModule> m) {
  // dropped modules are released after the lock (declared before it, the vector is destroyed
  // after it): an unload takes ~0.1-0.6 ms, and the lanes of a server share this cache
  std::vector<std::shared_ptr<JitModule>> release;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = cache_.find(key);
  if (it != cache_.end()) return it->second;  // compiled concurrently by another thread: keep the first
  if (cache_.size() >= kCapacity) {
    auto victim = cache_.begin();
    for (auto jt = cache_.begin(); jt != cache_.end(); ++jt)
      if (jt->second->last_use < victim->second->last_use) victim = jt;
    drop(std::move(victim->second), release);
    cache_.erase(victim);
  }
  m->last_use = ++clock_;
  cache_[key] = m;
  return m;
}

std::shared_ptr<JitModule> Jit::get_asm(const std::string& asm_text, const std::string& kernel, double* compile_ms) {
  const std::string tag("asm");
  std::string key = text_digest({&tag, &kernel, &asm_text});
  if (auto hit = lookup(key, compile_ms)) return hit;
  auto t0 = std::chrono::steady_clock::now();
  std::vector<char> code = assemble_native(asm_text);
  auto t1 = std::chrono::steady_clock::now();
  auto m = std::make_shared<JitModule>();
  NS_HIP(hipModuleLoadData(&m->mod, code.data()));
  NS_HIP(hipModuleGetFunction(&m->fn, m->mod, kernel.c_str()));
  auto t2 = std::chrono::steady_clock::now();
  m->compile_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
  if (trace_jit())
    std::fprintf(stderr, "[ns jit] %s assemble_ms=%.3f load_ms=%.3f bytes=%zu\n", kernel.c_str(),
                 std::chrono::duration<double, std::milli>(t1 - t0).count(),
                 std::chrono::duration<double, std::milli>(t2 - t1).count(), code.size());
  if (compile_ms) *compile_ms = m->compile_ms;
  return insert(key, m);
}

bool Jit::find_built(const std::string& fp, BuiltKernel* out) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = built_.find(fp);
  if (it == built_.end()) return false;
  it->second.second = ++clock_;
  *out = it->second.first;
  return true;
}

void Jit::put_built(const std::string& fp, const BuiltKernel& k) {
  std::vector<std::shared_ptr<JitModule>> release;  // see insert
  std::lock_guard<std::mutex> lk(mu_);
  auto it = built_.find(fp);
  if (it != built_.end()) {
    drop(std::move(it->second.first.mod), release);  // the replaced entry's reference
    it->second = {k, ++clock_};
    return;
  }
  if (built_.size() >= kCapacity) {
    auto victim = built_.begin();
    for (auto jt = built_.begin(); jt != built_.end(); ++jt)
      if (jt->second.second < victim->second.second) victim = jt;
    drop(std::move(victim->second.first.mod), release);
    built_.erase(victim);
  }
  built_[fp] = {k, ++clock_};
}

std::shared_ptr<JitModule> Jit::get_native(const std::string& kernel, const std::string& body,
                                           const std::string& asm_text, double* compile_ms) {
  const std::string tag("nat");
  std::string key = text_digest({&tag, &kernel, &body});
  if (auto hit = lookup(key, compile_ms)) return hit;
  auto t0 = std::chrono::steady_clock::now();
  std::vector<char> code = patch_native(kernel, body);
  const bool patched = !code.empty();
  if (!patched) code = assemble_native(asm_text);
  auto t1 = std::chrono::steady_clock::now();
  auto m = std::make_shared<JitModule>();
  NS_HIP(hipModuleLoadData(&m->mod, code.data()));
  NS_HIP(hipModuleGetFunction(&m->fn, m->mod, kernel.c_str()));
  auto t2 = std::chrono::steady_clock::now();
  m->compile_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
  if (compile_ms) *compile_ms = m->compile_ms;
  if (trace_jit())
    std::fprintf(stderr, "[ns jit] %s %s_ms=%.3f load_ms=%.3f bytes=%zu\n", kernel.c_str(),
                 patched ? "patch" : "assemble", std::chrono::duration<double, std::milli>(t1 - t0).count(),
                 std::chrono::duration<double, std::milli>(t2 - t1).count(), code.size());
  return insert(key, m);
}

Jit::~Jit() = default;

void DevBuf::ensure(size_t bytes) {
  if (bytes <= cap && p) return;
  if (!owned) throw Error(ST_INTERNAL, "a lent device buffer is smaller than the solve needs");
  release();
  if (bytes == 0) bytes = 256;
  NS_HIP(hipMalloc(&p, bytes));
  cap = bytes;
}

void DevBuf::release() {
  if (p && owned) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  owned = true;
}

Context::Context(int dev) : device(dev) {
  int count = 0;
  NS_HIP(hipGetDeviceCount(&count));
  if (dev < 0 || dev >= count) throw Error(ST_INVALID, "no HIP device " + std::to_string(dev));
  NS_HIP(hipSetDevice(dev));
  hipDeviceProp_t prop;
  NS_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    throw Error(ST_UNSUPPORTED, std::string("device arch ") + prop.gcnArchName +
                                    " is not gfx950 (kernels are built for gfx950 only)");
  // the allocation-free SAT entry point's counters: their address on this device
  NS_HIP(noalloc_scratch_setup(dev));
  main = std::make_unique<Lane>(dev, jit);
}

Context::~Context() {
  batcher_.reset();  // its kept batches (modules, graphs, buffers) before the lanes and the cache
}

Batcher& Context::batcher() {
  std::lock_guard<std::mutex> lk(batcher_mu_);
  if (!batcher_) {
    ns_serve_opts o{};
    o.batch_max = 256;
    batcher_ = std::make_unique<Batcher>(*this, o, true);
  }
  return *batcher_;
}

Lane::Lane(int dev, Jit& j) : device(dev), jit(j) {
  NS_HIP(hipSetDevice(dev));
  NS_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  NS_HIP(hipEventCreate(&ev0));
  NS_HIP(hipEventCreate(&ev1));
  stats.ensure(16384);  // counters and flags; search-kernel count slots from byte 4096
  NS_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned), kPinnedBytes, hipHostMallocDefault));
}

void Lane::put_graph(const std::string& key, hipGraphExec_t exec, std::shared_ptr<JitModule> mod) {
  if (graphs.size() >= 32) drop_graphs();
  graphs[key] = CapturedGraph{exec, std::move(mod)};
}

void Lane::drop_graphs() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto& kv : graphs)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  graphs.clear();
}

Lane::~Lane() {
  (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  drop_graphs();
  buf_a.release();
  buf_b.release();
  buf_c.release();
  bstats.release();
  bdesc.release();
  bchunk.release();
  trim_tables();
  stats.release();
  if (pinned) (void)hipHostFree(pinned);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

float Lane::elapsed_ms() {
  float ms = 0.f;
  NS_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  return ms;
}

}  // namespace ns
