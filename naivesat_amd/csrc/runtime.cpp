#include "runtime.h"

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

std::shared_ptr<JitModule> Jit::get(const std::string& src, const char* kernel, double* compile_ms) {
  std::lock_guard<std::mutex> lk(mu_);
  std::string key = std::string(kernel) + '\n' + src;
  auto it = cache_.find(key);
  if (it != cache_.end()) {
    if (compile_ms) *compile_ms = 0.0;
    return it->second;
  }
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
  cache_[key] = m;
  return m;
}

Jit::~Jit() {
  for (auto& kv : cache_)
    if (kv.second->mod) (void)hipModuleUnload(kv.second->mod);
}

void DevBuf::ensure(size_t bytes) {
  if (bytes <= cap && p) return;
  release();
  if (bytes == 0) bytes = 256;
  NS_HIP(hipMalloc(&p, bytes));
  cap = bytes;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
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
  NS_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  NS_HIP(hipEventCreate(&ev0));
  NS_HIP(hipEventCreate(&ev1));
  stats.ensure(4096);
}

Context::~Context() {
  (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  buf_a.release();
  buf_b.release();
  buf_c.release();
  buf_d.release();
  stats.release();
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

float Context::elapsed_ms() {
  float ms = 0.f;
  NS_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  return ms;
}

}  // namespace ns
