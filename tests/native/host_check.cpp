// Host-only run of the code generator for the sanitizer builds (AddressSanitizer +
// UndefinedBehaviorSanitizer, ThreadSanitizer): LUT mapping, scheduling, register allocation
// (with its threaded output-group refinement on the 36-input workload), machine-code
// encoding, template patching and batch code objects, for every kernel kind, on random and
// workload circuits. No device is touched. Exit status 0 and "host_check ok" when every
// step ran; sanitizer reports make the process fail.
#include <atomic>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "../../naivesat_amd/csrc/asmenc.h"
#include "../../naivesat_amd/csrc/asmgen.h"
#include "../../naivesat_amd/csrc/circuit.h"

using namespace ns;

namespace {

std::atomic<int> g_fail{0};

void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "host_check: %s\n", what);
    g_fail = 1;
  }
}

// builds, encodes and places one body; bodies that fit no register class are skipped
void run(const Circuit& c, NativeKind kind, uint32_t lane_shift, bool full) {
  std::vector<uint32_t> vb(c.n), outs;
  const bool qsat = kind == NativeKind::QSAT_A || kind == NativeKind::QSAT_B;
  for (uint32_t k = 0; k < c.n; k++) vb[k] = qsat ? c.n - 1 - k : k;
  if (kind == NativeKind::TABLE)
    for (uint32_t k = 0; k < c.outputs.size(); k++) outs.push_back(k);
  else
    outs.push_back(0);
  NativeKernel nk;
  try {
    nk = build_native(c, kind, vb, lane_shift, outs);
  } catch (const NativeUnfit&) {
    return;
  }
  std::vector<uint32_t> words;
  std::string err;
  check(encode_body(nk.body, words, &err), "encode_body failed");
  if (!full) return;
  if (kind == NativeKind::BATCH) {
    std::vector<std::vector<uint32_t>> bodies{words, words};
    auto t = batch_template(batch_bytes(bodies));
    check(t != nullptr, "no batch template");
    if (t) {
      std::vector<uint64_t> off;
      std::vector<char> elf = compose_batch(*t, bodies, off);
      check(off.size() == 2 && off[1] > off[0] && elf.size() == t->elf.size(), "batch layout");
    }
  } else {
    size_t slot = 0;
    std::vector<char> elf = patch_native(nk.kernel_name, nk.body, &slot);
    check(!elf.empty() && slot > 0, "patch_native");
  }
}

Circuit random_circuit(Rng& rng, uint32_t n, uint32_t gates) {
  Circuit c;
  c.n = n;
  for (uint32_t j = 0; j < gates; j++) {
    const uint32_t w = n + j;
    c.gates.push_back({(uint32_t)(rng.next() % w), (uint32_t)(rng.next() % w), (uint8_t)(rng.next() & 3)});
  }
  c.outputs.push_back({n + gates - 1, (rng.next() & 1) != 0});
  return c;
}

}  // namespace

int main() {
  Rng rng(20261204);
  for (int i = 0; i < 24; i++) {
    const uint32_t n = 6 + (uint32_t)(rng.next() % 26), g = 20 + (uint32_t)(rng.next() % 600);
    Circuit c = random_circuit(rng, n, g);
    const bool full = i < 4;  // the template assembly is slow under the sanitizers
    run(c, NativeKind::SAT, 0, full);
    run(c, NativeKind::BATCH, 0, full);
    if (n >= 13) {
      run(c, NativeKind::QSAT_A, 2, false);
      run(c, NativeKind::QSAT_B, 0, false);
    }
  }
  uint64_t key = 0, target = 0;
  run(gen_sat_circuit(1, 24, 2000, 32, &key), NativeKind::SAT, 0, true);
  run(gen_qsat_circuit(3, 32, 10000, 200), NativeKind::QSAT_A, 2, false);
  run(gen_satostop_circuit(4, 15, 6, 300, 14, &target), NativeKind::TABLE, 0, true);
  // the 36-input, 20,000-gate workload: output groups refined on threads
  run(gen_sat_circuit(2, 36, 20000, 40, &key), NativeKind::SAT, 0, false);
  // batch bodies generated concurrently, as a batch's thread pool does
  std::vector<Circuit> cs;
  for (int i = 0; i < 16; i++) cs.push_back(gen_sat_circuit(100 + i, 16 + i, 300 + 200 * i, 16, &key));
  std::vector<std::thread> th;
  for (int t = 0; t < 4; t++)
    th.emplace_back([&, t] {
      for (size_t i = t; i < cs.size(); i += 4) run(cs[i], NativeKind::BATCH, 0, i < 4);
    });
  for (auto& x : th) x.join();
  if (!g_fail) std::printf("host_check ok\n");
  return g_fail;
}
