// Batched solving: many SAT / QSAT circuits in one persistent kernel launch. Each request's
// circuit becomes a register-allocated body; the bodies of a batch are placed as subroutines
// in one code object (one module load for the batch instead of one per request), and the
// persistent kernels take chunks of every request's assignment space from a work counter,
// reading the request's work descriptor (body entry, space, output) for each chunk.
#pragma once
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/naivesat_hip.h"
#include "circuit.h"
#include "runtime.h"

namespace ns {

struct ServeRequest;

// Launch shape of the batch kernels for one register class, from the autotuner.
struct BatchTuning {
  uint32_t blocks_per_cu = 0;  // persistent blocks of 256 threads per CU (0 = the class's occupancy)
  uint32_t chunks_per_wave = 8;  // work chunks per resident wave (balance vs counter traffic)
};

struct BatchStats {
  uint32_t requests = 0, fallbacks = 0, vgpr_class = 0;
  uint64_t code_bytes = 0;
  double gen_ms = 0, load_ms = 0, kernel_ms = 0, wall_ms = 0;
};

class Batcher {
 public:
  Batcher(Context& ctx, const ns_serve_opts& o);
  // requests the batch kernels take: SAT, and QSAT whose levels below the outermost block
  // reach the thread's bit (at least 5 + jb of them), at most 36 inputs
  static bool batchable(const ServeRequest& r);
  uint32_t limit() const { return limit_; }
  // Solves the requests together on `lane`; res[i] gets request i's answer (v0..v3,
  // status, compile_ms = its code generation plus a share of the batch's load, kernel_ms =
  // the batch's device time). A request whose body does not fit the batch kernels is solved
  // alone on the lane.
  void run(Lane& lane, const std::vector<const ServeRequest*>& reqs, std::vector<ns_request_result>& res,
           BatchStats* stats = nullptr);
  BatchTuning tuning(uint32_t vgpr_class) const;
  void set_tuning(uint32_t vgpr_class, const BatchTuning& t);

 private:
  Context& ctx_;
  uint32_t limit_ = 32;
  uint32_t cus_ = 0;
  mutable std::mutex mu_;
  BatchTuning tune_[4];  // r64, r128, r256, r512
};

// One-shot batch on the context's own lane (C ABI ns_batch_solve): the same path as served
// batches, for callers that hold many circuits at once.
void batch_solve(Context& ctx, const std::vector<const ServeRequest*>& reqs, std::vector<ns_request_result>& res,
                 BatchStats* stats);

}  // namespace ns
