// Batched solving: many SAT / QSAT circuits in one persistent kernel launch. Each request's
// circuit becomes a register-allocated body; the bodies of a batch are placed as subroutines
// in one code object (one module load for the batch instead of one per request), and the
// persistent kernels' waves take chunks of every request's assignment space round-robin,
// reading the request's work descriptor (body entry, space, output) for each chunk.
#pragma once
#include <cstdint>
#include <list>
#include <memory>
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
  uint32_t blocks_per_cu = 0;    // persistent blocks of 256 threads per CU
  uint32_t chunks_per_wave = 8;  // work chunks per resident wave (balance vs per-chunk overhead)
};

struct BatchStats {
  uint32_t requests = 0, fallbacks = 0;
  uint32_t vgpr_class = 0;  // the largest register class of its bodies
  uint32_t launches = 0;    // batch kernel launches: one per kind and register class present
  uint32_t graph_runs = 0;  // a kept batch: its HIP-graph launches so far (0: eager launches)
  uint64_t code_bytes = 0;
  double gen_ms = 0;  // code generation of the batch's bodies (wall time)
  double load_ms = 0, kernel_ms = 0, wall_ms = 0;
};

struct BatchPlan;

class Batcher {
 public:
  // keep_plans: remember the last few batches (per lane) so that the same batch again reuses
  // its code object and runs as a captured HIP graph
  Batcher(Context& ctx, const ns_serve_opts& o, bool keep_plans = false);
  ~Batcher();
  // requests the batch kernels take: SAT, and QSAT whose levels below the outermost block
  // reach the thread's bit (at least 5 + jb of them), at most 36 inputs
  static bool batchable(const ServeRequest& r);
  uint32_t limit() const { return limit_; }
  // Solves the requests together on `lane`; res[i] gets request i's answer (v0..v3,
  // status, compile_ms = its code generation plus a share of the batch's load, kernel_ms =
  // the batch's device time). A request whose body does not fit the batch kernels is solved
  // alone on the lane. Answers do not depend on the launch shape the autotuner picks.
  void run(Lane& lane, const std::vector<const ServeRequest*>& reqs, std::vector<ns_request_result>& res,
           BatchStats* stats = nullptr);
  // The autotuner: for each register class, every candidate launch shape is tried on two
  // batches of real work (kernel time per unit of body cost), then the best is kept.
  BatchTuning tuning(uint32_t vgpr_class, uint32_t* candidate);
  void report(uint32_t vgpr_class, uint32_t candidate, double kernel_ms, double total_cost);
  BatchTuning chosen_tuning(uint32_t vgpr_class, bool* settled) const;

 private:
  void replay(Lane& lane, BatchPlan& p, std::vector<ns_request_result>& res, BatchStats* bs);
  std::shared_ptr<BatchPlan> find_plan(const Lane* lane, const std::string& key);
  void keep_plan(std::shared_ptr<BatchPlan> p);
  static constexpr uint32_t kDefaultCandidate = 2;  // the class's occupancy, 8 chunks per wave
  struct TuneState {
    std::vector<BatchTuning> cand;
    std::vector<double> sum;  // kernel ms per unit of cost, summed
    std::vector<int> n;
    int best = -1;
  };
  Context& ctx_;
  bool keep_plans_ = false;
  bool fixed_tuning_ = false;
  uint32_t limit_ = 32;
  uint32_t cus_ = 0;
  uint32_t gen_threads_ = 1;
  mutable std::mutex mu_;
  TuneState tune_[4];  // r64, r128, r256, r512
  std::mutex plan_mu_;
  std::list<std::shared_ptr<BatchPlan>> plans_;  // most recent first
};

// A batch on the context's own lane (C ABI ns_batch_solve), through the context's batcher,
// which keeps recent batches for repeats: the path of served batches, for callers that hold
// many circuits at once.
void batch_solve(Context& ctx, const std::vector<const ServeRequest*>& reqs, std::vector<ns_request_result>& res,
                 BatchStats* stats);

}  // namespace ns
