// Serving: SAT / QSAT / SATOSTOP requests taken as they arrive and solved by worker lanes
// that run concurrently on one device.
#pragma once
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/naivesat_hip.h"
#include "circuit.h"
#include "runtime.h"

namespace ns {

struct ServeRequest {
  uint32_t kind = 0;  // NS_REQ_SAT / NS_REQ_QSAT / NS_REQ_SATOSTOP
  uint32_t unknowns = 0;
  std::shared_ptr<const Circuit> circuit;
  std::vector<uint8_t> quants;
  double arrival_s = 0.0;  // trace replay only: arrival after the start of the run
};

// Device bytes a request needs besides the lane's small buffers (admission control).
uint64_t request_hbm_bytes(const ServeRequest& r);

class Arena;
class Batcher;

// A running server: `workers` lanes (a host thread, a HIP stream and buffers each) take
// submitted requests in submission order. A request is admitted when a lane is free and,
// for SATOSTOP, when its tables fit in the device arena; a request that waits for memory
// holds back only the memory-needing requests behind it, so the order among those is kept
// while SAT and QSAT requests go on. Each request's answer does not depend on what else is
// in flight. With batching on, the SAT requests waiting when a lane frees up are solved
// together by one persistent kernel.
class Server {
 public:
  Server(Context& ctx, const ns_serve_opts& o, uint64_t budget_bytes);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Queues a request; its arrival (submit_s) is now. Returns the ticket. A request the
  // server can never admit completes at once with its error status.
  uint64_t submit(ServeRequest r);
  // true and the result once the ticket's request has finished; a finished ticket's result
  // is handed out once, after which the ticket is unknown. Throws for an unknown ticket.
  bool poll(uint64_t ticket, ns_request_result* out);
  // waits up to timeout_s (< 0: without limit) for the ticket's request
  bool wait(uint64_t ticket, double timeout_s, ns_request_result* out);
  // waits for every submitted request, then stops the workers
  void close();
  double now_s() const;
  uint32_t workers() const { return (uint32_t)lanes_.size(); }

 private:
  struct Pending {
    uint64_t ticket = 0;
    ServeRequest req;
    uint64_t need = 0;
    double submit_s = 0;
  };
  void work(uint32_t wid);
  // the next admissible requests (under mu_): one, or several SAT requests for a batch
  bool take(std::vector<Pending>& got, uint64_t* arena_off);
  void run_one(Lane& lane, uint32_t wid, Pending& p, ns_request_result& res);
  void finish(const Pending& p, const ns_request_result& res);

  Context& ctx_;
  ns_serve_opts opts_;
  std::unique_ptr<Arena> arena_;
  std::vector<std::unique_ptr<Lane>> lanes_;
  std::unique_ptr<Batcher> batcher_;
  std::vector<std::thread> threads_;
  mutable std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Pending> queue_;
  std::unordered_map<uint64_t, ns_request_result> done_;
  std::unordered_map<uint64_t, bool> open_tickets_;  // submitted, result not yet handed out
  uint64_t next_ticket_ = 1;
  uint64_t running_ = 0;
  bool closing_ = false;
  bool parking_ = false;  // the context's kernel cache parks dropped modules while this is open
  std::chrono::steady_clock::time_point t0_;
};

// Replays a trace: requests are submitted at their arrival times (realtime) or all at once,
// and every result is waited for.
void serve(Context& ctx, const std::vector<ServeRequest>& reqs, const ns_serve_opts& o,
           std::vector<ns_request_result>& out, double* total_s);

}  // namespace ns
