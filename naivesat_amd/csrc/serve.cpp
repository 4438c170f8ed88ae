// Serving loop: worker threads, each owning a lane (HIP stream + buffers), pull requests
// in arrival order. Compilation (native path, milliseconds) runs on the worker's host
// thread while other lanes keep the device busy; device memory for SATOSTOP tables is
// admitted against a budget.
#include "serve.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>

#include "solvers.h"

namespace ns {

uint64_t request_hbm_bytes(const ServeRequest& r) {
  const uint32_t n = r.circuit->n;
  if (r.kind == NS_REQ_SATOSTOP) return (1ull << n) * 20ull + (1ull << r.unknowns) * 8ull;  // u32 + 2 x u64 tables
  if (r.kind == NS_REQ_QSAT) return n > 10 ? (1ull << (n - 8)) : 64;  // reduction bit arrays
  return 64;
}

namespace {

using clk = std::chrono::steady_clock;

struct Shared {
  std::mutex mu;
  std::condition_variable cv;
  size_t next = 0;        // next request to hand out (arrival order)
  uint64_t hbm_used = 0;  // admitted request bytes
};

}  // namespace

void serve(Context& ctx, const std::vector<ServeRequest>& reqs, const ns_serve_opts& o,
           std::vector<ns_request_result>& out, double* total_s) {
  const uint32_t workers = o.workers ? o.workers : 8;
  const uint64_t budget = o.hbm_budget_bytes ? o.hbm_budget_bytes : (64ull << 30);
  out.assign(reqs.size(), ns_request_result{});
  // requests are served in arrival order
  std::vector<size_t> order(reqs.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
  for (const auto& r : reqs) {
    if (!r.circuit) throw Error(ST_INVALID, "request without a circuit");
    if (request_hbm_bytes(r) > budget) throw Error(ST_UNSUPPORTED, "a request needs more memory than the budget");
  }
  NS_HIP(hipSetDevice(ctx.device));
  std::vector<std::unique_ptr<Lane>> lanes;
  for (uint32_t w = 0; w < workers; w++) lanes.push_back(std::make_unique<Lane>(ctx.device, ctx.jit));
  Shared sh;
  const auto t0 = clk::now();
  auto now_s = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
  auto work = [&](uint32_t wid) {
    Lane& lane = *lanes[wid];
    (void)hipSetDevice(ctx.device);
    while (true) {
      size_t idx;
      uint64_t need;
      {
        std::unique_lock<std::mutex> lk(sh.mu);
        if (sh.next >= order.size()) return;
        idx = order[sh.next];
        need = request_hbm_bytes(reqs[idx]);
        // admission: wait until the table fits the budget (requests keep arrival order)
        sh.cv.wait(lk, [&] {
          return sh.next >= order.size() || order[sh.next] != idx || sh.hbm_used + need <= budget;
        });
        if (sh.next >= order.size()) return;
        if (order[sh.next] != idx) continue;  // taken meanwhile
        sh.next++;
        sh.hbm_used += need;
      }
      const ServeRequest& r = reqs[idx];
      if (o.realtime) {
        double wait = r.arrival_s - now_s();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      }
      ns_request_result& res = out[idx];
      res.worker = wid;
      res.start_s = now_s();
      try {
        if (r.kind == NS_REQ_SAT) {
          ns_sat_opts so{};
          ns_sat_result sr;
          sat_solve(lane, *r.circuit, so, sr);
          res.v0 = sr.count;
          res.v1 = sr.lowest;
          res.v2 = sr.found;
          res.compile_ms = sr.compile_ms;
          res.kernel_ms = sr.kernel_ms;
        } else if (r.kind == NS_REQ_QSAT) {
          ns_qsat_opts qo{};
          ns_qsat_result qr;
          qsat_solve(lane, *r.circuit, r.quants, qo, qr, nullptr, 0);
          res.v0 = qr.truth;
          res.v1 = qr.witness;
          res.v2 = qr.witness_count;
          res.v3 = qr.has_witness;
          res.compile_ms = qr.compile_ms;
          res.kernel_ms = qr.kernel_ms;
        } else {
          ns_satostop_opts to{};
          ns_satostop_result tr;
          satostop_solve(lane, *r.circuit, r.unknowns, to, tr, nullptr, nullptr);
          res.v0 = tr.n_stopped;
          res.v1 = tr.lowest_u;
          res.v2 = tr.steps;
          res.v3 = tr.end;
          res.compile_ms = tr.compile_ms;
          res.kernel_ms = tr.build_ms + tr.jump_ms;
          lane.trim();  // table memory goes back to the budget
        }
        res.status = ST_OK;
      } catch (const Error& e) {
        res.status = e.code;
      } catch (const std::exception&) {
        res.status = ST_INTERNAL;
      }
      res.finish_s = now_s();
      {
        std::lock_guard<std::mutex> lk(sh.mu);
        sh.hbm_used -= need;
      }
      sh.cv.notify_all();
    }
  };
  std::vector<std::thread> th;
  for (uint32_t w = 0; w < workers; w++) th.emplace_back(work, w);
  for (auto& t : th) t.join();
  if (total_s) *total_s = now_s();
}

}  // namespace ns
