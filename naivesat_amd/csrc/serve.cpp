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
#include <map>
#include <mutex>
#include <thread>

#include "solvers.h"

namespace ns {

namespace {

uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// SATOSTOP buffers, in the order the solver uses them: u32 next-state table, two u64
// jump tables, per-start results (u64 steps and ends), and for more than 31 state bits
// the table of outputs 32..n
constexpr int kStopBufs = 5;
void stop_sizes(const ServeRequest& r, uint64_t sz[kStopBufs]) {
  const uint64_t states = 1ull << r.circuit->n;
  sz[0] = align256((states < 32 ? 32 : states) * 4);
  sz[1] = align256((states < 2 ? 2 : states) * 8);
  sz[2] = sz[1];
  sz[3] = align256((1ull << r.unknowns) * 16 + 64);
  sz[4] = r.circuit->n > 31 ? sz[0] : 0;
}

}  // namespace

uint64_t request_hbm_bytes(const ServeRequest& r) {
  if (r.kind != NS_REQ_SATOSTOP) return 0;  // SAT / QSAT use the lane's small buffers
  uint64_t sz[kStopBufs];
  stop_sizes(r, sz);
  uint64_t total = 0;
  for (uint64_t b : sz) total += b;
  return total;
}

namespace {

using clk = std::chrono::steady_clock;

// First-fit sub-allocator over one device allocation: SATOSTOP tables are carved out of
// it, so serving makes no hipMalloc / hipFree calls per request.
class Arena {
 public:
  explicit Arena(uint64_t bytes) : size_(bytes) {
    if (bytes) {
      NS_HIP(hipMalloc(&base_, bytes));
      holes_[0] = bytes;
    }
  }
  ~Arena() {
    if (base_) (void)hipFree(base_);
  }
  char* base() const { return static_cast<char*>(base_); }
  bool alloc(uint64_t bytes, uint64_t* off) {
    for (auto it = holes_.begin(); it != holes_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      uint64_t rest = it->second - bytes, at = it->first + bytes;
      holes_.erase(it);
      if (rest) holes_[at] = rest;
      return true;
    }
    return false;
  }
  void free(uint64_t off, uint64_t bytes) {
    auto it = holes_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != holes_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      holes_.erase(nx);
    }
    if (it != holes_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        holes_.erase(it);
      }
    }
  }

 private:
  void* base_ = nullptr;
  uint64_t size_ = 0;
  std::map<uint64_t, uint64_t> holes_;  // offset -> size
};

struct Shared {
  std::mutex mu;
  std::condition_variable cv;
  size_t next = 0;  // next request to hand out (arrival order)
};

}  // namespace

void serve(Context& ctx, const std::vector<ServeRequest>& reqs, const ns_serve_opts& o,
           std::vector<ns_request_result>& out, double* total_s) {
  const uint32_t workers = o.workers ? o.workers : 8;
  uint64_t biggest = 0;
  for (const auto& r : reqs) {
    if (!r.circuit) throw Error(ST_INVALID, "request without a circuit");
    biggest = std::max(biggest, request_hbm_bytes(r));
  }
  // default budget: room for every worker's largest table, at most 16 GiB
  const uint64_t budget =
      o.hbm_budget_bytes ? o.hbm_budget_bytes : std::max(biggest, std::min<uint64_t>(biggest * workers, 16ull << 30));
  out.assign(reqs.size(), ns_request_result{});
  // requests are served in arrival order
  std::vector<size_t> order(reqs.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
  if (biggest > budget) throw Error(ST_UNSUPPORTED, "a request needs more memory than the budget");
  NS_HIP(hipSetDevice(ctx.device));
  Arena arena(biggest ? budget : 0);
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
      uint64_t need, off = 0;
      {
        std::unique_lock<std::mutex> lk(sh.mu);
        if (sh.next >= order.size()) return;
        idx = order[sh.next];
        need = request_hbm_bytes(reqs[idx]);
        // admission: the head request waits until its tables fit in the arena; requests
        // behind it wait too, so they keep arrival order
        sh.cv.wait(lk, [&] {
          return sh.next >= order.size() || order[sh.next] != idx || need == 0 || arena.alloc(need, &off);
        });
        if (sh.next >= order.size()) return;
        if (order[sh.next] != idx) continue;  // taken by another worker meanwhile
        sh.next++;
      }
      const ServeRequest& r = reqs[idx];
      if (need) {
        uint64_t sz[kStopBufs], at = off;
        stop_sizes(r, sz);
        DevBuf* bufs[kStopBufs] = {&lane.tab_a, &lane.tab_b, &lane.tab_c, &lane.tab_d, &lane.tab_e};
        for (int k = 0; k < kStopBufs; k++) {
          if (!sz[k]) continue;
          bufs[k]->borrow(arena.base() + at, sz[k]);
          at += sz[k];
        }
      }
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
          res.kernel_ms = tr.build_ms + tr.jump_ms + tr.walk_ms;
        }
        res.status = ST_OK;
      } catch (const Error& e) {
        res.status = e.code;
      } catch (const std::exception&) {
        res.status = ST_INTERNAL;
      }
      res.finish_s = now_s();
      if (need) {
        lane.trim_tables();  // forgets the lent slices
        std::lock_guard<std::mutex> lk(sh.mu);
        arena.free(off, need);
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
