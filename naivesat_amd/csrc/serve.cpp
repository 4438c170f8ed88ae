// Serving: worker threads, each owning a lane (HIP stream + buffers), take submitted
// requests in submission order. Code generation (native path, about a millisecond) runs on
// the worker's host thread while other lanes keep the device busy; device memory for
// SATOSTOP tables is admitted against the arena's budget. With batching on, a worker takes
// every waiting SAT / QSAT request it can (up to the batch limit) and solves them with one
// persistent kernel whose code object holds all their circuit bodies (batch.cpp).
#include "serve.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>

#include "batch.h"
#include "solvers.h"

namespace ns {

namespace {

uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// SATOSTOP buffers lent to the lane, in the order of Lane::tab_a .. tab_e, as the solver's
// own plan for its default search sizes them (next-state table, waypoint records or jump
// tables, per-start results). Served SATOSTOP requests take at most 31 state bits.
constexpr int kStopBufs = 5;
void stop_sizes(const ServeRequest& r, uint64_t sz[kStopBufs]) {
  if (r.circuit->n > 31 || r.unknowns > r.circuit->n)
    throw Error(ST_UNSUPPORTED, "served SATOSTOP requests take 1..31 state bits");
  satostop_buffer_plan(r.circuit->n, r.unknowns, sz);
  for (int k = 0; k < kStopBufs; k++) sz[k] = align256(sz[k]);
}

using clk = std::chrono::steady_clock;

}  // namespace

uint64_t request_hbm_bytes(const ServeRequest& r) {
  if (r.kind != NS_REQ_SATOSTOP) return 0;  // SAT / QSAT use the lane's small buffers
  uint64_t sz[kStopBufs];
  stop_sizes(r, sz);
  uint64_t total = 0;
  for (uint64_t b : sz) total += b;
  return total;
}

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
  uint64_t size() const { return size_; }
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

Server::Server(Context& ctx, const ns_serve_opts& o, uint64_t budget_bytes) : ctx_(ctx), opts_(o) {
  const uint32_t workers = o.workers ? o.workers : 8;
  if (workers > 64) throw Error(ST_INVALID, "at most 64 workers");
  NS_HIP(hipSetDevice(ctx.device));
  arena_ = std::make_unique<Arena>(budget_bytes);
  for (uint32_t w = 0; w < workers; w++) lanes_.push_back(std::make_unique<Lane>(ctx.device, ctx.jit));
  if (o.batch_max > 1) batcher_ = std::make_unique<Batcher>(ctx, o, false);
  ctx.jit.park(true);  // modules the cache drops wait for idle lanes (Jit::reap)
  parking_ = true;
  t0_ = clk::now();
  for (uint32_t w = 0; w < workers; w++) threads_.emplace_back([this, w] { work(w); });
}

Server::~Server() {
  try {
    close();
  } catch (...) {
  }
}

double Server::now_s() const { return std::chrono::duration<double>(clk::now() - t0_).count(); }

uint64_t Server::submit(ServeRequest r) {
  if (!r.circuit) throw Error(ST_INVALID, "request without a circuit");
  Pending p;
  p.req = std::move(r);
  int status = ST_OK;
  try {
    p.need = request_hbm_bytes(p.req);
    if (p.need > arena_->size())
      throw Error(ST_UNSUPPORTED, "a request needs more device memory than the serving budget");
  } catch (const Error& e) {
    status = e.code;
  }
  std::lock_guard<std::mutex> lk(mu_);
  if (closing_) throw Error(ST_INVALID, "the server is closed");
  p.ticket = next_ticket_++;
  p.submit_s = now_s();
  open_tickets_[p.ticket] = true;
  if (status != ST_OK) {  // never admissible: done at once
    ns_request_result res{};
    res.status = status;
    res.submit_s = res.start_s = res.finish_s = p.submit_s;
    done_[p.ticket] = res;
    cv_done_.notify_all();
    return p.ticket;
  }
  const uint64_t t = p.ticket;
  queue_.push_back(std::move(p));
  cv_work_.notify_one();
  return t;
}

bool Server::poll(uint64_t ticket, ns_request_result* out) {
  std::lock_guard<std::mutex> lk(mu_);
  if (!open_tickets_.count(ticket)) throw Error(ST_INVALID, "unknown ticket");
  auto it = done_.find(ticket);
  if (it == done_.end()) return false;
  if (out) *out = it->second;
  done_.erase(it);
  open_tickets_.erase(ticket);
  return true;
}

bool Server::wait(uint64_t ticket, double timeout_s, ns_request_result* out) {
  std::unique_lock<std::mutex> lk(mu_);
  if (!open_tickets_.count(ticket)) throw Error(ST_INVALID, "unknown ticket");
  auto ready = [&] { return done_.count(ticket) != 0; };
  if (timeout_s < 0) {
    cv_done_.wait(lk, ready);
  } else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
    return false;
  }
  auto it = done_.find(ticket);
  if (out) *out = it->second;
  done_.erase(it);
  open_tickets_.erase(ticket);
  return true;
}

void Server::close() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (closing_ && threads_.empty()) return;
    closing_ = true;
  }
  cv_work_.notify_all();
  for (auto& t : threads_) t.join();
  threads_.clear();
  batcher_.reset();
  lanes_.clear();
  arena_.reset();
  if (parking_) {
    parking_ = false;
    ctx_.jit.park(false);
    ctx_.jit.reap(0);
  }
}

// Under mu_. Admission in submission order: a request that needs arena memory and cannot
// get it holds back the memory-needing requests behind it; requests that need none pass.
bool Server::take(std::vector<Pending>& got, uint64_t* arena_off) {
  got.clear();
  bool memory_blocked = false;
  for (auto it = queue_.begin(); it != queue_.end(); ++it) {
    if (it->need) {
      if (memory_blocked) continue;
      uint64_t off = 0;
      if (!arena_->alloc(it->need, &off)) {
        memory_blocked = true;
        continue;
      }
      *arena_off = off;
      got.push_back(std::move(*it));
      queue_.erase(it);
      return true;
    }
    if (batcher_ && Batcher::batchable(it->req)) {
      // this and every later waiting request the batch kernel takes, up to the limit
      const uint32_t lim = batcher_->limit();
      for (auto jt = it; jt != queue_.end() && got.size() < lim;) {
        if (!jt->need && Batcher::batchable(jt->req)) {
          got.push_back(std::move(*jt));
          jt = queue_.erase(jt);
        } else {
          ++jt;
        }
      }
      return true;
    }
    got.push_back(std::move(*it));
    queue_.erase(it);
    return true;
  }
  return false;
}

void Server::finish(const Pending& p, const ns_request_result& res) {
  std::lock_guard<std::mutex> lk(mu_);
  done_[p.ticket] = res;
  cv_done_.notify_all();
}

void Server::run_one(Lane& lane, uint32_t wid, Pending& p, ns_request_result& res) {
  const ServeRequest& r = p.req;
  res.worker = wid;
  res.submit_s = p.submit_s;
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
}

void Server::work(uint32_t wid) {
  Lane& lane = *lanes_[wid];
  (void)hipSetDevice(ctx_.device);
  std::vector<Pending> got;
  while (true) {
    uint64_t off = 0;
    // the whole server idle (no request waiting, no lane at work): unload parked modules, one
    // at a time, so that a request arriving meanwhile waits for one unload at most (an unload
    // while other lanes load and launch slows them)
    while (true) {
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (!queue_.empty() || running_ != 0 || closing_) break;
      }
      if (!ctx_.jit.reap(1)) break;
    }
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait(lk, [&] { return take(got, &off) || (closing_ && queue_.empty()); });
      if (got.empty()) return;  // closing and nothing left
      running_++;
    }
    if (got.size() == 1 && (got[0].need || !batcher_ || !Batcher::batchable(got[0].req))) {
      Pending& p = got[0];
      if (p.need) {
        uint64_t sz[kStopBufs], at = off;
        stop_sizes(p.req, sz);
        DevBuf* bufs[kStopBufs] = {&lane.tab_a, &lane.tab_b, &lane.tab_c, &lane.tab_d, &lane.tab_e};
        for (int k = 0; k < kStopBufs; k++) {
          if (!sz[k]) continue;
          bufs[k]->borrow(arena_->base() + at, sz[k]);
          at += sz[k];
        }
      }
      ns_request_result res{};
      run_one(lane, wid, p, res);
      if (p.need) {
        lane.trim_tables();  // forgets the lent slices
        std::lock_guard<std::mutex> lk(mu_);
        arena_->free(off, p.need);
      }
      finish(p, res);
    } else {
      std::vector<const ServeRequest*> reqs;
      std::vector<ns_request_result> res(got.size());
      for (size_t i = 0; i < got.size(); i++) {
        reqs.push_back(&got[i].req);
        res[i].worker = wid;
        res[i].submit_s = got[i].submit_s;
      }
      const double start = now_s();
      for (auto& x : res) x.start_s = start;
      try {
        batcher_->run(lane, reqs, res);
      } catch (...) {
        // a batch that fails as a whole (a device or code-object error): every request
        // alone, each with its own status
        for (size_t i = 0; i < got.size(); i++) {
          res[i] = ns_request_result{};
          run_one(lane, wid, got[i], res[i]);
        }
      }
      const double fin = now_s();
      for (size_t i = 0; i < got.size(); i++) {
        res[i].finish_s = fin;
        finish(got[i], res[i]);
      }
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      running_--;
    }
    cv_work_.notify_all();  // freed memory or a lane: others may be admissible now
  }
}

void serve(Context& ctx, const std::vector<ServeRequest>& reqs, const ns_serve_opts& o,
           std::vector<ns_request_result>& out, double* total_s) {
  const uint32_t workers = o.workers ? o.workers : 8;
  uint64_t biggest = 0;
  for (const auto& r : reqs) {
    if (!r.circuit) throw Error(ST_INVALID, "request without a circuit");
    try {
      biggest = std::max(biggest, request_hbm_bytes(r));
    } catch (const Error&) {
      // such a request fails on its own when submitted
    }
  }
  // default budget: room for every worker's largest table, at most 16 GiB
  const uint64_t budget =
      o.hbm_budget_bytes ? o.hbm_budget_bytes : std::max(biggest, std::min<uint64_t>(biggest * workers, 16ull << 30));
  std::vector<size_t> order(reqs.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
  out.assign(reqs.size(), ns_request_result{});
  Server srv(ctx, o, biggest ? budget : 0);
  std::vector<uint64_t> tickets(reqs.size());
  for (size_t i : order) {
    if (o.realtime) {
      const double wait = reqs[i].arrival_s - srv.now_s();
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
    }
    tickets[i] = srv.submit(reqs[i]);
  }
  for (size_t i = 0; i < reqs.size(); i++) srv.wait(tickets[i], -1.0, &out[i]);
  if (total_s) *total_s = srv.now_s();
  srv.close();
}

}  // namespace ns
