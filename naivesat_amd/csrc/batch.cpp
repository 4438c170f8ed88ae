// Batched SAT / QSAT solving with the persistent batch kernels (templates/gen_templates.py:
// ns_nat_bsat_rC, ns_nat_bqsat_rC). One batch = one code object holding every request's
// body and the batch kernels of every register class, one persistent launch per kind and
// register class present (each at its class's occupancy; plus the fold of the SAT chunk
// results and the QSAT reduction passes of each QSAT request), one read-back of all results. A batcher that keeps plans (the context's own, behind ns_batch_solve)
// remembers its recent batches: the same batch again skips code generation and the module
// load, is captured as a HIP graph the second time and replayed from then on.
#include "batch.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <thread>

#include "asmenc.h"
#include "asmgen.h"
#include "serve.h"
#include "solvers.h"
#include "static_kernels.h"

namespace ns {

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t0) { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

// the templates' NsBatchDesc
struct Desc {
  uint32_t wend, chunk0;
  uint64_t out;
  uint32_t entry, words, jb, opm, valid;
  uint32_t pad[7];
};
static_assert(sizeof(Desc) == 64 && sizeof(Desc) == sizeof(BatchFoldDesc), "batch descriptor layout");

// per-request results: [0] count, [1] lowest (SAT: written by the fold of its chunks;
// QSAT: the reduction passes' counters, with their 64 count slots from [512], 16 apart)
constexpr uint64_t kStatsWords = 512 + 64 * 16;
constexpr uint64_t kStatsBytes = kStatsWords * 8;
constexpr uint32_t kMaxChunkWords = 1u << 24;
constexpr size_t kKeptPlans = 4;

// waves per SIMD the batch kernels reach in each class (SGPRs limit r64 to 7)
const uint32_t kOccupancy[4] = {7, 4, 2, 1};
const uint32_t kClassOf[4] = {64, 128, 256, 512};

uint32_t c_valid_mask(uint32_t n) { return n >= 5 ? 0xffffffffu : (uint32_t)((1ull << (1u << n)) - 1); }

// Waits for the lane's stream when a batch leaves early (an error after its launches), so
// that the batch's module is never unloaded while a kernel of it may still run.
struct StreamDrain {
  hipStream_t s;
  ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

uint32_t qsat_first_bits(const std::vector<uint8_t>& q) {
  uint32_t fb = 0;
  while (fb < q.size() && q[fb] == q[0]) fb++;
  return fb;
}

void solve_alone(Lane& lane, const ServeRequest& r, ns_request_result& res) {
  if (r.kind == NS_REQ_SAT) {
    ns_sat_opts so{};
    ns_sat_result sr;
    sat_solve(lane, *r.circuit, so, sr);
    res.v0 = sr.count;
    res.v1 = sr.lowest;
    res.v2 = sr.found;
    res.v3 = 0;
    res.compile_ms = sr.compile_ms;
    res.kernel_ms = sr.kernel_ms;
  } else {
    ns_qsat_opts qo{};
    ns_qsat_result qr;
    qsat_solve(lane, *r.circuit, r.quants, qo, qr, nullptr, 0);
    res.v0 = qr.truth;
    res.v1 = qr.witness;
    res.v2 = qr.witness_count;
    res.v3 = qr.has_witness;
    res.compile_ms = qr.compile_ms;
    res.kernel_ms = qr.kernel_ms;
  }
}

}  // namespace

// One request of a batch: its body and its place in the batch's layout.
struct BatchItem {
  size_t idx = 0;  // position in the caller's request list
  bool qsat = false, fallback = false, q0all = false;
  uint32_t n = 0, jb = 0, vgpr = 0, cost = 0, m0 = 0, L = 0, opm = 0;
  uint64_t words = 0, chunk_w = 0, chunks = 0;
  uint64_t bits_off = 0;        // QSAT: byte offset of its bit array
  std::vector<uint8_t> quants;  // QSAT: its prefix (reduction-pass operations)
  std::vector<uint32_t> code;
  double gen_ms = 0;
};

// The requests of one kind and register class in a batch: one persistent launch of that
// class's kernel, at that class's occupancy.
struct BatchGroup {
  bool qsat = false;
  int ci = 0;                   // register class index (64, 128, 256, 512)
  size_t first = 0, count = 0;  // its descriptors (and items)
  uint64_t chunks = 0;
  uint64_t chunk_base = 0;      // SAT: its first chunk result (in chunk_out, 16 bytes each)
  uint32_t grid = 0, tune_idx = 0;
  double cost = 0;
  hipFunction_t fn = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};  // its launch, timed on eager runs
};

// Everything a batch launch needs, and for a kept plan the device memory it owns.
struct BatchPlan {
  std::string key;
  const Lane* lane = nullptr;
  JitModule mod;
  std::vector<BatchItem> items;  // the batched requests, in descriptor order (by group)
  std::vector<Desc> descs;       // out bound per buffer set
  std::vector<BatchGroup> groups;
  uint64_t csat = 0, qbytes = 0, pp_bytes = 0, K = 0;
  uint32_t cls = 64;             // the largest register class of its bodies
  double load_ms = 0, gen_ms = 0, gen_wall_ms = 0;
  uint64_t code_bytes = 0;
  // kept plans: their own buffers and, from the second run, their graph
  DevBuf d_desc, d_stats, d_chunk, d_bits, d_pp_a, d_pp_b;
  hipGraphExec_t graph = nullptr;
  uint32_t runs = 0;
  ~BatchPlan() {
    if (graph) (void)hipGraphExecDestroy(graph);
    for (BatchGroup& g : groups)
      for (hipEvent_t e : g.ev)
        if (e) (void)hipEventDestroy(e);
  }
};

namespace {

// device buffers a batch launch writes
struct BatchBufs {
  Desc* desc;
  unsigned long long* stats;
  unsigned long long* chunk;
  char* bits;
  unsigned int* pp[2];
};

// descriptors with their out pointers bound to the buffers
std::vector<Desc> bind(const BatchPlan& p, const BatchBufs& b) {
  std::vector<Desc> d = p.descs;
  for (size_t i = 0; i < d.size(); i++) {
    const BatchItem& it = p.items[i];
    d[i].out = it.qsat ? (uint64_t)(b.bits + it.bits_off) : (uint64_t)(b.stats + it.idx * kStatsWords);
  }
  return d;
}

// the batch's device work on `s` (also what a capture records): counter fills, each group's
// batch kernel, the fold of the SAT chunks, the QSAT reduction passes; `timed` records each
// group's launch between its events (eager runs)
void enqueue(const BatchPlan& p, const BatchBufs& b, hipStream_t s, bool timed) {
  NS_HIP(launch_fill_u64(s, b.stats, 0ull, p.K * kStatsWords));
  NS_HIP(launch_fill_strided_u64(s, b.stats + 1, ~0ull, p.K, kStatsWords));
  for (const BatchGroup& g : p.groups) {
    const Desc* dp = b.desc + g.first;
    uint32_t nd = (uint32_t)g.count, ch = (uint32_t)g.chunks;
    if (timed) NS_HIP(hipEventRecord(g.ev[0], s));
    if (!g.qsat) {
      unsigned long long* chunk_out = b.chunk + 2 * g.chunk_base;
      void* args[4] = {&dp, &nd, &ch, &chunk_out};
      NS_HIP(hipModuleLaunchKernel(g.fn, g.grid, 1, 1, 256, 1, 1, 0, s, args, nullptr));
    } else {
      void* args[3] = {&dp, &nd, &ch};
      NS_HIP(hipModuleLaunchKernel(g.fn, g.grid, 1, 1, 256, 1, 1, 0, s, args, nullptr));
    }
    if (timed) NS_HIP(hipEventRecord(g.ev[1], s));
    if (!g.qsat)
      NS_HIP(launch_batch_fold(s, reinterpret_cast<const BatchFoldDesc*>(dp), nd, b.chunk + 2 * g.chunk_base));
  }
  // the reduction passes of each QSAT request over its bit array, in prefix order
  for (const BatchItem& it : p.items) {
    if (!it.qsat) continue;
    const std::vector<uint8_t>& q = it.quants;
    uint32_t m = it.m0, remaining = it.L - 5 - it.jb, level = 5 + it.jb;
    const unsigned int* in = reinterpret_cast<const unsigned int*>(b.bits + it.bits_off);
    int which = 0;
    while (true) {
      const uint32_t k = remaining < 11 ? remaining : 11;
      const bool final_pass = k == remaining;
      const uint32_t mo = m - k;
      const uint64_t owords = mo >= 5 ? (1ull << (mo - 5)) : 1ull;
      unsigned int* o = b.pp[which];
      NS_HIP(launch_fill_u32(s, o, 0u, owords + 1));
      uint32_t opmask = 0;
      for (uint32_t l = 0; l < k; l++)
        if (q[it.n - 1 - (level + l)]) opmask |= 1u << l;
      NS_HIP(launch_qsat_reduce(s, in, m, k, opmask, o, final_pass, it.q0all, b.stats + it.idx * kStatsWords));
      if (final_pass) break;
      in = o;
      m = mo;
      remaining -= k;
      level += k;
      which ^= 1;
    }
  }
}

// request results from the read-back counters
void decode(const BatchPlan& p, const std::vector<unsigned long long>& host, double kernel_ms,
            std::vector<ns_request_result>& res) {
  for (const BatchItem& it : p.items) {
    const unsigned long long* h = host.data() + it.idx * kStatsWords;
    unsigned long long count = h[0];
    for (uint64_t s = 0; s < 64; s++) count += h[512 + s * 16];
    const bool hit = count != 0;
    ns_request_result& rr = res[it.idx];
    if (it.qsat) {
      rr.v0 = (hit ? 1u : 0u) ^ (it.q0all ? 1u : 0u);
      rr.v1 = hit ? h[1] : 0;
      rr.v2 = count;
      rr.v3 = hit ? 1 : 0;
    } else {
      rr.v0 = count;
      rr.v1 = hit ? h[1] : 0;
      rr.v2 = hit ? 1 : 0;
      rr.v3 = 0;
    }
    rr.status = ST_OK;
    rr.kernel_ms = kernel_ms;
  }
}

// the batch's identity: every request's circuit, kind and prefix, in order
std::string batch_key(const std::vector<const ServeRequest*>& reqs) {
  std::string k;
  for (const ServeRequest* r : reqs) {
    k += circuit_fingerprint(*r->circuit);
    k += '/' + std::to_string(r->kind) + '/';
    for (uint8_t q : r->quants) k += q ? 'a' : 'e';
    k += ';';
  }
  return k;
}

}  // namespace

Batcher::Batcher(Context& ctx, const ns_serve_opts& o, bool keep_plans) : ctx_(ctx), keep_plans_(keep_plans) {
  limit_ = o.batch_max > 1 ? std::min<uint32_t>(o.batch_max, 256) : 32;
  // body generation threads per batch: the host's cores (at most 16), shared among the
  // server's workers when the batcher serves several
  const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  gen_threads_ = keep_plans ? hw : std::max(1u, hw / std::max(1u, o.workers ? o.workers : 8u));
  hipDeviceProp_t prop;
  NS_HIP(hipGetDeviceProperties(&prop, ctx.device));
  cus_ = (uint32_t)prop.multiProcessorCount;
  // the autotuner's candidates per class: persistent blocks per CU (the class's occupancy, or
  // twice it: the second half waits for free slots) x chunks per resident wave
  for (int c = 0; c < 4; c++) {
    for (uint32_t f : {1u, 2u})
      for (uint32_t cpw : {2u, 4u, 8u, 16u}) tune_[c].cand.push_back({kOccupancy[c] * f, cpw});
    tune_[c].sum.assign(tune_[c].cand.size(), 0.0);
    tune_[c].n.assign(tune_[c].cand.size(), 0);
  }
  // NS_BATCH_TUNE=0 keeps the default launch shape (no search)
  if (const char* e = std::getenv("NS_BATCH_TUNE")) fixed_tuning_ = e[0] == '0';
}

Batcher::~Batcher() = default;

bool Batcher::batchable(const ServeRequest& r) {
  const uint32_t n = r.circuit->n;
  if (n == 0 || n > 36 || r.circuit->outputs.size() != 1) return false;
  if (r.kind == NS_REQ_SAT) return true;
  if (r.kind != NS_REQ_QSAT || r.quants.size() != n) return false;
  const uint32_t L = n - qsat_first_bits(r.quants);
  // one reduced bit per thread: the levels below the outermost block cover the word and
  // the thread's words (jb >= 0), and the bit array holds at least one ballot word
  return L >= 5 && n >= 11 + std::min<uint32_t>(2, L - 5);
}

BatchTuning Batcher::tuning(uint32_t vgpr_class, uint32_t* idx) {
  std::lock_guard<std::mutex> lk(mu_);
  TuneState& t = tune_[batch_class_index(vgpr_class)];
  uint32_t k = kDefaultCandidate;
  if (!fixed_tuning_) {
    if (t.best >= 0) {
      k = (uint32_t)t.best;
    } else {
      // every candidate measured twice, in turn; then the best mean is kept
      k = 0;
      for (uint32_t i = 1; i < t.cand.size(); i++)
        if (t.n[i] < t.n[k]) k = i;
    }
  }
  if (idx) *idx = k;
  return t.cand[k];
}

void Batcher::report(uint32_t vgpr_class, uint32_t idx, double kernel_ms, double total_cost) {
  // batches shorter than this say little about the launch shape
  if (fixed_tuning_ || kernel_ms < 0.2 || total_cost <= 0) return;
  std::lock_guard<std::mutex> lk(mu_);
  TuneState& t = tune_[batch_class_index(vgpr_class)];
  if (t.best >= 0 || idx >= t.cand.size()) return;
  t.sum[idx] += kernel_ms / total_cost;
  t.n[idx]++;
  for (int v : t.n)
    if (v < 2) return;
  int best = 0;
  for (size_t i = 1; i < t.cand.size(); i++)
    if (t.sum[i] / t.n[i] < t.sum[best] / t.n[best]) best = (int)i;
  t.best = best;
}

BatchTuning Batcher::chosen_tuning(uint32_t vgpr_class, bool* settled) const {
  std::lock_guard<std::mutex> lk(mu_);
  const TuneState& t = tune_[batch_class_index(vgpr_class)];
  if (settled) *settled = t.best >= 0;
  return t.cand[t.best >= 0 ? (size_t)t.best : kDefaultCandidate];
}

std::shared_ptr<BatchPlan> Batcher::find_plan(const Lane* lane, const std::string& key) {
  std::lock_guard<std::mutex> lk(plan_mu_);
  for (auto it = plans_.begin(); it != plans_.end(); ++it)
    if ((*it)->lane == lane && (*it)->key == key) {
      auto p = *it;
      plans_.erase(it);
      plans_.push_front(p);  // most recent first
      return p;
    }
  return nullptr;
}

void Batcher::keep_plan(std::shared_ptr<BatchPlan> p) {
  std::lock_guard<std::mutex> lk(plan_mu_);
  plans_.push_front(std::move(p));
  while (plans_.size() > kKeptPlans) plans_.pop_back();
}

void Batcher::run(Lane& lane, const std::vector<const ServeRequest*>& reqs, std::vector<ns_request_result>& res,
                  BatchStats* bs) {
  const auto t0 = clk::now();
  NS_HIP(hipSetDevice(lane.device));
  const size_t K = reqs.size();
  std::string key;
  if (keep_plans_) {
    key = batch_key(reqs);
    if (auto p = find_plan(&lane, key)) {
      replay(lane, *p, res, bs);
      if (bs) bs->wall_ms = ms_since(t0);
      return;
    }
  }
  auto plan = std::make_shared<BatchPlan>();
  BatchPlan& p = *plan;
  p.key = key;
  p.lane = &lane;
  p.K = K;
  // code generation: one body per request, on up to gen_threads_ threads
  std::vector<BatchItem> all(K);
  auto gen_one = [&](size_t i) {
    BatchItem& it = all[i];
    const ServeRequest& r = *reqs[i];
    const Circuit& c = *r.circuit;
    it.idx = i;
    it.n = c.n;
    it.qsat = r.kind == NS_REQ_QSAT;
    it.words = c.n >= 5 ? (1ull << (c.n - 5)) : 1ull;
    if (!batchable(r)) {
      it.fallback = true;
      return;
    }
    const auto tg = clk::now();
    std::vector<uint32_t> vb(c.n);
    uint32_t lane_shift = 0;
    if (it.qsat) {
      it.L = c.n - qsat_first_bits(r.quants);
      it.jb = std::min<uint32_t>(2, it.L - 5);
      it.m0 = c.n - 5 - it.jb;
      it.q0all = r.quants[0] != 0;
      it.quants = r.quants;
      for (uint32_t l = 0; l < 32 && l < c.n; l++)
        if (r.quants[c.n - 1 - l]) it.opm |= 1u << l;
      lane_shift = it.jb;
      for (uint32_t k = 0; k < c.n; k++) vb[k] = c.n - 1 - k;
    } else {
      for (uint32_t k = 0; k < c.n; k++) vb[k] = k;
    }
    try {
      NativeKernel nk = build_native(c, NativeKind::BATCH, vb, lane_shift, {0});
      std::string err;
      if (!encode_body(nk.body, it.code, &err)) {
        it.fallback = true;
      } else {
        it.vgpr = nk.vgpr_class;
        it.cost = nk.info.valu_ops + nk.info.scalar_ops + 16;  // issue slots per word
      }
    } catch (const NativeUnfit&) {
      it.fallback = true;
    }
    it.gen_ms = ms_since(tg);
  };
  const size_t nth = std::min<size_t>(gen_threads_, K);
  if (nth <= 1) {
    for (size_t i = 0; i < K; i++) gen_one(i);
  } else {
    std::atomic<size_t> next{0};
    std::vector<std::exception_ptr> err(nth);
    std::vector<std::thread> th;
    for (size_t t = 0; t < nth; t++)
      th.emplace_back([&, t] {
        try {
          for (size_t i; (i = next.fetch_add(1)) < K;) gen_one(i);
        } catch (...) {
          err[t] = std::current_exception();
          next = K;
        }
      });
    for (auto& x : th) x.join();
    for (auto& e : err)
      if (e) std::rethrow_exception(e);
  }
  for (const BatchItem& it : all) p.gen_ms += it.gen_ms;
  p.gen_wall_ms = ms_since(t0);
  // bodies past the largest batch region are solved alone, like the requests the batch
  // kernels do not take (NS_BATCH_REGION_BYTES lowers that size, for tests)
  {
    size_t at = 0;
    size_t cap = max_batch_region_bytes();
    if (const char* e = std::getenv("NS_BATCH_REGION_BYTES")) cap = std::min<size_t>(cap, std::strtoull(e, nullptr, 0));
    for (BatchItem& it : all) {
      if (it.fallback) continue;
      const size_t next = (at + it.code.size() * 4 + 4 + 255) & ~(size_t)255;
      if (next > cap) {
        it.fallback = true;
        std::vector<uint32_t>().swap(it.code);
      } else {
        at = next;
      }
    }
  }
  // descriptor order: by kind (SAT, QSAT), then by register class: one group each
  for (int kind = 0; kind < 2; kind++)
    for (int ci = 0; ci < 4; ci++) {
      BatchGroup g;
      g.qsat = kind == 1;
      g.ci = ci;
      g.first = p.items.size();
      for (BatchItem& it : all)
        if (!it.fallback && it.qsat == g.qsat && batch_class_index(it.vgpr) == ci) {
          p.cls = std::max(p.cls, it.vgpr);
          p.items.push_back(it);
        }
      g.count = p.items.size() - g.first;
      if (g.count) p.groups.push_back(g);
    }
  double kernel_ms = 0;
  if (!p.items.empty()) {
    std::vector<std::vector<uint32_t>> bodies;
    for (const BatchItem& it : p.items) bodies.push_back(it.code);
    auto tpl = batch_template(batch_bytes(bodies));
    if (!tpl) throw Error(ST_UNSUPPORTED, "the batch's circuit bodies exceed the largest batch region");
    const auto tl = clk::now();
    std::vector<uint64_t> off;
    std::vector<char> elf = compose_batch(*tpl, bodies, off);
    p.code_bytes = elf.size();
    NS_HIP(hipModuleLoadData(&p.mod.mod, elf.data()));
    StreamDrain drain{lane.stream};
    for (BatchGroup& g : p.groups) {
      NS_HIP(hipModuleGetFunction(&g.fn, p.mod.mod, tpl->kernel[g.qsat ? 1 : 0][g.ci].c_str()));
      for (hipEvent_t& e : g.ev) NS_HIP(hipEventCreate(&e));
    }
    p.load_ms = ms_since(tl);
    for (BatchGroup& g : p.groups) {
      // launch shape and chunks: about chunks_per_wave chunks per resident wave over the
      // group, each request's chunks of similar cost (words x body cost)
      const BatchTuning tu = tuning(kClassOf[g.ci], &g.tune_idx);
      g.grid = cus_ * tu.blocks_per_cu;
      const uint64_t waves = (uint64_t)g.grid * 4;
      for (size_t b = g.first; b < g.first + g.count; b++) g.cost += (double)p.items[b].words * p.items[b].cost;
      const double target = std::max(1.0, g.cost / (double)(waves * std::max<uint32_t>(1, tu.chunks_per_wave)));
      if (!g.qsat) g.chunk_base = p.csat;
      for (size_t b = g.first; b < g.first + g.count; b++) {
        BatchItem& it = p.items[b];
        const uint64_t tile = 64ull << it.jb;
        uint64_t cw = (uint64_t)(target / it.cost);
        cw = (cw + tile - 1) / tile * tile;
        cw = std::max(cw, tile);
        cw = std::min<uint64_t>(cw, (it.words + tile - 1) / tile * tile);
        cw = std::min<uint64_t>(cw, kMaxChunkWords);
        it.chunk_w = cw;
        it.chunks = (it.words + cw - 1) / cw;
        Desc d{};
        d.wend = (uint32_t)it.words;
        d.words = (uint32_t)it.chunk_w;
        const uint64_t entry = tpl->region + off[b] - tpl->pc_base[it.qsat ? 1 : 0][g.ci];
        if (entry > 0xffffffffull) throw Error(ST_INTERNAL, "batch body entry out of range");
        d.entry = (uint32_t)entry;
        d.jb = it.jb;
        d.opm = it.opm;
        d.valid = c_valid_mask(it.n);
        d.chunk0 = (uint32_t)g.chunks;  // chunks numbered within the group
        g.chunks += it.chunks;
        if (it.qsat) {
          it.bits_off = p.qbytes;
          p.qbytes += ((1ull << it.m0) / 8 + 255) & ~255ull;
          // ping-pong outputs of the reduction passes: the first is the largest
          const uint32_t k = std::min<uint32_t>(it.L - 5 - it.jb, 11);
          const uint32_t mo = it.m0 - k;
          const uint64_t ow = mo >= 5 ? (1ull << (mo - 5)) : 1ull;
          p.pp_bytes = std::max(p.pp_bytes, (ow + 1) * 4);
        }
        p.descs.push_back(d);
      }
      if (g.chunks > 0xffffffffull) throw Error(ST_INTERNAL, "too many batch chunks");
      if (!g.qsat) p.csat += g.chunks;
    }
    // a first run: the lane's own buffers
    lane.bstats.ensure(K * kStatsBytes);
    lane.bdesc.ensure(p.descs.size() * sizeof(Desc));
    if (p.csat) lane.bchunk.ensure(p.csat * 16);
    if (p.qbytes) {
      lane.buf_a.ensure(p.qbytes);
      lane.buf_b.ensure(p.pp_bytes);
      lane.buf_c.ensure(p.pp_bytes);
    }
    const BatchBufs bufs{lane.bdesc.as<Desc>(), lane.bstats.as<unsigned long long>(),
                         lane.bchunk.as<unsigned long long>(), lane.buf_a.as<char>(),
                         {lane.buf_b.as<unsigned int>(), lane.buf_c.as<unsigned int>()}};
    const std::vector<Desc> bound = bind(p, bufs);
    NS_HIP(hipMemcpyAsync(bufs.desc, bound.data(), bound.size() * sizeof(Desc), hipMemcpyHostToDevice, lane.stream));
    NS_HIP(hipEventRecord(lane.ev0, lane.stream));
    enqueue(p, bufs, lane.stream, true);
    NS_HIP(hipEventRecord(lane.ev1, lane.stream));
    std::vector<unsigned long long> host(K * kStatsWords);
    NS_HIP(hipMemcpyAsync(host.data(), bufs.stats, K * kStatsBytes, hipMemcpyDeviceToHost, lane.stream));
    NS_HIP(hipStreamSynchronize(lane.stream));
    kernel_ms = lane.elapsed_ms();
    for (const BatchGroup& g : p.groups) {
      float ms = 0.f;
      NS_HIP(hipEventElapsedTime(&ms, g.ev[0], g.ev[1]));
      report(kClassOf[g.ci], g.tune_idx, ms, g.cost);
    }
    decode(p, host, kernel_ms, res);
    for (const BatchItem& it : p.items) res[it.idx].compile_ms = it.gen_ms + p.load_ms / (double)p.items.size();
  }
  // the requests the batch kernels could not take, one at a time on the lane
  uint32_t fallbacks = 0;
  for (size_t i = 0; i < K; i++) {
    if (!all[i].fallback) continue;
    fallbacks++;
    try {
      solve_alone(lane, *reqs[i], res[i]);
      res[i].status = ST_OK;
    } catch (const Error& e) {
      res[i].status = e.code;
    } catch (const std::exception&) {
      res[i].status = ST_INTERNAL;
    }
  }
  if (bs) {
    *bs = BatchStats{};
    bs->requests = (uint32_t)K;
    bs->fallbacks = fallbacks;
    bs->vgpr_class = p.items.empty() ? 0 : p.cls;
    bs->launches = (uint32_t)p.groups.size();
    bs->code_bytes = p.code_bytes;
    bs->gen_ms = p.gen_wall_ms;
    bs->load_ms = p.load_ms;
    bs->kernel_ms = kernel_ms;
    bs->wall_ms = ms_since(t0);
  }
  p.runs = 1;
  // kept for a repeat when every request ran in the batch kernels
  if (keep_plans_ && fallbacks == 0 && !p.items.empty()) {
    for (BatchItem& it : p.items) std::vector<uint32_t>().swap(it.code);  // the module holds the code
    keep_plan(std::move(plan));
  }
}

// A kept batch again: its own buffers and graph (captured on the first repeat, replayed
// afterwards); no code generation, no module load.
void Batcher::replay(Lane& lane, BatchPlan& p, std::vector<ns_request_result>& res, BatchStats* bs) {
  StreamDrain drain{lane.stream};
  if (!p.graph) {
    p.d_stats.ensure(p.K * kStatsBytes);
    p.d_desc.ensure(p.descs.size() * sizeof(Desc));
    if (p.csat) p.d_chunk.ensure(p.csat * 16);
    if (p.qbytes) {
      p.d_bits.ensure(p.qbytes);
      p.d_pp_a.ensure(p.pp_bytes);
      p.d_pp_b.ensure(p.pp_bytes);
    }
    const BatchBufs bufs{p.d_desc.as<Desc>(), p.d_stats.as<unsigned long long>(), p.d_chunk.as<unsigned long long>(),
                         p.d_bits.as<char>(), {p.d_pp_a.as<unsigned int>(), p.d_pp_b.as<unsigned int>()}};
    const std::vector<Desc> bound = bind(p, bufs);
    NS_HIP(hipMemcpyAsync(bufs.desc, bound.data(), bound.size() * sizeof(Desc), hipMemcpyHostToDevice, lane.stream));
    NS_HIP(hipStreamSynchronize(lane.stream));
    hipGraph_t graph = nullptr;
    NS_HIP(hipStreamBeginCapture(lane.stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue(p, bufs, lane.stream, false);
    } catch (...) {
      (void)hipStreamEndCapture(lane.stream, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    NS_HIP(hipStreamEndCapture(lane.stream, &graph));
    const hipError_t ie = hipGraphInstantiate(&p.graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    NS_HIP(ie);
  }
  NS_HIP(hipEventRecord(lane.ev0, lane.stream));
  NS_HIP(hipGraphLaunch(p.graph, lane.stream));
  NS_HIP(hipEventRecord(lane.ev1, lane.stream));
  std::vector<unsigned long long> host(p.K * kStatsWords);
  NS_HIP(hipMemcpyAsync(host.data(), p.d_stats.p, p.K * kStatsBytes, hipMemcpyDeviceToHost, lane.stream));
  NS_HIP(hipStreamSynchronize(lane.stream));
  const double kernel_ms = lane.elapsed_ms();
  decode(p, host, kernel_ms, res);
  for (const BatchItem& it : p.items) res[it.idx].compile_ms = 0;
  p.runs++;
  if (bs) {
    *bs = BatchStats{};
    bs->requests = (uint32_t)p.K;
    bs->vgpr_class = p.cls;
    bs->launches = (uint32_t)p.groups.size();
    bs->code_bytes = p.code_bytes;
    bs->kernel_ms = kernel_ms;
    bs->graph_runs = p.runs - 1;
  }
}

void batch_solve(Context& ctx, const std::vector<const ServeRequest*>& reqs, std::vector<ns_request_result>& res,
                 BatchStats* stats) {
  res.assign(reqs.size(), ns_request_result{});
  ctx.batcher().run(*ctx.main, reqs, res, stats);
}

}  // namespace ns
