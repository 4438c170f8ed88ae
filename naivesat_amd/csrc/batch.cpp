// Batched SAT / QSAT solving with the persistent batch kernels (templates/gen_templates.py:
// ns_nat_bsat_rC, ns_nat_bqsat_rC). One batch = one code object holding every request's
// body, one descriptor array per kernel, at most two batch kernel launches (plus the fold of
// the SAT chunk results and the QSAT reduction passes of each QSAT request), one read-back
// of all results.
#include "batch.h"

#include <algorithm>
#include <chrono>
#include <cstring>

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

int class_index(uint32_t vgpr_class) {
  switch (vgpr_class) {
    case 64: return 0;
    case 128: return 1;
    case 256: return 2;
    default: return 3;
  }
}

// waves per SIMD the batch kernels reach in each class (SGPRs limit r64 to 7)
const uint32_t kOccupancy[4] = {7, 4, 2, 1};

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

struct Item {
  size_t idx = 0;
  bool qsat = false, fallback = false;
  uint32_t n = 0, jb = 0, vgpr = 0, cost = 0, m0 = 0;
  uint64_t words = 0;     // words of the space
  uint64_t chunk_w = 0;   // words per chunk
  uint64_t chunks = 0;
  uint64_t bits_off = 0;  // QSAT: byte offset of its bit array in lane.buf_a
  std::vector<uint32_t> code;
  double gen_ms = 0;
};

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

Batcher::Batcher(Context& ctx, const ns_serve_opts& o) : ctx_(ctx) {
  limit_ = o.batch_max > 1 ? std::min<uint32_t>(o.batch_max, 256) : 32;
  hipDeviceProp_t prop;
  NS_HIP(hipGetDeviceProperties(&prop, ctx.device));
  cus_ = (uint32_t)prop.multiProcessorCount;
}

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

BatchTuning Batcher::tuning(uint32_t vgpr_class) const {
  std::lock_guard<std::mutex> lk(mu_);
  return tune_[class_index(vgpr_class)];
}

void Batcher::set_tuning(uint32_t vgpr_class, const BatchTuning& t) {
  std::lock_guard<std::mutex> lk(mu_);
  tune_[class_index(vgpr_class)] = t;
}

void Batcher::run(Lane& lane, const std::vector<const ServeRequest*>& reqs, std::vector<ns_request_result>& res,
                  BatchStats* bs) {
  const auto t0 = clk::now();
  NS_HIP(hipSetDevice(lane.device));
  const size_t K = reqs.size();
  std::vector<Item> items(K);
  // code generation: one body per request (the caller's thread)
  for (size_t i = 0; i < K; i++) {
    Item& it = items[i];
    const ServeRequest& r = *reqs[i];
    const Circuit& c = *r.circuit;
    it.idx = i;
    it.n = c.n;
    it.qsat = r.kind == NS_REQ_QSAT;
    it.words = c.n >= 5 ? (1ull << (c.n - 5)) : 1ull;
    if (!batchable(r)) {
      it.fallback = true;
      continue;
    }
    const auto tg = clk::now();
    std::vector<uint32_t> vb(c.n);
    uint32_t lane_shift = 0;
    if (it.qsat) {
      const uint32_t L = c.n - qsat_first_bits(r.quants);
      it.jb = std::min<uint32_t>(2, L - 5);
      it.m0 = c.n - 5 - it.jb;
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
        it.cost = nk.info.valu_ops + nk.info.scalar_ops / 2 + 16;
      }
    } catch (const NativeUnfit&) {
      it.fallback = true;
    }
    it.gen_ms = ms_since(tg);
  }
  std::vector<size_t> batched;
  uint32_t cls = 64;
  for (size_t i = 0; i < K; i++)
    if (!items[i].fallback) {
      batched.push_back(i);
      cls = std::max(cls, items[i].vgpr);
    }
  double load_ms = 0, kernel_ms = 0;
  uint64_t code_bytes = 0;
  if (!batched.empty()) {
    std::vector<std::vector<uint32_t>> bodies;
    for (size_t i : batched) bodies.push_back(items[i].code);
    auto tpl = batch_template(cls, batch_bytes(bodies));
    if (!tpl) throw Error(ST_UNSUPPORTED, "the batch's circuit bodies exceed the largest batch region");
    const auto tl = clk::now();
    std::vector<uint64_t> off;
    std::vector<char> elf = compose_batch(*tpl, bodies, off);
    code_bytes = elf.size();
    JitModule mod;  // unloaded when the batch is done, after the drain below
    NS_HIP(hipModuleLoadData(&mod.mod, elf.data()));
    StreamDrain drain{lane.stream};
    hipFunction_t fsat = nullptr, fqsat = nullptr;
    NS_HIP(hipModuleGetFunction(&fsat, mod.mod, tpl->sat_kernel.c_str()));
    NS_HIP(hipModuleGetFunction(&fqsat, mod.mod, tpl->qsat_kernel.c_str()));
    load_ms = ms_since(tl);
    // launch shape and chunks: about chunks_per_wave chunks per resident wave over the
    // batch, each request's chunks of similar cost (words x body cost)
    const BatchTuning tu = tuning(cls);
    const uint32_t bpc = tu.blocks_per_cu ? tu.blocks_per_cu : kOccupancy[class_index(cls)];
    const uint64_t waves = (uint64_t)cus_ * bpc * 4;
    double total_cost = 0;
    for (size_t i : batched) total_cost += (double)items[i].words * items[i].cost;
    const double target = std::max(1.0, total_cost / (double)(waves * std::max<uint32_t>(1, tu.chunks_per_wave)));
    std::vector<Desc> dsat, dqsat;
    uint64_t qbytes = 0, pp_bytes = 0;
    for (size_t b = 0; b < batched.size(); b++) {
      Item& it = items[batched[b]];
      const uint64_t tile = 64ull << it.jb;
      uint64_t cw = (uint64_t)(target / it.cost);
      cw = (cw + tile - 1) / tile * tile;
      cw = std::max(cw, tile);
      cw = std::min<uint64_t>(cw, (it.words + tile - 1) / tile * tile);
      cw = std::min<uint64_t>(cw, kMaxChunkWords);
      it.chunk_w = cw;
      it.chunks = (it.words + cw - 1) / cw;
      if (it.qsat) {
        it.bits_off = qbytes;
        qbytes += ((1ull << it.m0) / 8 + 255) & ~255ull;
        // ping-pong outputs of the reduction passes: the first is the largest
        const uint32_t L = it.n - qsat_first_bits(reqs[it.idx]->quants);
        const uint32_t k = std::min<uint32_t>(L - 5 - it.jb, 11);
        const uint32_t mo = it.m0 - k;
        const uint64_t ow = mo >= 5 ? (1ull << (mo - 5)) : 1ull;
        pp_bytes = std::max(pp_bytes, (ow + 1) * 4);
      }
    }
    lane.bstats.ensure(K * kStatsBytes);
    lane.bdesc.ensure(K * sizeof(Desc));
    if (qbytes) {
      lane.buf_a.ensure(qbytes);
      lane.buf_b.ensure(pp_bytes);
      lane.buf_c.ensure(pp_bytes);
    }
    unsigned long long* st = lane.bstats.as<unsigned long long>();
    uint64_t csat = 0, cqsat = 0;
    for (size_t b = 0; b < batched.size(); b++) {
      Item& it = items[batched[b]];
      const ServeRequest& r = *reqs[it.idx];
      Desc d{};
      d.wend = (uint32_t)it.words;
      d.words = (uint32_t)it.chunk_w;
      const size_t pc = tpl->pc_base[it.qsat ? 1 : 0];
      const uint64_t entry = tpl->region + off[b] - pc;
      if (entry > 0xffffffffull) throw Error(ST_INTERNAL, "batch body entry out of range");
      d.entry = (uint32_t)entry;
      d.jb = it.jb;
      d.valid = c_valid_mask(it.n);
      if (it.qsat) {
        d.out = (uint64_t)(lane.buf_a.as<char>() + it.bits_off);
        for (uint32_t l = 0; l < 32 && l < it.n; l++)
          if (r.quants[it.n - 1 - l]) d.opm |= 1u << l;
        d.chunk0 = (uint32_t)cqsat;
        cqsat += it.chunks;
        dqsat.push_back(d);
      } else {
        d.out = (uint64_t)(st + it.idx * kStatsWords);
        d.chunk0 = (uint32_t)csat;
        csat += it.chunks;
        dsat.push_back(d);
      }
    }
    if (csat > 0xffffffffull || cqsat > 0xffffffffull) throw Error(ST_INTERNAL, "too many batch chunks");
    Desc* ddev = lane.bdesc.as<Desc>();
    std::vector<Desc> all = dsat;
    all.insert(all.end(), dqsat.begin(), dqsat.end());
    NS_HIP(hipMemcpyAsync(ddev, all.data(), all.size() * sizeof(Desc), hipMemcpyHostToDevice, lane.stream));
    NS_HIP(launch_fill_u64(lane.stream, st, 0ull, K * kStatsWords));
    NS_HIP(launch_fill_strided_u64(lane.stream, st + 1, ~0ull, K, kStatsWords));
    if (csat) lane.bchunk.ensure(csat * 16);
    NS_HIP(hipEventRecord(lane.ev0, lane.stream));
    // a persistent grid: every wave resident at once, each taking its chunks round-robin
    const unsigned grid = (unsigned)(cus_ * bpc);
    if (!dsat.empty()) {
      const Desc* dp = ddev;
      uint32_t nd = (uint32_t)dsat.size(), ch = (uint32_t)csat;
      unsigned long long* chunk_out = lane.bchunk.as<unsigned long long>();
      void* args[4] = {&dp, &nd, &ch, &chunk_out};
      NS_HIP(hipModuleLaunchKernel(fsat, grid, 1, 1, 256, 1, 1, 0, lane.stream, args, nullptr));
      NS_HIP(launch_batch_fold(lane.stream, reinterpret_cast<const BatchFoldDesc*>(dp), nd, chunk_out));
    }
    if (!dqsat.empty()) {
      const Desc* dp = ddev + dsat.size();
      uint32_t nd = (uint32_t)dqsat.size(), ch = (uint32_t)cqsat;
      void* args[3] = {&dp, &nd, &ch};
      NS_HIP(hipModuleLaunchKernel(fqsat, grid, 1, 1, 256, 1, 1, 0, lane.stream, args, nullptr));
      // the reduction passes of each QSAT request over its bit array, in prefix order
      for (size_t i : batched) {
        const Item& it = items[i];
        if (!it.qsat) continue;
        const std::vector<uint8_t>& q = reqs[it.idx]->quants;
        const uint32_t n = it.n, L = n - qsat_first_bits(q);
        const bool q0all = q[0] != 0;
        uint32_t m = it.m0, remaining = L - 5 - it.jb, level = 5 + it.jb;
        const unsigned int* in = reinterpret_cast<const unsigned int*>(lane.buf_a.as<char>() + it.bits_off);
        DevBuf* outs[2] = {&lane.buf_b, &lane.buf_c};
        int which = 0;
        while (true) {
          const uint32_t k = remaining < 11 ? remaining : 11;
          const bool final_pass = k == remaining;
          const uint32_t mo = m - k;
          const uint64_t owords = mo >= 5 ? (1ull << (mo - 5)) : 1ull;
          unsigned int* o = outs[which]->as<unsigned int>();
          NS_HIP(launch_fill_u32(lane.stream, o, 0u, owords + 1));
          uint32_t opmask = 0;
          for (uint32_t l = 0; l < k; l++)
            if (q[n - 1 - (level + l)]) opmask |= 1u << l;
          NS_HIP(launch_qsat_reduce(lane.stream, in, m, k, opmask, o, final_pass, q0all, st + it.idx * kStatsWords));
          if (final_pass) break;
          in = o;
          m = mo;
          remaining -= k;
          level += k;
          which ^= 1;
        }
      }
    }
    NS_HIP(hipEventRecord(lane.ev1, lane.stream));
    std::vector<unsigned long long> host(K * kStatsWords);
    NS_HIP(hipMemcpyAsync(host.data(), st, K * kStatsBytes, hipMemcpyDeviceToHost, lane.stream));
    NS_HIP(hipStreamSynchronize(lane.stream));
    kernel_ms = lane.elapsed_ms();
    for (size_t b = 0; b < batched.size(); b++) {
      const Item& it = items[batched[b]];
      const unsigned long long* h = host.data() + it.idx * kStatsWords;
      unsigned long long count = h[0];
      for (uint64_t s = 0; s < 64; s++) count += h[512 + s * 16];
      const bool hit = count != 0;
      ns_request_result& rr = res[it.idx];
      if (it.qsat) {
        const bool q0all = reqs[it.idx]->quants[0] != 0;
        rr.v0 = (hit ? 1u : 0u) ^ (q0all ? 1u : 0u);
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
      rr.compile_ms = it.gen_ms + load_ms / (double)batched.size();
    }
  }
  // the requests the batch kernels could not take, one at a time on the lane
  uint32_t fallbacks = 0;
  for (size_t i = 0; i < K; i++) {
    if (!items[i].fallback) continue;
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
    bs->requests = (uint32_t)K;
    bs->fallbacks = fallbacks;
    bs->vgpr_class = batched.empty() ? 0 : cls;
    bs->code_bytes = code_bytes;
    bs->load_ms = load_ms;
    bs->kernel_ms = kernel_ms;
    bs->gen_ms = 0;
    for (const Item& it : items) bs->gen_ms += it.gen_ms;
    bs->wall_ms = ms_since(t0);
  }
}

void batch_solve(Context& ctx, const std::vector<const ServeRequest*>& reqs, std::vector<ns_request_result>& res,
                 BatchStats* stats) {
  ns_serve_opts o{};
  o.batch_max = (uint32_t)std::max<size_t>(2, reqs.size());
  Batcher b(ctx, o);
  res.assign(reqs.size(), ns_request_result{});
  b.run(*ctx.main, reqs, res, stats);
}

}  // namespace ns
