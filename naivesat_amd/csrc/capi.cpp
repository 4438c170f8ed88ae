// extern "C" boundary: converts exceptions to status codes and keeps the last error
// message per thread.
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/naivesat_hip.h"
#include "asmenc.h"
#include "asmgen.h"
#include "circuit.h"
#include "runtime.h"
#include "file_image.h"
#include "batch.h"
#include "serve.h"
#include "static_kernels.h"
#include "solvers.h"
#include "static_kernels.h"

struct ns_circuit {
  ns::Circuit c;
};
struct ns_ctx {
  ns::Context ctx;
  explicit ns_ctx(int d) : ctx(d) {}
};

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
  try {
    f();
    g_err.clear();
    return NS_OK;
  } catch (const ns::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory";
    return NS_ERR_NOMEM;
  } catch (const std::out_of_range& e) {
    g_err = e.what();
    return NS_ERR_INVALID;
  } catch (const std::runtime_error& e) {
    g_err = e.what();
    return NS_ERR_INVALID;
  } catch (const std::exception& e) {
    g_err = e.what();
    return NS_ERR_INTERNAL;
  } catch (...) {
    g_err = "unknown error";
    return NS_ERR_INTERNAL;
  }
}

void need(bool ok, const char* what) {
  if (!ok) throw ns::Error(NS_ERR_INVALID, what);
}

// native kernel of a diagnostic kind: 0 SAT, 1 QSAT-A, 2 QSAT-B, 3 table, 4 batch body in
// the SAT input order, 5 batch body in the QSAT input order, 6 table-free walk body
// (lane_shift = waypoint bits), 7 SAT with group exits
ns::NativeKernel native_kernel(const ns::Circuit& cc, uint32_t kind, uint32_t lane_shift) {
  if (kind == 6) {
    try {
      return ns::build_native_walk(cc, lane_shift);
    } catch (const ns::NativeUnfit& e) {
      throw ns::Error(NS_ERR_UNSUPPORTED, e.what());
    }
  }
  std::vector<uint32_t> vb(cc.n), outs;
  for (uint32_t k = 0; k < cc.n; k++) vb[k] = (kind == 1 || kind == 2 || kind == 5) ? cc.n - 1 - k : k;
  if (kind == 3)
    for (uint32_t k = 0; k < cc.outputs.size(); k++) outs.push_back(k);
  else
    outs.push_back(0);
  static const ns::NativeKind kinds[8] = {ns::NativeKind::SAT,   ns::NativeKind::QSAT_A, ns::NativeKind::QSAT_B,
                                          ns::NativeKind::TABLE, ns::NativeKind::BATCH,  ns::NativeKind::BATCH,
                                          ns::NativeKind::SAT,   ns::NativeKind::SAT};
  try {
    return ns::build_native(cc, kinds[kind], vb, lane_shift, outs, kind == 7);
  } catch (const ns::NativeUnfit& e) {
    throw ns::Error(NS_ERR_UNSUPPORTED, e.what());
  }
}
}  // namespace

extern "C" {

const char* ns_last_error(void) { return g_err.c_str(); }
const char* ns_version(void) { return "naivesat_hip 0.1.0 (gfx950)"; }

int ns_circuit_parse(const char* text, size_t len, ns_circuit** out) {
  return guard([&] {
    need(text && out, "null argument");
    auto* h = new ns_circuit{ns::parse_circuit(std::string(text, len))};
    *out = h;
  });
}

int ns_quant_circuit_parse(const char* text, size_t len, ns_circuit** out, uint8_t* quants_all,
                           size_t quants_cap, int* is_quant) {
  return guard([&] {
    need(text && out && is_quant, "null argument");
    std::string s(text, len);
    std::vector<uint8_t> q;
    ns::Circuit c;
    if (ns::parse_quant_circuit(s, q, c)) {
      need(quants_all && quants_cap >= q.size(), "quantifier buffer too small");
      std::memcpy(quants_all, q.data(), q.size());
      *is_quant = 1;
    } else {
      c = ns::parse_circuit(s);
      if (quants_all) {
        need(quants_cap >= c.n, "quantifier buffer too small");
        std::memset(quants_all, 0, c.n);
      }
      *is_quant = 0;
    }
    *out = new ns_circuit{std::move(c)};
  });
}

int ns_circuit_from_arrays(uint32_t n_inputs, uint32_t n_gates, const uint32_t* in0, const uint32_t* in1,
                           const uint8_t* func, uint32_t n_outputs, const uint32_t* out_wire,
                           const uint8_t* out_neg, ns_circuit** out) {
  return guard([&] {
    need(out, "null argument");
    need(n_gates == 0 || (in0 && in1 && func), "null gate arrays");
    need(n_outputs == 0 || out_wire, "null output array");
    ns::Circuit c;
    c.n = n_inputs;
    c.gates.resize(n_gates);
    for (uint32_t j = 0; j < n_gates; j++) c.gates[j] = ns::Gate{in0[j], in1[j], func[j]};
    c.outputs.resize(n_outputs);
    for (uint32_t i = 0; i < n_outputs; i++) c.outputs[i] = ns::Output{out_wire[i], out_neg ? out_neg[i] != 0 : false};
    c.validate();
    *out = new ns_circuit{std::move(c)};
  });
}

int ns_circuit_info(const ns_circuit* c, uint32_t* n_inputs, uint32_t* n_gates, uint32_t* n_outputs) {
  return guard([&] {
    need(c, "null circuit");
    if (n_inputs) *n_inputs = c->c.n;
    if (n_gates) *n_gates = (uint32_t)c->c.gates.size();
    if (n_outputs) *n_outputs = (uint32_t)c->c.outputs.size();
  });
}

int ns_circuit_to_arrays(const ns_circuit* c, uint32_t* in0, uint32_t* in1, uint8_t* func,
                         uint32_t* out_wire, uint8_t* out_neg) {
  return guard([&] {
    need(c, "null circuit");
    for (size_t j = 0; j < c->c.gates.size(); j++) {
      if (in0) in0[j] = c->c.gates[j].i0;
      if (in1) in1[j] = c->c.gates[j].i1;
      if (func) func[j] = c->c.gates[j].func;
    }
    for (size_t i = 0; i < c->c.outputs.size(); i++) {
      if (out_wire) out_wire[i] = c->c.outputs[i].wire;
      if (out_neg) out_neg[i] = c->c.outputs[i].neg ? 1 : 0;
    }
  });
}

int ns_circuit_format(const ns_circuit* c, char* buf, size_t cap, size_t* needed) {
  return guard([&] {
    need(c, "null circuit");
    std::string s = ns::format_circuit(c->c);
    if (needed) *needed = s.size();
    if (buf && cap) {
      need(cap > s.size(), "text buffer too small");
      std::memcpy(buf, s.c_str(), s.size() + 1);
    }
  });
}

int ns_circuit_eval(const ns_circuit* c, const uint8_t* inputs, uint8_t* outputs) {
  return guard([&] {
    need(c && inputs && outputs, "null argument");
    std::vector<bool> in(c->c.n);
    for (uint32_t k = 0; k < c->c.n; k++) in[k] = inputs[k] != 0;
    auto o = c->c.eval(in);
    for (size_t i = 0; i < o.size(); i++) outputs[i] = o[i] ? 1 : 0;
  });
}

int ns_circuit_cone(const ns_circuit* c, uint32_t* cone_gates, uint32_t* peak) {
  return guard([&] {
    need(c, "null circuit");
    auto cone = ns::output_cone(c->c);
    auto order = ns::schedule_gates(c->c, cone);
    if (cone_gates) *cone_gates = (uint32_t)order.size();
    if (peak) *peak = ns::peak_live(c->c, order);
  });
}

void ns_circuit_free(ns_circuit* c) { delete c; }

int ns_gen_sat(uint64_t seed, uint32_t n, uint32_t gates, uint32_t word_bits, ns_circuit** out, uint64_t* key) {
  return guard([&] {
    need(out, "null argument");
    *out = new ns_circuit{ns::gen_sat_circuit(seed, n, gates, word_bits, key)};
  });
}

int ns_gen_factor(uint64_t product, uint32_t bits, ns_circuit** out) {
  return guard([&] {
    need(out, "null argument");
    *out = new ns_circuit{ns::gen_factor_circuit(product, bits)};
  });
}

int ns_gen_qsat(uint64_t seed, uint32_t n, uint32_t gates, uint32_t last, ns_circuit** out) {
  return guard([&] {
    need(out, "null argument");
    *out = new ns_circuit{ns::gen_qsat_circuit(seed, n, gates, last)};
  });
}

int ns_gen_satostop(uint64_t seed, uint32_t half, uint32_t rounds, uint32_t round_gates, uint32_t unknowns,
                    ns_circuit** out, uint64_t* target) {
  return guard([&] {
    need(out, "null argument");
    *out = new ns_circuit{ns::gen_satostop_circuit(seed, half, rounds, round_gates, unknowns, target)};
  });
}

int ns_device_list(uint32_t* count, char* names, size_t cap) {
  return guard([&] {
    need(count, "null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    (void)hipGetLastError();
    std::string all;
    for (int d = 0; d < n; d++) {
      hipDeviceProp_t prop;
      std::string name = "unknown";
      if (hipGetDeviceProperties(&prop, d) == hipSuccess) {
        name = prop.gcnArchName;
        size_t colon = name.find(':');
        if (colon != std::string::npos) name = name.substr(0, colon);
      }
      if (d) all += ',';
      all += name;
    }
    *count = (uint32_t)n;
    if (names && cap) {
      size_t k = std::min(cap - 1, all.size());
      std::memcpy(names, all.data(), k);
      names[k] = 0;
    }
  });
}

int ns_ctx_create(int device, ns_ctx** out) {
  return guard([&] {
    need(out, "null argument");
    *out = new ns_ctx(device);
  });
}

void ns_ctx_destroy(ns_ctx* ctx) {
  try {
    delete ctx;
  } catch (...) {
  }
}

int ns_ctx_trim(ns_ctx* ctx) {
  return guard([&] {
    need(ctx, "null argument");
    NS_HIP(hipSetDevice(ctx->ctx.main->device));
    NS_HIP(hipStreamSynchronize(ctx->ctx.main->stream));
    ctx->ctx.main->trim_tables();
  });
}

int ns_sat_solve(ns_ctx* ctx, const ns_circuit* c, const ns_sat_opts* opts, ns_sat_result* res) {
  return guard([&] {
    need(ctx && c && res, "null argument");
    ns_sat_opts o{};
    if (opts) o = *opts;
    ns::sat_solve(*ctx->ctx.main, c->c, o, *res);
  });
}

int ns_sat_solve_progress(ns_ctx* ctx, const ns_circuit* c, const ns_sat_opts* opts, ns_sat_progress_fn fn,
                          void* user, ns_sat_result* res) {
  return guard([&] {
    need(ctx && c && res && fn, "null argument");
    ns_sat_opts o{};
    if (opts) o = *opts;
    ns::sat_solve(*ctx->ctx.main, c->c, o, *res, fn, user);
  });
}

int ns_sat_solve_noalloc(ns_ctx* ctx, const ns_circuit* c, const ns_sat_opts* opts, ns_sat_result* res) {
  return guard([&] {
    need(ctx && c && res, "null argument");
    ns_sat_opts o{};
    if (opts) o = *opts;
    ns::sat_solve_noalloc(*ctx->ctx.main, c->c, o, *res, ns::noalloc_scratch(ctx->ctx.device));
  });
}

int ns_qsat_solve(ns_ctx* ctx, const ns_circuit* c, const uint8_t* quants_all, const ns_qsat_opts* opts,
                  ns_qsat_result* res, uint32_t* bitmap, uint64_t bitmap_words) {
  return guard([&] {
    need(ctx && c && res && quants_all, "null argument");
    ns_qsat_opts o{};
    if (opts) o = *opts;
    std::vector<uint8_t> q(quants_all, quants_all + c->c.n);
    for (auto& v : q) v = v ? 1 : 0;
    ns::qsat_solve(*ctx->ctx.main, c->c, q, o, *res, bitmap, bitmap_words);
  });
}

int ns_qsat_solve_progress(ns_ctx* ctx, const ns_circuit* c, const uint8_t* quants_all,
                           const ns_qsat_opts* opts, ns_sat_progress_fn fn, void* user, ns_qsat_result* res,
                           uint32_t* bitmap, uint64_t bitmap_words) {
  return guard([&] {
    need(ctx && c && res && quants_all && fn, "null argument");
    ns_qsat_opts o{};
    if (opts) o = *opts;
    std::vector<uint8_t> q(quants_all, quants_all + c->c.n);
    for (auto& v : q) v = v ? 1 : 0;
    ns::qsat_solve(*ctx->ctx.main, c->c, q, o, *res, bitmap, bitmap_words, fn, user);
  });
}

int ns_satostop_solve(ns_ctx* ctx, const ns_circuit* c, uint32_t unknowns, const ns_satostop_opts* opts,
                      ns_satostop_result* res, uint32_t* steps_out, uint32_t* ends_out) {
  return guard([&] {
    need(ctx && c && res, "null argument");
    need(c->c.n <= 31 || (!steps_out && !ends_out),
         "32-bit arrays hold states of at most 31 bits: use ns_satostop_solve_wide");
    ns_satostop_opts o{};
    if (opts) o = *opts;
    const size_t starts = unknowns <= c->c.n && unknowns < 40 ? (size_t)1 << unknowns : 0;
    std::vector<uint64_t> st(steps_out ? starts : 0), en(ends_out ? starts : 0);
    ns::satostop_solve(*ctx->ctx.main, c->c, unknowns, o, *res, steps_out ? st.data() : nullptr,
                       ends_out ? en.data() : nullptr);
    // n <= 31: steps <= 2^31 and states < 2^31 fit 32 bits
    for (size_t i = 0; i < st.size(); i++) steps_out[i] = (uint32_t)st[i];
    for (size_t i = 0; i < en.size(); i++) ends_out[i] = (uint32_t)en[i];
  });
}

int ns_satostop_solve_wide(ns_ctx* ctx, const ns_circuit* c, uint32_t unknowns, const ns_satostop_opts* opts,
                           ns_satostop_result* res, uint64_t* steps_out, uint64_t* ends_out) {
  return guard([&] {
    need(ctx && c && res, "null argument");
    ns_satostop_opts o{};
    if (opts) o = *opts;
    ns::satostop_solve(*ctx->ctx.main, c->c, unknowns, o, *res, steps_out, ends_out);
  });
}

int ns_satostop_hashmap_solve(ns_ctx* ctx, const ns_circuit* c, uint32_t unknowns, const ns_hashmap_opts* opts,
                              ns_hashmap_result* res) {
  return guard([&] {
    need(ctx && c && res, "null argument");
    ns_hashmap_opts o{};
    o.unknown_fill_bits = -1;
    if (opts) o = *opts;
    ns::satostop_hashmap_solve(*ctx->ctx.main, c->c, unknowns, o, *res);
  });
}

int ns_satostop_hashmap_solve_progress(ns_ctx* ctx, const ns_circuit* c, uint32_t unknowns,
                                       const ns_hashmap_opts* opts, ns_hashmap_progress_fn fn, void* user,
                                       ns_hashmap_result* res) {
  return guard([&] {
    need(ctx && c && res && fn, "null argument");
    ns_hashmap_opts o{};
    o.unknown_fill_bits = -1;
    if (opts) o = *opts;
    ns::satostop_hashmap_solve(*ctx->ctx.main, c->c, unknowns, o, *res, fn, user);
  });
}

int ns_hashmap_step(ns_ctx* ctx, uint32_t which, uint32_t state_len, uint32_t arg_bit_place, uint64_t arg,
                    uint32_t unknown_bits, uint32_t unknown_fill_bits, uint32_t max_predecessors, uint32_t test,
                    const uint32_t* outputs, uint64_t output_words, void* hm, void* hm_out, uint64_t hm_len,
                    uint32_t* fills, uint64_t fills_len, void* sol) {
  return guard([&] {
    need(ctx && hm && fills && sol && (which != 1 || hm_out) && (which == 1 || outputs), "null argument");
    ns::hashmap_step(*ctx->ctx.main, which, state_len, arg_bit_place, arg, unknown_bits, unknown_fill_bits,
                     max_predecessors, test != 0, outputs, output_words, hm, hm_out, hm_len, fills, fills_len, sol);
  });
}

int ns_table_build(ns_ctx* ctx, const ns_circuit* c, uint32_t* table_out, uint64_t table_len, double* kernel_ms) {
  return guard([&] {
    need(ctx && c, "null argument");
    double ms = ns::table_build(*ctx->ctx.main, c->c, table_out, table_len);
    if (kernel_ms) *kernel_ms = ms;
  });
}

namespace {
// request of the C ABI -> server request; `own` copies the circuit (online submit)
ns::ServeRequest to_request(const ns_request& q, bool own) {
  need(q.circuit != nullptr, "request without a circuit");
  need(q.kind <= NS_REQ_SATOSTOP, "bad request kind");
  ns::ServeRequest r;
  r.kind = q.kind;
  r.unknowns = q.unknowns;
  if (own)
    r.circuit = std::make_shared<const ns::Circuit>(q.circuit->c);
  else
    r.circuit = std::shared_ptr<const ns::Circuit>(&q.circuit->c, [](const ns::Circuit*) {});
  r.arrival_s = q.arrival_s;
  if (q.kind == NS_REQ_QSAT) {
    need(q.quants_all != nullptr, "QSAT request without quantifiers");
    r.quants.assign(q.quants_all, q.quants_all + r.circuit->n);
    for (auto& v : r.quants) v = v ? 1 : 0;
  }
  return r;
}
}  // namespace

int ns_serve(ns_ctx* ctx, const ns_request* reqs, uint32_t n, const ns_serve_opts* opts,
             ns_request_result* results, double* total_s) {
  return guard([&] {
    need(ctx && (n == 0 || (reqs && results)), "null argument");
    std::vector<ns::ServeRequest> rq;
    rq.reserve(n);
    for (uint32_t i = 0; i < n; i++) rq.push_back(to_request(reqs[i], false));
    ns_serve_opts o{};
    if (opts) o = *opts;
    std::vector<ns_request_result> out;
    ns::serve(ctx->ctx, rq, o, out, total_s);
    if (n) std::memcpy(results, out.data(), sizeof(ns_request_result) * n);
  });
}

struct ns_server {
  std::unique_ptr<ns::Server> srv;
};

int ns_serve_open(ns_ctx* ctx, const ns_serve_opts* opts, ns_server** out) {
  return guard([&] {
    need(ctx && out, "null argument");
    ns_serve_opts o{};
    if (opts) o = *opts;
    const uint64_t budget = o.hbm_budget_bytes ? o.hbm_budget_bytes : (4ull << 30);
    auto h = std::make_unique<ns_server>();
    h->srv = std::make_unique<ns::Server>(ctx->ctx, o, budget);
    *out = h.release();
  });
}

int ns_serve_submit(ns_server* srv, const ns_request* req, uint64_t* ticket) {
  return guard([&] {
    need(srv && req && ticket, "null argument");
    *ticket = srv->srv->submit(to_request(*req, true));
  });
}

int ns_serve_poll(ns_server* srv, uint64_t ticket, ns_request_result* res, int* done) {
  return guard([&] {
    need(srv && done, "null argument");
    *done = srv->srv->poll(ticket, res) ? 1 : 0;
  });
}

int ns_serve_wait(ns_server* srv, uint64_t ticket, double timeout_s, ns_request_result* res, int* done) {
  return guard([&] {
    need(srv && done, "null argument");
    *done = srv->srv->wait(ticket, timeout_s, res) ? 1 : 0;
  });
}

int ns_serve_close(ns_server* srv) {
  return guard([&] {
    if (!srv) return;
    std::unique_ptr<ns_server> h(srv);
    h->srv->close();
  });
}

int ns_batch_solve(ns_ctx* ctx, const ns_request* reqs, uint32_t n, ns_request_result* results, double* info) {
  return guard([&] {
    need(ctx && (n == 0 || (reqs && results)), "null argument");
    std::vector<ns::ServeRequest> rq;
    rq.reserve(n);
    for (uint32_t i = 0; i < n; i++) rq.push_back(to_request(reqs[i], false));
    std::vector<const ns::ServeRequest*> batchable, alone;
    std::vector<size_t> bidx;
    for (uint32_t i = 0; i < n; i++) {
      if (ns::Batcher::batchable(rq[i])) {
        batchable.push_back(&rq[i]);
        bidx.push_back(i);
      }
    }
    std::vector<ns_request_result> res(n);
    ns::BatchStats bs;
    if (!batchable.empty()) {
      std::vector<ns_request_result> br;
      ns::batch_solve(ctx->ctx, batchable, br, &bs);
      for (size_t k = 0; k < bidx.size(); k++) res[bidx[k]] = br[k];
    }
    // the others (SATOSTOP and circuits outside the batch kernels) one at a time
    for (uint32_t i = 0; i < n; i++) {
      if (ns::Batcher::batchable(rq[i])) continue;
      bs.fallbacks++;
      ns_request_result& r = res[i];
      const ns::ServeRequest& q = rq[i];
      if (q.kind == NS_REQ_SATOSTOP) {
        ns_satostop_opts to{};
        ns_satostop_result tr;
        ns::satostop_solve(*ctx->ctx.main, *q.circuit, q.unknowns, to, tr, nullptr, nullptr);
        r.v0 = tr.n_stopped;
        r.v1 = tr.lowest_u;
        r.v2 = tr.steps;
        r.v3 = tr.end;
        r.kernel_ms = tr.build_ms + tr.jump_ms + tr.walk_ms;
        r.compile_ms = tr.compile_ms;
      } else if (q.kind == NS_REQ_SAT) {
        ns_sat_opts so{};
        ns_sat_result sr;
        ns::sat_solve(*ctx->ctx.main, *q.circuit, so, sr);
        r.v0 = sr.count;
        r.v1 = sr.lowest;
        r.v2 = sr.found;
        r.kernel_ms = sr.kernel_ms;
        r.compile_ms = sr.compile_ms;
      } else {
        ns_qsat_opts qo{};
        ns_qsat_result qr;
        ns::qsat_solve(*ctx->ctx.main, *q.circuit, q.quants, qo, qr, nullptr, 0);
        r.v0 = qr.truth;
        r.v1 = qr.witness;
        r.v2 = qr.witness_count;
        r.v3 = qr.has_witness;
        r.kernel_ms = qr.kernel_ms;
        r.compile_ms = qr.compile_ms;
      }
      r.status = NS_OK;
    }
    if (n) std::memcpy(results, res.data(), sizeof(ns_request_result) * n);
    if (info) {
      const double v[10] = {(double)n, (double)bs.fallbacks, (double)bs.vgpr_class, (double)bs.code_bytes,
                            bs.gen_ms, bs.load_ms, bs.kernel_ms, bs.wall_ms, (double)bs.graph_runs,
                            (double)bs.launches};
      std::memcpy(info, v, sizeof(v));
    }
  });
}

int ns_native_body(const ns_circuit* c, uint32_t kind, uint32_t lane_shift, char* body, size_t cap,
                   size_t* needed, uint32_t* vgpr_class, uint64_t* code_bytes) {
  return guard([&] {
    need(c && kind <= 7, "bad argument");
    const ns::Circuit& cc = c->c;
    ns::NativeKernel nk = native_kernel(cc, kind, lane_shift);
    std::string text;
    if (kind >= 4) {
      std::vector<uint32_t> w;
      std::string err;
      if (!ns::encode_body(nk.body, w, &err)) throw ns::Error(NS_ERR_UNSUPPORTED, err);
      text = "; circuit body\n" + nk.body;
      if (code_bytes) *code_bytes = w.size() * 4;
    } else {
      std::vector<char> code = ns::assemble_native(nk.asm_text);
      // the body sits between the marker comment and the template's ";;#ASMEND"
      size_t a = nk.asm_text.find("; circuit body\n");
      size_t b = nk.asm_text.find(";;#ASMEND", a);
      text = nk.asm_text.substr(a, b - a);
      if (code_bytes) *code_bytes = code.size();
    }
    if (needed) *needed = text.size();
    if (vgpr_class) *vgpr_class = nk.vgpr_class;
    if (body && cap) {
      need(cap > text.size(), "body buffer too small");
      std::memcpy(body, text.c_str(), text.size() + 1);
    }
  });
}

int ns_native_encode(const char* body, size_t len, uint32_t* words, size_t cap, size_t* n_words) {
  return guard([&] {
    need(body != nullptr || len == 0, "bad argument");
    std::vector<uint32_t> w;
    std::string err;
    if (!ns::encode_body(std::string(body ? body : "", len), w, &err)) throw ns::Error(NS_ERR_UNSUPPORTED, err);
    if (n_words) *n_words = w.size();
    if (words && cap) {
      need(cap >= w.size(), "word buffer too small");
      std::memcpy(words, w.data(), w.size() * 4);
    }
  });
}

int ns_native_code(const ns_circuit* c, uint32_t kind, uint32_t lane_shift, void* code, size_t cap,
                   size_t* needed, uint32_t* patched, uint64_t* slot_offset) {
  return guard([&] {
    need(c && kind <= 7, "bad argument");
    ns::NativeKernel nk = native_kernel(c->c, kind, lane_shift);
    size_t off = 0;
    std::vector<char> obj;
    bool p = true;
    if (kind == 4 || kind == 5) {
      std::vector<std::vector<uint32_t>> bodies(1);
      std::string err;
      if (!ns::encode_body(nk.body, bodies[0], &err)) throw ns::Error(NS_ERR_UNSUPPORTED, err);
      auto tpl = ns::batch_template(ns::batch_bytes(bodies));
      if (!tpl) throw ns::Error(NS_ERR_UNSUPPORTED, "body exceeds the largest batch region");
      std::vector<uint64_t> bo;
      obj = ns::compose_batch(*tpl, bodies, bo);
      off = tpl->region + bo[0];
    } else {
      obj = ns::patch_native(nk.kernel_name, nk.body, &off);
      p = !obj.empty();
      if (!p) obj = ns::assemble_native(nk.asm_text);
    }
    if (needed) *needed = obj.size();
    if (patched) *patched = p ? 1u : 0u;
    if (slot_offset) *slot_offset = p ? off : 0;
    if (code && cap) {
      need(cap >= obj.size(), "code buffer too small");
      std::memcpy(code, obj.data(), obj.size());
    }
  });
}

int ns_satostop_pass_times(ns_ctx* ctx, double* pass_ms, uint32_t cap, uint32_t* written) {
  return guard([&] {
    need(ctx && written, "null argument");
    const auto& v = ns::last_pass_times(*ctx->ctx.main);
    uint32_t k = 0;
    for (; k < v.size() && k < cap; k++)
      if (pass_ms) pass_ms[k] = v[k];
    *written = k;
  });
}

int ns_greedy_partitioned_solve(ns_ctx* ctx, const ns_circuit* c, uint32_t unknowns, uint64_t partitions,
                                const char* file_prefix, ns_partitioned_result* res) {
  return guard([&] {
    need(ctx && c && res, "null argument");
    ns::greedy_partitioned_solve(*ctx->ctx.main, c->c, unknowns, partitions, file_prefix ? file_prefix : "", *res);
  });
}

}  // extern "C"

struct ns_file_image {
  ns::FileImage fi;
};

extern "C" {

int ns_file_image_create(uint32_t state_len, uint64_t partitions, const char* prefix, ns_file_image** out) {
  return guard([&] {
    need(out, "null argument");
    *out = new ns_file_image{ns::FileImage(state_len, partitions, prefix ? prefix : "")};
  });
}

int ns_file_image_save_chunk(ns_file_image* fi, const uint64_t* words, uint64_t nwords) {
  return guard([&] {
    need(fi && words, "null argument");
    fi->fi.save_chunk(words, nwords);
  });
}

int ns_file_image_load_partition(ns_file_image* fi, uint64_t part, uint64_t* words, uint64_t words_cap) {
  return guard([&] {
    need(fi && words, "null argument");
    need(words_cap >= fi->fi.partition_words(), "words array too small");
    fi->fi.load_partition(part, words);
  });
}

int ns_file_image_save_partition(ns_file_image* fi, uint64_t part, const uint64_t* words, uint64_t nwords) {
  return guard([&] {
    need(fi && words, "null argument");
    need(nwords == fi->fi.partition_words(), "a partition has partition_words words");
    fi->fi.save_partition(part, words);
  });
}

void ns_file_image_free(ns_file_image* fi) {
  try {
    delete fi;
  } catch (...) {
  }
}

int ns_pack_entries(ns_ctx* ctx, uint32_t width, const uint64_t* values, uint64_t count, uint64_t* words,
                    uint64_t words_cap) {
  return guard([&] {
    need(ctx && values && words, "null argument");
    need(width >= 1 && width <= 64, "width must be in 1..64");
    need(count > 0 && count % 64 == 0, "count must be a positive multiple of 64");
    const uint64_t nwords = count * width / 64;
    need(words_cap >= nwords, "words array too small");
    ns::Lane& l = *ctx->ctx.main;
    NS_HIP(hipSetDevice(l.device));
    std::vector<uint32_t> lo(count), hi(count);
    for (uint64_t i = 0; i < count; i++) {
      lo[i] = (uint32_t)values[i];
      hi[i] = (uint32_t)(values[i] >> 32);
    }
    ns::DevBuf dlo, dhi, dout;
    dlo.ensure(count * 4);
    dhi.ensure(count * 4);
    dout.ensure(nwords * 8);
    NS_HIP(hipMemcpyAsync(dlo.p, lo.data(), count * 4, hipMemcpyHostToDevice, l.stream));
    NS_HIP(hipMemcpyAsync(dhi.p, hi.data(), count * 4, hipMemcpyHostToDevice, l.stream));
    NS_HIP(ns::launch_pack(l.stream, dlo.as<unsigned int>(), width > 32 ? dhi.as<unsigned int>() : nullptr, 32, width,
                           count, dout.as<unsigned long long>()));
    NS_HIP(hipMemcpyAsync(words, dout.p, nwords * 8, hipMemcpyDeviceToHost, l.stream));
    NS_HIP(hipStreamSynchronize(l.stream));
  });
}

int ns_probe_gather(ns_ctx* ctx, uint32_t table_bits, uint32_t reads_per_lane, uint32_t lanes, uint32_t flags,
                    double* ms, uint64_t* reads) {
  return guard([&] {
    need(ctx && ms && reads, "null argument");
    need(table_bits >= 8 && table_bits <= 33, "table_bits must be in 8..33");
    need(lanes >= 256 && lanes % 256 == 0 && reads_per_lane > 0, "lanes: a positive multiple of 256");
    ns::Lane& l = *ctx->ctx.main;
    NS_HIP(hipSetDevice(l.device));
    ns::DevBuf table, out;
    table.ensure((1ull << table_bits) * 4);
    out.ensure((uint64_t)lanes * 4);
    NS_HIP(ns::launch_fill_u32(l.stream, table.as<unsigned int>(), 0x5a5a5a5au, 1ull << table_bits));
    NS_HIP(ns::launch_gather_probe(l.stream, table.as<unsigned int>(), table_bits, reads_per_lane, lanes,
                                   out.as<unsigned int>(), (flags & 1u) != 0));  // warm-up
    NS_HIP(hipEventRecord(l.ev0, l.stream));
    NS_HIP(ns::launch_gather_probe(l.stream, table.as<unsigned int>(), table_bits, reads_per_lane, lanes,
                                   out.as<unsigned int>(), (flags & 1u) != 0));
    NS_HIP(hipEventRecord(l.ev1, l.stream));
    NS_HIP(hipStreamSynchronize(l.stream));
    *ms = l.elapsed_ms();
    *reads = (uint64_t)lanes * reads_per_lane;
  });
}

}  // extern "C"
