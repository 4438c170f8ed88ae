/*
 * CPU oracle for the naivesat hot path. TEST INFRASTRUCTURE ONLY: it is the checker the
 * tests and the benchmark's CPU baseline use; nothing in the product links or calls it.
 *
 * Plain C restatement of the reference's algorithms (matszpk/naivesat 0.1.3):
 *   - circuit evaluation with gatesim semantics (And, Nor, Nimpl, Xor; negated outputs),
 *     used by the reference at src/main.rs:365,374 and src/qsat.rs:1291-1304;
 *   - SAT brute force over all 2^n assignments, input k = bit k of x
 *     (src/main.rs:362-371 scalar scan; src/main.rs:76-155 bit-sliced mapper run);
 *   - QuantReducer streaming reduction (src/qsat.rs:172-291) and a level-by-level QSAT
 *     reduction with input 0 as the most significant index bit (src/qsat.rs:1991-1998);
 *   - greedysatostop next-state table + pointer jumping + start-state check
 *     (src/greedysatostop.rs:142-302, 634-690) and the scalar SATOSTOP walk
 *     (src/satostop.rs:1721-1782).
 * Pinning of each function is described in the header of oracle.h.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* circuit evaluation                                                         */
/* ------------------------------------------------------------------------- */

static inline uint64_t gate64(uint8_t f, uint64_t a, uint64_t b) {
  switch (f) {
    case 0: return a & b;      /* And */
    case 1: return ~(a | b);   /* Nor */
    case 2: return a & ~b;     /* Nimpl */
    default: return a ^ b;     /* Xor */
  }
}

uint64_t or_eval_u64(const or_circuit* c, uint64_t x) {
  uint32_t nw = c->n + c->ngates;
  uint8_t* v = (uint8_t*)malloc(nw ? nw : 1);
  for (uint32_t k = 0; k < c->n; k++) v[k] = (uint8_t)((x >> k) & 1);
  for (uint32_t j = 0; j < c->ngates; j++)
    v[c->n + j] = (uint8_t)(gate64(c->func[j], v[c->in0[j]], v[c->in1[j]]) & 1);
  uint64_t r = 0;
  for (uint32_t i = 0; i < c->nout && i < 64; i++)
    r |= (uint64_t)((v[c->out_wire[i]] ^ (c->out_neg ? c->out_neg[i] : 0)) & 1) << i;
  free(v);
  return r;
}

/* gates needed by the outputs in `outs` (all when outs == NULL), in wire order */
static uint32_t* cone_list(const or_circuit* c, const uint32_t* outs, uint32_t nouts, uint32_t* count) {
  uint32_t nw = c->n + c->ngates;
  uint8_t* need = (uint8_t*)calloc(nw ? nw : 1, 1);
  for (uint32_t i = 0; i < (outs ? nouts : c->nout); i++) need[c->out_wire[outs ? outs[i] : i]] = 1;
  for (int64_t w = (int64_t)nw - 1; w >= (int64_t)c->n; w--)
    if (need[w]) {
      need[c->in0[w - c->n]] = 1;
      need[c->in1[w - c->n]] = 1;
    }
  uint32_t cnt = 0;
  for (uint32_t w = c->n; w < nw; w++) cnt += need[w];
  uint32_t* list = (uint32_t*)malloc(sizeof(uint32_t) * (cnt ? cnt : 1));
  cnt = 0;
  for (uint32_t w = c->n; w < nw; w++)
    if (need[w]) list[cnt++] = w;
  free(need);
  *count = cnt;
  return list;
}

/* 64 assignments per word: assignment index v = 64 * w + b. vbit[k] = position of
 * input k in v. Fills out[i] with the selected outputs' words. `val` has room for all
 * wires. */
static const uint64_t kMask64[6] = {0xAAAAAAAAAAAAAAAAull, 0xCCCCCCCCCCCCCCCCull, 0xF0F0F0F0F0F0F0F0ull,
                                    0xFF00FF00FF00FF00ull, 0xFFFF0000FFFF0000ull, 0xFFFFFFFF00000000ull};

static inline void eval_word(const or_circuit* c, const uint32_t* cone, uint32_t ncone,
                             const uint32_t* vbit, uint64_t w, uint64_t* val) {
  for (uint32_t k = 0; k < c->n; k++) {
    uint32_t b = vbit[k];
    val[k] = b < 6 ? kMask64[b] : (((w >> (b - 6)) & 1) ? ~0ull : 0ull);
  }
  for (uint32_t i = 0; i < ncone; i++) {
    uint32_t wi = cone[i], j = wi - c->n;
    val[wi] = gate64(c->func[j], val[c->in0[j]], val[c->in1[j]]);
  }
}

static inline uint64_t out_word(const or_circuit* c, uint32_t i, const uint64_t* val) {
  return val[c->out_wire[i]] ^ ((c->out_neg && c->out_neg[i]) ? ~0ull : 0ull);
}

static inline uint64_t valid_mask(uint32_t nbits) {
  return nbits >= 6 ? ~0ull : ((1ull << (1u << nbits)) - 1ull);
}

/* ------------------------------------------------------------------------- */
/* SAT                                                                        */
/* ------------------------------------------------------------------------- */

typedef struct {
  const or_circuit* c;
  const uint32_t* cone;
  uint32_t ncone;
  uint64_t w_lo, w_hi; /* word range */
  uint64_t count, lowest;
} sat_job;

/* 8 consecutive words at once (the inner loops vectorise): val[wire * 8 + i]; cloned for
   AVX-512 (one 512-bit op per gate) and the AVX2 build default, picked at load time by the
   CPU's features */
#define OR_BW 8
__attribute__((target_clones("avx512f", "default"))) static void eval_block8(const or_circuit* c, const uint32_t* cone, uint32_t ncone, uint64_t w0, uint64_t* val) {
  for (uint32_t k = 0; k < c->n; k++)
    for (int i = 0; i < OR_BW; i++)
      val[k * OR_BW + i] = k < 6 ? kMask64[k] : ((((w0 + (uint64_t)i) >> (k - 6)) & 1) ? ~0ull : 0ull);
  for (uint32_t g = 0; g < ncone; g++) {
    uint32_t wi = cone[g], j = wi - c->n;
    uint64_t* r = val + (size_t)wi * OR_BW;
    const uint64_t* a = val + (size_t)c->in0[j] * OR_BW;
    const uint64_t* b = val + (size_t)c->in1[j] * OR_BW;
    switch (c->func[j]) {
      case 0: for (int i = 0; i < OR_BW; i++) r[i] = a[i] & b[i]; break;
      case 1: for (int i = 0; i < OR_BW; i++) r[i] = ~(a[i] | b[i]); break;
      case 2: for (int i = 0; i < OR_BW; i++) r[i] = a[i] & ~b[i]; break;
      default: for (int i = 0; i < OR_BW; i++) r[i] = a[i] ^ b[i]; break;
    }
  }
}

static void* sat_worker(void* arg) {
  sat_job* jb = (sat_job*)arg;
  const or_circuit* c = jb->c;
  uint64_t* val = (uint64_t*)malloc(sizeof(uint64_t) * OR_BW * (c->n + c->ngates + 1));
  uint32_t* vbit = (uint32_t*)malloc(sizeof(uint32_t) * (c->n + 1));
  for (uint32_t k = 0; k < c->n; k++) vbit[k] = k;
  uint64_t vm = valid_mask(c->n);
  const uint32_t ow = c->out_wire[0];
  const uint64_t on = (c->out_neg && c->out_neg[0]) ? ~0ull : 0ull;
  jb->count = 0;
  jb->lowest = ~0ull;
  uint64_t w = jb->w_lo;
  for (; w + OR_BW <= jb->w_hi; w += OR_BW) {
    eval_block8(c, jb->cone, jb->ncone, w, val);
    for (int i = 0; i < OR_BW; i++) {
      uint64_t o = (val[(size_t)ow * OR_BW + i] ^ on) & vm;
      if (o) {
        jb->count += (uint64_t)__builtin_popcountll(o);
        if (jb->lowest == ~0ull) jb->lowest = ((w + (uint64_t)i) << 6) | (uint64_t)__builtin_ctzll(o);
      }
    }
  }
  for (; w < jb->w_hi; w++) {
    eval_word(c, jb->cone, jb->ncone, vbit, w, val);
    uint64_t o = out_word(c, 0, val) & vm;
    if (o) {
      jb->count += (uint64_t)__builtin_popcountll(o);
      if (jb->lowest == ~0ull) jb->lowest = (w << 6) | (uint64_t)__builtin_ctzll(o);
    }
  }
  free(val);
  free(vbit);
  return NULL;
}

const char* or_simd_isa(void) {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512f") ? "avx512f" : "avx2";
}

int or_sat_count(const or_circuit* c, int threads, uint64_t* count, uint64_t* lowest) {
  if (c->nout < 1 || c->n == 0 || c->n > 62) return -1;
  uint64_t words = c->n >= 6 ? (1ull << (c->n - 6)) : 1ull;
  return or_sat_count_range(c, 0, words, threads, count, lowest);
}

int or_sat_count_range(const or_circuit* c, uint64_t w_lo, uint64_t w_hi, int threads, uint64_t* count,
                       uint64_t* lowest) {
  if (threads < 1) threads = 1;
  uint32_t ncone;
  uint32_t out0 = 0;
  uint32_t* cone = cone_list(c, &out0, 1, &ncone);
  sat_job* jobs = (sat_job*)calloc((size_t)threads, sizeof(sat_job));
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  uint64_t span = w_hi - w_lo;
  for (int t = 0; t < threads; t++) {
    jobs[t].c = c;
    jobs[t].cone = cone;
    jobs[t].ncone = ncone;
    jobs[t].w_lo = w_lo + span * (uint64_t)t / (uint64_t)threads;
    jobs[t].w_hi = w_lo + span * (uint64_t)(t + 1) / (uint64_t)threads;
    pthread_create(&th[t], NULL, sat_worker, &jobs[t]);
  }
  uint64_t cnt = 0, low = ~0ull;
  for (int t = 0; t < threads; t++) {
    pthread_join(th[t], NULL);
    cnt += jobs[t].count;
    if (jobs[t].lowest < low) low = jobs[t].lowest;
  }
  free(jobs);
  free(th);
  free(cone);
  *count = cnt;
  *lowest = low;
  return 0;
}

/* ------------------------------------------------------------------------- */
/* QuantReducer (src/qsat.rs:172-291)                                         */
/* ------------------------------------------------------------------------- */

struct or_qr {
  uint32_t len;
  uint8_t* quants; /* reversed list: quants[0] = innermost; 1 = All */
  uint8_t* result;
  int started;
  uint64_t start, all_mask, first_mask, other_mask;
  int has_solution;
  uint64_t solution;
  /* out-of-order items: min-heap on index */
  uint64_t* heap_idx;
  uint8_t* heap_val;
  size_t heap_n, heap_cap;
};

or_qr* or_qr_new(const uint8_t* quants_all, uint32_t len) {
  if (len == 0 || len > 63) return NULL;
  or_qr* q = (or_qr*)calloc(1, sizeof(or_qr));
  q->len = len;
  uint32_t fb = 0;
  while (fb < len && quants_all[fb] == quants_all[0]) fb++;
  q->quants = (uint8_t*)malloc(len);
  q->result = (uint8_t*)malloc(len);
  for (uint32_t i = 0; i < len; i++) q->quants[i] = quants_all[len - 1 - i] ? 1 : 0;
  memcpy(q->result, q->quants, len);
  q->all_mask = (1ull << len) - 1;
  q->first_mask = ((1ull << fb) - 1) << (len - fb);
  q->other_mask = q->all_mask & ~q->first_mask;
  return q;
}

void or_qr_free(or_qr* q) {
  if (!q) return;
  free(q->quants);
  free(q->result);
  free(q->heap_idx);
  free(q->heap_val);
  free(q);
}

static int qr_is_end(const or_qr* q) { return q->started && (q->start & q->all_mask) == 0; }

static void heap_push(or_qr* q, uint64_t idx, uint8_t v) {
  if (q->heap_n == q->heap_cap) {
    q->heap_cap = q->heap_cap ? q->heap_cap * 2 : 64;
    q->heap_idx = (uint64_t*)realloc(q->heap_idx, q->heap_cap * sizeof(uint64_t));
    q->heap_val = (uint8_t*)realloc(q->heap_val, q->heap_cap);
  }
  size_t i = q->heap_n++;
  while (i > 0) {
    size_t p = (i - 1) / 2;
    if (q->heap_idx[p] <= idx) break;
    q->heap_idx[i] = q->heap_idx[p];
    q->heap_val[i] = q->heap_val[p];
    i = p;
  }
  q->heap_idx[i] = idx;
  q->heap_val[i] = v;
}

static void heap_pop(or_qr* q) {
  size_t n = --q->heap_n;
  uint64_t idx = q->heap_idx[n];
  uint8_t v = q->heap_val[n];
  size_t i = 0;
  while (1) {
    size_t l = 2 * i + 1, r = l + 1, m = i;
    uint64_t mi = idx;
    if (l < n && q->heap_idx[l] < mi) {
      m = l;
      mi = q->heap_idx[l];
    }
    if (r < n && q->heap_idx[r] < mi) m = r;
    if (m == i) break;
    q->heap_idx[i] = q->heap_idx[m];
    q->heap_val[i] = q->heap_val[m];
    i = m;
  }
  if (n) {
    q->heap_idx[i] = idx;
    q->heap_val[i] = v;
  }
}

static uint64_t rev_bits(uint64_t x, uint32_t bits) {
  uint64_t r = 0;
  for (uint32_t i = 0; i < bits; i++) r |= ((x >> i) & 1ull) << (bits - 1 - i);
  return r;
}

static void qr_apply(or_qr* q, uint8_t item) {
  q->started = 1;
  uint64_t index = q->start;
  uint8_t prev = item;
  for (uint32_t i = 0; i < q->len; i++) {
    uint8_t r = q->result[i], qq = q->quants[i];
    /* All: result & item, Exists: result | item */
    prev = (uint8_t)((r & prev) | ((r ^ prev) & (uint8_t)!qq));
    q->result[i] = prev;
    if ((index & 1) == 0) break;
    index >>= 1;
    q->result[i] = qq;
  }
  if (!q->has_solution && (q->start & q->other_mask) == q->other_mask && (prev ^ q->quants[q->len - 1])) {
    uint32_t fb = (uint32_t)__builtin_popcountll(q->first_mask);
    q->has_solution = 1;
    q->solution = rev_bits(q->start >> (q->len - fb), fb);
  }
  q->start = q->start + 1; /* wrapping, as the reference's overflowing_add */
}

int or_qr_push(or_qr* q, uint64_t index, int item) {
  if (qr_is_end(q)) return -1;
  heap_push(q, index, (uint8_t)(item != 0));
  while (q->heap_n && !qr_is_end(q)) {
    if (q->heap_idx[0] != q->start) break;
    uint8_t v = q->heap_val[0];
    heap_pop(q);
    int had = q->has_solution;
    qr_apply(q, v);
    if (!had && q->has_solution) break; /* the reference stops flushing at the first solution */
  }
  return 0;
}

int or_qr_final(const or_qr* q, int* reversed, uint32_t* solution_bits, int* has_solution, uint64_t* solution) {
  if (!(qr_is_end(q) || q->has_solution)) return 0;
  *reversed = q->quants[q->len - 1];
  *solution_bits = (uint32_t)__builtin_popcountll(q->first_mask);
  *has_solution = q->has_solution;
  *solution = q->has_solution ? q->solution : 0;
  return 1;
}

/* ------------------------------------------------------------------------- */
/* QSAT by level reduction                                                     */
/* ------------------------------------------------------------------------- */

/* values: one byte per index of a 2^m array; reduce levels [lvl, lvl+k): level l pairs
 * the indices differing in bit (l - lvl) of the array index; op from quants (input n-1-l). */
static void reduce_bytes(uint8_t* v, uint32_t m, uint32_t lvl, uint32_t k, const uint8_t* q, uint32_t n) {
  for (uint32_t i = 0; i < k; i++) {
    uint8_t all = q[n - 1 - (lvl + i)];
    uint64_t cnt = 1ull << (m - i - 1);
    for (uint64_t t = 0; t < cnt; t++) {
      uint8_t a = v[2 * t], b = v[2 * t + 1];
      v[t] = all ? (a & b) : (a | b);
    }
  }
}

typedef struct {
  const or_circuit* c;
  const uint8_t* q;
  const uint32_t* cone;
  uint32_t ncone;
  uint32_t C;           /* log2 of assignments per chunk */
  uint32_t L;           /* levels to reduce below the outermost block */
  uint64_t chunk_lo, chunk_hi;
  uint8_t* out;         /* per-chunk results: 2^(C-min(L,C)) bytes per chunk */
} qsat_job;

static void* qsat_worker(void* arg) {
  qsat_job* jb = (qsat_job*)arg;
  const or_circuit* c = jb->c;
  uint32_t n = c->n;
  uint64_t* val = (uint64_t*)malloc(sizeof(uint64_t) * (c->n + c->ngates + 1));
  uint32_t* vbit = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
  for (uint32_t k = 0; k < n; k++) vbit[k] = n - 1 - k;
  uint64_t csize = 1ull << jb->C;
  uint8_t* bytes = (uint8_t*)malloc(csize);
  uint32_t red = jb->L < jb->C ? jb->L : jb->C;
  uint64_t per_chunk = 1ull << (jb->C - red);
  for (uint64_t ch = jb->chunk_lo; ch < jb->chunk_hi; ch++) {
    uint64_t v0 = ch << jb->C;
    uint64_t words = csize >= 64 ? csize / 64 : 1;
    for (uint64_t wi = 0; wi < words; wi++) {
      uint64_t w = (v0 >> 6) + wi;
      eval_word(c, jb->cone, jb->ncone, vbit, w, val);
      uint64_t o = out_word(c, 0, val);
      uint64_t lim = csize < 64 ? csize : 64;
      for (uint64_t b = 0; b < lim; b++) bytes[wi * 64 + b] = (uint8_t)((o >> b) & 1);
    }
    reduce_bytes(bytes, jb->C, 0, red, jb->q, n);
    memcpy(jb->out + ch * per_chunk, bytes, per_chunk);
  }
  free(bytes);
  free(val);
  free(vbit);
  return NULL;
}

int or_qsat(const or_circuit* c, const uint8_t* quants_all, int threads, uint32_t* truth, uint64_t* wcount,
            uint64_t* lowest, uint8_t* bitmap_bytes) {
  uint32_t n = c->n;
  if (n == 0 || n > 40 || c->nout < 1) return -1;
  if (threads < 1) threads = 1;
  uint8_t* q = (uint8_t*)malloc(n);
  for (uint32_t k = 0; k < n; k++) q[k] = quants_all[k] ? 1 : 0;
  uint32_t fb = 0;
  while (fb < n && q[fb] == q[0]) fb++;
  uint32_t L = n - fb;
  uint32_t C = n < 20 ? n : 20;
  uint64_t chunks = 1ull << (n - C);
  uint32_t red = L < C ? L : C;
  uint64_t per_chunk = 1ull << (C - red);
  uint8_t* arr = (uint8_t*)malloc(chunks * per_chunk);
  uint32_t ncone;
  uint32_t out0 = 0;
  uint32_t* cone = cone_list(c, &out0, 1, &ncone);
  qsat_job* jobs = (qsat_job*)calloc((size_t)threads, sizeof(qsat_job));
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 0; t < threads; t++) {
    jobs[t] = (qsat_job){c, q, cone, ncone, C, L, chunks * (uint64_t)t / (uint64_t)threads,
                         chunks * (uint64_t)(t + 1) / (uint64_t)threads, arr};
    pthread_create(&th[t], NULL, qsat_worker, &jobs[t]);
  }
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  /* remaining levels above the chunk */
  uint32_t m = (uint32_t)(n - red);
  if (L > red) {
    reduce_bytes(arr, m, red, L - red, q, n);
    m -= (L - red);
  }
  /* arr now holds 2^fb outermost-block values */
  uint64_t cnt = 0, low = ~0ull;
  uint64_t nu = 1ull << m;
  for (uint64_t u = 0; u < nu; u++) {
    uint8_t wit = q[0] ? (uint8_t)!arr[u] : arr[u];
    if (wit) {
      cnt++;
      if (low == ~0ull) low = u;
    }
    if (bitmap_bytes) {
      if (arr[u]) bitmap_bytes[u >> 3] |= (uint8_t)(1u << (u & 7));
      else bitmap_bytes[u >> 3] &= (uint8_t)~(1u << (u & 7));
    }
  }
  *wcount = cnt;
  *lowest = low;
  *truth = (cnt != 0) ^ q[0];
  free(jobs);
  free(th);
  free(cone);
  free(arr);
  free(q);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* SATOSTOP                                                                   */
/* ------------------------------------------------------------------------- */

uint64_t or_satostop_walk(const or_circuit* c, uint64_t start, uint64_t max_steps, uint64_t* end) {
  uint32_t n = c->n;
  uint64_t s = start;
  for (uint64_t step = 1; step <= max_steps; step++) {
    uint64_t o = or_eval_u64(c, s);
    uint64_t next = o & ((n >= 64) ? ~0ull : ((1ull << n) - 1));
    if ((o >> n) & 1) {
      *end = next;
      return step;
    }
    s = next;
  }
  *end = 0;
  return 0;
}

typedef struct {
  const or_circuit* c;
  const uint32_t* cone;
  uint32_t ncone;
  uint64_t w_lo, w_hi;
  uint64_t* tab;
} table_job;

static void* table_worker(void* arg) {
  table_job* jb = (table_job*)arg;
  const or_circuit* c = jb->c;
  uint32_t n = c->n;
  uint64_t* val = (uint64_t*)malloc(sizeof(uint64_t) * (c->n + c->ngates + 1));
  uint32_t* vbit = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
  for (uint32_t k = 0; k < n; k++) vbit[k] = k;
  uint64_t states = 1ull << n;
  for (uint64_t w = jb->w_lo; w < jb->w_hi; w++) {
    eval_word(c, jb->cone, jb->ncone, vbit, w, val);
    uint64_t lim = states < 64 ? states : 64;
    for (uint64_t b = 0; b < lim; b++) {
      uint64_t e = 0;
      for (uint32_t i = 0; i <= n; i++) e |= ((out_word(c, i, val) >> b) & 1ull) << i;
      jb->tab[w * 64 + b] = e | (1ull << 32); /* steps = 1 */
    }
  }
  free(val);
  free(vbit);
  return NULL;
}

typedef struct {
  const uint64_t* src;
  uint64_t* dst;
  uint64_t lo, hi;
  uint32_t n;
} jump_job;

static void* jump_worker(void* arg) {
  jump_job* jb = (jump_job*)arg;
  uint64_t stop = 1ull << jb->n, mask = stop - 1;
  for (uint64_t i = jb->lo; i < jb->hi; i++) {
    uint64_t e = jb->src[i];
    if (!(e & stop)) {
      uint64_t f = jb->src[e & mask];
      e = (f & 0xffffffffull) | (((e >> 32) + (f >> 32)) << 32);
    }
    jb->dst[i] = e;
  }
  return NULL;
}

int or_satostop_table(const or_circuit* c, uint32_t unknowns, int threads, uint32_t* steps_out,
                      uint32_t* ends_out, uint64_t* n_stopped, uint64_t* lowest_u) {
  uint32_t n = c->n;
  if (n == 0 || n > 31 || c->nout != n + 1 || unknowns > n) return -1;
  if (threads < 1) threads = 1;
  uint64_t states = 1ull << n;
  uint64_t* a = (uint64_t*)malloc(states * sizeof(uint64_t));
  uint64_t* b = (uint64_t*)malloc(states * sizeof(uint64_t));
  if (!a || !b) {
    free(a);
    free(b);
    return -2;
  }
  uint32_t ncone;
  uint32_t* cone = cone_list(c, NULL, 0, &ncone);
  uint64_t words = n >= 6 ? (1ull << (n - 6)) : 1ull;
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  table_job* tj = (table_job*)calloc((size_t)threads, sizeof(table_job));
  for (int t = 0; t < threads; t++) {
    tj[t] = (table_job){c, cone, ncone, words * (uint64_t)t / (uint64_t)threads,
                        words * (uint64_t)(t + 1) / (uint64_t)threads, a};
    pthread_create(&th[t], NULL, table_worker, &tj[t]);
  }
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  /* n doubling passes: afterwards each entry covers 2^n evaluations or ends at STOP */
  jump_job* jj = (jump_job*)calloc((size_t)threads, sizeof(jump_job));
  for (uint32_t p = 0; p < n; p++) {
    for (int t = 0; t < threads; t++) {
      jj[t] = (jump_job){a, b, states * (uint64_t)t / (uint64_t)threads,
                         states * (uint64_t)(t + 1) / (uint64_t)threads, n};
      pthread_create(&th[t], NULL, jump_worker, &jj[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    uint64_t* tmp = a;
    a = b;
    b = tmp;
  }
  uint64_t starts = 1ull << unknowns, cnt = 0, low = ~0ull;
  uint64_t stop = 1ull << n, mask = stop - 1;
  for (uint64_t u = 0; u < starts; u++) {
    uint64_t e = a[u << (n - unknowns)];
    int st = (e & stop) != 0;
    if (steps_out) steps_out[u] = st ? (uint32_t)(e >> 32) : 0u;
    if (ends_out) ends_out[u] = st ? (uint32_t)(e & mask) : 0u;
    if (st) {
      cnt++;
      if (low == ~0ull) low = u;
    }
  }
  *n_stopped = cnt;
  *lowest_u = low;
  free(a);
  free(b);
  free(cone);
  free(th);
  free(tj);
  free(jj);
  return 0;
}

/* reference-style u32 table: entry s = sum_i output_i(s) << i (at most 32 outputs) */
int or_table_u32(const or_circuit* c, uint32_t* table) {
  uint32_t n = c->n;
  if (n == 0 || n > 26 || c->nout == 0 || c->nout > 32) return -1;
  for (uint64_t s = 0; s < (1ull << n); s++) table[s] = (uint32_t)or_eval_u64(c, s);
  return 0;
}

/* ------------------------------------------------------------------------------------
 * naivesatostop hash map (src/satostop.rs:127-1150), restated sequentially: entries are
 * processed in index order, so where the reference's parallel loops race (which of two
 * starts becomes `solution` first, the order of unknown-fill increments), this is one of
 * the orders the reference may take.
 * ---------------------------------------------------------------------------------- */

/* hash_function_64, src/satostop.rs:127-136 (u64 wrapping arithmetic) */
uint64_t or_hash64(uint32_t bits, uint64_t value) {
  const uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
  const uint32_t half = bits >> 1;
  uint64_t temp = value * 9615409803190489167ull;
  temp = (temp << half) | (bits - half >= 64 ? 0 : (temp >> (bits - half)));
  const uint64_t hash = (value * 6171710485021949031ull) ^ temp ^ 0xb89d2ecda078ca1full;
  return hash & mask;
}

enum { HS_UNUSED = 0, HS_USED = 1, HS_STOPPED = 2, HS_LOOPED = 3, HS_RESERVED = 4 };

static uint64_t output_at(const uint32_t* outputs, uint32_t state_len, uint64_t i) {
  const uint32_t wpe = (state_len + 1 + 31) >> 5;
  return wpe == 2 ? ((uint64_t)outputs[2 * i] | ((uint64_t)outputs[2 * i + 1] << 32)) : (uint64_t)outputs[i];
}

/* join_to_hashmap_cpu, src/satostop.rs:186-235 */
void or_hm_join_to(uint32_t state_len, uint32_t arg_bit_place, uint64_t arg, const uint32_t* outputs,
                   or_hash_entry* hm, uint64_t len) {
  const uint64_t arg_start = arg << arg_bit_place, arg_end = arg_start + (1ull << arg_bit_place);
  const uint64_t state_mask = (1ull << state_len) - 1ull;
  for (uint64_t i = 0; i < len; i++) {
    or_hash_entry* he = &hm[i];
    if (he->state == HS_USED && arg_start <= he->next && he->next < arg_end) {
      const uint64_t output = output_at(outputs, state_len, he->next - arg_start);
      const uint64_t old_next = he->next;
      he->next = output & state_mask;
      if ((output >> state_len) & 1ull)
        he->state = HS_STOPPED;
      else if (state_mask <= he->steps || he->next == he->current || he->next == old_next)
        he->state = HS_LOOPED;
      else
        he->state = HS_USED;
      he->steps += 1;
    }
  }
}

/* resolve_unknowns, src/satostop.rs:330-381 */
static void resolve_unknowns(uint32_t state_len, uint32_t unknown_bits, uint32_t unknown_fill_bits, uint64_t current,
                             uint64_t next, uint64_t steps, uint32_t entry_state, uint32_t* unknown_fills,
                             uint64_t* resolved, or_hm_solution* sol) {
  if ((current & ((1ull << (state_len - unknown_bits)) - 1ull)) == 0 &&
      (entry_state == HS_LOOPED || entry_state == HS_STOPPED)) {
    if (entry_state == HS_STOPPED && !sol->defined) {
      sol->defined = 1;
      sol->start = current;
      sol->end = next;
      sol->steps = steps;
    }
    const uint64_t idx = current >> (state_len - unknown_bits + unknown_fill_bits);
    const uint32_t fill_mask = (1u << unknown_fill_bits) - 1u;
    const uint32_t value = (uint32_t)((current >> (state_len - unknown_bits)) & fill_mask);
    if (unknown_fills[idx] == value) {
      unknown_fills[idx] = value + 1;
      if (value == fill_mask) (*resolved)++;
    }
  }
}

/* join_hashmap_itself_and_check_solution, the device kernel of src/satostop.rs:540-594
 * (its CPU twin, 391-479, bounds the summed steps by 2^(state_len-1) - 1 instead of
 * 2^state_len - 1; the reference's data set gives the same result under both) */
void or_hm_join_itself(uint32_t state_len, const or_hash_entry* in, or_hash_entry* out, uint64_t len,
                       uint32_t unknown_bits, uint32_t unknown_fill_bits, uint32_t* unknown_fills,
                       uint64_t* resolved, or_hm_solution* sol) {
  uint32_t hbits = 0;
  while ((1ull << hbits) < len) hbits++;
  const uint32_t shift = state_len - hbits;
  const uint64_t state_mask = state_len >= 64 ? ~0ull : (1ull << state_len) - 1ull;
  uint32_t* preds = (uint32_t*)calloc(len, sizeof(uint32_t));
  for (uint64_t i = 0; i < len; i++) {
    const or_hash_entry* inhe = &in[i];
    or_hash_entry* outhe = &out[i];
    if (inhe->state == HS_USED) {
      const uint64_t nh = or_hash64(state_len, inhe->next) >> shift;
      const or_hash_entry* nexthe = &in[nh];
      if (nexthe->state != HS_UNUSED && nexthe->current == inhe->next) {
        outhe->current = inhe->current;
        outhe->next = nexthe->next;
        const uint64_t res = inhe->steps + nexthe->steps;
        const int ov = res < inhe->steps;
        outhe->steps = res;
        outhe->state = nexthe->state;
        if (outhe->state == HS_USED && (inhe->current == nexthe->next || ov || res > state_mask))
          outhe->state = HS_LOOPED;
        outhe->predecessors = inhe->predecessors;
        preds[nh]++;
      } else {
        *outhe = *inhe;
      }
    } else {
      *outhe = *inhe;
    }
    resolve_unknowns(state_len, unknown_bits, unknown_fill_bits, outhe->current, outhe->next, outhe->steps,
                     outhe->state, unknown_fills, resolved, sol);
  }
  for (uint64_t i = 0; i < len; i++) out[i].predecessors += preds[i];
  free(preds);
}

/* add_to_hashmap_and_check_solution_cpu, src/satostop.rs:713-860 (test: skip elements
 * whose next state is 0, as the reference's test routine does) */
void or_hm_add(uint32_t state_len, uint32_t arg_bit_place, uint64_t arg, const uint32_t* outputs,
               or_hash_entry* hm, uint64_t len, uint32_t unknown_bits, uint32_t unknown_fill_bits,
               uint32_t* unknown_fills, uint64_t* resolved, or_hm_solution* sol, uint32_t max_predecessors,
               int test) {
  uint32_t hbits = 0;
  while ((1ull << hbits) < len) hbits++;
  const uint32_t shift = state_len - hbits;
  const uint64_t arg_start = arg << arg_bit_place;
  const uint64_t state_mask = (1ull << state_len) - 1ull;
  const uint64_t fill_mask = (1ull << unknown_fill_bits) - 1ull;
  const uint64_t before_unknowns = (1ull << (state_len - unknown_bits)) - 1ull;
  const uint32_t fill_shift = state_len - unknown_bits + unknown_fill_bits;
  for (uint64_t i = 0; i < (1ull << arg_bit_place); i++) {
    const uint64_t output = output_at(outputs, state_len, i);
    const uint64_t current = i + arg_start;
    const uint64_t next = output & state_mask;
    const uint32_t state = ((output >> state_len) & 1ull) ? HS_STOPPED : (next == current ? HS_LOOPED : HS_USED);
    if (test && next == 0) continue;
    or_hash_entry* cur = &hm[or_hash64(state_len, current) >> shift];
    const int cur_solved = (current & before_unknowns) == 0 &&
                           unknown_fills[current >> fill_shift] ==
                               (uint32_t)((current >> (state_len - unknown_bits)) & fill_mask);
    const int old_solved = cur->state != HS_UNUSED && (cur->current & before_unknowns) == 0 &&
                           unknown_fills[cur->current >> fill_shift] ==
                               (uint32_t)((cur->current >> (state_len - unknown_bits)) & fill_mask);
    /* one thread at a time: the reservation flag is always free, so the first trial decides */
    if ((cur_solved && !old_solved) || (!old_solved && cur->predecessors <= max_predecessors)) {
      cur->current = current;
      cur->next = next;
      cur->steps = 1;
      cur->predecessors = 0;
      cur->state = state;
    }
    resolve_unknowns(state_len, unknown_bits, unknown_fill_bits, current, next, 1, state, unknown_fills, resolved,
                     sol);
  }
}

/* ---- MemImage: packed (state_len + 1)-bit entries (src/greedysatostop.rs:339-475) ---- */
static uint64_t mem_mask(uint32_t width) { return width >= 64 ? ~0ull : ((1ull << width) - 1ull); }

/* MemImage::get, src/greedysatostop.rs:373-382 */
uint64_t or_mem_get(const uint64_t* data, uint32_t width, uint64_t i) {
  const uint64_t mask = mem_mask(width);
  const uint64_t idx = (i * width) >> 6;
  const uint32_t bit = (uint32_t)((i * width) & 63);
  if (64 - bit >= width) return (data[idx] >> bit) & mask;
  return ((data[idx] >> bit) | (data[idx + 1] << (64 - bit))) & mask;
}

/* MemImage::set, src/greedysatostop.rs:384-394 */
void or_mem_set(uint64_t* data, uint32_t width, uint64_t i, uint64_t value) {
  const uint64_t mask = mem_mask(width);
  const uint64_t idx = (i * width) >> 6;
  const uint32_t bit = (uint32_t)((i * width) & 63);
  value &= mask;
  data[idx] = (data[idx] & ~(mask << bit)) | (value << bit);
  if (64 - bit < width) data[idx + 1] = (data[idx + 1] & ~(mask >> (64 - bit))) | (value >> (64 - bit));
}

void or_mem_pack(const uint64_t* values, uint64_t count, uint32_t width, uint64_t* words) {
  for (uint64_t i = 0; i < count; i++) or_mem_set(words, width, i, values[i]);
}

void or_mem_unpack(const uint64_t* words, uint64_t count, uint32_t width, uint64_t* values) {
  for (uint64_t i = 0; i < count; i++) values[i] = or_mem_get(words, width, i);
}

/* MemImage::join_nexts, src/greedysatostop.rs:419-444 (the chunks there only split the
 * index range among threads; each entry's update reads `second` alone) */
void or_mem_join_nexts(uint64_t* dest, uint64_t dest_start, uint64_t len, const uint64_t* second,
                       uint64_t second_start, uint64_t second_len, uint32_t state_len) {
  (void)dest_start;
  const uint32_t width = state_len + 1;
  const uint64_t state_mask = mem_mask(width) >> 1, stop_mask = 1ull << state_len;
  const uint64_t end = second_start + second_len;
  for (uint64_t i = 0; i < len; i++) {
    const uint64_t old = or_mem_get(dest, width, i);
    const uint64_t old_next = old & state_mask;
    if ((old & stop_mask) == 0 && second_start <= old_next && old_next < end)
      or_mem_set(dest, width, i, or_mem_get(second, width, old_next - second_start));
  }
}

/* MemImage::find_solution, src/greedysatostop.rs:446-475 */
int or_mem_find_solution(const uint64_t* data, uint64_t start, uint64_t len, uint32_t state_len,
                         uint32_t unknowns, uint64_t* sol_start, uint64_t* sol_end) {
  const uint32_t width = state_len + 1;
  const uint64_t before = (1ull << (state_len - unknowns)) - 1ull, step = before + 1ull;
  uint64_t u = (start & before) != 0 ? (start & ~before) + step : (start & ~before);
  const uint64_t end = start + len;
  const uint64_t state_mask = mem_mask(width) >> 1, stop_mask = 1ull << state_len;
  for (; u < end; u += step) {
    const uint64_t v = or_mem_get(data, width, u - start);
    if (v & stop_mask) {
      *sol_start = u;
      *sol_end = v & state_mask;
      return 1;
    }
  }
  return 0;
}
