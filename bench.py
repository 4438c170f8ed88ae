"""Benchmark: brute-force SAT gate-evaluation rate on one MI355X.

A step is one complete search of the 36-input, 20,000-gate SAT workload (seed 2, 40-bit
output word, planted 36-bit key): all 2^36 assignments evaluated bit-sliced on the GPU,
satisfying count and lowest assignment reduced on the device. value = gate-evals/s =
2^36 x (gates in the output cone) / seconds per step.

Also reported on the same line:
  roofline      the search kernel's algorithmic 32-bit op rate (one op per cone gate per
                32 assignments) against the gfx950 VALU peak (256 CUs x 128 lanes/clk x
                2.4 GHz = 78.6e12 32-bit ops/s, the published 157.3 TFLOP/s FP32 vector rate
                counted as one op per FMA), from HIP events around the launches on the
                engine's stream; `traffic` = HBM bytes per launch measured with rocprofv3
                FETCH_SIZE / WRITE_SIZE (profiles/traffic.txt), when recorded;
  cpu_baseline  the CPU oracle (C, AVX2, pthreads over the host cores) on a bounded slice
                of the same search;
  extra         the same search with group exits (NS_SAT_GROUP_EXITS: waves leave a word once
                the output's groups so far are zero; time to the same answer), the SATOSTOP
                search over the 30-bit Feistel's 2^30-entry table (waypoint walks,
                their scattered-read rate against a probe of the same table size, and the
                passes over the whole table with their HBM rate, and the walks that evaluate
                the circuit instead of reading a table), the same without a table at 36 state
                bits, one sweep of the hash-map search at 34 state bits, the
                QSAT search time (32 inputs, 10,000 gates) and the served request stream
                (4,096 SAT/QSAT/SATOSTOP requests, Poisson 200/s, seed 5): solves/s and
                p50/p99 latency with every request queued at once and with requests
                submitted online as they arrive (200/s and 10x that), one at a time against
                8 concurrent lanes with and without batching into persistent kernels; every
                answer compared with the oracle's (tests/golden/cfg4_answers.json).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra] [--no-serve]
       python bench.py --only {sat,exits,factor,table,stop36,hashmap,qsat,serve,copy,batch}   (one part, for profiling runs;
       `copy` times a plain 1 GiB device copy used to calibrate the HBM byte counters)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "gate-evals/s vs VALU peak; batched solves/s + p99 latency vs one-at-a-time"
VALU_PEAK_OPS = 256 * 128 * 2.4e9  # 32-bit VALU lane-ops/s (FP32 vector peak / 2)
HBM_PEAK_GBS = 8000.0
SAT_WORKLOAD = ("SAT n=36 inputs, 20000 random AND/OR/XOR gates (seed 2), output [W(x)==W(k)] over a "
                "40-gate word, planted 36-bit key")
TABLE_WORKLOAD = "SATOSTOP n=30 Feistel (6 rounds x 300 gates, seed 4), 14 unknowns, 2^30-entry table"


def traffic_commit():
    """the commit profiles/traffic.txt names for its SAT kernel line, or None"""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if os.path.exists(path):
        for line in open(path):
            if "re-measured at commit" in line:
                return line.split("commit", 1)[1].split()[0]
    return None


def measured_traffic():
    """kernel -> HBM bytes per launch, from profiles/traffic.txt (lines: kernel workload bytes)"""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            t = line.split()
            if len(t) == 3 and not line.startswith("#"):
                out[(t[0], t[1])] = float(t[2])
    return out


def measured_valu_insts():
    """kernel -> wave-level VALU instructions per launch (rocprofv3 SQ_INSTS_VALU), from
    profiles/sq_insts.txt (lines: kernel workload count)"""
    path = os.path.join(ROOT, "profiles", "sq_insts.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            t = line.split()
            if len(t) == 3 and not line.startswith("#"):
                out[(t[0], t[1])] = float(t[2])
    return out


def sat_part(ns, ctx, steps, warmup):
    c, key = ns.Circuit.gen_sat(2, 36, 20000, 40)
    cone, peak_live = c.cone()
    n = 36
    first_compile_ms = None
    for _ in range(warmup):
        r = ctx.sat_solve(c)
        if first_compile_ms is None:
            first_compile_ms = r["compile_ms"]
    t0 = time.perf_counter()
    kernel_ms, launches = 0.0, 0
    for _ in range(steps):
        r = ctx.sat_solve(c)
        kernel_ms += r["kernel_ms"]
        launches += r["launches"]
    t1 = time.perf_counter()
    if first_compile_ms is None:
        first_compile_ms = r["compile_ms"]
    ms_per_step = (t1 - t0) * 1000.0 / steps
    value = (1 << n) * cone / (ms_per_step / 1000.0)
    # roofline of the search kernel: algorithmic ops per launch / average launch duration
    avg_launch_ms = kernel_ms / launches
    words_per_launch = (1 << (n - 5)) / (launches / steps)
    ops_per_launch = words_per_launch * cone
    achieved = ops_per_launch / (avg_launch_ms / 1000.0) / 1e12
    peak = VALU_PEAK_OPS / 1e12
    kernel = f"ns_nat_sat_r{r['vgpr_class']}" if r["backend"] else "ns_sat"
    traffic = measured_traffic().get((kernel, "sat_n36"))
    # issued VALU work: the recorded wave-level VALU instruction count of one launch (64 lanes
    # each) over the live launch time, against the same lane-op peak
    insts = measured_valu_insts().get((kernel, "sat_n36"))
    issued = None if insts is None else insts * 64 / (avg_launch_ms / 1000.0) / VALU_PEAK_OPS
    lut_ops_per_s = words_per_launch * r["valu_ops"] / (avg_launch_ms / 1000.0)
    return c, cone, {
        "ms_per_step": ms_per_step, "value": value,
        "config": {"workload": SAT_WORKLOAD, "cone_gates": cone, "gates_nominal": 20000,
                   "peak_live_values": peak_live, "assignments_per_step": 1 << n,
                   "launches_per_step": launches // steps, "count": r["count"], "lowest": r["lowest"],
                   "key": key, "backend": "native" if r["backend"] else "hiprtc",
                   "luts_per_word": r["valu_ops"], "scalar_ops_per_word": r["scalar_ops"],
                   "vgpr_class": r["vgpr_class"], "compile_ms_first": first_compile_ms},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TOP/s",
                     "frac": achieved / peak, "traffic": traffic, "kernel": kernel,
                     "traffic_source": "copied from profiles/traffic.txt (not measured in this run): rocprofv3 "
                                       "FETCH_SIZE x 1.99994 + WRITE_SIZE x 1.0 per dispatch (copy-kernel "
                                       f"calibration), measured at commit {traffic_commit()}"
                     if traffic is not None else None,
                     "avg_launch_ms": avg_launch_ms, "ops_per_launch": ops_per_launch,
                     "lut_ops_per_s": lut_ops_per_s, "lut_frac": lut_ops_per_s / VALU_PEAK_OPS,
                     "valu_issue_frac": issued,
                     "valu_issue_source": "profiles/sq_insts.txt: rocprofv3 SQ_INSTS_VALU per dispatch x 64 lanes"
                     if issued is not None else None},
    }


def group_exits_part(ns, ctx, c, full_ms_per_step, full_answer, searches=3):
    """The same 2^36 search with NS_SAT_GROUP_EXITS: a wave leaves a 32-assignment word once
    the output's groups so far are zero in all 64 lanes. Time to the same answer; the gates
    skipped are not evaluated, so this is not a gate-evaluation rate."""
    ctx.sat_solve(c, group_exits=True)  # kernel build (a body of its own)
    t0 = time.perf_counter()
    kernel_ms = 0.0
    for _ in range(searches):
        r = ctx.sat_solve(c, group_exits=True)
        kernel_ms += r["kernel_ms"]
    ms = (time.perf_counter() - t0) * 1000.0 / searches
    return {"workload": SAT_WORKLOAD + ", NS_SAT_GROUP_EXITS", "ms_per_search": ms,
            "kernel_ms_per_search": kernel_ms / searches,
            "speedup_vs_full_evaluation": full_ms_per_step / ms,
            "assignments_per_s": (1 << 36) / (ms / 1000.0),
            "count": r["count"], "lowest": r["lowest"],
            "answers_equal": (r["count"], r["lowest"]) == full_answer}


REFERENCE_FACTOR_ANSWER = "000000000000000110111010101000000000000100110011100001"


def factor_part(ns):
    """The search of the reference's published naivesat runs (benchmarks.txt: 54-input
    circuit for 69,690,421, answer printed above, 14.959 s wall with -e 32 on a Radeon Vega
    56): the naivesat tool on the rebuilt circuit, a = inputs 0..26, b = 27..53, output
    [a * b == 69,690,421] and a != 1 and b != 1, stopping at the first launch range of 2^32
    assignments that holds a solution, with group exits (the tool's default) and with every
    gate evaluated (-X). Wall time of the whole process, as `time naivesat` measures it."""
    import subprocess
    import tempfile

    c = ns.Circuit.gen_factor(69690421, 27)
    out = {"workload": "naivesat -e 32 on [a * b == 69690421], a != 1, b != 1 (54 inputs, "
                       f"{c.arrays()['func'].size} gates, rebuilt: the reference's circuit file is not "
                       "in the reference)",
           "reference_published_s": 14.959, "reference_hardware": "Radeon Vega 56 (benchmarks.txt:15-19)"}
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "prime_69690421.circuit")
        with open(f, "w") as fh:
            fh.write(c.text())
        for name, extra in (("group_exits", []), ("every_gate", ["-X"])):
            t0 = time.perf_counter()
            r = subprocess.run([os.path.join(ns.BIN_DIR, "naivesat"), "-e", "32", "-t", "opencl:0", *extra, f],
                               capture_output=True, text=True, timeout=300)
            wall = time.perf_counter() - t0
            lines = r.stdout.splitlines()
            tline = [x for x in lines if x.startswith("Time: ")]
            steps = [x for x in lines if x.startswith("Step: ")]
            out[name] = {"wall_s": wall, "search_s": float(tline[-1].split()[1]) if tline else None,
                         "launch_ranges": len(steps),
                         "answer_equals_reference": f"Found Input: {REFERENCE_FACTOR_ANSWER}" in lines,
                         "speedup_vs_reference_wall": 14.959 / wall}
    return out


def cpu_part(c, cone, threads):
    import oracle

    oc = oracle.OracleCircuit.from_dict(c.arrays())
    words = 1 << 28  # 2^34 assignments (64 per oracle word): ~13 s with 16 AVX-512 threads
    tc0 = time.perf_counter()
    oc.sat_range(0, words, threads=threads)
    tc = time.perf_counter() - tc0
    return {"value": (words * 64) * cone / tc, "unit": "gate-evals/s", "cores": threads, "kind": "port",
            "sample": f"first 2^34 of the 2^36 assignments, {tc:.1f} s", "cpu_model": oracle.cpu_model(),
            "isa": oracle.simd_isa(), "implementation": "oracle/oracle.c: C, bit-sliced 8 x 64-bit words per gate, "
                                                          "pthreads over the assignment range"}


def table_part(ns, ctx):
    """The 2^30-state SATOSTOP search over its next-state table: waypoint walks over the
    table and, beside them, the passes over the whole table; and the default full search,
    walks without a table (walks_without_table); all with the same answers."""
    cs, target = ns.Circuit.gen_satostop(4, 15, 6, 300, 14)
    ctx.satostop_solve(cs, 14, arrays=False, method="table")  # kernel build
    rs = ctx.satostop_solve(cs, 14, arrays=False, method="table")
    rp = ctx.satostop_solve(cs, 14, arrays=False, passes=True)
    jump = ctx.pass_times()[1:]  # u64 passes (the first reads the u32 table)
    rb0 = ctx.satostop_solve(cs, 14, arrays=False, method="bits")  # kernel build
    rb = ctx.satostop_solve(cs, 14, arrays=False, method="bits")
    for r in (rp, rb):
        assert (rs["n_stopped"], rs["lowest_u"], rs["steps"], rs["end"]) == \
            (r["n_stopped"], r["lowest_u"], r["steps"], r["end"])
    entries = 1 << 30
    bytes_pass = entries * 24  # read entry + gathered entry + write, 8 B each
    avg = sum(jump) / len(jump)
    gbs = bytes_pass / (avg / 1000.0) / 1e9
    # the walks read one 4-byte table entry per evaluation at a scattered index
    probe = ctx.probe_gather(30, 256, 1 << 20)
    walk_s = rs["walk_ms"] / 1000.0
    reads_per_s = rs["walk_evaluations"] / walk_s
    walk_gbs = rs["walk_evaluations"] * 4 / walk_s / 1e9
    traffic = measured_traffic()
    walk_traffic = traffic.get(("k_wp_walk", "satostop_n30"))
    pass_traffic = traffic.get(("k_jump_pass", "satostop_n30"))

    def traffic_rate(t, seconds):
        # measured HBM bytes of one launch (rocprofv3 counters, profiles/traffic.txt) over its time
        return None if t is None else t / seconds / 1e9
    walk_tr = traffic_rate(walk_traffic, walk_s)
    pass_tr = traffic_rate(pass_traffic, avg / 1000.0)
    return {"workload": TABLE_WORKLOAD, "method": "waypoint walks", "waypoint_bits": rs["waypoint_bits"],
            "build_ms": rs["build_ms"], "walk_ms": rs["walk_ms"], "walk_evaluations": rs["walk_evaluations"],
            "waypoint_passes": rs["waypoint_passes"], "waypoint_jump_ms": rs["jump_ms"],
            "device_ms": rs["build_ms"] + rs["walk_ms"] + rs["jump_ms"], "wall_ms": rs["wall_ms"],
            "n_stopped": rs["n_stopped"], "lowest_u": rs["lowest_u"], "steps": rs["steps"], "end": rs["end"],
            "roofline": {"bound": "hbm", "achieved": walk_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": walk_gbs / HBM_PEAK_GBS, "traffic": walk_traffic,
                         "traffic_gbs": walk_tr, "traffic_frac": None if walk_tr is None else walk_tr / HBM_PEAK_GBS,
                         "kernel": "k_wp_walk", "bytes_per_evaluation": 4,
                         "scattered_reads_per_s": reads_per_s,
                         "scattered_read_ceiling_per_s": probe["reads_per_s"],
                         "frac_of_scattered_read_ceiling": reads_per_s / probe["reads_per_s"]},
            "passes_method": {"device_ms": rp["build_ms"] + rp["jump_ms"] + rp["walk_ms"],
                              "passes": rp["passes"], "walk_jumps": rp["walk_jumps"], "walk_ms": rp["walk_ms"],
                              "avg_pass_ms": avg,
                              "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                           "frac": gbs / HBM_PEAK_GBS, "traffic": pass_traffic,
                                           "traffic_gbs": pass_tr,
                                           "traffic_frac": None if pass_tr is None else pass_tr / HBM_PEAK_GBS,
                                           "kernel": "k_jump_pass", "bytes_per_entry": 24}},
            "speedup_vs_passes": (rp["build_ms"] + rp["jump_ms"] + rp["walk_ms"]) /
                                 (rs["build_ms"] + rs["walk_ms"] + rs["jump_ms"]),
            "walks_without_table": walks_without_table(rb, rb0)}


def walks_without_table(r, first=None):
    """the search by walks that evaluate the circuit (32 per lane per step): VALU-bound, one
    op per cone gate per 32 evaluations. first: the run that built the kernel"""
    kernel = f"ns_nat_walk*_r{r['vgpr_class']}" if r["backend"] else "ns_walk"
    walk_s = r["walk_ms"] / 1000.0
    tops = r["walk_evaluations"] * r["cone_gates"] / 32 / walk_s / 1e12
    return {"method": r["method"], "waypoint_bits": r["waypoint_bits"], "walk_ms": r["walk_ms"],
            "waypoint_jump_ms": r["jump_ms"], "device_ms": r["walk_ms"] + r["jump_ms"],
            "walk_evaluations": r["walk_evaluations"], "evaluations_per_s": r["walk_evaluations"] / walk_s,
            "cone_gates": r["cone_gates"], "luts_per_step": r["valu_ops"], "n_stopped": r["n_stopped"],
            "lowest_u": r["lowest_u"], "backend": "native" if r["backend"] else "hiprtc",
            "vgpr_class": r["vgpr_class"], "compile_ms_first": None if first is None else first["compile_ms"],
            "roofline": {"bound": "valu", "achieved": tops, "peak": VALU_PEAK_OPS / 1e12, "unit": "TOP/s",
                         "frac": tops * 1e12 / VALU_PEAK_OPS, "kernel": kernel}}


STOP36_WORKLOAD = ("SATOSTOP n=36 Feistel (6 rounds x 200 gates, seed 10601), 14 unknowns: every unknown's "
                   "answer by walks without a table")


def stop36_part(ns, ctx):
    c, target = ns.Circuit.gen_satostop(10601, 18, 6, 200, 14)
    r0 = ctx.satostop_solve(c, 14, arrays=False)  # kernel build
    r = ctx.satostop_solve(c, 14, arrays=False)
    ctx.trim()
    d = walks_without_table(r, r0)
    d.update({"workload": STOP36_WORKLOAD, "wall_ms": r["wall_ms"], "steps": r["steps"], "end": r["end"]})
    return d


HASHMAP_WORKLOAD = ("SATOSTOP n=34 Feistel (6 rounds x 136 gates, seed 9001), 26 unknowns, hash-map search "
                    "(2^24 states per block, 2^24-entry map), one sweep of 1,024 blocks")


def hashmap_part(ns, ctx):
    """One sweep of naivesatostop's hash-map search past the table's 33 state bits."""
    c, _ = ns.Circuit.gen_satostop(9001, 17, 6, 136, 26)
    r = ctx.satostop_hashmap_solve(c, 26, hashmap_len_bits=24, elem_bits=24, max_sweeps=1)
    per_block = r["kernel_ms"] / max(1, r["blocks"])
    return {"workload": HASHMAP_WORKLOAD, "found": r["found"], "blocks": r["blocks"], "kernel_ms": r["kernel_ms"],
            "ms_per_block": per_block, "states_per_s": (1 << 24) / (per_block / 1000.0),
            "resolved_unknowns": r["resolved_unknowns"]}


def qsat_part(ns, ctx):
    cq = ns.Circuit.gen_qsat(3, 32, 10000, 200)
    q = [0] * 8 + [1] * 8 + [0] * 8 + [1] * 8
    ctx.qsat_solve(cq, q)
    rq = ctx.qsat_solve(cq, q)
    # one op per cone gate per 32-assignment word, over the whole search (the circuit kernel
    # and the reduction passes after it)
    words = (1 << 32) // 32
    tops = words * rq["cone_gates"] / (rq["kernel_ms"] / 1000.0) / 1e12
    return {"workload": "QSAT n=32, 10000 gates (seed 3), E8 A8 E8 A8", "kernel_ms": rq["kernel_ms"],
            "truth": rq["truth"], "result": ns.qsat_result_string(rq), "cone_gates": rq["cone_gates"],
            "gate_evals_per_s": (1 << 32) * rq["cone_gates"] / (rq["kernel_ms"] / 1000.0),
            "roofline": {"bound": "valu", "achieved": tops, "peak": VALU_PEAK_OPS / 1e12, "unit": "TOP/s",
                         "frac": tops * 1e12 / VALU_PEAK_OPS, "kernel": "ns_nat_qsat* + k_qsat_reduce",
                         "luts_per_word": rq["valu_ops"], "scalar_ops_per_word": rq["scalar_ops"],
                         "vgpr_class": rq["vgpr_class"],
                         "lut_ops_per_s": words * rq["valu_ops"] / (rq["kernel_ms"] / 1000.0)}}


def batch_repeat_part(ns, n_requests=64, repeats=5):
    """64 SAT / QSAT requests of the served stream: one at a time on a fresh context, as one
    batch the first time (code generation, one module load, eager launches) and as the same
    batch again (its kept plan replayed as a HIP graph); answers compared across all three."""
    from naivesat_amd import trace

    reqs = [rq for rq in trace.make_trace(seed=5, n_requests=200) if rq["kind"] != ns.REQ_SATOSTOP][:n_requests]
    ctx = ns.Context(0)
    t0 = time.perf_counter()
    alone = []
    for rq in reqs:
        if rq["kind"] == ns.REQ_SAT:
            r = ctx.sat_solve(rq["circuit"])
            alone.append((r["count"], r["lowest"], r["found"], 0))
        else:
            r = ctx.qsat_solve(rq["circuit"], rq["quants"])
            alone.append((r["truth"], r["witness"], r["witness_count"], r["has_witness"]))
    alone_ms = (time.perf_counter() - t0) * 1e3
    ctx.close()
    ctx = ns.Context(0)
    t0 = time.perf_counter()
    first, info0 = ctx.batch_solve(reqs)
    first_ms = (time.perf_counter() - t0) * 1e3
    rep_ms, rep_kernel = [], []
    same = [(r["v0"], r["v1"], r["v2"], r["v3"]) for r in first] == alone
    for _ in range(repeats):
        t0 = time.perf_counter()
        res, info = ctx.batch_solve(reqs)
        rep_ms.append((time.perf_counter() - t0) * 1e3)
        rep_kernel.append(info["kernel_ms"])
        same &= [(r["v0"], r["v1"], r["v2"], r["v3"]) for r in res] == alone
    ctx.close()
    # algorithmic work of the batch: one op per cone gate per 32-assignment word, every request
    gate_ops = sum(float(1 << max(0, rq["circuit"].n_inputs - 5)) * rq["circuit"].cone()[0] for rq in reqs)
    rep_rate = gate_ops / (min(rep_kernel) / 1e3)
    out = {"workload": f"{len(reqs)} SAT/QSAT requests of the served stream (seed 5)", "answers_equal": same,
           "one_at_a_time_ms": alone_ms, "batch_first_ms": first_ms, "batch_first_kernel_ms": info0["kernel_ms"],
           "batch_first_gen_ms": info0["gen_ms"], "batch_first_load_ms": info0["load_ms"],
           "batch_code_bytes": info0["code_bytes"], "graph_replay_ms": min(rep_ms),
           "graph_replay_kernel_ms": min(rep_kernel), "graph_runs": info["graph_runs"],
           "solves_per_s_one_at_a_time": len(reqs) / (alone_ms / 1e3),
           "solves_per_s_batch_first": len(reqs) / (first_ms / 1e3),
           "solves_per_s_graph_replay": len(reqs) / (min(rep_ms) / 1e3),
           "roofline": {"bound": "valu", "kernel": "ns_nat_bsat_r* / ns_nat_bqsat_r* (one launch per kind and "
                                                   "register class)",
                        "gate_ops_per_batch": gate_ops, "achieved": rep_rate / 1e12, "peak": VALU_PEAK_OPS / 1e12,
                        "unit": "TOP/s", "frac": rep_rate / VALU_PEAK_OPS,
                        "note": "graph replay's kernel time (HIP events around the batch's launches)"}}
    print(f"batch_repeat: {json.dumps(out)}", file=sys.stderr, flush=True)
    return out


GOLDEN_CFG4 = os.path.join(ROOT, "tests", "golden", "cfg4_answers.json")


def golden_answers(n_requests):
    """(v0, v1, v2, v3) per request of the served stream from the committed oracle answers
    (tests/golden/cfg4_answers.json), or None when absent or for another trace length"""
    if not os.path.exists(GOLDEN_CFG4):
        return None
    with open(GOLDEN_CFG4) as f:
        doc = json.load(f)
    rows = doc["rows"][:n_requests]
    if len(rows) < n_requests:
        return None
    out = []
    for kind, n, a, b, c, d in rows:
        if kind == 0:
            out.append((a, b if a else 0, int(a > 0), 0))
        elif kind == 1:
            out.append((a, c if b else 0, b, int(b > 0)))
        else:
            out.append((a, b or 0, c or 0, d or 0))
    return out


def online_run(ns, reqs, workers, batch_max, speedup):
    """The online API driven the way a service would: an arrival thread submits each request
    at its Poisson arrival time (divided by `speedup`), the main thread waits for every
    ticket. Returns (results, seconds from the first submit to the last completion)."""
    import threading

    ctx = ns.Context(0)
    srv = ctx.open_server(workers=workers, batch_max=batch_max)
    tickets = [None] * len(reqs)
    t0 = time.perf_counter()

    def feed():
        for i, rq in enumerate(reqs):
            delay = rq["arrival_s"] / speedup - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets[i] = srv.submit(rq)

    th = threading.Thread(target=feed)
    th.start()
    res = []
    for i in range(len(reqs)):
        while tickets[i] is None:
            time.sleep(0.0002)
        res.append(srv.wait(tickets[i]))
    th.join()
    total = max(r["finish_s"] for r in res) - min(r["submit_s"] for r in res)
    srv.close()
    ctx.close()
    return res, total


def serve_part(ns, n_requests, workers, batch_max=32):
    """The served stream (cfg4): saturated (every request queued at once) and online (arrivals
    submitted as they happen), each one at a time and served (concurrent lanes, with and
    without batching); every answer checked against the oracle's. Every run starts with an
    empty kernel cache."""
    import numpy as np
    from naivesat_amd import trace

    reqs = trace.make_trace(seed=5, n_requests=n_requests, rate=200.0)
    gold = golden_answers(n_requests)
    out = {"workload": f"{n_requests} requests (60% SAT / 25% QSAT / 15% SATOSTOP, n 16..32, gates "
                       "200..20000 log-uniform), Poisson 200/s, seed 5",
           "answers_checked_against": "tests/golden/cfg4_answers.json (oracle)" if gold else None}
    mismatches = 0

    def summarize(name, res, total, w, b, mode):
        nonlocal mismatches
        lat = np.array([r["finish_s"] - r["submit_s"] for r in res])
        bad = sum(1 for r in res if r["status"] != 0)
        wrong = None
        if gold is not None:
            wrong = sum(1 for r, g in zip(res, gold) if (r["v0"], r["v1"], r["v2"], r["v3"]) != g)
            mismatches += wrong
        d = {"mode": mode, "workers": w, "batch_max": b, "seconds": total, "solves_per_s": len(res) / total,
             "failed": bad, "wrong_answers": wrong, "p50_ms": float(np.percentile(lat, 50)) * 1e3,
             "p99_ms": float(np.percentile(lat, 99)) * 1e3,
             "compile_ms_total": sum(r["compile_ms"] for r in res)}
        out[name] = d
        print(f"serve {name}: {json.dumps(d)}", file=sys.stderr, flush=True)

    # saturated: all 4,096 requests at t = 0 (latency here is queueing)
    for name, w, b in (("saturated_one_at_a_time", 1, 0), ("saturated_lanes", workers, 0),
                       ("saturated_batched", workers, batch_max)):
        ctx = ns.Context(0)
        res, total = ctx.serve(reqs, workers=w, realtime=False, batch_max=b)
        ctx.close()
        summarize(name, res, total, w, b, "all queued at t=0")
    # online at the trace's rate (200/s) and 10x compressed (~2,000/s, beyond one lane)
    for name, w, b, sp in (("online_200_one_at_a_time", 1, 0, 1.0), ("online_200_served", workers, batch_max, 1.0),
                           ("online_2000_one_at_a_time", 1, 0, 10.0), ("online_2000_lanes", workers, 0, 10.0),
                           ("online_2000_served", workers, batch_max, 10.0)):
        res, total = online_run(ns, reqs, w, b, sp)
        summarize(name, res, total, w, b, f"online submit at arrival/{sp:g} (Poisson {200 * sp:g}/s)")
    out["saturated_speedup"] = out["saturated_batched"]["solves_per_s"] / out["saturated_one_at_a_time"]["solves_per_s"]
    out["online_200_p99_speedup"] = out["online_200_one_at_a_time"]["p99_ms"] / out["online_200_served"]["p99_ms"]
    out["online_2000_p99_speedup"] = out["online_2000_one_at_a_time"]["p99_ms"] / out["online_2000_served"]["p99_ms"]
    out["online_2000_solves_speedup"] = (out["online_2000_served"]["solves_per_s"] /
                                         out["online_2000_one_at_a_time"]["solves_per_s"])
    out["wrong_answers_total"] = mismatches if gold is not None else None
    return out


def copy_part():
    """A plain device copy of 1 GiB (torch), the reference point for the byte counters."""
    import torch

    x = torch.empty(1 << 27, dtype=torch.int64, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    y.copy_(x)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    return {"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30, "ms": ms,
            "gbs": 2 * (1 << 30) / (ms / 1000.0) / 1e9}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--serve-requests", type=int, default=4096)
    ap.add_argument("--serve-workers", type=int, default=8)
    ap.add_argument("--serve-batch", type=int, default=32)
    ap.add_argument("--only", choices=["sat", "exits", "factor", "table", "stop36", "hashmap", "qsat", "serve",
                                       "copy", "batch"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(one GPU per process; the searches shard by assignment range across processes)",
              file=sys.stderr)
        sys.exit(2)

    import naivesat_amd as ns

    if args.only:
        if args.only == "copy":
            part = copy_part()
        elif args.only == "batch":
            part = batch_repeat_part(ns)
        elif args.only == "factor":
            part = factor_part(ns)
        elif args.only == "serve":
            part = serve_part(ns, args.serve_requests, args.serve_workers, args.serve_batch)
        else:
            ctx = ns.Context(0)
            if args.only == "sat":
                part = sat_part(ns, ctx, args.steps, args.warmup)[2]
            elif args.only == "exits":  # the full evaluation, then the same with group exits
                c, _, sat = sat_part(ns, ctx, args.steps, args.warmup)
                part = group_exits_part(ns, ctx, c, sat["ms_per_step"],
                                        (sat["config"]["count"], sat["config"]["lowest"]), args.steps)
            elif args.only == "table":
                part = table_part(ns, ctx)
            elif args.only == "hashmap":
                part = hashmap_part(ns, ctx)
            elif args.only == "stop36":
                part = stop36_part(ns, ctx)
            else:
                part = qsat_part(ns, ctx)
        print(json.dumps({"only": args.only, **part}))
        return

    ctx = ns.Context(0)
    c, cone, sat = sat_part(ns, ctx, args.steps, args.warmup)
    line = {
        "metric": METRIC, "value": sat["value"], "unit": "gate-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": sat["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic", "config": sat["config"],
        "roofline": sat["roofline"],
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_part(c, cone, args.cpu_threads or min(os.cpu_count() or 1, 64))
    if not args.no_extra:
        extra = {"sat_group_exits": group_exits_part(ns, ctx, c, sat["ms_per_step"],
                                                     (sat["config"]["count"], sat["config"]["lowest"])),
                 "reference_factoring_run": factor_part(ns),
                 "satostop_table": table_part(ns, ctx), "satostop_36": stop36_part(ns, ctx),
                 "satostop_hashmap": hashmap_part(ns, ctx), "qsat": qsat_part(ns, ctx)}
        extra["batch_repeat"] = batch_repeat_part(ns)
        if not args.no_serve:
            extra["serve"] = serve_part(ns, args.serve_requests, args.serve_workers, args.serve_batch)
        line["extra"] = extra
    print(json.dumps(line))
    wrong = line.get("extra", {}).get("serve", {}).get("wrong_answers_total")
    if wrong:
        print(f"bench.py: {wrong} served answers differ from the oracle's", file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
