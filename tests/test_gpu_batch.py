"""Persistent batch kernels: many SAT / QSAT circuits in one launch give exactly the
answers each gets alone, and the small ones match the oracle."""
import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import random_circuit
from naivesat_amd import trace

pytestmark = pytest.mark.gpu


def alone(ctx, rq):
    c = rq["circuit"]
    if rq["kind"] == ns.REQ_SAT:
        r = ctx.sat_solve(c)
        return (r["count"], r["lowest"], r["found"], 0)
    if rq["kind"] == ns.REQ_QSAT:
        r = ctx.qsat_solve(c, rq["quants"])
        return (r["truth"], r["witness"], r["witness_count"], r["has_witness"])
    r = ctx.satostop_solve(c, rq["unknowns"], arrays=False)
    return (r["n_stopped"], r["lowest_u"], r["steps"], r["end"])


def oracle_answer(rq):
    oc = oracle.OracleCircuit.from_dict(rq["circuit"].arrays())
    if rq["kind"] == ns.REQ_SAT:
        count, lowest = oc.sat(threads=8)
        return (count, lowest if count else 0, int(count > 0), 0)
    o = oc.qsat(list(rq["quants"]), threads=8, bitmap=False)
    has = o["witness_count"] > 0
    return (o["truth"], o["witness"] if has else 0, o["witness_count"], int(has))


def random_requests(seed, count, n_lo, n_hi):
    """SAT and QSAT requests over random circuits, random alternating prefixes"""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        d = random_circuit(rng, n, int(rng.integers(20, 400)), recent=int(rng.integers(4, 40)))
        c = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])
        if i % 3 == 2:
            cut = int(rng.integers(1, n - 5))  # at least 5 levels below the outermost block
            q = [int(rng.integers(0, 2))] * cut
            while len(q) < n:
                q += [1 - q[-1]] * int(rng.integers(1, 4))
            out.append({"kind": ns.REQ_QSAT, "circuit": c, "quants": np.array(q[:n], np.uint8), "n": n})
        else:
            out.append({"kind": ns.REQ_SAT, "circuit": c, "n": n})
    return out


def test_batch_matches_oracle_small(ctx):
    reqs = random_requests(51, 24, 13, 18)
    res, info = ctx.batch_solve(reqs)
    assert info["fallbacks"] == 0
    for i, (rq, r) in enumerate(zip(reqs, res)):
        assert r["status"] == 0, (i, r)
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == oracle_answer(rq), (i, rq["kind"], rq["n"])


def test_batch_matches_alone_trace_prefix(ctx):
    # SAT / QSAT requests of the served stream (n 16..32, 200..20,000 gates)
    reqs = [rq for rq in trace.make_trace(seed=5, n_requests=160) if rq["kind"] != ns.REQ_SATOSTOP]
    res, info = ctx.batch_solve(reqs)
    print("batch info", info, flush=True)
    for i, (rq, r) in enumerate(zip(reqs, res)):
        assert r["status"] == 0, (i, r)
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, rq), (i, rq["kind"], rq["n"], rq["gates"])


def test_batch_mixed_register_classes(ctx):
    # one circuit whose body needs a larger register class beside small ones: each class's
    # requests run in their own class's kernel, launched at that class's occupancy
    reqs = random_requests(77, 6, 14, 20)
    big, _ = ns.Circuit.gen_sat(2, 26, 8000, 40)
    reqs.append({"kind": ns.REQ_SAT, "circuit": big, "n": 26})
    res, info = ctx.batch_solve(reqs)
    print("batch info", info, flush=True)
    assert info["vgpr_class"] >= 128 and info["launches"] >= 2
    for i, (rq, r) in enumerate(zip(reqs, res)):
        assert r["status"] == 0
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, rq), i


def test_batch_dense_hits_and_empty_spaces(ctx):
    # an always-true and an always-false circuit and a tiny space: counts from every chunk
    n = 20
    c_true = ns.Circuit.from_arrays(n, [0], [0], [3], [n], [1])   # ~(x0 ^ x0) = 1
    c_false = ns.Circuit.from_arrays(n, [0], [0], [3], [n], [0])  # x0 ^ x0 = 0
    c_small = ns.Circuit.from_arrays(3, [0], [1], [0], [3], [0])  # x0 & x1 over 8 assignments
    reqs = [{"kind": ns.REQ_SAT, "circuit": c, "n": c.n_inputs} for c in (c_true, c_false, c_small)]
    res, _ = ctx.batch_solve(reqs)
    assert (res[0]["v0"], res[0]["v1"], res[0]["v2"]) == (1 << n, 0, 1)
    assert (res[1]["v0"], res[1]["v2"]) == (0, 0)
    assert (res[2]["v0"], res[2]["v1"], res[2]["v2"]) == (2, 3, 1)


def test_batch_satostop_and_unbatchable_run_alone(ctx):
    reqs = random_requests(91, 3, 14, 16)
    cs, _ = ns.Circuit.gen_satostop(4, 8, 6, 80, 8)
    reqs.append({"kind": ns.REQ_SATOSTOP, "circuit": cs, "unknowns": 8, "n": 16})
    # QSAT whose levels below the outermost block do not reach the thread's bit
    cq = ns.Circuit.gen_qsat(3, 14, 200, 50)
    reqs.append({"kind": ns.REQ_QSAT, "circuit": cq, "quants": np.array([0] * 11 + [1] * 3, np.uint8), "n": 14})
    res, info = ctx.batch_solve(reqs)
    assert info["fallbacks"] == 2
    for i, (rq, r) in enumerate(zip(reqs, res)):
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, rq), i


def test_bodies_past_the_region_run_alone(ctx, monkeypatch):
    """Bodies that do not fit the batch region (its size lowered here to 4 KiB) are solved
    alone instead of failing the batch; the others still run in the batch kernels."""
    monkeypatch.setenv("NS_BATCH_REGION_BYTES", "4096")
    reqs = random_requests(97, 10, 14, 17)
    res, info = ctx.batch_solve(reqs)
    assert 0 < info["fallbacks"] < len(reqs)
    for i, (rq, r) in enumerate(zip(reqs, res)):
        assert r["status"] == 0, (i, r)
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, rq), i


def test_served_batches_past_the_region(ctx, monkeypatch):
    """The same through the server's batches (one lane)."""
    monkeypatch.setenv("NS_BATCH_REGION_BYTES", "4096")
    reqs = random_requests(98, 12, 14, 17)
    res, total = ctx.serve(reqs, workers=1, realtime=False, batch_max=16)
    for i, (rq, r) in enumerate(zip(reqs, res)):
        assert r["status"] == 0, (i, r)
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, rq), i


@pytest.mark.parametrize("workers", [1, 4])
def test_served_batches_equal_alone(ctx, workers):
    reqs = trace.make_trace(seed=5, n_requests=128)
    res, total = ctx.serve(reqs, workers=workers, realtime=False, batch_max=16)
    for i, (rq, r) in enumerate(zip(reqs, res)):
        assert r["status"] == 0, (i, r)
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, rq), (i, rq["kind"], rq["n"])


def test_repeated_batch_replays_its_graph(ctx):
    """The same batch again reuses its code object and runs as a HIP graph (captured on the
    first repeat, replayed after): every replay, with the program's own eager solves in
    between, gives the eager run's answers."""
    reqs = random_requests(123, 12, 13, 20)
    first, info0 = ctx.batch_solve(reqs)
    assert info0["graph_runs"] == 0 and info0["fallbacks"] == 0
    ref = [(r["v0"], r["v1"], r["v2"], r["v3"]) for r in first]
    for rep in range(1, 5):
        c, _ = ns.Circuit.gen_sat(500 + rep, 18, 200, 16)
        ctx.sat_solve(c)  # eager work on the same stream between replays
        res, info = ctx.batch_solve(reqs)
        assert info["graph_runs"] == rep and info["gen_ms"] == 0 and info["load_ms"] == 0
        assert [(r["v0"], r["v1"], r["v2"], r["v3"]) for r in res] == ref, rep
    for rq, r in zip(reqs, ref):
        assert r == oracle_answer(rq)


def test_tuned_launch_shapes_keep_answers(ctx):
    """Batches of one register class run under every candidate launch shape of the autotuner
    (blocks per CU x chunks per wave) while it explores: the answers never change."""
    reqs = random_requests(321, 10, 28, 30)  # about a millisecond of device time per batch
    ref = None
    for rep in range(18):
        # ten rotations of the list: new batches each time (more than the four kept)
        res, _ = ctx.batch_solve(reqs[rep % 10:] + reqs[: rep % 10])
        got = sorted((r["v0"], r["v1"], r["v2"], r["v3"]) for r in res)
        ref = ref or got
        assert got == ref, rep
