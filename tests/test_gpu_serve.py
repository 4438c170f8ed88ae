"""Serving: requests solved concurrently on worker lanes (and in batches) give exactly the
answers they give when solved alone; the full served stream matches the oracle's answers."""
import json
import os
import threading
import time

import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from naivesat_amd import trace

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cfg4_answers.json")


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


def golden_tuple(row):
    """(v0, v1, v2, v3) a served request reports, from an oracle row of cfg4_answers.json"""
    kind, n, a, b, c, d = row
    if kind == ns.REQ_SAT:
        return (a, b if a else 0, int(a > 0), 0)
    if kind == ns.REQ_QSAT:
        return (a, c if b else 0, b, int(b > 0))
    return (a, b or 0, c or 0, d or 0)


@pytest.fixture(scope="module")
def mixed():
    return trace.make_trace(seed=5, n_requests=96)


@pytest.fixture(scope="module")
def stream():
    return trace.make_trace(seed=5)


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        doc = json.load(f)
    assert doc["trace"] == {"seed": 5, "n_requests": trace.STREAM_LEN, "rate": 200.0}
    return doc["rows"]


@pytest.mark.parametrize("workers", [1, 8])
def test_mix_equals_alone(ctx, mixed, workers):
    res, total = ctx.serve(mixed, workers=workers, realtime=False)
    assert total > 0
    for i, (rq, r) in enumerate(zip(mixed, res)):
        assert r["status"] == 0, (i, r)
        assert r["finish_s"] >= r["start_s"] >= r["submit_s"]
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, rq), (i, rq["kind"], rq["n"])


def test_prefix_matches_golden(mixed, golden):
    # the 96-request trace is the head of the full stream, and its answers are the oracle's
    for rq, row in zip(mixed, golden):
        assert (rq["kind"], rq["n"]) == (row[0], row[1])


@pytest.mark.parametrize("workers,batch", [(8, 0), (1, 0), (8, 32)])
def test_full_stream_bit_exact(ctx, stream, golden, workers, batch):
    """All 4,096 requests of the served stream (seed 5), saturated: every answer equals the
    oracle's (tests/golden/cfg4_answers.json)."""
    print(f"full stream workers={workers} batch={batch}", flush=True)
    res, total = ctx.serve(stream, workers=workers, realtime=False, batch_max=batch)
    print(f"  {len(stream) / total:.0f} solves/s", flush=True)
    bad = [(i, stream[i]["kind"], stream[i]["n"]) for i, (r, row) in enumerate(zip(res, golden))
           if r["status"] != 0 or (r["v0"], r["v1"], r["v2"], r["v3"]) != golden_tuple(row)]
    assert not bad, bad[:20]


def test_small_requests_match_oracle(ctx, mixed):
    res, _ = ctx.serve(mixed, workers=4, realtime=False)
    checked = 0
    for rq, r in zip(mixed, res):
        if rq["n"] > 19:
            continue
        oc = oracle.OracleCircuit.from_dict(rq["circuit"].arrays())
        if rq["kind"] == ns.REQ_SAT:
            count, lowest = oc.sat(threads=8)
            assert (r["v0"], r["v2"]) == (count, int(count > 0))
            if count:
                assert r["v1"] == lowest
        elif rq["kind"] == ns.REQ_QSAT:
            o = oc.qsat(list(rq["quants"]))
            assert (r["v0"], r["v2"]) == (o["truth"], o["witness_count"])
            assert (r["v1"] if r["v3"] else None) == o["witness"]
        else:
            o = oc.satostop(rq["unknowns"], threads=8)
            assert r["v0"] == o["n_stopped"]
            if o["n_stopped"]:
                lo = o["lowest_u"]
                assert (r["v1"], r["v2"], r["v3"]) == (lo, o["steps"][lo], o["ends"][lo])
        checked += 1
    assert checked >= 10


def test_hbm_budget_admission(ctx, mixed):
    # a budget that admits one SATOSTOP table at a time still serves every request
    stop = [rq for rq in mixed if rq["kind"] == ns.REQ_SATOSTOP][:6]
    # u32 table + two u64 jump tables + per-start u64 steps and ends, each padded to 256 bytes
    biggest = max((1 << rq["n"]) * 20 + (1 << rq["unknowns"]) * 16 + 2048 for rq in stop)
    res, _ = ctx.serve(stop, workers=4, realtime=False, hbm_budget_bytes=biggest)
    for rq, r in zip(stop, res):
        assert r["status"] == 0
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, rq)


def test_every_state_unknown_is_served(ctx):
    # unknowns == state bits: every state is a start, so the waypoint records take 16 bytes
    # per state; the lent buffers follow the solver's own plan
    reqs = []
    for seed, half in ((21, 5), (22, 7), (23, 9)):
        c, _ = ns.Circuit.gen_satostop(seed, half, 6, 40, 2 * half)
        reqs.append({"kind": ns.REQ_SATOSTOP, "circuit": c, "unknowns": 2 * half, "n": 2 * half})
    res, _ = ctx.serve(reqs, workers=2, realtime=False)
    for rq, r in zip(reqs, res):
        assert r["status"] == 0, r
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, rq)


def test_oversized_satostop_fails_alone(ctx):
    c, _ = ns.Circuit.gen_satostop(5, 16, 6, 40, 8)  # 32 state bits: outside the served range
    small, _ = ns.Circuit.gen_sat(3, 16, 100, 16)
    reqs = [{"kind": ns.REQ_SATOSTOP, "circuit": c, "unknowns": 8, "n": 32},
            {"kind": ns.REQ_SAT, "circuit": small, "n": 16}]
    res, _ = ctx.serve(reqs, workers=2, realtime=False)
    assert res[0]["status"] == 4  # NS_ERR_UNSUPPORTED
    assert res[1]["status"] == 0 and (res[1]["v0"], res[1]["v1"], res[1]["v2"], res[1]["v3"]) == alone(ctx, reqs[1])


@pytest.mark.parametrize("batch", [0, 8])
def test_online_server(ctx, mixed, batch):
    """Requests submitted from another thread as they arrive (compressed arrival times),
    waited for by ticket: the same answers as alone, each ticket answered once."""
    srv = ctx.open_server(workers=4, batch_max=batch)
    tickets = [None] * len(mixed)
    t0 = time.perf_counter()

    def feed():
        for i, rq in enumerate(mixed):
            delay = rq["arrival_s"] / 10.0 - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets[i] = srv.submit(rq)

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    res = [srv.wait(t) for t in tickets]
    for i, (rq, r) in enumerate(zip(mixed, res)):
        assert r is not None and r["status"] == 0, (i, r)
        assert r["finish_s"] >= r["start_s"] >= r["submit_s"] >= 0
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, rq), (i, rq["kind"])
    with pytest.raises(ns.NaivesatError):
        srv.poll(tickets[0])  # already handed out
    late = srv.submit(mixed[0])
    r = None
    while r is None:
        r = srv.poll(late)
        time.sleep(0.001)
    assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, mixed[0])
    srv.close()
