"""Serving: requests solved concurrently on worker lanes give exactly the answers they give
when solved alone, and the small ones match the oracle."""
import numpy as np
import pytest

import naivesat_amd as ns
import oracle
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


@pytest.fixture(scope="module")
def mixed():
    return trace.make_trace(seed=5, n_requests=96)


@pytest.mark.parametrize("workers", [1, 8])
def test_mix_equals_alone(ctx, mixed, workers):
    res, total = ctx.serve(mixed, workers=workers, realtime=False)
    assert total > 0
    for i, (rq, r) in enumerate(zip(mixed, res)):
        assert r["status"] == 0, (i, r)
        assert r["finish_s"] >= r["start_s"]
        assert (r["v0"], r["v1"], r["v2"], r["v3"]) == alone(ctx, rq), (i, rq["kind"], rq["n"])


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
