"""The hiprtc backend (used when a circuit body does not fit the native templates) gives
the same answers as the native backend and the oracle."""
import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import random_circuit

pytestmark = pytest.mark.gpu


@pytest.fixture
def hiprtc_backend(monkeypatch):
    monkeypatch.setenv("NS_BACKEND", "hiprtc")


def mk(d):
    return ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])


def test_sat_hiprtc_matches_native_and_oracle(ctx, monkeypatch):
    c, key = ns.Circuit.gen_sat(11, 22, 600, 22)
    native = ctx.sat_solve(c)
    assert native["backend"] == 1
    monkeypatch.setenv("NS_BACKEND", "hiprtc")
    rtc = ctx.sat_solve(c)
    assert rtc["backend"] == 0
    count, lowest = oracle.OracleCircuit.from_dict(c.arrays()).sat(threads=8)
    for r in (native, rtc):
        assert (r["count"], r["lowest"]) == (count, lowest)


def test_qsat_hiprtc_matches_oracle(ctx, hiprtc_backend):
    c = ns.Circuit.gen_qsat(12, 20, 400, 60)
    q = [0] * 5 + [1] * 5 + [0] * 5 + [1] * 5
    r = ctx.qsat_solve(c, q)
    assert r["backend"] == 0
    o = oracle.OracleCircuit.from_dict(c.arrays()).qsat(q)
    assert (r["truth"], r["witness_count"]) == (o["truth"], o["witness_count"])
    assert (r["witness"] if r["has_witness"] else None) == o["witness"]


def test_table_hiprtc_matches_oracle(ctx, hiprtc_backend):
    rng = np.random.default_rng(77)
    d = random_circuit(rng, 14, 120, nout=15, recent=30)
    t = ctx.table_build(mk(d))
    assert np.array_equal(t, oracle.OracleCircuit.from_dict(d).table_u32())
