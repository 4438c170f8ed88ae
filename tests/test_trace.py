"""The served workload's request stream is deterministic and has the stated mix."""
import collections

import numpy as np

import naivesat_amd as ns
from naivesat_amd import trace


def test_trace_is_deterministic():
    a = trace.make_trace(seed=5, n_requests=64)
    b = trace.make_trace(seed=5, n_requests=64)
    for x, y in zip(a, b):
        assert x["kind"] == y["kind"] and x["arrival_s"] == y["arrival_s"] and x["n"] == y["n"]
        dx, dy = x["circuit"].arrays(), y["circuit"].arrays()
        for k in dx:
            assert np.array_equal(np.asarray(dx[k]), np.asarray(dy[k]))


def test_trace_mix_and_ranges():
    t = trace.make_trace(seed=5, n_requests=600)
    kinds = collections.Counter(x["kind"] for x in t)
    assert abs(kinds[ns.REQ_SAT] / 600 - 0.60) < 0.07
    assert abs(kinds[ns.REQ_QSAT] / 600 - 0.25) < 0.06
    assert abs(kinds[ns.REQ_SATOSTOP] / 600 - 0.15) < 0.05
    arr = np.array([x["arrival_s"] for x in t])
    assert np.all(np.diff(arr) > 0)
    assert 2.0 < arr[-1] < 4.0  # 600 requests at 200 per second
    for x in t:
        if x["kind"] == ns.REQ_SATOSTOP:
            assert 16 <= x["n"] <= 26 and x["n"] % 2 == 0
        else:
            assert 16 <= x["n"] <= 32
        if x["kind"] == ns.REQ_QSAT:
            q = x["quants"]
            assert len(q) == x["n"]
            blocks = 1 + int(np.count_nonzero(np.diff(q)))
            assert 2 <= blocks <= 5
