"""The allocation-free SAT entry point (ns_sat_solve_noalloc), called through the C ABI:
the same answers as ns_sat_solve on two calls in a row, the second with a smaller input,
with its counters in the library's per-device array."""
import ctypes as C

import numpy as np
import pytest

import naivesat_amd as ns
import oracle

pytestmark = pytest.mark.gpu

KEYS = ("count", "lowest", "found", "complete", "launches", "cone_gates", "valu_ops", "backend", "vgpr_class")


def noalloc(ctx, c, **kw):
    """through the raw C-ABI symbol"""
    lib = ns.load_library()
    o = ns._SatOpts(kw.get("elem_bits", 0), kw.get("group_len", 0), kw.get("words_per_lane", 0),
                    ns.NS_SAT_STOP_AT_FIRST if kw.get("stop_at_first") else 0)
    r = ns._SatResult()
    st = lib.ns_sat_solve_noalloc(ctx._h, c._h, C.byref(o), C.byref(r))
    if st != 0:
        raise ns.NaivesatError(st, lib.ns_last_error().decode())
    return {k: getattr(r, k) for k, _ in ns._SatResult._fields_}


def test_noalloc_matches_sat_solve_then_smaller():
    with ns.Context(0) as ctx:
        big, key = ns.Circuit.gen_sat(1, 24, 2000, 32)      # 2^24 assignments
        small, _ = ns.Circuit.gen_sat(7, 16, 300, 16)       # 2^16 assignments
        ref_big, ref_small = ctx.sat_solve(big), ctx.sat_solve(small)  # builds both kernels
        a = noalloc(ctx, big)
        b = noalloc(ctx, small)  # the second call, with the smaller input, right after
        for k in KEYS:
            assert a[k] == ref_big[k], k
            assert b[k] == ref_small[k], k
        assert a["compile_ms"] == 0.0 and b["compile_ms"] == 0.0
        cnt, low = oracle.OracleCircuit.from_dict(small.arrays()).sat(threads=8)
        assert (b["count"], b["lowest"] if b["count"] else None) == (cnt, low)
        assert a["found"] and a["lowest"] <= key
        # the counters start from zero on every call: the big search again gives its answer
        assert noalloc(ctx, big)["count"] == ref_big["count"]


def test_noalloc_dense_hits_and_stop_at_first():
    n = 22
    c_true = ns.Circuit.from_arrays(n, [0], [0], [3], [n], [1])  # every assignment satisfies
    with ns.Context(0) as ctx:
        ref = ctx.sat_solve(c_true, elem_bits=18)
        r = noalloc(ctx, c_true, elem_bits=18)
        assert (r["count"], r["lowest"], r["found"], r["launches"]) == (1 << n, 0, 1, ref["launches"])
        ctx.sat_solve(c_true, elem_bits=18, stop_at_first=True)
        f = noalloc(ctx, c_true, elem_bits=18, stop_at_first=True)
        assert f["found"] and f["lowest"] == 0 and f["launches"] == 1 and f["complete"] == 0


def test_noalloc_needs_a_built_kernel():
    with ns.Context(0) as ctx:
        c, _ = ns.Circuit.gen_sat(11, 18, 200, 16)
        with pytest.raises(ns.NaivesatError) as e:
            noalloc(ctx, c)
        assert e.value.status == 4  # NS_ERR_UNSUPPORTED
        ctx.sat_solve(c)
        assert noalloc(ctx, c)["count"] == ctx.sat_solve(c)["count"]
