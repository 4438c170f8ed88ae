"""The packed (n+1)-bit next-state image on the device (the reference's MemImage layout,
src/greedysatostop.rs:339-400) and the SATOSTOP search over it."""
import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import point_stop_circuit, random_circuit
from test_oracle_memimage import MEM_IMAGE_CASES, M64, bits_of

pytestmark = pytest.mark.gpu


def mk(d):
    return ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])


def test_pack_matches_reference_layout(ctx):
    # test_mem_image's values (src/greedysatostop.rs:1145-1201) packed by the device's kernel
    for k, (sl, mult, add, length) in enumerate(MEM_IMAGE_CASES):
        width = sl + 1
        vals = np.array([(i * mult + add) & M64 for i in range(length)], np.uint64)
        words = ctx.pack_entries(width, vals)
        want = vals & np.uint64((1 << width) - 1)
        assert np.array_equal(bits_of(words, width, length), want), k
        assert np.array_equal(words, oracle.mem_pack(width, vals)), k


def check_packed(ctx, d, U):
    r = ctx.satostop_solve(mk(d), U, method="table")
    o = oracle.OracleCircuit.from_dict(d).satostop(U, threads=8)
    assert np.array_equal(r["steps_arr"], o["steps"]), (d["n"], U)
    assert np.array_equal(r["ends_arr"], o["ends"]), (d["n"], U)
    assert r["n_stopped"] == o["n_stopped"]
    if o["n_stopped"]:
        assert (r["lowest_u"], r["steps"], r["end"]) == (o["lowest_u"], o["steps"][o["lowest_u"]],
                                                         o["ends"][o["lowest_u"]])
    return r


@pytest.mark.parametrize("n", [6, 9, 13, 17, 20])
def test_search_over_packed_image_matches_oracle(ctx, monkeypatch, n):
    monkeypatch.setenv("NS_PACKED_TABLE", "1")
    rng = np.random.default_rng(6100 + n)
    for _ in range(2):
        d = random_circuit(rng, n, int(rng.integers(n + 1, 4 * n + 20)), nout=n + 1, recent=3 * n + 4)
        r = check_packed(ctx, d, int(rng.integers(0, n + 1)))
        if r["any_stop"]:
            assert r["method"] == 1
    c, _ = ns.Circuit.gen_satostop(6200 + n, n // 2, 6, 10 * (n // 2), min(n - (n & 1), 12))
    r = check_packed(ctx, c.arrays(), min(n - (n & 1), 12))
    assert r["method"] == 1


def test_packed_image_walk_limit_falls_back(ctx, monkeypatch):
    # over 31 bits the split tables and passes still answer when a walk runs too long; here
    # the same path at 14 bits
    monkeypatch.setenv("NS_PACKED_TABLE", "1")
    monkeypatch.setenv("NS_WAYPOINT_LIMIT", "3")
    c, _ = ns.Circuit.gen_satostop(6301, 7, 6, 70, 8)
    r = check_packed(ctx, c.arrays(), 8)
    assert r["method"] == 0


def test_34_bit_state_space(ctx):
    # n = 34: the packed image (2^34 entries of 35 bits, 75 GiB) and walks from 2^22
    # waypoints; the Feistel map is a permutation with one STOP transition, so every
    # stopping start ends at the target after a distinct number of evaluations
    n, U = 34, 14
    c, target = ns.Circuit.gen_satostop(7501, 17, 6, 100, U)
    ctx.trim()
    try:
        r = ctx.satostop_solve(c, U, method="table")
        rb = ctx.satostop_solve(c, U)  # the default there: walks without a table
    finally:
        ctx.trim()  # 83 GiB of image and chunk tables
    assert (r["method"], r["waypoint_bits"]) == (1, 12) and rb["method"] == 2
    assert np.array_equal(rb["steps_arr"], r["steps_arr"]) and np.array_equal(rb["ends_arr"], r["ends_arr"])
    steps, ends = r["steps_arr"], r["ends_arr"]
    stopped = steps > 0
    assert r["n_stopped"] == int(stopped.sum()) > 0
    assert np.all(ends[stopped] == target)
    assert len(set(steps[stopped].tolist())) == int(stopped.sum())
    lo = int(np.nonzero(stopped)[0][0])
    assert (r["lowest_u"], r["start"], r["steps"], r["end"]) == (lo, lo << (n - U), int(steps[lo]), target)
    oc = oracle.OracleCircuit.from_dict(c.arrays())
    short = 0
    for u in range(1 << U):
        s, e = oc.walk(u << (n - U), 64)
        if s:
            short += 1
            assert (int(steps[u]), int(ends[u])) == (s, e), u
        else:
            assert steps[u] == 0 or steps[u] > 64, u
    assert short >= 1  # the planted start
    for u in sorted(np.nonzero(steps > 64)[0], key=lambda u: steps[u])[:3]:
        if steps[u] <= 1 << 21:
            assert oc.walk(int(u) << (n - U), int(steps[u])) == (int(steps[u]), target)


def test_35_bits_need_the_full_search(ctx):
    # 35 state bits fit only as the packed image, so greedysatostop's first-pass rule and pass
    # limits are refused there
    c35 = mk(point_stop_circuit(35, 5 << 31, 12345))
    with pytest.raises(ns.NaivesatError, match="34 and 35 state bits"):
        ctx.satostop_solve(c35, 4, first_pass=True, arrays=False)
    with pytest.raises(ns.NaivesatError, match="34 and 35 state bits"):
        ctx.satostop_solve(c35, 4, max_passes=20, arrays=False)
