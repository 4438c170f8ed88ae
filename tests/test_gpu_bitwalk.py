"""SATOSTOP's full search without a table: waypoint walks that evaluate the circuit for 32
walks per lane at each step (the generated ns_walk kernel), against the oracle and against
the walks over the table."""
import hashlib
import json
import os

import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import point_stop_circuit, random_circuit

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "configs.json")


def mk(d):
    return ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])


def check(ctx, d, U, **kw):
    r = ctx.satostop_solve(mk(d), U, method="bits", **kw)
    o = oracle.OracleCircuit.from_dict(d).satostop(U, threads=8)
    assert np.array_equal(r["steps_arr"], o["steps"]), (d["n"], U)
    assert np.array_equal(r["ends_arr"], o["ends"]), (d["n"], U)
    assert r["n_stopped"] == o["n_stopped"]
    if o["n_stopped"]:
        lo = o["lowest_u"]
        assert (r["lowest_u"], r["start"], r["steps"], r["end"]) == (lo, lo << (d["n"] - U), o["steps"][lo],
                                                                     o["ends"][lo])
    return r


@pytest.mark.parametrize("n", [2, 5, 8, 11, 14, 17, 20])
def test_random_next_state_circuits(ctx, n):
    rng = np.random.default_rng(10100 + n)
    for _ in range(2):
        d = random_circuit(rng, n, int(rng.integers(n + 1, 4 * n + 20)), nout=n + 1, recent=3 * n + 4)
        r = check(ctx, d, int(rng.integers(0, n + 1)))
        assert r["method"] == 2 and r["build_ms"] == 0


@pytest.mark.parametrize("half", [4, 7, 10, 12])
def test_feistel_workloads(ctx, half):
    c, target = ns.Circuit.gen_satostop(10200 + half, half, 6, 12 * half, min(2 * half, 12))
    r = check(ctx, c.arrays(), min(2 * half, 12))
    assert r["method"] == 2


def test_spacings_and_cycles(ctx, monkeypatch):
    # waypoints from every state up to the starts' spacing; sparse waypoints leave cycles
    # without any, which the walks' Brent tests find
    rng = np.random.default_rng(10300)
    d = random_circuit(rng, 13, 60, nout=14, recent=40)
    for k in (0, 1, 4, 8):
        monkeypatch.setenv("NS_WAYPOINT_BITS", str(k))
        r = check(ctx, d, 5)
        assert (r["method"], r["waypoint_bits"]) == (2, k)


def test_walk_outcomes(ctx, monkeypatch):
    # next(s) = s | 1, STOP = [s == 769]: 768 -> 769 stops at its second evaluation, the other
    # starts reach fixed points that are no waypoints (cycles of length one)
    monkeypatch.setenv("NS_WAYPOINT_BITS", "8")
    n, p = 10, 769
    in0, in1, func = [0], [0], [2]
    acc, pos = 0, True
    for w in range(1, n):
        pw = (p >> w) & 1 == 1
        if pos and pw:
            f, a, b = 0, acc, w
        elif pos:
            f, a, b = 2, acc, w
        elif pw:
            f, a, b = 2, w, acc
        else:
            f, a, b = 1, acc, w
        in0.append(a)
        in1.append(b)
        func.append(f)
        acc, pos = n + len(func) - 1, True
    d = {"n": n, "in0": np.array(in0, np.uint32), "in1": np.array(in1, np.uint32), "func": np.array(func, np.uint8),
         "out_wire": np.array([n] + list(range(1, n)) + [acc], np.uint32),
         "out_neg": np.array([1] + [0] * (n - 1) + [0], np.uint8)}
    r = check(ctx, d, 2)
    assert list(r["steps_arr"]) == [0, 0, 0, 2] and int(r["ends_arr"][3]) == 769


def test_limit_falls_back_to_the_table(ctx, monkeypatch):
    # waypoints 2^8 states apart, walks limited to 3 evaluations: the table answers instead
    monkeypatch.setenv("NS_WAYPOINT_BITS", "8")
    monkeypatch.setenv("NS_WAYPOINT_LIMIT", "3")
    c, _ = ns.Circuit.gen_satostop(10401, 8, 6, 80, 8)
    r = check(ctx, c.arrays(), 8)
    assert r["method"] != 2


def test_mem_image_cases(ctx):
    with open(os.path.join(os.path.dirname(__file__), "golden", "mem_image_find_solution.json")) as f:
        cases = json.load(f)["cases"]
    for case in cases:
        n, U = case["state_len"], case["unknowns"]
        p = case["start"] + case["pos_sol"]
        r = ctx.satostop_solve(mk(point_stop_circuit(n, p, case["next"])), U, arrays=False, method="bits")
        exp = case["expected"]
        if exp is None:
            assert r["n_stopped"] == 0, case
        else:
            assert (r["n_stopped"], r["start"], r["end"], r["steps"]) == (1, exp["start"], exp["end"], 1), case


def test_config3_full_size_golden(ctx):
    if not os.path.exists(GOLDEN) or "cfg3" not in json.load(open(GOLDEN)):
        pytest.skip("cfg3 golden not generated")
    g = json.load(open(GOLDEN))["cfg3"]
    c, target = ns.Circuit.gen_satostop(4, 15, 6, 300, 14)
    r = ctx.satostop_solve(c, 14, wide=False, method="bits")
    assert (r["method"], r["waypoint_bits"]) == (2, 5)
    # every state once, but those on the few short cycles without a waypoint
    assert (1 << 30) - (1 << 12) < r["walk_evaluations"] <= 1 << 30
    assert r["n_stopped"] == g["n_stopped"]
    assert hashlib.sha256(r["steps_arr"].tobytes()).hexdigest() == g["steps_sha256"]
    assert hashlib.sha256(r["ends_arr"].tobytes()).hexdigest() == g["ends_sha256"]


def test_36_bit_state_space(ctx):
    # n = 36: past the packed image; 2^28 waypoints 2^8 states apart. The Feistel map is a
    # permutation with one STOP transition, so every stopping start ends at the target after a
    # distinct number of evaluations; the short ones are checked by walking on the host
    n, U = 36, 14
    c, target = ns.Circuit.gen_satostop(10501, 18, 6, 100, U)
    try:
        r = ctx.satostop_solve(c, U)
    finally:
        ctx.trim()
    assert (r["method"], r["waypoint_bits"]) == (2, 8)
    steps, ends = r["steps_arr"], r["ends_arr"]
    stopped = steps > 0
    assert r["n_stopped"] == int(stopped.sum()) > 0
    assert np.all(ends[stopped] == target)
    assert len(set(steps[stopped].tolist())) == int(stopped.sum())
    oc = oracle.OracleCircuit.from_dict(c.arrays())
    short = 0
    for u in range(1 << U):
        s, e = oc.walk(u << (n - U), 64)
        if s:
            short += 1
            assert (int(steps[u]), int(ends[u])) == (s, e), u
        else:
            assert steps[u] == 0 or steps[u] > 64, u
    assert short >= 1


def test_42_bit_state_space_widest_tier(ctx):
    # n = 42: the widest walk template (48 state bits per slot), 2^28 waypoints 2^14 states
    # apart, 2^42 evaluations; checked like the 36-bit search
    n, U = 42, 14
    c, target = ns.Circuit.gen_satostop(10511, 21, 6, 60, U)
    try:
        r = ctx.satostop_solve(c, U)
    finally:
        ctx.trim()
    assert (r["method"], r["waypoint_bits"], r["backend"]) == (2, 14, 1)
    steps, ends = r["steps_arr"], r["ends_arr"]
    stopped = steps > 0
    assert r["n_stopped"] == int(stopped.sum()) > 0
    assert np.all(ends[stopped] == target)
    assert len(set(steps[stopped].tolist())) == int(stopped.sum())
    oc = oracle.OracleCircuit.from_dict(c.arrays())
    short = 0
    for u in range(1 << U):
        s, e = oc.walk(u << (n - U), 64)
        if s:
            short += 1
            assert (int(steps[u]), int(ends[u])) == (s, e), u
        else:
            assert steps[u] == 0 or steps[u] > 64, u
    assert short >= 1


def test_native_and_compiled_walks_agree(ctx, monkeypatch):
    # the native walk template and the compiled ns_walk kernel: same records, so the same
    # per-start answers and the same evaluation count
    c, _ = ns.Circuit.gen_satostop(10601, 12, 6, 150, 12)
    r1 = check(ctx, c.arrays(), 12)
    assert r1["backend"] == 1 and r1["vgpr_class"] in (256, 512)
    monkeypatch.setenv("NS_BACKEND", "hiprtc")
    r2 = ctx.satostop_solve(c, 12, method="bits")
    assert r2["backend"] == 0
    assert np.array_equal(r1["steps_arr"], r2["steps_arr"]) and np.array_equal(r1["ends_arr"], r2["ends_arr"])
    assert r1["walk_evaluations"] == r2["walk_evaluations"]


@pytest.mark.parametrize("n,bits", [(22, 1), (24, 1), (24, 3)])
def test_chase_levels_on_random_maps(ctx, monkeypatch, n, bits):
    # dense waypoints: 2^(n-bits) records, coarsened by chase passes (one waypoint in 16 per
    # level) before the pointer jumping; random maps run into cycles of fine waypoints
    # without a coarse one, which the chase's Brent test ends. Oracle answers, and the same
    # records as pointer jumping alone
    rng = np.random.default_rng(10700 + n + bits)
    d = random_circuit(rng, n, 4 * n + 40, nout=n + 1, recent=3 * n + 4)
    monkeypatch.setenv("NS_WAYPOINT_BITS", str(bits))
    r = check(ctx, d, 8)
    assert (r["method"], r["waypoint_bits"]) == (2, bits)
    monkeypatch.setenv("NS_WAYPOINT_CHASE", "0")
    r0 = ctx.satostop_solve(mk(d), 8, method="bits")
    assert np.array_equal(r["steps_arr"], r0["steps_arr"]) and np.array_equal(r["ends_arr"], r0["ends_arr"])
    assert r0["waypoint_passes"] != r["waypoint_passes"]
