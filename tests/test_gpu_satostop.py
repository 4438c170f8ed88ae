"""SATOSTOP parity: device table + pointer jumping vs the CPU oracle (per unknown: stops
or not, evaluations until STOP and the end state, bit-exact)."""
import hashlib
import json
import os

import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import identity_circuit, point_stop_circuit, random_circuit

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "configs.json")


def mk(d):
    return ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])


def check(ctx, d, U, **kw):
    # these tests are of the search over the table (its walks and passes); the default full
    # search (walks without a table) has its own in test_gpu_bitwalk.py
    kw.setdefault("method", "table")
    r = ctx.satostop_solve(mk(d), U, **kw)
    o = oracle.OracleCircuit.from_dict(d).satostop(U, threads=8)
    assert np.array_equal(r["steps_arr"], o["steps"]), (d["n"], U)
    assert np.array_equal(r["ends_arr"], o["ends"]), (d["n"], U)
    assert r["n_stopped"] == o["n_stopped"]
    if o["n_stopped"]:
        lo = o["lowest_u"]
        assert r["lowest_u"] == lo
        assert r["start"] == lo << (d["n"] - U)
        assert (r["steps"], r["end"]) == (o["steps"][lo], o["ends"][lo])
    return r, o


@pytest.mark.parametrize("passes", [False, True])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 8, 11, 14, 17, 20])
def test_random_next_state_circuits(ctx, n, passes):
    rng = np.random.default_rng(5000 + n)
    d = random_circuit(rng, n, int(rng.integers(n + 1, 4 * n + 20)), nout=n + 1, recent=3 * n + 4)
    r, o = check(ctx, d, int(rng.integers(0, n + 1)), passes=passes)
    if passes or not r["any_stop"]:
        assert r["method"] == 0


@pytest.mark.parametrize("passes", [False, True])
@pytest.mark.parametrize("half", [4, 6, 8, 10])
def test_feistel_workloads(ctx, half, passes):
    c, target = ns.Circuit.gen_satostop(70 + half, half, 6, 12 * half, min(2 * half, 12))
    d = c.arrays()
    r, o = check(ctx, d, min(2 * half, 12), passes=passes)
    assert r["any_stop"] == 1
    assert r["method"] == (0 if passes else 1)


def test_waypoint_spacings(ctx, monkeypatch):
    # every spacing from one waypoint per state up to the starts' own spacing gives the
    # oracle's answers (sparse waypoints leave cycles without any, found by the walks)
    rng = np.random.default_rng(5100)
    cases = [(random_circuit(rng, 13, 60, nout=14, recent=40), 5),
             (ns.Circuit.gen_satostop(5101, 7, 6, 70, 6)[0].arrays(), 6)]
    for d, U in cases:
        for k in sorted({0, 1, 3, d["n"] - U}):
            monkeypatch.setenv("NS_WAYPOINT_BITS", str(k))
            r, _ = check(ctx, d, U)
            if r["any_stop"]:
                assert (r["method"], r["waypoint_bits"]) == (1, k)


def test_waypoint_walk_limit_falls_back_to_passes(ctx, monkeypatch):
    # walks longer than the limit hand the search to the passes, with the same answers
    c, _ = ns.Circuit.gen_satostop(5201, 8, 6, 80, 8)
    monkeypatch.setenv("NS_WAYPOINT_LIMIT", "3")
    r, _ = check(ctx, c.arrays(), 8)
    assert r["method"] == 0 and r["passes"] > 0


def test_waypoint_walk_outcomes(ctx, monkeypatch):
    # next(s) = s | 1, STOP = [s == 769], n = 10, starts 0, 256, 512, 768 (waypoints 2^8
    # apart): 768 -> 769 stops at its second evaluation; the others reach a fixed point
    # that is no waypoint and never stop
    n, p = 10, 769
    in0, in1, func = [0], [0], [2]  # wire 10 = x0 & ~x0 = 0
    lits = [(i, (p >> i) & 1 == 1) for i in range(n)]
    acc, pos = lits[0]
    for w, pw in lits[1:]:
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
         "out_neg": np.array([1] + [0] * (n - 1) + [0 if pos else 1], np.uint8)}
    monkeypatch.setenv("NS_WAYPOINT_BITS", "8")
    r, o = check(ctx, d, 2)
    assert (r["method"], r["waypoint_bits"]) == (1, 8)
    assert list(o["steps"]) == [0, 0, 0, 2] and int(o["ends"][3]) == 769


def test_first_pass_mode_matches_reference_rule(ctx):
    # greedysatostop ends after the first jump pass p where some start has stopped; the
    # entries then cover 2^p evaluations, so the candidates are starts with steps <= 2^p
    c, _ = ns.Circuit.gen_satostop(91, 9, 6, 100, 12)
    d = c.arrays()
    o = oracle.OracleCircuit.from_dict(d).satostop(12, threads=8)
    r = ctx.satostop_solve(c, 12, first_pass=True)
    steps = o["steps"]
    assert o["n_stopped"] > 0
    p = max(1, int(np.ceil(np.log2(steps[steps > 0].min()))))
    cand = np.nonzero((steps > 0) & (steps <= (1 << p)))[0]
    assert r["passes"] == p
    assert r["lowest_u"] == cand[0]
    assert r["steps"] == steps[cand[0]]


def test_fused_gate_pattern_regression(ctx):
    # nine gates over lane-mask inputs and a constant mask whose compiler-fused form
    # computed a wrong table; each LUT is now one explicit v_bitop3
    c = ns.Circuit.parse("{0 1 2 3 4 5 6 7 8 9 10 11 xor(9,10) nimpl(6,7) xor(13,12) nimpl(12,13) "
                         "nor(8,3) xor(15,14) nimpl(16,15) nimpl(17,16) xor(19,18):0}")
    d = c.arrays()
    t = ctx.table_build(c)
    assert np.array_equal(t, oracle.OracleCircuit.from_dict(d).table_u32())
    r = ctx.sat_solve(c)
    assert (r["count"], r["lowest"]) == oracle.OracleCircuit.from_dict(d).sat(threads=4)


def test_no_stop_anywhere(ctx):
    # STOP output tied to 0: no start stops, no passes run
    n = 12
    d = random_circuit(np.random.default_rng(1), n, 40, nout=n, recent=20)
    d["in0"] = np.append(d["in0"], 0).astype(np.uint32)
    d["in1"] = np.append(d["in1"], 0).astype(np.uint32)
    d["func"] = np.append(d["func"], 2).astype(np.uint8)  # x0 & ~x0 = 0
    d["out_wire"] = np.append(d["out_wire"], n + len(d["func"]) - 1).astype(np.uint32)
    d["out_neg"] = np.append(d["out_neg"], 0).astype(np.uint8)
    r, o = check(ctx, d, 6)
    assert r["any_stop"] == 0 and r["n_stopped"] == 0 and r["passes"] == 0


@pytest.mark.parametrize("m", [5, 20, 24])
def test_table_identity_layout(ctx, m):
    t = ctx.table_build(mk(identity_circuit(m)))
    assert np.array_equal(t, np.arange(1 << m, dtype=np.uint32))


def test_table_matches_oracle(ctx):
    rng = np.random.default_rng(9)
    d = random_circuit(rng, 18, 200, nout=19, recent=40)
    t = ctx.table_build(mk(d))
    assert np.array_equal(t, oracle.OracleCircuit.from_dict(d).table_u32())


def test_config3_full_size_golden(ctx):
    if not os.path.exists(GOLDEN) or "cfg3" not in json.load(open(GOLDEN)):
        pytest.skip("cfg3 golden not generated")
    g = json.load(open(GOLDEN))["cfg3"]
    c, target = ns.Circuit.gen_satostop(4, 15, 6, 300, 14)
    assert target == g["target"]
    for passes in (False, True):
        r = ctx.satostop_solve(c, 14, wide=False, passes=passes, method="table")  # golden hashes: uint32 arrays
        if passes:
            # P passes over the table plus walks of 2^(30-P) jumps cover 2^30 evaluations
            assert r["method"] == 0 and r["walk_jumps"] << r["passes"] == 1 << 30
        else:
            assert (r["method"], r["waypoint_bits"]) == (1, 8) and r["waypoint_passes"] <= 22
        assert r["n_stopped"] == g["n_stopped"]
        assert hashlib.sha256(r["steps_arr"].tobytes()).hexdigest() == g["steps_sha256"]
        assert hashlib.sha256(r["ends_arr"].tobytes()).hexdigest() == g["ends_sha256"]
        if g["lowest_u"] is not None:
            assert (r["lowest_u"], r["steps"], r["end"]) == (g["lowest_u"], g["steps"], g["end"])


def test_mem_image_find_solution_vectors(ctx):
    # the reference's test_mem_image_find_solution cases (src/greedysatostop.rs:1203-1268)
    # as full next-state tables: identity except one STOP entry at start + pos_sol
    with open(os.path.join(os.path.dirname(__file__), "golden", "mem_image_find_solution.json")) as f:
        cases = json.load(f)["cases"]
    for case in cases:
        n, U = case["state_len"], case["unknowns"]
        p = case["start"] + case["pos_sol"]
        r = ctx.satostop_solve(mk(point_stop_circuit(n, p, case["next"])), U, arrays=False)
        exp = case["expected"]
        if exp is None:
            assert r["n_stopped"] == 0, case
        else:
            assert r["n_stopped"] == 1, case
            assert (r["start"], r["end"], r["steps"]) == (exp["start"], exp["end"], 1), case
        # greedysatostop's rule: stop after the first pass in which a start has stopped
        g = ctx.satostop_solve(mk(point_stop_circuit(n, p, case["next"])), U, first_pass=True, arrays=False)
        assert (g["n_stopped"] > 0) == (exp is not None), case
        if exp is not None:
            assert (g["start"], g["end"], g["passes"]) == (exp["start"], exp["end"], 1), case


@pytest.mark.parametrize("passes", [1, 2, 5, 9])
def test_walks_after_fewer_passes(ctx, monkeypatch, passes):
    # P passes over the table and 2^(n-P) jumps per start give the n-pass answers
    monkeypatch.setenv("NS_JUMP_PASSES", str(passes))
    for half, U in ((6, 10), (8, 9)):
        c, _ = ns.Circuit.gen_satostop(120 + half, half, 6, 10 * half, U)
        r, o = check(ctx, c.arrays(), U, passes=True)
        assert r["passes"] == passes and r["walk_jumps"] == 1 << (2 * half - passes)
    rng = np.random.default_rng(3100 + passes)
    d = random_circuit(rng, 15, 70, nout=16, recent=40)
    r, o = check(ctx, d, 11, passes=True)
    assert r["walk_jumps"] == 1 << (15 - passes)


def test_walk_coverage_with_max_passes(ctx, monkeypatch):
    # max_passes = C limits every walk to 2^C evaluations, with or without walks
    c, _ = ns.Circuit.gen_satostop(131, 8, 6, 80, 10)
    full = oracle.OracleCircuit.from_dict(c.arrays()).satostop(10, threads=8)
    for C in (4, 9):
        for P in ("", "2"):
            monkeypatch.setenv("NS_JUMP_PASSES", P)
            r = ctx.satostop_solve(c, 10, max_passes=C)
            within = (full["steps"] > 0) & (full["steps"] <= (1 << C))
            assert np.array_equal(r["steps_arr"], np.where(within, full["steps"], 0)), (C, P)
            assert np.array_equal(r["ends_arr"], np.where(within, full["ends"], 0)), (C, P)


def test_wide_layout_matches_oracle(ctx, monkeypatch):
    # the code path of state spaces over 31 bits (split table, wide jump entries, 64-bit
    # results) taken for small n, against the oracle
    monkeypatch.setenv("NS_WIDE_TABLE", "1")
    rng = np.random.default_rng(7100)
    for n in (3, 8, 14, 19):
        d = random_circuit(rng, n, int(rng.integers(n + 1, 4 * n + 20)), nout=n + 1, recent=3 * n + 4)
        check(ctx, d, int(rng.integers(0, n + 1)))
    for half in (6, 8):
        c, _ = ns.Circuit.gen_satostop(7200 + half, half, 6, 10 * half, 12)
        r, _ = check(ctx, c.arrays(), 12)
        assert r["method"] == 1
        check(ctx, c.arrays(), 12, passes=True)
        monkeypatch.setenv("NS_JUMP_PASSES", "3")
        check(ctx, c.arrays(), 12, passes=True)
        monkeypatch.delenv("NS_JUMP_PASSES")


def test_u32_and_u64_entry_points_agree(ctx):
    c, _ = ns.Circuit.gen_satostop(7301, 8, 6, 70, 11)
    a = ctx.satostop_solve(c, 11, wide=False)
    b = ctx.satostop_solve(c, 11)
    assert a["steps_arr"].dtype == np.uint32 and b["steps_arr"].dtype == np.uint64
    assert np.array_equal(a["steps_arr"], b["steps_arr"]) and np.array_equal(a["ends_arr"], b["ends_arr"])
    assert (a["n_stopped"], a["lowest_u"], a["steps"], a["end"]) == (b["n_stopped"], b["lowest_u"], b["steps"],
                                                                     b["end"])


def test_32_bit_state_space(ctx, monkeypatch):
    # n = 32: a 2^32-entry table, walked as the packed 33-bit image (17 GiB), then by the
    # passes (16 GiB per output half, 32 GiB per jump table). The Feistel map is a permutation with one STOP transition (into the planted target), so
    # every stopping start ends at the target after a distinct number of evaluations; the
    # starts that stop within 64 evaluations are found independently by walking on the host
    n, U = 32, 14
    c, target = ns.Circuit.gen_satostop(7401, 16, 6, 100, U)
    ctx.trim()
    r = ctx.satostop_solve(c, U, method="table")
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
    # the few stopping starts with the smallest counts above 64, walked on the host
    for u in sorted(np.nonzero(steps > 64)[0], key=lambda u: steps[u])[:3]:
        if steps[u] <= 1 << 21:
            assert oc.walk(int(u) << (n - U), int(steps[u])) == (int(steps[u]), target)
    # another split between passes and walks gives the same arrays
    assert (r["method"], r["waypoint_bits"]) == (1, 10)
    # the passes give the same arrays, also with another split between passes and walks
    r2 = ctx.satostop_solve(c, U, passes=True)
    assert r2["method"] == 0
    assert np.array_equal(r2["steps_arr"], steps) and np.array_equal(r2["ends_arr"], ends)
    monkeypatch.setenv("NS_JUMP_PASSES", "12")
    try:
        r2 = ctx.satostop_solve(c, U, passes=True)
    finally:
        ctx.trim()  # 114 GiB of tables
    assert r2["passes"] == 12
    assert np.array_equal(r2["steps_arr"], steps) and np.array_equal(r2["ends_arr"], ends)
