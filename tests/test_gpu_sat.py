"""SAT parity: HIP engine vs the CPU oracle (count and lowest assignment, bit-exact)."""
import json
import os

import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import random_circuit

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "configs.json")


def check(ctx, d, **kw):
    c = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])
    r = ctx.sat_solve(c, **kw)
    cnt, low = oracle.OracleCircuit.from_dict(d).sat(threads=8)
    assert r["count"] == cnt, (d["n"], kw)
    assert bool(r["found"]) == (cnt > 0)
    if cnt:
        assert r["lowest"] == low, (d["n"], kw)
    return r


@pytest.mark.parametrize("n", list(range(1, 21)))
def test_random_circuits_all_sizes(ctx, n):
    rng = np.random.default_rng(1000 + n)
    d = random_circuit(rng, n, int(rng.integers(1, 120)), recent=int(rng.integers(4, 40)))
    check(ctx, d)


@pytest.mark.parametrize("group_len,wpl", [(64, 1), (128, 2), (256, 4), (512, 8), (1024, 1), (64, 16), (256, 32)])
def test_launch_geometries(ctx, group_len, wpl):
    rng = np.random.default_rng(77)
    d = random_circuit(rng, 22, 150, recent=30)
    check(ctx, d, group_len=group_len, words_per_lane=wpl)


def test_many_launch_ranges(ctx):
    rng = np.random.default_rng(5)
    d = random_circuit(rng, 21, 90, recent=25)
    r = check(ctx, d, elem_bits=14)
    assert r["launches"] > 1


def test_stop_at_first_returns_global_lowest(ctx):
    c, key = ns.Circuit.gen_sat(21, 26, 600, 20)
    d = c.arrays()
    r = ctx.sat_solve(c, elem_bits=20, stop_at_first=True)
    full = ctx.sat_solve(c)
    assert r["found"] and r["lowest"] == full["lowest"] <= key
    assert r["count"] <= full["count"]
    lo = full["lowest"]
    # the oracle checks the words up to the lowest hit only
    cnt, low = oracle.OracleCircuit.from_dict(d).sat_range(0, (lo >> 6) + 1, threads=8)
    assert low == lo


def test_unsatisfiable_and_tautology(ctx):
    # x0 & ~x0 (never) and x0 | ~x0 (always) on 20 inputs
    never = ns.Circuit.from_arrays(20, [0], [0], [2], [20], [0])
    always = ns.Circuit.from_arrays(20, [0], [0], [2], [20], [1])
    r = ctx.sat_solve(never)
    assert r["count"] == 0 and not r["found"]
    r = ctx.sat_solve(always)
    assert r["count"] == 1 << 20 and r["lowest"] == 0


def test_output_is_an_input(ctx):
    c = ns.Circuit.from_arrays(17, [], [], [], [16], [0])
    r = ctx.sat_solve(c)
    assert r["count"] == 1 << 16 and r["lowest"] == 1 << 16


def test_planted_key_circuits(ctx):
    for seed in range(3):
        c, key = ns.Circuit.gen_sat(50 + seed, 24, 1200, 24)
        r = ctx.sat_solve(c)
        assert r["found"] and r["lowest"] <= key
        cnt, low = oracle.OracleCircuit.from_dict(c.arrays()).sat(threads=8)
        assert (r["count"], r["lowest"]) == (cnt, low)


def golden(name):
    if not os.path.exists(GOLDEN):
        pytest.skip("configs.json not generated")
    g = json.load(open(GOLDEN))
    if name not in g:
        pytest.skip(f"{name} golden not generated")
    return g[name]


def test_config0_golden_and_oracle(ctx):
    g = golden("cfg0")
    c, key = ns.Circuit.gen_sat(1, 24, 2000, 32)
    assert key == g["key"]
    r = ctx.sat_solve(c)
    assert (r["count"], r["lowest"]) == (g["count"], g["lowest"])
    cnt, low = oracle.OracleCircuit.from_dict(c.arrays()).sat(threads=8)
    assert (cnt, low) == (g["count"], g["lowest"])


def test_config1_full_size_golden(ctx):
    g = golden("cfg1")
    c, key = ns.Circuit.gen_sat(2, 36, 20000, 40)
    assert key == g["key"]
    r = ctx.sat_solve(c)
    assert r["complete"] == 1
    assert (r["count"], r["lowest"]) == (g["count"], g["lowest"])
    # the planted key satisfies, so it bounds the lowest answer
    assert r["lowest"] <= key


def test_words_per_lane_tuning_keeps_answers(ctx):
    # 64 launches of 2^15 words: the first solve times words-per-lane 1..16 on its first
    # launches, later solves reuse the choice; every setting must give the same answers
    c, key = ns.Circuit.gen_sat(21, 26, 800, 26)
    count, lowest = oracle.OracleCircuit.from_dict(c.arrays()).sat(threads=8)
    for kw in ({}, {}, {"words_per_lane": 2}, {"words_per_lane": 16}):
        r = ctx.sat_solve(c, elem_bits=20, **kw)
        assert (r["count"], r["lowest"]) == (count, lowest), kw
        assert r["launches"] == 64


def test_graph_replay_matches_eager(ctx):
    # the first solve runs eagerly, the second captures a HIP graph, later ones replay it;
    # other work between the replays must not disturb them
    c, key = ns.Circuit.gen_sat(31, 27, 700, 27)
    other, _ = ns.Circuit.gen_sat(32, 22, 300, 22)
    count, lowest = oracle.OracleCircuit.from_dict(c.arrays()).sat(threads=8)
    ocount, olowest = oracle.OracleCircuit.from_dict(other.arrays()).sat(threads=8)
    for i in range(5):
        r = ctx.sat_solve(c, elem_bits=24)
        assert (r["count"], r["lowest"]) == (count, lowest), i
        assert r["launches"] == 8  # 2^22 words in launches of 2^19
        ro = ctx.sat_solve(other)
        assert (ro["count"], ro["lowest"]) == (ocount, olowest), i


def test_more_circuits_than_the_kernel_cache(ctx):
    # 300 distinct circuits against a cache of 256 modules: every answer is the oracle's, and
    # the first circuit solved again after its kernel was dropped is built and answered anew
    cs = [ns.Circuit.gen_sat(7000 + i, 16, 60, 16)[0] for i in range(300)]
    want = [oracle.OracleCircuit.from_dict(c.arrays()).sat(threads=8) for c in cs]
    for i, (c, (cnt, low)) in enumerate(zip(cs, want)):
        r = ctx.sat_solve(c)
        assert r["count"] == cnt, i
        if cnt:
            assert r["lowest"] == low, i
    r = ctx.sat_solve(cs[0])
    assert r["compile_ms"] > 0.0
    assert r["count"] == want[0][0]


def test_dense_solutions_count_exactly(ctx):
    # most waves of these searches hit: their counts spread over the device's count slots
    # must add up exactly, on the eager, tuning and graph-replay paths alike
    n = 30
    xor = ns.Circuit.from_arrays(n, [0], [1], [3], [n], [0])  # x0 ^ x1: 2^29 hits, lowest 1
    conj = ns.Circuit.from_arrays(n, [0], [1], [0], [n], [0])  # x0 & x1: 2^28 hits, lowest 3
    for c, count, lowest in ((xor, 1 << 29, 1), (conj, 1 << 28, 3)):
        for i in range(4):
            r = ctx.sat_solve(c, elem_bits=26)
            assert (r["count"], r["lowest"]) == (count, lowest), i
    # a random circuit with about half of its assignments satisfying, against the oracle
    rng = np.random.default_rng(77)
    d = random_circuit(rng, 24, 120, recent=30)
    d["in0"] = np.append(d["in0"], d["out_wire"][0]).astype(np.uint32)
    d["in1"] = np.append(d["in1"], 0).astype(np.uint32)
    d["func"] = np.append(d["func"], 3).astype(np.uint8)
    d["out_wire"] = np.array([24 + len(d["func"]) - 1], np.uint32)
    check(ctx, d)


@pytest.mark.parametrize("K", [2, 3, 5])
def test_output_groups_keep_answers(ctx, monkeypatch, K):
    # a conjunction output evaluated as K groups of its literals (the shared gates recomputed
    # per group) gives the oracle's count and lowest hit, for SAT and QSAT kernels
    monkeypatch.setenv("NS_OUTPUT_GROUPS", str(K))
    for seed in range(3):
        c, key = ns.Circuit.gen_sat(4100 + 10 * K + seed, 22, 500, 16)
        body, cls, _ = c.native_body(0)
        r = ctx.sat_solve(c)
        assert (r["count"], r["lowest"]) == oracle.OracleCircuit.from_dict(c.arrays()).sat(threads=8), seed
    rng = np.random.default_rng(4200 + K)
    d = random_circuit(rng, 18, 120, recent=30)
    # AND the output with three more wires (one negated) so it is a conjunction of 4+ literals
    g = len(d["func"])
    w = [int(d["out_wire"][0]), 18 + g - 5, 18 + g - 9, 18 + g - 13]
    d["in0"] = np.append(d["in0"], [w[0], 18 + g, 18 + g + 1]).astype(np.uint32)
    d["in1"] = np.append(d["in1"], [w[1], w[2], w[3]]).astype(np.uint32)
    d["func"] = np.append(d["func"], [0, 2, 0]).astype(np.uint8)
    d["out_wire"] = np.array([18 + g + 2], np.uint32)
    check(ctx, d)
    c, key = ns.Circuit.gen_sat(4300 + K, 16, 300, 12)
    q = [0] * 4 + [1] * 4 + [0] * 4 + [1] * 4
    o = oracle.OracleCircuit.from_dict(c.arrays()).qsat(q)
    r = ctx.qsat_solve(c, q)
    assert (r["truth"], r["witness_count"]) == (o["truth"], o["witness_count"])


def test_groups_for_a_circuit_no_class_holds(ctx):
    # 40,000 gates with a 48-gate word: ~800 values live at once fit no register class as one
    # group, so the native body takes output groups (refined) instead of the compiler path
    c, key = ns.Circuit.gen_sat(2, 26, 40000, 48)
    body, cls, _ = c.native_body(0)
    assert cls in (256, 512)
    r = ctx.sat_solve(c)
    assert r["backend"] == 1 and r["vgpr_class"] == cls
    assert (r["count"], r["lowest"]) == oracle.OracleCircuit.from_dict(c.arrays()).sat(threads=8)
    assert r["lowest"] <= key


def test_group_exits_config1_golden(ctx):
    """NS_SAT_GROUP_EXITS on the full 2^36 search: waves leave a word once the output groups
    so far are zero in all lanes; count and lowest stay the golden ones."""
    g = golden("cfg1")
    c, key = ns.Circuit.gen_sat(2, 36, 20000, 40)
    r = ctx.sat_solve(c, group_exits=True)
    assert r["complete"] == 1 and r["backend"] == 1
    assert (r["count"], r["lowest"]) == (g["count"], g["lowest"])


def test_group_exits_with_a_negated_output(ctx):
    # ~(conjunction of the literals): almost every assignment satisfies; no group exits
    c, _ = ns.Circuit.gen_sat(2, 26, 40000, 48)
    d = c.arrays()
    neg = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], 1 - d["out_neg"])
    want = oracle.OracleCircuit.from_dict(neg.arrays()).sat(threads=8)
    r = ctx.sat_solve(neg, group_exits=True)
    assert (r["count"], r["lowest"]) == want


@pytest.mark.parametrize("groups", ["", "3"])
def test_group_exits_keep_answers(ctx, monkeypatch, groups):
    """Group exits against the oracle: a 40,000-gate circuit whose output takes groups
    (48-bit word, few solutions: most waves leave early) and, with three groups forced, one
    with a 12-bit word (many solutions: waves that hit run every group)."""
    if groups:
        monkeypatch.setenv("NS_OUTPUT_GROUPS", groups)
    for args in ((2, 26, 40000, 48), (3, 26, 40000, 12)):
        c, key = ns.Circuit.gen_sat(*args)
        body, cls, _ = c.native_body(7)
        if "s_cbranch_vccz" not in body:
            continue  # no groups for this circuit here: nothing to exit from
        want = oracle.OracleCircuit.from_dict(c.arrays()).sat(threads=8)
        r = ctx.sat_solve(c, group_exits=True)
        assert (r["count"], r["lowest"]) == want, args
        assert r["lowest"] <= key
        r = ctx.sat_solve(c, group_exits=True, words_per_lane=1)  # also one word per lane
        assert (r["count"], r["lowest"]) == want, args
