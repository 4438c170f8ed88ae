"""naivesatostop's hash-map search on the device: each step against the reference's data
sets, and whole searches whose answers are checked by walking the circuit and against the
device table."""
import json
import os

import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from hashmap_cases import hashmap, load, u32_array
from helpers import point_stop_circuit

pytestmark = pytest.mark.gpu


def mk(d):
    return ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])


def sol_array(resolved=0):
    s = np.zeros(5, np.uint64)  # resolved_unknowns, start, end, steps, defined
    s[0] = resolved
    return s


def sol_dict(s):
    return {"start": int(s[1]), "end": int(s[2]), "steps": int(s[3])} if int(s[4]) & 0xFFFFFFFF else None


def test_join_to_hashmap_reference_data(ctx):
    for case in load()["join_to_hashmap"]:
        hm = hashmap(case["hashmap"])
        fills = np.zeros(1, np.uint32)
        ctx.hashmap_step(0, case["state_len"], case["arg_bit_place"], case["arg"], 0, 0, 0, False,
                         u32_array(case["outputs"]), hm, None, fills, sol_array())
        assert np.array_equal(hm, hashmap(case["expected_hashmap"])), case["state_len"]


def test_join_hashmap_itself_reference_data(ctx):
    case = load()["join_hashmap_itself"]
    hm = hashmap(case["hashmap"])
    out = np.zeros_like(hm)
    out[:] = (1115774, 494022, 450589390239, 489481, 8459021)  # every field is overwritten
    fills = u32_array(case["unknown_fills"])
    sol = sol_array(case["resolved_unknowns"])
    ctx.hashmap_step(1, case["state_len"], 0, 0, case["unknown_bits"], case["unknown_fill_bits"], 0, False, None,
                     hm, out, fills, sol)
    assert np.array_equal(out, hashmap(case["expected_hashmap"]))
    assert np.array_equal(fills, u32_array(case["expected_unknown_fills"]))
    assert int(sol[0]) == case["expected_resolved_unknowns"]
    assert sol_dict(sol) == case["expected_solution"]


def test_fused_join_equals_join_to_then_join_itself(ctx):
    # the search's one-pass join against the oracle's two steps on the reference's join_to
    # data sets, with every unknown start (the low half of the state bits zero) resolving
    for case in load()["join_to_hashmap"]:
        sl, U = case["state_len"], case["state_len"] // 2
        outputs = u32_array(case["outputs"])
        hm = hashmap(case["hashmap"])
        mid = hm.copy()
        oracle.hm_join_to(sl, case["arg_bit_place"], case["arg"], outputs, mid)
        want = np.zeros_like(hm)
        st = oracle.HashMapState(np.zeros(1 << U, np.uint32))
        oracle.hm_join_itself(sl, mid, want, U, 0, st)
        out = np.zeros_like(hm)
        fills = np.zeros(1 << U, np.uint32)
        sol = sol_array()
        ctx.hashmap_step(3, sl, case["arg_bit_place"], case["arg"], U, 0, 0, False, outputs, hm, out, fills, sol)
        assert np.array_equal(out, want), sl
        assert np.array_equal(fills, st.unknown_fills) and int(sol[0]) == st.resolved.value
        assert (sol_dict(sol) is None) == (st.solution() is None)
        # the device's own two steps agree
        two = hm.copy()
        ctx.hashmap_step(0, sl, case["arg_bit_place"], case["arg"], U, 0, 0, False, outputs, two, None,
                         np.zeros(1 << U, np.uint32), sol_array())
        out2 = np.zeros_like(hm)
        ctx.hashmap_step(1, sl, 0, 0, U, 0, 0, False, None, two, out2, np.zeros(1 << U, np.uint32), sol_array())
        assert np.array_equal(out2, want), sl


def test_add_to_hashmap_reference_data(ctx):
    for case in load()["add_to_hashmap"]:
        hm = hashmap(case["hashmap"])
        fills = u32_array(case["unknown_fills"])
        sol = sol_array(case["resolved_unknowns"])
        ctx.hashmap_step(2, case["state_len"], case["arg_bit_place"], case["arg"], case["unknown_bits"],
                         case["unknown_fill_bits"], case["max_predecessors"], True, u32_array(case["outputs"]), hm,
                         None, fills, sol)
        assert np.array_equal(hm, hashmap(case["expected_hashmap"])), case["state_len"]
        assert np.array_equal(fills, u32_array(case["expected_unknown_fills"]))
        assert int(sol[0]) == case["expected_resolved_unknowns"]
        assert sol_dict(sol) == case["expected_solution"]


@pytest.mark.parametrize("half,U,seed", [(9, 9, 8101), (10, 10, 8102), (12, 8, 8103)])
def test_search_finds_a_verified_solution(ctx, half, U, seed):
    n = 2 * half
    c, target = ns.Circuit.gen_satostop(seed, half, 6, 8 * half, U)
    table = ctx.satostop_solve(c, U)
    assert table["n_stopped"] > 0  # the generator plants a stopping start
    r = ctx.satostop_hashmap_solve(c, U, hashmap_len_bits=n - 2, elem_bits=n - 6, max_sweeps=64)
    assert r["found"] == 1, r
    start, steps, end = r["start"], r["steps"], r["end"]
    assert start & ((1 << (n - U)) - 1) == 0  # an unknown's start state
    u = start >> (n - U)
    # the same start, evaluation count and end state as the table and as a walk
    assert (int(table["steps_arr"][u]), int(table["ends_arr"][u])) == (steps, end)
    assert oracle.OracleCircuit.from_dict(c.arrays()).walk(start, steps) == (steps, end)


def test_search_mem_image_cases(ctx):
    # the reference's test_mem_image_find_solution cases (src/greedysatostop.rs:1203-1268) as
    # next-state circuits: identity except one STOP entry at start + pos_sol. Without a
    # stopping unknown start every unknown start is a fixed point, so all resolve as looping.
    with open(os.path.join(os.path.dirname(__file__), "golden", "mem_image_find_solution.json")) as f:
        cases = json.load(f)["cases"]
    for case in cases:
        n, U = case["state_len"], case["unknowns"]
        c = mk(point_stop_circuit(n, case["start"] + case["pos_sol"], case["next"]))
        r = ctx.satostop_hashmap_solve(c, U, hashmap_len_bits=18, elem_bits=16, max_sweeps=200)
        exp = case["expected"]
        if exp is None:
            assert r["found"] == 0 and r["resolved_unknowns"] == 1 << (U - max(0, U - 17)), (case, r)
        else:
            assert r["found"] == 1, (case, r)
            assert (r["start"], r["end"], r["steps"]) == (exp["start"], exp["end"], 1), case


def test_search_progress_and_cancel(ctx):
    # per-block progress (the reference's "Step: arg / arg_steps" callback): blocks in order
    # within each sweep, the same answer as without a callback; a stop ends the search
    c, _ = ns.Circuit.gen_satostop(8102, 10, 6, 80, 10)
    plain = ctx.satostop_hashmap_solve(c, 10, hashmap_len_bits=18, elem_bits=14, max_sweeps=64)
    calls = []
    r = ctx.satostop_hashmap_solve(c, 10, hashmap_len_bits=18, elem_bits=14, max_sweeps=64,
                                   progress=lambda s, n, ru: calls.append((s, n, ru)) and False)
    # (which stopping start is found may depend on the order of the device's work: each
    # answer is checked by walking the circuit)
    assert r["found"] == plain["found"] == 1
    assert oracle.OracleCircuit.from_dict(c.arrays()).walk(r["start"], r["steps"]) == (r["steps"], r["end"])
    assert len(calls) == r["blocks"] and all(n == 1 << 6 for _, n, _ in calls)
    assert [s for s, _, _ in calls] == [b % 64 for b in range(len(calls))]
    with pytest.raises(ns.NaivesatError) as e:
        ctx.satostop_hashmap_solve(c, 10, hashmap_len_bits=18, elem_bits=14, max_sweeps=64,
                                   progress=lambda s, n, ru: s == 3)
    assert e.value.status == ns.NS_ERR_CANCELLED
    tool = os.path.join(ns.BIN_DIR, "naivesatostop")
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "s.circuit")
        open(f, "w").write(c.text())
        out = subprocess.run([tool, "-t", "opencl:0", "-M", "-H", "18", "-e", "14", f, "10"], capture_output=True,
                             text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    steps = [x for x in lines if x.startswith("Step: ")]
    assert steps[:3] == ["Step: 0 / 64", "Step: 1 / 64", "Step: 2 / 64"]
    assert sum(1 for x in lines if x.startswith("Resolved unknowns entries: ")) == len(steps)
