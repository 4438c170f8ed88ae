"""The four command-line tools (the process boundary the reference's users script against)
print the reference's lines with the oracle's answers."""
import json
import os
import subprocess

import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import point_stop_circuit

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "configs.json")


def run(tool, *args):
    r = subprocess.run([os.path.join(ns.BIN_DIR, tool), *map(str, args)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def mk(d):
    return ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])


def test_naivesat_config0(tmp_path):
    c, key = ns.Circuit.gen_sat(1, 24, 2000, 32)
    f = tmp_path / "cfg0.circuit"
    f.write_text(c.text())
    with open(GOLDEN) as g:
        lowest = json.load(g)["cfg0"]["lowest"]
    out = run("naivesat", "-t", "opencl:0", f)
    assert f"Found Input: {lowest:024b}" in out
    assert "Start execution" in out and any(x.startswith("Time: ") for x in out)


def test_naivesat_group_exits_and_full_evaluation(tmp_path):
    # a circuit whose output takes groups: the default (group exits) and -X (every gate
    # evaluated) print the oracle's count and lowest
    c, key = ns.Circuit.gen_sat(2, 26, 40000, 48)
    f = tmp_path / "big.circuit"
    f.write_text(c.text())
    count, lowest = oracle.OracleCircuit.from_dict(c.arrays()).sat(threads=8)
    for extra in ([], ["-X"]):
        out = run("naivesat", "-t", "opencl:0", "-c", *extra, f)
        assert f"Count: {count}" in out, extra
        assert f"Found Input: {lowest:026b}" in out, extra


# The reference's published runs (benchmarks.txt:3-19) print this for prime_69690421.circuit
# (54 inputs); that file is not in the reference, so the same search is rebuilt here as
# [a * b == 69,690,421], a != 1, b != 1 over a = inputs 0..26, b = inputs 27..53.
REFERENCE_FACTOR_ANSWER = "000000000000000110111010101000000000000100110011100001"


@pytest.mark.parametrize("extra", [[], ["-X"]])
def test_naivesat_prints_the_reference_factoring_answer(tmp_path, extra):
    c = ns.Circuit.gen_factor(69690421, 27)
    f = tmp_path / "prime_69690421.circuit"
    f.write_text(c.text())
    out = run("naivesat", "-e", "32", "-t", "opencl:0", *extra, f)
    assert f"Found Input: {REFERENCE_FACTOR_ANSWER}" in out, out[-3:]
    x = int(REFERENCE_FACTOR_ANSWER, 2)
    assert (x & ((1 << 27) - 1)) * (x >> 27) == 69690421


def test_naivesat_small_and_unsatisfiable(tmp_path):
    f = tmp_path / "small.circuit"
    # x0 & ~x1 & x2: lowest satisfying assignment is 0b101
    f.write_text(ns.Circuit.from_arrays(3, [0, 3], [1, 2], [2, 0], [4], [0]).text())
    assert "Found Input: 101" in run("naivesat", "-t", "opencl:0", f)
    u = tmp_path / "unsat.circuit"
    u.write_text(ns.Circuit.from_arrays(2, [0, 0, 2], [1, 1, 3], [0, 1, 0], [4], [0]).text())
    assert "Unsatisfiable!" in run("naivesat", "-t", "opencl:0", u)


def test_naiveqsat_prefix_and_plain_circuit(tmp_path):
    c = ns.Circuit.gen_qsat(12, 20, 400, 60)
    q = [0] * 5 + [1] * 5 + [0] * 5 + [1] * 5
    f = tmp_path / "q.circuit"
    f.write_text("".join("ea"[x] for x in q) + " " + c.text())
    o = oracle.OracleCircuit.from_dict(c.arrays()).qsat(q)
    want = ns.qsat_result_string({"first_bits": o["first_bits"], "first_all": o["first_all"],
                                  "has_witness": int(o["witness"] is not None), "witness": o["witness"] or 0})
    assert f"Result: {want}" in run("naiveqsat", "-t", "opencl:0", f)
    # a circuit without a prefix is read with every input existential
    p = tmp_path / "plain.circuit"
    p.write_text(c.text())
    out = run("naiveqsat", "-t", "opencl:0", p)
    assert "Load as circuit with quantifiers" in out
    count, lowest = oracle.OracleCircuit.from_dict(c.arrays()).sat(threads=8)
    assert (count > 0) == any(x.startswith("Result: (Satisfiable: ") for x in out)


def test_satostop_tools(tmp_path):
    # the reference's mem-image case 2 (src/greedysatostop.rs:1203-1268) as a full table
    n, U, p, nxt = 24, 11, (3 << 20) + 79 * 8192, 58591
    f = tmp_path / "stop.circuit"
    f.write_text(mk(point_stop_circuit(n, p, nxt)).text())
    out = run("naivesatostop", "-t", "opencl:0", f, U)
    assert f"Solution: Solution {{ start: {p}, end: {nxt}, steps: 1 }}" in out
    out = run("naivesatostop", "-t", "cpu", "-S", f, U)
    assert f"Stopped at {p >> (n - U):0{U}d} {nxt:0{n}d}" in out
    out = run("greedysatostop", f, U)
    assert f"Solution: Solution {{ start: {p}, end: {nxt} }}" in out
    # a start state that never stops
    g = tmp_path / "nostop.circuit"
    g.write_text(mk(point_stop_circuit(n, p + 1, nxt)).text())
    assert "Unsatisfiable!" in run("naivesatostop", "-t", "opencl:0", g, U)
    assert "Unsatisfiable!" in run("greedysatostop", g, U)


def test_naivesatostop_verify_walk(tmp_path):
    c, target = ns.Circuit.gen_satostop(6, 8, 6, 40, 8)
    f = tmp_path / "f.circuit"
    f.write_text(c.text())
    out = run("naivesatostop", "-t", "opencl:0", "-v", f, 8)
    r = oracle.OracleCircuit.from_dict(c.arrays()).satostop(8)
    if r["n_stopped"]:
        lo = r["lowest_u"]
        assert (f"Solution: Solution {{ start: {lo << 8}, end: {r['ends'][lo]}, steps: {r['steps'][lo]} }}"
                in out)
        assert f"Verified: end={r['ends'][lo]} steps={r['steps'][lo]}" in out and "INCORRECT" not in out
    else:
        assert "Unsatisfiable!" in out


def test_naivesatostop_hashmap_search(tmp_path):
    # 34 state bits: the reference's hash-map search (-M; without it the packed table
    # answers), next-state outputs in two 32-bit words; unknown starts are multiples of 2^32
    n, U, nxt = 34, 2, 0x2_1234_5678
    p = 1 << (n - U)
    f = tmp_path / "stop34.circuit"
    f.write_text(mk(point_stop_circuit(n, p, nxt)).text())
    out = run("naivesatostop", "-t", "opencl:0", "-M", "-e", 30, "-H", 20, "-v", f, U)
    assert "HashMap length: 33554432 bytes, bits=20" in out and "Unknown fill bits: 0" in out
    assert f"Solution: Solution {{ start: {p}, end: {nxt}, steps: 1 }}" in out
    assert "Verified: end=%d steps=1" % nxt in out and "INCORRECT" not in out
    out = run("naivesatostop", "-t", "opencl:0", f, U)
    assert "HashMap length" not in "".join(out)
    assert f"Solution: Solution {{ start: {p}, end: {nxt}, steps: 1 }}" in out
    # STOP at a state that is no unknown start: every unknown start is a fixed point
    g = tmp_path / "nostop34.circuit"
    g.write_text(mk(point_stop_circuit(n, p + 1, nxt)).text())
    assert "Unsatisfiable!" in run("naivesatostop", "-t", "opencl:0", "-M", "-e", 30, "-H", 20, g, U)
    # 36 state bits: past the table, the walks without one answer; -M runs the hash map
    n, U, nxt = 36, 2, 0x9_1234_5678
    p = 2 << (n - U)
    h36 = tmp_path / "stop36.circuit"
    h36.write_text(mk(point_stop_circuit(n, p, nxt)).text())
    out = run("naivesatostop", "-t", "opencl:0", h36, U)
    assert "HashMap length" not in "".join(out)
    assert f"Solution: Solution {{ start: {p}, end: {nxt}, steps: 1 }}" in out
    out = run("naivesatostop", "-t", "opencl:0", "-M", "-e", 30, "-H", 20, h36, U)
    assert "HashMap length: 33554432 bytes, bits=20" in out
    assert f"Solution: Solution {{ start: {p}, end: {nxt}, steps: 1 }}" in out
    # --hashmap-search at 24 state bits: the reference's mem-image case 2 again
    n, U, p, nxt = 24, 11, (3 << 20) + 79 * 8192, 58591
    h = tmp_path / "stop24.circuit"
    h.write_text(mk(point_stop_circuit(n, p, nxt)).text())
    out = run("naivesatostop", "-t", "opencl:0", "-M", "-e", 16, "-H", 18, h, U)
    assert "Unknown fill bits: 0" in out
    assert f"Solution: Solution {{ start: {p}, end: {nxt}, steps: 1 }}" in out


def test_greedysatostop_partitions(tmp_path):
    # -p: the reference's partitioned search with the image on disk under -x
    n, U, p, nxt = 24, 11, (3 << 20) + 79 * 8192, 58591
    f = tmp_path / "stop.circuit"
    f.write_text(mk(point_stop_circuit(n, p, nxt)).text())
    img = tmp_path / "img"
    img.mkdir()
    out = run("greedysatostop", "-p", 8, "-x", str(img) + "/", "-v", f, U)
    assert "Load partition bits: 5" in out and "Stage 0 / 24" in out
    assert f"Solution: Solution {{ start: {p}, end: {nxt} }}" in out
    assert f"Verified: end={nxt}" in out and "INCORRECT" not in out
    assert not any(img.iterdir())  # the image file is removed afterwards
    g = tmp_path / "nostop.circuit"
    g.write_text(mk(point_stop_circuit(n, p + 1, nxt)).text())
    assert "Unsatisfiable!" in run("greedysatostop", "-p", 4, "-x", str(img) + "/", g, U)
