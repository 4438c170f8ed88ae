"""Per-range progress and cancellation (ns_sat_solve_progress / ns_qsat_solve_progress, the
reference's per-arg callback and stop predicate): the ranges are reported in order, a
requested range size is kept exactly, answers equal the plain solve's, and a stopped search's
answer is the oracle's over the ranges it searched."""
import os
import signal
import subprocess

import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import random_circuit

pytestmark = pytest.mark.gpu


def mk(d):
    return ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])


def dense_circuit(seed, n, gates=300):
    # a circuit with many satisfying assignments spread over the space
    rng = np.random.default_rng(seed)
    while True:
        d = random_circuit(rng, n, gates, recent=30)
        cnt, _ = oracle.OracleCircuit.from_dict(d).sat_range(0, 1 << 12, threads=8)
        if 0 < cnt < (1 << 17):
            return d


@pytest.mark.parametrize("elem_bits", [13, 18, 20, 24])
def test_ranges_reported_in_order_and_sized_exactly(ctx, elem_bits):
    d = dense_circuit(3, 26)
    c = mk(d)
    calls = []
    r = ctx.sat_solve_progress(c, lambda s, n, b: calls.append((s, n, b)) and False, elem_bits=elem_bits)
    steps = 1 << (26 - elem_bits)
    assert calls[0] == (ns.SAT_STEP_START, steps, elem_bits)
    assert [x[0] for x in calls[1:]] == list(range(steps))
    assert all(x[1:] == (steps, elem_bits) for x in calls)
    assert r["launches"] == steps and r["complete"] == 1
    plain = ctx.sat_solve(c, elem_bits=elem_bits)
    cnt, low = oracle.OracleCircuit.from_dict(d).sat(threads=8)
    assert (r["count"], r["lowest"]) == (plain["count"], plain["lowest"]) == (cnt, low)


def test_stop_at_first_reports_until_the_hit(ctx):
    d = dense_circuit(5, 28)
    c = mk(d)
    cnt, low = oracle.OracleCircuit.from_dict(d).sat(threads=8)
    calls = []
    r = ctx.sat_solve_progress(c, lambda s, n, b: calls.append(s) and False, elem_bits=16, stop_at_first=True)
    hit = low >> 16
    assert calls == [ns.SAT_STEP_START] + list(range(hit + 1))
    assert r["found"] and r["lowest"] == low and r["launches"] == hit + 1


@pytest.mark.parametrize("stop_after", [0, 3, 40])
def test_cancelled_count_is_the_prefix_count(ctx, stop_after):
    d = dense_circuit(7, 28)
    c = mk(d)
    seen = []

    def cb(s, n, b):
        if s != ns.SAT_STEP_START:
            seen.append(s)
        return s == stop_after

    r = ctx.sat_solve_progress(c, cb, elem_bits=18)
    assert r["complete"] == 0 and seen[-1] == stop_after
    assert seen == list(range(stop_after + 1))
    # the ranges already launched when the callback stopped the search were searched too
    done = r["launches"]
    assert stop_after + 1 <= done < 1 << 10
    cnt, low = oracle.OracleCircuit.from_dict(d).sat_range(0, done << (18 - 6), threads=8)
    assert r["count"] == cnt and bool(r["found"]) == (cnt > 0)
    if cnt:
        assert r["lowest"] == low


def test_cancel_before_the_first_range(ctx):
    c = mk(dense_circuit(9, 24))
    r = ctx.sat_solve_progress(c, lambda s, n, b: True, elem_bits=16)
    assert r["launches"] == 0 and r["complete"] == 0 and r["count"] == 0


def test_callback_exception_stops_and_is_raised(ctx):
    c = mk(dense_circuit(11, 24))

    def cb(s, n, b):
        if s == 2:
            raise KeyError("stop")
        return False

    with pytest.raises(KeyError):
        ctx.sat_solve_progress(c, cb, elem_bits=16)
    # the context still works afterwards
    assert ctx.sat_solve(c)["complete"] == 1


def test_qsat_progress_and_cancel(ctx):
    c = ns.Circuit.gen_qsat(12, 24, 400, 60)
    q = [0] * 6 + [1] * 6 + [0] * 6 + [1] * 6
    calls = []
    r = ctx.qsat_solve(c, q, elem_bits=18, progress=lambda s, n, b: calls.append(s) and False)
    assert calls == [ns.SAT_STEP_START] + list(range(1 << 6))
    plain = ctx.qsat_solve(c, q)
    keys = ("truth", "has_witness", "witness", "witness_count")
    assert {k: r[k] for k in keys} == {k: plain[k] for k in keys}
    with pytest.raises(ns.NaivesatError) as e:
        ctx.qsat_solve(c, q, elem_bits=18, progress=lambda s, n, b: s == 5)
    assert e.value.status == ns.NS_ERR_CANCELLED
    assert {k: ctx.qsat_solve(c, q)[k] for k in keys} == {k: plain[k] for k in keys}


def test_cli_step_lines_and_interrupt(tmp_path):
    d = dense_circuit(13, 26)
    f = tmp_path / "dense.circuit"
    f.write_text(mk(d).text())
    tool = os.path.join(ns.BIN_DIR, "naivesat")
    cnt, low = oracle.OracleCircuit.from_dict(d).sat(threads=8)
    out = subprocess.run([tool, "-t", "opencl:0", "-e", "18", "-c", str(f)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert lines[0].startswith("OpenCL devices: [hip:0 gfx")
    assert "Elem inputs: 18" in lines
    i = lines.index("Start execution")
    assert lines[i + 1: i + 1 + 256] == [f"Step: {k} / 256" for k in range(256)]
    assert lines[i + 257].startswith("Time: ")
    assert f"Count: {cnt}" in lines and f"Found Input: {low:026b}" in lines
    # default -e: 2^32 per range on a larger space; SIGINT after the first range stops it
    big = tmp_path / "big.circuit"
    cb, _ = ns.Circuit.gen_sat(3, 44, 3000, 40)
    big.write_text(cb.text())
    p = subprocess.Popen([tool, "-t", "opencl:0", "-c", str(big)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                         text=True)
    try:
        got = []
        for line in p.stdout:
            got.append(line.rstrip("\n"))
            if line.startswith("Step: 0 / 4096"):
                p.send_signal(signal.SIGINT)
                break
        rest, err = p.communicate(timeout=120)
    finally:
        if p.poll() is None:
            p.kill()
    got += rest.splitlines()
    assert p.returncode == 130, (got[-5:], err)
    assert "Elem inputs: 32" in got
    cancel = [x for x in got if x.startswith("Cancelled after ")]
    assert len(cancel) == 1 and " / 4096 launch ranges" in cancel[0]
    assert any(x.startswith("Count: ") and x.endswith("(searched ranges only)") for x in got)
