"""Host-side checks that need no GPU: the C-ABI library, circuit I/O and generators."""
import os
import re
import subprocess

import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import py_eval, random_circuit

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "naivesat_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ns_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ns.load_library()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", ns.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (ns_[a-z0-9_]+)", out))
    assert set(syms) <= exported
    assert set(ns.exported_symbols()) == set(syms)


def test_version_string():
    assert b"gfx950" in ns.load_library().ns_version()


def test_parse_format_round_trip():
    text = "{0 1 2 and(0,1) nor(2,3):0n xor(3,4):1 nimpl(4,0):2}"
    c = ns.Circuit.parse(text)
    assert (c.n_inputs, c.n_gates, c.n_outputs) == (3, 4, 3)
    assert c.text() == text
    assert ns.Circuit.parse(c.text()).text() == text
    d = c.arrays()
    for x in range(8):
        got = c.eval([(x >> k) & 1 for k in range(3)])
        assert got == py_eval(d, x)


def test_parse_output_on_input_and_spacing():
    c = ns.Circuit.parse(" { 0:1 1 and( 0 ,1 ):0 } ")
    assert c.n_outputs == 2
    assert c.eval([1, 1]) == [1, 1]
    assert c.eval([0, 1]) == [0, 0]


@pytest.mark.parametrize("bad", ["{0 2}", "{0 1 and(0,5):0}", "{0 1 foo(0,1):0}", "{0 1 and(0,1):1}",
                                 "{0 1 and(0,1):0", "{0 and(0,0):0 1}"])
def test_parse_errors_are_reported(bad):
    with pytest.raises(ns.NaivesatError) as e:
        ns.Circuit.parse(bad)
    assert e.value.status == 1


def test_quant_circuit_parse():
    c, q = ns.Circuit.parse_quant("eaE {0 1 2 xor(0,1) and(3,2):0}")
    assert list(q) == [0, 1, 0]
    c2, q2 = ns.Circuit.parse_quant("{0 1 xor(0,1):0}")
    assert q2 is None and c2.n_inputs == 2


def test_from_arrays_validation():
    with pytest.raises(ns.NaivesatError):
        ns.Circuit.from_arrays(2, [0], [2], [0], [2])  # operand refers to itself
    c = ns.Circuit.from_arrays(2, [0], [1], [3], [2], [1])
    assert [c.eval([a, b])[0] for a in (0, 1) for b in (0, 1)] == [1, 0, 0, 1]


def test_sat_generator_plants_key():
    for seed, n, gates, wb in [(1, 24, 2000, 32), (11, 16, 300, 12), (12, 10, 100, 8)]:
        c, key = ns.Circuit.gen_sat(seed, n, gates, wb)
        assert c.n_inputs == n and c.n_outputs == 1
        assert c.eval([(key >> k) & 1 for k in range(n)]) == [1]
        # deterministic in the seed
        c2, key2 = ns.Circuit.gen_sat(seed, n, gates, wb)
        assert key2 == key and c2.text() == c.text()
    assert ns.Circuit.gen_sat(2, 24, 2000, 32)[0].text() != ns.Circuit.gen_sat(1, 24, 2000, 32)[0].text()


def test_generators_match_independent_evaluation():
    c, key = ns.Circuit.gen_sat(5, 12, 200, 10)
    d = c.arrays()
    oc = oracle.OracleCircuit.from_dict(d)
    for x in range(0, 1 << 12, 37):
        assert oc.eval(x) & 1 == py_eval(d, x)[0] == c.eval([(x >> k) & 1 for k in range(12)])[0]


def test_satostop_generator_is_feistel_permutation():
    c, target = ns.Circuit.gen_satostop(9, 5, 6, 40, 4)
    n = 10
    assert c.n_inputs == n and c.n_outputs == n + 1
    oc = oracle.OracleCircuit.from_dict(c.arrays())
    nexts, stops = [], []
    for s in range(1 << n):
        o = oc.eval(s)
        nexts.append(o & ((1 << n) - 1))
        stops.append((o >> n) & 1)
    assert sorted(nexts) == list(range(1 << n))  # Feistel rounds are a bijection
    assert sum(stops) == 1 and nexts[stops.index(1)] == target
    # planted: the target is reached from some unknown start
    assert any(oc.walk(u << (n - 4), 1 << n)[0] for u in range(16))


def test_qsat_generator_shape():
    c = ns.Circuit.gen_qsat(3, 32, 10000, 200)
    assert c.n_inputs == 32 and c.n_outputs == 1 and c.n_gates == 10002
    cone, peak = c.cone()
    assert 0 < cone <= 10002 and 0 < peak <= cone


def test_context_without_device_fails_cleanly():
    import torch  # noqa: F401  (device discovery only)

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(ns.NaivesatError) as e:
        ns.Context(0)
    assert e.value.status in (1, 2, 4)


def test_device_list_and_cli_device_line(tmp_path):
    import torch  # noqa: F401  (device discovery only)

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    # no device: an empty list, and the tools' first line in the reference's form
    assert ns.device_list() == []
    f = tmp_path / "c.circuit"
    f.write_text(ns.Circuit.from_arrays(3, [0, 3], [1, 2], [2, 0], [4], [0]).text())
    r = subprocess.run([os.path.join(ns.BIN_DIR, "naivesat"), "-t", "opencl:0", str(f)], capture_output=True,
                       text=True)
    assert r.stdout.splitlines()[0] == "OpenCL devices: []"
    assert r.returncode != 0  # no device to run on


def test_cli_usage_and_errors(tmp_path):
    exe = os.path.join(ns.BIN_DIR, "naivesat")
    r = subprocess.run([exe, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "exec-type" in r.stdout
    bad = tmp_path / "bad.circuit"
    bad.write_text("{0 1 and(0,7):0}")
    r = subprocess.run([exe, "-t", "hip:0", str(bad)], capture_output=True, text=True)
    assert r.returncode != 0 and "circuit" in r.stderr
    r = subprocess.run([exe, "-t", "gpu:zz", str(bad)], capture_output=True, text=True)
    assert r.returncode != 0
