"""Native code path without an assembler per kernel: the engine's own gfx950 encoder gives
the same machine code as LLVM's assembler for every body it generates, and the code object
it loads carries that code in the template's body slot."""
import os
import re
import subprocess

import numpy as np
import pytest

import naivesat_amd as ns
from helpers import identity_circuit, random_circuit

LLVM_MC = "/opt/rocm/llvm/bin/llvm-mc"


LLVM_OBJCOPY = "/opt/rocm/llvm/bin/llvm-objcopy"


def llvm_words(text, tmp_path=None):
    if re.search(r"^\S+:\s*$", text, re.M):
        # branches to local labels: -show-encoding prints them as unresolved fixups, so the
        # words come from an assembled object's .text
        obj, raw = tmp_path / "body.o", tmp_path / "body.bin"
        subprocess.run([LLVM_MC, "-triple=amdgcn-amd-amdhsa", "-mcpu=gfx950", "-filetype=obj", "-o", str(obj)],
                       input=text, text=True, check=True)
        subprocess.run([LLVM_OBJCOPY, "-O", "binary", "--only-section=.text", str(obj), str(raw)], check=True)
        return np.frombuffer(raw.read_bytes(), dtype="<u4")
    out = subprocess.run([LLVM_MC, "-triple=amdgcn-amd-amdhsa", "-mcpu=gfx950", "-show-encoding"],
                         input=text, capture_output=True, text=True, check=True).stdout
    b = []
    for line in out.splitlines():
        m = re.search(r"encoding: \[([^\]]*)\]", line)
        if m:
            b += [int(x, 16) for x in m.group(1).split(",")]
    assert len(b) % 4 == 0
    return np.frombuffer(bytes(b), dtype="<u4")


def mk(d):
    return ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])


def bodies():
    rng = np.random.default_rng(5)
    yield "cfg0", ns.Circuit.gen_sat(1, 24, 2000, 32)[0], 0, 0
    yield "cfg1 (AGPR copies, lane slots)", ns.Circuit.gen_sat(2, 36, 20000, 40)[0], 0, 0
    yield "qsat", ns.Circuit.gen_qsat(3, 32, 10000, 200), 1, 2
    yield "table", ns.Circuit.gen_satostop(33, 8, 6, 60, 8)[0], 3, 0
    yield "identity table", mk(identity_circuit(20)), 3, 0
    yield "random table", mk(random_circuit(rng, 14, 120, nout=15, recent=30)), 3, 0
    yield "cfg1 with group exits (labels, branches, v_cmp)", ns.Circuit.gen_sat(2, 36, 20000, 40)[0], 7, 0


@pytest.mark.skipif(not os.path.exists(LLVM_MC), reason="llvm-mc is not installed")
@pytest.mark.parametrize("case", list(range(7)))
def test_encoder_matches_llvm_mc(case, tmp_path):
    name, c, kind, ls = list(bodies())[case]
    body, cls, _ = c.native_body(kind, ls)
    ours = ns.native_encode(body)
    ref = llvm_words(body, tmp_path)
    assert ours.size == ref.size, name
    bad = np.nonzero(ours != ref)[0]
    assert bad.size == 0, (name, bad[:5], [hex(x) for x in ours[bad[:5]]], [hex(x) for x in ref[bad[:5]]])


def test_encoder_rejects_unknown_instructions():
    with pytest.raises(ns.NaivesatError):
        ns.native_encode("\tv_add_u32 v1, v2, v3\n")


def test_encoder_branches():
    # a SOPP branch holds the distance in words from the instruction after it, backwards too
    w = ns.native_encode("\ts_branch .La\n\tv_mov_b32 v0, 0\n.La:\n\ts_cbranch_vccz .La\n")
    assert [hex(x) for x in w] == ["0xbf820001", "0x7e000280", "0xbf86ffff"]
    for bad in ["\ts_branch .Lnowhere\n", ".La:\n.La:\n", "\tv_cmp_ne_u32 s4, 0, v1\n"]:
        with pytest.raises(ns.NaivesatError):
            ns.native_encode(bad)


@pytest.mark.parametrize("case", [0, 1, 3])
def test_code_object_slot(case):
    name, c, kind, ls = list(bodies())[case]
    body, _, _ = c.native_body(kind, ls)
    code, patched, off = c.native_code(kind, ls)
    assert patched, name
    words = ns.native_encode(body)
    got = np.frombuffer(code[off: off + 4 * words.size + 4], dtype="<u4")
    assert np.array_equal(got[:-1], words), name
    br = int(got[-1])
    assert br >> 16 == 0xBF82  # s_branch
    # the branch lands on the first byte after the slot: the rest of the slot is s_nop 0
    skip = br & 0xFFFF
    tail = np.frombuffer(code[off + 4 * words.size + 4: off + 4 * words.size + 4 + 4 * skip], dtype="<u4")
    assert tail.size == skip and np.all(tail == 0xBF800000)
    assert (4 * words.size + 4 + 4 * skip) in (16384, 65536, 131008)
    assert b"MBSNBODY" not in code
