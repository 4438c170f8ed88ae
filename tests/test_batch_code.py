"""Batch code objects, checked on the host: the persistent batch kernels call each circuit
body at (pc base + entry), where the pc base is what s_getpc_b64 returns (the address of
the marker instruction after it) and the body sits in a region appended to .text. These
tests read the code object the engine would load (ELF sections and segments, the
disassembly of the kernels) and check that every offset the host computes lands where the
kernels jump."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import naivesat_amd as ns
from helpers import random_circuit

OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"
S_SETPC_RET = 0xBE801D26  # s_setpc_b64 s[38:39]
S_MOV_S36_LIT = 0xBEA400FF  # s_mov_b32 s36, <literal>
MARK_SAT, MARK_QSAT = 0x534253C0, 0x514253C0  # + register class index (64, 128, 256, 512)
CLASSES = (64, 128, 256, 512)


def words_of(code: bytes):
    return struct.unpack(f"<{len(code) // 4}I", code[: len(code) // 4 * 4])


def elf_layout(code: bytes):
    """(.text offset, .text size, [(p_offset, p_filesz, p_vaddr, p_flags) of PT_LOAD])"""
    assert code[:4] == b"\x7fELF" and code[4] == 2  # ELF64
    e_phoff, = struct.unpack_from("<Q", code, 0x20)
    e_shoff, = struct.unpack_from("<Q", code, 0x28)
    e_phentsize, e_phnum, e_shentsize, e_shnum, e_shstrndx = struct.unpack_from("<HHHHH", code, 0x36)
    secs = []
    for i in range(e_shnum):
        name, typ, flags, addr, off, size = struct.unpack_from("<IIQQQQ", code, e_shoff + i * e_shentsize)
        secs.append((name, off, size))
    stro = secs[e_shstrndx][1]
    text = None
    for name, off, size in secs:
        nm = code[stro + name: code.index(b"\0", stro + name)].decode()
        if nm == ".text":
            text = (off, size)
    loads = []
    for i in range(e_phnum):
        p_type, p_flags, p_off, p_vaddr, p_paddr, p_filesz = struct.unpack_from("<IIQQQQ", code,
                                                                                e_phoff + i * e_phentsize)
        if p_type == 1:
            loads.append((p_off, p_filesz, p_vaddr, p_flags))
    return text, loads


def pc_base(W, lit):
    hits = [i * 4 for i in range(1, len(W) - 1)
            if W[i] == S_MOV_S36_LIT and W[i + 1] == lit and (W[i - 1] & 0xFF80FF00) == 0xBE801C00]
    assert len(hits) == 1, hits
    return hits[0]


def check_batch_code(c, kind, lane_shift=0):
    body, cls, size = c.native_body(kind, lane_shift)
    code, patched, start = c.native_code(kind, lane_shift)
    code = bytes(code)
    W = words_of(code)
    text, loads = elf_layout(code)
    t0, t1 = text[0], text[0] + text[1]
    ci = CLASSES.index(cls)
    ps, pq = pc_base(W, MARK_SAT + ci), pc_base(W, MARK_QSAT + ci)
    for other in range(4):  # every class's kernels sit before the region
        assert pc_base(W, MARK_SAT + other) < start and pc_base(W, MARK_QSAT + other) < start
    # both pc bases, the body and its return lie in .text, in one executable load segment
    for off in (ps, pq, start, start + size):
        assert t0 <= off < t1
    seg = [s for s in loads if s[0] <= t0 and t1 <= s[0] + s[1]]
    assert len(seg) == 1 and seg[0][3] & 1  # PF_X
    # distances in the file are distances in memory within the segment
    assert seg[0][2] - seg[0][0] == seg[0][2] - seg[0][0]
    assert ps < start and pq < start  # entries are unsigned offsets from the pc bases
    # the body words are the encoded body, followed by the return
    enc = ns.native_encode(body)
    assert list(W[start // 4: start // 4 + len(enc)]) == list(enc)
    assert W[start // 4 + len(enc)] == S_SETPC_RET
    assert size == 4 * len(enc)
    return code, cls, ps, pq, start


def test_kernels_call_through_the_pc_base(tmp_path):
    """The disassembled kernels (every register class) add the descriptor's entry to the very
    registers s_getpc_b64 wrote, and call s[36:37] with the return address in s[38:39]."""
    c, _ = ns.Circuit.gen_sat(1, 24, 2000, 32)
    code, cls, ps, pq, start = check_batch_code(c, 4)
    f = tmp_path / "batch.co"
    f.write_bytes(code)
    dis = subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", str(f)], capture_output=True, text=True, check=True).stdout
    for kern in [f"ns_nat_b{k}_r{c}" for c in CLASSES for k in ("sat", "qsat")]:
        # the kernel's instructions up to its s_endpgm (the region follows the last kernel)
        m = re.search(rf"<{kern}>:\n(.*?s_endpgm)", dis, re.S)
        assert m, kern
        lines = [ln.split("//")[0].strip() for ln in m.group(1).splitlines()]
        gp = [re.match(r"s_getpc_b64 s\[(\d+):(\d+)\]", ln) for ln in lines]
        gp = [g for g in gp if g]
        assert len(gp) == 1
        lo, hi = int(gp[0].group(1)), int(gp[0].group(2))
        calls = [i for i, ln in enumerate(lines) if ln.startswith("s_swappc_b64 s[38:39], s[36:37]")]
        assert len(calls) == 1
        i = calls[0]
        add = re.match(r"s_add_u32 s36, s(\d+), s(\d+)", lines[i - 2])
        addc = re.match(r"s_addc_u32 s37, s(\d+), 0", lines[i - 1])
        assert add and addc, lines[i - 3: i + 1]
        assert int(add.group(1)) == lo and int(addc.group(1)) == hi
        # no divergent control flow: nothing in the kernel writes exec, so every lane of a
        # wave runs every loop iteration (the chunk loop's cross-lane reductions and the
        # calls always see the whole wave)
        for ln in lines:
            assert not re.match(r"\S+ exec[, ]", ln + " ") and "saveexec" not in ln, ln
        # descriptor fields and pointers reach SGPRs as unsigned halves: no sign extension
        # (a sign-extended low half once turned the QSAT bit array's address into a wild one)
        for ln in lines:
            assert not ln.startswith(("s_bfe_i64", "s_ashr_i32", "s_ashr_i64", "s_sext")), ln
        # the pc base registers are not written between s_getpc_b64 and the call
        gi = next(k for k, ln in enumerate(lines) if ln.startswith("s_getpc_b64"))
        for ln in lines[gi + 1: i - 2]:
            dst = re.match(r"\S+ (s\[(\d+):(\d+)\]|s(\d+))[, ]", ln + " ")
            if not dst or ln.startswith(("s_cmp", "s_cbranch", "s_branch", "s_waitcnt", "global_", "s_load",
                                         "v_", "ds_", "s_nop", "buffer_", "flat_", "s_endpgm")):
                continue
            regs = ({int(dst.group(2)), int(dst.group(3))} if dst.group(2) else {int(dst.group(4))})
            assert lo not in regs and hi not in regs, ln


@pytest.mark.parametrize("seed", range(4))
def test_batch_code_layout(seed):
    rng = np.random.default_rng(3100 + seed)
    n = int(rng.integers(12, 30))
    d = random_circuit(rng, n, int(rng.integers(40, 400)), recent=int(rng.integers(6, 40)))
    c = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])
    check_batch_code(c, 4)
    check_batch_code(c, 5, 2)


def test_batch_bodies_keep_the_callers_registers():
    # the body pool starts at s48: s36..s39 (call target, return address) and the kernels'
    # own s0..s35 / s40..s47 are never written by a body; VGPRs stay in v0 and the pool
    for seed, (n, g) in enumerate([(24, 2000), (32, 8000), (18, 300)]):
        c, _ = ns.Circuit.gen_sat(40 + seed, n, g, min(32, g))
        body, cls, size = c.native_body(4)
        for ln in body.splitlines():
            ln = ln.strip()
            if not ln or ln.startswith(";"):
                continue
            op, rest = ln.split(None, 1)
            dst = rest.split(",")[0].strip()
            if dst.startswith("s"):
                assert int(dst[1:]) >= 48, ln
            elif dst.startswith("v"):
                r = int(dst[1:])
                assert r == 0 or 8 <= r < cls if cls < 512 else r == 0 or 8 <= r < 240, ln
