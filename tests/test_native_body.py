"""Native code path, checked on the host: the generated circuit bodies (register-allocated
gfx950 assembly) are executed instruction by instruction by a small interpreter of the
instruction subset they use, and their outputs compared with the oracle's evaluation."""
import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import random_circuit

M32 = 0xFFFFFFFF
LANES = np.arange(64, dtype=np.uint32)


def run_body(body, lm_base, wu, vregs=None, trace=None):
    """vregs: initial VGPRs (register -> 64 lane values) instead of the lane masks; trace
    (a dict): trace["branches"] lists the labels the taken branches went to"""
    v, s, acc = {}, {20: wu & M32, 21: (wu >> 32) & M32}, {}
    if vregs is not None:
        v = {r: x.copy() for r, x in vregs.items()}
    else:
        for i in range(6):
            v[lm_base + i] = np.where((LANES >> i) & 1, np.uint32(M32), np.uint32(0)).astype(np.uint32)

    def sidx(tok):  # SGPR number: sN, or vcc_lo / vcc_hi (106, 107)
        return {"vcc_lo": 106, "vcc_hi": 107}.get(tok) or int(tok[1:])

    def is_s(tok):
        return tok.startswith("s") or tok in ("vcc_lo", "vcc_hi")

    def val(tok):
        if is_s(tok):
            return np.full(64, s[sidx(tok)], np.uint32)
        if tok.startswith("v"):
            return v[int(tok[1:])]
        return np.full(64, int(tok, 0) & M32, np.uint32)  # constants (-1: all ones)

    def sval(tok):
        return s[sidx(tok)] if is_s(tok) else int(tok, 0) & M32

    lines = [ln.strip() for ln in body.splitlines()]
    lines = [ln for ln in lines if ln and not ln.startswith(";")]
    labels = {ln[:-1]: i for i, ln in enumerate(lines) if ln.endswith(":")}
    vcc = 0
    pc = 0
    while pc < len(lines):
        line = lines[pc]
        pc += 1
        if line.endswith(":"):
            continue
        op, rest = line.split(None, 1)
        args = [a.strip() for a in rest.split(",")]
        if op == "v_cmp_ne_u32":
            assert args[0] == "vcc"
            a, b = val(args[1]), val(args[2])
            vcc = sum(1 << i for i in range(64) if a[i] != b[i])
        elif op == "s_cbranch_vccz":
            if vcc == 0:
                pc = labels[args[0]]
                if trace is not None:
                    trace.setdefault("branches", []).append(args[0])
        elif op == "s_branch":
            pc = labels[args[0]]
            if trace is not None:
                trace.setdefault("branches", []).append(args[0])
        elif op == "s_mov_b32":
            s[sidx(args[0])] = sval(args[1])
        elif op == "s_bfe_i32":
            src, ctl = sval(args[1]), sval(args[2])
            off, width = ctl & 31, (ctl >> 16) & 0x7F
            assert width == 1
            s[sidx(args[0])] = M32 if (src >> off) & 1 else 0
        elif op in ("s_and_b32", "s_nor_b32", "s_andn2_b32", "s_xor_b32"):
            a, b = sval(args[1]), sval(args[2])
            r = {"s_and_b32": a & b, "s_nor_b32": ~(a | b), "s_andn2_b32": a & ~b, "s_xor_b32": a ^ b}[op]
            s[sidx(args[0])] = r & M32
        elif op == "v_bitop3_b32":
            src2, tt = args[3].split()
            tt = int(tt.split(":")[1], 0)
            A, B, Cc = val(args[1]), val(args[2]), val(src2)
            r = np.zeros(64, np.uint32)
            for idx in range(8):
                if (tt >> idx) & 1:
                    r |= (A if idx & 4 else ~A) & (B if idx & 2 else ~B) & (Cc if idx & 1 else ~Cc)
            v[int(args[0][1:])] = r
        elif op == "v_mov_b32":
            v[int(args[0][1:])] = val(args[1]).copy()
        elif op == "v_not_b32":
            v[int(args[0][1:])] = ~val(args[1])
        elif op == "v_accvgpr_write_b32":
            acc[int(args[0][1:])] = val(args[1]).copy()
        elif op == "v_accvgpr_read_b32":
            v[int(args[0][1:])] = acc[int(args[1][1:])].copy()
        elif op == "v_writelane_b32":
            reg = int(args[0][1:])
            v.setdefault(reg, np.zeros(64, np.uint32))
            v[reg] = v[reg].copy()
            v[reg][int(args[2])] = sval(args[1])
        elif op == "v_readlane_b32":
            s[sidx(args[0])] = int(v[int(args[1][1:])][int(args[2])])
        elif op == "s_nop":
            pass
        else:
            raise AssertionError("unexpected instruction " + op)
    return v


def expected_words(oc, n, outs, vbit, lane_shift, wu):
    """outputs as 32-bit words per lane: bit b of lane l = output for index 32*w + b,
    w = wu + (l << lane_shift)"""
    res = np.zeros((len(outs), 64), np.uint32)
    for lane in range(64):
        w = wu + (lane << lane_shift)
        for b in range(32):
            idx = 32 * w + b
            if n < 64 and idx >> n:
                continue
            x = 0
            for k in range(n):
                if (idx >> vbit[k]) & 1:
                    x |= 1 << k
            o = oc.eval(x)
            for i, out in enumerate(outs):
                if (o >> out) & 1:
                    res[i, lane] |= np.uint32(1 << b)
    return res


def check(c, kind, lane_shift, wus, traces=None):
    body, cls, size = c.native_body(kind, lane_shift)
    assert size > 0 and cls in (64, 128, 256, 512)
    d = c.arrays()
    n = d["n"]
    oc = oracle.OracleCircuit.from_dict(d)
    outs = list(range(len(d["out_wire"]))) if kind == 3 else [0]
    vbit = [n - 1 - k for k in range(n)] if kind in (1, 2, 5) else list(range(n))
    lm = 32 if kind == 3 else 1
    for wu in wus:
        tr = {}
        v = run_body(body, lm, wu, trace=tr)
        if traces is not None:
            traces.append(tr.get("branches", []))
        exp = expected_words(oc, n, outs, vbit, lane_shift, wu)
        for i in range(len(outs)):
            got = v[i]
            if n < 5:
                got = got & np.uint32((1 << (1 << n)) - 1)
            if n < 11:  # lanes beyond the space carry no assignments
                nl = max(1, 1 << max(0, n - 5)) if lane_shift == 0 else 64
                got, e = got[:nl], exp[i][:nl]
            else:
                e = exp[i]
            assert np.array_equal(got, e), (kind, lane_shift, wu, i)
        if kind == 3:  # table entries pack 32 registers: the ones without an output are zero
            for k in range(len(outs), 32):
                assert not v[k].any(), (wu, k)


@pytest.mark.parametrize("seed", range(5))
def test_sat_bodies(seed):
    rng = np.random.default_rng(900 + seed)
    n = int(rng.integers(12, 22))
    d = random_circuit(rng, n, int(rng.integers(30, 200)), recent=int(rng.integers(6, 40)))
    c = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])
    wus = [0, 64 * 3, (1 << (n - 5)) - 64] if n > 11 else [0]
    check(c, 0, 0, wus)


def test_config0_body():
    c, _ = ns.Circuit.gen_sat(1, 24, 2000, 32)
    check(c, 0, 0, [0, 64 * 1000 + 128])


@pytest.mark.parametrize("lane_shift", [0, 2, 3])
def test_qsat_bodies(lane_shift):
    c = ns.Circuit.gen_qsat(17, 20, 300, 60)
    n = 20
    base = [0, 1 << (lane_shift + 6), (1 << (n - 5)) - (1 << (lane_shift + 6))]
    check(c, 1, lane_shift, [b | j for b in base for j in range(min(2, 1 << lane_shift))])


def test_table_bodies():
    c, _ = ns.Circuit.gen_satostop(33, 8, 6, 60, 8)
    check(c, 3, 0, [0, 64 * 5, (1 << 11) - 64])
    rng = np.random.default_rng(4)
    d = random_circuit(rng, 14, 120, nout=15, recent=30)
    c2 = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])
    check(c2, 3, 0, [0, 64 * 2])


def test_table_outputs_are_inputs():
    # identity next-state function: every output is an input, most of them wave-uniform
    for m in (5, 12, 20):
        c = ns.Circuit.from_arrays(m, [], [], [], list(range(m)), [0] * m)
        check(c, 3, 0, [0] if m < 12 else [0, 64])


def test_uniform_and_negated_outputs():
    n = 17
    c = ns.Circuit.from_arrays(n, [0], [16], [3], [16, 17, 3, 17], [0, 1, 1, 0])
    check(c, 3, 0, [0, 64, 128 + 64])
    check(c, 0, 0, [0, 64])


def test_small_spaces():
    rng = np.random.default_rng(12)
    for n in (1, 3, 6, 9):
        d = random_circuit(rng, n, 20)
        c = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])
        check(c, 0, 0, [0])


def test_config1_body_with_agpr_spills(monkeypatch):
    # 2,102 cone gates with ~370 live LUT values: as one group the 512-register class spills
    # to AGPRs
    monkeypatch.setenv("NS_OUTPUT_GROUPS", "1")
    c, _ = ns.Circuit.gen_sat(2, 36, 20000, 40)
    body, cls, size = c.native_body(0)
    assert cls == 512 and "v_accvgpr_write_b32" in body
    check(c, 0, 0, [64 * 12345])


def test_config1_body_in_output_groups():
    # its output is the conjunction of 40 comparisons: evaluated as five groups of its
    # literals (shared gates recomputed), every group fits 256 VGPRs, two waves per SIMD; the
    # refined groups take fewer LUTs than the 2,165 of the greedy ones
    c, key = ns.Circuit.gen_sat(2, 36, 20000, 40)
    body, cls, size = c.native_body(0)
    assert cls == 256 and "v_accvgpr" not in body
    assert body.count("v_bitop3_b32") < 1700
    # also the waves holding the planted key and the lowest solution (798,579,149, the
    # full-size golden): words with satisfying bits, not only all-zero outputs
    check(c, 0, 0, [64 * 12345, (1 << 31) - 64, (key >> 5) & ~63, (798579149 >> 5) & ~63])


def test_config1_body_with_group_exits():
    """NS_SAT_GROUP_EXITS (kind 7): after each output group but the last the body leaves with
    output 0 when the running conjunction is zero in all 64 lanes. Same outputs as the oracle
    on waves that leave early and on the waves holding the key and the lowest solution, which
    run every group."""
    c, key = ns.Circuit.gen_sat(2, 36, 20000, 40)
    body, cls, size = c.native_body(7)
    assert cls == 256 and body.count("s_cbranch_vccz .Lns_exit") >= 1
    # exit-first groups: the first exit comes within the first third of the body
    ins = [ln for ln in body.splitlines() if ln.startswith("\t")]
    first = next(i for i, ln in enumerate(ins) if "s_cbranch_vccz" in ln)
    assert first < len(ins) / 3, (first, len(ins))
    traces = []
    wus = [64 * 12345, 64 * 999_999, (1 << 31) - 64, (key >> 5) & ~63, (798579149 >> 5) & ~63]
    check(c, 7, 0, wus, traces)
    assert any(t == [".Lns_exit"] for t in traces[:3])  # a zero wave leaves at a group exit
    assert traces[3] == [".Lns_done"] and traces[4] == [".Lns_done"]  # solutions: every group


def test_no_group_exits_for_a_negated_output():
    # the output ~(conjunction) is 1 where the conjunction is 0: no exit may write 0 there
    c, _ = ns.Circuit.gen_sat(2, 26, 40000, 48)
    d = c.arrays()
    assert "s_cbranch_vccz" in c.native_body(7)[0]
    neg = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], 1 - d["out_neg"])
    body, cls, _ = neg.native_body(7)
    assert "s_cbranch_vccz" not in body
    check(neg, 7, 0, [64 * 4321])


def test_groups_when_one_group_fits_no_class():
    # 3,204 cone gates with ~800 live values fit no register class as one group (the
    # compiler path would take them); as output groups they fit the AGPR class
    c, _ = ns.Circuit.gen_sat(2, 36, 40000, 48)
    with pytest.MonkeyPatch.context() as mp:
        mp.setenv("NS_OUTPUT_GROUPS", "1")
        with pytest.raises(ns.NaivesatError, match="does not fit"):
            c.native_body(0)
    body, cls, size = c.native_body(0)
    assert cls in (256, 512)
    check(c, 0, 0, [64 * 777])


@pytest.mark.parametrize("K", [2, 3, 4, 7])
def test_forced_output_groups(monkeypatch, K):
    monkeypatch.setenv("NS_OUTPUT_GROUPS", str(K))
    c, _ = ns.Circuit.gen_sat(600 + K, 20, 400, 12)
    check(c, 0, 0, [0, 64 * 77])
    # negated and Nor/Nimpl literals, and a negated output over the conjunction
    rng = np.random.default_rng(650 + K)
    d = random_circuit(rng, 16, 90, recent=25)
    g = len(d["func"])
    top = [int(d["out_wire"][0]), 16 + g - 3, 16 + g - 7, 16 + g - 11, 16 + g - 15]
    d["in0"] = np.append(d["in0"], [top[0], top[2], 16 + g, 16 + g + 2]).astype(np.uint32)
    d["in1"] = np.append(d["in1"], [top[1], top[3], top[4], 16 + g + 1]).astype(np.uint32)
    d["func"] = np.append(d["func"], [2, 1, 0, 0]).astype(np.uint8)
    for neg in (0, 1):
        c2 = ns.Circuit.from_arrays(16, d["in0"], d["in1"], d["func"], [16 + g + 3], [neg])
        check(c2, 0, 0, [0, 64 * 9])
        check(c2, 1, 2, [0, 1 << 8])


@pytest.mark.parametrize("seed", range(4))
def test_batch_bodies(seed):
    # bodies of the persistent batch kernels (kinds 4 and 5: SAT and QSAT input orders, SGPRs
    # from s48 on) compute the same words as the circuit
    rng = np.random.default_rng(1700 + seed)
    n = int(rng.integers(13, 24))  # the batch kernels take QSAT with n >= 11 + jb
    d = random_circuit(rng, n, int(rng.integers(40, 300)), recent=int(rng.integers(6, 40)))
    c = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])
    words = 1 << (n - 5)
    check(c, 4, 0, sorted({0, min(64 * 3, words - 64), words - 64}))
    for ls in (0, 2):
        tile = 1 << (ls + 6)
        base = sorted({0, min(tile, words - tile), words - tile})
        check(c, 5, ls, [b | j for b in base for j in range(min(2, 1 << ls))])


def test_config0_batch_body():
    c, _ = ns.Circuit.gen_sat(1, 24, 2000, 32)
    check(c, 4, 0, [0, 64 * 1000 + 128])


def test_register_budget_is_reported():
    # a body that needs more than 232 VGPRs + 256 AGPRs of simultaneously live values
    n = 40
    rng = np.random.default_rng(3)
    gates = 1400
    in0 = np.arange(n, n + gates, dtype=np.uint32) - 1
    in0[0] = 0
    in1 = rng.integers(0, n, gates).astype(np.uint32)
    func = np.full(gates, 3, np.uint8)
    # every gate feeds the final XOR chain, so all of them stay live until the end
    tail0 = np.arange(n, n + gates - 1, dtype=np.uint32)
    all0 = np.concatenate([in0, tail0 * 0 + (n + gates - 1)])
    c = ns.Circuit.from_arrays(n, in0, in1, func, [n + gates - 1], [0])
    body, cls, size = c.native_body(0)
    assert cls in (64, 128, 256, 512)
