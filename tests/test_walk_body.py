"""Walk bodies of the table-free SATOSTOP walks (ns_nat_walk<T> templates), checked on the
host: the register-allocated body (two walk steps and their per-slot bookkeeping) runs in the
instruction interpreter on random register contents, and every register it must produce is
compared with the walk iteration computed slot by slot from the oracle's circuit evaluation:
next state, STOP, waypoint and saved-state tests, final states, ages and masks."""
import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import point_stop_circuit, random_circuit
from test_native_body import run_body

M32 = np.uint32(0xFFFFFFFF)
TIERS = (16, 24, 32, 40, 48)
MASKS = ("active", "fin", "fstop", "fwp", "floop", "a0", "a1", "a2", "a3")


def regs_of(T):
    return {"s": 8, "f": 8 + T, "v": 8 + 2 * T, "m": 8 + 3 * T}


def next_words(oc, n, words):
    """bit-sliced next state: words[i][lane] bit b = state bit i of slot (lane, b); returns
    (Y[0..n), ON) in the same layout"""
    Y = [np.zeros(64, np.uint32) for _ in range(n + 1)]
    for lane in range(64):
        for b in range(32):
            x = 0
            for i in range(n):
                x |= int((words[i][lane] >> np.uint32(b)) & np.uint32(1)) << i
            y = oc.eval(x)
            for i in range(n + 1):
                if (y >> i) & 1:
                    Y[i][lane] |= np.uint32(1 << b)
    return Y


def model(oc, n, k, st):
    s, f, v = list(st["s"]), list(st["f"]), list(st["v"])
    m = dict(st["m"])
    for half in range(2):
        Y = next_words(oc, n, s)
        on = Y[n]
        act = m["active"]
        stop = on & act
        nz = np.zeros(64, np.uint32)
        for i in range(k):
            nz |= Y[i]
        wp = ~nz & act & ~stop
        done = stop | wp
        if half == 1:
            eq = np.full(64, M32, np.uint32)
            for i in range(n):
                eq &= ~(Y[i] ^ v[i])
            loop = eq & act & ~stop & ~wp
            done = done | loop
            m["floop"] = m["floop"] | loop
        for i in range(n):
            f[i] = (f[i] & ~done) | (Y[i] & done)
        cy = m["fin"]
        m["a0"] = m["a0"] ^ cy
        cy = cy & ~m["a0"]
        m["a1"] = m["a1"] ^ cy
        cy = cy & ~m["a1"]
        m["a2"] = m["a2"] ^ cy
        cy = cy & ~m["a2"]
        m["a3"] = m["a3"] ^ cy
        m["fin"] = m["fin"] | done
        m["fstop"] = m["fstop"] | stop
        m["fwp"] = m["fwp"] | wp
        m["active"] = act & ~done
        s = Y[:n]
    return s, f, m


def check_walk_body(c, k, seed):
    body, cls, size = c.native_body(6, k)
    assert cls in (256, 512) and size > 0
    d = c.arrays()
    n = d["n"]
    T = next(t for t in TIERS if n <= t)
    R = regs_of(T)
    oc = oracle.OracleCircuit.from_dict(d)
    rng = np.random.default_rng(seed)
    rnd = lambda: rng.integers(0, 1 << 32, 64, dtype=np.uint64).astype(np.uint32)
    st = {"s": [rnd() for _ in range(n)], "f": [rnd() for _ in range(n)], "v": [rnd() for _ in range(n)],
          "m": {x: rnd() for x in MASKS}}
    vr = {}
    for i in range(n):
        vr[R["s"] + i] = st["s"][i]
        vr[R["f"] + i] = st["f"][i]
        vr[R["v"] + i] = st["v"][i]
    for j, x in enumerate(MASKS):
        vr[R["m"] + j] = st["m"][x]
    out = run_body(body, 0, 0, vr)
    s, f, m = model(oc, n, k, st)
    for i in range(n):
        assert np.array_equal(out[R["s"] + i], s[i]), ("s", i)
        assert np.array_equal(out[R["f"] + i], f[i]), ("f", i)
        assert np.array_equal(out[R["v"] + i], st["v"][i]), ("v kept", i)
    for j, x in enumerate(MASKS):
        assert np.array_equal(out[R["m"] + j], m[x]), x
    # the body writes no template register outside its contract (the state bits above n and
    # the registers below the template's)
    for r in out:
        assert not (r < R["s"] or R["s"] + n <= r < R["f"] or R["f"] + n <= r < R["v"] or
                    R["v"] + n <= r < R["m"]), f"v{r} written"
    return body


@pytest.mark.parametrize("seed", range(3))
def test_walk_body_random_circuits(seed):
    rng = np.random.default_rng(4100 + seed)
    n = int(rng.integers(6, 14))
    d = random_circuit(rng, n, int(rng.integers(20, 120)), nout=n + 1, recent=20)
    c = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])
    check_walk_body(c, int(rng.integers(1, 5)), seed)


def test_walk_body_feistel_and_saved_state_hits():
    c, _ = ns.Circuit.gen_satostop(3, 8, 3, 40, 6)  # 16 state bits
    d = c.arrays()
    n = d["n"]
    body = check_walk_body(c, 4, 11)
    # loop detection: saved states equal to the state two steps on
    oc = oracle.OracleCircuit.from_dict(d)
    rng = np.random.default_rng(5)
    R = regs_of(16)
    s0 = [rng.integers(0, 1 << 32, 64, dtype=np.uint64).astype(np.uint32) for _ in range(n)]
    two = next_words(oc, n, next_words(oc, n, s0)[:n])[:n]
    vr = {R["s"] + i: s0[i] for i in range(n)}
    for i in range(n):
        vr[R["f"] + i] = np.zeros(64, np.uint32)
        vr[R["v"] + i] = two[i]
    for j, x in enumerate(MASKS):
        vr[R["m"] + j] = np.full(64, M32 if x == "active" else 0, np.uint32)
    out = run_body(body, 0, 0, vr)
    st = {"s": s0, "f": [vr[R["f"] + i] for i in range(n)], "v": two, "m": {x: vr[R["m"] + j] for j, x in
                                                                             enumerate(MASKS)}}
    _, _, m = model(oc, n, 4, st)
    assert np.array_equal(out[R["m"] + MASKS.index("floop")], m["floop"])
    assert int(np.count_nonzero(m["floop"])) > 0


@pytest.mark.parametrize("n", [16, 24, 30, 36, 48])
def test_walk_body_fits_each_tier(n):
    half = n // 2
    c, _ = ns.Circuit.gen_satostop(n, half, 4, 30, min(8, n - 4))
    body, cls, size = c.native_body(6, 6)
    T = next(t for t in TIERS if n <= t)
    # no instruction writes a VGPR below the template registers, and no scalar instruction
    # (the walk templates keep their own scalars in every SGPR across the body)
    for ln in body.splitlines():
        ln = ln.strip()
        if not ln or ln.startswith(";"):
            continue
        assert ln.startswith("v_") and not ln.startswith(("v_readlane", "v_readfirstlane")), ln
        dst = ln.split(None, 1)[1].split(",")[0].strip()
        if dst.startswith("v"):
            assert int(dst[1:]) >= 8, ln


def test_walk_code_object_holds_the_body():
    # the code object the engine loads: the walk template's with the encoded body in its slot
    import struct
    c, _ = ns.Circuit.gen_satostop(21, 12, 6, 120, 8)
    body, cls, size = c.native_body(6, 5)
    code, patched, off = c.native_code(6, 5)
    code = bytes(code)
    assert patched and code[:4] == b"\x7fELF"
    enc = ns.native_encode(body)
    got = struct.unpack_from(f"<{len(enc)}I", code, off)
    assert list(got) == list(enc)
