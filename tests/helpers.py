"""Shared test helpers: random circuits and an independent pure-Python evaluator."""
import numpy as np

# gate truth tables indexed by (a << 1) | b: And, Nor, Nimpl (a & ~b), Xor
FUNC_TT = {0: (0, 0, 0, 1), 1: (1, 0, 0, 0), 2: (0, 0, 1, 0), 3: (0, 1, 1, 0)}


def py_eval(d, x):
    """Evaluates a circuit dict (n, in0, in1, func, out_wire, out_neg) on assignment x
    (input k = bit k of x); returns the list of output bits."""
    n = int(d["n"])
    v = [(x >> k) & 1 for k in range(n)]
    for a, b, f in zip(d["in0"], d["in1"], d["func"]):
        v.append(FUNC_TT[int(f)][(v[int(a)] << 1) | v[int(b)]])
    return [v[int(w)] ^ int(ng) for w, ng in zip(d["out_wire"], d["out_neg"])]


def random_circuit(rng, n, gates, nout=1, recent=None):
    """Random netlist: operands uniform over earlier wires (or the last `recent` ones)."""
    in0 = np.zeros(gates, np.uint32)
    in1 = np.zeros(gates, np.uint32)
    func = rng.integers(0, 4, gates).astype(np.uint8)
    for j in range(gates):
        lo = 0 if recent is None else max(0, n + j - recent)
        in0[j] = rng.integers(lo, n + j)
        in1[j] = rng.integers(lo, n + j)
    out_wire = np.array([n + gates - 1 - i for i in range(nout)], np.uint32) if gates else \
        np.array(rng.integers(0, n, nout), np.uint32)
    out_neg = rng.integers(0, 2, nout).astype(np.uint8)
    return {"n": n, "in0": in0, "in1": in1, "func": func, "out_wire": out_wire, "out_neg": out_neg}


def qsat_recursive(d, q):
    """Value of Q0 x0 ... Q(n-1) x(n-1) C(x) by direct recursion (q[k] = 1 for 'all')."""
    n = int(d["n"])

    def rec(k, x):
        if k == n:
            return py_eval(d, x)[0]
        a = rec(k + 1, x)
        b = rec(k + 1, x | (1 << k))
        return (a & b) if q[k] else (a | b)

    return rec(0, 0)


def block_values(d, q):
    """Values of the formula below the outermost block, for each block value u
    (u's most significant bit = input 0)."""
    n = int(d["n"])
    fb = 0
    while fb < n and q[fb] == q[0]:
        fb += 1

    def rec(k, x):
        if k == n:
            return py_eval(d, x)[0]
        a = rec(k + 1, x)
        b = rec(k + 1, x | (1 << k))
        return (a & b) if q[k] else (a | b)

    vals = []
    for u in range(1 << fb):
        x = 0
        for i in range(fb):  # input i is bit (fb-1-i) of u
            if (u >> (fb - 1 - i)) & 1:
                x |= 1 << i
        vals.append(rec(fb, x))
    return vals


def identity_circuit(m):
    return {"n": m, "in0": np.zeros(0, np.uint32), "in1": np.zeros(0, np.uint32),
            "func": np.zeros(0, np.uint8), "out_wire": np.arange(m, dtype=np.uint32),
            "out_neg": np.zeros(m, np.uint8)}


def point_stop_circuit(n, p, nxt):
    """Next-state circuit whose table is the identity except at state p, whose entry is
    STOP with next state nxt: outputs 0..n-1 = s XOR ([s == p] & (p XOR nxt)), output n =
    [s == p]. Built from And / Nor / Nimpl / Xor gates only."""
    in0, in1, func = [], [], []

    def gate(f, a, b):
        in0.append(a)
        in1.append(b)
        func.append(f)
        return n + len(func) - 1

    # literals (wire, positive) of s_i == p_i, folded into one positive AND
    lits = [(i, (p >> i) & 1 == 1) for i in range(n)]
    while len(lits) > 1:
        nxt_lits = []
        for k in range(0, len(lits) - 1, 2):
            (a, pa), (b, pb) = lits[k], lits[k + 1]
            if pa and pb:
                w = gate(0, a, b)
            elif pa:
                w = gate(2, a, b)
            elif pb:
                w = gate(2, b, a)
            else:
                w = gate(1, a, b)
            nxt_lits.append((w, True))
        if len(lits) % 2:
            nxt_lits.append(lits[-1])
        lits = nxt_lits
    eq, pos = lits[0]
    if not pos:  # n == 1 and p_0 == 0: eq = NOR(s0, s0)
        eq = gate(1, eq, eq)
    outs = []
    for i in range(n):
        outs.append(gate(3, i, eq) if ((p ^ nxt) >> i) & 1 else i)
    outs.append(eq)
    return {"n": n, "in0": np.array(in0, np.uint32), "in1": np.array(in1, np.uint32),
            "func": np.array(func, np.uint8), "out_wire": np.array(outs, np.uint32),
            "out_neg": np.zeros(n + 1, np.uint8)}
