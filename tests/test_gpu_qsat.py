"""QSAT parity: HIP engine vs the CPU oracle (truth, lowest witness, witness count and
the outermost block's value bitmap, bit-exact)."""
import hashlib
import json
import os

import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import random_circuit

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "configs.json")


def check(ctx, d, q, **kw):
    c = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])
    r = ctx.qsat_solve(c, q, bitmap=True, **kw)
    o = oracle.OracleCircuit.from_dict(d).qsat(q, threads=8)
    ctx_info = (d["n"], "".join("a" if v else "e" for v in q), kw)
    assert r["truth"] == o["truth"], ctx_info
    assert r["first_bits"] == o["first_bits"]
    assert r["witness_count"] == o["witness_count"], ctx_info
    assert (r["witness"] if r["has_witness"] else None) == o["witness"], ctx_info
    assert np.array_equal(r["bitmap"], o["bitmap"]), ctx_info
    return r


def prefix(rng, n, blocks):
    """Alternating quantifier blocks of random lengths."""
    cuts = sorted(rng.choice(np.arange(1, n), size=min(blocks - 1, n - 1), replace=False)) if n > 1 else []
    q, first, prev = [], int(rng.integers(0, 2)), 0
    for i, cut in enumerate(list(cuts) + [n]):
        q += [first ^ (i & 1)] * (cut - prev)
        prev = cut
    return q


@pytest.mark.parametrize("n", list(range(1, 21)))
def test_random_prefixes_all_sizes(ctx, n):
    rng = np.random.default_rng(3000 + n)
    d = random_circuit(rng, n, int(rng.integers(2, 80)), recent=int(rng.integers(4, 30)))
    check(ctx, d, prefix(rng, n, int(rng.integers(1, 6))))


@pytest.mark.parametrize("fb", list(range(0, 17)))
def test_outer_block_boundary(ctx, fb):
    # inner levels L = 16 - fb sweep across the in-word / in-thread / reduction-pass split
    rng = np.random.default_rng(4000 + fb)
    n = 16
    d = random_circuit(rng, n, 60, recent=16)
    first = int(rng.integers(0, 2))
    fbb = max(fb, 1)
    q = [first] * fbb + [1 - first] * (n - fbb)
    check(ctx, d, q)


@pytest.mark.parametrize("group_len,wpl", [(64, 1), (128, 4), (256, 8), (512, 2), (64, 32)])
def test_geometries(ctx, group_len, wpl):
    rng = np.random.default_rng(11)
    n = 22
    d = random_circuit(rng, n, 120, recent=24)
    q = [0] * 3 + [1] * 7 + [0] * 6 + [1] * 6
    check(ctx, d, q, group_len=group_len, words_per_lane=wpl)


def test_all_exists_equals_sat(ctx):
    rng = np.random.default_rng(8)
    n = 18
    d = random_circuit(rng, n, 70, recent=20)
    r = check(ctx, d, [0] * n)
    c = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])
    s = ctx.sat_solve(c)
    assert r["witness_count"] == s["count"]
    assert r["truth"] == int(s["count"] > 0)


def test_result_string_matches_reference_format():
    assert ns.qsat_result_string({"first_bits": 3, "first_all": 0, "has_witness": 1, "witness": 0b110}) == \
        "(Satisfiable: 011)"
    assert ns.qsat_result_string({"first_bits": 2, "first_all": 1, "has_witness": 0, "witness": 0}) == \
        "(Satisfiable)"


def test_config2_golden(ctx):
    if not os.path.exists(GOLDEN) or "cfg2" not in json.load(open(GOLDEN)):
        pytest.skip("cfg2 golden not generated")
    g = json.load(open(GOLDEN))["cfg2"]
    c = ns.Circuit.gen_qsat(3, 32, 10000, 200)
    q = ns.quants_from_string(g["quants"])
    r = ctx.qsat_solve(c, q, bitmap=True)
    assert r["truth"] == g["truth"]
    assert (r["witness"] if r["has_witness"] else None) == g["witness"]
    assert r["witness_count"] == g["witness_count"]
    bm = np.packbits(r["bitmap"].astype(np.uint8), bitorder="little").tobytes().hex()
    assert bm == g["bitmap_hex"]


QSAT_FULL = os.path.join(os.path.dirname(__file__), "golden", "qsat_full_goldens.json")


def full_cases():
    with open(QSAT_FULL) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("idx", range(4))
def test_full_size_mixed_bitmap_goldens(ctx, idx):
    """2^32-assignment searches whose outermost-block bitmap is mixed (oracle goldens,
    tests/golden/qsat_full_goldens.json): truth, witness count, lowest witness and the bitmap
    bit for bit, alone and through the batch kernels."""
    g = full_cases()[idx]
    assert 0 < g["bitmap_ones"] < (1 << g["first_bits"])
    c = ns.Circuit.gen_qsat(g["seed"], g["n"], g["gates"], g["last"])
    q = ns.quants_from_string(g["quants"])
    r = ctx.qsat_solve(c, q, bitmap=True)
    assert r["truth"] == g["truth"]
    assert r["witness_count"] == g["witness_count"]
    assert (r["witness"] if r["has_witness"] else None) == g["witness"]
    packed = np.packbits(r["bitmap"].astype(np.uint8), bitorder="little").tobytes()
    assert int(r["bitmap"].sum()) == g["bitmap_ones"]
    assert hashlib.sha256(packed).hexdigest() == g["bitmap_sha256"]
    if "bitmap_hex" in g:
        assert packed.hex() == g["bitmap_hex"]
    res, info = ctx.batch_solve([{"kind": ns.REQ_QSAT, "circuit": c, "quants": q}])
    assert info["fallbacks"] == 0
    b = res[0]
    assert (b["v0"], b["v2"], b["v1"] if b["v3"] else None) == (g["truth"], g["witness_count"], g["witness"])


def test_multi_launch_and_multi_pass(ctx):
    """elem_bits far below the space forces 256 search launches into one bit array and two
    reduction passes over it; the same answer as one launch and as the oracle."""
    rng = np.random.default_rng(77)
    n = 28
    d = random_circuit(rng, n, 200, recent=40)
    q = [0] * 6 + [1, 0] * 11  # outermost block 6 inputs, 22 alternating levels below
    r = check(ctx, d, q, elem_bits=20)
    assert r["passes"] >= 2
    c = ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])
    one = ctx.qsat_solve(c, q, bitmap=True)
    assert (one["truth"], one["witness_count"], one["witness"]) == (r["truth"], r["witness_count"], r["witness"])
    assert np.array_equal(one["bitmap"], r["bitmap"])
