"""Oracle answers for full-size QSAT searches whose outermost-block bitmap is mixed (some
block values 1, some 0), written to tests/golden/qsat_full_goldens.json.

The 32-input, 10,000-gate circuit of the QSAT workload (gen_qsat seed 3, output = XOR of 3
of the last 200 gates) under prefixes with a large outermost block and a few short inner
blocks: each block value is then a shallow AND/OR tree, so the bitmap is neither all zeros
nor all ones (the workload's own prefix, whose innermost block is universal over 8 inputs,
makes every value 0; long alternations drive the values to all 0 or all 1 as well). Each
case records truth, witness count, lowest witness, the bitmap's popcount and the SHA-256 of
the packed bitmap (little-endian bit order, bit u = value of block value u), plus the
bitmap itself in hex up to 2^16 bits.

The oracle is the C restatement under oracle/ (test infrastructure); about 10 s per case on
eight cores.

Usage: python tests/golden/make_qsat_goldens.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import naivesat_amd as ns  # noqa: E402
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "qsat_full_goldens.json")


def blocks(spec):
    """[("e", 24), ("a", 1), ...] -> one quantifier per input (1 = all)"""
    q = []
    for ch, k in spec:
        q += [1 if ch == "a" else 0] * k
    return q


CASES = [
    # (name, seed, n, gates, last, prefix)
    ("e24_ae4", 3, 32, 10000, 200, blocks([("e", 24)] + [("a", 1), ("e", 1)] * 4)),
    ("a24_ea4", 3, 32, 10000, 200, blocks([("a", 24)] + [("e", 1), ("a", 1)] * 4)),
    ("a20_e3a3e3a3", 3, 32, 10000, 200, blocks([("a", 20)] + [("e", 3), ("a", 3)] * 2)),
    ("a16_e2a2x4", 3, 32, 10000, 200, blocks([("a", 16)] + [("e", 2), ("a", 2)] * 4)),
]


def main():
    cases = []
    for name, seed, n, gates, last, q in CASES:
        c = ns.Circuit.gen_qsat(seed, n, gates, last)
        o = oracle.OracleCircuit.from_dict(c.arrays()).qsat(q, threads=os.cpu_count() or 8)
        bm = np.asarray(o["bitmap"], np.uint8)
        ones = int(bm.sum())
        assert 0 < ones < bm.size, (name, ones)  # mixed
        packed = np.packbits(bm, bitorder="little").tobytes()
        case = {"name": name, "seed": seed, "n": n, "gates": gates, "last": last,
                "quants": "".join("a" if x else "e" for x in q), "truth": o["truth"],
                "witness_count": o["witness_count"], "witness": o["witness"],
                "first_bits": o["first_bits"], "bitmap_ones": ones,
                "bitmap_sha256": hashlib.sha256(packed).hexdigest()}
        if bm.size <= 1 << 16:
            case["bitmap_hex"] = packed.hex()
        cases.append(case)
        print(name, o["truth"], o["witness_count"], o["witness"], ones, flush=True)
    with open(OUT, "w") as f:
        json.dump({"cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
