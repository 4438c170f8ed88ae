"""Computes the benchmark configurations' answers with the CPU oracle and writes
tests/golden/configs.json (run once; takes minutes on 8 cores and ~17 GB of RAM).

    python tests/golden/make_config_goldens.py [cfg0 cfg1 cfg2 cfg3]
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import naivesat_amd as ns  # noqa: E402  (generators only; no device use)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs.json")
THREADS = os.cpu_count() or 8
CFG2_QUANTS = [0] * 8 + [1] * 8 + [0] * 8 + [1] * 8  # exists x0..7, all x8..15, exists x16..23, all x24..31


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def cfg_sat(seed, n, gates, wbits):
    c, key = ns.Circuit.gen_sat(seed, n, gates, wbits)
    oc = oracle.OracleCircuit.from_dict(c.arrays())
    t = time.time()
    cnt, low = oc.sat(threads=THREADS)
    return {"seed": seed, "n": n, "gates": gates, "word_bits": wbits, "key": key, "count": cnt, "lowest": low,
            "cone_gates": c.cone()[0], "text_sha256": hashlib.sha256(c.text().encode()).hexdigest(),
            "oracle_s": round(time.time() - t, 2)}


def cfg2():
    c = ns.Circuit.gen_qsat(3, 32, 10000, 200)
    oc = oracle.OracleCircuit.from_dict(c.arrays())
    t = time.time()
    r = oc.qsat(CFG2_QUANTS, threads=THREADS)
    bm = np.packbits(r["bitmap"].astype(np.uint8), bitorder="little")
    return {"seed": 3, "n": 32, "gates": 10000, "last": 200, "quants": "".join("a" if q else "e" for q in CFG2_QUANTS),
            "truth": r["truth"], "witness": r["witness"], "witness_count": r["witness_count"],
            "bitmap_hex": bm.tobytes().hex(), "text_sha256": hashlib.sha256(c.text().encode()).hexdigest(),
            "oracle_s": round(time.time() - t, 2)}


def cfg3():
    c, target = ns.Circuit.gen_satostop(4, 15, 6, 300, 14)
    oc = oracle.OracleCircuit.from_dict(c.arrays())
    t = time.time()
    r = oc.satostop(14, threads=THREADS)
    lo = r["lowest_u"]
    return {"seed": 4, "half": 15, "rounds": 6, "round_gates": 300, "unknowns": 14, "target": target,
            "n_stopped": r["n_stopped"], "lowest_u": lo,
            "steps": int(r["steps"][lo]) if lo is not None else None,
            "end": int(r["ends"][lo]) if lo is not None else None,
            "max_steps": int(r["steps"].max()),
            "steps_sha256": sha(r["steps"]), "ends_sha256": sha(r["ends"]),
            "text_sha256": hashlib.sha256(c.text().encode()).hexdigest(),
            "oracle_s": round(time.time() - t, 2)}


def main():
    want = sys.argv[1:] or ["cfg0", "cfg1", "cfg2", "cfg3"]
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    makers = {"cfg0": lambda: cfg_sat(1, 24, 2000, 32), "cfg1": lambda: cfg_sat(2, 36, 20000, 40),
              "cfg2": cfg2, "cfg3": cfg3}
    for k in want:
        data[k] = makers[k]()
        print(k, data[k], flush=True)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
