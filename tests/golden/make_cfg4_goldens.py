"""Oracle answers for every request of the served stream (4,096 requests, seed 5,
naivesat_amd.trace.make_trace), written to tests/golden/cfg4_answers.json.

One row per request in trace order: [kind, n, a, b, c, d] with
  SAT       a = satisfying count, b = lowest assignment (null when none)
  QSAT      a = truth, b = witness count, c = lowest witness (null when none)
  SATOSTOP  a = stopping unknowns, b = lowest stopping unknown, c = its evaluation count,
            d = its end state (b, c, d null when none stops)
The oracle is the C restatement under oracle/ (test infrastructure). About half an hour on
eight cores.

Usage: python tests/golden/make_cfg4_goldens.py [--limit N] [--threads T]
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import naivesat_amd as ns  # noqa: E402
import oracle  # noqa: E402
from naivesat_amd import trace  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg4_answers.json")


def answer(rq, threads):
    oc = oracle.OracleCircuit.from_dict(rq["circuit"].arrays())
    if rq["kind"] == ns.REQ_SAT:
        count, lowest = oc.sat(threads=threads)
        return [count, lowest, None, None]
    if rq["kind"] == ns.REQ_QSAT:
        o = oc.qsat(list(rq["quants"]), threads=threads, bitmap=False)
        return [o["truth"], o["witness_count"], o["witness"], None]
    o = oc.satostop(rq["unknowns"], threads=threads)
    lo = o["lowest_u"]
    if lo is None:
        return [o["n_stopped"], None, None, None]
    return [o["n_stopped"], lo, int(o["steps"][lo]), int(o["ends"][lo])]


def rows_digest(rows):
    return hashlib.sha256(json.dumps(rows, separators=(",", ":")).encode()).hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--limit", type=int, default=trace.STREAM_LEN)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 8)
    args = ap.parse_args()
    reqs = trace.make_trace(seed=5, n_requests=args.limit)
    rows = []
    t0 = time.time()
    for i, rq in enumerate(reqs):
        rows.append([rq["kind"], rq["n"]] + answer(rq, args.threads))
        if i % 256 == 255:
            print(f"{i + 1} requests, {time.time() - t0:.0f} s", flush=True)
    doc = {"trace": {"seed": 5, "n_requests": len(reqs), "rate": 200.0},
           "columns": ["kind", "n", "a", "b", "c", "d"], "rows": rows, "sha256": rows_digest(rows)}
    with open(OUT + ".tmp", "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    os.replace(OUT + ".tmp", OUT)
    print(f"wrote {OUT}: {len(rows)} rows in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
