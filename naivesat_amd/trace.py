"""Synthetic request stream of the served workload (4,096 requests, seed 5).

Arrivals are Poisson at `rate` requests/s. Each request is SAT (60%), QSAT (25%) or
SATOSTOP (15%). SAT and QSAT circuits have n ~ U{16..32} inputs; SATOSTOP circuits have
2*h state bits with h ~ U{8..13} (16..26 bits). Gate counts are log-uniform in
[200, 20000]. Circuits come from the SAT / QSAT / Feistel SATOSTOP generators with a
per-request seed; QSAT prefixes are 2..5 alternating blocks of random lengths. The draws
come from numpy's PCG64 seeded with `seed`, in the order listed in `make_trace`.
"""
import numpy as np

from . import REQ_QSAT, REQ_SAT, REQ_SATOSTOP, Circuit


STREAM_LEN = 4096


def make_trace(seed: int = 5, n_requests: int = STREAM_LEN, rate: float = 200.0):
    """The first n_requests requests of the stream: arrivals and kinds are drawn for at
    least STREAM_LEN requests, so a shorter trace is a prefix of the full one."""
    rng = np.random.default_rng(seed)
    drawn = max(n_requests, STREAM_LEN)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, drawn))
    kinds = rng.choice([REQ_SAT, REQ_QSAT, REQ_SATOSTOP], size=drawn, p=[0.60, 0.25, 0.15])
    reqs = []
    for i in range(n_requests):
        kind = int(kinds[i])
        gates = int(round(float(np.exp(rng.uniform(np.log(200.0), np.log(20000.0))))))
        cseed = (seed << 32) | i
        if kind == REQ_SATOSTOP:
            half = int(rng.integers(8, 14))
            round_gates = max(half + 1, gates // 6)
            unknowns = half
            c, _ = Circuit.gen_satostop(cseed, half, 6, round_gates, unknowns)
            reqs.append({"kind": kind, "circuit": c, "unknowns": unknowns, "arrival_s": float(arrivals[i]),
                         "n": 2 * half, "gates": 6 * (round_gates + half)})
            continue
        n = int(rng.integers(16, 33))
        if kind == REQ_SAT:
            c, _ = Circuit.gen_sat(cseed, n, gates, min(32, gates))
            reqs.append({"kind": kind, "circuit": c, "arrival_s": float(arrivals[i]), "n": n, "gates": gates})
        else:
            c = Circuit.gen_qsat(cseed, n, gates, min(200, gates))
            blocks = int(rng.integers(2, 6))
            cuts = np.sort(rng.choice(np.arange(1, n), size=blocks - 1, replace=False))
            first = int(rng.integers(0, 2))
            q, prev = [], 0
            for b, cut in enumerate(list(cuts) + [n]):
                q += [first ^ (b & 1)] * int(cut - prev)
                prev = int(cut)
            reqs.append({"kind": kind, "circuit": c, "quants": np.array(q, np.uint8), "arrival_s": float(arrivals[i]),
                         "n": n, "gates": gates})
    return reqs


def latency_stats(requests, results):
    """p50 / p99 of (finish - arrival) in seconds and solves/s over the run."""
    lat = np.array([r["finish_s"] - rq["arrival_s"] for rq, r in zip(requests, results)])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "max_s": float(lat.max())}
