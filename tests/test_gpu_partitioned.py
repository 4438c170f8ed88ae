"""greedysatostop's partitioned search with the image on disk (src/greedysatostop.rs:692-837):
device joins against the oracle's restatement of the same partition order."""
import numpy as np
import pytest

import naivesat_amd as ns
import oracle
from helpers import point_stop_circuit, random_circuit

pytestmark = pytest.mark.gpu


def mk(d):
    return ns.Circuit.from_arrays(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])


def expect(d, U, P):
    table = oracle.OracleCircuit.from_dict(d).table_u32().astype(np.uint64)
    return oracle.greedy_partitioned(table, d["n"], U, P)


def got(r):
    return {"start": r["start"], "end": r["end"]} if r["found"] else None


@pytest.mark.parametrize("P", [2, 4, 8])
def test_random_circuits_match_the_restatement(ctx, tmp_path, P):
    rng = np.random.default_rng(9100 + P)
    for t in range(6):
        n = int(rng.integers(10, 17))
        U = int(rng.integers(1, n - 3))
        d = random_circuit(rng, n, int(rng.integers(n + 1, 4 * n + 20)), nout=n + 1, recent=3 * n + 4)
        r = ctx.greedy_partitioned_solve(mk(d), U, P, str(tmp_path) + "/")
        assert got(r) == expect(d, U, P), (t, n, U)
        assert r["partitions"] == P


def test_feistel_and_no_stop(ctx, tmp_path):
    for half, U, P in ((6, 8, 4), (8, 10, 16)):
        c, target = ns.Circuit.gen_satostop(9200 + half, half, 6, 10 * half, U)
        d = c.arrays()
        r = ctx.greedy_partitioned_solve(c, U, P, str(tmp_path) + "/")
        want = expect(d, U, P)
        assert got(r) == want
        if want is not None:
            assert want["end"] == target
    # STOP at a state that is no unknown start: none found, after all n stages
    n, U = 14, 4
    d = point_stop_circuit(n, (3 << (n - U)) + 1, 77)
    r = ctx.greedy_partitioned_solve(mk(d), U, 4, str(tmp_path) + "/")
    assert r["found"] == 0 and r["stages"] == n


def test_reference_mem_image_case_at_26_bits(ctx, tmp_path):
    # the mem-image find_solution case 2 position (src/greedysatostop.rs:1203-1268) in a
    # 26-bit space: a 226 MiB image in 16 partitions on disk
    n, U = 26, 13
    p, nxt = (3 << 22) + 79 * 8192 * 4, 58591
    r = ctx.greedy_partitioned_solve(mk(point_stop_circuit(n, p, nxt)), U, 16, str(tmp_path) + "/")
    assert got(r) == {"start": p, "end": nxt} and r["stages"] == 1
