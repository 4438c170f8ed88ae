"""greedysatostop's on-disk image (FileImage, src/greedysatostop.rs:478-594) through the C ABI
(host only), the reference's test_file_image (1270-1352), and the oracle's restatement of the
partitioned search (692-837) against properties that pin its answers."""
import json
import os

import numpy as np

import naivesat_amd as ns
import oracle
from helpers import point_stop_circuit, random_circuit

M64 = (1 << 64) - 1


def seq(start, count, mult, add):
    i = np.arange(start, start + count, dtype=np.uint64)
    return i * np.uint64(mult) + np.uint64(add)  # wrapping u64 arithmetic


def test_file_image_reference_case(tmp_path):
    # (state_len, mult, add, partitions, changer) = (26, 4849217, 3455641, 16, 9450290114)
    sl, mult, add, partitions, changer = 26, 4849217, 3455641, 16, 9450290114
    width, mask = sl + 1, np.uint64((1 << (sl + 1)) - 1)
    fi = ns.FileImage(sl, partitions, str(tmp_path) + "/")
    assert os.path.basename(fi_path(tmp_path)).startswith("greedy_temp_")
    save_parts = partitions << 1
    chunk_len = (1 << sl) // save_parts
    for part in range(save_parts):
        fi.save_chunk(oracle.mem_pack(width, seq(chunk_len * part, chunk_len, mult, add)))
    chunk_len = (1 << sl) // partitions
    for part in reversed(range(partitions)):
        got = oracle.mem_unpack(width, fi.load_partition(part), chunk_len)
        assert np.array_equal(got, seq(chunk_len * part, chunk_len, mult, add) & mask), part
    for part in reversed(range(partitions)):
        fi.save_partition(part, oracle.mem_pack(width, seq(chunk_len * part + changer, chunk_len, mult, add)))
    for part in reversed(range(partitions)):
        got = oracle.mem_unpack(width, fi.load_partition(part), chunk_len)
        assert np.array_equal(got, seq(chunk_len * part + changer, chunk_len, mult, add) & mask), part
    # the file holds the reference's bytes: partition p at byte p * partition_len
    path = fi_path(tmp_path)
    raw = np.fromfile(path, dtype="<u8")
    assert raw.size == partitions * fi.partition_words
    assert np.array_equal(raw[3 * fi.partition_words:4 * fi.partition_words], fi.load_partition(3))
    fi.close()
    assert not os.path.exists(path)  # removed with the image


def fi_path(tmp_path):
    names = [n for n in os.listdir(tmp_path) if n.startswith("greedy_temp_")]
    assert len(names) == 1
    return os.path.join(tmp_path, names[0])


def test_file_image_rejects_bad_shapes(tmp_path):
    for sl, parts in ((20, 3), (20, 0), (10, 32)):
        try:
            ns.FileImage(sl, parts, str(tmp_path) + "/")
        except ns.NaivesatError:
            continue
        raise AssertionError((sl, parts))
    fi = ns.FileImage(12, 4, str(tmp_path) + "/")
    try:
        fi.load_partition(0)  # nothing written yet
    except ns.NaivesatError:
        pass
    else:
        raise AssertionError("loaded an incomplete image")


def table_of(d):
    return oracle.OracleCircuit.from_dict(d).table_u32().astype(np.uint64)


def test_partitioned_restatement_mem_image_cases():
    # the reference's find_solution cases as full tables: the one stopping state is found in
    # the first stage when it is an unknown start, never otherwise
    with open(os.path.join(os.path.dirname(__file__), "golden", "mem_image_find_solution.json")) as f:
        cases = json.load(f)["cases"]
    for case in cases[:4]:
        n, U = case["state_len"], case["unknowns"]
        p = case["start"] + case["pos_sol"]
        if n > 20:  # the same positions in a 20-bit space
            p >>= n - 20
            U = min(U, 20)
            n = 20
        table = table_of(point_stop_circuit(n, p, case["next"] & ((1 << n) - 1)))
        sol = oracle.greedy_partitioned(table, n, U, 4)
        is_start = p & ((1 << (n - U)) - 1) == 0
        assert sol == ({"start": p, "end": case["next"] & ((1 << n) - 1)} if is_start else None), case


def test_partitioned_restatement_answers_are_stopping_starts():
    # whatever start the partition order reaches first, it stops at the reported end (a walk
    # on the circuit), and there is none exactly when no unknown start stops
    rng = np.random.default_rng(8800)
    for t in range(12):
        n = int(rng.integers(10, 15))
        U = int(rng.integers(1, n - 4))
        d = random_circuit(rng, n, int(rng.integers(n + 1, 4 * n + 20)), nout=n + 1, recent=3 * n + 4)
        oc = oracle.OracleCircuit.from_dict(d)
        full = oc.satostop(U, threads=4)
        for P in (2, 4, 8):
            sol = oracle.greedy_partitioned(table_of(d), n, U, P)
            assert (sol is None) == (full["n_stopped"] == 0), (t, P)
            if sol is not None:
                u = sol["start"] >> (n - U)
                assert sol["start"] == u << (n - U) and full["steps"][u] > 0, (t, P)
                assert sol["end"] == full["ends"][u], (t, P)
