"""The oracle's MemImage restatement (packed (n+1)-bit next-state entries, the reference's
src/greedysatostop.rs:339-475) against the reference's own tests and closed forms."""
import json
import os

import numpy as np

import oracle

M64 = (1 << 64) - 1
# test_mem_image (src/greedysatostop.rs:1145-1201): (state_len, mult, add, len)
MEM_IMAGE_CASES = [(10, 3, 25, 256), (16, 67, 441, 1024), (19, 49414, 9481778, 64 * 80),
                   (31, 96069201, 9669093056, 64 * 222), (32, 139696069201, 9669093056, 64 * 222),
                   (40, 166139696069201, 966771239093056, 64 * 320), (45, 16171069201, 966771239093056, 64 * 400)]


def messy(nwords):
    # the reference's initial content: i * 5884491582115956921 + 5560939029013487713 (wrapping)
    return np.array([(i * 5884491582115956921 + 5560939029013487713) & M64 for i in range(nwords)], np.uint64)


def bits_of(words, width, length):
    """entry i read bit by bit at global bit i*width + j (the reference's closed-form check)"""
    g = np.arange(length, dtype=np.uint64)[:, None] * np.uint64(width) + np.arange(width, dtype=np.uint64)[None, :]
    b = (words[(g >> np.uint64(6)).astype(np.int64)] >> (g & np.uint64(63))) & np.uint64(1)
    return (b << np.arange(width, dtype=np.uint64)[None, :]).sum(axis=1, dtype=np.uint64)


def test_mem_image_reference_cases():
    for k, (sl, mult, add, length) in enumerate(MEM_IMAGE_CASES):
        width = sl + 1
        words = messy(width * (length >> 6))
        assert words.size == (sl + 1) * (length >> 6)
        vals = [(i * mult + add) & M64 for i in range(length)]
        for i, v in enumerate(vals):
            oracle.mem_set(words, width, i, v)
        want = np.array(vals, np.uint64) & np.uint64((1 << width) - 1)
        assert np.array_equal(bits_of(words, width, length), want), k
        assert [oracle.mem_get(words, width, i) for i in range(length)] == want.tolist(), k


def test_mem_image_find_solution_reference_cases():
    with open(os.path.join(os.path.dirname(__file__), "golden", "mem_image_find_solution.json")) as f:
        cases = json.load(f)["cases"]
    for case in cases:
        sl, start, length = case["state_len"], case["start"], case["len"]
        words = oracle.mem_words(sl + 1, length)
        oracle.mem_set(words, sl + 1, case["pos_sol"], (1 << sl) | (case["next"] & ((1 << sl) - 1)))
        assert oracle.mem_find_solution(words, start, length, sl, case["unknowns"]) == case["expected"], case


def test_join_nexts_doubles_a_rotation():
    # next(s) = s + c mod 2^n, nothing stops: one join with the whole image gives s + 2c; a
    # join with a partition moves only the entries whose next state lies in it
    sl, c = 12, 37
    L = 1 << sl
    vals = [(i + c) % L for i in range(L)]
    dest = oracle.mem_pack(sl + 1, vals)
    second = dest.copy()
    oracle.mem_join_nexts(dest, 0, L, second, 0, L, sl)
    assert [oracle.mem_get(dest, sl + 1, i) for i in range(L)] == [(i + 2 * c) % L for i in range(L)]
    part_start, part_len = 1024, 512
    dest = oracle.mem_pack(sl + 1, vals)
    second = oracle.mem_pack(sl + 1, vals[part_start:part_start + part_len])
    oracle.mem_join_nexts(dest, 0, L, second, part_start, part_len, sl)
    for i in range(L):
        moved = part_start <= (i + c) % L < part_start + part_len
        assert oracle.mem_get(dest, sl + 1, i) == ((i + 2 * c) % L if moved else (i + c) % L), i
    # a stopped entry keeps its value
    vals2 = list(vals)
    vals2[5] = (1 << sl) | 99
    dest = oracle.mem_pack(sl + 1, vals2)
    oracle.mem_join_nexts(dest, 0, L, dest.copy(), 0, L, sl)
    assert oracle.mem_get(dest, sl + 1, 5) == (1 << sl) | 99
    assert oracle.mem_get(dest, sl + 1, (5 - c) % L) == (1 << sl) | 99  # its predecessor takes the STOP entry
