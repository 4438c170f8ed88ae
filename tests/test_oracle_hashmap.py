"""The oracle's restatement of naivesatostop's hash-map functions against the reference's
own data sets (src/satostop.rs:1917-4144), transcribed to tests/golden/hashmap_cases.json."""
import numpy as np

import oracle
from hashmap_cases import hashmap, load, u32_array


def test_hash_function_closed_form():
    # src/satostop.rs:127-136 for a few widths, with u64 wrapping done by hand
    M = (1 << 64) - 1
    for bits in (24, 32, 40, 44, 63):
        for v in (0, 1, 0x123456, (1 << bits) - 1, 0xdeadbeef & ((1 << bits) - 1)):
            t = (v * 9615409803190489167) & M
            t = ((t << (bits >> 1)) & M) | (t >> (bits - (bits >> 1)))
            want = (((v * 6171710485021949031) & M) ^ t ^ 0xb89d2ecda078ca1f) & ((1 << bits) - 1)
            assert oracle.hash64(bits, v) == want


def test_join_to_hashmap_reference_data():
    for case in load()["join_to_hashmap"]:
        hm = hashmap(case["hashmap"])
        oracle.hm_join_to(case["state_len"], case["arg_bit_place"], case["arg"], u32_array(case["outputs"]), hm)
        assert np.array_equal(hm, hashmap(case["expected_hashmap"])), case["state_len"]


def test_join_hashmap_itself_reference_data():
    case = load()["join_hashmap_itself"]
    hm = hashmap(case["hashmap"])
    out = np.zeros_like(hm)
    out[:] = (1115774, 494022, 450589390239, 489481, 8459021)  # every field is overwritten
    st = oracle.HashMapState(u32_array(case["unknown_fills"]), case["resolved_unknowns"])
    oracle.hm_join_itself(case["state_len"], hm, out, case["unknown_bits"], case["unknown_fill_bits"], st)
    assert np.array_equal(out, hashmap(case["expected_hashmap"]))
    assert np.array_equal(st.unknown_fills, u32_array(case["expected_unknown_fills"]))
    assert st.resolved.value == case["expected_resolved_unknowns"]
    assert st.solution() == case["expected_solution"]


def test_add_to_hashmap_reference_data():
    for case in load()["add_to_hashmap"]:
        hm = hashmap(case["hashmap"])
        st = oracle.HashMapState(u32_array(case["unknown_fills"]), case["resolved_unknowns"])
        oracle.hm_add(case["state_len"], case["arg_bit_place"], case["arg"], u32_array(case["outputs"]), hm,
                      case["unknown_bits"], case["unknown_fill_bits"], st, case["max_predecessors"], True)
        assert np.array_equal(hm, hashmap(case["expected_hashmap"])), case["state_len"]
        assert np.array_equal(st.unknown_fills, u32_array(case["expected_unknown_fills"]))
        assert st.resolved.value == case["expected_resolved_unknowns"]
        assert st.solution() == case["expected_solution"]


def test_add_data_depends_on_the_hash():
    # some inserted states land on slots already holding other states: the expected
    # "not replaced" entries hold only when the hash places them there, so the data set
    # pins hash_function_64 beyond its closed form
    case = load()["add_to_hashmap"][0]
    sl, hb, arg, abp = case["state_len"], case["hbits"], case["arg"], case["arg_bit_place"]
    occupied = {int(k) for k in case["hashmap"]["entries"]}
    hits = [i for i in map(int, case["outputs"]["values"])
            if oracle.hash64(sl, i | (arg << abp)) >> (sl - hb) in occupied]
    assert len(hits) >= 2
