"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker
and by bench.py's CPU-baseline leg. The product (naivesat_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class _Circ(C.Structure):
    _fields_ = [("n", C.c_uint32), ("ngates", C.c_uint32), ("nout", C.c_uint32),
                ("in0", C.c_void_p), ("in1", C.c_void_p), ("func", C.c_void_p),
                ("out_wire", C.c_void_p), ("out_neg", C.c_void_p)]


class _HmSolution(C.Structure):
    _fields_ = [("start", C.c_uint64), ("end", C.c_uint64), ("steps", C.c_uint64), ("defined", C.c_uint32)]


# naivesatostop hash entry (src/satostop.rs:168-174): 32 bytes
HASH_ENTRY = np.dtype([("current", "<u8"), ("next", "<u8"), ("steps", "<u8"), ("predecessors", "<u4"),
                       ("state", "<u4")])

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P, U32, U64, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
        L.or_eval_u64.restype = U64
        L.or_eval_u64.argtypes = [C.POINTER(_Circ), U64]
        L.or_simd_isa.restype = C.c_char_p
        L.or_simd_isa.argtypes = []
        L.or_sat_count.restype = I
        L.or_sat_count.argtypes = [C.POINTER(_Circ), I, C.POINTER(U64), C.POINTER(U64)]
        L.or_sat_count_range.restype = I
        L.or_sat_count_range.argtypes = [C.POINTER(_Circ), U64, U64, I, C.POINTER(U64), C.POINTER(U64)]
        L.or_qr_new.restype = P
        L.or_qr_new.argtypes = [P, U32]
        L.or_qr_push.restype = I
        L.or_qr_push.argtypes = [P, U64, I]
        L.or_qr_final.restype = I
        L.or_qr_final.argtypes = [P, C.POINTER(I), C.POINTER(U32), C.POINTER(I), C.POINTER(U64)]
        L.or_qr_free.restype = None
        L.or_qr_free.argtypes = [P]
        L.or_qsat.restype = I
        L.or_qsat.argtypes = [C.POINTER(_Circ), P, I, C.POINTER(U32), C.POINTER(U64), C.POINTER(U64), P]
        L.or_satostop_walk.restype = U64
        L.or_satostop_walk.argtypes = [C.POINTER(_Circ), U64, U64, C.POINTER(U64)]
        L.or_satostop_table.restype = I
        L.or_satostop_table.argtypes = [C.POINTER(_Circ), U32, I, P, P, C.POINTER(U64), C.POINTER(U64)]
        L.or_table_u32.restype = I
        L.or_table_u32.argtypes = [C.POINTER(_Circ), P]
        L.or_hash64.restype = U64
        L.or_hash64.argtypes = [U32, U64]
        L.or_mem_get.restype = U64
        L.or_mem_get.argtypes = [P, U32, U64]
        L.or_mem_set.restype = None
        L.or_mem_set.argtypes = [P, U32, U64, U64]
        L.or_mem_pack.restype = None
        L.or_mem_pack.argtypes = [P, U64, U32, P]
        L.or_mem_unpack.restype = None
        L.or_mem_unpack.argtypes = [P, U64, U32, P]
        L.or_mem_join_nexts.restype = None
        L.or_mem_join_nexts.argtypes = [P, U64, U64, P, U64, U64, U32]
        L.or_mem_find_solution.restype = C.c_int
        L.or_mem_find_solution.argtypes = [P, U64, U64, U32, U32, C.POINTER(U64), C.POINTER(U64)]
        L.or_hm_join_to.restype = None
        L.or_hm_join_to.argtypes = [U32, U32, U64, P, P, U64]
        L.or_hm_join_itself.restype = None
        L.or_hm_join_itself.argtypes = [U32, P, P, U64, U32, U32, P, C.POINTER(U64), C.POINTER(_HmSolution)]
        L.or_hm_add.restype = None
        L.or_hm_add.argtypes = [U32, U32, U64, P, P, U64, U32, U32, P, C.POINTER(U64), C.POINTER(_HmSolution),
                                U32, I]
        _lib = L
    return _lib


def simd_isa() -> str:
    """the vector ISA of the oracle's bit-sliced SAT evaluation on this CPU"""
    return lib().or_simd_isa().decode()


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class OracleCircuit:
    """Circuit held as numpy arrays (the oracle's own copy of the netlist)."""

    def __init__(self, n, in0, in1, func, out_wire, out_neg=None):
        self.n = int(n)
        self.in0 = np.ascontiguousarray(in0, np.uint32)
        self.in1 = np.ascontiguousarray(in1, np.uint32)
        self.func = np.ascontiguousarray(func, np.uint8)
        self.out_wire = np.ascontiguousarray(out_wire, np.uint32)
        self.out_neg = np.ascontiguousarray(out_neg if out_neg is not None else np.zeros(len(out_wire)),
                                            np.uint8)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        self._c = _Circ(self.n, len(self.in0), len(self.out_wire), p(self.in0), p(self.in1), p(self.func),
                        p(self.out_wire), p(self.out_neg))

    @classmethod
    def from_dict(cls, d):
        return cls(d["n"], d["in0"], d["in1"], d["func"], d["out_wire"], d["out_neg"])

    def eval(self, x: int) -> int:
        return int(lib().or_eval_u64(C.byref(self._c), x))

    def sat(self, threads: int = 8):
        cnt, low = C.c_uint64(), C.c_uint64()
        if lib().or_sat_count(C.byref(self._c), threads, C.byref(cnt), C.byref(low)) != 0:
            raise ValueError("unsupported circuit for the SAT oracle")
        return cnt.value, (low.value if cnt.value else None)

    def sat_range(self, w_lo: int, w_hi: int, threads: int = 8):
        cnt, low = C.c_uint64(), C.c_uint64()
        lib().or_sat_count_range(C.byref(self._c), w_lo, w_hi, threads, C.byref(cnt), C.byref(low))
        return cnt.value, (low.value if cnt.value else None)

    def qsat(self, quants_all, threads: int = 8, bitmap: bool = True):
        q = np.ascontiguousarray(quants_all, np.uint8)
        fb = 0
        while fb < len(q) and q[fb] == q[0]:
            fb += 1
        bm = np.zeros(max(1, (1 << fb) // 8), np.uint8) if bitmap else None
        t, wc, lo = C.c_uint32(), C.c_uint64(), C.c_uint64()
        if lib().or_qsat(C.byref(self._c), q.ctypes.data_as(C.c_void_p), threads, C.byref(t), C.byref(wc),
                         C.byref(lo), None if bm is None else bm.ctypes.data_as(C.c_void_p)) != 0:
            raise ValueError("unsupported circuit for the QSAT oracle")
        d = {"truth": t.value, "witness_count": wc.value, "witness": lo.value if wc.value else None,
             "first_bits": fb, "first_all": int(q[0])}
        if bm is not None:
            d["bitmap"] = np.unpackbits(bm, bitorder="little")[: 1 << fb]
        return d

    def walk(self, start: int, max_steps: int):
        end = C.c_uint64()
        s = lib().or_satostop_walk(C.byref(self._c), start, max_steps, C.byref(end))
        return int(s), int(end.value)

    def satostop(self, unknowns: int, threads: int = 8):
        steps = np.zeros(1 << unknowns, np.uint32)
        ends = np.zeros(1 << unknowns, np.uint32)
        ns_, lo = C.c_uint64(), C.c_uint64()
        r = lib().or_satostop_table(C.byref(self._c), unknowns, threads, steps.ctypes.data_as(C.c_void_p),
                                    ends.ctypes.data_as(C.c_void_p), C.byref(ns_), C.byref(lo))
        if r != 0:
            raise ValueError("unsupported circuit for the SATOSTOP oracle")
        return {"steps": steps, "ends": ends, "n_stopped": ns_.value,
                "lowest_u": lo.value if ns_.value else None}

    def table_u32(self):
        t = np.zeros(1 << self.n, np.uint32)
        if lib().or_table_u32(C.byref(self._c), t.ctypes.data_as(C.c_void_p)) != 0:
            raise ValueError("unsupported circuit for the table oracle")
        return t


class QuantReducer:
    """The reference's streaming QuantReducer (src/qsat.rs:172-291), restated in C."""

    def __init__(self, quants_all):
        q = np.ascontiguousarray(quants_all, np.uint8)
        self._q = q
        self._h = lib().or_qr_new(q.ctypes.data_as(C.c_void_p), len(q))
        if not self._h:
            raise ValueError("bad quantifier list")

    def push(self, index: int, item: bool) -> None:
        if lib().or_qr_push(self._h, index, 1 if item else 0) != 0:
            raise AssertionError("push after the end")

    def final_result(self):
        rev, bits, has, sol = C.c_int(), C.c_uint32(), C.c_int(), C.c_uint64()
        if not lib().or_qr_final(self._h, C.byref(rev), C.byref(bits), C.byref(has), C.byref(sol)):
            return None
        return {"reversed": bool(rev.value), "solution_bits": bits.value,
                "solution": sol.value if has.value else None}

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_qr_free(self._h)
            self._h = None


def hash64(bits: int, value: int) -> int:
    return int(lib().or_hash64(bits, value))


class HashMapState:
    """Unknown-fill counters, resolved-unknown count and solution of the hash-map path."""

    def __init__(self, unknown_fills: np.ndarray, resolved: int = 0):
        self.unknown_fills = np.ascontiguousarray(unknown_fills, np.uint32)
        self.resolved = C.c_uint64(resolved)
        self.sol = _HmSolution()

    def solution(self):
        s = self.sol
        return {"start": s.start, "end": s.end, "steps": s.steps} if s.defined else None


def hm_join_to(state_len, arg_bit_place, arg, outputs, hm):
    lib().or_hm_join_to(state_len, arg_bit_place, arg, outputs.ctypes.data_as(C.c_void_p),
                        hm.ctypes.data_as(C.c_void_p), hm.size)


def hm_join_itself(state_len, hm_in, hm_out, unknown_bits, unknown_fill_bits, st: HashMapState):
    lib().or_hm_join_itself(state_len, hm_in.ctypes.data_as(C.c_void_p), hm_out.ctypes.data_as(C.c_void_p),
                            hm_in.size, unknown_bits, unknown_fill_bits,
                            st.unknown_fills.ctypes.data_as(C.c_void_p), C.byref(st.resolved), C.byref(st.sol))


def hm_add(state_len, arg_bit_place, arg, outputs, hm, unknown_bits, unknown_fill_bits, st: HashMapState,
           max_predecessors, test):
    lib().or_hm_add(state_len, arg_bit_place, arg, outputs.ctypes.data_as(C.c_void_p),
                    hm.ctypes.data_as(C.c_void_p), hm.size, unknown_bits, unknown_fill_bits,
                    st.unknown_fills.ctypes.data_as(C.c_void_p), C.byref(st.resolved), C.byref(st.sol),
                    max_predecessors, 1 if test else 0)


# ---- MemImage (src/greedysatostop.rs:339-475) ----
def mem_words(width: int, length: int) -> np.ndarray:
    """zeroed storage of `length` entries (a multiple of 64) of `width` bits"""
    assert length % 64 == 0
    return np.zeros(length * width // 64, np.uint64)


def mem_set(words: np.ndarray, width: int, i: int, value: int) -> None:
    lib().or_mem_set(words.ctypes.data_as(C.c_void_p), width, i, value & ((1 << 64) - 1))


def mem_get(words: np.ndarray, width: int, i: int) -> int:
    return int(lib().or_mem_get(words.ctypes.data_as(C.c_void_p), width, i))


def mem_pack(width: int, values) -> np.ndarray:
    v = np.ascontiguousarray(values, np.uint64)
    w = mem_words(width, v.size)
    lib().or_mem_pack(v.ctypes.data_as(C.c_void_p), v.size, width, w.ctypes.data_as(C.c_void_p))
    return w


def mem_unpack(width: int, words: np.ndarray, count: int) -> np.ndarray:
    v = np.zeros(count, np.uint64)
    lib().or_mem_unpack(words.ctypes.data_as(C.c_void_p), count, width, v.ctypes.data_as(C.c_void_p))
    return v


def greedy_partitioned(table: np.ndarray, state_len: int, unknowns: int, partitions: int):
    """do_solve_with_cpu_builder_with_partitions (src/greedysatostop.rs:692-837) on a full
    next-state table (entries next | STOP << n): the image split into partitions, every
    stage joining each partition in place with every partition in order (itself as a copy
    of its current state, the others as last saved), then searching it. Returns
    {start, end} or None."""
    width = state_len + 1
    stop = 1 << state_len
    if not np.any(np.asarray(table, np.uint64) & np.uint64(stop)):
        return None
    part_len = (1 << state_len) // partitions
    parts = [mem_pack(width, np.asarray(table[i * part_len:(i + 1) * part_len], np.uint64))
             for i in range(partitions)]
    for _stage in range(state_len):
        for i in range(partitions):
            dest = parts[i].copy()
            for j in range(partitions):
                second = parts[j].copy() if j != i else dest.copy()
                mem_join_nexts(dest, i * part_len, part_len, second, j * part_len, part_len, state_len)
            sol = mem_find_solution(dest, i * part_len, part_len, state_len, unknowns)
            if sol is not None:
                return sol
            parts[i] = dest
    return None


def mem_join_nexts(dest: np.ndarray, dest_start: int, length: int, second: np.ndarray, second_start: int,
                   second_len: int, state_len: int) -> None:
    lib().or_mem_join_nexts(dest.ctypes.data_as(C.c_void_p), dest_start, length,
                            second.ctypes.data_as(C.c_void_p), second_start, second_len, state_len)


def mem_find_solution(words: np.ndarray, start: int, length: int, state_len: int, unknowns: int):
    s, e = C.c_uint64(), C.c_uint64()
    ok = lib().or_mem_find_solution(words.ctypes.data_as(C.c_void_p), start, length, state_len, unknowns,
                                    C.byref(s), C.byref(e))
    return {"start": s.value, "end": e.value} if ok else None
