"""MI355X (gfx950) brute-force SAT / QSAT / SATOSTOP engine — Python host binding.

Thin ctypes layer over ``libnaivesat_hip.so`` (the C ABI declared in
``include/naivesat_hip.h``). There is no CPU fallback: every solve runs the HIP
kernels, and a missing library or device raises.

The solver entry points mirror the reference's search runs:
  * ``Context.sat_solve``      — naivesat (src/main.rs:157-201, 266-382)
  * ``Context.qsat_solve``     — naiveqsat (src/qsat.rs:1658-1720, 1838-2000)
  * ``Context.satostop_solve`` — greedysatostop's table + pointer jumping
    (src/greedysatostop.rs:1042-1099) and naivesatostop's answer
    (src/satostop.rs:322-327, 1650-1782)
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libnaivesat_hip.so")
BIN_DIR = os.path.join(_HERE, "bin")

NS_SAT_STOP_AT_FIRST = 1
NS_SAT_GROUP_EXITS = 2
NS_ERR_CANCELLED = 7
SAT_STEP_START = 0xFFFFFFFFFFFFFFFF  # NS_SAT_STEP_START: the progress call before the first launch
NS_STOP_FIRST_PASS = 1
NS_STOP_PASSES = 2
NS_STOP_TABLE_WALKS = 4
NS_STOP_BIT_WALKS = 8
FUNC_NAMES = ("and", "nor", "nimpl", "xor")


class NaivesatError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"naivesat_hip status {status}: {msg}")
        self.status = status


class _SatOpts(C.Structure):
    _fields_ = [("elem_bits", C.c_uint32), ("group_len", C.c_uint32),
                ("words_per_lane", C.c_uint32), ("flags", C.c_uint32)]


_SatProgressFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32)
_HmProgressFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64)


class _SatResult(C.Structure):
    _fields_ = [("count", C.c_uint64), ("lowest", C.c_uint64), ("found", C.c_uint32),
                ("complete", C.c_uint32), ("launches", C.c_uint32), ("cone_gates", C.c_uint32),
                ("valu_ops", C.c_uint32), ("scalar_ops", C.c_uint32), ("backend", C.c_uint32),
                ("vgpr_class", C.c_uint32), ("kernel_ms", C.c_double), ("compile_ms", C.c_double), ("wall_ms", C.c_double)]


_QsatOpts = _SatOpts


class _QsatResult(C.Structure):
    _fields_ = [("truth", C.c_uint32), ("first_bits", C.c_uint32), ("first_all", C.c_uint32),
                ("has_witness", C.c_uint32), ("witness", C.c_uint64), ("witness_count", C.c_uint64),
                ("passes", C.c_uint32), ("cone_gates", C.c_uint32), ("valu_ops", C.c_uint32),
                ("scalar_ops", C.c_uint32), ("backend", C.c_uint32), ("vgpr_class", C.c_uint32),
                ("kernel_ms", C.c_double),
                ("compile_ms", C.c_double), ("wall_ms", C.c_double)]


class _StopOpts(C.Structure):
    _fields_ = [("flags", C.c_uint32), ("max_passes", C.c_uint32), ("group_len", C.c_uint32),
                ("words_per_lane", C.c_uint32)]


class _StopResult(C.Structure):
    _fields_ = [("any_stop", C.c_uint32), ("passes", C.c_uint32), ("n_stopped", C.c_uint64),
                ("lowest_u", C.c_uint64), ("start", C.c_uint64), ("end", C.c_uint64),
                ("steps", C.c_uint64), ("cone_gates", C.c_uint32), ("valu_ops", C.c_uint32),
                ("backend", C.c_uint32), ("vgpr_class", C.c_uint32),
                ("build_ms", C.c_double), ("jump_ms", C.c_double), ("compile_ms", C.c_double),
                ("wall_ms", C.c_double), ("walk_jumps", C.c_uint64), ("walk_ms", C.c_double),
                ("method", C.c_uint32), ("waypoint_bits", C.c_uint32), ("waypoint_passes", C.c_uint32),
                ("reserved", C.c_uint32), ("walk_evaluations", C.c_uint64)]


class _PartResult(C.Structure):
    _fields_ = [("found", C.c_uint32), ("stages", C.c_uint32), ("start", C.c_uint64), ("end", C.c_uint64),
                ("partitions", C.c_uint64), ("build_ms", C.c_double), ("join_ms", C.c_double),
                ("io_ms", C.c_double), ("compile_ms", C.c_double), ("wall_ms", C.c_double)]


class _HmOpts(C.Structure):
    _fields_ = [("hashmap_len_bits", C.c_uint32), ("elem_bits", C.c_uint32), ("unknown_fill_bits", C.c_int32),
                ("max_predecessors", C.c_uint32), ("iters_per_clear_predecessors", C.c_uint32),
                ("max_sweeps", C.c_uint32)]


class _HmResult(C.Structure):
    _fields_ = [("found", C.c_uint32), ("sweeps", C.c_uint32), ("start", C.c_uint64), ("end", C.c_uint64),
                ("steps", C.c_uint64), ("resolved_unknowns", C.c_uint64), ("blocks", C.c_uint64),
                ("compile_ms", C.c_double), ("kernel_ms", C.c_double), ("wall_ms", C.c_double)]


class _Request(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("unknowns", C.c_uint32), ("circuit", C.c_void_p),
                ("quants_all", C.c_void_p), ("arrival_s", C.c_double)]


class _RequestResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("worker", C.c_uint32), ("v0", C.c_uint64), ("v1", C.c_uint64),
                ("v2", C.c_uint64), ("v3", C.c_uint64), ("start_s", C.c_double), ("finish_s", C.c_double),
                ("compile_ms", C.c_double), ("kernel_ms", C.c_double), ("submit_s", C.c_double)]


class _ServeOpts(C.Structure):
    _fields_ = [("workers", C.c_uint32), ("realtime", C.c_uint32), ("hbm_budget_bytes", C.c_uint64),
                ("batch_max", C.c_uint32), ("reserved", C.c_uint32)]


REQ_SAT, REQ_QSAT, REQ_SATOSTOP = 0, 1, 2

_lib = None


def load_library() -> C.CDLL:
    """Loads libnaivesat_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    P, U32, U64, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    sig = {
        "ns_last_error": (C.c_char_p, []),
        "ns_version": (C.c_char_p, []),
        "ns_circuit_parse": (I, [C.c_char_p, C.c_size_t, C.POINTER(P)]),
        "ns_quant_circuit_parse": (I, [C.c_char_p, C.c_size_t, C.POINTER(P), P, C.c_size_t, C.POINTER(I)]),
        "ns_circuit_from_arrays": (I, [U32, U32, P, P, P, U32, P, P, C.POINTER(P)]),
        "ns_circuit_info": (I, [P, C.POINTER(U32), C.POINTER(U32), C.POINTER(U32)]),
        "ns_circuit_to_arrays": (I, [P, P, P, P, P, P]),
        "ns_circuit_format": (I, [P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "ns_circuit_eval": (I, [P, P, P]),
        "ns_circuit_cone": (I, [P, C.POINTER(U32), C.POINTER(U32)]),
        "ns_circuit_free": (None, [P]),
        "ns_gen_sat": (I, [U64, U32, U32, U32, C.POINTER(P), C.POINTER(U64)]),
        "ns_gen_qsat": (I, [U64, U32, U32, U32, C.POINTER(P)]),
        "ns_gen_factor": (I, [U64, U32, C.POINTER(P)]),
        "ns_gen_satostop": (I, [U64, U32, U32, U32, U32, C.POINTER(P), C.POINTER(U64)]),
        "ns_ctx_create": (I, [I, C.POINTER(P)]),
        "ns_ctx_destroy": (None, [P]),
        "ns_sat_solve": (I, [P, P, C.POINTER(_SatOpts), C.POINTER(_SatResult)]),
        "ns_sat_solve_noalloc": (I, [P, P, C.POINTER(_SatOpts), C.POINTER(_SatResult)]),
        "ns_sat_solve_progress": (I, [P, P, C.POINTER(_SatOpts), _SatProgressFn, P, C.POINTER(_SatResult)]),
        "ns_qsat_solve": (I, [P, P, P, C.POINTER(_QsatOpts), C.POINTER(_QsatResult), P, U64]),
        "ns_qsat_solve_progress": (I, [P, P, P, C.POINTER(_QsatOpts), _SatProgressFn, P, C.POINTER(_QsatResult),
                                       P, U64]),
        "ns_device_list": (I, [C.POINTER(U32), C.c_char_p, C.c_size_t]),
        "ns_satostop_solve": (I, [P, P, U32, C.POINTER(_StopOpts), C.POINTER(_StopResult), P, P]),
        "ns_satostop_solve_wide": (I, [P, P, U32, C.POINTER(_StopOpts), C.POINTER(_StopResult), P, P]),
        "ns_satostop_hashmap_solve": (I, [P, P, U32, C.POINTER(_HmOpts), C.POINTER(_HmResult)]),
        "ns_satostop_hashmap_solve_progress": (I, [P, P, U32, C.POINTER(_HmOpts), _HmProgressFn, P,
                                                   C.POINTER(_HmResult)]),
        "ns_hashmap_step": (I, [P, U32, U32, U32, U64, U32, U32, U32, U32, P, U64, P, P, U64, P, U64, P]),
        "ns_table_build": (I, [P, P, P, U64, C.POINTER(C.c_double)]),
        "ns_satostop_pass_times": (I, [P, P, U32, C.POINTER(U32)]),
        "ns_native_body": (I, [P, U32, U32, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(U32),
                               C.POINTER(U64)]),
        "ns_serve": (I, [P, C.POINTER(_Request), U32, C.POINTER(_ServeOpts), C.POINTER(_RequestResult),
                         C.POINTER(C.c_double)]),
        "ns_serve_open": (I, [P, C.POINTER(_ServeOpts), C.POINTER(P)]),
        "ns_serve_submit": (I, [P, C.POINTER(_Request), C.POINTER(U64)]),
        "ns_serve_poll": (I, [P, U64, C.POINTER(_RequestResult), C.POINTER(I)]),
        "ns_serve_wait": (I, [P, U64, C.c_double, C.POINTER(_RequestResult), C.POINTER(I)]),
        "ns_serve_close": (I, [P]),
        "ns_batch_solve": (I, [P, C.POINTER(_Request), U32, C.POINTER(_RequestResult), P]),
        "ns_native_encode": (I, [C.c_char_p, C.c_size_t, P, C.c_size_t, C.POINTER(C.c_size_t)]),
        "ns_native_code": (I, [P, U32, U32, P, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(U32),
                               C.POINTER(U64)]),
        "ns_probe_gather": (I, [P, U32, U32, U32, U32, C.POINTER(C.c_double), C.POINTER(U64)]),
        "ns_pack_entries": (I, [P, U32, P, U64, P, U64]),
        "ns_ctx_trim": (I, [P]),
        "ns_greedy_partitioned_solve": (I, [P, P, U32, U64, C.c_char_p, C.POINTER(_PartResult)]),
        "ns_file_image_create": (I, [U32, U64, C.c_char_p, C.POINTER(P)]),
        "ns_file_image_save_chunk": (I, [P, P, U64]),
        "ns_file_image_load_partition": (I, [P, U64, P, U64]),
        "ns_file_image_save_partition": (I, [P, U64, P, U64]),
        "ns_file_image_free": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list:
    """Names of the C-ABI entry points this binding declares."""
    return ["ns_last_error", "ns_version", "ns_circuit_parse", "ns_quant_circuit_parse",
            "ns_circuit_from_arrays", "ns_circuit_info", "ns_circuit_to_arrays", "ns_circuit_format",
            "ns_circuit_eval", "ns_circuit_cone", "ns_circuit_free", "ns_gen_sat", "ns_gen_qsat", "ns_gen_factor",
            "ns_gen_satostop", "ns_ctx_create", "ns_ctx_destroy", "ns_sat_solve", "ns_qsat_solve",
            "ns_satostop_solve", "ns_table_build", "ns_satostop_pass_times", "ns_native_body", "ns_serve", "ns_native_encode",
            "ns_native_code", "ns_satostop_solve_wide", "ns_satostop_hashmap_solve", "ns_hashmap_step",
            "ns_probe_gather", "ns_pack_entries", "ns_ctx_trim", "ns_greedy_partitioned_solve",
            "ns_file_image_create", "ns_file_image_save_chunk", "ns_file_image_load_partition",
            "ns_file_image_save_partition", "ns_file_image_free", "ns_serve_open", "ns_serve_submit",
            "ns_serve_poll", "ns_serve_wait", "ns_serve_close", "ns_batch_solve", "ns_sat_solve_noalloc",
            "ns_sat_solve_progress", "ns_device_list", "ns_qsat_solve_progress",
            "ns_satostop_hashmap_solve_progress"]


def device_list() -> list:
    """ns_device_list: the gfx architecture name of each HIP device ([] without one)."""
    n = C.c_uint32(0)
    buf = C.create_string_buffer(4096)
    _check(load_library().ns_device_list(C.byref(n), buf, 4096))
    names = buf.value.decode()
    return names.split(",") if n.value else []


def _check(status: int) -> None:
    if status != 0:
        msg = load_library().ns_last_error().decode(errors="replace")
        raise NaivesatError(status, msg)


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Circuit:
    """Gate circuit (gatesim model: inputs 0..n-1, gate j is wire n+j, outputs (wire, neg))."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.ns_circuit_free(h)
            self._h = None

    # ---- constructors ----
    @classmethod
    def parse(cls, text: str) -> "Circuit":
        lib = load_library()
        h = C.c_void_p()
        b = text.encode()
        _check(lib.ns_circuit_parse(b, len(b), C.byref(h)))
        return cls(h.value)

    @classmethod
    def parse_quant(cls, text: str):
        """Returns (circuit, quants_all or None) for "eaa {..}" / plain text."""
        lib = load_library()
        h = C.c_void_p()
        b = text.encode()
        q = np.zeros(4096, dtype=np.uint8)
        isq = C.c_int(0)
        _check(lib.ns_quant_circuit_parse(b, len(b), C.byref(h), _ptr(q), q.size, C.byref(isq)))
        c = cls(h.value)
        return c, (q[: c.n_inputs].copy() if isq.value else None)

    @classmethod
    def from_arrays(cls, n: int, in0: Sequence[int], in1: Sequence[int], func: Sequence[int],
                    out_wire: Sequence[int], out_neg: Optional[Sequence[int]] = None) -> "Circuit":
        lib = load_library()
        a0 = np.ascontiguousarray(in0, dtype=np.uint32)
        a1 = np.ascontiguousarray(in1, dtype=np.uint32)
        f = np.ascontiguousarray(func, dtype=np.uint8)
        ow = np.ascontiguousarray(out_wire, dtype=np.uint32)
        on = np.ascontiguousarray(out_neg if out_neg is not None else np.zeros(len(ow)), dtype=np.uint8)
        if not (len(a0) == len(a1) == len(f)) or len(on) != len(ow):
            raise ValueError("array lengths differ")
        h = C.c_void_p()
        _check(lib.ns_circuit_from_arrays(n, len(a0), _ptr(a0), _ptr(a1), _ptr(f), len(ow), _ptr(ow),
                                          _ptr(on), C.byref(h)))
        return cls(h.value)

    @classmethod
    def gen_sat(cls, seed: int, n: int, gates: int, word_bits: int):
        """SAT workload: output = [W(x) == W(key)], W = last word_bits gates. Returns (c, key)."""
        lib = load_library()
        h = C.c_void_p()
        key = C.c_uint64()
        _check(lib.ns_gen_sat(seed, n, gates, word_bits, C.byref(h), C.byref(key)))
        return cls(h.value), key.value

    @classmethod
    def gen_qsat(cls, seed: int, n: int, gates: int, last: int) -> "Circuit":
        lib = load_library()
        h = C.c_void_p()
        _check(lib.ns_gen_qsat(seed, n, gates, last, C.byref(h)))
        return cls(h.value)

    @classmethod
    def gen_factor(cls, product: int, bits: int):
        """Factoring circuit: output = [a * b == product], a != 1, b != 1 for a = inputs
        0..bits-1 and b = inputs bits..2 bits - 1."""
        h = C.c_void_p()
        _check(load_library().ns_gen_factor(product, bits, C.byref(h)))
        return cls(h.value)

    @classmethod
    def gen_satostop(cls, seed: int, half: int, rounds: int, round_gates: int, unknowns: int):
        """Feistel next-state circuit with STOP = [next == target]. Returns (c, target)."""
        lib = load_library()
        h = C.c_void_p()
        t = C.c_uint64()
        _check(lib.ns_gen_satostop(seed, half, rounds, round_gates, unknowns, C.byref(h), C.byref(t)))
        return cls(h.value), t.value

    # ---- accessors ----
    def _info(self):
        n, g, o = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(load_library().ns_circuit_info(self._h, C.byref(n), C.byref(g), C.byref(o)))
        return n.value, g.value, o.value

    @property
    def n_inputs(self) -> int:
        return self._info()[0]

    @property
    def n_gates(self) -> int:
        return self._info()[1]

    @property
    def n_outputs(self) -> int:
        return self._info()[2]

    def arrays(self) -> dict:
        n, g, o = self._info()
        in0 = np.zeros(g, np.uint32)
        in1 = np.zeros(g, np.uint32)
        func = np.zeros(g, np.uint8)
        ow = np.zeros(o, np.uint32)
        on = np.zeros(o, np.uint8)
        _check(load_library().ns_circuit_to_arrays(self._h, _ptr(in0), _ptr(in1), _ptr(func), _ptr(ow),
                                                   _ptr(on)))
        return {"n": n, "in0": in0, "in1": in1, "func": func, "out_wire": ow, "out_neg": on}

    def text(self) -> str:
        lib = load_library()
        need = C.c_size_t()
        _check(lib.ns_circuit_format(self._h, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value + 1)
        _check(lib.ns_circuit_format(self._h, buf, need.value + 1, C.byref(need)))
        return buf.value.decode()

    def eval(self, inputs: Sequence[int]) -> list:
        n, _, o = self._info()
        a = np.ascontiguousarray(inputs, dtype=np.uint8)
        if a.size != n:
            raise ValueError("one value per input is required")
        out = np.zeros(o, np.uint8)
        _check(load_library().ns_circuit_eval(self._h, _ptr(a), _ptr(out)))
        return [int(v) for v in out]

    def native_body(self, kind: int = 0, lane_shift: int = 0):
        """Assembly of the circuit body on the native path and its register class
        (kind 0 SAT, 1/2 QSAT, 3 table); assembled and linked on the host as a check."""
        lib = load_library()
        need = C.c_size_t()
        cls = C.c_uint32()
        size = C.c_uint64()
        _check(lib.ns_native_body(self._h, kind, lane_shift, None, 0, C.byref(need), C.byref(cls), C.byref(size)))
        buf = C.create_string_buffer(need.value + 1)
        _check(lib.ns_native_body(self._h, kind, lane_shift, buf, need.value + 1, C.byref(need), C.byref(cls),
                                  C.byref(size)))
        return buf.value.decode(), cls.value, size.value

    def native_code(self, kind: int = 0, lane_shift: int = 0):
        """Code object the engine loads for this circuit's native kernel:
        (bytes, patched into the template's body slot, slot byte offset)."""
        lib = load_library()
        need, patched, off = C.c_size_t(), C.c_uint32(), C.c_uint64()
        _check(lib.ns_native_code(self._h, kind, lane_shift, None, 0, C.byref(need), C.byref(patched), C.byref(off)))
        buf = (C.c_char * need.value)()
        _check(lib.ns_native_code(self._h, kind, lane_shift, buf, need.value, C.byref(need), C.byref(patched),
                                  C.byref(off)))
        return bytes(buf), bool(patched.value), off.value

    def cone(self):
        """(gates in the output cone, peak live values of the kernel's order)."""
        g, p = C.c_uint32(), C.c_uint32()
        _check(load_library().ns_circuit_cone(self._h, C.byref(g), C.byref(p)))
        return g.value, p.value


def native_encode(body: str) -> np.ndarray:
    """gfx950 machine-code words of native body text (the engine's own encoder)."""
    lib = load_library()
    raw = body.encode()
    n = C.c_size_t()
    _check(lib.ns_native_encode(raw, len(raw), None, 0, C.byref(n)))
    w = np.zeros(max(1, n.value), np.uint32)
    _check(lib.ns_native_encode(raw, len(raw), _ptr(w), w.size, C.byref(n)))
    return w[: n.value]


class Context:
    """One HIP device context (stream + device buffers + compiled-kernel cache)."""

    def __init__(self, device: int = 0):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.ns_ctx_create(device, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value and _lib is not None:
            _lib.ns_ctx_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sat_solve(self, c: Circuit, elem_bits: int = 0, group_len: int = 0, words_per_lane: int = 0,
                  stop_at_first: bool = False, group_exits: bool = False) -> dict:
        """group_exits: NS_SAT_GROUP_EXITS (a wave leaves a word once the output's groups so far
        are zero in all lanes; same answers)."""
        o = _SatOpts(elem_bits, group_len, words_per_lane, _sat_flags(stop_at_first, group_exits))
        r = _SatResult()
        _check(load_library().ns_sat_solve(self._h, c._h, C.byref(o), C.byref(r)))
        return {k: getattr(r, k) for k, _ in _SatResult._fields_}

    def sat_solve_progress(self, c: Circuit, callback, elem_bits: int = 0, group_len: int = 0,
                           words_per_lane: int = 0, stop_at_first: bool = False, group_exits: bool = False) -> dict:
        """ns_sat_solve_progress: callback(step, steps, range_bits) -> bool (True stops the
        search), called once with step = SAT_STEP_START before the first launch, then after
        each launch range in order. An exception raised by the callback stops the search and is
        raised again afterwards."""
        o = _SatOpts(elem_bits, group_len, words_per_lane, _sat_flags(stop_at_first, group_exits))
        r = _SatResult()
        err = []

        def tramp(user, step, steps, bits):
            try:
                return 1 if callback(step, steps, bits) else 0
            except BaseException as e:  # noqa: BLE001 - re-raised after the search
                err.append(e)
                return 1

        fn = _SatProgressFn(tramp)
        st = load_library().ns_sat_solve_progress(self._h, c._h, C.byref(o), fn, None, C.byref(r))
        if err:
            raise err[0]
        _check(st)
        return {k: getattr(r, k) for k, _ in _SatResult._fields_}

    def sat_solve_noalloc(self, c: Circuit, elem_bits: int = 0, group_len: int = 0, words_per_lane: int = 0,
                          stop_at_first: bool = False, group_exits: bool = False) -> dict:
        """ns_sat_solve_noalloc: the search without any device allocation; the circuit's kernel
        must have been built by an earlier sat_solve with the same options."""
        o = _SatOpts(elem_bits, group_len, words_per_lane, _sat_flags(stop_at_first, group_exits))
        r = _SatResult()
        _check(load_library().ns_sat_solve_noalloc(self._h, c._h, C.byref(o), C.byref(r)))
        return {k: getattr(r, k) for k, _ in _SatResult._fields_}

    def qsat_solve(self, c: Circuit, quants_all: Sequence[int], elem_bits: int = 0, group_len: int = 0,
                   words_per_lane: int = 0, bitmap: bool = False, progress=None) -> dict:
        """progress: callback(step, steps, range_bits) -> bool as for sat_solve_progress
        (ns_qsat_solve_progress); True stops the search, which then raises NaivesatError
        (status NS_ERR_CANCELLED)."""
        q = np.ascontiguousarray(quants_all, dtype=np.uint8)
        if q.size != c.n_inputs:
            raise ValueError("one quantifier per input is required")
        o = _QsatOpts(elem_bits, group_len, words_per_lane, 0)
        r = _QsatResult()
        fb = 0
        while fb < q.size and q[fb] == q[0]:
            fb += 1
        bm = np.zeros(max(1, (1 << fb) // 32), np.uint32) if bitmap else None
        if progress is None:
            _check(load_library().ns_qsat_solve(self._h, c._h, _ptr(q), C.byref(o), C.byref(r), _ptr(bm),
                                                0 if bm is None else bm.size))
        else:
            err = []

            def tramp(user, step, steps, bits):
                try:
                    return 1 if progress(step, steps, bits) else 0
                except BaseException as e:  # noqa: BLE001 - re-raised after the search
                    err.append(e)
                    return 1

            fn = _SatProgressFn(tramp)
            st = load_library().ns_qsat_solve_progress(self._h, c._h, _ptr(q), C.byref(o), fn, None, C.byref(r),
                                                       _ptr(bm), 0 if bm is None else bm.size)
            if err:
                raise err[0]
            _check(st)
        d = {k: getattr(r, k) for k, _ in _QsatResult._fields_}
        if bm is not None:
            bits = np.unpackbits(bm.view(np.uint8), bitorder="little")[: 1 << fb]
            d["bitmap"] = bits
        return d

    def satostop_solve(self, c: Circuit, unknowns: int, first_pass: bool = False, max_passes: int = 0,
                       group_len: int = 0, words_per_lane: int = 0, arrays: bool = True,
                       wide: bool = True, passes: bool = False, method: str = "") -> dict:
        """Per-start steps/ends arrays are uint64 (ns_satostop_solve_wide, up to 62 state
        bits) or, with wide=False, uint32 (ns_satostop_solve, up to 31). passes=True runs the
        full search as passes over the whole table instead of waypoint walks; method
        "table" / "bits" picks the waypoint walks over the table / without one."""
        flags = (NS_STOP_FIRST_PASS if first_pass else 0) | (NS_STOP_PASSES if passes else 0)
        flags |= {"": 0, "table": NS_STOP_TABLE_WALKS, "bits": NS_STOP_BIT_WALKS}[method]
        o = _StopOpts(flags, max_passes, group_len, words_per_lane)
        r = _StopResult()
        dt = np.uint64 if wide else np.uint32
        steps = np.zeros(1 << unknowns, dt) if arrays else None
        ends = np.zeros(1 << unknowns, dt) if arrays else None
        fn = load_library().ns_satostop_solve_wide if wide else load_library().ns_satostop_solve
        _check(fn(self._h, c._h, unknowns, C.byref(o), C.byref(r), _ptr(steps), _ptr(ends)))
        d = {k: getattr(r, k) for k, _ in _StopResult._fields_}
        if arrays:
            d["steps_arr"] = steps
            d["ends_arr"] = ends
        return d

    def satostop_hashmap_solve(self, c: Circuit, unknowns: int, hashmap_len_bits: int = 0, elem_bits: int = 0,
                               unknown_fill_bits: int = -1, max_predecessors: int = 0,
                               iters_per_clear_predecessors: int = 0, max_sweeps: int = 0, progress=None) -> dict:
        """progress: callback(step, steps, resolved_unknowns) -> bool after each block
        (ns_satostop_hashmap_solve_progress); True stops the search (NaivesatError,
        NS_ERR_CANCELLED)."""
        o = _HmOpts(hashmap_len_bits, elem_bits, unknown_fill_bits, max_predecessors,
                    iters_per_clear_predecessors, max_sweeps)
        r = _HmResult()
        if progress is None:
            _check(load_library().ns_satostop_hashmap_solve(self._h, c._h, unknowns, C.byref(o), C.byref(r)))
        else:
            err = []

            def tramp(user, step, steps, resolved):
                try:
                    return 1 if progress(step, steps, resolved) else 0
                except BaseException as e:  # noqa: BLE001 - re-raised after the search
                    err.append(e)
                    return 1

            fn = _HmProgressFn(tramp)
            st = load_library().ns_satostop_hashmap_solve_progress(self._h, c._h, unknowns, C.byref(o), fn, None,
                                                                   C.byref(r))
            if err:
                raise err[0]
            _check(st)
        return {k: getattr(r, k) for k, _ in _HmResult._fields_}

    def hashmap_step(self, which: int, state_len: int, arg_bit_place: int, arg: int, unknown_bits: int,
                     unknown_fill_bits: int, max_predecessors: int, test: bool, outputs, hm, hm_out, fills, sol):
        """One device step of the hash-map search on numpy arrays: which = 0 join_to, 1
        join_itself (hm -> hm_out), 2 add, 3 join_to + join_itself in one pass (hm -> hm_out).
        hm: 32-byte entries, sol: 40-byte {resolved_unknowns, start, end, steps, defined};
        updated in place."""
        out_ptr = _ptr(outputs) if outputs is not None else None
        _check(load_library().ns_hashmap_step(
            self._h, which, state_len, arg_bit_place, arg, unknown_bits, unknown_fill_bits, max_predecessors,
            1 if test else 0, out_ptr, 0 if outputs is None else outputs.size, _ptr(hm),
            _ptr(hm_out) if hm_out is not None else None, hm.size, _ptr(fills), fills.size, _ptr(sol)))

    def table_build(self, c: Circuit) -> np.ndarray:
        n = c.n_inputs
        t = np.zeros(1 << n, np.uint32)
        ms = C.c_double()
        _check(load_library().ns_table_build(self._h, c._h, _ptr(t), t.size, C.byref(ms)))
        return t

    def serve(self, requests, workers: int = 8, realtime: bool = True, hbm_budget_bytes: int = 0,
              batch_max: int = 0):
        """Solves a stream of requests on concurrent lanes. Each request is a dict with
        kind (REQ_SAT / REQ_QSAT / REQ_SATOSTOP), circuit, quants (QSAT), unknowns
        (SATOSTOP) and arrival_s. batch_max > 1 batches waiting SAT / QSAT requests into
        persistent kernels. Returns (list of result dicts, total seconds)."""
        n = len(requests)
        arr, keep = _requests(requests)
        res = (_RequestResult * max(1, n))()
        total = C.c_double()
        o = _ServeOpts(workers, 1 if realtime else 0, hbm_budget_bytes, batch_max, 0)
        _check(load_library().ns_serve(self._h, arr, n, C.byref(o), res, C.byref(total)))
        out = [_result_dict(res[i]) for i in range(n)]
        return out, total.value

    def open_server(self, workers: int = 8, batch_max: int = 0, hbm_budget_bytes: int = 0) -> "Server":
        """An online server on this context (ns_serve_open)."""
        return Server(self, workers, batch_max, hbm_budget_bytes)

    def batch_solve(self, requests):
        """Solves SAT / QSAT requests together in the persistent batch kernels (SATOSTOP and
        circuits they cannot take one at a time). Returns (result dicts, info dict)."""
        n = len(requests)
        arr, keep = _requests(requests)
        res = (_RequestResult * max(1, n))()
        info = np.zeros(10, np.float64)
        _check(load_library().ns_batch_solve(self._h, arr, n, res, _ptr(info)))
        names = ["requests", "fallbacks", "vgpr_class", "code_bytes", "gen_ms", "load_ms", "kernel_ms", "wall_ms",
                 "graph_runs", "launches"]
        return [_result_dict(res[i]) for i in range(n)], dict(zip(names, info.tolist()))

    def greedy_partitioned_solve(self, c: Circuit, unknowns: int, partitions: int, file_prefix: str = "") -> dict:
        """greedysatostop's partitioned search with the image on disk (see the C header)."""
        r = _PartResult()
        _check(load_library().ns_greedy_partitioned_solve(self._h, c._h, unknowns, partitions,
                                                          file_prefix.encode(), C.byref(r)))
        return {k: getattr(r, k) for k, _ in _PartResult._fields_}

    def trim(self) -> None:
        """Releases the SATOSTOP tables and images the context keeps between solves."""
        _check(load_library().ns_ctx_trim(self._h))

    def pack_entries(self, width: int, values) -> np.ndarray:
        """The device's MemImage packing of `values` (uint64, a multiple of 64 entries)."""
        v = np.ascontiguousarray(values, np.uint64)
        words = np.zeros(v.size * width // 64, np.uint64)
        _check(load_library().ns_pack_entries(self._h, width, _ptr(v), v.size, _ptr(words), words.size))
        return words

    def probe_gather(self, table_bits: int = 30, reads_per_lane: int = 256, lanes: int = 1 << 20,
                     nontemporal: bool = False) -> dict:
        """Scattered 4-byte read rate from a 2^table_bits-entry device table."""
        ms, reads = C.c_double(), C.c_uint64()
        _check(load_library().ns_probe_gather(self._h, table_bits, reads_per_lane, lanes, 1 if nontemporal else 0,
                                              C.byref(ms), C.byref(reads)))
        return {"ms": ms.value, "reads": reads.value, "reads_per_s": reads.value / (ms.value / 1000.0)}

    def pass_times(self) -> list:
        buf = np.zeros(64, np.float64)
        w = C.c_uint32()
        _check(load_library().ns_satostop_pass_times(self._h, _ptr(buf), 64, C.byref(w)))
        return [float(x) for x in buf[: w.value]]


def _sat_flags(stop_at_first, group_exits):
    return (NS_SAT_STOP_AT_FIRST if stop_at_first else 0) | (NS_SAT_GROUP_EXITS if group_exits else 0)


def _requests(requests):
    arr = (_Request * max(1, len(requests)))()
    keep = []
    for i, rq in enumerate(requests):
        arr[i] = _request(rq, keep)
    return arr, keep


def _request(rq, keep):
    q = None
    if rq["kind"] == REQ_QSAT:
        q = np.ascontiguousarray(rq["quants"], dtype=np.uint8)
        keep.append(q)
    return _Request(rq["kind"], rq.get("unknowns", 0), rq["circuit"]._h, _ptr(q), rq.get("arrival_s", 0.0))


def _result_dict(r) -> dict:
    return {k: getattr(r, k) for k, _ in _RequestResult._fields_}


class Server:
    """Online serving (ns_serve_open / submit / poll / wait / close): requests are taken as
    they are submitted; results are handed out once per ticket."""

    def __init__(self, ctx: "Context", workers: int = 8, batch_max: int = 0, hbm_budget_bytes: int = 0):
        h = C.c_void_p()
        o = _ServeOpts(workers, 0, hbm_budget_bytes, batch_max, 0)
        _check(load_library().ns_serve_open(ctx._h, C.byref(o), C.byref(h)))
        self._h = h.value
        self._ctx = ctx  # the context outlives the server

    def submit(self, rq: dict) -> int:
        keep = []
        t = C.c_uint64()
        _check(load_library().ns_serve_submit(self._h, C.byref(_request(rq, keep)), C.byref(t)))
        return t.value

    def poll(self, ticket: int):
        """The result dict once the ticket's request has finished, else None."""
        r, done = _RequestResult(), C.c_int()
        _check(load_library().ns_serve_poll(self._h, ticket, C.byref(r), C.byref(done)))
        return _result_dict(r) if done.value else None

    def wait(self, ticket: int, timeout_s: float = -1.0):
        """The result dict; None when timeout_s (>= 0) passes first."""
        r, done = _RequestResult(), C.c_int()
        _check(load_library().ns_serve_wait(self._h, ticket, timeout_s, C.byref(r), C.byref(done)))
        return _result_dict(r) if done.value else None

    def close(self) -> None:
        if self._h:
            h, self._h = self._h, None
            _check(load_library().ns_serve_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def quants_from_string(s: str) -> np.ndarray:
    """'eeaa' -> [0,0,1,1] (1 = all); other characters are ignored."""
    return np.array([1 if ch == "a" else 0 for ch in s.lower() if ch in "ae"], dtype=np.uint8)


def qsat_result_string(r: dict) -> str:
    """The reference's printed FinalResult (src/qsat.rs:156-170): the witness is printed
    with bit b = input b, most significant character first."""
    fb = r["first_bits"]
    sol = None
    if r["has_witness"]:
        u = r["witness"]
        sol = int(format(u, f"0{fb}b")[::-1], 2) if fb else 0
    if r["first_all"]:
        return f"(Unsatisfiable: {sol:0{fb}b})" if sol is not None else "(Satisfiable)"
    return f"(Satisfiable: {sol:0{fb}b})" if sol is not None else "(Unsatisfiable)"


class FileImage:
    """greedysatostop's on-disk image (host only): packed words in, partitions out."""

    def __init__(self, state_len: int, partitions: int, prefix: str = ""):
        h = C.c_void_p()
        _check(load_library().ns_file_image_create(state_len, partitions, prefix.encode(), C.byref(h)))
        self._h = h.value
        self.state_len, self.partitions = state_len, partitions
        self.partition_words = ((1 << state_len) // partitions) * (state_len + 1) // 64

    def save_chunk(self, words: np.ndarray) -> None:
        w = np.ascontiguousarray(words, np.uint64)
        _check(load_library().ns_file_image_save_chunk(self._h, _ptr(w), w.size))

    def load_partition(self, part: int) -> np.ndarray:
        w = np.zeros(self.partition_words, np.uint64)
        _check(load_library().ns_file_image_load_partition(self._h, part, _ptr(w), w.size))
        return w

    def save_partition(self, part: int, words: np.ndarray) -> None:
        w = np.ascontiguousarray(words, np.uint64)
        _check(load_library().ns_file_image_save_partition(self._h, part, _ptr(w), w.size))

    def close(self) -> None:
        if self._h:
            load_library().ns_file_image_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
