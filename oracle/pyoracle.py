"""ORACLE -- test infrastructure only.

ctypes bindings for the CPU restatement (oracle/liboracle.so, built from
oracle/oracle.cpp) and for the reference's own CPU engines
(oracle/_ref/libhgmref.so, built from the reference's sources by
oracle/Makefile when the reference tree is present).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker; the HIP library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhgmref.so")
# the same reference engine with its GPU stub left out and this repository's
# libhgm.so linked in its place (oracle/Makefile, target ref-hgm)
REF_HGM_SO = os.path.join(HERE, "_ref", "libhgmref_hgm.so")

CAT_DOUBLE = 0
CAT_REFERENCE = 1
T_CPU = 0
T_PAR_CPU = 1
T_GPU = 2
T_HYBRID = 3

_lib = None
_ref = {}


def build(ref_tree: str | None = None) -> None:
    """Compile liboracle.so (and _ref when the reference tree is readable)."""
    args = ["make", "-C", HERE, "-j8", "-s"]
    if ref_tree:
        args.append(f"REF={ref_tree}")
    subprocess.run(args, check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        lib = C.CDLL(ORACLE_SO)
        lib.orc_new.restype = C.c_void_p
        lib.orc_new.argtypes = [C.c_int, C.c_int, C.c_int]
        lib.orc_free.argtypes = [C.c_void_p]
        lib.orc_last_error.restype = C.c_char_p
        lib.orc_add_gate.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]
        lib.orc_get_gate.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]
        lib.orc_set_function.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int]
        lib.orc_set_function_steps.argtypes = [C.c_void_p, C.POINTER(C.c_char_p), C.c_int, C.c_int, C.c_int]
        lib.orc_qubits.argtypes = [C.c_void_p]
        lib.orc_num_gates.argtypes = [C.c_void_p]
        lib.orc_num_gates.restype = C.c_int64
        lib.orc_get_gates.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_set_gates.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(C.c_char_p)]
        lib.orc_set_basis.argtypes = [C.c_void_p, C.c_int64]
        lib.orc_set_state.argtypes = [C.c_void_p, C.c_void_p]
        lib.orc_get_state.argtypes = [C.c_void_p, C.c_void_p]
        lib.orc_execute.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.orc_seed.argtypes = [C.c_void_p, C.c_ulong]
        lib.orc_measure.argtypes = [C.c_void_p, C.c_int]
        _lib = lib
    return _lib


def _steps_arg(steps):
    arr = (C.c_char_p * len(steps))(*[s.encode() for s in steps])
    return arr


class Oracle:
    """CPU restatement of the DGM semantics (precision 64 or 32)."""

    def __init__(self, qubits: int, precision: int = 64, catalog: int = CAT_DOUBLE):
        self.lib = _load()
        self.h = self.lib.orc_new(qubits, precision, catalog)
        if not self.h:
            raise ValueError("bad oracle arguments")
        self.qubits = qubits
        self.precision = precision
        self.set_basis(0)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_free(self.h)
            self.h = None

    @property
    def dtype(self):
        return np.complex128 if self.precision == 64 else np.complex64

    def add_gate(self, name: str, m) -> bool:
        m = np.asarray(m, dtype=np.complex128).reshape(4)
        buf = (C.c_double * 8)(*np.column_stack([m.real, m.imag]).ravel())
        return bool(self.lib.orc_add_gate(self.h, name.encode(), buf))

    def get_gate(self, name: str):
        buf = (C.c_double * 8)()
        if not self.lib.orc_get_gate(self.h, name.encode(), buf):
            return None
        a = np.frombuffer(buf, dtype=np.float64)
        return a[0::2] + 1j * a[1::2]

    def set_function(self, function, iterations: int = 1, reset: bool = True) -> None:
        if isinstance(function, str):
            self.lib.orc_set_function(self.h, function.encode(), iterations, int(reset))
        else:
            steps = list(function)
            self.lib.orc_set_function_steps(self.h, _steps_arg(steps), len(steps), iterations, int(reset))

    def gates(self):
        n = self.lib.orc_num_gates(self.h)
        t = np.zeros(n, np.int32)
        cm = np.zeros(n, np.int64)
        cv = np.zeros(n, np.int64)
        m = np.zeros((n, 8), np.float64)
        self.lib.orc_get_gates(self.h, t.ctypes.data, cm.ctypes.data, cv.ctypes.data, m.ctypes.data)
        return t, cm, cv, m

    def set_basis(self, index: int) -> None:
        self.lib.orc_set_basis(self.h, index)

    def set_state(self, amps) -> None:
        a = np.ascontiguousarray(amps, dtype=self.dtype)
        assert a.size == 1 << self.qubits
        self.lib.orc_set_state(self.h, a.ctypes.data)

    def state(self):
        out = np.empty(1 << self.qubits, dtype=self.dtype)
        self.lib.orc_get_state(self.h, out.ctypes.data)
        return out

    def execute(self, engine: int = T_CPU, iterations: int = 1, threads: int = 1,
                coalesced_bits: int = 11, region_bits: int = 14) -> None:
        rc = self.lib.orc_execute(self.h, engine, threads, coalesced_bits, region_bits, iterations)
        if rc:
            raise RuntimeError(self.lib.orc_last_error().decode())

    def seed(self, s: int) -> None:
        self.lib.orc_seed(self.h, s)

    def measure(self, qubit: int) -> int:
        return int(self.lib.orc_measure(self.h, qubit))


def run(qubits, function, *, gates=None, basis=0, precision=64, iterations=1,
        engine=T_CPU, threads=1, catalog=CAT_DOUBLE, reset_steps=None):
    """One-shot helper: register gates, parse, execute, return the state."""
    o = Oracle(qubits, precision, catalog)
    for name, m in (gates or {}).items():
        o.add_gate(name, m)
    o.set_function(function)
    if reset_steps:
        for steps, it in reset_steps:
            o.set_function(steps, it, reset=False)
    o.set_basis(basis)
    o.execute(engine, iterations=iterations, threads=threads)
    return o.state()


# --------------------------------------------------------------------------
# the reference's own CPU engines (oracle/_ref/libhgmref.so)
# --------------------------------------------------------------------------
def ref_available(so: str = REF_SO) -> bool:
    return os.path.exists(so)


def _load_ref(so: str = REF_SO):
    if so not in _ref:
        lib = C.CDLL(so)
        lib.ref_new.restype = C.c_void_p
        lib.ref_new.argtypes = [C.c_int]
        lib.ref_free.argtypes = [C.c_void_p]
        lib.ref_add_gate.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_float)]
        lib.ref_get_gate.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_float)]
        lib.ref_set_function_steps.argtypes = [C.c_void_p, C.POINTER(C.c_char_p), C.c_int, C.c_int, C.c_int]
        lib.ref_qubits.argtypes = [C.c_void_p]
        lib.ref_num_pts.argtypes = [C.c_void_p]
        lib.ref_num_pts.restype = C.c_long
        lib.ref_get_pts.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ref_set_basis.argtypes = [C.c_void_p, C.c_long]
        lib.ref_set_state.argtypes = [C.c_void_p, C.c_void_p]
        lib.ref_get_state.argtypes = [C.c_void_p, C.c_void_p]
        lib.ref_execute.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.ref_seed.argtypes = [C.c_ulong]
        lib.ref_measure.argtypes = [C.c_void_p, C.c_int]
        lib.ref_gen.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_long), C.c_int, C.c_char_p,
                                C.c_char_p, C.c_size_t]
        lib.ref_grover_iteration.argtypes = [C.c_long, C.c_long, C.c_char_p, C.c_size_t]
        lib.ref_frac_approx.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]
        lib.ref_mul_inv.argtypes = [C.c_long, C.c_long]
        lib.ref_mul_inv.restype = C.c_long
        lib.ref_revert_bits.argtypes = [C.c_int, C.c_int]
        lib.ref_shor.argtypes = [C.c_long, C.POINTER(C.c_int)]
        lib.ref_print_pts.argtypes = [C.c_void_p]
        lib.ref_colapse.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.ref_measure_many.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_long]
        lib.ref_measure_many.restype = C.c_long
        lib.ref_set_gpu_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.ref_general_hadamard.argtypes = [C.c_long, C.c_int, C.c_int]
        lib.ref_grover_run.argtypes = [C.c_long, C.c_long, C.c_int, C.c_int, C.c_ulong]
        lib.ref_grover_run.restype = C.c_long
        lib.ref_shor_run.argtypes = [C.c_long, C.c_int, C.c_int, C.c_ulong, C.POINTER(C.c_int)]
        _ref[so] = lib
    return _ref[so]


class RefDGM:
    """The reference's DGM (FP32), driven through oracle/ref_driver.cpp. so picks
    the build: REF_SO (GPU stub) or REF_HGM_SO (libhgm.so behind t_GPU and t_HYBRID)."""

    def __init__(self, qubits: int, so: str = REF_SO):
        self.lib = _load_ref(so)
        self.h = self.lib.ref_new(qubits)
        self.qubits = qubits

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_free(self.h)
            self.h = None

    def add_gate(self, name, m) -> bool:
        m = np.asarray(m, dtype=np.complex128).reshape(4)
        buf = (C.c_float * 8)(*np.column_stack([m.real, m.imag]).ravel())
        return bool(self.lib.ref_add_gate(self.h, name.encode(), buf))

    def get_gate(self, name):
        buf = (C.c_float * 8)()
        if not self.lib.ref_get_gate(self.h, name.encode(), buf):
            return None
        a = np.frombuffer(buf, dtype=np.float32).astype(np.float64)
        return a[0::2] + 1j * a[1::2]

    def set_function(self, steps, iterations=1, reset=True):
        if isinstance(steps, str):
            steps = [s for s in steps.split(";") if s]
        steps = list(steps)
        self.lib.ref_set_function_steps(self.h, _steps_arg(steps), len(steps), iterations, int(reset))

    def pts(self):
        n = self.lib.ref_num_pts(self.h)
        t = np.zeros(n, np.int32)
        cm = np.zeros(n, np.int64)
        cv = np.zeros(n, np.int64)
        m = np.zeros((n, 8), np.float32)
        self.lib.ref_get_pts(self.h, t.ctypes.data, cm.ctypes.data, cv.ctypes.data, m.ctypes.data)
        return t, cm, cv, m

    def set_gpu_tuning(self, gpu_region_bits, gpu_coalesced_bits, block_size, repeat_count):
        self.lib.ref_set_gpu_tuning(self.h, gpu_region_bits, gpu_coalesced_bits, block_size, repeat_count)

    def set_basis(self, index):
        self.lib.ref_set_basis(self.h, index)

    def set_state(self, amps):
        a = np.ascontiguousarray(amps, dtype=np.complex64)
        self.lib.ref_set_state(self.h, a.ctypes.data)

    def state(self):
        out = np.empty(1 << self.qubits, dtype=np.complex64)
        self.lib.ref_get_state(self.h, out.ctypes.data)
        return out

    def execute(self, exec_type=T_PAR_CPU, threads=1, coalesced_bits=11, region_bits=14, iterations=1):
        self.lib.ref_execute(self.h, exec_type, threads, coalesced_bits, region_bits, iterations)

    def measure(self, qubit):
        return int(self.lib.ref_measure(self.h, qubit))

    def gen(self, which, args, op=""):
        arr = (C.c_long * max(1, len(args)))(*args)
        n = self.lib.ref_gen(self.h, which.encode(), arr, len(args), op.encode(), None, 0)
        if n < 0:
            raise KeyError(which)
        buf = C.create_string_buffer(n + 1)
        self.lib.ref_gen(self.h, which.encode(), arr, len(args), op.encode(), buf, n + 1)
        return buf.value.decode()


def ref_seed(seed: int) -> None:
    _load_ref().ref_seed(seed)


def ref_grover_iteration(qubits: int, search_value: int) -> list[str]:
    lib = _load_ref()
    n = lib.ref_grover_iteration(qubits, search_value, None, 0)
    buf = C.create_string_buffer(n + 1)
    lib.ref_grover_iteration(qubits, search_value, buf, n + 1)
    return buf.value.decode().split(";")


def ref_frac_approx(num: int, den: int, width: int) -> tuple[int, int]:
    """The reference's quantum_frac_approx (lib_shor_number_theory.cpp:43-71)."""
    n, d = C.c_int(num), C.c_int(den)
    _load_ref().ref_frac_approx(C.byref(n), C.byref(d), width)
    return n.value, d.value


def ref_mul_inv(v: int, m: int) -> int:
    return int(_load_ref().ref_mul_inv(v, m))


def ref_revert_bits(v: int, bits: int) -> int:
    return int(_load_ref().ref_revert_bits(v, bits))


def ref_shor(number: int, seed: int) -> tuple[int, int]:
    """One attempt of the reference's Shor() on its t_CPU engine after seed_rng(seed)."""
    ref_seed(seed)
    out = (C.c_int * 2)()
    _load_ref().ref_shor(number, out)
    return out[0], out[1]
