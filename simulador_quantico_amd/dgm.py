"""Python mirror of the reference's DGM executor (include/core/dgm.h:65-163).

Every method forwards to the C++ executor inside libhgm.so through the C ABI
(include/hgm_gpu.h); only exec_type == T_GPU executes. Method names, argument
meanings and defaults follow the reference so that a reference test reads
the same here::

    dgm = DGM()
    dgm.qubits = 3
    dgm.exec_type = T_GPU
    dgm.allocateMemory()
    dgm.setMemoryValue(5)
    dgm.executeFunction(Toffoli(3, 0, 1, 2, 3))
    dgm.state[7]
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._native import HGMError, Op, Stats, check, lib

T_CPU, T_PAR_CPU, T_GPU, T_HYBRID = 0, 1, 2, 3      # include/core/dgm.h:42-47
DENSE, DIAG_PRI, DIAG_SEC = 0, 1, 2                  # include/core/common.h:48-52

_INT_FIELDS = ("qubits", "exec_type", "gpu_count", "block_size", "repeat_count", "gpu_region_bits",
               "gpu_coalesced_bits", "ops_block", "planner", "kernel", "thread_count", "cpu_region_bits",
               "cpu_coalesced_bits", "resident")


def _m8(m) -> "C.Array":
    m = np.asarray(m, dtype=np.complex128).reshape(4)
    return (C.c_double * 8)(*np.column_stack([m.real, m.imag]).ravel())


class _Gates:
    """The DGM's gate catalog (include/core/gates.h:40-54)."""

    def __init__(self, owner: "DGM"):
        self._o = owner

    def addGate(self, name: str, *m) -> bool:
        mat = m[0] if len(m) == 1 else m
        added = C.c_int(0)
        check(lib.hgm_dgm_add_gate(self._o._h, name.encode(), _m8(mat), C.byref(added)))
        return bool(added.value)

    def getMatrix(self, name: str):
        buf = (C.c_double * 8)()
        found = C.c_int(0)
        check(lib.hgm_dgm_get_gate(self._o._h, name.encode(), buf, C.byref(found)))
        if not found.value:
            return None
        a = np.frombuffer(buf, dtype=np.float64).copy()
        return a[0::2] + 1j * a[1::2]


class DGM:
    """Executor handle. dtype: np.complex128 (default) or np.complex64."""

    def __init__(self, dtype=np.complex128):
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.complex128, np.complex64):
            raise ValueError("dtype must be complex128 or complex64")
        h = C.c_void_p()
        check(lib.hgm_dgm_create(N.HGM_C128 if self.dtype == np.complex128 else N.HGM_C64, C.byref(h)))
        self._h = h
        self.gates = _Gates(self)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.hgm_dgm_destroy(h)
            self._h = None

    def __getattr__(self, name):
        if name in _INT_FIELDS:
            v = C.c_int64()
            check(lib.hgm_dgm_get(self.__dict__["_h"], name.encode(), C.byref(v)))
            return v.value
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _INT_FIELDS:
            check(lib.hgm_dgm_set(self._h, name.encode(), int(value)))
        else:
            object.__setattr__(self, name, value)

    # ---- configuration (include/core/dgm.h:113-116, src/core/dgm_core.cpp:30-32,302-318) ----
    def setExecType(self, exec_type: int) -> None:
        self.exec_type = exec_type

    def setCpuStructure(self, cpu_region_bits: int, cpu_coalesced_bits: int) -> None:
        self.cpu_region_bits = cpu_region_bits
        self.cpu_coalesced_bits = cpu_coalesced_bits

    def setGpuStructure(self, gpu_coalesced_bits: int, gpu_region_bits: int, repeat_count: int = 1) -> None:
        """block_size follows from the tile invariant 2*block_size*repeat_count == 2^gpu_region_bits
        (the reference writes 1 << r/2/R, which C parses as 1 << (r/2/R), src/core/dgm_core.cpp:317)."""
        self.gpu_coalesced_bits = gpu_coalesced_bits
        self.gpu_region_bits = gpu_region_bits
        self.repeat_count = repeat_count
        self.block_size = (1 << gpu_region_bits) // 2 // repeat_count

    def validateTuning(self) -> None:
        check(lib.hgm_dgm_validate_tuning(self._h))

    def erase(self) -> None:
        """Drop the compiled PT list (src/core/dgm_core.cpp:41-52)."""
        check(lib.hgm_dgm_erase(self._h))

    # ---- state -------------------------------------------------------
    def allocateMemory(self) -> None:
        check(lib.hgm_dgm_allocate_memory(self._h))

    def freeMemory(self) -> None:
        check(lib.hgm_dgm_free_memory(self._h))

    def setMemoryValue(self, pos: int) -> None:
        check(lib.hgm_dgm_set_memory_value(self._h, int(pos)))

    @property
    def state(self) -> np.ndarray:
        """Zero-copy view of the host amplitudes (valid until freeMemory). Brought up
        to date first; in resident mode the device copy is then treated as stale,
        so writes through the view reach the next execute()."""
        p = lib.hgm_dgm_state(self._h)
        if not p:
            raise HGMError(6, "no state allocated")
        n = 1 << self.qubits
        buf = (C.c_byte * (n * self.dtype.itemsize)).from_address(p)
        return np.frombuffer(buf, dtype=self.dtype)

    def syncHost(self) -> None:
        check(lib.hgm_dgm_sync_host(self._h))

    # ---- circuits ----------------------------------------------------
    def setFunction(self, function, iterations: int = 1, reset: bool = True) -> None:
        if isinstance(function, str):
            check(lib.hgm_dgm_set_function(self._h, function.encode(), iterations, int(reset)))
        else:
            steps = [s.encode() for s in function]
            arr = (C.c_char_p * max(1, len(steps)))(*steps)
            check(lib.hgm_dgm_set_function_steps(self._h, arr, len(steps), iterations, int(reset)))

    def ops(self) -> "C.Array":
        n = C.c_int64()
        check(lib.hgm_dgm_num_pts(self._h, C.byref(n)))
        arr = (Op * max(1, n.value))()
        check(lib.hgm_dgm_get_ops(self._h, arr, n.value))
        return arr if n.value else (Op * 0)()

    def pts(self):
        """The compiled PT list as (target_bits, control_masks, control_values, matrices[n,4])."""
        return ops_to_arrays(self.ops())

    def execute(self, iterations: int = 1):
        """Run the PT list on the GPU (src/core/dgm_core.cpp:243-267) and return the
        host state; in resident mode the result stays on the device and None is
        returned (read `state`, which copies it back, or measure on the device)."""
        check(lib.hgm_dgm_execute(self._h, iterations))
        return None if self.resident else self.state

    def executeFunction(self, function, iterations: int = 1) -> None:
        if isinstance(function, str):
            if function == "":
                return
            check(lib.hgm_dgm_execute_function(self._h, function.encode(), iterations))
        else:
            self.setFunction(function)
            self.execute(iterations)

    # ---- readout -----------------------------------------------------
    def measure(self, qubit):
        if isinstance(qubit, (list, tuple)):
            qs = (C.c_int * max(1, len(qubit)))(*qubit)
            n = C.c_int64(0)
            check(lib.hgm_dgm_measure_many(self._h, qs, len(qubit), None, None, C.byref(n)))
            vals = (C.c_int64 * max(1, n.value))()
            probs = (C.c_double * max(1, n.value))()
            check(lib.hgm_dgm_measure_many(self._h, qs, len(qubit), vals, probs, C.byref(n)))
            return {int(vals[i]): float(probs[i]) for i in range(n.value)}
        out = C.c_int()
        check(lib.hgm_dgm_measure(self._h, int(qubit), C.byref(out)))
        return out.value

    def colapse(self, qubit: int, value: int) -> None:
        """Keep the amplitudes whose qubit equals value, renormalised; zero the rest
        (src/core/dgm_core.cpp:129-146, which also prints the kept probability)."""
        check(lib.hgm_dgm_colapse(self._h, int(qubit), int(value)))

    def printPTs(self) -> None:
        """Every PT's fields to stdout as the reference prints them (src/core/dgm_core.cpp:35-39)."""
        check(lib.hgm_dgm_print_pts(self._h))

    def CountOps(self, iterations: int = 1) -> dict:
        c = (C.c_int64 * 7)()
        check(lib.hgm_dgm_count_ops(self._h, iterations, c))
        keys = ("total_op", "dense", "main_diag", "sec_diag", "c_dense", "c_main_diag", "c_sec_diag")
        return dict(zip(keys, list(c)))

    @property
    def stats(self) -> dict:
        s = Stats()
        check(lib.hgm_dgm_stats(self._h, C.byref(s)))
        return s.as_dict()


def seed_rng(seed: int) -> None:
    lib.hgm_seed_rng(int(seed))


def ops_to_arrays(ops):
    n = len(ops)
    t = np.array([ops[i].target_bit for i in range(n)], dtype=np.int32)
    cm = np.array([ops[i].control_mask for i in range(n)], dtype=np.uint64)
    cv = np.array([ops[i].control_value for i in range(n)], dtype=np.uint64)
    m = np.array([list(ops[i].m) for i in range(n)], dtype=np.float64).reshape(n, 8)
    return t, cm, cv, m[:, 0::2] + 1j * m[:, 1::2]


def make_ops(targets, masks, values, mats) -> "C.Array":
    """Build an hgm_op array from parallel sequences (matrices as 4 complex each)."""
    n = len(targets)
    arr = (Op * max(1, n))()
    for i in range(n):
        arr[i].target_bit = int(targets[i])
        arr[i].control_mask = int(masks[i])
        arr[i].control_value = int(values[i])
        m = np.asarray(mats[i], dtype=np.complex128).reshape(4)
        for k in range(4):
            arr[i].m[2 * k] = m[k].real
            arr[i].m[2 * k + 1] = m[k].imag
    return arr if n else (Op * 0)()


class State:
    """A device-resident state (hgm_state): apply gate lists without host copies."""

    def __init__(self, qubits: int, dtype=np.complex128):
        self.dtype = np.dtype(dtype)
        self.qubits = qubits
        h = C.c_void_p()
        check(lib.hgm_state_create(qubits, N.HGM_C128 if self.dtype == np.complex128 else N.HGM_C64, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.hgm_state_destroy(h)
            self._h = None

    def set_basis(self, index: int) -> None:
        check(lib.hgm_state_set_basis(self._h, int(index)))

    def upload(self, amps) -> None:
        a = np.ascontiguousarray(amps, dtype=self.dtype)
        assert a.size == 1 << self.qubits
        check(lib.hgm_state_upload(self._h, a.ctypes.data))

    def copy_from(self, other: "State") -> None:
        check(lib.hgm_state_copy(self._h, other._h))

    def download(self) -> np.ndarray:
        out = np.empty(1 << self.qubits, dtype=self.dtype)
        check(lib.hgm_state_download(self._h, out.ctypes.data))
        return out

    def _region_host(self, host) -> int:
        if not (isinstance(host, np.ndarray) and host.dtype == self.dtype and host.flags.c_contiguous):
            raise HGMError(N.HGM_E_ARG, f"host state must be a contiguous {self.dtype} array")
        n = int(host.size).bit_length() - 1
        if host.size != 1 << n:
            raise HGMError(N.HGM_E_ARG, "host state size is not a power of two")
        return n

    def gather(self, host: np.ndarray, region_id: int, region_mask: int) -> None:
        """Fill this state with the host amplitudes region_id | deposit(j, region_mask)
        (the region projection of src/core/kernel.cu:321-382)."""
        n = self._region_host(host)
        check(lib.hgm_state_gather(self._h, host.ctypes.data, n, int(region_id), int(region_mask)))

    def scatter(self, host: np.ndarray, region_id: int, region_mask: int) -> None:
        """Write this state back to the host amplitudes region_id | deposit(j, region_mask)
        (src/core/kernel.cu:384-435); the other host amplitudes are untouched."""
        n = self._region_host(host)
        check(lib.hgm_state_scatter(self._h, host.ctypes.data, n, int(region_id), int(region_mask)))

    def apply(self, ops, tuning=None, iterations: int = 1) -> dict:
        st = Stats()
        t = tuning if tuning is not None else N.default_tuning()
        check(lib.hgm_state_apply(self._h, ops, len(ops), C.byref(t), iterations, C.byref(st)))
        return st.as_dict()

    def norm2(self) -> float:
        v = C.c_double()
        check(lib.hgm_state_norm2(self._h, C.byref(v)))
        return v.value

    def prob_one(self, qubit: int) -> float:
        v = C.c_double()
        check(lib.hgm_state_prob_one(self._h, qubit, C.byref(v)))
        return v.value

    def prob_halves(self, qubit: int):
        """(P(qubit == 0), P(qubit == 1)), each summed on the device in FP64."""
        z, o = C.c_double(), C.c_double()
        check(lib.hgm_state_prob_halves(self._h, qubit, C.byref(z), C.byref(o)))
        return z.value, o.value

    def collapse_reset(self, qubit: int, outcome: int, p: float) -> None:
        check(lib.hgm_state_collapse_reset(self._h, qubit, outcome, p))

    def argmax(self):
        i = C.c_uint64()
        p = C.c_double()
        check(lib.hgm_state_argmax(self._h, C.byref(i), C.byref(p)))
        return i.value, p.value

    def compare(self, other: "State"):
        """(fidelity, max |a - b|) against another state, computed on the device."""
        f = C.c_double()
        m = C.c_double()
        check(lib.hgm_state_compare(self._h, other._h, C.byref(f), C.byref(m)))
        return f.value, m.value


class Plan:
    """A pass plan (hgm_plan) built on the host; inspectable without a GPU."""

    def __init__(self, qubits: int, ops, tuning=None, iterations: int = 1, dtype=np.complex128):
        h = C.c_void_p()
        t = tuning if tuning is not None else N.default_tuning()
        dt = N.HGM_C128 if np.dtype(dtype) == np.complex128 else N.HGM_C64
        check(lib.hgm_plan_build(qubits, dt, ops, len(ops), iterations, C.byref(t), C.byref(h)))
        self._h = h
        self.qubits = qubits

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.hgm_plan_destroy(h)
            self._h = None

    def __len__(self) -> int:
        return int(lib.hgm_plan_num_passes(self._h))

    def shape(self):
        r, b, rc, c = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        check(lib.hgm_plan_shape(self._h, C.byref(r), C.byref(b), C.byref(rc), C.byref(c)))
        return r.value, b.value, rc.value, c.value

    def passes(self):
        """List of (local_mask, [gate indices in execution order]) per pass."""
        out = []
        for i in range(len(self)):
            mask = C.c_uint64()
            n = C.c_int64()
            check(lib.hgm_plan_pass(self._h, i, C.byref(mask), C.byref(n), None))
            idx = np.zeros(max(1, n.value), dtype=np.int64)
            check(lib.hgm_plan_pass(self._h, i, C.byref(mask), C.byref(n), idx.ctypes.data))
            out.append((mask.value, idx[: n.value].tolist()))
        return out

    def run(self, state: State) -> dict:
        st = Stats()
        check(lib.hgm_state_run_plan(state._h, self._h, C.byref(st)))
        return st.as_dict()

    @property
    def kernel(self) -> int:
        return int(lib.hgm_plan_kernel(self._h))

    def set_graph(self, enable: bool = True) -> None:
        """Replay the passes from a captured hipGraph (plan-compiled plans)."""
        check(lib.hgm_plan_set_graph(self._h, int(enable)))

    @property
    def compile_seconds(self) -> float:
        return float(lib.hgm_plan_compile_seconds(self._h))

    def jit_source(self) -> str:
        n = C.c_size_t()
        check(lib.hgm_plan_jit_source(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib.hgm_plan_jit_source(self._h, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def program(self, i: int):
        """Register-kernel program of pass i: (structured ops, first map, last map, local_pos)."""
        n = C.c_int64()
        f, l = C.c_int32(), C.c_int32()
        check(lib.hgm_plan_program(self._h, i, None, C.byref(n), C.byref(f), C.byref(l), None))
        ops = np.zeros(n.value, dtype=REGOP_DTYPE)
        pos = np.zeros(16, dtype=np.int8)
        check(lib.hgm_plan_program(self._h, i, ops.ctypes.data, C.byref(n), C.byref(f), C.byref(l),
                                   pos.ctypes.data))
        return ops, f.value, l.value, pos

    def maps(self):
        n = C.c_int64()
        check(lib.hgm_plan_maps(self._h, None, C.byref(n)))
        maps = np.zeros(n.value, dtype=PHASEMAP_DTYPE)
        check(lib.hgm_plan_maps(self._h, maps.ctypes.data, C.byref(n)))
        return maps


class Batch:
    """States of any qubit counts, each advanced by its plan (any kernel
    family), replayed as one captured hipGraph (hgm_batch_*): run() is one
    graph launch for the whole batch. Keeps references to the states and
    plans."""

    def __init__(self, states, plans):
        if len(states) != len(plans):
            raise ValueError("one plan per state")
        self._states, self._plans = list(states), list(plans)
        n = len(self._states)
        sh = (C.c_void_p * max(1, n))(*[s._h for s in self._states])
        ph = (C.c_void_p * max(1, n))(*[p._h for p in self._plans])
        h = C.c_void_p()
        check(lib.hgm_batch_create(sh, ph, n, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.hgm_batch_destroy(h)
            self._h = None

    def run(self) -> dict:
        st = Stats()
        check(lib.hgm_batch_run(self._h, C.byref(st)))
        return st.as_dict()


# record layouts of csrc/planner.hpp (RegOp, PhaseMap)
REGOP_DTYPE = np.dtype([("kind", np.int32), ("rbit", np.int32), ("rmask", np.uint32), ("rval", np.uint32),
                        ("tmask", np.uint32), ("tval", np.uint32), ("gmask", np.uint64), ("gval", np.uint64),
                        ("m", np.float64, 8)])
PHASEMAP_DTYPE = np.dtype([("reg", np.uint16, 64), ("thr", np.uint16, (2, 32))])
assert REGOP_DTYPE.itemsize == 104 and PHASEMAP_DTYPE.itemsize == 256
