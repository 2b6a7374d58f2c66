"""Serving front end: the request mix, the batched server (hgm_server_* of
include/hgm_gpu.h) and the sequential one-circuit-per-call baseline.

The mix: requests with Poisson arrivals at a fixed rate, qubit counts uniform
in [qmin, qmax], circuit type uniform over {Hadamard layer, Grover with at
most 64 iterations, QFT (Draper) adder, random-layer circuit of depth 10},
parameters drawn from one seeded generator. Circuits are built with the same
generators as the executor API (src/core/gates.cpp, src/algorithms/lib_grover.cpp,
src/algorithms/lib_shor_circuits.cpp in the reference).
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from ._native import check, lib
from .circuits import DraperAdder, GroverIteration, Hadamard, RandomLayers
from .dgm import DGM

KINDS = ("hadamard", "grover", "adder", "random")

# numpy mirror of hgm_op (88 bytes)
OP_NP = np.dtype([("target_bit", "<i4"), ("reserved", "<i4"), ("control_mask", "<u8"), ("control_value", "<u8"),
                  ("m", "<f8", (8,))])
assert OP_NP.itemsize == C.sizeof(N.Op)


def ops_np(ops) -> np.ndarray:
    """A ctypes Op array as a (copied) numpy record array."""
    return np.frombuffer(bytes(ops), dtype=OP_NP).copy() if len(ops) else np.zeros(0, OP_NP)


def _op_ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(N.Op))


@dataclass
class Job:
    index: int
    arrival_s: float
    kind: str
    qubits: int
    ops: np.ndarray
    basis: int = 0
    expect: int | None = None          # index the readout must return (adder: a, a+b)
    params: dict = field(default_factory=dict)


def _circuit_ops(steps, dgm=None) -> np.ndarray:
    d = dgm if dgm is not None else DGM()
    d.setFunction(steps)
    return ops_np(d.ops())


def build_job(index: int, arrival_s: float, kind: str, n: int, rng: np.random.Generator,
              grover_max_iter: int = 64, random_depth: int = 10) -> Job:
    if kind == "hadamard":
        return Job(index, arrival_s, kind, n, _circuit_ops([Hadamard(n, 0, n)]), 0, None, {})
    if kind == "grover":
        marked = int(rng.integers(0, 1 << n))
        k = int(rng.integers(1, grover_max_iter + 1))
        h = _circuit_ops([Hadamard(n, 0, n)])
        it = _circuit_ops(GroverIteration(n, marked))
        return Job(index, arrival_s, kind, n, np.concatenate([h] + [it] * k), 0, None,
                   {"marked": marked, "iterations": k})
    if kind == "adder":
        w = (n - 1) // 2
        d = DGM()
        ops = _circuit_ops(DraperAdder(w, d), d)     # acts on the low 2w+1 index bits
        a, b = (int(v) for v in rng.integers(0, 1 << w, 2))
        basis = (a << (w + 1)) | b
        return Job(index, arrival_s, kind, n, ops, basis, (a << (w + 1)) | (a + b), {"a": a, "b": b, "width": w})
    if kind == "random":
        seed = int(rng.integers(0, 2 ** 31))
        d = DGM()
        return Job(index, arrival_s, kind, n, _circuit_ops(RandomLayers(n, random_depth, seed, d), d), 0, None,
                   {"seed": seed, "depth": random_depth})
    raise KeyError(kind)


def mix_shapes(n: int, grover_iterations=(1, 2, 3, 4), random_depth: int = 10):
    """One circuit of each shape the mix draws at `n` qubits, for
    Server.precompile: the Hadamard layer, Grover with 1-4 iterations (their
    passes already hold every pass structure of 1-64 iterations), the Draper
    adder and random layers. Matrix entries and control values do not matter:
    kernels are shared by structure."""
    h = _circuit_ops([Hadamard(n, 0, n)])
    it = _circuit_ops(GroverIteration(n, (1 << n) - 1))
    out = [(h, 1)] + [(np.concatenate([h] + [it] * k), 1) for k in grover_iterations]
    w = (n - 1) // 2
    if w >= 1:
        d = DGM()
        out.append((_circuit_ops(DraperAdder(w, d), d), 1))
    d = DGM()
    out.append((_circuit_ops(RandomLayers(n, random_depth, 1, d), d), 1))
    return out


def make_mix(n_requests: int = 4096, rate: float = 1000.0, seed: int = 2026, qmin: int = 10, qmax: int = 26,
             grover_max_iter: int = 64, random_depth: int = 10) -> list[Job]:
    """The serving mix: Poisson arrivals at `rate` requests/s (first at t=0)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n_requests)
    arrivals = np.concatenate([[0.0], np.cumsum(gaps[1:])])
    qubits = rng.integers(qmin, qmax + 1, n_requests)
    kinds = rng.integers(0, len(KINDS), n_requests)
    return [build_job(i, float(arrivals[i]), KINDS[int(kinds[i])], int(qubits[i]), rng, grover_max_iter,
                      random_depth) for i in range(n_requests)]


class Server:
    """Batched serving engine on the current device (hgm_server_*)."""

    def __init__(self, dtype=np.complex64, max_active: int = 256, hbm_budget_gib: float = 32.0,
                 planner_threads: int = 4, autotune: bool = False, qubit_range=(10, 26)):
        self.dtype = np.dtype(dtype)
        cfg = N.ServerConfig()
        lib.hgm_server_default_config(C.byref(cfg))
        cfg.dtype = N.HGM_C64 if self.dtype == np.complex64 else N.HGM_C128
        cfg.max_active = max_active
        cfg.hbm_budget_bytes = int(hbm_budget_gib * (1 << 30))
        cfg.planner_threads = planner_threads
        cfg.autotune = int(autotune)
        cfg.min_qubits, cfg.max_qubits = qubit_range
        h = C.c_void_p()
        check(lib.hgm_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._keep = {}      # id -> host buffer receiving the final state
        self._lock = threading.Lock()

    def submit(self, qubits: int, ops: np.ndarray, basis: int = 0, iterations: int = 1,
               state_out: np.ndarray | None = None) -> int:
        ops = np.ascontiguousarray(ops, dtype=OP_NP)
        rq = N.Request()
        rq.qubits = qubits
        rq.iterations = iterations
        rq.ops = _op_ptr(ops)
        rq.n_ops = len(ops)
        rq.basis = basis
        if state_out is not None:
            if state_out.dtype != self.dtype or state_out.size != (1 << qubits) or not state_out.flags.c_contiguous:
                raise ValueError("state_out must be a contiguous array of 2^qubits amplitudes of the server dtype")
            rq.state_out = state_out.ctypes.data
        i = C.c_int64()
        with self._lock:
            check(lib.hgm_server_submit(self._h, C.byref(rq), C.byref(i)))
            if state_out is not None:
                self._keep[i.value] = state_out
        return i.value

    def poll(self, timeout_ms: int = 0, cap: int = 256) -> list[dict]:
        buf = (N.Result * cap)()
        n = C.c_int()
        check(lib.hgm_server_poll(self._h, buf, cap, C.byref(n), int(timeout_ms)))
        out = [buf[k].as_dict() for k in range(n.value)]
        with self._lock:
            for r in out:
                self._keep.pop(r["id"], None)
        return out

    def wait_all(self) -> None:
        check(lib.hgm_server_wait_all(self._h))

    def tuning(self, qubits: int) -> dict:
        t = N.Tuning()
        check(lib.hgm_server_tuning(self._h, qubits, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def set_tuning(self, qubits: int, **tuning) -> None:
        """Tile shape for requests of `qubits` qubits submitted from now on
        (fields of hgm_tuning; kernel is the register kernel)."""
        t = N.Tuning()
        check(lib.hgm_server_tuning(self._h, qubits, C.byref(t)))
        for k, v in tuning.items():
            setattr(t, k, int(v))
        check(lib.hgm_server_set_tuning(self._h, qubits, C.byref(t)))

    def precompile(self, qubits: int, ops: np.ndarray, iterations: int = 1) -> int:
        """Compiles the pass kernels of this circuit shape (hgm_server_precompile);
        returns the number of kernels added."""
        ops = np.ascontiguousarray(ops, dtype=OP_NP)
        n = C.c_int()
        check(lib.hgm_server_precompile(self._h, qubits, _op_ptr(ops), len(ops), iterations, C.byref(n)))
        return n.value

    def clock_ms(self) -> float:
        return float(lib.hgm_server_clock_ms(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.hgm_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def run_served(jobs: list[Job], server: Server, time_scale: float = 1.0) -> list[dict]:
    """Submits every job at its arrival time (scaled) from a feeder thread and
    collects the results; each result gets latency_ms = done - submitted."""
    results: dict[int, dict] = {}
    ids: dict[int, int] = {}
    err: list[BaseException] = []

    def feeder():
        try:
            t0 = time.perf_counter()
            for j in jobs:
                delay = j.arrival_s * time_scale - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids[server.submit(j.qubits, j.ops, j.basis)] = j.index
        except BaseException as e:   # surfaced by the caller
            err.append(e)

    th = threading.Thread(target=feeder)
    th.start()
    while len(results) < len(jobs):
        for r in server.poll(timeout_ms=50):
            results[r["id"]] = r
        if err:
            break
        if not th.is_alive() and len(results) < len(jobs) and len(ids) < len(jobs):
            break
    th.join()
    if err:
        raise err[0]
    out = []
    for rid, r in sorted(results.items()):
        r = dict(r)
        r["job"] = ids[rid]
        r["latency_ms"] = r["t_done_ms"] - r["t_submit_ms"]
        out.append(r)
    out.sort(key=lambda r: r["job"])
    return out


def sequential_tuning(n: int):
    """Tile shape of the one-circuit-per-call FP64 baseline."""
    return N.default_tuning(block_size=256, repeat_count=4, gpu_region_bits=11, gpu_coalesced_bits=2,
                            ops_block=1 << 30)


def run_sequential(jobs: list[Job], dtype=np.complex128) -> list[dict]:
    """The plain drop-in path, one request after the other: host state in,
    hgm_execute (upload, plan, passes, download), readout on the host. Returns
    the service time of each request; queueing is applied by `queue_latency`."""
    out = []
    code = N.HGM_C128 if np.dtype(dtype) == np.complex128 else N.HGM_C64
    for j in jobs:
        t0 = time.perf_counter()
        psi = np.zeros(1 << j.qubits, dtype=dtype)
        psi[j.basis] = 1
        st = N.Stats()
        tun = sequential_tuning(j.qubits)
        check(lib.hgm_execute(psi.ctypes.data, j.qubits, code, _op_ptr(j.ops), len(j.ops), C.byref(tun), 1,
                              C.byref(st)))
        p = psi.real ** 2 + psi.imag ** 2
        idx = int(np.argmax(p))
        out.append({"job": j.index, "service_ms": (time.perf_counter() - t0) * 1e3, "argmax": idx,
                    "p_max": float(p[idx]), "norm2": float(p.sum())})
    return out


def queue_latency(arrivals_s, service_ms) -> np.ndarray:
    """Latency (ms) of a single server taking requests in arrival order."""
    t = 0.0
    lat = []
    for a, s in zip(arrivals_s, service_ms):
        a_ms = a * 1e3
        t = max(t, a_ms) + s
        lat.append(t - a_ms)
    return np.array(lat)


def fidelity(a: np.ndarray, b: np.ndarray) -> float:
    a = a.astype(np.complex128)
    b = b.astype(np.complex128)
    return float(abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real))
