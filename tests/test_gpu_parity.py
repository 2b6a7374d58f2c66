"""Parity of the HIP t_GPU path (through the C ABI) with the oracle and with
the reference's own outputs. Tolerances: FP64 state vs FP64 oracle 1e-12
max-abs; FP32 state vs FP64 oracle 2e-5 max-abs (single-precision rounding of
~10^2 gates); measured/argmax indices exact."""
import ctypes as C
import sys

import numpy as np
import pytest

import simulador_quantico_amd as hq
from simulador_quantico_amd import _native as N
from oracle import pyoracle as po
from tests.test_oracle import _random_gate_list, grover_closed_form, oracle_run_gates, truth_table_cases

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dgm_for(n, dtype=np.complex128, **tun):
    d = hq.DGM(dtype)
    d.qubits = n
    d.exec_type = hq.T_GPU
    for k, v in tun.items():
        setattr(d, k, v)
    d.allocateMemory()
    return d


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_truth_tables(dtype):
    say("truth tables", dtype)
    for n, circ, start, exp in truth_table_cases():
        d = dgm_for(n, dtype)
        d.setMemoryValue(start)
        d.executeFunction(circ)
        want = np.zeros(1 << n)
        want[exp] = 1
        assert np.abs(d.state - want).max() < 1e-6, (circ, start)


@pytest.mark.parametrize("planner", [N.HGM_PLAN_GREEDY, N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT])
def test_golden_circuits(golden, planner):
    """FP32 path vs the reference's FP32 states; FP64 path vs the FP64 oracle;
    with and without the fusion and row-scaling rewrites."""
    for c in golden["circuits"]:
        say("golden", c["name"], planner)
        n = c["qubits"]
        ref32 = golden["states"][f"{c['name']}__tcpu"]
        for dtype in (np.complex64, np.complex128):
            d = dgm_for(n, dtype, planner=planner)
            for name, m in c["gates"].items():
                d.gates.addGate(name, [complex(a, b) for a, b in m])
            d.setMemoryValue(c["basis"])
            d.setFunction(c["steps"], c["iterations"])
            got = d.execute(1).copy()
            if dtype == np.complex64:
                assert np.abs(got - ref32).max() < TOL32, c["name"]
            else:
                o = po.Oracle(n, 64)
                for name, m in c["gates"].items():
                    o.add_gate(name, [complex(a, b) for a, b in m])
                o.set_function(c["steps"], c["iterations"])
                o.set_basis(c["basis"])
                o.execute(po.T_CPU)
                assert np.abs(got - o.state()).max() < TOL64, c["name"]


TUNINGS = [  # block_size, repeat_count, gpu_region_bits, gpu_coalesced_bits
    (64, 2, 8, 4), (128, 2, 9, 4), (256, 8, 12, 4), (256, 16, 13, 2), (64, 1, 7, 0), (32, 4, 8, 6),
    (1024, 4, 13, 5), (16, 1, 5, 1),
]


@pytest.mark.parametrize("n", [5, 9, 13, 16])
def test_random_gates_tuning_sweep(n):
    rng = np.random.default_rng(1000 + n)
    t, cm, cv, mats = _random_gate_list(rng, n, 120)
    ops = hq.make_ops(t, cm, cv, mats)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    want = oracle_run_gates(n, t, cm, cv, mats, state=psi0)
    for B, R, r, c in TUNINGS:
        for planner in (N.HGM_PLAN_GREEDY, N.HGM_PLAN_WINDOW,
                        N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_FOLD_PHASES,
                        N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT):
            for ops_block in (300, 5):
                say("sweep", n, (B, R, r, c), planner, ops_block)
                tun = N.default_tuning(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=c,
                                       planner=planner, ops_block=ops_block)
                for dtype, tol in ((np.complex128, TOL64), (np.complex64, TOL32)):
                    s = hq.State(n, dtype)
                    s.upload(psi0)
                    s.apply(ops, tun)
                    assert np.abs(s.download() - want).max() < tol, (B, R, r, c, planner, ops_block, dtype)


def test_iterations_replay_the_list():
    """t_GPU honours `iterations` (src/core/kernel.cu:192)."""
    n = 10
    rng = np.random.default_rng(3)
    t, cm, cv, mats = _random_gate_list(rng, n, 40)
    ops = hq.make_ops(t, cm, cv, mats)
    want = oracle_run_gates(n, list(t) * 3, list(cm) * 3, list(cv) * 3, list(mats) * 3, basis=7)
    s = hq.State(n)
    s.set_basis(7)
    st = s.apply(ops, N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9), iterations=3)
    assert st["gates"] == 120
    assert np.abs(s.download() - want).max() < TOL64


def test_hadamard_20q_config():
    """Hadamard benchmark, 20 qubits, FP64: every amplitude 2^-10 (1 layer and general.out's 3 layers)."""
    for layers in (1, 3):
        say("hadamard20", layers)
        d = dgm_for(20, block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=4)
        d.setMemoryValue(0)
        d.setFunction(hq.Hadamard(20, 0, 20), layers)
        st = d.execute(1)
        assert np.abs(st - 2.0 ** -10).max() < TOL64


def test_grover_22q_config():
    """Grover, 22 qubits, FP64, 1608 iterations: closed form within 1e-10, argmax exact."""
    n = 22
    marked = int(np.random.default_rng(22).integers(0, 1 << n))
    k = hq.GroverIterations(n)
    assert k == 1608
    d = hq.DGM()
    d.qubits = n
    d.setFunction(hq.GroverIteration(n, marked))
    ops = d.ops()
    d.setFunction(hq.Hadamard(n, 0, n))
    hops = d.ops()
    for planner in (N.HGM_PLAN_GREEDY, N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_FOLD_PHASES,
                    N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT):
        tun = N.default_tuning(block_size=256, repeat_count=16, gpu_region_bits=13, gpu_coalesced_bits=4,
                               planner=planner)
        s = hq.State(n)
        s.set_basis(0)
        s.apply(hops, tun)
        st = s.apply(ops, tun, iterations=k)
        say("grover22 planner", planner, "passes", st["passes"], "kernel_ms", st["kernel_ms"])
        idx, p = s.argmax()
        assert idx == marked
        psi = s.download()
        assert np.abs(psi - grover_closed_form(n, marked, k)).max() < 1e-10
        assert abs(p - np.sin((2 * k + 1) * np.arcsin(2 ** (-n / 2))) ** 2) < 1e-10


def test_draper_adder_29q_config():
    """2x14-bit Draper adder + carry (29 qubits): sum read bit-exact, FP64 and FP32."""
    w = 14
    rng = np.random.default_rng(29)
    a, b = (int(v) for v in rng.integers(0, 1 << w, 2))
    d = hq.DGM()
    d.setFunction(hq.DraperAdder(w, d))
    ops = d.ops()
    tun = N.default_tuning(block_size=256, repeat_count=16, gpu_region_bits=13, gpu_coalesced_bits=4)
    for dtype, floor in ((np.complex128, 1 - 1e-10), (np.complex64, 1 - 1e-5)):
        s = hq.State(2 * w + 1, dtype)
        s.set_basis((a << (w + 1)) | b)
        st = s.apply(ops, tun)
        idx, p = s.argmax()
        say("draper29", dtype, "passes", st["passes"], "p", p)
        assert idx == (a << (w + 1)) | (a + b)
        assert (idx & ((1 << (w + 1)) - 1)) == a + b
        assert p > floor
        assert abs(s.norm2() - 1) < (1e-12 if dtype == np.complex128 else 1e-5)


def inverse_ops(ops):
    """U^dagger of a gate list: reversed order, conjugate-transposed matrices."""
    t, cm, cv, mats = hq.ops_to_arrays(ops)
    inv = [m.reshape(2, 2).conj().T.reshape(4) for m in mats[::-1]]
    return hq.make_ops(t[::-1], cm[::-1], cv[::-1], inv)


def test_random_layers_20q_vs_oracle():
    n, depth = 20, 20
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, depth, 5, d))
    ops = d.ops()
    t, cm, cv, mats = hq.ops_to_arrays(ops)
    want = oracle_run_gates(n, t, cm, cv, mats, basis=0, engine=po.T_PAR_CPU)
    for tun in (N.default_tuning(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=4),
                N.default_tuning(block_size=256, repeat_count=16, gpu_region_bits=13, gpu_coalesced_bits=2),
                N.default_tuning(block_size=256, repeat_count=4, gpu_region_bits=11, gpu_coalesced_bits=3,
                                 kernel=N.HGM_KERNEL_JIT, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FOLD_PHASES),
                N.default_tuning(block_size=256, repeat_count=4, gpu_region_bits=11, gpu_coalesced_bits=3,
                                 kernel=N.HGM_KERNEL_JIT, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE |
                                 N.HGM_PLAN_PIVOT),
                N.default_tuning(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=3,
                                 kernel=N.HGM_KERNEL_JIT, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE |
                                 N.HGM_PLAN_ROTATE | N.HGM_PLAN_BALANCE),
                N.default_tuning(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=4,
                                 kernel=N.HGM_KERNEL_REG, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE |
                                 N.HGM_PLAN_ROTATE)):
        s = hq.State(n)
        s.set_basis(0)
        st = s.apply(ops, tun)
        say("random20 passes", st["passes"])
        assert np.abs(s.download() - want).max() < TOL64


PIVOT_JIT = dict(kernel=N.HGM_KERNEL_JIT, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT)


@pytest.mark.parametrize("n,dtype,tol,tuning", [
    (30, np.complex128, 1e-12, (256, 16, 13, 2, {})),
    (28, np.complex64, 2e-5, (256, 16, 13, 2, {})),
    (28, np.complex64, 2e-5, (512, 8, 13, 4, PIVOT_JIT))])
def test_random_layers_full_size_round_trip(n, dtype, tol, tuning):
    """U then U^dagger returns |0>; the norm stays 1 (size-independent checks at full size)."""
    depth = 20
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, depth, 1, d))
    ops = d.ops()
    B, R, r, c, extra = tuning
    tun = N.default_tuning(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=c, **extra)
    s = hq.State(n, dtype)
    s.set_basis(0)
    st = s.apply(ops, tun)
    say("random", n, "passes", st["passes"], "kernel_ms", st["kernel_ms"])
    assert abs(s.norm2() - 1) < 10 * tol
    s.apply(inverse_ops(ops), tun)
    idx, p = s.argmax()
    assert idx == 0
    assert abs(p - 1) < 10 * tol
    del s


def test_measure_matches_reference(golden):
    """Grover (reference form) measured on the t_GPU path: the reference's bits at its seeds."""
    for resident in (0, 1):
        for case in golden["measure"]:
            n, v = case["qubits"], case["search_value"]
            d = dgm_for(n, resident=resident)
            d.setMemoryValue(1 << (n - 1))
            d.setFunction(hq.Hadamard(n, 0, n))
            d.setFunction(hq.GroverIterationRef(n, v), case["iterations"], reset=False)
            d.execute(1)
            hq.seed_rng(case["seed"])
            bits = [d.measure(q) for q in range(1, n)]
            assert bits == case["bits"], (resident, case)


def _edge_amps(pairs, dtype):
    return np.array([complex(re, im) for re, im in pairs], dtype=dtype)


def test_measure_float_comparison(golden):
    """measure(qubit) on states and seeds written by the reference whose draw falls in a
    float rounding window of P(0) (src/core/dgm_core.cpp:75-99): the FP32 executor gives
    the reference's bit and collapsed state bit for bit on the host path; where P(0) is
    exact in float the device-resident path and the FP64 executor agree too."""
    say("measure edges")
    for case in golden["measure_edges"]:
        n, q = case["qubits"], case["qubit"]
        want_after = _edge_amps(case["after"], np.complex64)
        exact_p0 = case["name"].startswith("draw_rounding")
        runs = [(np.complex64, 0)] + ([(np.complex64, 1), (np.complex128, 0), (np.complex128, 1)] if exact_p0 else [])
        for dtype, resident in runs:
            d = dgm_for(n, dtype, resident=resident)
            d.state[:] = _edge_amps(case["amps"], dtype)
            if resident:   # X twice on the last qubit (exactly the identity) puts the state on the device
                x_last = ",".join(["ID"] * (n - 1) + ["X"])
                d.setFunction([x_last, x_last])
                assert d.execute(1) is None
            hq.seed_rng(case["seed"])
            bit = d.measure(q)
            assert bit == case["bit"], (case["name"], dtype, resident)
            got = d.state
            if dtype == np.complex64:
                assert np.array_equal(got, want_after), (case["name"], resident)
            else:
                assert np.abs(got - want_after).max() < 1e-7, (case["name"], resident)


def test_resident_state_view():
    """Resident mode: execute() keeps the result on the device and returns None; the
    state view brings it back, and writes through the view reach the next execute()."""
    n = 6
    for resident in (0, 1):
        d = dgm_for(n, resident=resident)
        d.setMemoryValue(0)
        d.setFunction(hq.Hadamard(n, 0, n))
        r = d.execute(1)
        assert (r is None) == bool(resident)
        assert np.abs(d.state - 2.0 ** (-n / 2)).max() < 1e-15
        d.state[:] = 0
        d.state[5] = 1
        d.setFunction(",".join(["X"] + ["ID"] * (n - 1)))
        d.execute(1)
        want = np.zeros(1 << n)
        want[5 | (1 << (n - 1))] = 1
        assert np.abs(d.state - want).max() == 0


def test_device_readout():
    n = 14
    rng = np.random.default_rng(9)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    s = hq.State(n)
    s.upload(psi)
    assert abs(s.norm2() - 1) < 1e-13
    for q in (0, 5, 13):
        bit = n - 1 - q
        p1 = np.sum(np.abs(psi[(np.arange(1 << n) >> bit) & 1 == 1]) ** 2)
        assert abs(s.prob_one(q) - p1) < 1e-13
    idx, p = s.argmax()
    assert idx == int(np.argmax(np.abs(psi) ** 2))
    # collapse-and-reset against the reference's host formula (src/core/dgm_core.cpp:115-122)
    q, out = 3, 1
    p1 = s.prob_one(q)
    s.collapse_reset(q, out, p1)
    bit = n - 1 - q
    want = np.zeros_like(psi)
    for i in range(1 << (n - 1)):
        pos = ((i >> bit) << (bit + 1)) | (i & ((1 << bit) - 1))
        want[pos] = psi[pos | (out << bit)] / np.sqrt(p1)
    assert np.abs(s.download() - want).max() < 1e-13


class RefPT(C.Structure):
    """The reference's struct PT on LP64 (include/core/common.h:57-90)."""
    _fields_ = [("qubits", C.c_int), ("matrix", C.c_void_p), ("matrix_size", C.c_int),
                ("span_start_bit", C.c_int), ("target_bit", C.c_int), ("affected", C.c_bool),
                ("control_value", C.c_long), ("control_mask", C.c_long), ("control_bit_positions", C.c_void_p),
                ("control_count", C.c_long), ("control_rest", C.c_void_p), ("control_rest_count", C.c_long)]


def test_gpu_execution_wrapper_reference_abi():
    """The reference-compatible entry point on an FP32 host state and a PT** list."""
    assert C.sizeof(RefPT) == 80
    n = 11
    rng = np.random.default_rng(77)
    t, cm, cv, mats = _random_gate_list(rng, n, 50)
    mats32 = [np.asarray(m, np.complex64) for m in mats]
    keep = []
    pts = (C.POINTER(RefPT) * (len(t) + 1))()
    for i in range(len(t)):
        m = np.ascontiguousarray(mats32[i])
        keep.append(m)
        p = RefPT(qubits=1, matrix=m.ctypes.data, matrix_size=2, span_start_bit=int(t[i]) + 1,
                  target_bit=int(t[i]), affected=False, control_value=int(cv[i]), control_mask=int(cm[i]),
                  control_count=bin(int(cm[i])).count("1"))
        keep.append(p)
        pts[i] = C.pointer(p)
    pts[len(t)] = None
    state = np.zeros(1 << n, np.complex64)
    state[5] = 1
    ret = N.lib.GpuExecutionWrapper(state.ctypes.data, C.cast(pts, C.c_void_p), n, 4, 8, 1, 64, 2, 2)
    assert ret == state.ctypes.data
    want = oracle_run_gates(n, list(t) * 2, list(cm) * 2, list(cv) * 2,
                            [m.astype(np.complex128) for m in mats32] * 2, basis=5)
    assert np.abs(state - want).max() < TOL32


@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_jit_kernel_parity(dtype, tol):
    """Plan-time compiled pass kernels against the oracle (controls on register,
    thread and tile bits, diagonal and anti-diagonal gates, several remaps)."""
    n = 14
    rng = np.random.default_rng(4242)
    t, cm, cv, mats = _random_gate_list(rng, n, 160)
    ops = hq.make_ops(t, cm, cv, mats)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    want = oracle_run_gates(n, t, cm, cv, mats, state=psi0)
    for B, R, r, c, planner in [(64, 8, 10, 3, N.HGM_PLAN_GREEDY), (32, 4, 8, 2, N.HGM_PLAN_GREEDY),
                                (64, 4, 9, 3, N.HGM_PLAN_GREEDY | N.HGM_PLAN_FOLD_PHASES),
                                (64, 4, 9, 3, N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_FOLD_PHASES),
                                (64, 4, 9, 3, N.HGM_PLAN_GREEDY | N.HGM_PLAN_PIVOT),
                                (64, 4, 9, 3, N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT),
                                (64, 8, 10, 3, N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT |
                                 N.HGM_PLAN_BALANCE),
                                (64, 4, 9, 3, N.HGM_PLAN_GREEDY | N.HGM_PLAN_PIVOT | N.HGM_PLAN_LOOKAHEAD),
                                (64, 4, 9, 3, N.HGM_PLAN_GREEDY | N.HGM_PLAN_ROTATE),
                                (64, 8, 10, 3, N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_ROTATE |
                                 N.HGM_PLAN_BALANCE),
                                (32, 16, 10, 2, N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_ROTATE)]:
        tun = N.default_tuning(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=c,
                               kernel=N.HGM_KERNEL_JIT, ops_block=60, planner=planner)
        plan = hq.Plan(n, ops, tun, 1, dtype)
        assert plan.kernel == N.HGM_KERNEL_JIT
        s = hq.State(n, dtype)
        s.upload(psi0)
        plan.run(s)
        say("jit", (B, R, r, c, planner), dtype, "passes", len(plan))
        assert np.abs(s.download() - want).max() < tol


@pytest.mark.parametrize("flag,tile", [(N.HGM_PLAN_FOLD_PHASES, (256, 4, 11, 3)),
                                       (N.HGM_PLAN_PIVOT | N.HGM_PLAN_BALANCE, (256, 8, 12, 3)),
                                       (N.HGM_PLAN_ROTATE | N.HGM_PLAN_BALANCE, (256, 8, 12, 3))])
def test_jit_random_layers_30q_round_trip(flag, tile):
    """The benchmark's 30-qubit circuit and configuration: U then U^dagger
    returns |0> and the norm stays 1 (size-independent checks at full size)."""
    n = 30
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, 20, 1, d))
    ops = d.ops()
    B, R, r, c = tile
    tun = N.default_tuning(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=c,
                           kernel=N.HGM_KERNEL_JIT, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | flag)
    s = hq.State(n)
    s.set_basis(0)
    st = hq.Plan(n, ops, tun).run(s)
    say("jit random30 passes", st["passes"], "kernel_ms", st["kernel_ms"])
    assert abs(s.norm2() - 1) < 1e-11
    hq.Plan(n, inverse_ops(ops), tun).run(s)
    idx, p = s.argmax()
    assert idx == 0 and abs(p - 1) < 1e-11


def test_jit_graph_replay_matches_direct_launches():
    """hipGraph replay of a plan-compiled plan: bit-identical to direct launches, recaptured for another state,
    several replays in a row with eager work in between."""
    n = 16
    rng = np.random.default_rng(77)
    t, cm, cv, mats = _random_gate_list(rng, n, 120)
    ops = hq.make_ops(t, cm, cv, mats)
    tun = N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9, gpu_coalesced_bits=3,
                           kernel=N.HGM_KERNEL_JIT)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    direct = hq.Plan(n, ops, tun)
    graph = hq.Plan(n, ops, tun)
    graph.set_graph(True)
    a, b, c = hq.State(n), hq.State(n), hq.State(n)
    for s in (a, b, c):
        s.upload(psi0)
    for rep in range(3):
        direct.run(a)
        graph.run(b)
        assert np.array_equal(a.download(), b.download()), rep
        assert abs(b.norm2() - 1) < 1e-12          # eager work between replays
    graph.run(c)                                    # another state: captured again
    direct_c = hq.State(n)
    direct_c.upload(psi0)
    direct.run(direct_c)
    assert np.array_equal(c.download(), direct_c.download())
    with pytest.raises(N.HGMError):
        hq.Plan(n, ops, N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9,
                                         kernel=N.HGM_KERNEL_REG)).set_graph(True)


def test_jit_modules_beyond_the_cache_bound():
    """More distinct plan-compiled plans than the library keeps loaded (16):
    each still matches the oracle when run again after its module went out,
    and a graph captured before the eviction is captured again, not replayed
    with unloaded code."""
    n = 12
    rng = np.random.default_rng(91)
    tun = N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9, gpu_coalesced_bits=3,
                           kernel=N.HGM_KERNEL_JIT)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    cases = []
    for _ in range(18):
        t, cm, cv, mats = _random_gate_list(rng, n, 24)
        cases.append((hq.Plan(n, hq.make_ops(t, cm, cv, mats), tun),
                      oracle_run_gates(n, t, cm, cv, mats, state=psi0)))
    cases[0][0].set_graph(True)
    for rep in range(2):
        for k, (plan, want) in enumerate(cases):
            s = hq.State(n)
            s.upload(psi0)
            plan.run(s)
            assert np.abs(s.download() - want).max() < TOL64, (rep, k)


def test_one_call_entry_reuses_plans_only_for_the_same_circuit():
    """hgm_execute keeps recent plans: the same gate list runs again without
    planning, and a list that differs in one matrix entry gets its own plan."""
    import ctypes as C
    n = 12
    rng = np.random.default_rng(31)
    t, cm, cv, mats = _random_gate_list(rng, n, 60)
    tun = N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9, gpu_coalesced_bits=3,
                           kernel=N.HGM_KERNEL_JIT, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)

    def run(mats_):
        ops = hq.make_ops(t, cm, cv, mats_)
        psi = psi0.astype(np.complex128).copy()
        st = N.Stats()
        N.check(N.lib.hgm_execute(psi.ctypes.data, n, N.HGM_C128, ops, len(ops), C.byref(tun), 2, C.byref(st)))
        return psi, st.plan_ms

    first, ms1 = run(mats)
    again, ms2 = run(mats)
    assert np.array_equal(first, again)
    assert ms2 < 0.2 * ms1 + 1.0, (ms1, ms2)
    mats2 = [m.copy() for m in mats]
    k = next(i for i, m in enumerate(mats2) if m[1] != 0)
    rot = np.array([[np.cos(0.3), -np.sin(0.3)], [np.sin(0.3), np.cos(0.3)]])
    mats2[k] = (np.asarray(mats2[k]).reshape(2, 2) @ rot).reshape(4)
    other, _ = run(mats2)
    want = oracle_run_gates(n, np.tile(t, 2), np.tile(cm, 2), np.tile(cv, 2), list(mats2) * 2, state=psi0)
    assert np.abs(other - want).max() < 1e-12


def test_release_cache_then_reuse():
    """hgm_release_cache frees the cached device buffers, drops cached plans
    and unloads compiled modules; the next calls rebuild them and give the
    same results, for the one-call path and for a plan built before."""
    import ctypes as C
    n = 13
    rng = np.random.default_rng(5)
    t, cm, cv, mats = _random_gate_list(rng, n, 40)
    ops = hq.make_ops(t, cm, cv, mats)
    tun = N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9, gpu_coalesced_bits=3,
                           kernel=N.HGM_KERNEL_JIT, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT)
    want = oracle_run_gates(n, t, cm, cv, mats, basis=3)
    plan = hq.Plan(n, ops, tun)
    for _ in range(2):
        psi = np.zeros(1 << n, np.complex128)
        psi[3] = 1
        N.check(N.lib.hgm_execute(psi.ctypes.data, n, N.HGM_C128, ops, len(ops), C.byref(tun), 1, None))
        assert np.abs(psi - want).max() < 1e-12
        s = hq.State(n)
        s.set_basis(3)
        plan.run(s)
        assert np.abs(s.download() - want).max() < 1e-12
        del s
        N.check(N.lib.hgm_release_cache())


def test_one_call_entry_from_several_threads():
    """Four threads run two different circuits through hgm_execute at once
    (the plan cache and the device context are shared); every result matches
    the oracle."""
    import ctypes as C
    import threading
    n = 11
    rng = np.random.default_rng(8)
    circuits = []
    for _ in range(2):
        t, cm, cv, mats = _random_gate_list(rng, n, 50)
        want = oracle_run_gates(n, t, cm, cv, mats, basis=1)
        circuits.append((hq.make_ops(t, cm, cv, mats), want))
    tun = N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9, gpu_coalesced_bits=3,
                           planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT)
    errors = []

    def worker(w):
        try:
            for i in range(6):
                ops, want = circuits[(w + i) % 2]
                psi = np.zeros(1 << n, np.complex128)
                psi[1] = 1
                N.check(N.lib.hgm_execute(psi.ctypes.data, n, N.HGM_C128, ops, len(ops), C.byref(tun), 1, None))
                err = float(np.abs(psi - want).max())
                if err >= 1e-12:
                    errors.append((w, i, err))
        except Exception as e:   # reported by the main thread
            errors.append((w, repr(e)))

    threads = [threading.Thread(target=worker, args=(w,)) for w in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads)
    assert not errors, errors


C_GPU_PROGRAM = r"""
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "hgm_gpu.h"

/* H on every qubit of a 20-qubit FP64 state that starts in |0x5a5a5>, through
   the POD API only: once gives amplitudes 2^-10 * (-1)^popcount(k & 0x5a5a5),
   twice gives the start state back. */
int main(void) {
    const int n = 20;
    const double h = 1.0 / sqrt(2.0);
    hgm_op ops[20];
    for (int q = 0; q < n; ++q) {
        memset(&ops[q], 0, sizeof(ops[q]));
        ops[q].target_bit = q;
        ops[q].m[0] = h; ops[q].m[2] = h; ops[q].m[4] = h; ops[q].m[6] = -h;
    }
    if (hgm_init(0)) { printf("init: %s\n", hgm_last_error()); return 1; }
    hgm_tuning t;
    hgm_default_tuning(&t);
    t.block_size = 256; t.repeat_count = 8; t.gpu_region_bits = 12; t.gpu_coalesced_bits = 3;
    t.kernel = HGM_KERNEL_JIT;
    t.planner = HGM_PLAN_GREEDY | HGM_PLAN_FUSE | HGM_PLAN_PIVOT | HGM_PLAN_BALANCE;
    hgm_state *s;
    hgm_plan *p;
    hgm_stats st;
    memset(&st, 0, sizeof(st));   /* calls add to the counts */
    if (hgm_state_create(n, HGM_C128, &s) || hgm_state_set_basis(s, 0x5a5a5)) return 2;
    if (hgm_plan_build(n, HGM_C128, ops, n, 1, &t, &p)) { printf("plan: %s\n", hgm_last_error()); return 3; }
    if (hgm_state_run_plan(s, p, &st)) { printf("run: %s\n", hgm_last_error()); return 4; }
    static double amps[2 << 20];
    if (hgm_state_download(s, amps)) return 5;
    double err = 0.0;
    for (unsigned k = 0; k < (1u << n); ++k) {
        double want = ldexp(__builtin_popcount(k & 0x5a5a5u) & 1 ? -1.0 : 1.0, -10);
        err = fmax(err, fmax(fabs(amps[2 * k] - want), fabs(amps[2 * k + 1])));
    }
    if (hgm_state_run_plan(s, p, &st)) return 6;
    uint64_t idx;
    double prob, norm;
    if (hgm_state_argmax(s, &idx, &prob) || hgm_state_norm2(s, &norm)) return 7;
    printf("%.3e %llu %.15f %.15f %lld\n", err, (unsigned long long)idx, prob, norm, (long long)st.passes);
    hgm_plan_destroy(p);
    hgm_state_destroy(s);
    return 0;
}
"""


def test_c_program_on_the_pod_api(tmp_path):
    """A plain C11 program drives the GPU through include/hgm_gpu.h alone:
    state, plan-compiled kernels, download, argmax, norm (no Python, no torch)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src, exe = tmp_path / "pod.c", tmp_path / "pod"
    src.write_text(C_GPU_PROGRAM)
    lib_dir = os.path.join(root, "simulador_quantico_amd")
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-I", os.path.join(root, "include"), str(src),
                    "-o", str(exe), "-L", lib_dir, "-lhgm", "-lm", f"-Wl,-rpath,{lib_dir}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    say("C program:", r.returncode, r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    err, idx, prob, norm, passes = r.stdout.split()
    assert float(err) < 1e-15
    assert int(idx) == 0x5a5a5 and abs(float(prob) - 1) < 1e-12 and abs(float(norm) - 1) < 1e-12
    assert int(passes) == 4   # two runs of the two-pass plan (20 bits, 12 in a tile), counts added up


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_one_call_entry_on_pageable_and_pinned_host_states(dtype):
    """hgm_execute on a pageable host state (a numpy array: staged through the
    library's pinned chunks) and on a pinned one (the DGM's state: straight
    DMA) gives the same bits; a Hadamard layer on |0..0> at 26 qubits is the
    uniform state. Prints the copy rates of both paths."""
    n = 26
    h = 1 / np.sqrt(2)
    ops = hq.make_ops(list(range(n)), [0] * n, [0] * n, [np.array([h, h, h, -h])] * n)
    code = N.HGM_C128 if dtype == np.complex128 else N.HGM_C64
    tun = N.default_tuning(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=3)
    pageable = np.zeros(1 << n, dtype)
    pageable[0] = 1
    d = dgm_for(n, dtype)                       # pinned host state
    d.setMemoryValue(0)
    pinned = d.state
    out = {}
    for name, arr in (("pageable", pageable), ("pinned", pinned)):
        for _ in range(2):                      # the second call is timed: staging and plan set up
            arr[:] = 0
            arr[0] = 1
            st = N.Stats()
            check = N.lib.hgm_execute(arr.ctypes.data, n, code, ops, len(ops), C.byref(tun), 1, C.byref(st))
            assert check == 0, N.lib.hgm_last_error()
        nbytes = arr.nbytes
        say(f"{name} {np.dtype(dtype).name}: H2D {nbytes / st.h2d_ms / 1e6:.1f} GB/s, "
            f"D2H {nbytes / st.d2h_ms / 1e6:.1f} GB/s, kernels {st.kernel_ms:.2f} ms")
        out[name] = arr.copy()
    assert np.array_equal(out["pageable"], out["pinned"])
    tol = 1e-6 if dtype == np.complex64 else 1e-15
    assert np.abs(out["pinned"] - 2.0 ** (-n / 2)).max() < tol


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_executor_t_hybrid_runs_every_region_on_the_gpu(dtype):
    """exec_type t_HYBRID in the executor: every region on the GPU (the
    reference's schedule at thread_count 1) and the PT list applied once,
    whatever execute()'s iterations (DGM::execute calls HybridExecution(pts),
    src/core/dgm_core.cpp:262-264): the same bits as t_GPU with one iteration."""
    n = 12
    rng = np.random.default_rng(12)
    t, cm, cv, mats = _random_gate_list(rng, n, 80)
    out = {}
    for et, iters in ((hq.T_GPU, 1), (hq.T_HYBRID, 3)):
        d = dgm_for(n, dtype)
        d.exec_type = et
        for i, m in enumerate(mats):
            d.gates.addGate(f"g{i}", list(m))
        steps = []
        for i, (tb, m_, v_) in enumerate(zip(t, cm, cv)):
            tok = ["ID"] * n
            tok[n - 1 - tb] = f"Target1(g{i})" if m_ else f"g{i}"
            for b in range(n):
                if (m_ >> b) & 1:
                    tok[n - 1 - b] = f"Control1({(v_ >> b) & 1})"
            steps.append(",".join(tok))
        d.setMemoryValue(5)
        d.setFunction(steps)
        out[et] = d.execute(iters).copy()
    assert np.array_equal(out[hq.T_GPU], out[hq.T_HYBRID])
    want = oracle_run_gates(n, t, cm, cv, mats, basis=5)
    assert np.abs(out[hq.T_GPU] - want).max() < (TOL32 if dtype == np.complex64 else TOL64)
