"""The batched serving engine on the GPU: every result is the same alone and
in the mix, FP32 fidelity against FP64 >= 1 - 1e-6, FP64 results within 1e-12
of the oracle, readouts exact, admission by HBM footprint."""
import sys

import numpy as np
import pytest

import simulador_quantico_amd as hq
from simulador_quantico_amd import _native as N
from simulador_quantico_amd import serving as S
from tests.test_serving_mix import oracle_run_job

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_mix(count, seed, qmin, qmax, grover_max_iter=6):
    return S.make_mix(count, rate=1000.0, seed=seed, qmin=qmin, qmax=qmax, grover_max_iter=grover_max_iter)


def run_collect(srv, jobs, dtype, together):
    """Final states and readouts of every job, submitted all at once or one by one."""
    states = {j.index: np.zeros(1 << j.qubits, dtype) for j in jobs}
    res = {}
    if together:
        ids = {srv.submit(j.qubits, j.ops, j.basis, state_out=states[j.index]): j.index for j in jobs}
        srv.wait_all()
        for r in srv.poll(cap=len(jobs) + 8):
            res[ids[r["id"]]] = r
    else:
        for j in jobs:
            rid = srv.submit(j.qubits, j.ops, j.basis, state_out=states[j.index])
            srv.wait_all()
            (r,) = srv.poll(cap=4)
            assert r["id"] == rid
            res[j.index] = r
    assert len(res) == len(jobs)
    return states, res


def test_served_results_same_alone_and_in_mix_fp32():
    jobs = small_mix(28, seed=5, qmin=10, qmax=17)
    say("serving alone vs mix:", [(j.kind, j.qubits) for j in jobs])
    with S.Server(np.complex64, max_active=64, hbm_budget_gib=2) as srv:
        alone, ra = run_collect(srv, jobs, np.complex64, together=False)
        mixed, rm = run_collect(srv, jobs, np.complex64, together=True)
    for j in jobs:
        a, m = alone[j.index], mixed[j.index]
        assert ra[j.index]["status"] == 0 and rm[j.index]["status"] == 0
        # same kernels and programs for a state whatever runs beside it: tolerance 1e-7 max-abs (expected 0)
        assert np.abs(a - m).max() <= 1e-7, (j.kind, j.qubits)
        assert ra[j.index]["argmax"] == rm[j.index]["argmax"]
        assert abs(ra[j.index]["norm2"] - rm[j.index]["norm2"]) <= 1e-12
        ref = oracle_run_job(j)
        f = S.fidelity(ref, m)
        assert f >= 1 - 1e-6, (j.kind, j.qubits, f)
        p = np.abs(m.astype(np.complex128)) ** 2
        assert abs(rm[j.index]["norm2"] - p.sum()) < 1e-9
        assert rm[j.index]["argmax"] == int(np.argmax(p))
        if j.kind == "adder":
            assert rm[j.index]["argmax"] == j.expect


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_compiled_pass_kernels_against_oracle_alone_and_in_mix(dtype):
    """Pass kernels compiled per circuit shape (hgm_server_precompile): requests
    of those shapes with other seeds, marked values and operands run them,
    beside requests that run the interpreter (other shapes, one-tile states),
    with the same results alone and in the mix and within FP32 (FP64: 1e-12)
    of the oracle."""
    jobs = small_mix(32, seed=21, qmin=12, qmax=16)
    tol = 2e-5 if dtype == np.complex64 else 1e-12
    with S.Server(dtype, max_active=64, hbm_budget_gib=2) as srv:
        added = 0
        for n in (14, 15):                      # 16 qubits stays on the interpreter
            for ops, iters in S.mix_shapes(n):
                added += srv.precompile(n, ops, iters)
        again = sum(srv.precompile(14, ops, iters) for ops, iters in S.mix_shapes(14))
        say("compiled serving kernels:", added, "again:", again, [(j.kind, j.qubits) for j in jobs])
        assert added > 0 and again == 0
        assert srv.precompile(12, jobs[0].ops) == 0     # one tile: the interpreter runs it whole
        with pytest.raises(N.HGMError):
            srv.precompile(15, jobs[0].ops[:0])           # no gate
        with pytest.raises(N.HGMError):
            srv.precompile(41, jobs[0].ops)               # qubit count out of range
        alone, ra = run_collect(srv, jobs, dtype, together=False)
        mixed, rm = run_collect(srv, jobs, dtype, together=True)
    assert {j.qubits for j in jobs} >= {12, 14, 15, 16}
    for j in jobs:
        a, m = alone[j.index], mixed[j.index]
        assert ra[j.index]["status"] == 0 and rm[j.index]["status"] == 0
        assert np.abs(a - m).max() <= 1e-7, (j.kind, j.qubits)      # expected 0
        assert ra[j.index]["argmax"] == rm[j.index]["argmax"]
        ref = oracle_run_job(j)
        f = S.fidelity(ref, m)
        assert f >= 1 - 1e-6, (j.kind, j.qubits, f)
        assert np.abs(m - ref).max() < tol, (j.kind, j.qubits)
        p = np.abs(m.astype(np.complex128)) ** 2
        assert abs(rm[j.index]["norm2"] - p.sum()) < 1e-9
        assert rm[j.index]["argmax"] == int(np.argmax(p))
        if j.kind == "adder":
            assert rm[j.index]["argmax"] == j.expect


def test_precompile_while_requests_are_in_flight():
    """Kernels added while requests of the same qubit counts are queued,
    planned and running: requests planned before run the interpreter, later
    ones the new kernels, and every result is right."""
    jobs = small_mix(24, seed=33, qmin=14, qmax=16)
    with S.Server(np.complex64, max_active=8, hbm_budget_gib=1) as srv:
        outs = {j.index: np.zeros(1 << j.qubits, np.complex64) for j in jobs}
        ids = {}
        for k, j in enumerate(jobs):
            ids[srv.submit(j.qubits, j.ops, j.basis, state_out=outs[j.index])] = j
            if k == 6:
                for n in (14, 15, 16):
                    for ops, iters in S.mix_shapes(n):
                        srv.precompile(n, ops, iters)
        srv.wait_all()
        res = {r["id"]: r for r in srv.poll(cap=64)}
    assert len(res) == len(jobs)
    for rid, j in ids.items():
        assert res[rid]["status"] == 0
        f = S.fidelity(oracle_run_job(j), outs[j.index])
        assert f >= 1 - 1e-6, (j.kind, j.qubits, f)


def _fp64_reference(job):
    """The job's final state through the plain FP64 drop-in path (hgm_execute)."""
    import ctypes as C
    psi = np.zeros(1 << job.qubits, np.complex128)
    psi[job.basis] = 1
    tun = S.sequential_tuning(job.qubits)
    st = N.Stats()
    N.check(N.lib.hgm_execute(psi.ctypes.data, job.qubits, N.HGM_C128, S._op_ptr(job.ops), len(job.ops),
                              C.byref(tun), 1, C.byref(st)))
    return psi


@pytest.mark.parametrize("compiled", [False, True])
def test_large_tiles_at_21_to_24_qubits(compiled):
    """FP32 states of 21-24 qubits default to 2^13 and 2^14 tiles (512 and 1024
    threads): interpreted and compiled, Grover, random layers and the adder
    match the FP64 drop-in path."""
    rng = np.random.default_rng(44)
    jobs = [S.build_job(0, 0.0, "grover", 21, rng, grover_max_iter=3),
            S.build_job(1, 0.0, "random", 22, rng),
            S.build_job(2, 0.0, "grover", 23, rng, grover_max_iter=3),
            S.build_job(3, 0.0, "adder", 24, rng),
            S.build_job(4, 0.0, "random", 24, rng)]
    with S.Server(np.complex64, max_active=8, hbm_budget_gib=4) as srv:
        assert srv.tuning(22)["gpu_region_bits"] == 13 and srv.tuning(24)["gpu_region_bits"] == 14
        assert srv.tuning(24)["block_size"] == 1024 and srv.tuning(26)["gpu_region_bits"] == 12
        if compiled:
            for n in {j.qubits for j in jobs}:
                for ops, iters in S.mix_shapes(n):
                    srv.precompile(n, ops, iters)
        states, res = run_collect(srv, jobs, np.complex64, together=True)
    for j in jobs:
        assert res[j.index]["status"] == 0
        ref = _fp64_reference(j)
        f = S.fidelity(ref, states[j.index])
        say("large tile", compiled, j.kind, j.qubits, f)
        assert f >= 1 - 1e-6, (j.kind, j.qubits, f)
        if j.kind == "adder":
            assert res[j.index]["argmax"] == j.expect


def test_2pow15_tiles_at_25_and_26_qubits():
    """With compiled kernels, FP32 requests of 25-26 qubits whose plan on 2^15
    tiles (1024 threads x 32 amplitudes) needs under 1/1.7 of the passes take
    that plan: Grover with 1-3 iterations (4-8 passes instead of 7-15) and the
    adder (1 instead of 3); the Hadamard layer and random layers keep 2^12
    tiles. Results match the FP64 drop-in path, alone and in the mix."""
    rng = np.random.default_rng(45)
    jobs = [S.build_job(0, 0.0, "grover", 26, rng, grover_max_iter=3),
            S.build_job(1, 0.0, "grover", 25, rng, grover_max_iter=3),
            S.build_job(2, 0.0, "adder", 26, rng),
            S.build_job(3, 0.0, "random", 26, rng),
            S.build_job(4, 0.0, "hadamard", 25, rng)]
    with S.Server(np.complex64, max_active=8, hbm_budget_gib=6) as srv:
        assert srv.tuning(26)["gpu_region_bits"] == 12
        interp, ri = run_collect(srv, jobs[:1], np.complex64, together=True)   # no compiled kernels yet
        for n in (25, 26):
            for ops, iters in S.mix_shapes(n):
                srv.precompile(n, ops, iters)
        alone, ra = run_collect(srv, jobs, np.complex64, together=False)
        mixed, rm = run_collect(srv, jobs, np.complex64, together=True)
    say("2^15 tiles passes", [(j.kind, j.qubits, ra[j.index]["passes"]) for j in jobs], "interpreted", ri[0]["passes"])
    for j in jobs[:2]:   # Grover with k iterations: 3 + 4k passes on 2^12 tiles, 2 + 2k on 2^15 tiles
        assert ra[j.index]["passes"] == 2 + 2 * j.params["iterations"], j.params
    assert ri[0]["passes"] == 3 + 4 * jobs[0].params["iterations"]
    assert ra[2]["passes"] == 1 and ra[4]["passes"] == 3
    for j in jobs:
        assert ra[j.index]["status"] == 0 and rm[j.index]["status"] == 0
        assert rm[j.index]["passes"] == ra[j.index]["passes"]
        assert np.abs(alone[j.index] - mixed[j.index]).max() <= 1e-7, (j.kind, j.qubits)
        ref = _fp64_reference(j)
        f = S.fidelity(ref, alone[j.index])
        say("2^15 tiles", j.kind, j.qubits, f)
        assert f >= 1 - 1e-6, (j.kind, j.qubits, f)
        if j.kind == "adder":
            assert ra[j.index]["argmax"] == j.expect
        if j.kind == "grover":
            assert ra[j.index]["argmax"] == j.params["marked"]
    assert S.fidelity(_fp64_reference(jobs[0]), interp[0]) >= 1 - 1e-6


def test_served_fp64_matches_oracle():
    jobs = small_mix(10, seed=8, qmin=10, qmax=15)
    with S.Server(np.complex128, max_active=16, hbm_budget_gib=1) as srv:
        states, res = run_collect(srv, jobs, np.complex128, together=True)
    for j in jobs:
        ref = oracle_run_job(j)
        assert np.abs(states[j.index] - ref).max() < 1e-12, (j.kind, j.qubits)
        pr = np.abs(ref) ** 2
        r = res[j.index]
        assert abs(r["norm2"] - 1) < 1e-12 and abs(r["p_max"] - pr.max()) < 1e-12
        assert pr[r["argmax"]] >= pr.max() - 1e-12


def test_admission_by_footprint_and_errors():
    n = 20                                 # 8 MiB per FP32 state
    jobs = [S.build_job(i, 0.0, "hadamard", n, np.random.default_rng(i)) for i in range(12)]
    with S.Server(np.complex64, max_active=64, hbm_budget_gib=48 / 1024) as srv:
        ids = [srv.submit(j.qubits, j.ops) for j in jobs]
        big = S.build_job(99, 0.0, "hadamard", 23, np.random.default_rng(0))   # 64 MiB > budget
        bid = srv.submit(big.qubits, big.ops)
        srv.wait_all()
        res = {r["id"]: r for r in srv.poll(cap=64)}
        with pytest.raises(N.HGMError):
            srv.submit(n, jobs[0].ops[:0])
    assert res[bid]["status"] == 1        # HGM_E_ARG: can never fit
    admits = sorted(res[i]["t_admit_ms"] for i in ids)
    dones = sorted(res[i]["t_done_ms"] for i in ids)
    for i in ids:
        r = res[i]
        assert r["status"] == 0 and r["argmax"] == 0 and abs(r["p_max"] - 2.0 ** -n) < 1e-9
        assert abs(r["norm2"] - 1) < 1e-5
    # at most 5 states (8 MiB + programs + partials) fit in 48 MiB: the 6th admission waits for a completion
    assert admits[5] >= dones[0]


def test_autotune_picks_a_valid_shape():
    import ctypes as C
    n = 18
    job = S.build_job(0, 0.0, "random", n, np.random.default_rng(1))
    best = N.Tuning()
    ms = C.c_double()
    N.check(N.lib.hgm_autotune(n, N.HGM_C64, S._op_ptr(job.ops), len(job.ops), 1, N.HGM_KERNEL_REG,
                               N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT,
                               C.byref(best), C.byref(ms)))
    say("autotune", n, best.block_size, best.repeat_count, best.gpu_region_bits, best.gpu_coalesced_bits, ms.value)
    assert ms.value > 0
    assert 2 * best.block_size * best.repeat_count == 1 << best.gpu_region_bits
    assert N.lib.hgm_validate_tuning(n, N.HGM_C64, C.byref(best)) == 0


def test_autotune_plan_compiled_family():
    """The autotuner compiles and times plan-compiled candidates with the
    given planner flags; the tile it returns runs the circuit correctly."""
    import ctypes as C
    from tests.test_oracle import oracle_run_gates
    n = 14
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, 4, 7, d))
    ops = d.ops()
    flags = N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT
    best = N.Tuning()
    ms = C.c_double()
    N.check(N.lib.hgm_autotune(n, N.HGM_C128, ops, len(ops), 1, N.HGM_KERNEL_JIT, flags, C.byref(best), C.byref(ms)))
    say("autotune jit", best.block_size, best.repeat_count, best.gpu_region_bits, best.gpu_coalesced_bits, ms.value)
    # FP64 candidates have 128-byte (c=3) or 256-byte (c=4) runs per wave segment
    assert best.kernel == N.HGM_KERNEL_JIT and best.planner == flags and best.gpu_coalesced_bits in (3, 4)
    assert 10 <= best.gpu_region_bits <= 13
    s = hq.State(n)
    s.set_basis(0)
    hq.Plan(n, ops, best).run(s)
    t, cm, cv, mats = hq.ops_to_arrays(ops)
    want = oracle_run_gates(n, t, cm, cv, mats, basis=0)
    assert np.abs(s.download() - want).max() < 1e-12


def test_tuning_override_per_qubit_count():
    jobs = small_mix(6, seed=12, qmin=16, qmax=16)
    with S.Server(np.complex64, max_active=16, hbm_budget_gib=1) as srv:
        srv.set_tuning(16, block_size=512, repeat_count=4, gpu_region_bits=12, gpu_coalesced_bits=4)
        assert srv.tuning(16)["block_size"] == 512
        with pytest.raises(N.HGMError):
            srv.set_tuning(16, block_size=100)
        with pytest.raises(N.HGMError):
            srv.set_tuning(16, kernel=N.HGM_KERNEL_JIT)
        states, res = run_collect(srv, jobs, np.complex64, together=True)
    for j in jobs:
        assert res[j.index]["status"] == 0
        assert S.fidelity(oracle_run_job(j), states[j.index]) >= 1 - 1e-6


def test_tiny_states_beside_others():
    """1-, 2- and 3-qubit requests (2 and 4 amplitudes per thread) run beside larger ones:
    every request succeeds, each matches the oracle, and wait_all reports no error."""
    rng = np.random.default_rng(17)
    jobs = [S.build_job(i, 0.0, kind, n, rng, 4, 3)
            for i, (kind, n) in enumerate([("hadamard", 1), ("random", 1), ("hadamard", 2), ("random", 2),
                                           ("random", 3), ("adder", 3), ("random", 12), ("grover", 15)])]
    say("tiny states:", [(j.kind, j.qubits) for j in jobs])
    with S.Server(np.complex128, max_active=16, hbm_budget_gib=1) as srv:
        states, res = run_collect(srv, jobs, np.complex128, together=True)
        srv.wait_all()
    for j in jobs:
        assert res[j.index]["status"] == 0, (j.kind, j.qubits, res[j.index])
        ref = oracle_run_job(j)
        assert np.abs(states[j.index] - ref).max() < 1e-12, (j.kind, j.qubits)
