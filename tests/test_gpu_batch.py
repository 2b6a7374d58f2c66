"""The batch graph (hgm_batch_*): resident states of uneven qubit counts, each
advanced by its plan (plan-compiled, interpreted register kernel or LDS
kernel), captured as one hipGraph with parallel branches. Every replay gives
each state exactly what running its plan alone gives (same kernels, same
launch shapes, so bit-identical), with eager work on the states between
replays."""
import sys

import numpy as np
import pytest

import simulador_quantico_amd as hq
from simulador_quantico_amd import _native as N
from simulador_quantico_amd import serving as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ops(a):
    return (N.Op * len(a)).from_buffer_copy(a.tobytes())


def _jit_plan(n, ops, dtype, kernel=N.HGM_KERNEL_JIT):
    c = 4 if dtype == np.complex64 else 3
    if kernel == N.HGM_KERNEL_LDS:
        t = N.default_tuning(block_size=256, repeat_count=4, gpu_region_bits=11, gpu_coalesced_bits=c,
                             kernel=kernel, planner=N.HGM_PLAN_GREEDY)
    else:
        t = N.default_tuning(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=c,
                             kernel=kernel, planner=N.HGM_PLAN_GREEDY | 32, ops_block=1 << 30)
    return hq.Plan(n, _ops(ops), t, 1, dtype)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_batch_graph_matches_each_plan_alone(dtype):
    rng = np.random.default_rng(61)
    jobs = [S.build_job(0, 0.0, "grover", 10, rng, grover_max_iter=8),
            S.build_job(1, 0.0, "grover", 13, rng, grover_max_iter=8),
            S.build_job(2, 0.0, "random", 14, rng, random_depth=4),
            S.build_job(3, 0.0, "grover", 16, rng, grover_max_iter=8),
            S.build_job(4, 0.0, "adder", 17, rng),
            S.build_job(5, 0.0, "grover", 18, rng, grover_max_iter=8),
            S.build_job(6, 0.0, "random", 19, rng, random_depth=4),
            S.build_job(7, 0.0, "grover", 20, rng, grover_max_iter=8)]
    # plan-compiled plans, with interpreted register-kernel plans (the random layers) and an LDS plan (the adder)
    kernels = {2: N.HGM_KERNEL_REG, 4: N.HGM_KERNEL_LDS, 6: N.HGM_KERNEL_REG}
    plans = [_jit_plan(j.qubits, j.ops, dtype, kernels.get(j.index, N.HGM_KERNEL_JIT)) for j in jobs]
    assert [p.kernel for p in plans][2:7] == [N.HGM_KERNEL_REG, N.HGM_KERNEL_JIT, N.HGM_KERNEL_LDS,
                                              N.HGM_KERNEL_JIT, N.HGM_KERNEL_REG]
    plans.append(plans[5])                     # one plan shared by two states
    jobs.append(jobs[5])
    say("batch graph:", [(j.kind, j.qubits) for j in jobs], np.dtype(dtype).name)
    refs, solo_ms = [], 0.0
    for j, p in zip(jobs, plans):
        s = hq.State(j.qubits, dtype)
        s.set_basis(j.basis)
        solo_ms += p.run(s)["kernel_ms"]
        refs.append(s.download())
        del s
    states = [hq.State(j.qubits, dtype) for j in jobs]
    batch = hq.Batch(states, plans)
    for replay in range(3):
        for j, s in zip(jobs, states):
            s.set_basis(j.basis)
        st = batch.run()
        assert st["passes"] == sum(len([1 for _, idx in p.passes() if len(idx)]) for p in plans)
        for k, (j, s) in enumerate(zip(jobs, states)):
            out = s.download()
            assert np.array_equal(out, refs[k]), (replay, j.kind, j.qubits)
        say("replay", replay, "batch graph ms", round(st["kernel_ms"], 3), "plans one after another ms",
            round(solo_ms, 3))
    for j, s in zip(jobs, states):
        if j.kind == "adder":
            assert s.argmax()[0] == j.expect
        if j.kind == "grover":
            assert s.argmax()[0] == j.params["marked"]
    del batch


def test_batch_graph_rejects_what_it_cannot_capture():
    ops = S.build_job(0, 0.0, "grover", 14, np.random.default_rng(1), grover_max_iter=2).ops
    p = _jit_plan(14, ops, np.complex64)
    s1, s2 = hq.State(14, np.complex64), hq.State(14, np.complex64)
    with pytest.raises(N.HGMError, match="appears twice"):
        hq.Batch([s1, s1], [p, p])
    with pytest.raises(N.HGMError, match="another qubit count or dtype"):
        hq.Batch([s1, hq.State(15, np.complex64)], [p, p])
    with pytest.raises(N.HGMError, match="another qubit count or dtype"):
        hq.Batch([hq.State(14, np.complex128)], [p])
    reg = _jit_plan(14, ops, np.complex64, N.HGM_KERNEL_REG)
    b = hq.Batch([s1, s2], [p, reg])           # a valid batch still builds and runs after the failures
    s1.set_basis(0)
    s2.set_basis(0)
    b.run()
    assert np.abs(s1.download() - s2.download()).max() < 1e-6   # same circuit, two kernel families


def test_batch_graph_beyond_64_states_and_of_one():
    """70 states share the 64 branches (some branches run two states in
    order); a batch of one state is one branch. Results equal the plans run
    alone."""
    rng = np.random.default_rng(62)
    job = S.build_job(0, 0.0, "grover", 12, rng, grover_max_iter=4)
    jit = _jit_plan(12, job.ops, np.complex64)
    reg = _jit_plan(12, job.ops, np.complex64, N.HGM_KERNEL_REG)
    ref = {}
    for name, p in (("jit", jit), ("reg", reg)):
        s = hq.State(12, np.complex64)
        s.set_basis(0)
        p.run(s)
        ref[name] = s.download()
    plans = [jit if k % 3 else reg for k in range(70)]
    states = [hq.State(12, np.complex64) for _ in plans]
    batch = hq.Batch(states, plans)
    for replay in range(2):
        for s in states:
            s.set_basis(0)
        batch.run()
        for k, s in enumerate(states):
            assert np.array_equal(s.download(), ref["jit" if k % 3 else "reg"]), (replay, k)
    one = hq.Batch([states[0]], [reg])
    states[0].set_basis(0)
    one.run()
    assert np.array_equal(states[0].download(), ref["reg"])
