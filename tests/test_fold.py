"""Phase folding (HGM_PLAN_FOLD_PHASES) keeps the unitary: the folded gate list
run on the oracle equals the original within FP64 rounding, every uncontrolled
dense gate of the folded list has a real non-negative first column, and the
list grows only by the carried diagonal factors: one before a controlled
non-diagonal gate on a qubit that carries one, and one per qubit at the end."""
import ctypes as C

import numpy as np
import pytest

import simulador_quantico_amd as hq
from simulador_quantico_amd import _native as N
from tests.test_oracle import _random_gate_list, oracle_run_gates


def fold(n, ops):
    cnt = C.c_int64()
    N.check(N.lib.hgm_fold_phases(n, ops, len(ops), None, 0, C.byref(cnt)))
    out = (N.Op * max(1, cnt.value))()
    N.check(N.lib.hgm_fold_phases(n, ops, len(ops), out, cnt.value, C.byref(cnt)))
    return out if cnt.value else (N.Op * 0)()


def run(n, ops, psi0):
    t, cm, cv, mats = hq.ops_to_arrays(ops)
    return oracle_run_gates(n, t, cm, cv, mats, state=psi0)


@pytest.mark.parametrize("seed", range(6))
def test_folded_list_is_the_same_unitary(seed):
    n = 7
    rng = np.random.default_rng(seed)
    t, cm, cv, mats = _random_gate_list(rng, n, 90)
    ops = hq.make_ops(t, cm, cv, mats)
    f = fold(n, ops)
    controlled = sum(1 for k in range(len(mats)) if cm[k] and not (mats[k][1] == 0 and mats[k][2] == 0))
    assert len(ops) <= len(f) <= len(ops) + controlled + n
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    assert np.abs(run(n, f, psi0) - run(n, ops, psi0)).max() < 1e-13


def test_random_layers_fold_to_real_first_columns():
    n = 8
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, 6, 3, d))
    ops = d.ops()
    f = fold(n, ops)
    dense = 0
    for i in range(len(f)):
        m = list(f[i].m)
        if f[i].control_mask == 0 and not (m[2] == m[3] == m[4] == m[5] == 0):
            dense += 1
            assert m[1] == 0.0 and m[5] == 0.0 and m[0] >= 0 and m[4] >= 0
    assert dense == 6 * n
    psi0 = np.zeros(1 << n, complex)
    psi0[0] = 1
    assert np.abs(run(n, f, psi0) - run(n, ops, psi0)).max() < 1e-13


def test_real_gates_fold_to_themselves():
    """H, X and controlled gates carry no phase: the list comes back unchanged."""
    n = 6
    d = hq.DGM()
    d.setFunction(hq.GroverIteration(n, 37))
    ops = d.ops()
    f = fold(n, ops)
    assert bytes(f) == bytes(ops)


def test_fold_flag_plans_and_rejects_bad_targets():
    n = 10
    rng = np.random.default_rng(9)
    t, cm, cv, mats = _random_gate_list(rng, n, 40)
    ops = hq.make_ops(t, cm, cv, mats)
    tun = N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9, gpu_coalesced_bits=2,
                           planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FOLD_PHASES, kernel=N.HGM_KERNEL_REG)
    plan = hq.Plan(n, ops, tun)
    assert len(plan) >= 1
    bad = hq.make_ops([n], [0], [0], [np.eye(2)])
    with pytest.raises(N.HGMError):
        hq.Plan(n, bad, tun)
    with pytest.raises(N.HGMError):
        fold(n, bad)


def fuse(n, ops):
    cnt = C.c_int64()
    N.check(N.lib.hgm_fuse_gates(n, ops, len(ops), None, 0, C.byref(cnt)))
    out = (N.Op * max(1, cnt.value))()
    N.check(N.lib.hgm_fuse_gates(n, ops, len(ops), out, cnt.value, C.byref(cnt)))
    return out if cnt.value else (N.Op * 0)()


@pytest.mark.parametrize("seed", range(6))
def test_fused_list_is_the_same_unitary(seed):
    """Random lists with many uncontrolled gates (fusable runs) and controlled ones."""
    n = 6
    rng = np.random.default_rng(100 + seed)
    t, cm, cv, mats = _random_gate_list(rng, n, 120)
    for k in range(0, len(cm), 2):       # half the gates uncontrolled
        cm[k] = cv[k] = 0
    ops = hq.make_ops(t, cm, cv, mats)
    f = fuse(n, ops)
    assert len(f) < len(ops)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    assert np.abs(run(n, f, psi0) - run(n, ops, psi0)).max() < 1e-13
    ff = fold(n, f)                      # fuse then fold, as the planner does
    assert np.abs(run(n, ff, psi0) - run(n, ops, psi0)).max() < 1e-13


def test_grover_iterations_fuse_to_one_gate_per_qubit_between_reflections():
    """Between two n-controlled Z gates every qubit keeps one fused real gate; X X pairs cancel."""
    n = 7
    marked = 0b0101100
    d = hq.DGM()
    d.setFunction(hq.GroverIteration(n, marked), 3)
    ops = d.ops()
    f = fuse(n, ops)
    mcz = sum(1 for i in range(len(ops)) if ops[i].control_mask)
    assert sum(1 for i in range(len(f)) if f[i].control_mask) == mcz == 6
    assert len(f) <= mcz + (mcz + 1) * n
    assert all(f[i].m[1] == f[i].m[3] == f[i].m[5] == f[i].m[7] == 0 for i in range(len(f)))
    psi0 = np.zeros(1 << n, complex)
    psi0[3] = 1
    assert np.abs(run(n, f, psi0) - run(n, ops, psi0)).max() < 1e-13
    x = hq.make_ops([2, 2], [0, 0], [0, 0], [np.array([[0, 1], [1, 0]])] * 2)
    assert len(fuse(n, x)) == 0


def test_fuse_flag_expands_iterations():
    """With HGM_PLAN_FUSE the planner fuses across iteration boundaries: fewer gates than n_ops * iterations."""
    n = 10
    d = hq.DGM()
    d.setFunction(hq.GroverIteration(n, 77))
    ops = d.ops()
    flags = N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_FOLD_PHASES
    tun = N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9, gpu_coalesced_bits=2,
                           planner=flags, kernel=N.HGM_KERNEL_REG)
    plan = hq.Plan(n, ops, tun, iterations=5)
    total = sum(len(idx) for _, idx in plan.passes())
    assert total < len(ops) * 5
