"""Phase folding (HGM_PLAN_FOLD_PHASES) keeps the unitary: the folded gate list
run on the oracle equals the original within FP64 rounding, every uncontrolled
dense gate of the folded list has a real non-negative first column, and the
list grows only by the carried diagonal factors: one before a controlled
non-diagonal gate on a qubit that carries one, and one per qubit at the end.
Fusion (HGM_PLAN_FUSE) and row scaling (HGM_PLAN_PIVOT) keep the unitary too."""
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


def pivot(n, ops):
    cnt = C.c_int64()
    N.check(N.lib.hgm_pivot_rows(n, ops, len(ops), None, 0, C.byref(cnt)))
    out = (N.Op * max(1, cnt.value))()
    N.check(N.lib.hgm_pivot_rows(n, ops, len(ops), out, cnt.value, C.byref(cnt)))
    return out if cnt.value else (N.Op * 0)()


def apply_list(n, ops, psi, dtype):
    """The gate list applied gate by gate in numpy at `dtype` (a check of the
    list's numerics, independent of the oracle)."""
    psi = psi.astype(dtype).copy()
    idx = np.arange(1 << n)
    for i in range(len(ops)):
        o = ops[i]
        m = (np.array(o.m[0::2]) + 1j * np.array(o.m[1::2])).astype(dtype)
        b = 1 << o.target_bit
        sel = ((idx & b) == 0) & ((idx & o.control_mask) == o.control_value)
        i0 = idx[sel]
        x, y = psi[i0].copy(), psi[i0 | b].copy()
        psi[i0] = m[0] * x + m[1] * y
        psi[i0 | b] = m[2] * x + m[3] * y
    return psi


@pytest.mark.parametrize("seed", range(6))
def test_pivoted_list_is_the_same_unitary(seed):
    """Row scaling (HGM_PLAN_PIVOT) keeps the unitary, global phase included;
    every uncontrolled dense gate it emits has an exact 1 in each row and an
    other entry of magnitude <= 1."""
    n = 7
    rng = np.random.default_rng(200 + seed)
    t, cm, cv, mats = _random_gate_list(rng, n, 120)
    for k in range(0, len(cm), 2):
        cm[k] = cv[k] = 0
    ops = hq.make_ops(t, cm, cv, mats)
    f = pivot(n, ops)
    assert len(f) <= 2 * len(ops) + n + 1
    dense = 0
    for i in range(len(f)):
        m = np.array(f[i].m[0::2]) + 1j * np.array(f[i].m[1::2])
        if f[i].control_mask or (m[1] == 0 and m[2] == 0) or (m[0] == 0 and m[3] == 0):
            continue
        dense += 1
        for row in (m[:2], m[2:]):
            assert 1 in (row[0], row[1])
            assert min(abs(row[0]), abs(row[1])) <= 1 + 1e-15
    assert dense >= 10
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    want = run(n, ops, psi0)
    assert np.abs(run(n, f, psi0) - want).max() < 1e-13
    ff = pivot(n, fuse(n, ops))           # fuse then scale rows, as the planner does
    assert np.abs(run(n, ff, psi0) - want).max() < 1e-13


def test_pivot_turns_x_into_a_swap_and_absorbs_diagonals():
    """An uncontrolled anti-diagonal gate becomes the exact X (a register swap),
    an uncontrolled diagonal gate disappears into the carried factors, and the
    carried factors come out at the end as diagonal gates."""
    n = 3
    y = np.array([[0, -1j], [1j, 0]])
    s = np.diag([1, 1j])
    ops = hq.make_ops([0, 1, 2], [0, 0, 0], [0, 0, 0], [y, s, y @ s])
    f = pivot(n, ops)
    mats = [np.array(f[i].m[0::2]) + 1j * np.array(f[i].m[1::2]) for i in range(len(f))]
    assert np.array_equal(mats[0], [0, 1, 1, 0]) and f[0].target_bit == 0
    assert all(m[1] == 0 and m[2] == 0 for m in mats[2:])       # the rest are diagonal
    psi0 = np.arange(1 << n) + 1j
    assert np.abs(run(n, f, psi0) - run(n, ops, psi0)).max() < 1e-13


def test_pivot_rescales_deep_circuits_in_fp32():
    """2000 Hadamard gates: every one scales the stored amplitudes by sqrt(2),
    so the list must carry uniform rescales; applied in complex64 it stays
    finite and equals the original within FP32 rounding."""
    n = 5
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    rng = np.random.default_rng(4)
    targets = list(rng.integers(0, n, 2000))
    ops = hq.make_ops(targets, [0] * 2000, [0] * 2000, [h] * 2000)
    f = pivot(n, ops)
    scales = [i for i in range(len(f)) if f[i].m[0] == f[i].m[6] and f[i].m[2] == f[i].m[4] == 0 and
              f[i].m[1] == 0 and f[i].m[0] != 1]
    assert len(scales) >= 2000 * 0.5 / 24 - 1
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    got = apply_list(n, f, psi0, np.complex64)
    assert np.all(np.isfinite(got))
    want = apply_list(n, ops, psi0, np.complex128)
    assert np.abs(got - want).max() < 1e-4


def test_pivot_flag_plans_with_fusion():
    n = 10
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, 4, 5, d))
    ops = d.ops()
    flags = N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT
    tun = N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9, gpu_coalesced_bits=2,
                           planner=flags, kernel=N.HGM_KERNEL_REG)
    assert len(hq.Plan(n, ops, tun, iterations=2)) >= 1
    bad = hq.make_ops([n], [0], [0], [np.eye(2)])
    with pytest.raises(N.HGMError):
        pivot(n, bad)


def rotate(n, ops):
    cnt = C.c_int64()
    N.check(N.lib.hgm_rotate_rows(n, ops, len(ops), None, 0, C.byref(cnt)))
    out = (N.Op * max(1, cnt.value))()
    N.check(N.lib.hgm_rotate_rows(n, ops, len(ops), out, cnt.value, C.byref(cnt)))
    return out if cnt.value else (N.Op * 0)()


def _mat(o):
    return np.array(o.m[0::2]) + 1j * np.array(o.m[1::2])


@pytest.mark.parametrize("seed", range(6))
def test_rotated_list_is_the_same_unitary(seed):
    """The rotation split (HGM_PLAN_ROTATE) keeps the unitary, global phase included.
    Every uncontrolled dense gate it emits for a unitary is real with an exact 1 in
    each row; every phase gate tagged to ride with a rotation is diag(1, y) and comes
    right before a real rotation on the same qubit."""
    n = 7
    rng = np.random.default_rng(300 + seed)
    t, cm, cv, mats = _random_gate_list(rng, n, 120)
    for k in range(0, len(cm), 2):
        cm[k] = cv[k] = 0
    ops = hq.make_ops(t, cm, cv, mats)
    f = rotate(n, ops)
    assert len(f) <= 2 * len(ops) + n + 1
    real_rot = tagged = 0
    for i in range(len(f)):
        m = _mat(f[i])
        if f[i].reserved:
            tagged += 1
            assert f[i].control_mask == 0 and m[0] == 1 and m[1] == 0 and m[2] == 0 and abs(abs(m[3]) - 1) < 1e-15
            nxt = _mat(f[i + 1])
            assert f[i + 1].target_bit == f[i].target_bit and f[i + 1].control_mask == 0
            assert np.all(nxt.imag == 0)
            continue
        if f[i].control_mask or (m[1] == 0 and m[2] == 0) or (m[0] == 0 and m[3] == 0):
            continue
        if np.all(m.imag == 0):
            real_rot += 1
            for row in (m[:2], m[2:]):
                assert 1 in (row[0], row[1]) and min(abs(row[0]), abs(row[1])) <= 1 + 1e-15
    assert real_rot >= 10 and tagged >= 10
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    want = run(n, ops, psi0)
    assert np.abs(run(n, f, psi0) - want).max() < 1e-13
    ff = rotate(n, fuse(n, ops))           # fuse then split, as the planner does
    assert np.abs(run(n, ff, psi0) - want).max() < 1e-13


def test_rotation_split_of_u3_layers():
    """Random U3 layers with CZ brickwork: every U3 becomes a tagged phase and a real
    rotation; the list is the same unitary."""
    n = 8
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, 6, 11, d))
    ops = d.ops()
    f = rotate(n, ops)
    dense_in = sum(1 for i in range(len(ops)) if not ops[i].control_mask)
    real_out = sum(1 for i in range(len(f)) if not f[i].control_mask and not f[i].reserved and
                   _mat(f[i])[1] != 0 and np.all(_mat(f[i]).imag == 0))
    assert real_out == dense_in
    psi0 = np.zeros(1 << n, complex)
    psi0[0] = 1
    assert np.abs(run(n, f, psi0) - run(n, ops, psi0)).max() < 1e-13


def test_rotation_split_leaves_non_unitary_gates_in_the_pivot_form():
    """A dense gate that is no unitary times a scale has no real middle factor: it keeps
    the row-scaled complex form, and the list is still the same linear map."""
    n = 3
    rng = np.random.default_rng(8)
    mats = [rng.normal(size=4) + 1j * rng.normal(size=4) for _ in range(6)]
    ops = hq.make_ops([0, 1, 2, 0, 1, 2], [0] * 6, [0] * 6, mats)
    f = rotate(n, ops)
    assert not any(f[i].reserved for i in range(len(f)))
    psi0 = np.arange(1 << n) + 0.5j
    want = run(n, ops, psi0)
    assert np.abs(run(n, f, psi0) - want).max() < 1e-12 * np.abs(want).max()


def test_rotation_split_rescales_deep_circuits_in_fp32():
    """2000 Hadamard gates through the rotation split stay finite in complex64."""
    n = 5
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    rng = np.random.default_rng(5)
    targets = list(rng.integers(0, n, 2000))
    ops = hq.make_ops(targets, [0] * 2000, [0] * 2000, [h] * 2000)
    f = rotate(n, ops)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    got = apply_list(n, f, psi0, np.complex64)
    assert np.all(np.isfinite(got))
    assert np.abs(got - apply_list(n, ops, psi0, np.complex128)).max() < 1e-4


def test_rotate_flag_keeps_phases_with_their_rotation():
    """With HGM_PLAN_ROTATE every tagged phase is planned into the pass of the rotation
    it precedes, with its target tile-local there."""
    n = 12
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, 6, 3, d))
    ops = d.ops()
    flags = N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_ROTATE
    tun = N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9, gpu_coalesced_bits=2,
                           planner=flags, kernel=N.HGM_KERNEL_REG)
    plan = hq.Plan(n, ops, tun)
    f = rotate(n, fuse(n, ops))
    where = {}
    for p, (local, idx) in enumerate(plan.passes()):
        for g in idx:
            where[g] = (p, local)
    assert len(where) == len(f)
    for g in range(len(f)):
        if f[g].reserved:
            p, local = where[g]
            assert (local >> f[g].target_bit) & 1
            assert where[g + 1][0] == p
