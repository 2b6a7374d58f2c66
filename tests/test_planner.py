"""Pass planner invariants (host only): every pass is executable inside one
tile, the schedule is a permutation that respects every non-commuting pair,
and replaying the schedule on the oracle reproduces the original order."""
import numpy as np
import pytest

import simulador_quantico_amd as hq
from simulador_quantico_amd import _native as N
from tests.test_oracle import _random_gate_list, oracle_run_gates


def classify(m):
    m = np.asarray(m).reshape(4)
    if m[1] == 0 and m[2] == 0:
        return "diag"
    return "other"


def commute(a, b):
    """The planner's commutation rule: writes vs reads/writes disjoint."""
    def wr(g):
        t, cm, cv, m = g
        tb = 1 << int(t)
        if classify(m) == "diag":
            return 0, tb | int(cm)
        return tb, int(cm)
    wa, ra = wr(a)
    wb, rb = wr(b)
    return not (wa & (wb | rb)) and not (wb & ra)


def check_plan(n, t, cm, cv, mats, tuning, iterations=1):
    ops = hq.make_ops(t, cm, cv, mats)
    plan = hq.Plan(n, ops, tuning, iterations)
    r, B, R, c = plan.shape()
    passes = plan.passes()
    order = []
    for mask, idx in passes:
        assert bin(mask).count("1") == min(r, n)
        assert mask & ((1 << c) - 1) == (1 << c) - 1
        for k in idx:
            i = k % len(t)
            if classify(mats[i]) != "diag":
                assert (mask >> int(t[i])) & 1, "non-diagonal gate scheduled with a non-local target"
        assert 0 < len(idx) <= tuning.ops_block
        order += idx
    N_ = len(t) * iterations
    assert sorted(order) == list(range(N_))
    pos = {k: p for p, k in enumerate(order)}
    for a in range(N_):
        for b in range(a + 1, N_):
            if pos[a] > pos[b]:
                ga = (t[a % len(t)], cm[a % len(t)], cv[a % len(t)], mats[a % len(t)])
                gb = (t[b % len(t)], cm[b % len(t)], cv[b % len(t)], mats[b % len(t)])
                assert commute(ga, gb), (a, b)
    return order, passes


@pytest.mark.parametrize("planner", [N.HGM_PLAN_GREEDY, N.HGM_PLAN_WINDOW])
@pytest.mark.parametrize("n,r,c,ops_block", [(6, 3, 1, 300), (9, 5, 2, 7), (12, 6, 3, 300), (10, 4, 0, 5),
                                             (5, 8, 4, 300)])
def test_plan_invariants_and_replay(planner, n, r, c, ops_block):
    rng = np.random.default_rng(n * 100 + r)
    t, cm, cv, mats = _random_gate_list(rng, n, 60)
    B = max(1, min(64, 1 << (r - 1)))
    tun = N.default_tuning(block_size=B, repeat_count=(1 << (r - 1)) // B, gpu_region_bits=r,
                           gpu_coalesced_bits=c, ops_block=ops_block, planner=planner)
    order, _ = check_plan(n, t, cm, cv, mats, tun, iterations=2)
    T, CM, CV, M = (np.array(list(x) * 2, dtype=object) for x in (t, cm, cv, mats))
    psi0 = np.zeros(1 << n, complex)
    psi0[5 % (1 << n)] = 1
    want = oracle_run_gates(n, list(T), list(CM), list(CV), list(M), state=psi0)
    got = oracle_run_gates(n, [T[k] for k in order], [CM[k] for k in order], [CV[k] for k in order],
                           [M[k] for k in order], state=psi0)
    assert np.abs(got - want).max() < 1e-12


def test_window_planner_follows_reference_contract():
    """In-order, one contiguous window of r-c bits above the c low bits."""
    n, r, c = 12, 6, 2
    t = [0, 1, 2, 5, 7, 8, 11, 3, 10]
    mats = [[2 ** -0.5, 2 ** -0.5, 2 ** -0.5, -2 ** -0.5]] * len(t)
    tun = N.default_tuning(block_size=16, repeat_count=2, gpu_region_bits=r, gpu_coalesced_bits=c,
                           planner=N.HGM_PLAN_WINDOW)
    order, passes = check_plan(n, t, [0] * len(t), [0] * len(t), mats, tun)
    assert order == list(range(len(t)))
    for mask, _ in passes:
        hi = mask >> c
        assert hi & (hi + (hi & -hi)) == 0 or bin(hi).count("1") == r - c   # contiguous window


def test_diagonal_gates_never_force_a_pass():
    n = 20
    d = hq.DGM()
    steps = hq.RandomLayers(n, 1, 3, d)
    d.setFunction(steps)
    ops = d.ops()
    tun = N.default_tuning(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=4)
    plan = hq.Plan(n, ops, tun)
    # 20 dense gates need ceil((20-4)/(12-4)) = 2 passes; the 10 CZ ride along
    assert len(plan) == 2


def test_random_layer_pass_counts():
    """Greedy (commutation lookahead) never needs more passes than the window planner."""
    n, depth = 30, 20
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, depth, 1, d))
    ops = d.ops()
    counts = {}
    for name, pl in [("greedy", N.HGM_PLAN_GREEDY), ("window", N.HGM_PLAN_WINDOW)]:
        tun = N.default_tuning(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=4, planner=pl)
        counts[name] = len(hq.Plan(n, ops, tun))
    assert counts["greedy"] <= counts["window"]
    assert counts["greedy"] <= 3 * depth


def test_grover_ops_plan_quickly():
    n = 22
    d = hq.DGM()
    d.setFunction([hq.Hadamard(n, 0, n)])
    d.setFunction(hq.GroverIteration(n, 123456), 50, reset=False)
    ops = d.ops()
    # 18 high bits in two groups of r-c = 9: one pass per n-controlled Z, two per iteration
    tun = N.default_tuning(block_size=256, repeat_count=16, gpu_region_bits=13, gpu_coalesced_bits=4)
    assert len(hq.Plan(n, ops, tun)) <= 2 * 50 + 2
    # three groups of 8 need two passes per n-controlled Z
    tun = N.default_tuning(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=4)
    assert len(hq.Plan(n, ops, tun)) <= 4 * 50 + 3


def test_jit_source_compiles_and_dedupes():
    """The plan-time generated pass kernels compile for gfx950 (hipRTC, no GPU
    needed); repeated iterations with identical passes share one kernel."""
    n = 16
    d = hq.DGM()
    d.setFunction(hq.GroverIteration(n, 12345))
    tun = N.default_tuning(block_size=64, repeat_count=8, gpu_region_bits=10, gpu_coalesced_bits=3,
                           kernel=N.HGM_KERNEL_JIT)
    plan = hq.Plan(n, d.ops(), tun, iterations=20)
    src = plan.jit_source()
    kernels = src.count("__global__")
    assert plan.kernel == N.HGM_KERNEL_JIT and 0 < kernels < len(plan)
    assert "v_pk" not in src and "hgm_pass_0" in src


def test_jit_source_of_pivoted_plan_drops_a_third_of_the_multiplies():
    """Folded dense updates take 3 products per output component (12 per pair);
    with rows scaled to an exact 1 (HGM_PLAN_PIVOT) the generated code has 2
    (8 per pair). Both plans compile for gfx950."""
    n = 14
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, 6, 11, d))
    ops = d.ops()
    counts = {}
    for name, flag in (("fold", N.HGM_PLAN_FOLD_PHASES), ("pivot", N.HGM_PLAN_PIVOT)):
        tun = N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9, gpu_coalesced_bits=3,
                               kernel=N.HGM_KERNEL_JIT, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | flag)
        plan = hq.Plan(n, ops, tun)
        counts[name] = plan.jit_source().count(" * ")
    assert counts["pivot"] < 0.72 * counts["fold"]


def test_jit_dispatch_limit():
    """A plan-compiled pass launches one workgroup per tile; a launch holds at
    most 2^32 - 1 work-items, so 33 qubits need repeat_count >= 2."""
    n = 33
    ops = hq.make_ops([0], [0], [0], [np.array([[0, 1], [1, 0]])])
    tun = N.default_tuning(block_size=1024, repeat_count=1, gpu_region_bits=11, gpu_coalesced_bits=3,
                           kernel=N.HGM_KERNEL_JIT)
    with pytest.raises(N.HGMError) as e:
        hq.Plan(n, ops, tun, 1, np.complex64)
    assert "dispatch" in str(e.value)


def test_balance_keeps_passes_and_gates():
    """HGM_PLAN_BALANCE moves gates between neighbouring passes only: the pass
    count and the multiset of gates stay, the heaviest pass gets no heavier,
    and every gate's target is local in the pass it lands in."""
    n = 24
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, 12, 3, d))
    ops = d.ops()
    base = N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT
    plans = {}
    for flags in (base, base | N.HGM_PLAN_BALANCE):
        tun = N.default_tuning(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=3,
                               kernel=N.HGM_KERNEL_REG, planner=flags)
        plans[flags] = hq.Plan(n, ops, tun)
    a, b = plans[base], plans[base | N.HGM_PLAN_BALANCE]
    assert len(a) == len(b)
    ga = sorted(i for _, idx in a.passes() for i in idx)
    gb = sorted(i for _, idx in b.passes() for i in idx)
    assert ga == gb
    from tests.test_fold import fuse, pivot
    t, cm, cv, mats = hq.ops_to_arrays(pivot(n, fuse(n, ops)))   # the list the planner schedules
    assert ga == list(range(len(t)))
    for mask, idx in b.passes():
        for i in idx:
            if not (mats[i][1] == 0 and mats[i][2] == 0):
                assert (mask >> int(t[i])) & 1


def test_lookahead_saves_passes_on_random_layers():
    """Choosing each pass's local set by what it and the next pass admit
    (HGM_PLAN_LOOKAHEAD) never needs more passes than plain greedy here, and
    needs fewer on the 30-qubit depth-20 random-layer circuit."""
    n = 30
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, 20, 1, d))
    ops = d.ops()
    base = N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT
    counts = {}
    for B, R, r in ((256, 8, 12), (512, 8, 13)):
        for flags in (base, base | N.HGM_PLAN_LOOKAHEAD):
            tun = N.default_tuning(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=3,
                                   kernel=N.HGM_KERNEL_REG, planner=flags)
            counts[(r, flags)] = len(hq.Plan(n, ops, tun))
        assert counts[(r, base | N.HGM_PLAN_LOOKAHEAD)] < counts[(r, base)], counts
    assert counts[(12, base | N.HGM_PLAN_LOOKAHEAD)] <= 11
