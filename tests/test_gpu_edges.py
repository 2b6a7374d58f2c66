"""Edge cases on the GPU, for every kernel family: states of 1-3 qubits
(a workgroup of one or two threads), an empty gate list, gates and controls on
the most significant qubit, and the fused/folded planner on each family."""
import sys

import numpy as np
import pytest

import simulador_quantico_amd as hq
from simulador_quantico_amd import _native as N
from tests.test_oracle import _random_gate_list, oracle_run_gates

pytestmark = pytest.mark.gpu

FAMILIES = [N.HGM_KERNEL_LDS, N.HGM_KERNEL_REG, N.HGM_KERNEL_JIT]
PLANNERS = [N.HGM_PLAN_GREEDY, N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_FOLD_PHASES,
            N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [1, 2, 3])
def test_tiny_states_every_family(n):
    rng = np.random.default_rng(70 + n)
    t, cm, cv, mats = _random_gate_list(rng, n, 30)
    ops = hq.make_ops(t, cm, cv, mats)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    want = oracle_run_gates(n, t, cm, cv, mats, state=psi0)
    for kern in FAMILIES:
        for planner in PLANNERS:
            say("tiny", n, kern, planner)
            tun = N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9, gpu_coalesced_bits=2,
                                   kernel=kern, planner=planner)
            for dtype, tol in ((np.complex128, 1e-12), (np.complex64, 2e-5)):
                s = hq.State(n, dtype)
                s.upload(psi0)
                hq.Plan(n, ops, tun, 1, dtype).run(s)
                assert np.abs(s.download() - want).max() < tol, (n, kern, planner, dtype)


def test_empty_gate_list_leaves_the_state():
    n = 12
    rng = np.random.default_rng(5)
    psi0 = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex128)
    empty = (N.Op * 0)()
    for kern in FAMILIES:
        s = hq.State(n)
        s.upload(psi0)
        st = hq.Plan(n, empty, N.default_tuning(block_size=64, repeat_count=4, gpu_region_bits=9,
                                                gpu_coalesced_bits=2, kernel=kern)).run(s)
        assert st["passes"] == 0
        assert np.array_equal(s.download(), psi0)


def test_msb_targets_and_controls_every_family():
    """Gates on index bit n-1 (qubit 0) and controls on it, with a state larger than one tile."""
    n = 15
    rng = np.random.default_rng(15)
    top = n - 1
    t, cm, cv, mats = [], [], [], []
    for k in range(40):
        z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        u, _ = np.linalg.qr(z)
        tgt = top if k % 2 == 0 else int(rng.integers(0, top))
        ctrl = 0 if k % 3 == 0 else (1 << top if tgt != top else 1 << int(rng.integers(0, top)))
        t.append(tgt)
        cm.append(ctrl)
        cv.append(ctrl if k % 4 else 0)
        mats.append(u.reshape(4))
    ops = hq.make_ops(t, cm, cv, mats)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    want = oracle_run_gates(n, t, cm, cv, mats, state=psi0)
    for kern in FAMILIES:
        for planner in PLANNERS:
            tun = N.default_tuning(block_size=128, repeat_count=4, gpu_region_bits=10, gpu_coalesced_bits=3,
                                   kernel=kern, planner=planner)
            s = hq.State(n)
            s.upload(psi0)
            hq.Plan(n, ops, tun).run(s)
            say("msb", kern, planner)
            assert np.abs(s.download() - want).max() < 1e-12, (kern, planner)


def test_state_of_2_pow_32_amplitudes():
    """32 qubits in FP32 (32 GiB): the clear, basis and readout launches stay
    within a dispatch's 2^32 - 1 work-items (grid-stride loops)."""
    say("state 32q fp32")
    s = hq.State(32, np.complex64)
    s.set_basis((1 << 32) - 5)
    idx, p = s.argmax()
    assert idx == (1 << 32) - 5 and p == 1.0
    assert abs(s.norm2() - 1) < 1e-6
    del s


def test_allocations_beyond_hbm_fail_cleanly():
    """A state or a serving arena larger than the device's memory is an error, and the library keeps working."""
    with pytest.raises(N.HGMError) as e:
        hq.State(36)                      # 1 TiB of FP64 amplitudes
    assert e.value.code == 4              # HGM_E_NOMEM
    from simulador_quantico_amd import serving as S
    with pytest.raises(N.HGMError):
        S.Server(np.complex64, hbm_budget_gib=1024)
    s = hq.State(10)                      # and the device is still usable
    s.set_basis(3)
    assert s.argmax()[0] == 3
