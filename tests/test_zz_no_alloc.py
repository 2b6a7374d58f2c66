"""hgm_state_apply_noalloc (no device allocation; plan tables in the library's
per-GPU workspace) against hgm_state_apply on the same states and circuits,
through the C ABI: two calls in a row, the second on a smaller state with
fewer gates, on the register and the LDS kernels. Same plans, same kernels:
the results are bit-identical."""
import ctypes as C
import sys

import numpy as np
import pytest

import simulador_quantico_amd as hq
from simulador_quantico_amd import _native as N

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_ops(n, depth, seed):
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, depth, seed, d))
    return d.ops()


def run(fn, n, ops, tun, dtype):
    s = hq.State(n, dtype)
    s.set_basis(0)
    st = N.Stats()
    N.check(fn(s._h, ops, len(ops), C.byref(tun), 1, C.byref(st)))
    return s.download(), st


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("kernel", [N.HGM_KERNEL_REG, N.HGM_KERNEL_LDS])
def test_noalloc_matches_apply_on_two_calls(kernel, dtype):
    for n, depth, seed in ((16, 8, 3), (12, 3, 4)):   # the second call is smaller
        ops = random_ops(n, depth, seed)
        tun = N.default_tuning(block_size=128, repeat_count=8, gpu_region_bits=11, gpu_coalesced_bits=3,
                               kernel=kernel, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT)
        want, st_a = run(N.lib.hgm_state_apply, n, ops, tun, dtype)
        got, st_n = run(N.lib.hgm_state_apply_noalloc, n, ops, tun, dtype)
        say("noalloc", kernel, dtype.__name__, n, "qubits", len(ops), "gates", st_n.passes, "passes")
        assert st_n.passes == st_a.passes
        assert np.array_equal(got, want), (n, kernel)


def test_noalloc_rejects_what_it_cannot_run_without_allocating():
    n = 12
    ops = random_ops(n, 2, 5)
    s = hq.State(n)
    s.set_basis(0)
    st = N.Stats()
    jit = N.default_tuning(block_size=128, repeat_count=8, gpu_region_bits=11, gpu_coalesced_bits=3,
                           kernel=N.HGM_KERNEL_JIT)
    assert N.lib.hgm_state_apply_noalloc(s._h, ops, len(ops), C.byref(jit), 1, C.byref(st)) == N.HGM_E_TUNING
    # one pass per gate: 6000 passes of 2 KiB pass tables exceed the 8 MiB workspace
    many = random_ops(n, 400, 6)
    assert len(many) > 6000
    small = N.default_tuning(block_size=128, repeat_count=8, gpu_region_bits=11, gpu_coalesced_bits=3,
                             kernel=N.HGM_KERNEL_REG, ops_block=1)
    assert N.lib.hgm_state_apply_noalloc(s._h, many, len(many), C.byref(small), 1, C.byref(st)) == N.HGM_E_ARG
    # the state is untouched by the rejected calls
    psi = s.download()
    assert psi[0] == 1 and np.count_nonzero(psi) == 1
