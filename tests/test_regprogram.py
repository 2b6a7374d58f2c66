"""The register kernel's host-generated programs, executed by a numpy model of
the kernel's semantics (test infrastructure), must reproduce the oracle.

The model follows k_pass_reg: thread t holds A amplitudes; element (t, k)
sits at tile index thr[0][t & 31] | thr[1][t >> 5] | reg[k] of the current
phase map; R_REMAP moves the tile to the next map; R_DENSE / R_ANTI update
register pairs differing in register bit `rbit`; R_PHASE multiplies where the
register, thread and global conditions hold; R_DENSE_PIV* apply rows with an
exact unit entry, out_i = x[p_i] + t_i x[1 - p_i], with the pivots in tval."""
import numpy as np
import pytest

import simulador_quantico_amd as hq
from simulador_quantico_amd import _native as N
from tests.test_oracle import _random_gate_list, oracle_run_gates

R_DENSE, R_ANTI, R_PHASE, R_REMAP, R_SWAP, R_PHASE_REAL, R_DENSE_REAL, R_DENSE_PIV, R_DENSE_PIV_REAL = range(9)


def pdep(v, mask):
    out, bit = 0, 0
    for b in range(64):
        if (mask >> b) & 1:
            if (v >> bit) & 1:
                out |= 1 << b
            bit += 1
    return out


def tile_index(m, B, A):
    """Maps hold LDS slots (swizzled tile indices); the swizzle is XOR-linear and an involution."""
    t = np.arange(B)
    tt = m["thr"][0][t & 31].astype(np.int64) ^ m["thr"][1][(t >> 5) & 31].astype(np.int64)
    s = tt[:, None] ^ m["reg"][:A].astype(np.int64)[None, :]          # [B, A]
    return s ^ ((s >> 4) & 15) ^ ((s >> 8) & 15) ^ ((s >> 12) & 15)


def emulate(n, plan, psi):
    psi = psi.astype(np.complex128).copy()
    r, B, R, c = plan.shape()
    A = 2 * R
    maps = plan.maps()
    allmask = (1 << n) - 1
    for i, (mask, _) in enumerate(plan.passes()):
        ops, first, last, local_pos = plan.program(i)
        to_state = np.array([1 << int(local_pos[b]) for b in range(r)], dtype=np.int64)

        def glob(T):
            g = np.zeros_like(T)
            for b in range(r):
                g |= ((T >> b) & 1) * to_state[b]
            return g

        T_first = tile_index(maps[first], B, A)
        T_last = tile_index(maps[last], B, A)
        G_first, G_last = glob(T_first), glob(T_last)
        t = np.arange(B)
        for tile in range(1 << (n - r)):
            base = pdep(tile, allmask & ~mask)
            a = psi[base | G_first].copy()
            cur = first
            for o in ops:
                kind = int(o["kind"])
                if kind == R_REMAP:
                    buf = np.zeros(1 << r, complex)
                    buf[tile_index(maps[cur], B, A)] = a
                    cur = int(o["rbit"])
                    a = buf[tile_index(maps[cur], B, A)]
                    continue
                if (base & int(o["gmask"])) != int(o["gval"]):
                    continue
                m = o["m"][0::2] + 1j * o["m"][1::2]
                if kind in (R_DENSE_PIV, R_DENSE_PIV_REAL):
                    assert int(o["tmask"]) == 0 and int(o["rmask"]) == 0
                    if kind == R_DENSE_PIV_REAL:
                        assert o["m"][1] == 0 and o["m"][3] == 0
                    p0, p1 = int(o["tval"]) & 1, int(o["tval"]) >> 1
                    bit = int(o["rbit"])
                    for k0 in range(A):
                        if (k0 >> bit) & 1:
                            continue
                        k1 = k0 | (1 << bit)
                        xy = (a[:, k0].copy(), a[:, k1].copy())
                        a[:, k0] = xy[p0] + m[0] * xy[1 - p0]
                        a[:, k1] = xy[p1] + m[1] * xy[1 - p1]
                    continue
                lane = (t & int(o["tmask"])) == int(o["tval"])
                if kind in (R_PHASE, R_PHASE_REAL):
                    for k in range(A):
                        if (k & int(o["rmask"])) == int(o["rval"]):
                            a[lane, k] *= m[0]
                    continue
                bit = int(o["rbit"])
                for k0 in range(A):
                    if (k0 >> bit) & 1 or (k0 & int(o["rmask"])) != int(o["rval"]):
                        continue
                    k1 = k0 | (1 << bit)
                    x, y = a[lane, k0].copy(), a[lane, k1].copy()
                    if kind in (R_DENSE, R_DENSE_REAL):
                        a[lane, k0] = m[0] * x + m[1] * y
                        a[lane, k1] = m[2] * x + m[3] * y
                    else:
                        a[lane, k0] = m[1] * y
                        a[lane, k1] = m[2] * x
            assert cur == last
            psi[base | G_last] = a
    return psi


SHAPES = [(7, 4, 2, 4, 1), (9, 8, 4, 6, 2), (10, 16, 4, 7, 2), (10, 4, 8, 6, 0), (8, 2, 16, 6, 3), (6, 32, 1, 6, 2),
          (11, 32, 8, 9, 4)]


@pytest.mark.parametrize("planner", [N.HGM_PLAN_GREEDY, N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT,
                                     N.HGM_PLAN_GREEDY | N.HGM_PLAN_PIVOT | N.HGM_PLAN_BALANCE,
                                     N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT | N.HGM_PLAN_LOOKAHEAD])
@pytest.mark.parametrize("n,B,R,r,c", SHAPES)
def test_register_program_matches_oracle(n, B, R, r, c, planner):
    rng = np.random.default_rng(n * 7 + r)
    t, cm, cv, mats = _random_gate_list(rng, n, 80)
    ops = hq.make_ops(t, cm, cv, mats)
    tun = N.default_tuning(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=c,
                           kernel=N.HGM_KERNEL_REG, ops_block=40, planner=planner)
    plan = hq.Plan(n, ops, tun)
    assert plan.kernel == N.HGM_KERNEL_REG
    kinds = np.concatenate([plan.program(i)[0]["kind"] for i in range(len(plan))])
    if planner & N.HGM_PLAN_PIVOT:
        assert np.any(kinds == R_DENSE_PIV)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    want = oracle_run_gates(n, t, cm, cv, mats, state=psi0)
    got = emulate(n, plan, psi0)
    assert np.abs(got - want).max() < 1e-12


def test_pivot_program_of_grover_is_real():
    """Fused Grover gates are real, and stay real under row scaling: every
    dense update of the program is R_DENSE_PIV_REAL (4 multiply-adds per pair)."""
    n, k = 9, 5
    d = hq.DGM()
    d.setFunction(hq.GroverIteration(n, 300))
    ops = d.ops()
    tun = N.default_tuning(block_size=16, repeat_count=4, gpu_region_bits=7, gpu_coalesced_bits=2,
                           kernel=N.HGM_KERNEL_REG, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT)
    plan = hq.Plan(n, ops, tun, iterations=k)
    kinds = np.concatenate([plan.program(i)[0]["kind"] for i in range(len(plan))])
    assert np.any(kinds == R_DENSE_PIV_REAL)
    assert not np.any(np.isin(kinds, (R_DENSE, R_DENSE_REAL, R_DENSE_PIV)))
    psi0 = np.full(1 << n, 2 ** (-n / 2), complex)
    t, cm, cv, mats = hq.ops_to_arrays(ops)
    want = oracle_run_gates(n, np.tile(t, k), np.tile(cm, k), np.tile(cv, k), list(mats) * k, state=psi0)
    assert abs(abs(want[300]) ** 2 - np.sin(11 * np.arcsin(2 ** (-n / 2))) ** 2) < 1e-12
    assert np.abs(emulate(n, plan, psi0) - want).max() < 1e-12


def test_register_program_keeps_loads_coalesced():
    """First and last phase keep the c coalesced tile bits on the thread index."""
    n = 30
    d = hq.DGM()
    d.setFunction(hq.RandomLayers(n, 20, 1, d))
    tun = N.default_tuning(block_size=512, repeat_count=8, gpu_region_bits=13, gpu_coalesced_bits=2)
    plan = hq.Plan(n, d.ops(), tun)
    maps = plan.maps()
    phases = 0
    for i in range(len(plan)):
        ops, first, last, _ = plan.program(i)
        for mi in (first, last):
            thr0 = maps[mi]["thr"][0]
            assert thr0[1] == 1 and thr0[2] == 2      # thread bits 0,1 -> tile bits 0,1
        phases += 1 + int(np.sum(ops["kind"] == R_REMAP))
    # 600 dense U3 over 8-10 passes: every phase holds 3 register bits (A = 16), so about 600 / 3 / 2
    # phases are needed at the least; a regression to one phase per gate would exceed this bound
    assert phases <= 100


FLAG_SETS = [N.HGM_PLAN_GREEDY, N.HGM_PLAN_WINDOW, N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE,
             N.HGM_PLAN_GREEDY | N.HGM_PLAN_FOLD_PHASES, N.HGM_PLAN_GREEDY | N.HGM_PLAN_PIVOT,
             N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT | N.HGM_PLAN_BALANCE | N.HGM_PLAN_LOOKAHEAD,
             N.HGM_PLAN_WINDOW | N.HGM_PLAN_FUSE | N.HGM_PLAN_PIVOT]


@pytest.mark.parametrize("seed", range(12))
def test_random_shapes_and_planner_flags(seed):
    """Random circuits (controls inside and outside the tile, diagonal, anti-
    diagonal and dense gates, repeated iterations), random valid tile shapes
    and every combination of planner flags: the register programs reproduce
    the oracle within 1e-12."""
    rng = np.random.default_rng(9000 + seed)
    n = int(rng.integers(5, 11))
    t, cm, cv, mats = _random_gate_list(rng, n, int(rng.integers(20, 70)))
    ops = hq.make_ops(t, cm, cv, mats)
    iterations = int(rng.integers(1, 4))
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    want = oracle_run_gates(n, list(t) * iterations, list(cm) * iterations, list(cv) * iterations,
                            list(mats) * iterations, state=psi0)
    for flags in FLAG_SETS:
        while True:
            r = int(rng.integers(3, min(n, 9) + 1))
            R = 1 << int(rng.integers(0, min(4, r - 1) + 1))
            B = (1 << r) // (2 * R)
            c = int(rng.integers(0, min(r - 1, 4) + 1)) if n > r else int(rng.integers(0, r + 1))
            if B >= 1 and 2 * R <= 32:
                break
        tun = N.default_tuning(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=c,
                               kernel=N.HGM_KERNEL_REG, planner=flags, ops_block=int(rng.integers(3, 40)))
        plan = hq.Plan(n, ops, tun, iterations)
        got = emulate(n, plan, psi0)
        assert np.abs(got - want).max() < 1e-12, (flags, B, R, r, c)
