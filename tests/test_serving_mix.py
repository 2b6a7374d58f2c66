"""The serving mix (host side): seeded, shaped as specified, and every circuit
type computes what it should (checked on the oracle)."""
import numpy as np
import pytest

from simulador_quantico_amd import serving as S
from tests.test_oracle import grover_closed_form, oracle_run_gates


def oracle_run_job(job, precision=64):
    ops = job.ops
    mats = [np.array(m[0::2] + 1j * m[1::2]) for m in ops["m"]]
    return oracle_run_gates(job.qubits, list(ops["target_bit"]), [int(x) for x in ops["control_mask"]],
                            [int(x) for x in ops["control_value"]], mats, basis=job.basis, precision=precision)


def test_mix_is_seeded_and_shaped():
    a = S.make_mix(300, rate=500.0, seed=3)
    b = S.make_mix(300, rate=500.0, seed=3)
    assert [(j.kind, j.qubits, j.arrival_s, j.basis) for j in a] == [(j.kind, j.qubits, j.arrival_s, j.basis) for j in b]
    assert all(np.array_equal(x.ops, y.ops) for x, y in zip(a, b))
    arr = np.array([j.arrival_s for j in a])
    assert arr[0] == 0 and np.all(np.diff(arr) >= 0)
    assert abs(np.diff(arr).mean() - 1 / 500.0) < 0.2 / 500.0
    assert {j.kind for j in a} == set(S.KINDS)
    q = np.array([j.qubits for j in a])
    assert q.min() >= 10 and q.max() <= 26 and len(set(q)) > 12
    g = [j for j in a if j.kind == "grover"]
    assert all(1 <= j.params["iterations"] <= 64 for j in g)


def test_mix_circuits_against_oracle():
    jobs = S.make_mix(32, rate=100.0, seed=11, qmin=5, qmax=9, grover_max_iter=6)
    seen = set()
    for j in jobs:
        psi = oracle_run_job(j)
        assert abs(np.vdot(psi, psi).real - 1) < 1e-12
        n = j.qubits
        if j.kind == "hadamard":
            assert np.abs(psi - 2.0 ** (-n / 2)).max() < 1e-12
        elif j.kind == "grover":
            want = grover_closed_form(n, j.params["marked"], j.params["iterations"])
            assert np.abs(psi - want).max() < 1e-10
        elif j.kind == "adder":
            p = np.abs(psi) ** 2
            assert int(np.argmax(p)) == j.expect and p[j.expect] > 1 - 1e-10
            w = j.params["width"]
            assert (j.expect & ((1 << (w + 1)) - 1)) == j.params["a"] + j.params["b"]
        seen.add(j.kind)
    assert seen == set(S.KINDS)


def test_queue_latency():
    lat = S.queue_latency([0.0, 0.001, 0.010], [5.0, 5.0, 1.0])
    # second waits for the first (done at 5 ms, then 5 ms more); third arrives after both
    assert np.allclose(lat, [5.0, 9.0, 1.0])


def test_fidelity():
    rng = np.random.default_rng(0)
    a = rng.normal(size=64) + 1j * rng.normal(size=64)
    assert S.fidelity(a, 2j * a) == pytest.approx(1.0)
    b = a.copy()
    b[0] += 1
    assert S.fidelity(a, b) < 1
