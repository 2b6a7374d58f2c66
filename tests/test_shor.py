"""Shor's algorithm: circuit builders byte-identical to the reference's, the
classical helpers equal to the reference's on every fixture vector, and (GPU)
whole attempts with the reference's seeds finding the same factors as the
reference's own t_CPU attempts."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import simulador_quantico_amd as hq
from simulador_quantico_amd import _native as N

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "shor.json")


@pytest.fixture(scope="module")
def shor_golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_builders_match_reference_strings(shor_golden):
    fns = {"C2AddMod": hq.C2AddMod, "C2SubMod": hq.C2SubMod, "CMultMod": hq.CMultMod, "CRMultMod": hq.CRMultMod}
    for case in shor_golden["strings"]:
        d = hq.DGM()
        got = ";".join(fns[case["which"]](*case["args"], d))
        assert got == case["text"], case["which"]


def test_classical_helpers_match_reference(shor_golden):
    for num, den, w, rn, rd in shor_golden["frac_approx"]:
        a, b = C.c_int(num), C.c_int(den)
        N.lib.hgm_frac_approx(C.byref(a), C.byref(b), w)
        assert (a.value, b.value) == (rn, rd), (num, den, w)
    for v, m, r in shor_golden["mul_inv"]:
        assert N.lib.hgm_mul_inv(v, m) == r
        assert (v * r) % m == 1
    for v, bits, r in shor_golden["revert_bits"]:
        assert N.lib.hgm_revert_bits(v, bits) == r


def test_modular_multiplication_circuit_on_the_oracle():
    """CMultMod then CRMultMod with the control set: |1>|x>|0> -> |1>|x>|a x mod N> -> back."""
    from tests.test_serving_mix import oracle_run_gates
    N_, a, w = 15, 7, 4
    qubits = 2 * w + 3
    ctrl, reg1, reg2, over, over_bool = 0, 1, w + 2, w + 1, qubits - 1
    d = hq.DGM()
    d.qubits = qubits
    for x in (1, 4, 11):
        steps = hq.CMultMod(qubits, ctrl, reg1, reg2, over, over_bool, w, a, N_, d)
        d.setFunction(steps)
        t, cm, cv, mats = hq.ops_to_arrays(d.ops())
        basis = (1 << (qubits - 1 - ctrl)) | (x << (qubits - reg1 - w))
        psi = oracle_run_gates(qubits, t, cm, cv, mats, basis=basis)
        idx = int(np.argmax(np.abs(psi)))
        assert np.abs(psi[idx]) > 1 - 1e-9
        y = (idx >> (qubits - reg2 - w)) & ((1 << w) - 1)
        assert y == (a * x) % N_ and (idx >> (qubits - reg1 - w)) & ((1 << w) - 1) == x


@pytest.mark.gpu
def test_shor_attempts_match_reference(shor_golden):
    """Same seed: same base, same measured phase bits, same factors as the reference's t_CPU."""
    for case in shor_golden["attempts"]:
        for resident in (True, False):
            d = hq.DGM()
            d.exec_type = hq.T_GPU
            d.resident = int(resident)
            hq.seed_rng(case["seed"])
            f, base, bits, st = hq.Shor(case["number"], d)
            want = tuple(case["factors"]) if case["factors"][0] else ()
            assert f == want, (case, base, bits)
            assert st["gates"] > 0
