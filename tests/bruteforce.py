"""Textbook dense-matrix simulation for tiny circuits (test infrastructure).

Builds each gate's full 2^n x 2^n unitary from Kronecker products in circuit
qubit order (qubit 0 is the leftmost factor, i.e. the most significant index
bit) and projectors on the control qubits, then multiplies. Shares no code
and no bit tricks with the oracle or the HIP library.
"""
from __future__ import annotations

import numpy as np

P0 = np.array([[1, 0], [0, 0]], dtype=np.complex128)
P1 = np.array([[0, 0], [0, 1]], dtype=np.complex128)
I2 = np.eye(2, dtype=np.complex128)


def kron_all(factors):
    out = np.array([[1.0 + 0j]])
    for f in factors:
        out = np.kron(out, f)
    return out


def gate_unitary(n, target_qubit, matrix, controls):
    """controls: dict qubit -> required value."""
    u = np.asarray(matrix, dtype=np.complex128).reshape(2, 2)
    fire = [I2] * n
    for q, v in controls.items():
        fire[q] = P1 if v else P0
    fire_t = list(fire)
    fire_t[target_qubit] = u
    full = kron_all(fire_t)
    # identity on the complement of the control pattern
    full += np.eye(2 ** n, dtype=np.complex128) - kron_all(fire)
    return full


def bit_to_qubit(n, bit):
    return n - 1 - bit


def apply_gates(n, state, target_bits, masks, values, mats):
    psi = np.asarray(state, dtype=np.complex128).copy()
    for t, cm, cv, m in zip(target_bits, masks, values, mats):
        ctrls = {}
        for b in range(n):
            if (int(cm) >> b) & 1:
                ctrls[bit_to_qubit(n, b)] = (int(cv) >> b) & 1
        psi = gate_unitary(n, bit_to_qubit(n, int(t)), m, ctrls) @ psi
    return psi
