"""Circuit step generators and builders, by the reference's names.

Thin wrappers over hgm_generate (include/hgm_gpu.h): the text is produced by
the C++ generators of libhgm.so (src/core/gates.cpp:69-163,
src/algorithms/lib_grover.cpp, src/algorithms/lib_shor_circuits.cpp in the
reference). Builders that register gates (QFT, AddF, RandomLayers, ...) take
the DGM whose catalog receives them.
"""
from __future__ import annotations

import ctypes as C

from ._native import check, lib


def _gen(which: str, args, op: str = "", dgm=None) -> str:
    a = (C.c_int64 * max(1, len(args)))(*[int(x) for x in args])
    h = dgm._h if dgm is not None else None
    n = C.c_size_t(0)
    check(lib.hgm_generate(h, which.encode(), a, len(args), op.encode(), None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(lib.hgm_generate(h, which.encode(), a, len(args), op.encode(), buf, n.value + 1, C.byref(n)))
    return buf.value.decode()


def _steps(s: str):
    return s.split(";") if s else []


def CNot(qubits, ctrl, target, control_value=1):
    return _gen("CNot", [qubits, ctrl, target, control_value])


def Toffoli(qubits, ctrl1, ctrl2, target, control_value=3):
    return _gen("Toffoli", [qubits, ctrl1, ctrl2, target, control_value])


def Controlled1(qubits, ctrl, target, op, control_value=1):
    return _gen("Controlled1", [qubits, ctrl, target, control_value], op)


def Controlled2(qubits, ctrl1, ctrl2, target, op, control_value=3):
    return _gen("Controlled2", [qubits, ctrl1, ctrl2, target, control_value], op)


def ControlledN(qubits, ctrls, target, op, control_value=-1):
    return _gen("ControlledN", [qubits, target, control_value, *ctrls], op)


def Pauli_X(qubits, reg, width=1):
    return _gen("Pauli_X", [qubits, reg, width])


def Pauli_Z(qubits, reg, width=1):
    return _gen("Pauli_Z", [qubits, reg, width])


def Hadamard(qubits, reg, width=1):
    return _gen("Hadamard", [qubits, reg, width])


def Oracle1(qubits, search_value):
    return _gen("Oracle1", [qubits, search_value])


def ControledZ(qubits):
    return _gen("ControledZ", [qubits])


def QFT(qubits, reg, over, width, dgm):
    return _steps(_gen("QFT", [qubits, reg, over, width], dgm=dgm))


def RQFT(qubits, reg, over, width, dgm):
    return _steps(_gen("RQFT", [qubits, reg, over, width], dgm=dgm))


def AddF(qubits, reg, over, value, width, dgm, controlled=None):
    if controlled is None:
        return _gen("AddF", [qubits, reg, over, value, width], dgm=dgm)
    return _steps(_gen("AddFc", [qubits, reg, over, value, width, int(controlled)], dgm=dgm))


def SubF(qubits, reg, over, value, width, dgm, controlled=None):
    if controlled is None:
        return _gen("SubF", [qubits, reg, over, value, width], dgm=dgm)
    return _steps(_gen("SubFc", [qubits, reg, over, value, width, int(controlled)], dgm=dgm))


def CAddF(qubits, ctrl1, reg, over, value, width, dgm):
    return _gen("CAddF", [qubits, ctrl1, reg, over, value, width], dgm=dgm)


def C2AddF(qubits, ctrl1, ctrl2, reg, over, value, width, dgm):
    return _gen("C2AddF", [qubits, ctrl1, ctrl2, reg, over, value, width], dgm=dgm)


def CSubF(qubits, ctrl1, reg, over, value, width, dgm):
    return _gen("CSubF", [qubits, ctrl1, reg, over, value, width], dgm=dgm)


def C2SubF(qubits, ctrl1, ctrl2, reg, over, value, width, dgm):
    return _gen("C2SubF", [qubits, ctrl1, ctrl2, reg, over, value, width], dgm=dgm)


def genRot(qubits, reg, phase_bits, dgm):
    return _gen("genRot", [qubits, reg, phase_bits], dgm=dgm)


def CSwapR(qubits, ctrl, reg1, reg2, width):
    return _steps(_gen("CSwapR", [qubits, ctrl, reg1, reg2, width]))


def SwapOver(qubits, reg, width):
    return _steps(_gen("SwapOver", [qubits, reg, width]))


def GroverIterationRef(qubits, search_value):
    """One iteration of the reference's ancilla-form Grover (after its H layer)."""
    return _steps(_gen("GroverIterationRef", [qubits, search_value]))


def GroverIteration(qubits, marked):
    """One ancilla-free Grover iteration: oracle (X, MCZ, X) then diffusion (H, X, MCZ, X, H)."""
    return _steps(_gen("GroverIteration", [qubits, marked]))


def GroverIterations(search_bits):
    return int(_gen("GroverIterations", [search_bits]))


def RandomLayers(qubits, depth, seed, dgm):
    """U3 on every qubit + CZ brickwork per layer; registers U3_<layer>_<qubit> in dgm."""
    return _steps(_gen("RandomLayers", [qubits, depth, seed], dgm=dgm))


def DraperAdder(width, dgm):
    """QFT-based adder on |a>_w |carry,b>_(w+1) -> |a, a+b>; registers R/R' rotations in dgm."""
    return _steps(_gen("DraperAdder", [width], dgm=dgm))


def C2AddMod(qubits, ctrl1, ctrl2, reg, over, over_bool, width, a, N, dgm):
    return _steps(_gen("C2AddMod", [qubits, ctrl1, ctrl2, reg, over, over_bool, width, a, N], dgm=dgm))


def C2SubMod(qubits, ctrl1, ctrl2, reg, over, over_bool, width, a, N, dgm):
    return _steps(_gen("C2SubMod", [qubits, ctrl1, ctrl2, reg, over, over_bool, width, a, N], dgm=dgm))


def CMultMod(qubits, ctrl, reg1, reg2, over, over_bool, width, a, N, dgm):
    return _steps(_gen("CMultMod", [qubits, ctrl, reg1, reg2, over, over_bool, width, a, N], dgm=dgm))


def CRMultMod(qubits, ctrl, reg1, reg2, over, over_bool, width, a, N, dgm):
    return _steps(_gen("CRMultMod", [qubits, ctrl, reg1, reg2, over, over_bool, width, a, N], dgm=dgm))


def Shor(number, dgm):
    """One attempt of Shor's algorithm on an executor configured like `dgm`
    (src/algorithms/lib_shor.cpp:17-163). Returns (factors, base, phase_bits,
    stats): factors is () when the attempt failed."""
    from ._native import Stats
    out = (C.c_int64 * 4)()
    st = Stats()
    check(lib.hgm_shor(dgm._h, int(number), out, C.byref(st)))
    f = (out[0], out[1]) if out[0] else ()
    return f, out[2], out[3], st.as_dict()
