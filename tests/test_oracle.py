"""Pins the CPU oracle (oracle/oracle.cpp) before anything is checked against it.

* the reference's own truth tables (tests/test_gates.cpp:45-90 of the reference)
* closed forms: Hadamard layer, Grover's iterate
* a textbook Kronecker-product brute force (tests/bruteforce.py) for n <= 7
* golden vectors written by the reference itself (tests/golden/): parsed PT
  lists, FP32 final states of t_CPU and t_PAR_CPU (bit-exact), measured bits
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import bruteforce as bf


def truth_table_cases():
    """(qubits, circuit, start, expected) of the reference's test_gates.cpp."""
    cases = []
    for s, e in enumerate([0, 1, 3, 2]):
        cases.append((2, "Control1(1),Target1(X)", s, e))                      # CNot(2,0,1,1)
    for s, e in enumerate([0, 1, 2, 3, 4, 5, 7, 6]):
        cases.append((3, "Control1(1),Control1(1),Target1(X)", s, e))         # Toffoli(3,0,1,2,3)
    for s in range(16):
        cases.append((4, "Control1(1),Control1(1),Control1(1),Target1(X)", s, s ^ 1 if (s >> 1) == 7 else s))
    for s, e in enumerate([1, 0, 2, 3]):
        cases.append((2, "Control1(0),Target1(X)", s, e))                      # Controlled1(2,0,1,"X",0)
    return cases


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("engine", [po.T_CPU, po.T_PAR_CPU])
def test_truth_tables(precision, engine):
    cases = truth_table_cases()
    assert len(cases) == 32
    for n, circ, start, exp in cases:
        st = po.run(n, circ, basis=start, precision=precision, engine=engine, threads=2)
        want = np.zeros(1 << n)
        want[exp] = 1
        assert np.abs(st - want).max() < 1e-6, (circ, start)


@pytest.mark.parametrize("n", [10, 14, 18])
def test_hadamard_closed_form(n):
    # general.out applies H three times (src/cli/general.cpp:44): the state is
    # the uniform superposition; smoke_test.sh:71 expects 0.03125 / 0.0078125 / 0.001953125
    st = po.run(n, ",".join(["H"] * n), precision=64, engine=po.T_PAR_CPU, threads=4, iterations=1,
                reset_steps=[([",".join(["H"] * n)], 2)])
    assert np.abs(st - 2.0 ** (-n / 2)).max() < 1e-12
    assert abs(st[0].real - {10: 0.03125, 14: 0.0078125, 18: 0.001953125}[n]) < 1e-12


def grover_closed_form(n, marked, k):
    """Ancilla-free Grover after k iterations from H^n|0>."""
    N = 2 ** n
    th = np.arcsin(1 / np.sqrt(N))
    sgn = (-1) ** k
    amp = np.full(N, sgn * np.cos((2 * k + 1) * th) / np.sqrt(N - 1))
    amp[marked] = sgn * np.sin((2 * k + 1) * th)
    return amp


def grover_steps(n, marked):
    xs = ["X" if not (marked >> (n - 1 - q)) & 1 else "ID" for q in range(n)]
    mcz = ",".join(["Control1(1)"] * (n - 1) + ["Target1(Z)"])
    h, x = ",".join(["H"] * n), ",".join(["X"] * n)
    return [",".join(xs), mcz, ",".join(xs), h, x, mcz, x, h]


@pytest.mark.parametrize("n,marked,k", [(6, 41, 3), (9, 300, 5), (10, 0, 2), (10, 1023, 25)])
def test_grover_closed_form(n, marked, k):
    o = po.Oracle(n, 64)
    o.set_function([",".join(["H"] * n)])
    o.set_function(grover_steps(n, marked), k, reset=False)
    o.set_basis(0)
    o.execute(po.T_CPU)
    assert np.abs(o.state() - grover_closed_form(n, marked, k)).max() < 1e-12


def _random_gate_list(rng, n, count):
    t, cm, cv, mats = [], [], [], []
    for _ in range(count):
        tb = int(rng.integers(0, n))
        others = [b for b in range(n) if b != tb]
        k = int(rng.integers(0, min(3, len(others)) + 1))
        ctrl = rng.choice(others, size=k, replace=False) if k else []
        mask = sum(1 << int(b) for b in ctrl)
        val = sum((int(rng.integers(0, 2)) << int(b)) for b in ctrl)
        kind = rng.integers(0, 3)
        z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        u, _ = np.linalg.qr(z)
        if kind == 1:
            u = np.diag(np.exp(1j * rng.uniform(0, 6.3, 2)))
        elif kind == 2:
            u = np.array([[0, np.exp(1j * rng.uniform(0, 6.3))], [np.exp(1j * rng.uniform(0, 6.3)), 0]])
        t.append(tb)
        cm.append(mask)
        cv.append(val)
        mats.append(u.reshape(4))
    return t, cm, cv, mats


def oracle_run_gates(n, t, cm, cv, mats, basis=None, state=None, engine=po.T_CPU, precision=64):
    o = po.Oracle(n, precision)
    names = []
    for i, m in enumerate(mats):
        o.add_gate(f"g{i}", m)
        names.append(f"g{i}")
    import ctypes as C
    arr = (C.c_char_p * max(1, len(names)))(*[s.encode() for s in names])
    ta, ma, va = np.asarray(t, np.int32), np.asarray(cm, np.int64), np.asarray(cv, np.int64)
    o.lib.orc_set_gates(o.h, len(t), ta.ctypes.data, ma.ctypes.data, va.ctypes.data, arr)
    if state is not None:
        o.set_state(state)
    else:
        o.set_basis(basis or 0)
    o.execute(engine, threads=2, coalesced_bits=2, region_bits=min(4, n))
    return o.state()


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7])
@pytest.mark.parametrize("engine", [po.T_CPU, po.T_PAR_CPU])
def test_bruteforce_random_gates(n, engine):
    rng = np.random.default_rng(100 + n)
    t, cm, cv, mats = _random_gate_list(rng, n, 30)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    want = bf.apply_gates(n, psi0, t, cm, cv, mats)
    got = oracle_run_gates(n, t, cm, cv, mats, state=psi0, engine=engine)
    assert np.abs(got - want).max() < 1e-12


def test_bruteforce_catches_planted_mistake():
    """A transposed matrix in the oracle's input must fail the brute force."""
    rng = np.random.default_rng(5)
    n = 4
    t, cm, cv, mats = _random_gate_list(rng, n, 10)
    psi0 = np.zeros(1 << n, complex)
    psi0[3] = 1
    want = bf.apply_gates(n, psi0, t, cm, cv, mats)
    bad = [np.asarray(m).reshape(2, 2).T.reshape(4) for m in mats]
    got = oracle_run_gates(n, t, cm, cv, bad, state=psi0)
    assert np.abs(got - want).max() > 1e-3


def test_parse_matches_reference(golden):
    for case in golden["parse"]:
        n = len(case["text"].split(";")[0].split(","))
        o = po.Oracle(n, 64)
        o.set_function(case["text"].split(";"), case.get("iterations", 1))
        if "append" in case:
            o.set_function(case["append"], case["append_iterations"], reset=False)
        t, cm, cv, _ = o.gates()
        assert t.tolist() == case["target_bits"], case["text"]
        assert cm.tolist() == case["control_masks"], case["text"]
        assert cv.tolist() == case["control_values"], case["text"]


def _oracle_for_circuit(c, precision, catalog):
    o = po.Oracle(c["qubits"], precision, catalog)
    for name, m in c["gates"].items():
        o.add_gate(name, [complex(a, b) for a, b in m])
    o.set_function(c["steps"], c["iterations"])
    o.set_basis(c["basis"])
    return o


@pytest.mark.parametrize("engine,tag", [(po.T_CPU, "tcpu"), (po.T_PAR_CPU, "tparcpu")])
def test_golden_states_fp32_bit_exact(golden, engine, tag):
    """The FP32 oracle reproduces the reference's FP32 output bit for bit."""
    for c in golden["circuits"]:
        o = _oracle_for_circuit(c, 32, po.CAT_REFERENCE)
        o.execute(engine, threads=2, coalesced_bits=min(3, c["qubits"] - 1), region_bits=min(6, c["qubits"]))
        ref = golden["states"][f"{c['name']}__{tag}"]
        got = o.state()
        # IEEE equality of every component: exact, except that -0.0 == +0.0
        assert np.array_equal(got, ref), c["name"]


def test_golden_states_fp64_close(golden):
    """The FP64 oracle (double catalog) agrees with the reference's FP32 output to FP32 accuracy."""
    for c in golden["circuits"]:
        o = _oracle_for_circuit(c, 64, po.CAT_DOUBLE)
        o.execute(po.T_CPU)
        ref = golden["states"][f"{c['name']}__tcpu"].astype(np.complex128)
        assert np.abs(o.state() - ref).max() < 2e-5, c["name"]


def test_measure_matches_reference(golden):
    for case in golden["measure"]:
        n, v = case["qubits"], case["search_value"]
        o = po.Oracle(n, 64)
        from oracle.pyoracle import _load  # noqa: F401
        steps = [",".join(["H"] * n)]
        o.set_function(steps)
        o.set_function(_grover_ref_iteration(n, v), case["iterations"], reset=False)
        o.set_basis(1 << (n - 1))
        o.execute(po.T_CPU)
        o.seed(case["seed"])
        bits = [o.measure(q) for q in range(1, n)]
        assert bits == case["bits"]
        assert int("".join(map(str, bits)), 2) == v


def _amps(pairs, dtype):
    return np.array([complex(re, im) for re, im in pairs], dtype=dtype)


@pytest.mark.parametrize("precision", [32, 64])
def test_measure_float_comparison(golden, precision):
    """measure(qubit) keeps the reference's float arithmetic (src/core/dgm_core.cpp:75-99):
    on states and seeds whose draw falls in a rounding window of P(0) (written by the
    reference, tests/golden/measure_edges.json) the FP32 oracle returns the reference's
    bit and collapsed state exactly. The cases discriminate: with P(0) and the draw
    compared in double, each would come out the other way."""
    for case in golden["measure_edges"]:
        n = case["qubits"]
        lo, hi = case["window"]
        if case["name"].startswith("draw_rounding"):
            p0 = sum(re * re + im * im for re, im in case["amps"][: 1 << (n - 1)])
            assert lo < p0 == hi and case["bit"] == 1      # double semantics: p0 > draw -> 0
            double_bit = 0
        else:
            zf, zd = case["zero_float"], case["zero_double"]
            double_bit = 0 if zd > hi else 1               # zd > every draw in the window, or none
            assert lo <= min(zf, zd) + 1e-18 and double_bit != case["bit"]
        o = po.Oracle(n, precision)
        o.set_state(_amps(case["amps"], np.complex64 if precision == 32 else np.complex128))
        o.seed(case["seed"])
        bit = o.measure(case["qubit"])
        if precision == 32:
            assert bit == case["bit"], case["name"]
            assert np.array_equal(o.state(), _amps(case["after"], np.complex64)), case["name"]
        elif case["name"].startswith("draw_rounding"):
            # P(0) is exact in double too: the float-rounded draw decides as in the reference
            assert bit == case["bit"], case["name"]
        else:
            # a double sum of P(0) sits on the other side of the draw
            assert bit == double_bit, case["name"]


def _grover_ref_iteration(n, v):
    """lib_grover.cpp:34-51, written out from the reference's documented form."""
    ctrl = ["Target1(X)"] + [f"Control1({(v >> (n - 1 - q)) & 1})" for q in range(1, n)]
    steps = [",".join(ctrl)]
    one = lambda q, g: ",".join(g if i == q else "ID" for i in range(n))  # noqa: E731
    for q in range(1, n):
        steps += [one(q, "H"), one(q, "X")]
    steps.append(",".join(["ID"] + ["Control1(1)"] * (n - 2) + ["Target1(Z)"]))
    for q in range(n - 1, 0, -1):
        steps += [one(q, "X"), one(q, "H")]
    return steps


def test_grover_ref_iteration_text_matches_reference(golden):
    for s in golden["strings"]:
        if s["which"] == "GroverIterationRef":
            n, v = s["args"]
            assert ";".join(_grover_ref_iteration(n, v)) == s["text"]


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
def test_live_reference_random_circuits():
    """Fresh random circuits: oracle FP32 == reference FP32 bit for bit (t_PAR_CPU)."""
    from tests.golden.make_golden import custom_gates, random_circuit
    rng = np.random.default_rng(7)
    for n in (5, 8, 12):
        g = custom_gates(rng)
        steps = random_circuit(rng, n, 12)
        r = po.RefDGM(n)
        o = po.Oracle(n, 32, po.CAT_REFERENCE)
        for k, m in g.items():
            r.add_gate(k, m)
            o.add_gate(k, m)
        r.set_function(steps)
        o.set_function(steps)
        r.set_basis(1)
        o.set_basis(1)
        r.execute(po.T_PAR_CPU, threads=3, coalesced_bits=2, region_bits=5)
        o.execute(po.T_PAR_CPU, threads=3, coalesced_bits=2, region_bits=5)
        assert np.array_equal(r.state(), o.state())
