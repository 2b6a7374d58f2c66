"""Writes the golden fixtures of tests/golden/ from the reference itself.

Runs the reference's own code -- its circuit generators, parser, t_CPU and
t_PAR_CPU engines and measure() -- compiled from its sources into
oracle/_ref/libhgmref.so by oracle/Makefile (`make -C oracle ref`). The
outputs are data only:

  strings.json  generator outputs (gates.cpp, lib_grover.cpp, lib_shor_circuits.cpp)
                for the parameter sets of the reference's tests/test_qft_addf.cpp
                and more
  parse.json    compiled PT lists (target bit, control mask/value) of circuit texts
  circuits.json circuit texts + custom FP32 gate matrices of the state fixtures
  states.npz    final FP32 states of those circuits under t_CPU and t_PAR_CPU
  measure.json  measured bits of Grover runs at fixed seeds
  grover23.json the reference's Grover() (src/algorithms/lib_grover.cpp:15-66) at
                qubits = 23 (22-bit register + ancilla, 1608 iterations) on
                its t_PAR_CPU: the bits it measures at a fixed seed
  measure_edges.json  measure(qubit) on states and seeds chosen so that the
                reference's float arithmetic (float sums of P(0), the draw
                rounded to float, src/core/dgm_core.cpp:75-99) decides the
                outcome: the draw lies in a rounding window of P(0)

Usage: python tests/golden/make_golden.py [shor | measure_edges | grover23]   (needs the reference tree)
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402

T_CPU, T_PAR_CPU = 0, 1


def gen_strings():
    r = po.RefDGM(4)
    out = []

    def add(which, args, op=""):
        out.append({"which": which, "args": list(args), "op": op, "text": r.gen(which, args, op)})

    for q, c, t, v in [(2, 0, 1, 1), (3, 2, 0, 0), (5, 1, 4, 1), (8, 7, 0, 1)]:
        add("CNot", [q, c, t, v])
    for q, c1, c2, t, v in [(3, 0, 1, 2, 3), (4, 3, 0, 1, 2), (5, 1, 2, 4, 1), (6, 5, 4, 0, 0)]:
        add("Toffoli", [q, c1, c2, t, v])
    for q, c, t, v, op in [(2, 0, 1, 0, "X"), (4, 3, 1, 1, "R2"), (6, 0, 5, 1, "Z")]:
        add("Controlled1", [q, c, t, v], op)
    for q, c1, c2, t, v, op in [(4, 0, 1, 3, 3, "X"), (5, 4, 2, 0, 2, "H"), (5, 1, 3, 2, 1, "Y")]:
        add("Controlled2", [q, c1, c2, t, v], op)
    # ControlledN: args = qubits, target, control_value, ctrls...
    for q, ctrls, t, v, op in [(4, [0, 1, 2], 3, -1, "X"), (6, [5, 0, 3], 1, 5, "Z"), (5, [4], 0, 0, "H"),
                               (7, [0, 1, 2, 3, 4, 5], 6, 42, "X")]:
        add("ControlledN", [q, t, v, *ctrls], op)
    for q, reg, w in [(4, 0, 4), (6, 2, 3), (5, 4, 1), (9, 1, 7)]:
        add("Pauli_X", [q, reg, w])
        add("Pauli_Z", [q, reg, w])
        add("Hadamard", [q, reg, w])
    for q, v in [(4, 5), (8, 10), (12, 10), (6, 0), (7, 63)]:
        add("Oracle1", [q, v])
    for q in [2, 3, 8, 12]:
        add("ControledZ", [q])
    # tests/test_qft_addf.cpp:176-202 parameter sets
    for q, reg, over, w in [(10, 1, 6, 4), (15, 2, 9, 6), (8, 0, 5, 3)]:
        add("QFT", [q, reg, over, w])
        add("RQFT", [q, reg, over, w])
    for v in [5, 13, 57, 0, 1023]:
        for w in [4, 6, 8]:
            q, reg, over = w + 3, 1, w + 1
            add("AddFc", [q, reg, over, v, w, 1])
            add("AddFc", [q, reg, over, v, w, 0])
            add("SubFc", [q, reg, over, v, w, 1])
            add("SubFc", [q, reg, over, v, w, 0])
            add("AddF", [q, reg, over, v, w])
            add("SubF", [q, reg, over, v, w])
    for q, c1, reg, over, v, w in [(9, 0, 2, 1, 7, 5), (11, 10, 1, 8, 21, 6)]:
        add("CAddF", [q, c1, reg, over, v, w])
    for q, reg, bits in [(7, 0, 1), (7, 0, 5), (9, 3, 11), (9, 0, 0)]:
        add("genRot", [q, reg, bits])
    for q, c, r1, r2, w in [(9, 0, 1, 5, 4), (11, 0, 1, 6, 5)]:
        add("CSwapR", [q, c, r1, r2, w])
    for q, v in [(5, 3), (8, 10)]:
        out.append({"which": "GroverIterationRef", "args": [q, v], "op": "",
                    "text": ";".join(po.ref_grover_iteration(q, v))})
    return out


def parse_cases():
    texts = [
        ("H,ID,ID", 1),
        ("Control1(1),ID,Target1(X)", 1),
        ("H,Control2(1),Target2(Z),H", 1),
        ("Control1(1),Control1(0),Target1(X);H,H,H", 2),
        ("Target1(X),Control1(1),Control2(0),Target2(Y),H", 1),
        ("H,Control0(1),H", 1),                       # a control in group 0 binds the plain gates
        ("Control1(2),Target1(X),ID", 1),             # any nonzero control value means 1
        ("X,Y,Z,H;Z,Y,X,H;H,H,H,H", 3),               # sort direction alternates per step
        ("Control3(1),Target3(R2),Target3(R3),Control3(0),ID", 1),   # two targets share a control
        ("Control12(1),ID,Target12(H),Control7(1),Target7(X)", 1),
    ]
    out = []
    for text, it in texts:
        n = len(text.split(";")[0].split(","))
        r = po.RefDGM(n)
        r.set_function(text.split(";"), it)
        t, cm, cv, _ = r.pts()
        out.append({"text": text, "iterations": it, "target_bits": t.tolist(), "control_masks": cm.tolist(),
                    "control_values": cv.tolist()})
    # reset=false appends to the previous list (src/core/dgm_parser.cpp:44)
    r = po.RefDGM(3)
    r.set_function(["H,H,H"], 1)
    r.set_function(["Control1(1),ID,Target1(X)", "X,ID,ID"], 2, reset=False)
    t, cm, cv, _ = r.pts()
    out.append({"text": "H,H,H", "append": ["Control1(1),ID,Target1(X)", "X,ID,ID"], "append_iterations": 2,
                "target_bits": t.tolist(), "control_masks": cm.tolist(), "control_values": cv.tolist()})
    return out


def random_circuit(rng, n, n_steps):
    """Steps with random gates, random groups and controls (one token per qubit)."""
    names = ["H", "X", "Y", "Z", "R1", "R2", "R3", "U0", "U1", "U2", "D0", "A0"]
    steps = []
    for _ in range(n_steps):
        tokens = ["ID"] * n
        free = list(rng.permutation(n))
        group = 1
        while free:
            kind = rng.random()
            if kind < 0.45 or len(free) < 2:
                q = free.pop()
                tokens[q] = str(rng.choice(names))
            elif kind < 0.8:
                k = int(rng.integers(1, min(3, len(free) - 1) + 1))
                ctrls = [free.pop() for _ in range(k)]
                tq = free.pop()
                for c in ctrls:
                    tokens[c] = f"Control{group}({int(rng.integers(0, 2))})"
                tokens[tq] = f"Target{group}({rng.choice(names)})"
                group += 1
            else:
                free.pop()
        steps.append(",".join(tokens))
    return steps


def custom_gates(rng):
    gates = {}
    for k in range(3):   # random unitaries (QR), FP32-representable
        z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        q, _ = np.linalg.qr(z)
        gates[f"U{k}"] = q.reshape(4).astype(np.complex64).astype(np.complex128)
    ph = np.exp(1j * rng.uniform(0, 2 * np.pi, 2))
    gates["D0"] = np.array([ph[0], 0, 0, ph[1]]).astype(np.complex64).astype(np.complex128)
    gates["A0"] = np.array([0, ph[0], ph[1], 0]).astype(np.complex64).astype(np.complex128)
    return gates


def state_fixtures():
    rng = np.random.default_rng(20261202)
    circuits = []
    states = {}

    def run(name, n, steps, basis, gates, iterations=1):
        for eng, tag in [(T_CPU, "tcpu"), (T_PAR_CPU, "tparcpu")]:
            r = po.RefDGM(n)
            for gname, m in gates.items():
                r.add_gate(gname, m)
            r.set_function(steps, iterations)
            r.set_basis(basis)
            # t_PAR_CPU with small regions so the region planner is exercised
            r.execute(eng, threads=2, coalesced_bits=min(3, n - 1), region_bits=min(6, n))
            states[f"{name}__{tag}"] = r.state()
        circuits.append({"name": name, "qubits": n, "steps": steps, "basis": int(basis), "iterations": iterations,
                         "gates": {k: [[float(v.real), float(v.imag)] for v in m] for k, m in gates.items()}})

    run("hadamard10x3", 10, [",".join(["H"] * 10)], 0, {}, iterations=3)
    for i, (n, depth) in enumerate([(6, 8), (8, 10), (9, 12), (10, 14), (11, 10)]):
        g = custom_gates(rng)
        steps = random_circuit(rng, n, depth)
        run(f"random{i}_{n}q", n, steps, int(rng.integers(0, 1 << n)), g)
    # Grover, reference ancilla form, 8 qubits: H layer then iterations
    n, v = 8, 37
    steps = [",".join(["H"] * n)] + po.ref_grover_iteration(n, v) * 2
    run("grover_ref8_v37_2it", n, steps, 1 << (n - 1), {})
    # QFT of the reference's builder on a basis state (uses the reference's R gates)
    r = po.RefDGM(7)
    qft = r.gen("QFT", [7, 1, 0, 6]).split(";")
    gates = {f"R{k}": r.get_gate(f"R{k}") for k in range(1, 8)}
    run("qft7_b45", 7, qft, 45, {k: v for k, v in gates.items() if k not in ("R1", "R2", "R3")})
    return circuits, states


def measure_fixtures():
    """Grover (reference form) measured qubit by qubit at fixed seeds."""
    out = []
    for n, v, seed in [(6, 11, 7), (8, 37, 1234), (9, 200, 99)]:
        it = int(np.pi / 4.0 * np.sqrt(1 << (n - 1)))
        r = po.RefDGM(n)
        r.set_function([",".join(["H"] * n)])
        r.set_function(po.ref_grover_iteration(n, v), it, reset=False)
        r.set_basis(1 << (n - 1))
        r.execute(T_PAR_CPU, threads=2, coalesced_bits=3, region_bits=6)
        po.ref_seed(seed)
        bits = [r.measure(q) for q in range(1, n)]
        out.append({"qubits": n, "search_value": v, "seed": seed, "iterations": it, "bits": bits})
    return out


def _find_seed(lo: float, hi: float, start: int = 1, count: int = 1 << 30) -> int:
    """Smallest seed whose first uniform_real_distribution<double> draw from
    std::mt19937 lies in [lo, hi) (searched by the oracle's helper, which uses
    the same libstdc++ generator and distribution as the reference)."""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    lib.orc_find_draw_seed.restype = C.c_long
    lib.orc_find_draw_seed.argtypes = [C.c_double, C.c_double, C.c_long, C.c_long, C.c_int]
    seed = lib.orc_find_draw_seed(lo, hi, start, count, os.cpu_count() or 1)
    if seed < 0:
        raise RuntimeError(f"no seed in [{start}, {start + count}) draws in [{lo}, {hi})")
    return int(seed)


def _float_sum(terms) -> float:
    """The reference's accumulation: float accumulator, double terms (dgm_core.cpp:79-84)."""
    acc = np.float32(0)
    for t in terms:
        acc = np.float32(np.float64(acc) + t)
    return float(acc)


def measure_edge_fixtures():
    """measure(0) where the reference's float comparison decides the outcome.

    draw_rounding: P(0) is exact in float (0.75 or 0.5) and the draw lies within
    half a float ulp below it, so the draw rounds up to P(0) and `zero > draw`
    is false in float (outcome 1) while true in double (outcome 0).
    float_sum: a seeded 12-qubit state whose float-accumulated P(0) differs
    from the double sum, with the draw between the two."""
    out = []
    for name, amps, p0 in (("draw_rounding_p075", [0.5, 0.5, 0.5, 0, 0.5, 0, 0, 0], 0.75),
                           ("draw_rounding_p05", [0.5, 0.5, 0, 0, 0.5, 0.5, 0, 0], 0.5)):
        a = np.array(amps, np.complex64)
        half_ulp = float(np.spacing(np.float32(p0 / 2)))   # float spacing just below p0
        seed = _find_seed(p0 - half_ulp / 2, p0)
        out.append(_measure_case(name, a, seed, window=[p0 - half_ulp / 2, p0]))
    rng = np.random.default_rng(4242)
    n = 12
    best = None
    for _ in range(64):
        v = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
        v = (v / np.float32(np.linalg.norm(v.astype(np.complex128)))).astype(np.complex64)
        terms = np.abs(v.astype(np.complex128)) ** 2
        zero_half = terms[: 1 << (n - 1)]
        zf, zd = _float_sum(zero_half), float(np.sum(zero_half))
        if best is None or abs(zf - zd) > abs(best[1] - best[2]):
            best = (v, zf, zd)
    v, zf, zd = best
    ulp = float(np.spacing(np.float32(zf)))
    if zf > zd:   # double: zd > r false -> 1; float: zf > float(r) true -> 0 while float(r) < zf
        lo, hi = zd, zf - ulp
    else:         # double: zd > r true -> 0; float: zf > float(r) false -> 1 once r >= zf
        lo, hi = zf, zd
    seed = _find_seed(lo, hi)
    out.append(_measure_case("float_sum_12q", v, seed, window=[lo, hi], zero_float=zf, zero_double=zd))
    return out


def _measure_case(name, amps, seed, **extra):
    n = int(np.log2(len(amps)))
    r = po.RefDGM(n)
    r.set_state(amps)
    po.ref_seed(seed)
    bit = r.measure(0)
    after = r.state()
    case = {"name": name, "qubits": n, "qubit": 0, "seed": seed,
            "amps": [[float(z.real), float(z.imag)] for z in amps], "bit": int(bit),
            "after": [[float(z.real), float(z.imag)] for z in after]}
    case.update(extra)
    return case


def grover23_fixture():
    """Grover's reference form at 23 qubits, built and measured as Grover() does it
    (src/algorithms/lib_grover.cpp:30-66): |1,0..0>, H on all, 1608 iterations of
    oracle + diffusion, then qubits 1..22 measured in order at a fixed seed."""
    n = 23
    v = int(np.random.default_rng(23).integers(0, 1 << (n - 1)))
    it = int(np.pi / 4.0 * np.sqrt(1 << (n - 1)))
    r = po.RefDGM(n)
    r.set_function([",".join(["H"] * n)])
    r.set_function(po.ref_grover_iteration(n, v), it, reset=False)
    r.set_basis(1 << (n - 1))
    r.execute(T_PAR_CPU, threads=os.cpu_count() or 1, coalesced_bits=11, region_bits=14)
    psi = r.state()
    p = np.abs(psi.astype(np.complex128)) ** 2
    reg_marginal = p.reshape(2, -1).sum(axis=0)     # qubit 0 (the ancilla) is the top index bit
    seed = 2323
    po.ref_seed(seed)
    bits = [r.measure(q) for q in range(1, n)]
    return {"qubits": n, "search_value": v, "iterations": it, "seed": seed, "bits": bits,
            "register_argmax": int(np.argmax(reg_marginal)), "p_register_max": float(reg_marginal.max())}


def shor_fixtures():
    """Shor's circuit builders, classical helpers and whole attempts (t_CPU) of the reference."""
    out = {"strings": [], "catalog": {}, "frac_approx": [], "mul_inv": [], "revert_bits": [], "attempts": []}
    r = po.RefDGM(4)
    for which, args in [("C2AddMod", [11, 0, 3, 6, 5, 10, 4, 7, 15]), ("C2SubMod", [11, 0, 2, 6, 5, 10, 4, 13, 15]),
                        ("CMultMod", [11, 0, 1, 6, 5, 10, 4, 7, 15]), ("CRMultMod", [11, 0, 1, 6, 5, 10, 4, 7, 15]),
                        ("CMultMod", [13, 0, 1, 7, 6, 12, 5, 11, 21]), ("CRMultMod", [13, 0, 1, 7, 6, 12, 5, 4, 21])]:
        out["strings"].append({"which": which, "args": args, "text": r.gen(which, args)})
    for name in r.gate_names() if hasattr(r, "gate_names") else []:
        out["catalog"][name] = [list(map(float, (z.real, z.imag))) for z in r.get_gate(name)]
    for w in (3, 4, 5):
        for num in range(1, 1 << (2 * w)):
            out["frac_approx"].append([num, 1 << (2 * w), w, *po.ref_frac_approx(num, 1 << (2 * w), w)])
    for m in (15, 21, 35, 57, 119):
        for v in range(2, m):
            if np.gcd(v, m) == 1:
                out["mul_inv"].append([v, m, po.ref_mul_inv(v, m)])
    for bits in (4, 8, 10):
        for v in range(0, 1 << bits, 7):
            out["revert_bits"].append([v, bits, po.ref_revert_bits(v, bits)])
    for number, seeds in ((15, range(1, 17)), (21, range(1, 9))):
        for seed in seeds:
            out["attempts"].append({"number": number, "seed": seed, "factors": list(po.ref_shor(number, seed))})
    return out


def main():
    if not po.ref_available():
        sys.exit("oracle/_ref/libhgmref.so is missing: run `make -C oracle ref` with the reference tree present")
    if sys.argv[1:] == ["grover23"]:
        with open(os.path.join(HERE, "grover23.json"), "w") as f:
            json.dump(grover23_fixture(), f, indent=0)
        print("grover23.json written to", HERE)
        return
    if sys.argv[1:] == ["measure_edges"]:
        with open(os.path.join(HERE, "measure_edges.json"), "w") as f:
            json.dump(measure_edge_fixtures(), f, indent=0)
        print("measure_edges.json written to", HERE)
        return
    if sys.argv[1:] == ["shor"]:
        with open(os.path.join(HERE, "shor.json"), "w") as f:
            json.dump(shor_fixtures(), f, indent=0)
        print("shor.json written to", HERE)
        return
    with open(os.path.join(HERE, "strings.json"), "w") as f:
        json.dump(gen_strings(), f, indent=0)
    with open(os.path.join(HERE, "parse.json"), "w") as f:
        json.dump(parse_cases(), f, indent=0)
    circuits, states = state_fixtures()
    with open(os.path.join(HERE, "circuits.json"), "w") as f:
        json.dump(circuits, f, indent=0)
    np.savez_compressed(os.path.join(HERE, "states.npz"), **states)
    with open(os.path.join(HERE, "measure.json"), "w") as f:
        json.dump(measure_fixtures(), f, indent=0)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
