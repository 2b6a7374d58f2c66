"""ORACLE -- test infrastructure only: the CPU baseline timed beside the GPU path.

Run by bench.py's cpu_baseline leg as a child process, so that the OpenMP
runtime starts with the placement the baseline asks for (OMP_NUM_THREADS = the
host's physical cores, OMP_PROC_BIND=close, OMP_PLACES=cores, set by the
caller before this process starts). Times, on one random layer of the
benchmark's random-layer circuit (U3 on every qubit + a CZ brickwork, the
first layer of the seed-1 circuit):

  * the reference's own t_PAR_CPU (src/core/dgm_par_exec.cpp:82-132, FP32, its
    precision), compiled from its sources into oracle/_ref/libhgmref.so, when
    that library is present ("reference");
  * this repository's FP64 restatement of the same engine (oracle.cpp, the
    precision of the GPU line) ("port").

Both with the reference's CLI region defaults (cpu_region_bits 14,
cpu_coalesced_bits 11, include/cli/cli_common.h:10-11), one untimed warm-up run
and the median of `runs` timed runs. Prints one JSON object.

Usage: python -m oracle.cpu_baseline [--qubits N] [--runs 3] [--budget-s 20]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def host_cpu() -> dict:
    """CPU model, physical cores (distinct core/socket pairs of `lscpu -p`) and
    logical CPUs of this host."""
    model, cores = "unknown", None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
        p = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        pairs = {ln for ln in p.splitlines() if ln and not ln.startswith("#")}
        cores = len(pairs) or None
    except (OSError, subprocess.SubprocessError):
        pass
    logical = os.cpu_count() or 1
    return {"model": model, "physical_cores": cores or logical, "logical_cpus": logical}


def _circuit(n: int):
    import simulador_quantico_amd as hq
    d = hq.DGM()
    steps = hq.RandomLayers(n, 1, 1, d)
    names = sorted({tok for s in steps for tok in s.split(",") if tok.startswith("U3_")})
    d.setFunction(steps)
    return steps, {nm: d.gates.getMatrix(nm) for nm in names}, len(d.pts()[0])


def _time_engine(kind: str, n: int, threads: int, runs: int):
    from oracle import pyoracle as po
    steps, mats, gates = _circuit(n)
    eng = po.RefDGM(n) if kind == "reference" else po.Oracle(n, 64)
    for nm, m in mats.items():
        eng.add_gate(nm, m)
    eng.set_function(steps)
    times = []
    for k in range(runs + 1):             # run 0 is the warm-up
        eng.set_basis(0)
        t0 = time.perf_counter()
        eng.execute(po.T_PAR_CPU, threads=threads, coalesced_bits=11, region_bits=14)
        dt = time.perf_counter() - t0
        if k:
            times.append(dt)
    times.sort()
    med = times[len(times) // 2]
    del eng
    return {"value": gates * float(1 << n) / med, "unit": "amp-updates/s", "seconds_median": med,
            "seconds_runs": times, "gates": gates, "qubits": n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=0, help="0: the largest count whose runs fit the budget")
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--budget-s", type=float, default=20.0, help="CPU time for all runs of both engines")
    args = ap.parse_args()
    from oracle import pyoracle as po
    cpu = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", cpu["physical_cores"]))
    have_ref = po.ref_available()
    n = args.qubits
    if n <= 0:   # calibrate on 22 qubits, then grow while (warm-up + runs) of both engines fit the budget
        probe = _time_engine("port", 22, threads, 1)
        rate = probe["value"] * (0.8 if have_ref else 1.0)
        n = 22
        per_layer = lambda q: 1.5 * q * float(1 << q) / rate   # noqa: E731  (U3 + CZ gates of one layer)
        while n < 30 and 2 * (args.runs + 1) * per_layer(n + 1) < args.budget_s:
            n += 1
    out = {"threads": threads, "cpu": cpu, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
           "omp_places": os.environ.get("OMP_PLACES"), "runs": args.runs, "warmup": 1,
           "circuit": f"one random layer (U3 on every qubit + CZ brickwork, seed 1) at {n} qubits"}
    if have_ref:
        out["reference_fp32"] = _time_engine("reference", n, threads, args.runs)
    out["port_fp64"] = _time_engine("port", n, threads, args.runs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
