"""Benchmark of the t_GPU path on one MI355X.

Default workload: the random-layer circuit at 30 qubits, depth 20, complex
double amplitudes (U3 on every qubit + CZ brickwork per layer, 890 gates),
on a device-resident state. One step = one execution of the whole circuit
(every pass of the plan). Prints one JSON line:

  value      amp-updates/s = steps * gates * 2^qubits / wall time
  roofline   HBM roofline of the dominant kernel (the pass kernel): algorithmic
             bytes per launch (2 * 2^qubits * 16 B) / mean launch duration
             measured with HIP events on the library's launch stream; beside
             the published 8 TB/s peak, copy_ceiling is the rate of a plain
             streaming copy of the same state measured in the same run
  cpu_baseline  the reference's t_PAR_CPU (FP32, its own precision) on a
             bounded sample of the same circuit family, timed on this host

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6     # MI355X dense FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3


def algorithmic_flops(ops, n: int, iterations: int) -> float:
    """Flops of one execution of the gate list, counted as the arithmetic the
    reference's kernel writes (src/core/kernel.cu:109-111): a dense gate does
    4 complex multiplies and 2 complex adds per amplitude pair (14 flops per
    amplitude); a diagonal or anti-diagonal gate one complex multiply (6 flops)
    per amplitude it changes; controls restrict both to the matching amplitudes."""
    total = 0.0
    for i in range(len(ops)):
        o = ops[i]
        m = o.m
        frac = 0.5 ** bin(int(o.control_mask)).count("1")
        if m[2] == 0 and m[3] == 0 and m[4] == 0 and m[5] == 0:
            # diagonal: a multiply where the factor is not exactly 1
            per = (0 if (m[0] == 1 and m[1] == 0) else 3) + (0 if (m[6] == 1 and m[7] == 0) else 3)
        elif m[0] == 0 and m[1] == 0 and m[6] == 0 and m[7] == 0:
            per = 6
        else:
            per = 14
        total += per * frac * float(1 << n)
    return total * iterations
METRIC = "amp-updates/s + %HBM roofline @30q; circuits/s and p99 latency on serving mix"

WORKLOADS = {
    # name: (qubits, dtype, builder, tuning); planner 96 = greedy + gate fusion + row scaling,
    # 224 = the same + pass balancing, 544 = greedy + gate fusion + rotation split, 672 = the same
    # + pass balancing, 928 = the same + two-pass lookahead
    "random30": dict(qubits=30, dtype="c128", kind="random", depth=20, seed=1,
                     tuning=dict(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=4,
                                 kernel=3, planner=928)),
    "u3layers30": dict(qubits=30, dtype="c128", kind="u3layers", depth=20, seed=1,
                       tuning=dict(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=3,
                                   kernel=3, planner=224)),
    "random32f": dict(qubits=32, dtype="c64", kind="random", depth=20, seed=1,
                      tuning=dict(block_size=512, repeat_count=8, gpu_region_bits=13, gpu_coalesced_bits=4,
                                  kernel=3, planner=544)),
    "hadamard20": dict(qubits=20, dtype="c128", kind="hadamard", layers=1,
                       tuning=dict(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=4,
                                   planner=96)),
    "hadamard30x3": dict(qubits=30, dtype="c128", kind="hadamard", layers=3,
                         tuning=dict(block_size=512, repeat_count=8, gpu_region_bits=13, gpu_coalesced_bits=3,
                                     kernel=3, planner=0)),
    "hadamard30x3f": dict(qubits=30, dtype="c64", kind="hadamard", layers=3,
                          tuning=dict(block_size=512, repeat_count=8, gpu_region_bits=13, gpu_coalesced_bits=4,
                                      kernel=3, planner=0)),
    "grover22": dict(qubits=22, dtype="c128", kind="grover", seed=22,
                     tuning=dict(block_size=512, repeat_count=8, gpu_region_bits=13, gpu_coalesced_bits=4,
                                 kernel=3, planner=96)),
    "draper29": dict(qubits=29, dtype="c128", kind="draper", seed=29,
                     tuning=dict(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=3,
                                 kernel=3, planner=96)),
}


def planner_name(flags: int) -> str:
    base = "window" if (flags & 15) == 1 else "greedy"
    extra = [n for bit, n in ((32, "gate fusion"), (512, "rotation split"), (64, "row scaling"), (16, "phase folding"),
                              (128, "pass balancing"), (256, "two-pass lookahead"))
             if flags & bit and not (bit == 16 and flags & (64 | 512)) and not (bit == 64 and flags & 512)]
    return " + ".join([base] + extra)


KERNEL_NAMES = {1: "k_pass_lds", 2: "k_pass_reg", 3: "hgm_pass_* (plan-compiled)"}


def measured_traffic(workload: str, tun: dict, kind: str = "traffic"):
    """A counter measurement stored beside the profile it comes from, or None
    when none exists for this workload and tile shape. kind "traffic": HBM
    bytes per pass-kernel launch from the FETCH_SIZE / WRITE_SIZE counters
    (rocprofv3 --pmc, one run each, scaled by the same counters on a plain copy
    of known size); kind "valu": VALU and FP64 wave instructions of one step."""
    path = os.path.join(ROOT, "profiles", f"{kind}_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if rec.get("tile") != {k: int(v) for k, v in tun.items()}:
        return None
    return rec


def build_circuit(w):
    import simulador_quantico_amd as hq
    n = w["qubits"]
    d = hq.DGM(np.complex64 if w["dtype"] == "c64" else np.complex128)
    d.qubits = n
    iterations = 1
    basis = 0
    if w["kind"] == "random":
        d.setFunction(hq.RandomLayers(n, w["depth"], w["seed"], d))
        name = f"random-layer {n}q depth {w['depth']}"
    elif w["kind"] == "u3layers":
        steps = hq.RandomLayers(n, w["depth"], w["seed"], d)
        d.setFunction([s for s in steps if s.startswith("U3_")])
        name = f"U3 layers without CZ {n}q depth {w['depth']}"
    elif w["kind"] == "hadamard":
        d.setFunction(hq.Hadamard(n, 0, n), w["layers"])
        name = f"Hadamard layer {n}q x{w['layers']}"
    elif w["kind"] == "grover":
        marked = int(np.random.default_rng(w["seed"]).integers(0, 1 << n))
        d.setFunction(hq.GroverIteration(n, marked))
        iterations = hq.GroverIterations(n)
        name = f"Grover {n}q, {iterations} iterations"
    elif w["kind"] == "draper":
        width = (n - 1) // 2
        d.setFunction(hq.DraperAdder(width, d))
        a, b = (int(v) for v in np.random.default_rng(w["seed"]).integers(0, 1 << width, 2))
        basis = (a << (width + 1)) | b
        name = f"Draper adder 2x{width} bits + carry ({n}q)"
    else:
        raise KeyError(w["kind"])
    return d, d.ops(), iterations, basis, name


def cpu_baseline(budget_s: float = 20.0, runs: int = 3) -> dict:
    """The reference's t_PAR_CPU (FP32, its own precision) and the oracle's FP64
    restatement of it on one random layer of the benchmark circuit, timed on every
    physical core of this host (OMP_PROC_BIND=close, OMP_PLACES=cores; one warm-up,
    median of `runs`), in a child process so that OpenMP starts with that placement."""
    import subprocess
    from oracle import cpu_baseline as cb
    cpu = cb.host_cpu()
    env = dict(os.environ, OMP_NUM_THREADS=str(cpu["physical_cores"]), OMP_PROC_BIND="close", OMP_PLACES="cores")
    res = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--runs", str(runs), "--budget-s",
                          str(budget_s)], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    if res.returncode:
        raise RuntimeError(f"CPU baseline failed: {res.stderr[-2000:]}")
    rec = json.loads(res.stdout.strip().splitlines()[-1])
    main = rec.get("reference_fp32") or rec["port_fp64"]
    kind = "reference" if "reference_fp32" in rec else "port"
    label = "the reference's own t_PAR_CPU in FP32" if kind == "reference" else "the oracle's t_PAR_CPU in FP64"
    out = {"value": main["value"], "unit": "amp-updates/s", "cores": rec["threads"], "kind": kind,
           "sample": f"{label} (cpu_region_bits=14, cpu_coalesced_bits=11) on {rec['circuit']} "
                     f"({main['gates']} gates), median of {rec['runs']} runs after one warm-up: "
                     f"{main['seconds_median']:.2f} s, {rec['threads']} threads "
                     f"(OMP_PROC_BIND={rec['omp_proc_bind']}, OMP_PLACES={rec['omp_places']})",
           "cpu_model": rec["cpu"]["model"], "physical_cores": rec["cpu"]["physical_cores"],
           "logical_cpus": rec["cpu"]["logical_cpus"]}
    if kind == "reference":
        port = rec["port_fp64"]
        out["oracle_fp64"] = {"value": port["value"], "unit": "amp-updates/s", "seconds_median": port["seconds_median"],
                              "what": "the oracle's FP64 restatement of t_PAR_CPU on the same layer and threads "
                                      "(the GPU line's precision)"}
    return out


def fidelity_report(names=("hadamard20", "grover22", "draper29", "random30", "random32f")) -> dict:
    """FP32 states against FP64 states of the same circuits, both on the GPU
    (the FP64 path is the one checked against the FP64 oracle within 1e-12):
    fidelity, max amplitude error, kernel times."""
    import simulador_quantico_amd as hq
    from simulador_quantico_amd import _native as N
    out = {}
    for name in names:
        w = WORKLOADS[name]
        n = w["qubits"]
        d, ops, iterations, basis, title = build_circuit(w)
        row = {"circuit": title}
        states = {}
        for dt, tun in ((np.complex128, dict(block_size=256, repeat_count=4, gpu_region_bits=11,
                                               gpu_coalesced_bits=2, kernel=2, planner=96)),
                        (np.complex64, dict(block_size=256, repeat_count=8, gpu_region_bits=12,
                                              gpu_coalesced_bits=4, kernel=2, planner=96))):
            s = hq.State(n, dt)
            s.set_basis(basis)
            st = s.apply(ops, N.default_tuning(**tun), iterations=iterations)
            row["kernel_ms_" + ("f64" if dt == np.complex128 else "f32")] = st["kernel_ms"]
            states[dt] = s
        f, m = states[np.complex128].compare(states[np.complex64])
        i64, p64 = states[np.complex128].argmax()
        i32, p32 = states[np.complex64].argmax()
        row.update(fidelity=f, max_abs_err=m, argmax_equal=bool(i64 == i32), p_max_f64=p64, p_max_f32=p32)
        del states
        print(f"bench: fidelity {name}: {row}", file=sys.stderr, flush=True)
        out[name] = row
    return out


def drop_in_rate(name: str = "random30", tuning: dict | None = None) -> dict:
    """The workload through the drop-in executor (DGM with exec_type t_GPU and a
    pinned host state; every execute() uploads the state, plans, runs the
    passes and downloads, as the reference's t_GPU does): the PCIe-inclusive
    rate, reported beside `value`, never as it."""
    import simulador_quantico_amd as hq
    w = WORKLOADS[name]
    d, ops, iterations, basis, title = build_circuit(w)
    d.exec_type = hq.T_GPU
    tun = dict(w["tuning"]) if tuning is None else tuning
    for k, v in tun.items():
        setattr(d, k, v)
    d.allocateMemory()
    grover = w["kind"] == "grover"

    def init():
        if grover:   # Grover starts from the uniform superposition (the H layer, set on the host)
            d.state[:] = 2.0 ** (-w["qubits"] / 2)
        else:
            d.setMemoryValue(basis)
    init()
    t0 = time.perf_counter()
    d.execute(iterations)                 # first call: device buffer, kernel compilation
    cold = time.perf_counter() - t0
    init()
    t0 = time.perf_counter()
    d.execute(iterations)
    wall = time.perf_counter() - t0
    st = d.stats
    gates = len(ops) * iterations
    out = {"workload": title, "dtype": "f64" if w["dtype"] == "c128" else "f32",
           "amp_updates_per_s": gates * float(1 << w["qubits"]) / wall, "wall_ms": wall * 1e3,
           "first_call_ms": cold * 1e3, "h2d_ms": st["h2d_ms"], "kernel_ms": st["kernel_ms"], "d2h_ms": st["d2h_ms"],
           "plan_ms": st["plan_ms"], "tuning": tun}
    if grover:   # readout on the host state the drop-in returns (untimed)
        prob = np.abs(d.state) ** 2
        idx = int(np.argmax(prob))
        marked = int(np.random.default_rng(w["seed"]).integers(0, 1 << w["qubits"]))
        out.update(argmax=idx, marked=marked, argmax_exact=idx == marked, p_max=float(prob[idx]))
    d.freeMemory()
    return out


def other_configs() -> dict:
    """Short checks of the other benchmark configurations, run after the timed
    steps: Hadamard 20q against its closed form (every amplitude 2^-10),
    Grover 22q through the drop-in executor (argmax = the marked value), and
    the Draper adder 29q (the sum register read bit-exact)."""
    import simulador_quantico_amd as hq
    from simulador_quantico_amd import _native as N
    out = {}
    w = WORKLOADS["hadamard20"]
    d, ops, iterations, basis, title = build_circuit(w)
    s = hq.State(w["qubits"])
    s.set_basis(basis)
    st = hq.Plan(w["qubits"], ops, N.default_tuning(**w["tuning"]), iterations).run(s)
    err = float(np.abs(s.download() - 2.0 ** (-w["qubits"] / 2)).max())
    out["hadamard20"] = {"circuit": title, "kernel_ms": st["kernel_ms"], "max_abs_err_vs_closed_form": err}
    del s
    out["grover22_drop_in"] = drop_in_rate("grover22")
    w = WORKLOADS["draper29"]
    d, ops, iterations, basis, title = build_circuit(w)
    width = (w["qubits"] - 1) // 2
    a, b = (int(v) for v in np.random.default_rng(w["seed"]).integers(0, 1 << width, 2))
    s = hq.State(w["qubits"])
    s.set_basis(basis)
    st = hq.Plan(w["qubits"], ops, N.default_tuning(**w["tuning"]), iterations).run(s)
    idx, p = s.argmax()
    out["draper29"] = {"circuit": title, "kernel_ms": st["kernel_ms"], "a": a, "b": b,
                       "sum_read": idx & ((1 << (width + 1)) - 1), "sum_exact": idx == ((a << (width + 1)) | (a + b)),
                       "p_max": p}
    del s
    return out


def general_out(qubits: int, tuning: dict) -> dict:
    """The reference's general.out timing (src/cli/general.cpp:42-52 around
    HadamardNQubits, src/algorithms/lib_general.cpp:15-47): executor creation,
    allocateMemory, setMemoryValue(0), setFunction(H on every qubit, 3 times),
    execute(1) through t_GPU with a host state, and a read of state[0], in FP32
    as the reference computes. In a fresh process this includes the HIP
    runtime start-up, as the reference's number includes CUDA's."""
    import simulador_quantico_amd as hq
    t0 = time.perf_counter()
    d = hq.DGM(np.complex64)
    d.qubits = qubits
    d.exec_type = hq.T_GPU
    for k, v in tuning.items():
        setattr(d, k, v)
    t1 = time.perf_counter()
    d.allocateMemory()
    alloc_ms = (time.perf_counter() - t1) * 1e3
    d.setMemoryValue(0)
    d.setFunction(hq.Hadamard(qubits, 0, qubits), 3)
    d.execute(1)
    amp0 = complex(d.state[0])
    wall = time.perf_counter() - t0
    st = d.stats
    return {"qubits": qubits, "dtype": "f32", "tuning": tuning, "seconds": wall, "amplitude_0": [amp0.real, amp0.imag],
            "expected_amplitude_0": 2.0 ** (-qubits / 2), "allocate_ms": alloc_ms, "h2d_ms": st["h2d_ms"],
            "kernel_ms": st["kernel_ms"],
            "d2h_ms": st["d2h_ms"], "plan_ms": st["plan_ms"], "passes": st["passes"]}


def shor_attempts(number: int, attempts: int, dtype, resident: bool, tuning: dict) -> dict:
    """Attempts of Shor's algorithm as the reference's shor.out runs them
    (src/cli/shor.cpp: one Shor() call timed per attempt), through t_GPU."""
    import simulador_quantico_amd as hq
    times, found = [], 0
    gates = 0
    for a in range(attempts):
        hq.seed_rng(1000 + a)
        t0 = time.perf_counter()
        d = hq.DGM(dtype)
        d.exec_type = hq.T_GPU
        d.resident = int(resident)
        for k, v in tuning.items():
            setattr(d, k, v)
        f, base, bits, st = hq.Shor(number, d)
        times.append(time.perf_counter() - t0)
        found += bool(f)
        gates = st["gates"]
    return {"number": number, "attempts": attempts, "found": found, "dtype": "f64" if dtype == np.complex128 else "f32",
            "resident": resident, "tuning": tuning, "seconds_mean": float(np.mean(times)),
            "seconds_first": times[0], "gates_per_attempt": gates}


def batch_graph_bench(copies: int = 8, iterations: int = 8, replays: int = 5) -> dict:
    """Resident FP32 states of 12-19 qubits (`copies` states per qubit count,
    Grover with `iterations` iterations, plan-compiled on 2^12 tiles), run
    once as one batch graph (hgm_batch_*) and once plan by plan."""
    import simulador_quantico_amd as hq
    from simulador_quantico_amd import _native as N
    from simulador_quantico_amd.serving import _circuit_ops
    plans, states = [], []
    for n in range(12, 20):
        a = np.concatenate([_circuit_ops([hq.Hadamard(n, 0, n)])] +
                           [_circuit_ops(hq.GroverIteration(n, ((1 << n) - 1) // 3))] * iterations)
        ops = (N.Op * len(a)).from_buffer_copy(a.tobytes())
        t = N.default_tuning(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=4,
                             kernel=N.HGM_KERNEL_JIT, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE, ops_block=1 << 30)
        p = hq.Plan(n, ops, t, 1, np.complex64)
        for _ in range(copies):
            plans.append(p)
            states.append(hq.State(n, np.complex64))
    for s in states:
        s.set_basis(0)
    for p, s in zip(plans, states):      # warm-up: module loads
        p.run(s)
    seq = []
    for _ in range(replays):
        t0 = time.perf_counter()
        for p, s in zip(plans, states):
            p.run(s)
        seq.append((time.perf_counter() - t0) * 1e3)
    b = hq.Batch(states, plans)
    b.run()
    wall, dev = [], []
    for _ in range(replays):
        t0 = time.perf_counter()
        st = b.run()
        wall.append((time.perf_counter() - t0) * 1e3)
        dev.append(st["kernel_ms"])
    passes = st["passes"]
    return {"states": len(states), "qubits": "12-19", "circuit": f"Grover, {iterations} iterations", "dtype": "f32",
            "passes_per_run": passes, "one_after_another_ms": float(np.median(seq)),
            "batch_graph_ms": float(np.median(wall)), "batch_graph_device_ms": float(np.median(dev)),
            "speedup": float(np.median(seq) / np.median(wall))}


def serving_bench(requests: int, rate: float, seq_sample: int, seed: int = 2026, autotune: bool = False,
                  fidelity_sample: int = 24, tile=None, planner=None, compiled: bool = True) -> dict:
    """The serving mix on the batched server (FP32 states) against sequential
    plain FP64 one-circuit-per-call runs of the same request stream."""
    from simulador_quantico_amd import _native as N
    from simulador_quantico_amd import serving as S
    t0 = time.perf_counter()
    jobs = S.make_mix(requests, rate=rate, seed=seed)
    warm = S.make_mix(64, rate=rate, seed=seed + 1, qmax=20)
    build_s = time.perf_counter() - t0
    print(f"bench: serving mix of {requests} requests built in {build_s:.1f} s", file=sys.stderr, flush=True)
    with S.Server(np.complex64, max_active=256, hbm_budget_gib=64, autotune=autotune) as srv:
        if tile:   # multi-tile states (14 qubits and up) on another tile shape
            B, R, r, c = tile
            for n in range(14, 27):
                srv.set_tuning(n, block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=c)
        if planner is not None:   # another gate-list rewrite for every qubit count of the mix
            for n in range(10, 27):
                srv.set_tuning(n, planner=planner)
        pre = {"kernels": 0, "seconds": 0.0}
        if compiled:   # pass kernels of the mix's circuit shapes, compiled before the timed run
            t_pc = time.perf_counter()
            for n in range(14, 27):
                for ops, iters in S.mix_shapes(n):
                    pre["kernels"] += srv.precompile(n, ops, iters)
            pre["seconds"] = time.perf_counter() - t_pc
            print(f"bench: serving precompiled {pre['kernels']} pass kernels in {pre['seconds']:.1f} s",
                  file=sys.stderr, flush=True)
        S.run_served(warm, srv, time_scale=0.0)
        res = S.run_served(jobs, srv)
        t_sat = (max(r["t_done_ms"] for r in res) - min(r["t_submit_ms"] for r in res)) / 1e3
        # the same stream offered below the measured capacity (0.8 x), then at half of it: the
        # engine's latency without a standing queue
        cap_rate = len(res) / t_sat
        below_rate, half_rate = 0.8 * cap_rate, 0.5 * cap_rate
        below = jobs[:min(len(jobs), 4096)]
        res_below = S.run_served(below, srv, time_scale=rate / below_rate)
        half = jobs[:min(len(jobs), 2048)]
        res_half = S.run_served(half, srv, time_scale=rate / half_rate)
        # fidelity of FP32 results against FP64 on a sample of the same stream (untimed)
        sample = [j for j in jobs if j.qubits <= 20][:fidelity_sample]
        outs = {j.index: np.zeros(1 << j.qubits, np.complex64) for j in sample}
        for j in sample:
            srv.submit(j.qubits, j.ops, j.basis, state_out=outs[j.index])
        srv.wait_all()
        srv.poll(cap=len(sample) + 8)
    bad = [r for r in res if r["status"] != 0]
    lat = np.array([r["latency_ms"] for r in res])
    t_first = min(r["t_submit_ms"] for r in res)
    t_last = max(r["t_done_ms"] for r in res)
    served_cps = len(res) / ((t_last - t_first) / 1e3)
    by_job = {r["job"]: r for r in res}
    adders = [j for j in jobs if j.kind == "adder"]
    adder_ok = sum(by_job[j.index]["argmax"] == j.expect for j in adders)
    # state traffic of the saturated run: every pass reads and writes the whole FP32 state
    traffic_by_kind = {}
    for r in res:
        j = jobs[r["job"]]
        traffic_by_kind[j.kind] = traffic_by_kind.get(j.kind, 0) + r["passes"] * (1 << j.qubits) * 2 * 8
    traffic = sum(traffic_by_kind.values())
    print(f"bench: served {len(res)} in {(t_last - t_first) / 1e3:.2f} s", file=sys.stderr, flush=True)

    S.run_sequential(warm[:4])
    seq_jobs = jobs[:seq_sample]
    seq = S.run_sequential(seq_jobs)
    service = np.array([r["service_ms"] for r in seq])
    seq_cps = len(seq) / (service.sum() / 1e3)
    seq_lat = S.queue_latency([j.arrival_s for j in seq_jobs], service)
    agree = sum(by_job[r["job"]]["argmax"] == r["argmax"] for r in seq)
    print(f"bench: sequential {len(seq)} in {service.sum() / 1e3:.2f} s", file=sys.stderr, flush=True)

    fids = []
    for j in sample:
        ref = np.zeros(1 << j.qubits, np.complex128)
        ref[j.basis] = 1
        tun = S.sequential_tuning(j.qubits)
        st = N.Stats()
        N.check(N.lib.hgm_execute(ref.ctypes.data, j.qubits, N.HGM_C128, S._op_ptr(j.ops), len(j.ops),
                                  C.byref(tun), 1, C.byref(st)))
        fids.append(S.fidelity(ref, outs[j.index]))
    return {
        "requests": len(res), "failed": len(bad), "arrival_rate_per_s": rate, "dtype": "f32",
        "circuits_per_s": served_cps, "p50_latency_ms": float(np.percentile(lat, 50)),
        "p99_latency_ms": float(np.percentile(lat, 99)), "makespan_s": (t_last - t_first) / 1e3,
        "sequential_fp64": {"sample": len(seq), "circuits_per_s": seq_cps,
                            "p50_latency_ms": float(np.percentile(seq_lat, 50)),
                            "p99_latency_ms": float(np.percentile(seq_lat, 99)),
                            "mean_service_ms": float(service.mean()), "argmax_agree": agree},
        "speedup_circuits_per_s": served_cps / seq_cps,
        "compiled_pass_kernels": dict(pre, note="one kernel per pass structure of the mix's circuit shapes at 14-26 "
                                      "qubits, compiled before the timed run; states of at most 13 qubits and "
                                      "passes of other structures run the batched interpreter"),
        "state_traffic": {"bytes": traffic, "achieved_TBps": traffic / ((t_last - t_first) / 1e3) / 1e12,
                          "share_by_kind": {k: v / traffic for k, v in sorted(traffic_by_kind.items())},
                          "counting": "passes x 2^qubits x 8 B, read and written once per pass"},
        "at_0.8_capacity": {
            "arrival_rate_per_s": below_rate, "requests": len(res_below),
            "p50_latency_ms": float(np.percentile([r["latency_ms"] for r in res_below], 50)),
            "p99_latency_ms": float(np.percentile([r["latency_ms"] for r in res_below], 99)),
            "circuits_per_s": len(res_below) / ((max(r["t_done_ms"] for r in res_below) -
                                                 min(r["t_submit_ms"] for r in res_below)) / 1e3),
            "failed": sum(r["status"] != 0 for r in res_below)},
        "at_half_capacity": {
            "arrival_rate_per_s": half_rate, "requests": len(res_half),
            "p50_latency_ms": float(np.percentile([r["latency_ms"] for r in res_half], 50)),
            "p99_latency_ms": float(np.percentile([r["latency_ms"] for r in res_half], 99)),
            "failed": sum(r["status"] != 0 for r in res_half)},
        "adder_sums_exact": f"{adder_ok}/{len(adders)}",
        "fidelity_fp32_vs_fp64_min": float(min(fids)) if fids else None, "fidelity_sample": len(fids),
        "mix": "Poisson arrivals; qubits uniform in [10, 26]; uniform over Hadamard layer, Grover (1-64 "
               "iterations), Draper adder, random layers depth 10; seed %d" % seed,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="random30", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--planner", type=int, default=None,
                    help="HGM_PLAN_* value (with --serving: for every qubit count of the mix)")
    ap.add_argument("--ops-block", type=int, default=None)
    ap.add_argument("--tuning", default=None, help="B,R,r,c override")
    ap.add_argument("--kernel", type=int, default=None, help="0 auto, 1 LDS, 2 register, 3 plan-compiled")
    ap.add_argument("--graph", action="store_true", help="replay plan-compiled passes from a hipGraph")
    ap.add_argument("--serving", action="store_true", help="run only the serving mix and print its JSON line")
    ap.add_argument("--fidelity", action="store_true",
                    help="run only the FP32-vs-FP64 fidelity report of the benchmark circuits")
    ap.add_argument("--requests", type=int, default=4096)
    ap.add_argument("--rate", type=float, default=2000.0, help="serving mix arrival rate (requests/s)")
    ap.add_argument("--seq-sample", type=int, default=256,
                    help="requests of the same stream run by the sequential FP64 baseline")
    ap.add_argument("--autotune", action="store_true", help="serving: autotune tile shapes per qubit count")
    ap.add_argument("--autotune-tile", action="store_true",
                    help="autotune the workload's tile shape (hgm_autotune, untimed) before the timed steps")
    ap.add_argument("--serving-tile", default=None, help="serving: B,R,r,c for states of 14 qubits and more")
    ap.add_argument("--interpreted-serving", action="store_true",
                    help="serving: no compiled pass kernels (every pass on the batched interpreter)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-mix leg of the default run")
    ap.add_argument("--general-out", type=int, default=None, metavar="QUBITS",
                    help="time the reference's general.out (H x3, FP32, t_GPU, host state) at QUBITS qubits")
    ap.add_argument("--shor", type=int, default=None, metavar="N",
                    help="time attempts of Shor's algorithm factoring N through t_GPU")
    ap.add_argument("--drop-in", action="store_true",
                    help="run only the workload through the drop-in executor (PCIe-inclusive rate)")
    ap.add_argument("--no-fidelity", action="store_true", help="skip the FP32-vs-FP64 leg of the default run")
    ap.add_argument("--batch-graph", action="store_true",
                    help="time 64 resident states of 12-19 qubits as one batch graph against plan by plan")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the checks of the Hadamard, Grover and Draper configurations in the default run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="time a plain device copy of the workload's state (the known-size reference for the "
                         "FETCH_SIZE / WRITE_SIZE counters) instead of the circuit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the multi-GPU path is a "
              f"design note, not built", file=sys.stderr)
        sys.exit(2)

    import simulador_quantico_amd as hq
    from simulador_quantico_amd import _native as N

    if args.general_out:
        tun = dict(block_size=64, repeat_count=2, gpu_region_bits=8, gpu_coalesced_bits=4)   # the CLI defaults
        if args.tuning:
            B, R, r, c = (int(x) for x in args.tuning.split(","))
            tun = dict(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=c)
        if args.kernel is not None:
            tun["kernel"] = args.kernel
        print(json.dumps({"general_out": general_out(args.general_out, tun)}))
        return

    if args.shor:
        tun = dict(block_size=64, repeat_count=2, gpu_region_bits=8, gpu_coalesced_bits=4)   # the CLI defaults
        if args.tuning:
            B, R, r, c = (int(x) for x in args.tuning.split(","))
            tun = dict(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=c)
        if args.kernel is not None:
            tun["kernel"] = args.kernel
        if args.planner is not None:
            tun["planner"] = args.planner
        rows = [shor_attempts(args.shor, 5, np.complex64, False, tun), shor_attempts(args.shor, 5, np.complex64, True, tun),
                shor_attempts(args.shor, 5, np.complex128, True, tun)]
        print(json.dumps({"shor": rows}))
        return

    if args.drop_in:
        tun = None
        if args.tuning or args.kernel is not None or args.planner is not None:
            tun = dict(WORKLOADS[args.workload]["tuning"])
            if args.tuning:
                B, R, r, c = (int(x) for x in args.tuning.split(","))
                tun.update(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=c)
            if args.kernel is not None:
                tun["kernel"] = args.kernel
            if args.planner is not None:
                tun["planner"] = args.planner
        print(json.dumps({"drop_in": drop_in_rate(args.workload, tun)}))
        return

    if args.fidelity:
        print(json.dumps({"fidelity_fp32_vs_fp64": fidelity_report()}))
        return

    if args.batch_graph:
        print(json.dumps({"batch_graph": batch_graph_bench()}))
        return

    if args.serving:
        tile = tuple(int(x) for x in args.serving_tile.split(",")) if args.serving_tile else None
        out = serving_bench(args.requests, args.rate, args.seq_sample, autotune=args.autotune, tile=tile,
                            planner=args.planner, compiled=not args.interpreted_serving)
        print(json.dumps({"metric": METRIC, "value": out["circuits_per_s"], "unit": "circuits/s", "n_gpus": 1,
                          "higher_is_better": True, "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": "serving mix"}, "serving": out}))
        return

    if args.copy_calibration:
        w = WORKLOADS[args.workload]
        dtype = np.complex128 if w["dtype"] == "c128" else np.complex64
        src, dst = hq.State(w["qubits"], dtype), hq.State(w["qubits"], dtype)
        src.set_basis(0)
        dst.copy_from(src)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            dst.copy_from(src)
        wall = time.perf_counter() - t0
        nbytes = (1 << w["qubits"]) * np.dtype(dtype).itemsize
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "steps": args.steps,
                          "ms_per_copy": wall * 1e3 / args.steps,
                          "copy_GBps": 2 * nbytes * args.steps / wall / 1e9}))
        return

    w = WORKLOADS[args.workload]
    tun = dict(w["tuning"])
    if args.tuning:
        B, R, r, c = (int(x) for x in args.tuning.split(","))
        tun.update(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=c)
    if args.planner is not None:
        tun["planner"] = args.planner
    if args.ops_block is not None:
        tun["ops_block"] = args.ops_block
    if args.kernel is not None:
        tun["kernel"] = args.kernel
    n = w["qubits"]
    dtype = np.complex128 if w["dtype"] == "c128" else np.complex64
    d, ops, iterations, basis, name = build_circuit(w)
    autotuned = None
    if args.autotune_tile:   # time the kernel family's tile shapes on this circuit and keep the fastest
        best, best_ms = N.Tuning(), C.c_double()
        t_at = time.perf_counter()
        N.check(N.lib.hgm_autotune(n, N.HGM_C128 if dtype == np.complex128 else N.HGM_C64, ops, len(ops), iterations,
                                   tun.get("kernel", 0), tun.get("planner", 0), C.byref(best), C.byref(best_ms)))
        tun.update(block_size=best.block_size, repeat_count=best.repeat_count, gpu_region_bits=best.gpu_region_bits,
                   gpu_coalesced_bits=best.gpu_coalesced_bits)
        autotuned = {"seconds": time.perf_counter() - t_at, "best_kernel_ms": best_ms.value}
        print(f"bench: autotuned tile {tun} in {autotuned['seconds']:.1f} s", file=sys.stderr, flush=True)
    tuning = N.default_tuning(**tun)
    plan = hq.Plan(n, ops, tuning, iterations, dtype)
    if args.graph:
        plan.set_graph(True)
    state = hq.State(n, dtype)
    state.set_basis(basis)
    print(f"bench: {name}, {len(ops) * iterations} gates, {len(plan)} passes, tuning {tun}", file=sys.stderr,
          flush=True)

    for _ in range(args.warmup):
        plan.run(state)
    kernel_ms = 0.0
    passes = 0
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = plan.run(state)
        kernel_ms += st["kernel_ms"]
        step_ms.append(st["kernel_ms"])
        passes += st["passes"]
    wall = time.perf_counter() - t0
    gates = len(ops) * iterations
    amp_updates = args.steps * gates * float(1 << n)
    value = amp_updates / wall
    amp_bytes = 16 if dtype == np.complex128 else 8
    bytes_per_launch = 2.0 * (1 << n) * amp_bytes
    mean_launch_ms = kernel_ms / max(1, passes)
    achieved = bytes_per_launch / (mean_launch_ms * 1e-3) / 1e9
    flops = algorithmic_flops(ops, n, iterations)
    tflops = flops * args.steps / (kernel_ms * 1e-3) / 1e12
    fpeak = FP64_PEAK_TFLOPS if dtype == np.complex128 else FP32_PEAK_TFLOPS
    norm = state.norm2()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "amp-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "kernel_ms_per_step_median": float(np.median(step_ms)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if dtype == np.complex128 else "f32",
        "data": "synthetic (seeded circuit, state |0..0> resident in HBM)",
        "config": {"workload": name, "qubits": n, "gates_per_step": gates, "passes_per_step": len(plan),
                   "tile": {k: v for k, v in tun.items()}, "planner": planner_name(tun.get("planner", 0)), "plan_compile_s": plan.compile_seconds, "state_norm_after": norm,
                   "tile_autotuned": autotuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": KERNEL_NAMES.get(plan.kernel, "?"),
                     "bytes_per_launch": bytes_per_launch,
                     "mean_launch_ms": mean_launch_ms,
                     "compute": {"achieved": None, "peak": fpeak, "unit": "TFLOP/s", "frac": None,
                                 "executed_flops_per_step": None,
                                 "reference_formula": {
                                     "tflops": tflops, "flops_per_step": flops,
                                     "counting": "dense gate 14 flop/amplitude, (anti)diagonal 6 flop per changed "
                                                 "amplitude (the reference kernel's arithmetic), over all passes' "
                                                 "kernel time"}}},
    }
    traffic = measured_traffic(args.workload, tun)
    if traffic is not None:
        out["roofline"]["traffic"] = traffic["traffic_bytes_per_launch"]
        out["roofline"]["traffic_source"] = traffic["source"]
    valu = measured_traffic(args.workload, tun, "valu")
    if valu is not None:   # executed FP64 work (counters) over this run's kernel time
        executed = valu["executed_fp64_flops_per_step"] * args.steps / (kernel_ms * 1e-3) / 1e12
        out["roofline"]["compute"].update(achieved=executed, frac=executed / fpeak,
                                          executed_flops_per_step=valu["executed_fp64_flops_per_step"],
                                          executed_source=valu["source"])
        if "floors_ms" in valu:
            # per pass, the VALU issue floor at the measured clock and the HBM floor at the copy rate, summed
            # over the step: the larger sum names the resource that bounds the step
            f = valu["floors_ms"]
            out["roofline"]["floors_ms_per_step"] = dict(f, source=valu["source"])
            out["roofline"]["bound"] = "hbm" if f["hbm_at_copy_rate_ms"] >= f["valu_ms"] else "fp64-valu"
    default_run = args.tuning is None and args.kernel is None and args.planner is None and args.ops_block is None
    # the copy ceiling beside the published peak: one streaming kernel that reads and writes the
    # whole state (hgm_state_copy), median of 5 after one untimed copy
    try:
        other = hq.State(n, dtype)
        other.copy_from(state)
        copy_ms = []
        for _ in range(5):
            t0 = time.perf_counter()
            other.copy_from(state)
            copy_ms.append((time.perf_counter() - t0) * 1e3)
        del other
        ceiling = 2 * (1 << n) * np.dtype(dtype).itemsize / (np.median(copy_ms) * 1e-3) / 1e9
        out["roofline"]["copy_ceiling"] = {"GBps": ceiling, "frac": achieved / ceiling,
                                           "how": "hgm_state_copy of the state (reads and writes it once), "
                                                  "median of 5, wall clock around the synchronous call"}
    except Exception as e:
        out["roofline"]["copy_ceiling"] = {"error": str(e)}
    out["config"]["passes_per_gate"] = len(plan) / gates
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # the baseline is reported, never the target
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    del plan, state
    if default_run and not args.no_serving:
        try:
            out["serving"] = serving_bench(args.requests, args.rate, min(args.seq_sample, 128),
                                           autotune=args.autotune, compiled=not args.interpreted_serving)
        except Exception as e:
            out["serving"] = {"error": str(e)}
    if default_run and not args.no_configs:
        try:
            out["other_configs"] = other_configs()
        except Exception as e:
            out["other_configs"] = {"error": str(e)}
    if default_run and not args.no_fidelity:
        try:
            out["fidelity_fp32_vs_fp64"] = fidelity_report()
        except Exception as e:
            out["fidelity_fp32_vs_fp64"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
