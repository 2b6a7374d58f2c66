"""The GPU side of t_HYBRID through the C ABI: region gather/scatter
(hgm_state_gather / hgm_state_scatter), ProjectState -> GpuExecutionWrapper(NULL,
...) -> GetState over every region of a state, the control test of the
reference's kernel on ill-formed PTs, and the reference's own DGM
(t_GPU and t_HYBRID, src/core/dgm_core.cpp:243-267, src/core/dgm_par_exec.cpp:259-476)
linked against libhgm.so in place of its GPU stub.

Bars: region copies are bit-exact (they move bytes); FP32 states against the
FP64 oracle or the reference's FP32 t_CPU within 2e-5 max-abs."""
import ctypes as C
import sys
import time

import numpy as np
import pytest

import simulador_quantico_amd as hq
from simulador_quantico_amd import _native as N
from oracle import pyoracle as po
from tests.test_gpu_parity import RefPT
from tests.test_oracle import oracle_run_gates

pytestmark = pytest.mark.gpu

TOL32 = 2e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_of(mask):
    return [b for b in range(64) if (mask >> b) & 1]


def region_index(region_id, mask):
    """Host index of every region amplitude j: region_id | deposit(j, mask),
    the bits of j placed on the set bits of mask, lowest first."""
    b = bits_of(mask)
    j = np.arange(1 << len(b), dtype=np.uint64)
    out = np.full(j.shape, region_id, dtype=np.uint64)
    for k, bit in enumerate(b):
        out |= ((j >> np.uint64(k)) & np.uint64(1)) << np.uint64(bit)
    return out.astype(np.int64)


def compress(x, mask):
    """Inverse of deposit on mask's bits: bit k of the result is the k-th set bit of mask in x."""
    r = 0
    for k, bit in enumerate(bits_of(mask)):
        r |= ((x >> bit) & 1) << k
    return r


def subsets(mask):
    """All subsets of mask in increasing order (the reference's region walk)."""
    s = 0
    while True:
        yield s
        s = (s - mask) & mask
        if s == 0:
            return


def random_host(rng, n, dtype):
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return a.astype(dtype)


# (host qubits, region mask, region id, dtype): gaps in the mask, a mask without
# bit 0 (one-amplitude portions), the whole state, and regions of several
# staging chunks in FP32 and FP64 (32, 64 and 256 MiB)
REGION_CASES = [
    (12, 0b010100111111, 0b001001000000, np.complex64),
    (12, 0b101010101010, 0b000000010001, np.complex128),
    (10, 0b1111111111, 0, np.complex64),
    (25, ((1 << 15) - 1) | (1 << 16) | (1 << 17) | (1 << 19) | (1 << 20) | (1 << 22) | (1 << 23) | (1 << 24),
     (1 << 18) | (1 << 21), np.complex64),
    (24, ((1 << 24) - 1) & ~((1 << 2) | (1 << 13)), 1 << 13, np.complex128),
    (27, ((1 << 27) - 1) & ~((1 << 20) | (1 << 26)), 1 << 20, np.complex64),
]


@pytest.mark.parametrize("n,mask,rid,dtype", REGION_CASES)
def test_region_gather_scatter_bit_exact(n, mask, rid, dtype):
    say("region copy", n, hex(mask), hex(rid), np.dtype(dtype).name)
    rng = np.random.default_rng(n * 7 + len(bits_of(mask)))
    host = random_host(rng, n, dtype)
    k = len(bits_of(mask))
    idx = region_index(rid, mask)
    s = hq.State(k, dtype)
    s.gather(host, rid, mask)
    assert np.array_equal(s.download(), host[idx])
    t0 = time.perf_counter()
    s.gather(host, rid, mask)              # timed again, staging set up
    t1 = time.perf_counter()
    y = random_host(rng, k, dtype)
    tu = time.perf_counter()
    s.upload(y)
    tu = time.perf_counter() - tu
    out = host.copy()
    t2 = time.perf_counter()
    s.scatter(out, rid, mask)
    t3 = time.perf_counter()
    nbytes = idx.size * np.dtype(dtype).itemsize
    say(f"  {nbytes / 2**20:.1f} MiB region: gather {nbytes / (t1 - t0) / 1e9:.1f} GB/s, "
        f"scatter {nbytes / (t3 - t2) / 1e9:.1f} GB/s, contiguous upload {nbytes / tu / 1e9:.1f} GB/s")
    want = host.copy()
    want[idx] = y
    assert np.array_equal(out, want)


def test_region_copy_argument_errors():
    s = hq.State(4, np.complex64)
    host = np.zeros(1 << 8, np.complex64)
    with pytest.raises(hq.HGMError, match="region_mask"):
        s.gather(host, 0, 0b111)               # 3 bits for a 4-qubit state
    with pytest.raises(hq.HGMError, match="region_mask"):
        s.gather(host, 0b1, 0b1111)            # id on a region bit
    with pytest.raises(hq.HGMError, match="region_mask"):
        s.scatter(host, 1 << 8, 0b1111)        # id outside the state
    with pytest.raises(hq.HGMError, match="contiguous"):
        s.gather(host.astype(np.complex128), 0, 0b1111)


def _pts(gates):
    """NULL-terminated PT** for [(target, cmask, cvalue, 2x2 complex64)]."""
    keep = []
    pts = (C.POINTER(RefPT) * (len(gates) + 1))()
    for i, (t, cm, cv, m) in enumerate(gates):
        m = np.ascontiguousarray(np.asarray(m, np.complex64).reshape(4))
        p = RefPT(qubits=1, matrix=m.ctypes.data, matrix_size=2, span_start_bit=t + 1, target_bit=t,
                  affected=False, control_value=cv, control_mask=cm, control_count=bin(cm).count("1"))
        keep += [m, p]
        pts[i] = C.pointer(p)
    pts[len(gates)] = None
    return pts, keep


def _hybrid_gates(rng, n, mask, count):
    """Random gates whose targets lie in mask, with controls anywhere."""
    inside = bits_of(mask)
    t, cm, cv, mats = [], [], [], []
    for _ in range(count):
        tb = int(rng.choice(inside))
        others = [b for b in range(n) if b != tb]
        ctrl = rng.choice(others, size=int(rng.integers(0, 3)), replace=False)
        z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        u = np.linalg.qr(z)[0].astype(np.complex64).astype(np.complex128)
        t.append(tb)
        cm.append(sum(1 << int(b) for b in ctrl))
        cv.append(sum(int(rng.integers(0, 2)) << int(b) for b in ctrl))
        mats.append(u.reshape(4))
    return t, cm, cv, mats


@pytest.mark.parametrize("n,mask,region_size", [
    (13, 0b1101101001111, 9),          # low run of 4, gaps: 16 regions
    (12, 0x7fff, 12),                  # mask past the state (HybridExecution below 15 qubits)
])
def test_project_execute_get_every_region(n, mask, region_size):
    """HybridExecution's GPU leg over every region: gates whose controls outside
    the region match region_id, renumbered onto the region, run on the projected
    state; the regions together give the full circuit (FP64 oracle)."""
    say("hybrid regions", n, hex(mask))
    rng = np.random.default_rng(5150 + n)
    inside = mask & ((1 << n) - 1)
    t, cm, cv, mats = _hybrid_gates(rng, n, inside, 60)
    psi0 = random_host(rng, n, np.complex64)
    psi0 /= np.linalg.norm(psi0)
    want = oracle_run_gates(n, t, cm, cv, mats, state=psi0.astype(np.complex128))
    state = psi0.copy()
    outside = ((1 << n) - 1) & ~inside
    for rid in subsets(outside):
        proj = []
        for tb, m_, v_, u in zip(t, cm, cv, mats):
            if (rid & m_ & outside) != (v_ & outside):
                continue
            proj.append((compress(1 << tb, inside).bit_length() - 1, compress(m_ & inside, inside),
                         compress(v_ & inside, inside), u))
        pts, keep = _pts(proj)
        assert N.lib.ProjectState(state.ctypes.data, n, region_size, rid, mask, 1) is True
        ret = N.lib.GpuExecutionWrapper(None, C.cast(pts, C.c_void_p), region_size, 4, 8, 1, 64, 2, 1)
        assert not ret
        assert N.lib.GetState(state.ctypes.data, n, region_size, rid, mask, 1) is True
    assert np.abs(state - want).max() < TOL32


def test_kernel_control_test_on_ill_formed_pts():
    """The reference's kernel applies a gate where (pos0 & control_mask) ==
    control_value, pos0 having the target bit clear (src/core/kernel.cu:102-108):
    a value bit outside the mask, a 1 asked of the target bit or of a bit beyond
    the state never fires; a 0 asked of those is no condition; a zero mask
    ignores the value (src/core/common.cpp:53-65)."""
    n = 8
    rng = np.random.default_rng(99)
    psi0 = random_host(rng, n, np.complex64)
    u = np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))[0]
    u = u.astype(np.complex64).astype(np.complex128)
    never = [(3, 1 << 2, (1 << 2) | (1 << 5), u),        # value bit outside the mask
             (1, (1 << 3) | (1 << 1), 1 << 1, u),        # 1 on the target bit
             (0, 1 << 9, 1 << 9, u)]                     # 1 beyond the state
    state = psi0.copy()
    pts, keep = _pts(never)
    N.lib.GpuExecutionWrapper(state.ctypes.data, C.cast(pts, C.c_void_p), n, 4, 8, 1, 64, 2, 1)
    assert np.array_equal(state, psi0)
    odd = [(4, 0, 0b101, u),                              # zero mask: no condition
           (1, (1 << 3) | (1 << 1), 1 << 3, u),           # 0 on the target bit: control on bit 3 only
           (6, (1 << 9) | (1 << 2), 1 << 2, u)]           # 0 beyond the state: control on bit 2 only
    state = psi0.copy()
    pts, keep = _pts(odd)
    N.lib.GpuExecutionWrapper(state.ctypes.data, C.cast(pts, C.c_void_p), n, 4, 8, 1, 64, 2, 1)
    want = oracle_run_gates(n, [4, 1, 6], [0, 1 << 3, 1 << 2], [0, 1 << 3, 1 << 2], [u.reshape(4)] * 3,
                            state=psi0.astype(np.complex128))
    assert np.abs(state - want).max() < TOL32


def _low_control_circuit(rng, n, n_steps, names):
    """Random steps in the reference's language whose controls sit on state bits
    below 15 (qubits n-15..n-1). HybridExecution renumbers controls on region
    bits from 15 up without carrying their values over
    (src/core/dgm_par_exec.cpp:426-439), so circuits with such controls differ
    from t_CPU under the reference's own kernel as well."""
    low = set(range(max(0, n - 15), n))
    steps = []
    for _ in range(n_steps):
        tokens = ["ID"] * n
        free = list(rng.permutation(n))
        group = 1
        while free:
            q = free.pop()
            ctrl_ok = [c for c in free if c in low]
            if rng.random() < 0.35 and ctrl_ok:
                c = ctrl_ok[0]
                free.remove(c)
                tokens[c] = f"Control{group}({int(rng.integers(0, 2))})"
                tokens[q] = f"Target{group}({rng.choice(names)})"
                group += 1
            else:
                tokens[q] = str(rng.choice(names))
        steps.append(",".join(tokens))
    return steps


@pytest.mark.skipif(not po.ref_available(po.REF_HGM_SO),
                    reason="oracle/_ref/libhgmref_hgm.so not built (needs the reference tree)")
@pytest.mark.parametrize("n,threads", [(21, 1), (21, 3), (16, 2), (23, 8)])
def test_reference_engine_linked_against_libhgm(n, threads):
    """The reference's DGM with this library in place of its GPU stub: t_GPU and
    t_HYBRID (thread 0 drives the GPU through ProjectState /
    GpuExecutionWrapper(NULL) / GetState, the other threads run the reference's
    CPU regions) against the reference's t_CPU on the same circuit."""
    say("reference engine + libhgm", n, threads)
    from tests.golden.make_golden import custom_gates
    rng = np.random.default_rng(2026 + n + threads)
    g = custom_gates(rng)
    steps = _low_control_circuit(rng, n, 6, ["H", "X", "Z", "R1", "U0", "U1", "U2", "D0", "A0"])
    psi0 = random_host(rng, n, np.complex64)
    psi0 /= np.linalg.norm(psi0)
    out = {}
    for exec_type in (po.T_CPU, po.T_GPU, po.T_HYBRID):
        r = po.RefDGM(n, po.REF_HGM_SO)
        for k, m in g.items():
            r.add_gate(k, m)
        r.set_function(steps)
        r.set_state(psi0)
        r.execute(exec_type, threads=threads, coalesced_bits=11, region_bits=14)
        out[exec_type] = r.state()
    assert np.abs(out[po.T_GPU] - out[po.T_CPU]).max() < TOL32
    assert np.abs(out[po.T_HYBRID] - out[po.T_CPU]).max() < TOL32


@pytest.mark.skipif(not po.ref_available(po.REF_HGM_SO),
                    reason="oracle/_ref/libhgmref_hgm.so not built (needs the reference tree)")
def test_reference_smoke_checks_on_libhgm(capfd):
    """The GPU half of the reference's own smoke test (tests/smoke_test.sh:86-101,114-118),
    through the reference's drivers linked against libhgm.so: general.out's
    uniform amplitude 1/sqrt(2^n) (|error| < 1e-4, :26) under t_GPU at 10/14/18
    qubits and t_HYBRID at 10/14; Grover at 12 qubits finds search_value 10 in
    at least 7 of 10 seeded runs; Shor at 15 qubits (N = 57) succeeds at least
    once in 16 seeded attempts. t_GPU runs with gpu_count 1, t_HYBRID with 2 threads."""
    lib = po._load_ref(po.REF_HGM_SO)
    for n, exec_type, threads in [(10, po.T_GPU, 1), (14, po.T_GPU, 1), (18, po.T_GPU, 1),
                                  (10, po.T_HYBRID, 2), (14, po.T_HYBRID, 2)]:
        capfd.readouterr()
        lib.ref_general_hadamard(n, exec_type, threads)
        line = capfd.readouterr().out.strip().splitlines()[-1]
        say("general", n, exec_type, line)
        assert abs(float(line.split()[1]) - 2.0 ** (-n / 2)) < 1e-4
    for exec_type, threads in [(po.T_GPU, 1), (po.T_HYBRID, 2)]:
        found = [lib.ref_grover_run(12, 10, exec_type, threads, seed) for seed in range(10)]
        say("grover", exec_type, found)
        assert sum(v == 10 for v in found) >= 7
    out = (C.c_int * 2)()
    for exec_type, threads in [(po.T_GPU, 1), (po.T_HYBRID, 2)]:
        tries = []
        for seed in range(16):
            lib.ref_shor_run(57, exec_type, threads, seed, out)
            tries.append(tuple(out))
            if out[0]:
                break
        say("shor", exec_type, tries)
        assert sorted(tries[-1]) == [3, 19]


@pytest.mark.skipif(not po.ref_available(po.REF_HGM_SO),
                    reason="oracle/_ref/libhgmref_hgm.so not built (needs the reference tree)")
def test_reference_engine_golden_circuits_on_libhgm(golden):
    """Every golden circuit through the reference's own DGM with t_GPU and t_HYBRID
    (2 threads) backed by libhgm.so, against the FP32 states the reference's t_CPU
    wrote into the fixtures (tests/golden/states.npz). Below 8 qubits the tile is
    the whole state (2^(n-1) threads of one pair each), since the reference's
    validateTuning clamps gpu_region_bits to the qubit count and then requires
    2*block_size*repeat_count == 2^gpu_region_bits."""
    for c in golden["circuits"]:
        ref32 = golden["states"][f"{c['name']}__tcpu"]
        for exec_type, threads in [(po.T_GPU, 1), (po.T_HYBRID, 2)]:
            say("golden via reference engine", c["name"], exec_type)
            r = po.RefDGM(c["qubits"], po.REF_HGM_SO)
            if c["qubits"] < 8:   # the CLI default tile (2^8) is larger than the state: validateTuning exits
                n = c["qubits"]
                r.set_gpu_tuning(n, min(4, n), 1 << (n - 1), 1)
            for name, m in c["gates"].items():
                r.add_gate(name, [complex(a, b) for a, b in m])
            r.set_function(c["steps"], c["iterations"])
            r.set_basis(c["basis"])
            r.execute(exec_type, threads=threads, coalesced_bits=11, region_bits=14, iterations=1)
            assert np.abs(r.state() - ref32).max() < TOL32, (c["name"], exec_type)


@pytest.mark.skipif(not po.ref_available(po.REF_HGM_SO),
                    reason="oracle/_ref/libhgmref_hgm.so not built (needs the reference tree)")
def test_reference_engine_with_auto_tuning(capfd):
    """HGM_TUNING=auto (hgm_auto_tuning: plan-compiled kernels, fusion, the rotation split,
    the measured tile) behind the reference's own engine: general.out at 30 qubits
    (src/cli/general.cpp:42-52) and Grover at 22 qubits (src/algorithms/lib_grover.cpp:15-66)
    give the same results as with the reference's CLI tuning; the warm times of both are
    printed (the first call of each compiles the pass kernels)."""
    import os
    import time
    lib = po._load_ref(po.REF_HGM_SO)
    times = {}
    try:
        for mode in ("cli", "auto"):
            if mode == "auto":
                os.environ["HGM_TUNING"] = "auto"
            else:
                os.environ.pop("HGM_TUNING", None)
            for rep in range(2):
                capfd.readouterr()
                t0 = time.perf_counter()
                lib.ref_general_hadamard(30, po.T_GPU, 1)
                dt = time.perf_counter() - t0
                line = capfd.readouterr().out.strip().splitlines()[-1]
                assert abs(float(line.split()[1]) - 2.0 ** -15) < 1e-6, (mode, line)
                times[("general30", mode, rep)] = dt
            for rep in range(2):
                t0 = time.perf_counter()
                found = lib.ref_grover_run(22, 1234567 % (1 << 21), po.T_GPU, 1, 7 + rep)
                times[("grover22", mode, rep)] = time.perf_counter() - t0
                assert found == 1234567 % (1 << 21), (mode, found)
    finally:
        os.environ.pop("HGM_TUNING", None)
    say("reference engine, warm seconds (CLI tuning -> HGM_TUNING=auto):",
        {w: (round(times[(w, "cli", 1)], 3), round(times[(w, "auto", 1)], 3)) for w in ("general30", "grover22")},
        "first calls:", {w: (round(times[(w, "cli", 0)], 3), round(times[(w, "auto", 0)], 3))
                         for w in ("general30", "grover22")})
