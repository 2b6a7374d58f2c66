"""Host-side (no GPU) checks of the library: generators, parser, tuning
validation, error behaviour, and the exported C ABI."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import simulador_quantico_amd as hq
from simulador_quantico_amd import _native as N
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ------------------------------------------------------------- C ABI surface
def declared_functions():
    text = open(os.path.join(ROOT, "include", "hgm_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(hgm_[a-z0-9_]+|setDevice|GpuExecutionWrapper|ProjectState|GetState)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 50
    lib = C.CDLL(N.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    bound = {s[0] for s in N.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_version_and_no_device_here():
    assert b"gfx950" in N.lib.hgm_version()


# ------------------------------------------------------------- generators
def _product_gen(which, args, op, dgm):
    return hq.circuits._gen(which, args, op, dgm)


def test_generators_match_reference_text(golden):
    """Every generator output equals the reference's byte for byte."""
    d = hq.DGM()
    n = 0
    for s in golden["strings"]:
        got = _product_gen(s["which"], s["args"], s["op"], d)
        assert got == s["text"], (s["which"], s["args"])
        n += 1
    assert n > 100


def test_qft_is_the_dft():
    """QFT steps of the builder on an m-qubit register = DFT with reversed output bits."""
    m = 5
    d = hq.DGM()
    steps = hq.QFT(m, 1, 0, m - 1, d)
    o = po.Oracle(m, 64)
    d.setFunction(steps)
    t, cm, cv, mats = d.pts()
    from tests.test_oracle import oracle_run_gates
    N_ = 2 ** m
    for x in [0, 1, 5, 17, 31]:
        got = oracle_run_gates(m, t, cm, cv, mats, basis=x)
        rev = lambda k: int(format(k, f"0{m}b")[::-1], 2)  # noqa: E731
        want = np.zeros(N_, complex)
        for k in range(N_):
            want[rev(k)] = np.exp(2j * np.pi * x * k / N_) / np.sqrt(N_)
        assert np.abs(got - want).max() < 1e-12
    del o


@pytest.mark.parametrize("width", [2, 3, 4])
def test_draper_adder_adds(width):
    d = hq.DGM()
    steps = hq.DraperAdder(width, d)
    d.setFunction(steps)
    n = 2 * width + 1
    t, cm, cv, mats = d.pts()
    from tests.test_oracle import oracle_run_gates
    rng = np.random.default_rng(width)
    for _ in range(6):
        a, b = (int(v) for v in rng.integers(0, 1 << width, 2))
        basis = (a << (width + 1)) | b
        st = oracle_run_gates(n, t, cm, cv, mats, basis=basis)
        k = int(np.argmax(np.abs(st)))
        assert k == (a << (width + 1)) | (a + b)
        assert abs(abs(st[k]) - 1) < 1e-12


def test_u3_and_random_layers():
    u = np.array(hq.dgm.DGM().gates.getMatrix("H"))
    assert np.allclose(u, [2 ** -0.5, 2 ** -0.5, 2 ** -0.5, -2 ** -0.5])
    d = hq.DGM()
    steps = hq.RandomLayers(30, 20, 7, d)
    d.setFunction(steps)
    t, cm, cv, mats = d.pts()
    assert len(t) == 600 + 290              # 30 U3 per layer + 15/14 CZ
    for m in mats[:600:37]:
        U = m.reshape(2, 2)
        assert np.allclose(U.conj().T @ U, np.eye(2), atol=1e-14)
    cz = mats[np.array([bin(int(x)).count("1") for x in cm]) == 1]
    assert len(cz) == 290 and np.allclose(cz, [1, 0, 0, -1])
    # same seed, same circuit
    d2 = hq.DGM()
    assert hq.RandomLayers(30, 20, 7, d2) == steps
    assert np.array_equal(d2.gates.getMatrix("U3_19_29"), d.gates.getMatrix("U3_19_29"))


def test_grover_iteration_count():
    assert hq.GroverIterations(22) == 1608


# ------------------------------------------------------------- parser
def test_parser_matches_reference(golden):
    for case in golden["parse"]:
        d = hq.DGM()
        d.setFunction(case["text"].split(";"), case.get("iterations", 1))
        if "append" in case:
            d.setFunction(case["append"], case["append_iterations"], reset=False)
        t, cm, cv, _ = d.pts()
        assert t.tolist() == case["target_bits"], case["text"]
        assert cm.astype(np.int64).tolist() == case["control_masks"], case["text"]
        assert cv.astype(np.int64).tolist() == case["control_values"], case["text"]


def test_parser_matches_oracle_on_random_text():
    from tests.golden.make_golden import custom_gates, random_circuit
    rng = np.random.default_rng(11)
    for n in (3, 6, 9, 16):
        g = custom_gates(rng)
        steps = random_circuit(rng, n, 8)
        d = hq.DGM()
        o = po.Oracle(n, 64)
        for k, m in g.items():
            d.gates.addGate(k, m)
            o.add_gate(k, m)
        d.setFunction(steps, 2)
        o.set_function(steps, 2)
        t, cm, cv, mats = d.pts()
        t2, cm2, cv2, m2 = o.gates()
        assert t.tolist() == t2.tolist()
        assert cm.astype(np.int64).tolist() == cm2.tolist() and cv.astype(np.int64).tolist() == cv2.tolist()
        assert np.array_equal(mats, m2[:, 0::2] + 1j * m2[:, 1::2])


def test_unknown_gate_is_an_error():
    d = hq.DGM()
    with pytest.raises(hq.HGMError, match="unknown gate 'FOO'"):
        d.setFunction("H,FOO")


def test_addgate_first_writer_wins():
    d = hq.DGM()
    assert d.gates.addGate("G", [1, 0, 0, 1j])
    assert not d.gates.addGate("G", [0, 1, 1, 0])
    assert np.allclose(d.gates.getMatrix("G"), [1, 0, 0, 1j])
    assert not d.gates.addGate("H", [1, 0, 0, 1])          # built-ins are not overwritten


def test_catalog_phases_in_double():
    d = hq.DGM()
    for k in (1, 2, 3):
        assert abs(d.gates.getMatrix(f"R{k}")[3] - np.exp(1j * np.pi / 2 ** (k - 1))) < 1e-16


def test_count_ops():
    d = hq.DGM()
    d.setFunction(["H,X,Z", "Control1(1),Target1(Y),ID", "Control1(0),ID,Target1(R2)"], 2)
    c = d.CountOps(3)
    assert c["dense"] == 2 * 3 * 1 and c["sec_diag"] == 2 * 3 and c["main_diag"] == 2 * 3
    assert c["c_sec_diag"] == 6 and c["c_main_diag"] == 6 and c["total_op"] == 30


# ------------------------------------------------------------- tuning / errors
def tuning_rc(q, **kw):
    t = N.default_tuning(**kw)
    return N.lib.hgm_validate_tuning(q, N.HGM_C128, C.byref(t))


def test_tuning_contract():
    assert tuning_rc(20) == 0                                           # reference defaults
    assert tuning_rc(20, block_size=128, repeat_count=2, gpu_region_bits=9) == 0
    # src/core/dgm_core.cpp:207: 2*128*4 != 2^8 (docs/04-gpu-cuda.md:139-142)
    assert tuning_rc(20, block_size=128, repeat_count=4, gpu_region_bits=8) == 2
    assert b"2*block_size*repeat_count" in N.lib.hgm_last_error()
    assert tuning_rc(20, gpu_coalesced_bits=9) == 2                     # r < c
    assert tuning_rc(20, gpu_coalesced_bits=8) == 2                     # r == c loops forever in the reference
    assert tuning_rc(8, gpu_coalesced_bits=8) == 0                      # fine when the state is one tile
    assert tuning_rc(20, block_size=2048, repeat_count=2, gpu_region_bits=13) == 2
    assert tuning_rc(20, block_size=256, repeat_count=32, gpu_region_bits=14) == 2   # LDS (FP64)
    t = N.default_tuning(block_size=256, repeat_count=32, gpu_region_bits=14)
    assert N.lib.hgm_validate_tuning(20, N.HGM_C64, C.byref(t)) == 0    # FP32 fits 160 KiB
    assert tuning_rc(20, ops_block=0) == 2
    assert tuning_rc(3) == 0                                            # small states run as one tile


def test_compiled_kernel_tile_limits():
    """The compiled kernels move tiles above 32 KiB through LDS in halves with
    1/16 + 1/256 padding, so FP32 tiles of 2^15 fit 160 KiB for them and not
    for the interpreter."""
    def rc(dtype, kernel, r, B, R):
        t = N.default_tuning(block_size=B, repeat_count=R, gpu_region_bits=r, gpu_coalesced_bits=4, kernel=kernel)
        return N.lib.hgm_validate_tuning(26, dtype, C.byref(t))
    assert rc(N.HGM_C64, N.HGM_KERNEL_JIT, 15, 1024, 16) == 0
    assert rc(N.HGM_C64, N.HGM_KERNEL_REG, 15, 1024, 16) == 2
    assert b"160 KiB" in N.lib.hgm_last_error()
    assert rc(N.HGM_C64, N.HGM_KERNEL_JIT, 16, 1024, 32) == 2         # 2^16 FP32: 279 KiB per half
    assert rc(N.HGM_C128, N.HGM_KERNEL_JIT, 15, 1024, 16) == 2        # 2^15 FP64: 279 KiB per half
    assert rc(N.HGM_C128, N.HGM_KERNEL_JIT, 14, 512, 16) == 0         # 2^14 FP64: 139 KiB per half


def test_cpu_exec_types_are_rejected():
    d = hq.DGM()
    d.qubits = 3
    d.allocateMemory()
    for et, msg in [(hq.T_CPU, "only t_GPU"), (hq.T_PAR_CPU, "only t_GPU"), (7, "Erro exec type")]:
        d.exec_type = et
        d.setFunction("H,H,H")
        with pytest.raises(hq.HGMError, match=msg):
            d.execute(1)
    d.exec_type = hq.T_GPU
    d.gpu_count = 2
    with pytest.raises(hq.HGMError, match="gpu_count"):
        d.execute(1)


def test_circuit_width_must_match_state():
    d = hq.DGM()
    d.qubits = 4
    d.allocateMemory()
    d.setFunction("H,H,H")
    with pytest.raises(hq.HGMError, match="3 qubits but the state was allocated for 4"):
        d.execute(1)


@pytest.mark.parametrize("call,message", [
    ("ProjectState(None, 4, 2, 0, 3, 1)", "ProjectState: bad state"),
    ("ProjectState(buf, 4, 2, 0, 3, 2)", "gpu_count 2"),
    ("ProjectState(buf, 4, 3, 0, 0b10011, 1)", "fewer than region_size bits"),
    ("ProjectState(buf, 4, 2, 0b0001, 0b0011, 1)", "region_id outside the state or on a region bit"),
    ("GetState(buf, 4, 2, 1 << 4, 0b0011, 1)", "region_id outside the state or on a region bit"),
    ("GetState(buf, 4, 5, 0, 0b1111, 1)", "GetState: bad state"),
])
def test_projection_argument_errors_exit(call, message):
    """ProjectState / GetState check their arguments before any device call and,
    like the reference's error(), print and exit(1) (run in a child process)."""
    import subprocess
    import sys
    code = ("import numpy as np; from simulador_quantico_amd import _native as N; "
            "buf = np.zeros(16, np.complex64).ctypes.data; N.lib." + call)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode == 1
    assert message in r.stdout


def test_region_copy_null_arguments():
    host = np.zeros(16, np.complex64)
    assert N.lib.hgm_state_gather(None, host.ctypes.data, 4, 0, 0b11) == N.HGM_E_ARG
    assert N.lib.hgm_state_scatter(None, host.ctypes.data, 4, 0, 0b11) == N.HGM_E_ARG


def test_measure_host_path_matches_reference(golden):
    """DGM.measure on a host state follows the reference's draw order (no GPU needed)."""
    for case in golden["measure"]:
        n, v = case["qubits"], case["search_value"]
        o = po.Oracle(n, 64)
        o.set_function([",".join(["H"] * n)])
        o.set_function(hq.GroverIterationRef(n, v), case["iterations"], reset=False)
        o.set_basis(1 << (n - 1))
        o.execute(po.T_CPU)
        d = hq.DGM()
        d.qubits = n
        d.allocateMemory()
        d.state[:] = o.state()
        hq.seed_rng(case["seed"])
        bits = [d.measure(q) for q in range(1, n)]
        assert bits == case["bits"]


C_PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "hgm_gpu.h"
int main(void) {
    hgm_tuning t; hgm_default_tuning(&t);
    t.block_size = 64; t.repeat_count = 4; t.gpu_region_bits = 9; t.gpu_coalesced_bits = 2;
    t.planner = HGM_PLAN_GREEDY | HGM_PLAN_FUSE | HGM_PLAN_FOLD_PHASES; t.kernel = HGM_KERNEL_REG;
    if (hgm_validate_tuning(12, HGM_C128, &t) != HGM_OK) return 1;
    hgm_op ops[3]; memset(ops, 0, sizeof ops);
    const double s = 0.7071067811865476;
    for (int i = 0; i < 3; ++i) { ops[i].target_bit = i * 4; ops[i].m[0] = s; ops[i].m[2] = s; ops[i].m[4] = s; ops[i].m[6] = -s; }
    hgm_plan *p = NULL;
    if (hgm_plan_build(12, HGM_C128, ops, 3, 2, &t, &p) != HGM_OK) return 2;
    long long passes = (long long)hgm_plan_num_passes(p);
    hgm_plan_destroy(p);
    t.block_size = 100;
    if (hgm_validate_tuning(12, HGM_C128, &t) != HGM_E_TUNING) return 3;
    printf("%lld %s\n", passes, hgm_version());
    return 0;
}
"""


def test_header_is_plain_c_and_links(tmp_path):
    """include/hgm_gpu.h compiles as C11 and a C program links libhgm.so and plans without a GPU."""
    import subprocess
    src = tmp_path / "t.c"
    src.write_text(C_PROGRAM)
    exe = tmp_path / "t"
    lib_dir = os.path.join(ROOT, "simulador_quantico_amd")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", lib_dir, "-lhgm", f"-Wl,-rpath,{lib_dir}"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert out[0] == "1" and out[1] == "hgm-mi355x"


def test_register_kernels_keep_amplitudes_in_registers():
    """The interpreted register kernels (per-pass and batched) compile with no
    scratch memory: the amplitude array lives in VGPRs. A runtime index into it
    (for instance from merged switch cases) would move it to scratch and cost
    several times the kernel time."""
    import shutil
    import subprocess
    if not shutil.which("hipcc") and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("no hipcc")
    out = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "simulador_quantico_amd", "csrc"), "resource-usage"],
                         check=True, capture_output=True, text=True, timeout=600)
    usage, name = {}, None
    for line in (out.stdout + out.stderr).splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            usage[name] = int(m.group(1))
    reg = {k: v for k, v in usage.items() if "k_pass_reg" in k or "k_batch_reg" in k}
    # k_pass_reg and k_batch_reg, each for 2, 4, 8, 16 and 32 amplitudes per thread, FP32 and FP64
    assert len(reg) == 20, sorted(usage)
    assert all(v == 0 for v in reg.values()), reg


def test_gpu_entry_points_fail_cleanly_without_a_device():
    """Without a visible device the serving and autotune entry points return an error, not a crash."""
    if N.device_count() > 0:
        pytest.skip("a device is visible")
    cfg = N.ServerConfig()
    N.lib.hgm_server_default_config(C.byref(cfg))
    h = C.c_void_p()
    rc = N.lib.hgm_server_create(C.byref(cfg), C.byref(h))
    assert rc != 0 and b"device" in N.lib.hgm_last_error()
    ops = hq.make_ops([0], [0], [0], [np.eye(2)])
    best = N.Tuning()
    assert N.lib.hgm_autotune(8, N.HGM_C128, ops, 1, 1, N.HGM_KERNEL_REG, N.HGM_PLAN_GREEDY, C.byref(best), None) != 0
    bad = N.ServerConfig()
    N.lib.hgm_server_default_config(C.byref(bad))
    bad.max_active = 0
    assert N.lib.hgm_server_create(C.byref(bad), C.byref(h)) == 1   # HGM_E_ARG: checked before any device call


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
def test_print_pts_matches_reference(capfd, golden):
    """printPTs prints what the reference's printPTs prints, byte for byte, on
    the golden circuits (the parsed PT lists, controls included)."""
    for c in golden["circuits"]:
        d = hq.DGM(np.complex64)
        d.qubits = c["qubits"]
        r = po.RefDGM(c["qubits"])
        for name, m in c["gates"].items():
            d.gates.addGate(name, [complex(a, b) for a, b in m])
            r.add_gate(name, [complex(a, b) for a, b in m])
        d.setFunction(c["steps"], c["iterations"])
        r.set_function(c["steps"], c["iterations"])
        capfd.readouterr()
        d.printPTs()
        ours = capfd.readouterr().out
        r.lib.ref_print_pts(r.h)
        theirs = capfd.readouterr().out
        assert ours == theirs and ours, c["name"]


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
def test_colapse_matches_reference(capfd):
    """colapse on an FP32 host state: the same amplitudes bit for bit and the
    same printed probability as the reference's DGM::colapse."""
    rng = np.random.default_rng(31)
    n = 9
    for qubit, value in [(0, 1), (4, 0), (8, 1)]:
        psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
        psi /= np.float32(np.linalg.norm(psi))
        d = hq.DGM(np.complex64)
        d.qubits = n
        d.allocateMemory()
        d.state[:] = psi
        r = po.RefDGM(n)
        r.set_state(psi)
        capfd.readouterr()
        d.colapse(qubit, value)
        ours = capfd.readouterr().out
        r.lib.ref_colapse(r.h, qubit, value)
        theirs = capfd.readouterr().out
        assert ours == theirs and ours.strip()
        assert np.array_equal(d.state.view(np.uint32), r.state().view(np.uint32)), (qubit, value)


def test_structure_setters_validate_and_erase():
    """setExecType / setCpuStructure / setGpuStructure / validateTuning / erase
    (include/core/dgm.h:113-116, src/core/dgm_core.cpp:30-52,177-241,302-318)."""
    d = hq.DGM()
    d.qubits = 14
    d.setExecType(hq.T_GPU)
    d.setCpuStructure(12, 9)
    assert (d.exec_type, d.cpu_region_bits, d.cpu_coalesced_bits) == (hq.T_GPU, 12, 9)
    d.setGpuStructure(3, 12, 8)
    assert (d.gpu_coalesced_bits, d.gpu_region_bits, d.repeat_count, d.block_size) == (3, 12, 8, 256)
    d.validateTuning()
    d.block_size = 128
    with pytest.raises(hq.HGMError, match="2\\*block_size\\*repeat_count"):
        d.validateTuning()
    d.setFunction("H,H,H,H,H,H,H,H,H,H,H,H,H,H")
    assert len(d.pts()[0]) == 14
    d.erase()
    assert len(d.pts()[0]) == 0


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
def test_measure_many_matches_reference():
    """measure(list of qubits) on an FP32 host state returns the reference's
    map<long, float> exactly: keys are the masked index values, values the float sums."""
    rng = np.random.default_rng(77)
    n = 10
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    psi /= np.float32(np.linalg.norm(psi))
    d = hq.DGM(np.complex64)
    d.qubits = n
    d.allocateMemory()
    d.state[:] = psi
    r = po.RefDGM(n)
    r.set_state(psi)
    for qs in ([0], [3, 7], [9, 0, 5], list(range(n))):
        ours = d.measure(qs)
        q = np.asarray(qs, np.int32)
        keys = np.zeros(1 << n, np.int64)
        vals = np.zeros(1 << n, np.float32)
        k = r.lib.ref_measure_many(r.h, q.ctypes.data, len(qs), keys.ctypes.data, vals.ctypes.data, 1 << n)
        theirs = {int(keys[i]): float(vals[i]) for i in range(k)}
        assert ours == theirs, qs

def _serving_sources(n, ops, dtype=np.complex64):
    arr = (N.Op * len(ops)).from_buffer_copy(ops.tobytes())
    t = N.default_tuning(block_size=256, repeat_count=8, gpu_region_bits=12, gpu_coalesced_bits=4,
                         kernel=N.HGM_KERNEL_REG, planner=N.HGM_PLAN_GREEDY | N.HGM_PLAN_FUSE, ops_block=1 << 30)
    p = hq.Plan(n, arr, t, 1, dtype)
    out = []
    for i in range(len(p)):
        ln = C.c_size_t()
        if N.lib.hgm_plan_serving_source(p._h, i, None, 0, C.byref(ln)) != 0:
            continue
        b = C.create_string_buffer(ln.value + 1)
        N.check(N.lib.hgm_plan_serving_source(p._h, i, b, ln.value + 1, C.byref(ln)))
        out.append(b.value.decode())
    return out


def test_serving_kernels_depend_on_the_circuit_shape_only():
    """The serving engine's compiled pass kernels hold no matrix entry or
    control value: circuits of one shape (another seed, another marked value)
    have the same kernel sources, other shapes and dtypes do not, and the
    values are read from the register program (ops[i].m, .rval, .gval)."""
    from simulador_quantico_amd import serving as S
    n = 16
    g = [np.concatenate([S._circuit_ops([hq.Hadamard(n, 0, n)])] +
                        [S._circuit_ops(hq.GroverIteration(n, v))] * 3) for v in (5, 40000)]
    r = []
    for seed in (1, 2):
        d = hq.DGM()
        r.append(S._circuit_ops(hq.RandomLayers(n, 10, seed, d), d))
    assert not np.array_equal(g[0]["m"], g[1]["m"]) or not np.array_equal(g[0]["control_value"], g[1]["control_value"])
    assert not np.array_equal(r[0]["m"], r[1]["m"])
    sg = [_serving_sources(n, x) for x in g]
    sr = [_serving_sources(n, x) for x in r]
    assert len(sg[0]) > 2 and sg[0] == sg[1]
    assert len(sr[0]) > 1 and sr[0] == sr[1]
    assert sr[0] != _serving_sources(n, r[0], np.complex128)
    d = hq.DGM()
    assert sr[0][0] != _serving_sources(n + 1, S._circuit_ops(hq.RandomLayers(n + 1, 10, 1, d), d))[0]
    src = "".join(sr[0] + sg[0])
    assert "ops[0].m[0]" in src and ".rval" in src and ".gval" in src
    assert "0x1.6a09e6" not in src and "0x1.6a09e667f3bcdp-1" not in src     # no 1/sqrt(2) literal
