"""ctypes bindings of libhgm.so (the C ABI declared in include/hgm_gpu.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C simulador_quantico_amd/csrc``). Importing this module without the
library raises ImportError: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhgm.so")

HGM_OK = 0
HGM_C128 = 0
HGM_C64 = 1
HGM_PLAN_GREEDY = 0
HGM_PLAN_WINDOW = 1
HGM_PLAN_FOLD_PHASES = 16
HGM_PLAN_FUSE = 32
HGM_PLAN_PIVOT = 64
HGM_PLAN_BALANCE = 128
HGM_PLAN_LOOKAHEAD = 256
HGM_PLAN_ROTATE = 512
HGM_KERNEL_AUTO = 0
HGM_KERNEL_LDS = 1
HGM_KERNEL_REG = 2
HGM_KERNEL_JIT = 3

HGM_E_ARG, HGM_E_TUNING, HGM_E_HIP, HGM_E_NOMEM, HGM_E_PARSE, HGM_E_STATE, HGM_E_UNSUPPORTED = range(1, 8)

ERROR_NAMES = {1: "HGM_E_ARG", 2: "HGM_E_TUNING", 3: "HGM_E_HIP", 4: "HGM_E_NOMEM", 5: "HGM_E_PARSE",
               6: "HGM_E_STATE", 7: "HGM_E_UNSUPPORTED"}


class HGMError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code


class Op(C.Structure):
    _fields_ = [("target_bit", C.c_int32), ("reserved", C.c_int32), ("control_mask", C.c_uint64),
                ("control_value", C.c_uint64), ("m", C.c_double * 8)]


class Tuning(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("repeat_count", C.c_int32), ("gpu_region_bits", C.c_int32),
                ("gpu_coalesced_bits", C.c_int32), ("ops_block", C.c_int32), ("planner", C.c_int32),
                ("kernel", C.c_int32), ("lookahead", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("gates", C.c_int64), ("passes", C.c_int64), ("amp_updates", C.c_int64), ("bytes_alg", C.c_double),
                ("plan_ms", C.c_double), ("h2d_ms", C.c_double), ("kernel_ms", C.c_double), ("d2h_ms", C.c_double),
                ("total_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("max_active", C.c_int32), ("hbm_budget_bytes", C.c_uint64),
                ("planner_threads", C.c_int32), ("autotune", C.c_int32), ("min_qubits", C.c_int32),
                ("max_qubits", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("qubits", C.c_int32), ("iterations", C.c_int32), ("ops", C.POINTER(Op)), ("n_ops", C.c_int64),
                ("basis", C.c_uint64), ("state_out", C.c_void_p)]


class Result(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("passes", C.c_int32), ("argmax", C.c_uint64),
                ("p_max", C.c_double), ("norm2", C.c_double), ("t_submit_ms", C.c_double),
                ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


assert C.sizeof(Op) == 88 and C.sizeof(Tuning) == 32
assert C.sizeof(ServerConfig) == 32 and C.sizeof(Request) == 40 and C.sizeof(Result) == 64

# (name, restype, argtypes) of every exported entry point of include/hgm_gpu.h
_P = C.c_void_p
SIGNATURES = [
    ("hgm_default_tuning", None, [C.POINTER(Tuning)]),
    ("hgm_last_error", C.c_char_p, []),
    ("hgm_version", C.c_char_p, []),
    ("hgm_init", C.c_int, [C.c_int]),
    ("hgm_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("hgm_release_cache", C.c_int, []),
    ("hgm_execute", C.c_int, [_P, C.c_int, C.c_int, C.POINTER(Op), C.c_int64, C.POINTER(Tuning), C.c_int,
                              C.POINTER(Stats)]),
    ("hgm_state_create", C.c_int, [C.c_int, C.c_int, C.POINTER(_P)]),
    ("hgm_state_wrap", C.c_int, [_P, C.c_int, C.c_int, C.POINTER(_P)]),
    ("hgm_state_destroy", None, [_P]),
    ("hgm_state_qubits", C.c_int, [_P]),
    ("hgm_state_dtype", C.c_int, [_P]),
    ("hgm_state_device_ptr", _P, [_P]),
    ("hgm_state_set_basis", C.c_int, [_P, C.c_uint64]),
    ("hgm_state_upload", C.c_int, [_P, _P]),
    ("hgm_state_download", C.c_int, [_P, _P]),
    ("hgm_state_gather", C.c_int, [_P, _P, C.c_int, C.c_uint64, C.c_uint64]),
    ("hgm_state_scatter", C.c_int, [_P, _P, C.c_int, C.c_uint64, C.c_uint64]),
    ("hgm_state_copy", C.c_int, [_P, _P]),
    ("hgm_state_apply", C.c_int, [_P, C.POINTER(Op), C.c_int64, C.POINTER(Tuning), C.c_int, C.POINTER(Stats)]),
    ("hgm_state_norm2", C.c_int, [_P, C.POINTER(C.c_double)]),
    ("hgm_state_apply_noalloc", C.c_int, [_P, C.POINTER(Op), C.c_int64, C.POINTER(Tuning), C.c_int,
                                          C.POINTER(Stats)]),
    ("hgm_state_prob_one", C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    ("hgm_state_prob_halves", C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("hgm_state_collapse_reset", C.c_int, [_P, C.c_int, C.c_int, C.c_double]),
    ("hgm_state_argmax", C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    ("hgm_state_compare", C.c_int, [_P, _P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("hgm_plan_build", C.c_int, [C.c_int, C.c_int, C.POINTER(Op), C.c_int64, C.c_int, C.POINTER(Tuning),
                                 C.POINTER(_P)]),
    ("hgm_plan_destroy", None, [_P]),
    ("hgm_plan_num_passes", C.c_int64, [_P]),
    ("hgm_plan_pass", C.c_int, [_P, C.c_int64, C.POINTER(C.c_uint64), C.POINTER(C.c_int64), _P]),
    ("hgm_plan_shape", C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32)]),
    ("hgm_state_run_plan", C.c_int, [_P, _P, C.POINTER(Stats)]),
    ("hgm_plan_kernel", C.c_int, [_P]),
    ("hgm_plan_serving_source", C.c_int, [_P, C.c_int64, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("hgm_plan_program", C.c_int, [_P, C.c_int64, _P, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), _P]),
    ("hgm_plan_maps", C.c_int, [_P, _P, C.POINTER(C.c_int64)]),
    ("hgm_plan_jit_source", C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("hgm_plan_compile_seconds", C.c_double, [_P]),
    ("hgm_plan_set_graph", C.c_int, [_P, C.c_int]),
    ("hgm_batch_create", C.c_int, [C.POINTER(_P), C.POINTER(_P), C.c_int, C.POINTER(_P)]),
    ("hgm_batch_run", C.c_int, [_P, C.POINTER(Stats)]),
    ("hgm_batch_destroy", None, [_P]),
    ("hgm_validate_tuning", C.c_int, [C.c_int, C.c_int, C.POINTER(Tuning)]),
    ("hgm_fold_phases", C.c_int, [C.c_int, C.POINTER(Op), C.c_int64, C.POINTER(Op), C.c_int64,
                                  C.POINTER(C.c_int64)]),
    ("hgm_fuse_gates", C.c_int, [C.c_int, C.POINTER(Op), C.c_int64, C.POINTER(Op), C.c_int64,
                                 C.POINTER(C.c_int64)]),
    ("hgm_pivot_rows", C.c_int, [C.c_int, C.POINTER(Op), C.c_int64, C.POINTER(Op), C.c_int64,
                                 C.POINTER(C.c_int64)]),
    ("hgm_rotate_rows", C.c_int, [C.c_int, C.POINTER(Op), C.c_int64, C.POINTER(Op), C.c_int64,
                                 C.POINTER(C.c_int64)]),
    ("hgm_dgm_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("hgm_dgm_destroy", None, [_P]),
    ("hgm_dgm_set", C.c_int, [_P, C.c_char_p, C.c_int64]),
    ("hgm_dgm_get", C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int64)]),
    ("hgm_dgm_allocate_memory", C.c_int, [_P]),
    ("hgm_dgm_free_memory", C.c_int, [_P]),
    ("hgm_dgm_set_memory_value", C.c_int, [_P, C.c_int64]),
    ("hgm_dgm_state", _P, [_P]),
    ("hgm_dgm_add_gate", C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    ("hgm_dgm_get_gate", C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    ("hgm_dgm_set_function", C.c_int, [_P, C.c_char_p, C.c_int, C.c_int]),
    ("hgm_dgm_set_function_steps", C.c_int, [_P, C.POINTER(C.c_char_p), C.c_int, C.c_int, C.c_int]),
    ("hgm_dgm_num_pts", C.c_int, [_P, C.POINTER(C.c_int64)]),
    ("hgm_dgm_get_ops", C.c_int, [_P, C.POINTER(Op), C.c_int64]),
    ("hgm_dgm_execute", C.c_int, [_P, C.c_int]),
    ("hgm_dgm_execute_function", C.c_int, [_P, C.c_char_p, C.c_int]),
    ("hgm_dgm_measure", C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    ("hgm_dgm_colapse", C.c_int, [_P, C.c_int, C.c_int]),
    ("hgm_dgm_measure_many", C.c_int, [_P, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int64),
                                       C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    ("hgm_dgm_sync_host", C.c_int, [_P]),
    ("hgm_dgm_stats", C.c_int, [_P, C.POINTER(Stats)]),
    ("hgm_dgm_count_ops", C.c_int, [_P, C.c_int, C.POINTER(C.c_int64)]),
    ("hgm_dgm_erase", C.c_int, [_P]),
    ("hgm_dgm_validate_tuning", C.c_int, [_P]),
    ("hgm_dgm_print_pts", C.c_int, [_P]),
    ("hgm_seed_rng", None, [C.c_uint64]),
    ("hgm_shor", C.c_int, [_P, C.c_int64, C.POINTER(C.c_int64), C.POINTER(Stats)]),
    ("hgm_frac_approx", None, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]),
    ("hgm_mul_inv", C.c_int64, [C.c_int64, C.c_int64]),
    ("hgm_revert_bits", C.c_int, [C.c_int, C.c_int]),
    ("hgm_generate", C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int64), C.c_int, C.c_char_p, C.c_char_p, C.c_size_t,
                               C.POINTER(C.c_size_t)]),
    ("hgm_auto_tuning", C.c_int, [C.c_int, C.c_int, C.POINTER(Tuning)]),
    ("hgm_autotune", C.c_int, [C.c_int, C.c_int, C.POINTER(Op), C.c_int64, C.c_int, C.c_int, C.c_int, C.POINTER(Tuning),
                               C.POINTER(C.c_double)]),
    ("hgm_server_default_config", None, [C.POINTER(ServerConfig)]),
    ("hgm_server_create", C.c_int, [C.POINTER(ServerConfig), C.POINTER(_P)]),
    ("hgm_server_submit", C.c_int, [_P, C.POINTER(Request), C.POINTER(C.c_int64)]),
    ("hgm_server_poll", C.c_int, [_P, C.POINTER(Result), C.c_int, C.POINTER(C.c_int), C.c_int]),
    ("hgm_server_wait_all", C.c_int, [_P]),
    ("hgm_server_tuning", C.c_int, [_P, C.c_int, C.POINTER(Tuning)]),
    ("hgm_server_set_tuning", C.c_int, [_P, C.c_int, C.POINTER(Tuning)]),
    ("hgm_server_clock_ms", C.c_double, [_P]),
    ("hgm_server_precompile", C.c_int, [_P, C.c_int, C.POINTER(Op), C.c_int64, C.c_int, C.POINTER(C.c_int)]),
    ("hgm_server_destroy", None, [_P]),
    ("setDevice", C.c_bool, [C.c_int]),
    ("GpuExecutionWrapper", _P, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("ProjectState", C.c_bool, [_P, C.c_int, C.c_int, C.c_long, C.c_long, C.c_int]),
    ("GetState", C.c_bool, [_P, C.c_int, C.c_int, C.c_long, C.c_long, C.c_int]),
]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built: run __graft_entry__.build() or make -C simulador_quantico_amd/csrc")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc: int) -> None:
    if rc != HGM_OK:
        raise HGMError(rc, lib.hgm_last_error().decode(errors="replace"))


def default_tuning(**kw) -> Tuning:
    t = Tuning()
    lib.hgm_default_tuning(C.byref(t))
    for k, v in kw.items():
        if not hasattr(t, k):
            raise KeyError(k)
        setattr(t, k, int(v))
    return t


def device_count() -> int:
    c = C.c_int(0)
    lib.hgm_device_count(C.byref(c))
    return c.value
