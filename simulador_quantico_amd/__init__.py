"""MI355X-native t_GPU path of the HybriD-GM state-vector simulator.

Python access to libhgm.so (HIP kernels for gfx950 + C++ executor) through its
C ABI (include/hgm_gpu.h). The executor mirrors the reference's DGM API; see
simulador_quantico_amd.dgm.
"""
from ._native import (HGM_C64, HGM_C128, HGM_KERNEL_AUTO, HGM_KERNEL_LDS, HGM_KERNEL_REG, HGM_PLAN_GREEDY,
                      HGM_PLAN_WINDOW, HGMError, default_tuning, device_count)
from .circuits import (AddF, C2AddF, C2AddMod, C2SubF, C2SubMod, CAddF, CMultMod, CNot, Controlled1, Controlled2,
                       ControlledN, ControledZ, CRMultMod, CSubF, CSwapR, DraperAdder, GroverIteration,
                       GroverIterationRef, GroverIterations, Hadamard, Oracle1, Pauli_X, Pauli_Z, QFT, RandomLayers,
                       RQFT, Shor, SubF, SwapOver, Toffoli, genRot)
from .dgm import (DENSE, DIAG_PRI, DIAG_SEC, DGM, T_CPU, T_GPU, T_HYBRID, T_PAR_CPU, Batch, Plan, State, make_ops,
                  ops_to_arrays, seed_rng)

__all__ = [name for name in dir() if not name.startswith("_")]
