// Internal interface between the HIP runtime (runtime.hip) and the host-side
// executor / C ABI layer (dgm.cpp, capi.cpp).
#pragma once

#include <cstddef>
#include <string>

#include "hgm_gpu.h"

namespace hgm {

void set_error(const std::string &msg);
const std::string &last_error();

// Initialise the device context (device < 0: keep the current one or use 0).
int gpu_init(int device);

// Pinned host allocation (zeroed) for host states; HGM_E_NOMEM on failure.
int alloc_pinned(void **p, size_t bytes);
void free_pinned(void *p);

// The gate-list rewrite of hgm_auto_tuning (the rotation split: real rotations
// and one gathered phase multiply per register slot in the compiled kernels).
constexpr int kAutoRowFlag = HGM_PLAN_ROTATE;

// The tuning a call runs with: the caller's, unless HGM_TUNING is set in the
// environment to "auto" (hgm_auto_tuning for the state's qubit count and
// dtype) or "autotune" (a tile autotuned once per qubit count and dtype, on the
// first circuit seen, with hgm_auto_tuning's kernel and planner flags).
// Results do not depend on the tuning beyond rounding.
hgm_tuning env_tuning(const hgm_tuning &caller, int qubits, int dtype, const hgm_op *ops, int64_t n_ops,
                      int iterations);

// Plan ops for state s and run every pass (the body of hgm_state_apply).
int plan_and_run(hgm_state *s, const hgm_op *ops, int64_t n_ops, const hgm_tuning *tuning, int iterations,
                 hgm_stats *st);

}  // namespace hgm
