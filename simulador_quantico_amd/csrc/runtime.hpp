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

// Plan ops for state s and run every pass (the body of hgm_state_apply).
int plan_and_run(hgm_state *s, const hgm_op *ops, int64_t n_ops, const hgm_tuning *tuning, int iterations,
                 hgm_stats *st);

}  // namespace hgm
