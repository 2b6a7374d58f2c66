// Launch wrappers of the gfx950 kernels (defined in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "planner.hpp"

namespace hgm {

// Per-pass launch arguments shared by the pass kernels.
struct PassArgs {
    uint64_t nonlocal_mask;            // state-index bits that number the tiles
    int64_t n_tiles;                   // 2^(qubits - r)
    int32_t r, c, B, R;                // tile shape
    int8_t local_pos[kMaxTileBits];    // tile bit j -> state-index bit
};

// One full-state pass of the LDS tile kernel: every tile is loaded into LDS,
// the pass's gates are applied there, and the tile is written back.
hipError_t launch_pass_lds(int dtype, void *state, const DevOp *ops, int n_ops, const PassArgs &a, hipStream_t s);

// Raises the dynamic LDS limit of the LDS and register pass kernels of one
// dtype (otherwise done on their first launch): call before capturing their
// launches into a graph.
hipError_t prepare_pass_kernels(int dtype);

// One full-state pass of the register kernel (A amplitudes per thread, B
// threads per tile of 2^r); lds_bytes is 0 for a pass without remaps.
hipError_t launch_pass_reg(int dtype, int A, void *state, const RegOp *ops, int n_ops, const PhaseMap *maps,
                           const PassMem *pm, uint64_t nonlocal, int64_t n_tiles, int r, int B, size_t lds_bytes,
                           hipStream_t s);

// Readout of one tile or of a whole state: FP64 sum of |a|^2 and the largest
// |a|^2 with its index (lowest index on ties).
struct ReadoutPartial {
    double sum;
    double pmax;
    uint64_t idx;
    uint64_t pad;
};
static_assert(sizeof(ReadoutPartial) == 32, "shared with the kernels");

// One pass of one state inside a batched register-kernel launch. Addresses
// are byte offsets into the serving arena, which the kernel receives as
// separate read-only (programs) and written (states, partials) bases, so that
// the compiler sees global memory and reads the programs with scalar loads.
struct BatchItem {
    uint64_t state_off;          // the state (2^qubits amplitudes)
    uint64_t ops_off;            // the pass's register program (RegOp[n_ops])
    uint64_t maps_off;           // the plan's phase maps
    uint64_t pm_off;             // the pass's load/store table (PassMem)
    uint64_t nonlocal_mask;      // state-index bits that number the tiles
    int64_t tile_begin;          // first workgroup of this item in the launch
    int64_t partials_off;        // >= 0: per-tile readout of the pass output (ReadoutPartial[])
    int32_t n_ops;
    int32_t pad;
};
static_assert(sizeof(BatchItem) == 64, "shared with the kernels");

struct ReadoutTask {
    const ReadoutPartial *partials;
    int64_t n;
    ReadoutPartial *out;
};

// One launch of the register kernel over several states (items sorted by
// tile_begin; workgroups [0, total_tiles)); every item's plan has the same
// A = 2*repeat_count and block size B. lds_bytes covers the largest tile of
// an item with remaps; the items' offsets are relative to `arena`.
hipError_t launch_batch_reg(int dtype, int A, int B, const BatchItem *items, int n_items, int64_t total_tiles,
                            size_t lds_bytes, void *arena, hipStream_t s);

// out of each task = combination of its n partials (one workgroup per task).
hipError_t launch_readout_final(const ReadoutTask *tasks, int n_tasks, hipStream_t s);

// dst[0..bytes) = src[0..bytes) (bytes a multiple of 16).
hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);

// state[i] = (i == index) for i < n_amps.
hipError_t launch_set_basis(int dtype, void *state, uint64_t n_amps, uint64_t index, hipStream_t s);

// Per-block FP64 partial sums of |a|^2 split by index bit `bit`
// (bit < 0: everything lands in the first sum). partial holds 2*blocks doubles.
hipError_t launch_prob_halves(int dtype, const void *state, uint64_t n_amps, int bit, double *partial, int blocks,
                              hipStream_t s);

// Per-block maximum of |a|^2 with the lowest index among ties.
hipError_t launch_argmax(int dtype, const void *state, uint64_t n_amps, double *pmax, uint64_t *pidx, int blocks,
                         hipStream_t s);

// Per-block partials (5 doubles per block) of the comparison of two states of
// n amplitudes: Re, Im of sum conj(a)*b, sum |a|^2, sum |b|^2, max |a - b|.
hipError_t launch_compare(int dtype_a, const void *a, int dtype_b, const void *b, uint64_t n, double *partial,
                          int blocks, hipStream_t s);

// Collapse `bit` to `outcome`, moving the kept half to the bit=0 slot divided
// by norm and zeroing the bit=1 slot (src/core/dgm_core.cpp:115-122).
hipError_t launch_collapse(int dtype, void *state, uint64_t n_amps, int bit, int outcome, double norm, hipStream_t s);

}  // namespace hgm
