// Pass planner: turns a gate list into full-state passes over tiles.
//
// A pass streams the whole state once through tiles of 2^r amplitudes. The r
// tile-local index bits are the low c bits (contiguous in memory, so loads
// coalesce) plus r-c more bits chosen per pass; the remaining bits number the
// tiles. Every gate scheduled into a pass must be executable inside one tile:
//   * a non-diagonal gate needs its target bit local;
//   * a diagonal gate needs nothing local (a non-local target just selects
//     m00 or m11 for the whole tile);
//   * controls may sit anywhere: local controls are tested per amplitude,
//     non-local ones once per tile against the tile's base index.
// This generalises the batching of the reference (src/core/kernel.cu:199-235),
// which requires target bits in one contiguous window and forces a new pass
// for any gate outside it, diagonal or not.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "hgm_gpu.h"

namespace hgm {

struct JitCode;

enum OpKind : int32_t {
    K_DENSE = 0,        // full 2x2 update of the pair
    K_DIAG = 1,         // m01 == m10 == 0, local target: per-amplitude phase
    K_ANTI = 2,         // m00 == m11 == 0: swap-and-scale of the pair
    K_DIAG_GLOBAL = 3   // diagonal gate on a non-local target: one phase for the tile
};

// Exact-zero classification of PT::matrixType (src/core/common.cpp:33-38).
int classify(const double m[8]);

// One gate in tile coordinates, as the pass kernels read it.
struct DevOp {
    int32_t kind;
    int32_t tbit;       // tile bit of the target, or its state-index bit for K_DIAG_GLOBAL
    uint32_t lmask;     // controls on tile-local bits (tile coordinates)
    uint32_t lval;
    uint64_t gmask;     // controls on non-local bits (state-index coordinates)
    uint64_t gval;
    double m[8];
};
static_assert(sizeof(DevOp) == 96, "DevOp layout is shared with the kernels");

constexpr int kMaxQubits = 40;
constexpr int kMaxTileBits = 16;
constexpr int kMaxRegAmps = 32;      // amplitudes per thread of the register kernel
constexpr int kMaxThreadBits = 10;   // 1024 threads per workgroup

// Workgroup size limit of the register kernel for A amplitudes per thread of
// amp_bytes bytes: the VGPR budget per lane is 512 / (waves per SIMD); A
// complex doubles take 4A VGPRs, A complex floats 2A (FP32 with 16 amplitudes
// per thread fits 128 VGPRs, so 1024 threads can hold a 2^14-amplitude tile).
constexpr int reg_kernel_max_threads(int A, size_t amp_bytes) {
    return amp_bytes <= 8 ? (A >= 32 ? 512 : 1024) : (A >= 32 ? 256 : (A >= 16 ? 512 : 1024));
}

// Workgroup size limit of the plan-compiled pass kernels (and the serving
// engine's compiled kernels): they keep no interpreter state, so FP32 tiles of
// 32 amplitudes per thread fit 1024 threads (Grover passes on 2^15-amplitude
// tiles take 95-128 VGPRs; denser passes may spill). FP64 as the interpreter.
constexpr int compiled_kernel_max_threads(int A, size_t amp_bytes) {
    return amp_bytes <= 8 ? 1024 : reg_kernel_max_threads(A, amp_bytes);
}

struct TileShape {
    int r;      // tile bits
    int B;      // threads per tile
    int R;      // amplitude pairs per thread
    int c;      // coalesced (always-local) low bits
};

struct Pass {
    uint64_t local_mask;              // state-index bits that are tile-local
    int8_t local_pos[kMaxTileBits];   // tile bit j -> state-index bit
    int64_t op_begin;                 // first DevOp of the pass
    int64_t op_count;
    int64_t rop_begin = 0;            // register program of the pass (RegOp range)
    int64_t rop_count = 0;
    int64_t mem = -1;                 // PassMem index of the register program
    int32_t phases = 0;               // register phases (LDS remaps = phases - 1)
};

// ---- register kernel program --------------------------------------------
// In the register kernel each thread holds A = 2R amplitudes of its tile in
// registers. A "phase" assigns log2(A) tile bits to the register index and the
// other tile bits, in ascending order, to the thread index. Gates are applied
// to registers only; a remap moves the tile through LDS into the next phase's
// assignment.
enum RegKind : int32_t {
    R_DENSE = 0,       // 2x2 update of the register pair differing in register bit `rbit`
    R_ANTI = 1,        // swap-and-scale of that pair (m00 == m11 == 0)
    R_PHASE = 2,       // multiply by (m[0], m[1]) where all three conditions hold
    R_REMAP = 3,       // move to phase map `rbit` through LDS
    R_SWAP = 4,        // exchange of the pair (the X matrix)
    R_PHASE_REAL = 5,  // R_PHASE with a real factor m[0] (m[1] == 0)
    R_DENSE_REAL = 6,  // R_DENSE with a real matrix (imaginary parts exactly 0)
    // Unconditioned 2x2 update whose rows each hold an exact 1 (HGM_PLAN_PIVOT):
    //   out0 = x[p0] + t0 * x[1 - p0],  out1 = x[p1] + t1 * x[1 - p1]
    // with (x[0], x[1]) the pair, t0 = (m[0], m[1]), t1 = (m[2], m[3]) and the
    // pivot columns p0 = tval & 1, p1 = tval >> 1 (rmask == tmask == 0).
    R_DENSE_PIV = 7,
    R_DENSE_PIV_REAL = 8   // R_DENSE_PIV with real t0, t1
};

struct RegOp {
    int32_t kind;
    int32_t rbit;       // register-index bit of the target, or the next PhaseMap (R_REMAP)
    uint32_t rmask;     // condition on register-index bits
    uint32_t rval;
    uint32_t tmask;     // condition on thread-index bits
    uint32_t tval;      // (the pivot columns for R_DENSE_PIV*)
    uint64_t gmask;     // condition on non-local state-index bits
    uint64_t gval;
    double m[8];
};
static_assert(sizeof(RegOp) == 104, "RegOp layout is shared with the kernels");

// Tile index of the element a thread holds in register k:
//   thr[0][t & 31] | thr[1][t >> 5] | reg[k]
struct PhaseMap {
    uint16_t reg[64];
    uint16_t thr[2][32];
};
static_assert(sizeof(PhaseMap) == 256, "PhaseMap layout is shared with the kernels");

// State-index offsets of the load (first phase) and store (last phase)
// assignments: element (t, k) of tile `base` lives at
//   base | lthr[0][t & 31] | lthr[1][t >> 5] | lreg[k]
struct PassMem {
    uint64_t lthr[2][32];
    uint64_t lreg[64];
    uint64_t sthr[2][32];
    uint64_t sreg[64];
    int32_t first_map;
    int32_t last_map;
};
static_assert(sizeof(PassMem) == 2056, "PassMem layout is shared with the kernels");

struct Plan {
    int qubits = 0;
    int dtype = HGM_C128;
    int kernel = HGM_KERNEL_LDS;
    TileShape shape{};
    std::vector<Pass> passes;
    std::vector<DevOp> ops;
    std::vector<int64_t> src_index;   // index of each DevOp in the iteration-expanded gate list
    std::vector<uint8_t> pre_phase;   // per DevOp: a kPrePhaseTag gate (kept with the next rotation)
    int64_t gates = 0;
    double plan_ms = 0;
    // register kernel program (kernel == HGM_KERNEL_REG)
    std::vector<RegOp> rops;
    std::vector<PhaseMap> maps;
    std::vector<PassMem> mems;
    int max_phases = 0;
    std::shared_ptr<JitCode> jit;     // HGM_KERNEL_JIT: generated and compiled pass kernels
};

// Builds the register program of every pass (fills rops/maps/mems).
void build_register_program(Plan &plan);

size_t amp_bytes(int dtype);

// Validates a tuning tuple (the reference's rules of src/core/dgm_core.cpp:177-241
// plus this library's limits) and derives the tile shape used for a state of
// `qubits` qubits. Returns an HGM_* status and fills *err on failure.
int make_shape(int qubits, int dtype, const hgm_tuning &t, TileShape *out, std::string *err);

// Equivalent gate list whose uncontrolled dense gates have a real,
// non-negative first column: each such gate U becomes diag(p0, p1) * M, and
// the diagonal factor is carried forward on its qubit (diagonal gates and
// controls commute with it) into the next uncontrolled gate on that qubit, or
// emitted as a diagonal gate before a controlled gate on it and at the end.
// A dense update then costs 12 real multiply-adds per amplitude pair instead
// of 16 (14 when m00 is already real).
void fold_phases(int qubits, const hgm_op *ops, int64_t n_ops, std::vector<hgm_op> *out);

// Equivalent gate list in which each row of every uncontrolled dense gate is
// divided by its larger-magnitude entry, so that entry is exactly 1 and the
// other has magnitude <= 1. The row factors d0, d1 are carried forward as
// d0 * diag(1, d1 / d0): the relative factor stays on the qubit until the next
// uncontrolled gate there absorbs it (uncontrolled diagonal gates are absorbed
// into it; controls commute with it; a controlled non-diagonal gate on the
// qubit gets it as a diagonal gate first), and d0 joins one global factor g.
// The emitted list therefore produces the state divided by g; a uniform scale
// by |g| is emitted whenever |g| leaves [2^-24, 2^24], and g itself with the
// last carried diagonal. A dense update then costs 8 real multiply-adds per
// amplitude pair (4 when the scaled matrix is real) instead of 16.
void pivot_rows(int qubits, const hgm_op *ops, int64_t n_ops, std::vector<hgm_op> *out);

// `reserved` tag of the phase gates rotate_rows puts in front of its real
// rotations: the planners keep such a gate in the pass and register phase of
// the rotation it precedes (its target must be local there).
constexpr int32_t kPrePhaseTag = 1;

// Equivalent gate list in which every uncontrolled dense gate V (after the
// carried factor is absorbed, as in pivot_rows) is split as
// V = diag(x0, x1) * O * diag(1, y) with O real: diag(1, y) is emitted in
// front of O, tagged kPrePhaseTag; O is emitted row-scaled (real, an exact 1
// per row); diag(x0, x1) times the row factors is carried forward. Gates
// without such a split keep pivot_rows' form.
void rotate_rows(int qubits, const hgm_op *ops, int64_t n_ops, std::vector<hgm_op> *out);

// Equivalent gate list in which every run of uncontrolled gates on one qubit
// (with nothing in between that touches the qubit, except diagonal gates
// while the run so far is diagonal) is one gate: the product of the run,
// placed where the run's first non-commuting successor needs it, or at the
// end; products equal to the identity are dropped.
void fuse_gates(int qubits, const hgm_op *ops, int64_t n_ops, std::vector<hgm_op> *out);

// Builds the pass list for ops[0..n_ops) repeated `iterations` times
// (after fuse_gates, then pivot_rows or fold_phases, when tuning.planner has
// HGM_PLAN_FUSE, HGM_PLAN_PIVOT or HGM_PLAN_FOLD_PHASES; fusing first expands
// the iterations). With program_only, a plan-compiled (HGM_KERNEL_JIT)
// tuning gets its tile shape checked and its register program built, but no
// kernels generated or compiled: the plan (kernel HGM_KERNEL_REG) is for
// kernels compiled from its passes' structure elsewhere (the serving engine).
int build_plan(int qubits, int dtype, const hgm_op *ops, int64_t n_ops, int iterations, const hgm_tuning &t,
               Plan *out, std::string *err, bool program_only = false);

}  // namespace hgm
