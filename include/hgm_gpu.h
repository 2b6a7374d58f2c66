/*
 * hgm_gpu.h -- C ABI of the MI355X (gfx950) t_GPU execution library.
 *
 * This is the drop-in boundary for HybriD-GM's GPU execution path. It has two
 * layers:
 *
 *   1. The reference-compatible entry points, with the exact names, argument
 *      order and meaning that the reference's DGM binds
 *      (include/core/dgm.h:22-32, defined in src/core/kernel.cu:31,303,321,384).
 *      A reference build links this library in place of kernel.o.
 *
 *   2. A POD ABI (hgm_*) with plain pointers, sizes and status codes, for
 *      callers that do not carry the reference's C++ PT struct: op arrays,
 *      device-resident states, plans and a DGM handle.
 *
 * Amplitude layout everywhere: interleaved (re, im) pairs, index bit
 * (qubits-1-q) <-> circuit qubit q (qubit 0 is the most significant bit,
 * src/core/dgm_parser.cpp:151-152). Gate matrices are row-major
 * [m00, m01, m10, m11] and every gate maps a pair (a0, a1) differing in the
 * target bit to (m00*a0 + m01*a1, m10*a0 + m11*a1) (src/core/kernel.cu:109-111).
 *
 * Errors: every hgm_* call returns HGM_OK (0) or an HGM_E_* code and records
 * a message readable through hgm_last_error() (per calling thread). The
 * reference-compatible entry points keep the reference's convention instead:
 * print the message and exit(1) (src/core/kernel.cu:439-447).
 */
#ifndef HGM_GPU_H
#define HGM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
typedef __complex__ float hgm_cfloat;
#else
#include <stdbool.h>
typedef float _Complex hgm_cfloat;
#endif

/* ---------------------------------------------------------------- codes */
enum {
    HGM_OK = 0,
    HGM_E_ARG = 1,         /* bad argument (null pointer, size, qubit index) */
    HGM_E_TUNING = 2,      /* tuning tuple violates the contract */
    HGM_E_HIP = 3,         /* HIP runtime error (message carries hipGetErrorString) */
    HGM_E_NOMEM = 4,       /* device or pinned host allocation failed */
    HGM_E_PARSE = 5,       /* circuit text names an unknown gate or is malformed */
    HGM_E_STATE = 6,       /* state handle / qubit count mismatch */
    HGM_E_UNSUPPORTED = 7  /* feature outside this library (the CPU engines behind the executor's
                              t_CPU / t_PAR_CPU / t_HYBRID, multi-GPU) */
};

/* amplitude types */
enum { HGM_C128 = 0, /* complex double, 16 B/amplitude (default) */
       HGM_C64 = 1   /* complex float, 8 B/amplitude (the reference's precision) */ };

/* pass planners */
enum { HGM_PLAN_GREEDY = 0, /* dependency-aware lookahead, any set of tile-local bits */
       HGM_PLAN_WINDOW = 1, /* in-order, contiguous window (the contract of src/core/kernel.cu:199-235) */
       /* flag, or-ed with a planner: rewrite each uncontrolled dense gate U as diag(p0, p1) * M with
        * M's first column real, carrying the diagonal factor forward on its qubit into the next gate
        * there (diagonal gates and controls commute with it). Same unitary; results differ from the
        * unfolded list by rounding only (the dense update drops from 16 to 12 multiply-adds per pair). */
       HGM_PLAN_FOLD_PHASES = 16,
       /* flag: fuse every run of uncontrolled gates on one qubit into one gate (their product),
        * placed before the first later gate on that qubit it does not commute with; products equal
        * to the identity vanish. Iterations are expanded first, so runs fuse across them. Same
        * unitary; results differ by rounding only. */
       HGM_PLAN_FUSE = 32,
       /* flag: scale each row of every uncontrolled dense gate by its larger entry, which becomes
        * exactly 1; the row factors are carried forward as a per-qubit relative factor diag(1, r)
        * (absorbed by the next uncontrolled gate on that qubit, like uncontrolled diagonal gates)
        * and one global factor (applied as a uniform scale when its magnitude leaves
        * [2^-24, 2^24], and at the end). Subsumes HGM_PLAN_FOLD_PHASES. Same unitary; results
        * differ by rounding only (a dense update drops to 8 multiply-adds per pair, 4 when the
        * scaled matrix is real). */
       HGM_PLAN_PIVOT = 64,
       /* flag (greedy planner): after packing, move gates between neighbouring passes from the one
        * with more arithmetic to the one with less, where the gate commutes with the gates it passes
        * and its target is tile-local there; the pass count is unchanged. */
       HGM_PLAN_BALANCE = 128,
       /* flag (greedy planner): choose each pass's tile-local set by the gates it admits plus the
        * most the next pass could then admit, instead of by the gates it admits alone. Fewer passes
        * on layered circuits (30-qubit random layers, 2^12 tiles: 11 instead of 12), at about
        * 20 times the planning work. */
       HGM_PLAN_LOOKAHEAD = 256,
       /* flag: like HGM_PLAN_PIVOT, but each uncontrolled dense gate V (after absorbing the
        * carried factor) is split as V = diag(x0, x1) * O * diag(1, y) with O real: the phase
        * diag(1, y) goes in front of O as a diagonal gate kept in O's pass, O is row-scaled
        * (real: 4 multiply-adds per pair), and diag(x0, x1) with the row factors is carried
        * forward. The plan-compiled kernels multiply each register slot once by the product of
        * the consecutive diagonal gates in front of a group of rotations, instead of folding
        * each phase into its gate. Gates that admit no such split (not a unitary times a scale)
        * keep the HGM_PLAN_PIVOT form. Subsumes HGM_PLAN_PIVOT. Same unitary; results differ
        * by rounding only. */
       HGM_PLAN_ROTATE = 512 };

/* kernel families */
enum { HGM_KERNEL_AUTO = 0,  /* the register kernel when the tile shape allows it, else LDS */
       HGM_KERNEL_LDS = 1,   /* tile in LDS, one LDS round trip per gate */
       HGM_KERNEL_REG = 2,   /* amplitudes in registers, interpreted per-pass program */
       HGM_KERNEL_JIT = 3    /* the register kernel's programs compiled to straight-line code at plan time */ };

/* One compiled single-qubit gate: the POD form of the reference's PT
 * (include/core/common.h:57-90). */
typedef struct hgm_op {
    int32_t target_bit;      /* state-index bit the gate acts on */
    int32_t reserved;        /* must be 0 */
    uint64_t control_mask;   /* state-index bits that must match control_value */
    uint64_t control_value;
    double m[8];             /* m00.re m00.im m01.re m01.im m10.re m10.im m11.re m11.im */
} hgm_op;

/* The tuning contract of the reference (include/cli/cli_common.h:8-17,
 * src/core/dgm_core.cpp:177-241, makefile:7): a pass loads tiles of
 * 2^gpu_region_bits amplitudes whose low gpu_coalesced_bits index bits are
 * contiguous in memory; block_size threads each own 2*repeat_count of them;
 * at most ops_block gates are fused into one pass.
 * Invariant: 2*block_size*repeat_count == 2^gpu_region_bits. */
typedef struct hgm_tuning {
    int32_t block_size;          /* threads per workgroup (reference default 64) */
    int32_t repeat_count;        /* amplitude pairs per thread (default 2) */
    int32_t gpu_region_bits;     /* log2 tile size (default 8) */
    int32_t gpu_coalesced_bits;  /* low index bits always in the tile (default 4) */
    int32_t ops_block;           /* max gates per pass (OPS_BLOCK, default 300) */
    int32_t planner;             /* HGM_PLAN_* */
    int32_t kernel;              /* HGM_KERNEL_* */
    int32_t lookahead;           /* greedy planner scan window in gates (0 = default 4096) */
} hgm_tuning;

/* Call statistics. Every call that takes an hgm_stats* adds its figures to
 * the fields, so zero the struct before the first call to count from there.
 * bytes_alg = sum over passes of 2 * tile amplitudes * sizeof(amplitude) =
 * the HBM bytes one read and one write of every loaded amplitude costs. */
typedef struct hgm_stats {
    int64_t gates;           /* gates applied (iterations included) */
    int64_t passes;          /* kernel launches over the state */
    int64_t amp_updates;     /* gates * 2^qubits */
    double bytes_alg;        /* algorithmic HBM bytes of all passes */
    double plan_ms;          /* host planning time */
    double h2d_ms;           /* host-to-device copy (0 for resident states) */
    double kernel_ms;        /* HIP-event time over all passes, on the launch stream */
    double d2h_ms;           /* device-to-host copy */
    double total_ms;         /* wall time of the call */
} hgm_stats;

/* Default tuning (the reference's CLI defaults, greedy planner, auto kernel). */
void hgm_default_tuning(hgm_tuning *t);

/* Last error message of the calling thread ("" if none). */
const char *hgm_last_error(void);

/* Library version string. */
const char *hgm_version(void);

/* ------------------------------------------------------------------ device */
/* Select and initialise device `device` (creates the context and the stream
 * the library launches on). Idempotent. */
int hgm_init(int device);
/* Number of visible HIP devices (0 when none). */
int hgm_device_count(int *count);
/* Free the device buffers the library caches between calls (the hgm_execute
 * state buffer, op tables, register programs), drop the plans that hgm_execute
 * and hgm_state_apply keep for the last 4 distinct (gate list, tuning,
 * iterations, qubits, dtype) they ran, and unload plan-compiled kernel
 * modules; all are rebuilt on demand. */
int hgm_release_cache(void);

/* ------------------------------------------------------------------ one-shot
 * Applies ops[0..n_ops) `iterations` times to a host state of 2^qubits
 * amplitudes of type dtype, in place: upload, run every pass, download. This
 * is what the reference's t_GPU does per execute() (src/core/kernel.cu:131-291),
 * with the device buffer cached across calls, and the plan too when the same
 * gate list, tuning and iterations ran recently (stats->plan_ms then holds the
 * lookup time). A pinned host state (hipHostMalloc) is copied by DMA directly;
 * a pageable one through the library's pinned staging chunks. */
int hgm_execute(void *host_state, int qubits, int dtype, const hgm_op *ops, int64_t n_ops,
                const hgm_tuning *tuning, int iterations, hgm_stats *stats);

/* ------------------------------------------------------------------ resident states */
typedef struct hgm_state hgm_state;

/* Allocate a device-resident state of 2^qubits amplitudes (zeroed). */
int hgm_state_create(int qubits, int dtype, hgm_state **out);
/* Borrow caller-owned device memory (e.g. a framework tensor); not freed by destroy. */
int hgm_state_wrap(void *device_ptr, int qubits, int dtype, hgm_state **out);
void hgm_state_destroy(hgm_state *s);
int hgm_state_qubits(const hgm_state *s);
int hgm_state_dtype(const hgm_state *s);
void *hgm_state_device_ptr(hgm_state *s);
/* |index> on the device, no host copy (setMemoryValue, src/core/dgm_core.cpp:63-65). */
int hgm_state_set_basis(hgm_state *s, uint64_t index);
int hgm_state_upload(hgm_state *s, const void *host);
/* Device-to-device copy between states of equal size and dtype (one streaming kernel). */
int hgm_state_copy(hgm_state *dst, const hgm_state *src);
int hgm_state_download(const hgm_state *s, void *host);
/* Region copies of a host state of host_qubits qubits (same dtype as s):
 * amplitude j of s is host amplitude region_id | deposit(j, region_mask),
 * deposit placing the bits of j, lowest first, on the set bits of region_mask.
 * region_mask holds exactly hgm_state_qubits(s) bits of the host state and
 * region_id none of them. Gather fills s, scatter writes s back; the other
 * host amplitudes are not touched. The region projection of t_HYBRID
 * (ProjectState / GetState, src/core/kernel.cu:321-435) in one pipelined copy. */
int hgm_state_gather(hgm_state *s, const void *host, int host_qubits, uint64_t region_id, uint64_t region_mask);
int hgm_state_scatter(const hgm_state *s, void *host, int host_qubits, uint64_t region_id, uint64_t region_mask);
/* Apply a gate list `iterations` times (plans, uploads the op tables once, launches). */
int hgm_state_apply(hgm_state *s, const hgm_op *ops, int64_t n_ops, const hgm_tuning *tuning,
                    int iterations, hgm_stats *stats);
/* hgm_state_apply without any device allocation: the plan's tables (register
 * programs, phase maps, pass tables, or op tables) go to an 8 MiB workspace
 * that the library holds per GPU for the life of the process, instead of to
 * buffers it allocates. No counterpart in the reference: for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc). Needs a device context (any hgm_state_create or hgm_init call
 * made it); runs the register or LDS kernels (HGM_KERNEL_JIT is rejected with
 * HGM_E_TUNING, since loading compiled code allocates); a plan whose tables
 * exceed 8 MiB is rejected with HGM_E_ARG. Every call on a GPU shares that
 * GPU's workspace: calls are serialised by the library's lock, and a caller
 * running other work on the same workspace concurrently would race. */
int hgm_state_apply_noalloc(hgm_state *s, const hgm_op *ops, int64_t n_ops, const hgm_tuning *tuning,
                            int iterations, hgm_stats *stats);
/* Sum of |a|^2 over the whole state (FP64 accumulation on the device). */
int hgm_state_norm2(const hgm_state *s, double *out);
/* P(qubit == 1) (FP64 accumulation on the device). qubit counts from the MSB. */
int hgm_state_prob_one(const hgm_state *s, int qubit, double *out);
/* P(qubit == 0) and P(qubit == 1), each summed directly over its half (FP64
 * accumulation on the device): the two sums of DGM::measure
 * (src/core/dgm_core.cpp:79-84). qubit counts from the MSB. */
int hgm_state_prob_halves(const hgm_state *s, int qubit, double *zero, double *one);
/* Collapse qubit to `outcome` and move the surviving half to the |0> slot,
 * dividing by sqrt(p) (src/core/dgm_core.cpp:115-122). */
int hgm_state_collapse_reset(hgm_state *s, int qubit, int outcome, double p);
/* Index of the largest |a|^2 (ties -> lowest index) and its probability. */
int hgm_state_argmax(const hgm_state *s, uint64_t *index, double *prob);

/* Fidelity |<a|b>|^2 / (<a|a><b|b>) and max |a_i - b_i| of two states of the
 * same qubit count and any amplitude types (FP64 accumulation on the device):
 * the FP32-against-FP64 check of large states without a host copy. */
int hgm_state_compare(const hgm_state *a, const hgm_state *b, double *fidelity, double *max_abs_err);

/* ------------------------------------------------------------------ plans (CPU only) */
typedef struct hgm_plan hgm_plan;
int hgm_plan_build(int qubits, int dtype, const hgm_op *ops, int64_t n_ops, int iterations,
                   const hgm_tuning *tuning, hgm_plan **out);
void hgm_plan_destroy(hgm_plan *p);
int64_t hgm_plan_num_passes(const hgm_plan *p);
/* For pass i: the tile-local bit mask, the number of gates, and (optionally)
 * their indices into the (iteration-expanded) op list, in execution order. */
int hgm_plan_pass(const hgm_plan *p, int64_t i, uint64_t *local_mask, int64_t *n_gates, int64_t *gate_index);
/* Effective tile shape of a plan: region bits, block size, repeat count, coalesced bits. */
int hgm_plan_shape(const hgm_plan *p, int32_t *region_bits, int32_t *block_size, int32_t *repeat_count,
                   int32_t *coalesced_bits);
int hgm_state_run_plan(hgm_state *s, const hgm_plan *p, hgm_stats *stats);
/* Kernel family the plan runs on (HGM_KERNEL_LDS, HGM_KERNEL_REG or HGM_KERNEL_JIT; never AUTO). */
int hgm_plan_kernel(const hgm_plan *p);
/* Register-kernel program of pass i, for inspection: rops receives n_rops
 * records of 104 bytes, local_pos the 16 tile-bit -> state-bit entries, and
 * first/last the phase maps of the load and store assignments. Passing NULL
 * buffers queries n_rops. */
int hgm_plan_program(const hgm_plan *p, int64_t i, void *rops, int64_t *n_rops, int32_t *first_map,
                     int32_t *last_map, int8_t *local_pos);
/* All phase maps of a register-kernel plan (records of 256 bytes). */
int hgm_plan_maps(const hgm_plan *p, void *maps, int64_t *n_maps);
/* Generated source of an HGM_KERNEL_JIT plan (empty for other kernels); out
 * may be NULL to query the length. */
int hgm_plan_jit_source(const hgm_plan *p, char *out, size_t cap, size_t *len);
/* Source of the serving engine's compiled kernel for pass `pass` of an
 * HGM_KERNEL_REG or HGM_KERNEL_JIT plan (see hgm_server_precompile): a
 * function of the pass's structure only, so plans of circuits that differ in
 * matrix entries or control values have equal sources. out may be NULL to
 * query the length. No counterpart in the reference (an inspection aid of
 * this library). */
int hgm_plan_serving_source(const hgm_plan *p, int64_t pass, char *out, size_t cap, size_t *len);
/* Replay the plan's passes from a hipGraph (HGM_KERNEL_JIT plans only): the
 * first hgm_state_run_plan on a state captures its launches, later runs on
 * the same state replay them with one graph launch. */
int hgm_plan_set_graph(hgm_plan *p, int enable);
/* hipRTC compile time of an HGM_KERNEL_JIT plan in seconds (0 when the code
 * came from the process cache or the plan is not JIT). */
double hgm_plan_compile_seconds(const hgm_plan *p);

/* A batch of n resident states of any qubit counts, state i advanced by the
 * plan plans[i] (any kernel family: the tables of register and LDS plans are
 * copied to the device once, here), captured as one hipGraph: the graph forks
 * into up to 64 branches (one per state; beyond 64 states, spread over them
 * by their plans' bytes, largest first), each branch runs its states' passes
 * in order, and the branches join, so passes of different states can run at
 * once. One
 * hgm_batch_run is one graph launch for the whole batch. States must be
 * distinct and match their plan's qubit count and dtype; the plan-compiled
 * plans' modules (at most 16 distinct) are kept loaded while the batch runs,
 * and the graph is captured again when one was unloaded in between. The states and
 * plans must outlive the batch. No counterpart in the reference (it runs one
 * state per call: src/core/kernel.cu:192-282). */
typedef struct hgm_batch hgm_batch;
int hgm_batch_create(hgm_state *const *states, const hgm_plan *const *plans, int n, hgm_batch **out);
/* Launches the batch's graph and waits for it; stats sums the plans' gates,
 * passes and bytes, with the graph's device time as kernel_ms. */
int hgm_batch_run(hgm_batch *b, hgm_stats *stats);
void hgm_batch_destroy(hgm_batch *b);

/* The gate list after phase folding (HGM_PLAN_FOLD_PHASES), for inspection
 * and checking: out may be NULL to query *n_out (at most 2 * n_ops + qubits). */
int hgm_fold_phases(int qubits, const hgm_op *ops, int64_t n_ops, hgm_op *out, int64_t cap, int64_t *n_out);
/* The gate list after fusion (HGM_PLAN_FUSE), same conventions (at most n_ops + qubits gates). */
int hgm_fuse_gates(int qubits, const hgm_op *ops, int64_t n_ops, hgm_op *out, int64_t cap, int64_t *n_out);
/* The gate list after row scaling (HGM_PLAN_PIVOT), same conventions (at most 2 * n_ops + qubits + 1
 * gates). */
int hgm_pivot_rows(int qubits, const hgm_op *ops, int64_t n_ops, hgm_op *out, int64_t cap, int64_t *n_out);
/* The gate list after the rotation split (HGM_PLAN_ROTATE), same conventions (at most
 * 2 * n_ops + qubits + 1 gates; the `reserved` field of the phase gates in front of the real
 * rotations is nonzero). */
int hgm_rotate_rows(int qubits, const hgm_op *ops, int64_t n_ops, hgm_op *out, int64_t cap, int64_t *n_out);

/* The fast tuning of this library for a state of `qubits` qubits: the
 * plan-compiled kernels, gate fusion, the rotation split and pass balancing on
 * the greedy planner, and the tile measured fastest on the 30- and 32-qubit
 * random-layer circuits (FP64: 256 threads x 8 pairs on 2^12 tiles; FP32:
 * 512 x 8 on 2^13; 4 coalesced bits). No counterpart in
 * the reference, whose tuning comes from the command line
 * (include/cli/cli_common.h:21-25). Callers of the reference's ABI
 * (GpuExecutionWrapper) and of the executor opt in with the environment
 * variable HGM_TUNING=auto, or HGM_TUNING=autotune for a tile autotuned once
 * per qubit count and dtype on the first circuit (hgm_autotune); results are
 * the same as with the caller's tuning up to rounding. */
int hgm_auto_tuning(int qubits, int dtype, hgm_tuning *out);

/* Tuning check with the reference's rules plus the gfx950 limits; returns
 * HGM_E_TUNING with a message when the tuple is rejected. */
int hgm_validate_tuning(int qubits, int dtype, const hgm_tuning *t);

/* ------------------------------------------------------------------ autotuner
 * Times every valid tile shape of the given kernel family on a scratch device
 * state running the given circuit planned with `planner` (HGM_PLAN_* flags),
 * and returns the fastest tuning and its kernel time. Shapes: 2*repeat_count
 * in {8, 16, 32} amplitudes per thread; for AUTO, REG and LDS gpu_region_bits
 * 8..13 and gpu_coalesced_bits 2..4; for JIT, whose every candidate is
 * compiled, gpu_region_bits 10..13 and the coalesced bits that make 128- and
 * 256-byte runs (3 and 4 for FP64, 4 and 5 for FP32). No counterpart in the reference, whose tuning
 * is given by hand on the command line (include/cli/cli_common.h:21-25). */
int hgm_autotune(int qubits, int dtype, const hgm_op *ops, int64_t n_ops, int iterations, int kernel, int planner,
                 hgm_tuning *best, double *best_ms);

/* ------------------------------------------------------------------ serving
 * Continuous batching of many circuit requests on one GPU. No counterpart in
 * the reference (one DGM executes one circuit per execute() call). Requests
 * are planned on host threads as they arrive, admitted in arrival order while
 * their HBM footprint fits the budget and a slot is free, and then advance
 * together: each step runs the next pass of every admitted state in batched
 * launches. A finished request reports its readout (FP64 norm and the argmax
 * of |a|^2) and, when asked, its final state. Each result is computed by the
 * same kernels and programs as when the request runs alone, so it does not
 * depend on what else is in flight. Thread-safe: any thread may submit and
 * poll. */
typedef struct hgm_server hgm_server;

typedef struct hgm_server_config {
    int32_t dtype;               /* amplitude type of every request (default HGM_C64) */
    int32_t max_active;          /* requests resident on the device at once (default 256) */
    uint64_t hbm_budget_bytes;   /* device memory for states, programs and readout (default 32 GiB) */
    int32_t planner_threads;     /* host threads that plan requests (default 4) */
    int32_t autotune;            /* 1: time tile shapes per qubit count at creation (hgm_autotune) */
    int32_t min_qubits;          /* qubit range to autotune (default 10..26) */
    int32_t max_qubits;
} hgm_server_config;

typedef struct hgm_request {
    int32_t qubits;
    int32_t iterations;          /* >= 1: the gate list is applied this many times */
    const hgm_op *ops;           /* copied at submission */
    int64_t n_ops;               /* >= 1 */
    uint64_t basis;              /* initial state |basis> */
    void *state_out;             /* optional host buffer for the final state (2^qubits amplitudes) */
} hgm_request;

typedef struct hgm_result {
    int64_t id;
    int32_t status;              /* HGM_OK or the planning / launch error */
    int32_t passes;              /* passes over the state */
    uint64_t argmax;             /* index of the largest |a|^2 (lowest index on ties) */
    double p_max;                /* that |a|^2 */
    double norm2;                /* sum of |a|^2 (FP64) */
    double t_submit_ms;          /* server clock: submission, admission, result ready on the host */
    double t_admit_ms;
    double t_done_ms;
} hgm_result;

void hgm_server_default_config(hgm_server_config *c);
int hgm_server_create(const hgm_server_config *c, hgm_server **out);
/* Queue a request; *id numbers requests in submission order from 0. */
int hgm_server_submit(hgm_server *s, const hgm_request *r, int64_t *id);
/* Up to cap finished results in completion order; waits up to timeout_ms for
 * the first one (0: return at once). */
int hgm_server_poll(hgm_server *s, hgm_result *out, int cap, int *n, int timeout_ms);
/* Block until every submitted request has finished (results stay pollable). */
int hgm_server_wait_all(hgm_server *s);
/* Tuning the server uses for states of `qubits` qubits. */
int hgm_server_tuning(hgm_server *s, int qubits, hgm_tuning *out);
/* Replace that tuning for requests submitted afterwards (kernel must be
 * HGM_KERNEL_REG with 2*repeat_count in {2, 4, 8, 16, 32}; the rest as
 * hgm_validate_tuning checks it). */
int hgm_server_set_tuning(hgm_server *s, int qubits, const hgm_tuning *t);
/* Milliseconds since the server was created (the clock of hgm_result). */
double hgm_server_clock_ms(const hgm_server *s);
/* Compiles the serving pass kernels of a circuit shape: `ops` (applied
 * `iterations` times) is planned as a request of `qubits` qubits would be,
 * and every pass structure without a kernel gets one (hgm_plan_serving_source;
 * hipRTC, up to 16 compilations at once). From then on, requests of that
 * qubit count are planned with HGM_PLAN_GREEDY | HGM_PLAN_FUSE (no row
 * scaling, whose pivot choice depends on the matrix entries), so a circuit of
 * the same shape with other matrices or control values (another seed, another
 * marked value) finds its passes' kernels; its passes run them, and passes of
 * other structures run the batched interpreter. States of one tile (whose
 * whole circuit is one pass) always run the interpreter. FP32 requests of
 * 25-26 qubits are also planned on tiles of 2^15 amplitudes (1024 threads x
 * 32, which only compiled kernels hold); a request takes that plan when each
 * of its passes has a kernel and it needs under 1/1.7 of the default plan's
 * passes, and precompile compiles those kernels where the example would take
 * it. hgm_server_set_tuning for a qubit count turns this off for it. *n_new:
 * kernels added. No counterpart in the reference (the serving engine is this
 * library's own). */
int hgm_server_precompile(hgm_server *s, int qubits, const hgm_op *ops, int64_t n_ops, int iterations,
                          int *n_new);
/* Finishes every submitted request, then frees the server. */
void hgm_server_destroy(hgm_server *s);

/* ------------------------------------------------------------------ DGM handle
 * The executor API of the reference (class DGM, include/core/dgm.h:65-163)
 * behind an opaque handle, for callers without C++ (ctypes, cgo). Fields are
 * set by name: "qubits", "exec_type", "gpu_count", "block_size",
 * "repeat_count", "gpu_region_bits", "gpu_coalesced_bits", "ops_block",
 * "planner", "kernel", "thread_count", "cpu_region_bits", "cpu_coalesced_bits",
 * "resident". */
typedef struct hgm_dgm hgm_dgm;
int hgm_dgm_create(int dtype, hgm_dgm **out);
void hgm_dgm_destroy(hgm_dgm *d);
int hgm_dgm_set(hgm_dgm *d, const char *field, int64_t value);
int hgm_dgm_get(const hgm_dgm *d, const char *field, int64_t *value);
int hgm_dgm_allocate_memory(hgm_dgm *d);                       /* dgm_core.cpp:54-56 */
int hgm_dgm_free_memory(hgm_dgm *d);
int hgm_dgm_set_memory_value(hgm_dgm *d, int64_t pos);         /* dgm_core.cpp:63-65 */
/* Host amplitudes, brought up to date first; in resident mode the device copy
 * is then treated as stale, so writes through the pointer reach the next
 * execute. NULL when no state is allocated or the copy fails. */
void *hgm_dgm_state(hgm_dgm *d);
int hgm_dgm_add_gate(hgm_dgm *d, const char *name, const double *m8, int *added); /* gates.cpp:42-60 */
int hgm_dgm_get_gate(const hgm_dgm *d, const char *name, double *m8, int *found);
int hgm_dgm_set_function(hgm_dgm *d, const char *function, int iterations, int reset); /* dgm_parser.cpp:29-35 */
int hgm_dgm_set_function_steps(hgm_dgm *d, const char *const *steps, int n_steps, int iterations, int reset);
int hgm_dgm_num_pts(const hgm_dgm *d, int64_t *n);
int hgm_dgm_get_ops(const hgm_dgm *d, hgm_op *out, int64_t cap);
int hgm_dgm_execute(hgm_dgm *d, int iterations);               /* dgm_core.cpp:243-267 */
int hgm_dgm_execute_function(hgm_dgm *d, const char *function, int iterations);
int hgm_dgm_measure(hgm_dgm *d, int qubit, int *outcome);      /* dgm_core.cpp:69-125 */
/* colapse (dgm_core.cpp:129-146): keep the amplitudes whose qubit equals
 * value, renormalise them, zero the rest, and print the kept probability. */
int hgm_dgm_colapse(hgm_dgm *d, int qubit, int value);
/* Marginal distribution over the given qubits: probs[v] for each value v of
 * the masked index bits, listed in ascending mask-value order (dgm_core.cpp:150-162). */
int hgm_dgm_measure_many(hgm_dgm *d, const int *qubits, int n, int64_t *values, double *probs, int64_t *n_out);
int hgm_dgm_sync_host(hgm_dgm *d);                             /* resident mode: download */
int hgm_dgm_stats(const hgm_dgm *d, hgm_stats *out);
int hgm_dgm_count_ops(hgm_dgm *d, int iterations, int64_t counts[7]); /* CountOps, dgm_core.cpp:270-299 */
int hgm_dgm_erase(hgm_dgm *d);                                 /* erase: drop the PT list, dgm_core.cpp:41-52 */
/* validateTuning (dgm_core.cpp:177-241): clamps the region bits to the qubit
 * count, then HGM_E_TUNING (with the reference's message) for a bad t_GPU tuple. */
int hgm_dgm_validate_tuning(hgm_dgm *d);
/* printPTs (dgm_core.cpp:35-39): every PT's fields to stdout, in PT::print's
 * order and formats (common.cpp:68-74). */
int hgm_dgm_print_pts(const hgm_dgm *d);
void hgm_seed_rng(uint64_t seed);                              /* seed_rng, common.cpp:12-14 */
/* Classical steps of Shor's algorithm: continued-fraction approximation of
 * num/den with denominator below 2^width (replaces quantum_frac_approx,
 * src/algorithms/lib_shor_number_theory.cpp:43-71, same mixed float/double
 * arithmetic), modular inverse (mul_inv, :77-90) and bit reversal
 * (revert_bits, :14-22). */
void hgm_frac_approx(int *num, int *den, int width);
int64_t hgm_mul_inv(int64_t value, int64_t modulus);
int hgm_revert_bits(int value, int bits);
/* One attempt of Shor's algorithm (src/algorithms/lib_shor.cpp:17-163) on an
 * executor configured like cfg (dtype, exec_type, tuning, resident; cfg's own
 * state is not touched): out = {factor, N / factor, base, measured phase
 * bits}, factors 0 when the attempt failed (the algorithm is probabilistic).
 * The base and the measurements draw from the generator seeded by hgm_seed_rng. */
int hgm_shor(hgm_dgm *cfg, int64_t number, int64_t out[4], hgm_stats *stats);

/* Circuit step generators and builders by name (src/core/gates.cpp:69-163,
 * src/algorithms/lib_grover.cpp, src/algorithms/lib_shor_circuits.cpp, plus
 * the benchmark circuits). Builders that register gates do so in d's catalog.
 * Multi-step results are joined with ';'. Returns HGM_OK and sets *len to the
 * text length; out may be NULL to query the length. */
int hgm_generate(hgm_dgm *d, const char *which, const int64_t *args, int n_args, const char *op,
                 char *out, size_t cap, size_t *len);

/* ------------------------------------------------------------------ reference ABI
 * Memory layout of the reference's struct PT (include/core/common.h:57-90) on
 * LP64, so that a reference build can hand its PT** straight to this library. */
typedef struct hgm_ref_pt {
    int qubits;
    hgm_cfloat *matrix;
    int matrix_size;
    int span_start_bit, target_bit;
    bool affected;
    long control_value, control_mask;
    long *control_bit_positions, control_count;
    long *control_rest, control_rest_count;
} hgm_ref_pt;

/* replaces setDevice (src/core/kernel.cu:31-33; include/core/dgm.h:22).
 * Like the reference it returns the runtime's error flag: false on success. */
bool setDevice(int device_id);

/* replaces GpuExecutionWrapper (src/core/kernel.cu:303-316; include/core/dgm.h:30).
 * Same argument order. Runs the NULL-terminated PT list on the FP32 host state
 * in place and returns it. With read_memory == NULL it runs on the region that
 * ProjectState placed on the device (qubits must be that region's size) and
 * returns NULL, as src/core/kernel.cu:178-187,285-290 do for t_HYBRID. A gate fires
 * where (index & control_mask) == control_value on the pair's first amplitude,
 * the test of src/core/kernel.cu:102-108 with the arguments of
 * src/core/common.cpp:43-66: a value bit outside the mask (or a 1 asked of
 * the target bit or of a bit beyond the state) never fires, and a zero mask
 * ignores the value. gpu_count must be 1. OPS_BLOCK comes from the
 * HGM_OPS_BLOCK environment variable (default 300). Errors print and exit(1). */
hgm_cfloat *GpuExecutionWrapper(hgm_cfloat *read_memory, hgm_ref_pt **pts, int qubits, int coalesced_bits,
                                int gpu_region_bits, int gpu_count, int block_size, int repeat_count,
                                int iterations);

/* replace ProjectState / GetState (src/core/kernel.cu:321-435; include/core/dgm.h:31-32),
 * the GPU side of t_HYBRID (src/core/dgm_par_exec.cpp:447-451). ProjectState
 * allocates an FP32 device state of 2^region_size amplitudes and gathers into
 * it the host amplitudes region_id | deposit(j, m), where m is the lowest
 * region_size set bits of region_mask inside the state (the reference's
 * portions of 2^coalesced_bits amplitudes, coalesced_bits being the low run of
 * region_mask); GetState scatters it back and frees it. gpu_count must be 1.
 * Both return true; bad arguments print and exit(1), as the reference's
 * error() does. Calls from one thread at a time, as t_HYBRID makes them. */
bool ProjectState(hgm_cfloat *state, int qubits, int region_size, long region_id, long region_mask, int gpu_count);
bool GetState(hgm_cfloat *state, int qubits, int region_size, long region_id, long region_mask, int gpu_count);

#ifdef __cplusplus
}
#endif

#endif /* HGM_GPU_H */
