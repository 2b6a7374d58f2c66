// HIP runtime of the t_GPU path: device context, device-resident states,
// op-table upload, pass launches with HIP-event timing, device readout, and
// the state / plan / one-shot part of the C ABI (include/hgm_gpu.h).
//
// Replaces the host orchestration of src/core/kernel.cu:131-291 (cudaMalloc +
// H2D per call, __constant__ op upload and two device syncs per pass): here
// all passes of a call are planned up front, their op tables go to the device
// in one copy, and the passes are launched back to back on one stream with a
// single synchronisation at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <list>
#include <atomic>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "hgm_gpu.h"
#include "kernels.hpp"
#include "planner.hpp"
#include "runtime.hpp"
#include "jit.hpp"

struct hgm_state {
    void *d = nullptr;
    int qubits = 0;
    int dtype = HGM_C128;
    bool owned = true;
    size_t bytes = 0;
};

struct hgm_plan {
    hgm::Plan plan;
    // hipGraph replay of a plan-compiled plan (hgm_plan_set_graph): the graph
    // holds the pass launches for one state pointer and one loaded module
    bool graph_on = false;
    mutable void *graph_state = nullptr;
    mutable uint64_t graph_module_gen = 0;
    mutable hipGraph_t graph = nullptr;
    mutable hipGraphExec_t graph_exec = nullptr;
    void drop_graph() const {
        if (graph_exec) hipGraphExecDestroy(graph_exec);
        if (graph) hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_state = nullptr;
    }
};

// A batch of states advanced by their plans as one captured graph
// (hgm_batch_create): branch[i] is the fork branch of state i. 64 states of
// 12-19 qubits (Grover, 1016 passes) replay in 3.78 ms on 4 branches, 3.14 ms
// on 16 and 2.91 ms on 64, against 7.1 ms plan by plan.
constexpr int kBatchBranches = 64;
struct hgm_batch {
    std::vector<hgm_state *> states;
    std::vector<const hgm_plan *> plans;
    std::vector<int> branch;
    int n_branches = 0;
    hipStream_t streams[kBatchBranches] = {};
    hipEvent_t fork = nullptr, join[kBatchBranches] = {};
    std::vector<uint64_t> gens;   // load generation of each plan's module when the graph was captured
    void *tables = nullptr;       // device copies of the register / LDS plans' tables
    std::vector<int64_t> table_off;   // per state: byte offset of its plan's tables (-1: plan-compiled)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void drop_graph() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

namespace hgm {

namespace {

thread_local std::string t_err;

struct Ctx {
    std::mutex mu;
    bool ready = false;
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr, e4 = nullptr, e5 = nullptr;
    int cus = 0;
    DevOp *ops = nullptr;       // op tables of the current call (grow-only)
    size_t ops_cap = 0;
    void *rops = nullptr;       // register-kernel programs, phase maps, pass tables (grow-only)
    size_t rops_cap = 0;
    void *maps = nullptr;
    size_t maps_cap = 0;
    void *mems = nullptr;
    size_t mems_cap = 0;
    void *cache = nullptr;      // device buffer of hgm_execute (grow-only)
    size_t cache_cap = 0;
    double *partial = nullptr;  // reduction partials (device, 8 doubles per block)
    uint64_t *pidx = nullptr;
    double *h_partial = nullptr;  // pinned host mirror
    uint64_t *h_pidx = nullptr;
    char *h_stage[2] = {nullptr, nullptr};          // pinned chunks of the region copies
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

struct JitModule {
    hipModule_t mod = nullptr;
    std::vector<hipFunction_t> fns;
    size_t binary_size = 0;     // with binary_hash, compared on lookup: the source hash only narrows the search
    uint64_t binary_hash = 0;
    uint64_t gen = 0;           // load generation: graphs captured with an older one are stale
    uint64_t last_use = 0;
};

// Loaded plan-compiled modules, least recently used first out beyond
// kMaxJitModules (unloaded after the library's stream drains; a plan whose
// module went out is loaded again on its next run, and its captured graph,
// tagged with the old load generation, is captured again).
constexpr size_t kMaxJitModules = 16;
uint64_t g_module_tick = 0;

uint64_t g_module_gen = 0;

// Workspace of hgm_state_apply_noalloc: the plan tables (register programs,
// phase maps and pass tables, or op tables) of one call. Each GPU holds its
// own copy for the life of the process. In host code `scratch` is only a
// stand-in for the symbol: its device address is looked up once per device
// (init_locked) and kept in g_scratch_ptr, and only that address is used.
constexpr size_t kScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};

Ctx &ctx() {
    static Ctx c;
    return c;
}

constexpr int kRedBlocks = 1024;

std::unordered_map<uint64_t, JitModule> &jit_modules() {
    static std::unordered_map<uint64_t, JitModule> m;
    return m;
}

int hip_fail(hipError_t e, const char *what) {
    t_err = std::string(what) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();   // a failed call leaves its error for the next launch check to find
    return HGM_E_HIP;
}

#define HGM_HIP(call, what)                            \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

int init_locked(int device) {
    Ctx &c = ctx();
    if (c.ready && (device < 0 || device == c.device)) return HGM_OK;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        t_err = "no HIP device is visible (the t_GPU path needs an MI355X)";
        return HGM_E_HIP;
    }
    if (device < 0) device = 0;
    if (device >= count) { t_err = "device index out of range"; return HGM_E_ARG; }
    HGM_HIP(hipSetDevice(device), "hipSetDevice");
    if (c.stream) {   // re-init on another device
        hipStreamDestroy(c.stream);
        c.stream = nullptr;
    }
    HGM_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate");
    for (hipEvent_t *e : {&c.e0, &c.e1, &c.e2, &c.e3, &c.e4, &c.e5}) {
        if (!*e) HGM_HIP(hipEventCreate(e), "hipEventCreate");
    }
    hipDeviceProp_t prop;
    HGM_HIP(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    c.cus = prop.multiProcessorCount;
    if (!c.partial) {
        HGM_HIP(hipMalloc(&c.partial, 8 * kRedBlocks * sizeof(double)), "hipMalloc(partials)");
        HGM_HIP(hipMalloc(&c.pidx, kRedBlocks * sizeof(uint64_t)), "hipMalloc(partials)");
        HGM_HIP(hipHostMalloc(&c.h_partial, 8 * kRedBlocks * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
        HGM_HIP(hipHostMalloc(&c.h_pidx, kRedBlocks * sizeof(uint64_t), hipHostMallocDefault), "hipHostMalloc");
    }
    if (device < kMaxDevices && !g_scratch_ptr[device]) {
        void *ptr = nullptr;
        HGM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = ptr;
    }
    c.device = device;
    c.ready = true;
    return HGM_OK;
}

// Loads a compiled pass module once per process (by source hash). Assumes the lock.
int jit_load_locked(const JitCode &jc, JitModule **out) {
    auto &mods = jit_modules();
    auto it = mods.find(jc.hash);
    if (it != mods.end() && (it->second.binary_size != jc.binary.size() || it->second.binary_hash != jc.binary_hash)) {
        // another plan's code under the same source hash
        hipStreamSynchronize(ctx().stream);
        hipModuleUnload(it->second.mod);
        mods.erase(it);
        it = mods.end();
    }
    if (it == mods.end()) {
        while (mods.size() >= kMaxJitModules) {
            auto lru = mods.begin();
            for (auto m = mods.begin(); m != mods.end(); ++m)
                if (m->second.last_use < lru->second.last_use) lru = m;
            hipStreamSynchronize(ctx().stream);   // no kernel of it is still queued
            hipModuleUnload(lru->second.mod);
            mods.erase(lru);
        }
        JitModule m;
        HGM_HIP(hipModuleLoadData(&m.mod, jc.binary.data()), "hipModuleLoadData(pass kernels)");
        for (const std::string &name : jc.names) {
            hipFunction_t f;
            const hipError_t e = hipModuleGetFunction(&f, m.mod, name.c_str());
            if (e != hipSuccess) {
                hipModuleUnload(m.mod);
                return hip_fail(e, "hipModuleGetFunction");
            }
            m.fns.push_back(f);
        }
        m.binary_size = jc.binary.size();
        m.binary_hash = jc.binary_hash;
        m.gen = ++g_module_gen;
        it = mods.emplace(jc.hash, std::move(m)).first;
    }
    it->second.last_use = ++g_module_tick;
    *out = &it->second;
    return HGM_OK;
}

int reserve_ops(size_t n) {
    Ctx &c = ctx();
    if (n <= c.ops_cap) return HGM_OK;
    if (c.ops) { hipStreamSynchronize(c.stream); hipFree(c.ops); c.ops = nullptr; c.ops_cap = 0; }
    size_t cap = n < 1024 ? 1024 : n + n / 2;
    HGM_HIP(hipMalloc(&c.ops, cap * sizeof(DevOp)), "hipMalloc(op tables)");
    c.ops_cap = cap;
    return HGM_OK;
}

int reserve(void **buf, size_t *cap, size_t bytes, const char *what) {
    Ctx &c = ctx();
    if (bytes <= *cap) return HGM_OK;
    if (*buf) { hipStreamSynchronize(c.stream); hipFree(*buf); *buf = nullptr; *cap = 0; }
    size_t want = bytes < 4096 ? 4096 : bytes + bytes / 2;
    HGM_HIP(hipMalloc(buf, want), what);
    *cap = want;
    return HGM_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    return ms;
}

// One pass of a plan-compiled plan: one tile per workgroup, LDS slots padded
// by 1/16 + 1/256. On the context stream unless `stream` is given. Assumes
// the lock.
int launch_jit_pass_locked(hgm_state *s, const Plan &p, JitModule *jm, size_t pi, hipStream_t stream = nullptr) {
    Ctx &c = ctx();
    const Pass &ps = p.passes[pi];
    const size_t jit_lds = p.jit->lds_bytes;
    const int64_t n_tiles = int64_t(1) << (p.qubits - p.shape.r);
    void *state = s->d;
    void *args[] = {&state};
    HGM_HIP(hipModuleLaunchKernel(jm->fns[p.jit->pass_kernel[pi]], (unsigned)n_tiles, 1, 1, (unsigned)p.shape.B, 1,
                                  1, ps.phases > 1 ? (unsigned)jit_lds : 0u, stream ? stream : c.stream, args,
                                  nullptr),
            "pass launch");
    return HGM_OK;
}

void add_stats(hgm_stats *st, const Plan &p, float kernel_ms, double plan_ms) {
    if (!st) return;
    int64_t launched = 0;
    for (const Pass &ps : p.passes) launched += ps.op_count > 0;
    st->gates += p.gates;
    st->passes += launched;
    st->amp_updates += p.gates * (int64_t(1) << p.qubits);
    st->bytes_alg += (double)launched * 2.0 * (double)(uint64_t(1) << p.qubits) * (double)amp_bytes(p.dtype);
    st->plan_ms += plan_ms;
    st->kernel_ms += kernel_ms;
}

// Replays the plan's captured graph on s, capturing it first when there is
// none for this state and this load of the plan's kernels. Assumes the lock.
int run_plan_graph_locked(hgm_state *s, const hgm_plan *hp, hgm_stats *st) {
    Ctx &c = ctx();
    const Plan &p = hp->plan;
    if (s->qubits != p.qubits || s->dtype != p.dtype) {
        t_err = "plan was built for another qubit count or dtype";
        return HGM_E_STATE;
    }
    JitModule *jm = nullptr;
    int rc = jit_load_locked(*p.jit, &jm);
    if (rc) return rc;
    if (!hp->graph_exec || hp->graph_state != s->d || hp->graph_module_gen != jm->gen) {
        hp->drop_graph();
        HGM_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        for (size_t pi = 0; pi < p.passes.size() && !rc; ++pi)
            if (p.passes[pi].op_count) rc = launch_jit_pass_locked(s, p, jm, pi);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c.stream, &g);
        if (rc || e != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return rc ? rc : hip_fail(e, "hipStreamEndCapture");
        }
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            hipGraphDestroy(g);
            return hip_fail(ei, "hipGraphInstantiate");
        }
        hp->graph = g;
        hp->graph_exec = x;
        hp->graph_state = s->d;
        hp->graph_module_gen = jm->gen;
    }
    HGM_HIP(hipEventRecord(c.e0, c.stream), "hipEventRecord");
    HGM_HIP(hipGraphLaunch(hp->graph_exec, c.stream), "hipGraphLaunch");
    HGM_HIP(hipEventRecord(c.e1, c.stream), "hipEventRecord");
    HGM_HIP(hipEventSynchronize(c.e1), "pass execution");
    add_stats(st, p, elapsed(c.e0, c.e1), p.plan_ms);
    return HGM_OK;
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Bytes of device tables the register or LDS kernels read for plan p.
size_t plan_table_bytes(const Plan &p) {
    if (p.kernel == HGM_KERNEL_REG)
        return align256(p.rops.size() * sizeof(RegOp)) + align256(p.maps.size() * sizeof(PhaseMap)) +
               align256(p.mems.size() * sizeof(PassMem));
    return align256(p.ops.size() * sizeof(DevOp));
}

// Device tables of a register or LDS plan: where its kernels read their
// programs (unused by plan-compiled plans).
struct PlanTables {
    const void *rops, *maps, *mems;
    const DevOp *dops;
};

// Launches every non-empty pass of the plan on `stream`, with no events or
// waits (jm: the plan's loaded module, for plan-compiled plans). Assumes the
// lock.
int launch_passes_locked(hgm_state *s, const Plan &p, JitModule *jm, const PlanTables &t, hipStream_t stream) {
    const bool reg = p.kernel == HGM_KERNEL_REG;
    const bool jit = p.kernel == HGM_KERNEL_JIT;
    PassArgs a;
    std::memset(&a, 0, sizeof(a));
    a.r = p.shape.r; a.c = p.shape.c; a.B = p.shape.B; a.R = p.shape.R;
    a.n_tiles = int64_t(1) << (p.qubits - p.shape.r);
    const uint64_t all = (uint64_t(1) << p.qubits) - 1;
    const size_t tile_bytes = (size_t(1) << p.shape.r) * amp_bytes(p.dtype);
    for (size_t pi = 0; pi < p.passes.size(); ++pi) {
        const Pass &ps = p.passes[pi];
        if (ps.op_count == 0) continue;
        a.nonlocal_mask = all & ~ps.local_mask;
        if (jit) {
            if (int rc = launch_jit_pass_locked(s, p, jm, pi, stream)) return rc;
        } else if (reg) {
            HGM_HIP(launch_pass_reg(p.dtype, 2 * p.shape.R, s->d, static_cast<const RegOp *>(t.rops) + ps.rop_begin,
                                    (int)ps.rop_count, static_cast<const PhaseMap *>(t.maps),
                                    static_cast<const PassMem *>(t.mems) + ps.mem, a.nonlocal_mask, a.n_tiles,
                                    p.shape.r, p.shape.B, ps.phases > 1 ? tile_bytes : 0, stream),
                    "pass launch");
        } else {
            std::memcpy(a.local_pos, ps.local_pos, sizeof(a.local_pos));
            HGM_HIP(launch_pass_lds(p.dtype, s->d, t.dops + ps.op_begin, (int)ps.op_count, a, stream), "pass launch");
        }
    }
    return HGM_OK;
}

// Launches every pass of the plan on the context stream. Assumes the lock.
// With `workspace` (at least plan_table_bytes(p) bytes of device memory) the
// register and LDS kernels read their tables from there instead of from the
// context's grow-only buffers, and nothing is allocated.
int run_plan_locked(hgm_state *s, const Plan &p, hgm_stats *st, double plan_ms, void *workspace = nullptr) {
    Ctx &c = ctx();
    if (s->qubits != p.qubits || s->dtype != p.dtype) {
        t_err = "plan was built for another qubit count or dtype";
        return HGM_E_STATE;
    }
    const bool reg = p.kernel == HGM_KERNEL_REG;
    const bool jit = p.kernel == HGM_KERNEL_JIT;
    int rc;
    JitModule *jm = nullptr;
    if (jit) {
        if ((rc = jit_load_locked(*p.jit, &jm))) return rc;
    }
    void *rops = c.rops, *maps = c.maps, *mems = c.mems;
    DevOp *dops = c.ops;
    if (workspace && !jit) {
        char *w = static_cast<char *>(workspace);
        if (reg) {
            rops = w;
            maps = w + align256(p.rops.size() * sizeof(RegOp));
            mems = static_cast<char *>(maps) + align256(p.maps.size() * sizeof(PhaseMap));
        } else {
            dops = reinterpret_cast<DevOp *>(w);
        }
    } else if (reg) {
        if ((rc = reserve(&c.rops, &c.rops_cap, p.rops.size() * sizeof(RegOp), "hipMalloc(programs)"))) return rc;
        if ((rc = reserve(&c.maps, &c.maps_cap, p.maps.size() * sizeof(PhaseMap), "hipMalloc(maps)"))) return rc;
        if ((rc = reserve(&c.mems, &c.mems_cap, p.mems.size() * sizeof(PassMem), "hipMalloc(pass tables)"))) return rc;
        rops = c.rops;
        maps = c.maps;
        mems = c.mems;
    } else if (!jit) {
        if ((rc = reserve_ops(p.ops.size()))) return rc;
        dops = c.ops;
    }
    if (reg) {
        if (!p.rops.empty())
            HGM_HIP(hipMemcpyAsync(rops, p.rops.data(), p.rops.size() * sizeof(RegOp), hipMemcpyHostToDevice,
                                   c.stream), "hipMemcpyAsync(programs)");
        if (!p.maps.empty())
            HGM_HIP(hipMemcpyAsync(maps, p.maps.data(), p.maps.size() * sizeof(PhaseMap), hipMemcpyHostToDevice,
                                   c.stream), "hipMemcpyAsync(maps)");
        if (!p.mems.empty())
            HGM_HIP(hipMemcpyAsync(mems, p.mems.data(), p.mems.size() * sizeof(PassMem), hipMemcpyHostToDevice,
                                   c.stream), "hipMemcpyAsync(pass tables)");
    } else if (!jit && !p.ops.empty()) {
        HGM_HIP(hipMemcpyAsync(dops, p.ops.data(), p.ops.size() * sizeof(DevOp), hipMemcpyHostToDevice, c.stream),
                "hipMemcpyAsync(op tables)");
    }
    HGM_HIP(hipEventRecord(c.e0, c.stream), "hipEventRecord");
    if ((rc = launch_passes_locked(s, p, jm, {rops, maps, mems, dops}, c.stream))) return rc;
    HGM_HIP(hipEventRecord(c.e1, c.stream), "hipEventRecord");
    HGM_HIP(hipEventSynchronize(c.e1), "pass execution");
    add_stats(st, p, elapsed(c.e0, c.e1), plan_ms);
    return HGM_OK;
}

int blocks_for(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    return (int)(b < (uint64_t)kRedBlocks ? (b ? b : 1) : kRedBlocks);
}

int prob_halves_locked(const hgm_state *s, int bit, double *zero, double *one) {
    Ctx &c = ctx();
    const uint64_t n = uint64_t(1) << s->qubits;
    const int blocks = blocks_for(n);
    HGM_HIP(launch_prob_halves(s->dtype, s->d, n, bit, c.partial, blocks, c.stream), "prob kernel");
    HGM_HIP(hipMemcpyAsync(c.h_partial, c.partial, 2 * blocks * sizeof(double), hipMemcpyDeviceToHost, c.stream),
            "hipMemcpyAsync(partials)");
    HGM_HIP(hipStreamSynchronize(c.stream), "prob kernel");
    double z = 0, o = 0;
    for (int b = 0; b < blocks; ++b) { z += c.h_partial[2 * b]; o += c.h_partial[2 * b + 1]; }
    *zero = z;
    *one = o;
    return HGM_OK;
}

// Loads the module of every plan of the batch and reports whether any load
// generation differs from the one its graph was captured with. Loading
// touches each module, so the batch's (at most kMaxJitModules distinct)
// modules are the most recently used ones and loading one never unloads
// another of them. Assumes the lock.
int batch_modules_locked(hgm_batch *b, std::vector<JitModule *> *jms, bool *stale) {
    jms->assign(b->plans.size(), nullptr);
    for (size_t i = 0; i < b->plans.size(); ++i)
        if (b->plans[i]->plan.kernel == HGM_KERNEL_JIT)
            if (int rc = jit_load_locked(*b->plans[i]->plan.jit, &(*jms)[i])) return rc;
    *stale = !b->exec || b->gens.size() != b->plans.size();
    for (size_t i = 0; i < b->plans.size() && !*stale; ++i) *stale = (*jms)[i] && b->gens[i] != (*jms)[i]->gen;
    return HGM_OK;
}

// Where the kernels of state i's register or LDS plan read their tables: the
// batch's device copy, laid out as plan_table_bytes counts it.
PlanTables batch_tables(const hgm_batch *b, size_t i) {
    PlanTables t{nullptr, nullptr, nullptr, nullptr};
    if (b->table_off[i] < 0) return t;
    const Plan &p = b->plans[i]->plan;
    char *base = static_cast<char *>(b->tables) + b->table_off[i];
    if (p.kernel == HGM_KERNEL_REG) {
        t.rops = base;
        t.maps = base + align256(p.rops.size() * sizeof(RegOp));
        t.mems = static_cast<const char *>(t.maps) + align256(p.maps.size() * sizeof(PhaseMap));
    } else {
        t.dops = reinterpret_cast<const DevOp *>(base);
    }
    return t;
}

// Captures the batch: a fork on the context stream, each state's passes in
// order on its branch stream, a join back. Assumes the lock.
int capture_batch_locked(hgm_batch *b, const std::vector<JitModule *> &jms) {
    Ctx &c = ctx();
    b->drop_graph();
    HGM_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    int rc = HGM_OK;
    hipError_t e = hipEventRecord(b->fork, c.stream);
    for (int k = 0; k < b->n_branches && e == hipSuccess; ++k) e = hipStreamWaitEvent(b->streams[k], b->fork, 0);
    for (size_t i = 0; i < b->plans.size() && e == hipSuccess && !rc; ++i)
        rc = launch_passes_locked(b->states[i], b->plans[i]->plan, jms[i], batch_tables(b, i), b->streams[b->branch[i]]);
    for (int k = 0; k < b->n_branches && e == hipSuccess && !rc; ++k) {
        e = hipEventRecord(b->join[k], b->streams[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(c.stream, b->join[k], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(c.stream, &g);
    if (rc || e != hipSuccess || ee != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return rc ? rc : hip_fail(e != hipSuccess ? e : ee, "batch graph capture");
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        hipGraphDestroy(g);
        return hip_fail(ei, "hipGraphInstantiate");
    }
    b->graph = g;
    b->exec = x;
    b->gens.resize(b->plans.size());
    for (size_t i = 0; i < b->plans.size(); ++i) b->gens[i] = jms[i] ? jms[i]->gen : 0;
    return HGM_OK;
}

}  // namespace

void set_error(const std::string &msg) { t_err = msg; }
const std::string &last_error() { return t_err; }

int gpu_init(int device) {
    std::lock_guard<std::mutex> g(ctx().mu);
    return init_locked(device);
}

int alloc_pinned(void **p, size_t bytes) {
    if (hipHostMalloc(p, bytes, hipHostMallocDefault) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return HGM_E_NOMEM;
    }
    // zero-fill like calloc; large states in parallel chunks (one thread
    // writes a few GB/s, an 8 GiB FP32 30-qubit state would take seconds)
    char *base = static_cast<char *>(*p);
    const size_t chunk = size_t(64) << 20;
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt < 1 ? 1 : (nt > 16 ? 16 : nt);
    if (bytes < 2 * chunk || nt == 1) {
        std::memset(base, 0, bytes);
        return HGM_OK;
    }
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t off; (off = next.fetch_add(chunk)) < bytes;)
            std::memset(base + off, 0, std::min(chunk, bytes - off));
    };
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < nt; ++i) pool.emplace_back(work);
    work();
    for (std::thread &t : pool) t.join();
    return HGM_OK;
}

void free_pinned(void *p) {
    if (p) hipHostFree(p);
}

// Plans built by the one-call entry points (hgm_execute, hgm_state_apply) for
// the last few (gate list, tuning, iterations, qubits, dtype): a caller that
// runs the same circuit again, as the drop-in executor does on every
// execute(), skips planning and code generation. Keys hold the whole gate
// list and are compared field by field.
struct CachedPlan {
    int qubits, dtype, iterations;
    hgm_tuning t;
    std::vector<hgm_op> ops;
    std::shared_ptr<const Plan> plan;
};
constexpr size_t kPlanCacheEntries = 4;
std::mutex g_plan_cache_mu;
std::list<CachedPlan> g_plan_cache;   // most recently used first

bool same_tuning(const hgm_tuning &a, const hgm_tuning &b) {
    return a.block_size == b.block_size && a.repeat_count == b.repeat_count &&
           a.gpu_region_bits == b.gpu_region_bits && a.gpu_coalesced_bits == b.gpu_coalesced_bits &&
           a.ops_block == b.ops_block && a.planner == b.planner && a.kernel == b.kernel && a.lookahead == b.lookahead;
}

bool same_ops(const std::vector<hgm_op> &a, const hgm_op *b, int64_t n) {
    if ((int64_t)a.size() != n) return false;
    for (int64_t i = 0; i < n; ++i)
        if (a[i].target_bit != b[i].target_bit || a[i].control_mask != b[i].control_mask ||
            a[i].control_value != b[i].control_value || std::memcmp(a[i].m, b[i].m, sizeof(a[i].m)) != 0)
            return false;
    return true;
}

int cached_plan(int qubits, int dtype, const hgm_op *ops, int64_t n_ops, int iterations, const hgm_tuning &t,
                std::shared_ptr<const Plan> *out, double *plan_ms) {
    const auto t0 = std::chrono::steady_clock::now();
    if (n_ops < 0 || (n_ops > 0 && !ops)) { t_err = "bad op array"; return HGM_E_ARG; }
    {
        std::lock_guard<std::mutex> g(g_plan_cache_mu);
        for (auto it = g_plan_cache.begin(); it != g_plan_cache.end(); ++it)
            if (it->qubits == qubits && it->dtype == dtype && it->iterations == iterations && same_tuning(it->t, t) &&
                same_ops(it->ops, ops, n_ops)) {
                g_plan_cache.splice(g_plan_cache.begin(), g_plan_cache, it);
                *out = g_plan_cache.front().plan;
                *plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                return HGM_OK;
            }
    }
    auto p = std::make_shared<Plan>();
    int rc = build_plan(qubits, dtype, ops, n_ops, iterations, t, p.get(), &t_err);
    if (rc) return rc;
    *plan_ms = p->plan_ms;
    std::lock_guard<std::mutex> g(g_plan_cache_mu);
    g_plan_cache.push_front(CachedPlan{qubits, dtype, iterations, t, std::vector<hgm_op>(ops, ops + n_ops), p});
    if (g_plan_cache.size() > kPlanCacheEntries) g_plan_cache.pop_back();
    *out = std::move(p);
    return HGM_OK;
}

int plan_and_run(hgm_state *s, const hgm_op *ops, int64_t n_ops, const hgm_tuning *tuning, int iterations,
                 hgm_stats *st) {
    hgm_tuning t;
    if (tuning) t = *tuning; else hgm_default_tuning(&t);
    std::shared_ptr<const Plan> p;
    double plan_ms = 0;
    int rc = cached_plan(s->qubits, s->dtype, ops, n_ops, iterations, t, &p, &plan_ms);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(ctx().mu);
    rc = init_locked(-1);
    if (rc) return rc;
    return run_plan_locked(s, *p, st, plan_ms);
}

}  // namespace hgm

using namespace hgm;

extern "C" {

void hgm_default_tuning(hgm_tuning *t) {
    std::memset(t, 0, sizeof(*t));
    t->block_size = 64;          // include/cli/cli_common.h:14-16
    t->repeat_count = 2;
    t->gpu_region_bits = 8;
    t->gpu_coalesced_bits = 4;
    t->ops_block = 300;          // makefile:7
    t->planner = HGM_PLAN_GREEDY;
    t->kernel = HGM_KERNEL_AUTO;
    t->lookahead = 0;
}

const char *hgm_last_error(void) { return t_err.c_str(); }

const char *hgm_version(void) { return "hgm-mi355x 0.1 (gfx950)"; }

int hgm_init(int device) { return gpu_init(device); }

int hgm_device_count(int *count) {
    if (!count) return HGM_E_ARG;
    if (hipGetDeviceCount(count) != hipSuccess) *count = 0;
    return HGM_OK;
}

int hgm_release_cache(void) {
    {
        std::lock_guard<std::mutex> g(g_plan_cache_mu);
        g_plan_cache.clear();
    }
    Ctx &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (!c.ready) return HGM_OK;
    hipStreamSynchronize(c.stream);
    auto drop = [](void **buf, size_t *cap) {
        if (*buf) hipFree(*buf);
        *buf = nullptr;
        *cap = 0;
    };
    drop(&c.cache, &c.cache_cap);
    drop(reinterpret_cast<void **>(&c.ops), &c.ops_cap);
    drop(&c.rops, &c.rops_cap);
    drop(&c.maps, &c.maps_cap);
    drop(&c.mems, &c.mems_cap);
    for (char *&h : c.h_stage) {
        if (h) hipHostFree(h);
        h = nullptr;
    }
    for (auto &kv : jit_modules()) hipModuleUnload(kv.second.mod);
    jit_modules().clear();
    return HGM_OK;
}

int hgm_validate_tuning(int qubits, int dtype, const hgm_tuning *t) {
    if (!t) { t_err = "null tuning"; return HGM_E_ARG; }
    TileShape s;
    return make_shape(qubits, dtype, *t, &s, &t_err);
}

int hgm_state_create(int qubits, int dtype, hgm_state **out) {
    if (!out || qubits < 1 || qubits > 40 || (dtype != HGM_C128 && dtype != HGM_C64)) {
        t_err = "hgm_state_create: bad arguments";
        return HGM_E_ARG;
    }
    Ctx &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = init_locked(-1);
    if (rc) return rc;
    hgm_state *s = new hgm_state();
    s->qubits = qubits;
    s->dtype = dtype;
    s->bytes = (size_t(1) << qubits) * amp_bytes(dtype);
    if (hipMalloc(&s->d, s->bytes) != hipSuccess) {
        delete s;
        hipGetLastError();
        t_err = "hipMalloc of a " + std::to_string(qubits) + "-qubit state failed";
        return HGM_E_NOMEM;
    }
    hipError_t e = launch_set_basis(dtype, s->d, uint64_t(1) << qubits, ~uint64_t(0), c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) { hipFree(s->d); delete s; return hip_fail(e, "state clear"); }
    *out = s;
    return HGM_OK;
}

int hgm_state_wrap(void *device_ptr, int qubits, int dtype, hgm_state **out) {
    if (!out || !device_ptr || qubits < 1 || qubits > 40 || (dtype != HGM_C128 && dtype != HGM_C64)) {
        t_err = "hgm_state_wrap: bad arguments";
        return HGM_E_ARG;
    }
    int rc = gpu_init(-1);
    if (rc) return rc;
    hgm_state *s = new hgm_state();
    s->d = device_ptr;
    s->qubits = qubits;
    s->dtype = dtype;
    s->owned = false;
    s->bytes = (size_t(1) << qubits) * amp_bytes(dtype);
    *out = s;
    return HGM_OK;
}

void hgm_state_destroy(hgm_state *s) {
    if (!s) return;
    if (s->owned && s->d) {
        std::lock_guard<std::mutex> g(ctx().mu);
        hipStreamSynchronize(ctx().stream);
        hipFree(s->d);
    }
    delete s;
}

int hgm_state_qubits(const hgm_state *s) { return s ? s->qubits : -1; }
int hgm_state_dtype(const hgm_state *s) { return s ? s->dtype : -1; }
void *hgm_state_device_ptr(hgm_state *s) { return s ? s->d : nullptr; }

int hgm_state_set_basis(hgm_state *s, uint64_t index) {
    if (!s || index >= (uint64_t(1) << s->qubits)) { t_err = "basis index outside the state"; return HGM_E_ARG; }
    Ctx &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    HGM_HIP(launch_set_basis(s->dtype, s->d, uint64_t(1) << s->qubits, index, c.stream), "set_basis");
    HGM_HIP(hipStreamSynchronize(c.stream), "set_basis");
    return HGM_OK;
}

int hgm_state_copy(hgm_state *dst, const hgm_state *src) {
    if (!dst || !src || dst->bytes != src->bytes || dst->dtype != src->dtype) {
        t_err = "hgm_state_copy: states differ in size or dtype";
        return HGM_E_ARG;
    }
    Ctx &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    HGM_HIP(launch_copy(dst->d, src->d, src->bytes, c.stream), "copy");
    HGM_HIP(hipStreamSynchronize(c.stream), "copy");
    return HGM_OK;
}

int hgm_state_upload(hgm_state *s, const void *host) {
    if (!s || !host) { t_err = "null argument"; return HGM_E_ARG; }
    Ctx &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    HGM_HIP(hipMemcpyAsync(s->d, host, s->bytes, hipMemcpyHostToDevice, c.stream), "upload");
    HGM_HIP(hipStreamSynchronize(c.stream), "upload");
    return HGM_OK;
}

int hgm_state_download(const hgm_state *s, void *host) {
    if (!s || !host) { t_err = "null argument"; return HGM_E_ARG; }
    Ctx &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    HGM_HIP(hipMemcpyAsync(host, s->d, s->bytes, hipMemcpyDeviceToHost, c.stream), "download");
    HGM_HIP(hipStreamSynchronize(c.stream), "download");
    return HGM_OK;
}

}  // extern "C"

namespace {

// Region copies between a host state and a smaller device state. Device
// amplitude j is host amplitude region_id | deposit(j, region_mask). The low
// run of region_mask makes portions of 2^run consecutive amplitudes (the
// reference's coalesced bits, src/core/kernel.cu:330-343); the portions are
// visited in device order by the subset walk base -> ((base | ~hi) + 1) & hi
// over the remaining mask bits. The host side is pageable memory, so the bytes
// go through two pinned chunks: the host copy into one chunk overlaps the DMA
// of the other.
constexpr size_t kStageBytes = size_t(32) << 20;

struct RegionWalk {
    const uint64_t id, hi;   // fixed bits; region bits above the run, shifted down by run
    const int run;
    const size_t amp, portion;
    uint64_t base = 0;       // deposited portion index
    size_t within = 0;       // byte offset inside the current portion
    // positioned at device byte offset start
    RegionWalk(uint64_t region_id, uint64_t mask, size_t amp_bytes, size_t start)
        : id(region_id), hi(mask >> __builtin_ctzll(~mask)), run(__builtin_ctzll(~mask)), amp(amp_bytes),
          portion(amp_bytes << __builtin_ctzll(~mask)) {
        uint64_t p = start / portion, m = hi;
        within = start % portion;
        for (; p && m; p >>= 1, m &= m - 1)   // deposit p on the bits of hi
            if (p & 1) base |= m & (~m + 1);
    }
    // the next n bytes in device order, as host runs: fn(host_byte_offset, bytes, offset_in_n)
    template <class F>
    void next(size_t n, F fn) {
        for (size_t w = 0; w < n;) {
            const size_t k = std::min(portion - within, n - w);
            fn(((base << run) | id) * amp + within, k, w);
            w += k;
            within += k;
            if (within == portion) {
                within = 0;
                base = ((base | ~hi) + 1) & hi;
            }
        }
    }
};

// Host side of one staged chunk: bytes [off, off + n) in device order, split
// over up to 8 threads when the chunk is large (one thread moves pageable
// memory at about 10 GB/s, below what the DMA of the other chunk takes).
template <class F>
void chunk_host_copy(uint64_t id, uint64_t mask, size_t amp, size_t off, size_t n, F fn) {
    constexpr size_t kPerThread = size_t(1) << 20;
    unsigned nt = std::thread::hardware_concurrency();
    nt = (unsigned)std::min<size_t>({nt ? nt : 1u, 8u, n / kPerThread});
    if (nt <= 1) {
        RegionWalk(id, mask, amp, off).next(n, fn);
        return;
    }
    const size_t part = n / nt / amp * amp;
    auto run_part = [&](unsigned t) {
        const size_t b = t * part, e = t + 1 == nt ? n : b + part;
        RegionWalk(id, mask, amp, off + b).next(e - b, [&](size_t h, size_t k, size_t at) { fn(h, k, b + at); });
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(run_part, t);
    run_part(0);
    for (std::thread &t : pool) t.join();
}

int region_check(const hgm_state *s, const void *host, int host_qubits, uint64_t region_id, uint64_t region_mask) {
    if (!s || !host) { t_err = "null argument"; return HGM_E_ARG; }
    if (host_qubits < s->qubits || host_qubits > 40) { t_err = "region copy: host state smaller than the region"; return HGM_E_ARG; }
    const uint64_t all = (uint64_t(1) << host_qubits) - 1;
    if ((region_mask & ~all) || (region_id & ~all) || (region_id & region_mask) ||
        __builtin_popcountll(region_mask) != s->qubits) {
        t_err = "region copy: region_mask must hold " + std::to_string(s->qubits) +
                " bits of the state and region_id none of them";
        return HGM_E_ARG;
    }
    return HGM_OK;
}

// Chunk of a staged copy of `bytes` (a power of two): 32 MiB for large copies,
// down to 8 MiB so that mid-size ones still pipeline over at least 4 chunks.
size_t stage_chunk(size_t bytes) {
    const size_t quarter = std::max<size_t>(bytes / 4, size_t(8) << 20);
    return std::min({kStageBytes, quarter, bytes});
}

int stage_locked() {
    Ctx &c = ctx();
    for (int i = 0; i < 2; ++i) {
        if (!c.stage_ev[i]) HGM_HIP(hipEventCreateWithFlags(&c.stage_ev[i], hipEventDisableTiming), "hipEventCreate");
        if (!c.h_stage[i])
            HGM_HIP(hipHostMalloc(reinterpret_cast<void **>(&c.h_stage[i]), kStageBytes, hipHostMallocDefault),
                    "hipHostMalloc(staging)");
    }
    return HGM_OK;
}

// Host -> device through the staging chunks: device bytes [0, bytes) in
// order are the host bytes the region walk (id, mask) names. Assumes the lock.
int staged_h2d_locked(char *d, const char *h, size_t bytes, uint64_t id, uint64_t mask, size_t amp) {
    Ctx &c = ctx();
    int rc = stage_locked();
    if (rc) return rc;
    const size_t chunk = stage_chunk(bytes);   // powers of two: chunks tile the state
    for (size_t off = 0, i = 0; off < bytes; off += chunk, ++i) {
        char *buf = c.h_stage[i & 1];
        if (i >= 2) HGM_HIP(hipEventSynchronize(c.stage_ev[i & 1]), "staged H2D");
        chunk_host_copy(id, mask, amp, off, chunk, [&](size_t src, size_t n, size_t at) { std::memcpy(buf + at, h + src, n); });
        HGM_HIP(hipMemcpyAsync(d + off, buf, chunk, hipMemcpyHostToDevice, c.stream), "staged H2D");
        HGM_HIP(hipEventRecord(c.stage_ev[i & 1], c.stream), "hipEventRecord");
    }
    HGM_HIP(hipStreamSynchronize(c.stream), "staged H2D");
    return HGM_OK;
}

// Device -> host, the inverse. The first fetches queue behind whatever the
// stream already holds. Assumes the lock.
int staged_d2h_locked(char *h, const char *d, size_t bytes, uint64_t id, uint64_t mask, size_t amp) {
    Ctx &c = ctx();
    int rc = stage_locked();
    if (rc) return rc;
    const size_t chunk = stage_chunk(bytes);
    const size_t n_chunks = bytes / chunk;
    auto fetch = [&](size_t i) -> int {
        HGM_HIP(hipMemcpyAsync(c.h_stage[i & 1], d + i * chunk, chunk, hipMemcpyDeviceToHost, c.stream), "staged D2H");
        HGM_HIP(hipEventRecord(c.stage_ev[i & 1], c.stream), "hipEventRecord");
        return HGM_OK;
    };
    for (size_t i = 0; i < std::min<size_t>(2, n_chunks); ++i)
        if ((rc = fetch(i))) return rc;
    for (size_t i = 0; i < n_chunks; ++i) {
        const char *buf = c.h_stage[i & 1];
        HGM_HIP(hipEventSynchronize(c.stage_ev[i & 1]), "staged D2H");
        chunk_host_copy(id, mask, amp, i * chunk, chunk, [&](size_t dst, size_t n, size_t at) { std::memcpy(h + dst, buf + at, n); });
        if (i + 2 < n_chunks && (rc = fetch(i + 2))) return rc;
    }
    return HGM_OK;
}

// Pinned (page-locked) host memory goes to the DMA engines directly; pageable
// memory is better staged: HIP's own path for it moves 28-30 GB/s to the device
// and 18 GB/s back on the MI355X box, the two pinned chunks filled by 8
// threads about 55 GB/s both ways.
bool host_is_pinned(const void *p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

}  // namespace

extern "C" {

int hgm_state_gather(hgm_state *s, const void *host, int host_qubits, uint64_t region_id, uint64_t region_mask) {
    int rc = region_check(s, host, host_qubits, region_id, region_mask);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(ctx().mu);
    return staged_h2d_locked(static_cast<char *>(s->d), static_cast<const char *>(host), s->bytes, region_id,
                             region_mask, amp_bytes(s->dtype));
}

int hgm_state_scatter(const hgm_state *s, void *host, int host_qubits, uint64_t region_id, uint64_t region_mask) {
    int rc = region_check(s, host, host_qubits, region_id, region_mask);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(ctx().mu);
    return staged_d2h_locked(static_cast<char *>(host), static_cast<const char *>(s->d), s->bytes, region_id,
                             region_mask, amp_bytes(s->dtype));
}

int hgm_state_apply(hgm_state *s, const hgm_op *ops, int64_t n_ops, const hgm_tuning *tuning, int iterations,
                    hgm_stats *stats) {
    if (!s) { t_err = "null state"; return HGM_E_ARG; }
    auto t0 = std::chrono::steady_clock::now();
    int rc = plan_and_run(s, ops, n_ops, tuning, iterations, stats);
    if (stats) stats->total_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int hgm_state_apply_noalloc(hgm_state *s, const hgm_op *ops, int64_t n_ops, const hgm_tuning *tuning,
                            int iterations, hgm_stats *stats) {
    if (!s) { t_err = "null state"; return HGM_E_ARG; }
    auto t0 = std::chrono::steady_clock::now();
    hgm_tuning t;
    if (tuning) t = *tuning; else hgm_default_tuning(&t);
    if (t.kernel == HGM_KERNEL_JIT) {
        t_err = "hgm_state_apply_noalloc runs the register or LDS kernels: plan-compiled kernels load a code object";
        return HGM_E_TUNING;
    }
    std::shared_ptr<const Plan> p;
    double plan_ms = 0;
    int rc = cached_plan(s->qubits, s->dtype, ops, n_ops, iterations, t, &p, &plan_ms);
    if (rc) return rc;
    const size_t need = plan_table_bytes(*p);
    if (need > kScratchBytes) {
        t_err = "hgm_state_apply_noalloc: the plan's tables take " + std::to_string(need) + " bytes, above the " +
                std::to_string(kScratchBytes) + "-byte workspace (fewer passes or gates per call)";
        return HGM_E_ARG;
    }
    Ctx &c = ctx();
    {
        std::lock_guard<std::mutex> g(c.mu);
        // a state exists, so the context and the workspace address were set up by its creation
        if (!c.ready || c.device < 0 || c.device >= kMaxDevices || !g_scratch_ptr[c.device]) {
            t_err = "hgm_state_apply_noalloc: no device context (hgm_init or hgm_state_create first)";
            return HGM_E_STATE;
        }
        rc = run_plan_locked(s, *p, stats, plan_ms, g_scratch_ptr[c.device]);
    }
    if (stats) stats->total_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int hgm_state_norm2(const hgm_state *s, double *out) {
    if (!s || !out) { t_err = "null argument"; return HGM_E_ARG; }
    std::lock_guard<std::mutex> g(ctx().mu);
    double z, o;
    int rc = prob_halves_locked(s, -1, &z, &o);
    if (!rc) *out = z + o;
    return rc;
}

int hgm_state_prob_one(const hgm_state *s, int qubit, double *out) {
    if (!s || !out || qubit < 0 || qubit >= s->qubits) { t_err = "bad qubit"; return HGM_E_ARG; }
    std::lock_guard<std::mutex> g(ctx().mu);
    double z, o;
    int rc = prob_halves_locked(s, s->qubits - 1 - qubit, &z, &o);
    if (!rc) *out = o;
    return rc;
}

int hgm_state_prob_halves(const hgm_state *s, int qubit, double *zero, double *one) {
    if (!s || !zero || !one || qubit < 0 || qubit >= s->qubits) { t_err = "bad qubit"; return HGM_E_ARG; }
    std::lock_guard<std::mutex> g(ctx().mu);
    return prob_halves_locked(s, s->qubits - 1 - qubit, zero, one);
}

int hgm_state_collapse_reset(hgm_state *s, int qubit, int outcome, double p) {
    if (!s || qubit < 0 || qubit >= s->qubits || (outcome != 0 && outcome != 1) || !(p > 0)) {
        t_err = "bad collapse arguments";
        return HGM_E_ARG;
    }
    Ctx &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    HGM_HIP(launch_collapse(s->dtype, s->d, uint64_t(1) << s->qubits, s->qubits - 1 - qubit, outcome, std::sqrt(p),
                            c.stream), "collapse");
    HGM_HIP(hipStreamSynchronize(c.stream), "collapse");
    return HGM_OK;
}

int hgm_state_argmax(const hgm_state *s, uint64_t *index, double *prob) {
    if (!s || !index) { t_err = "null argument"; return HGM_E_ARG; }
    Ctx &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    const uint64_t n = uint64_t(1) << s->qubits;
    const int blocks = blocks_for(n);
    HGM_HIP(launch_argmax(s->dtype, s->d, n, c.partial, c.pidx, blocks, c.stream), "argmax");
    HGM_HIP(hipMemcpyAsync(c.h_partial, c.partial, blocks * sizeof(double), hipMemcpyDeviceToHost, c.stream), "argmax");
    HGM_HIP(hipMemcpyAsync(c.h_pidx, c.pidx, blocks * sizeof(uint64_t), hipMemcpyDeviceToHost, c.stream), "argmax");
    HGM_HIP(hipStreamSynchronize(c.stream), "argmax");
    double best = -1;
    uint64_t bi = 0;
    for (int b = 0; b < blocks; ++b)
        if (c.h_partial[b] > best || (c.h_partial[b] == best && c.h_pidx[b] < bi)) { best = c.h_partial[b]; bi = c.h_pidx[b]; }
    *index = bi;
    if (prob) *prob = best;
    return HGM_OK;
}

int hgm_state_compare(const hgm_state *a, const hgm_state *b, double *fidelity, double *max_abs_err) {
    if (!a || !b) { t_err = "null argument"; return HGM_E_ARG; }
    if (a->qubits != b->qubits) { t_err = "states of different qubit counts"; return HGM_E_ARG; }
    Ctx &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    const uint64_t n = uint64_t(1) << a->qubits;
    const int blocks = blocks_for(n);
    HGM_HIP(launch_compare(a->dtype, a->d, b->dtype, b->d, n, c.partial, blocks, c.stream), "compare");
    HGM_HIP(hipMemcpyAsync(c.h_partial, c.partial, 5 * blocks * sizeof(double), hipMemcpyDeviceToHost, c.stream),
            "compare");
    HGM_HIP(hipStreamSynchronize(c.stream), "compare");
    double re = 0, im = 0, na = 0, nb = 0, mx = 0;
    for (int k = 0; k < blocks; ++k) {
        const double *q = c.h_partial + 5 * k;
        re += q[0]; im += q[1]; na += q[2]; nb += q[3];
        mx = std::max(mx, q[4]);
    }
    if (fidelity) *fidelity = (na > 0 && nb > 0) ? (re * re + im * im) / (na * nb) : 0.0;
    if (max_abs_err) *max_abs_err = mx;
    return HGM_OK;
}

int hgm_plan_build(int qubits, int dtype, const hgm_op *ops, int64_t n_ops, int iterations, const hgm_tuning *tuning,
                   hgm_plan **out) {
    if (!out) { t_err = "null output"; return HGM_E_ARG; }
    hgm_tuning t;
    if (tuning) t = *tuning; else hgm_default_tuning(&t);
    hgm_plan *p = new hgm_plan();
    int rc = build_plan(qubits, dtype, ops, n_ops, iterations, t, &p->plan, &t_err);
    if (rc) { delete p; return rc; }
    *out = p;
    return HGM_OK;
}

void hgm_plan_destroy(hgm_plan *p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> g(ctx().mu);
        p->drop_graph();
    }
    delete p;
}

namespace {
enum class Rewrite { fold, fuse, pivot, rotate };

int rewrite_ops(Rewrite how, int qubits, const hgm_op *ops, int64_t n_ops, hgm_op *out, int64_t cap, int64_t *n_out) {
    if (!n_out || n_ops < 0 || (n_ops > 0 && !ops) || qubits < 1 || qubits > kMaxQubits) {
        t_err = "bad argument";
        return HGM_E_ARG;
    }
    for (int64_t i = 0; i < n_ops; ++i)
        if (ops[i].target_bit < 0 || ops[i].target_bit >= qubits || (ops[i].control_mask >> ops[i].target_bit) & 1) {
            t_err = "target bit outside the state or among the controls";
            return HGM_E_ARG;
        }
    std::vector<hgm_op> v;
    if (how == Rewrite::fuse) fuse_gates(qubits, ops, n_ops, &v);
    else if (how == Rewrite::pivot) pivot_rows(qubits, ops, n_ops, &v);
    else if (how == Rewrite::rotate) rotate_rows(qubits, ops, n_ops, &v);
    else fold_phases(qubits, ops, n_ops, &v);
    *n_out = (int64_t)v.size();
    if (out) {
        if (cap < (int64_t)v.size()) { t_err = "output buffer too small"; return HGM_E_ARG; }
        std::memcpy(out, v.data(), v.size() * sizeof(hgm_op));
    }
    return HGM_OK;
}
}  // namespace

int hgm_fold_phases(int qubits, const hgm_op *ops, int64_t n_ops, hgm_op *out, int64_t cap, int64_t *n_out) {
    return rewrite_ops(Rewrite::fold, qubits, ops, n_ops, out, cap, n_out);
}

int hgm_fuse_gates(int qubits, const hgm_op *ops, int64_t n_ops, hgm_op *out, int64_t cap, int64_t *n_out) {
    return rewrite_ops(Rewrite::fuse, qubits, ops, n_ops, out, cap, n_out);
}

int hgm_pivot_rows(int qubits, const hgm_op *ops, int64_t n_ops, hgm_op *out, int64_t cap, int64_t *n_out) {
    return rewrite_ops(Rewrite::pivot, qubits, ops, n_ops, out, cap, n_out);
}

int hgm_rotate_rows(int qubits, const hgm_op *ops, int64_t n_ops, hgm_op *out, int64_t cap, int64_t *n_out) {
    return rewrite_ops(Rewrite::rotate, qubits, ops, n_ops, out, cap, n_out);
}

int64_t hgm_plan_num_passes(const hgm_plan *p) { return p ? (int64_t)p->plan.passes.size() : -1; }

int hgm_plan_pass(const hgm_plan *p, int64_t i, uint64_t *local_mask, int64_t *n_gates, int64_t *gate_index) {
    if (!p || i < 0 || i >= (int64_t)p->plan.passes.size()) { t_err = "pass index out of range"; return HGM_E_ARG; }
    const Pass &ps = p->plan.passes[i];
    if (local_mask) *local_mask = ps.local_mask;
    if (n_gates) *n_gates = ps.op_count;
    if (gate_index)
        for (int64_t k = 0; k < ps.op_count; ++k) gate_index[k] = p->plan.src_index[ps.op_begin + k];
    return HGM_OK;
}

int hgm_plan_shape(const hgm_plan *p, int32_t *region_bits, int32_t *block_size, int32_t *repeat_count,
                   int32_t *coalesced_bits) {
    if (!p) { t_err = "null plan"; return HGM_E_ARG; }
    if (region_bits) *region_bits = p->plan.shape.r;
    if (block_size) *block_size = p->plan.shape.B;
    if (repeat_count) *repeat_count = p->plan.shape.R;
    if (coalesced_bits) *coalesced_bits = p->plan.shape.c;
    return HGM_OK;
}

int hgm_plan_kernel(const hgm_plan *p) { return p ? p->plan.kernel : -1; }

int hgm_plan_jit_source(const hgm_plan *p, char *out, size_t cap, size_t *len) {
    if (!p || !len) { t_err = "bad arguments"; return HGM_E_ARG; }
    const std::string src = p->plan.jit ? p->plan.jit->source : std::string();
    *len = src.size();
    if (out && cap > src.size()) std::memcpy(out, src.c_str(), src.size() + 1);
    return HGM_OK;
}

int hgm_plan_serving_source(const hgm_plan *p, int64_t pass, char *out, size_t cap, size_t *len) {
    if (!p || !len || pass < 0 || pass >= (int64_t)p->plan.passes.size()) { t_err = "bad arguments"; return HGM_E_ARG; }
    if ((p->plan.kernel != HGM_KERNEL_REG && p->plan.kernel != HGM_KERNEL_JIT) || p->plan.passes[pass].op_count == 0) {
        t_err = "serving kernels are generated for the non-empty passes of register-program plans";
        return HGM_E_ARG;
    }
    size_t lds = 0;
    const std::string src = serving_kernel_source(p->plan, p->plan.passes[pass], "hgm_serve", &lds);
    *len = src.size();
    if (out && cap > src.size()) std::memcpy(out, src.c_str(), src.size() + 1);
    return HGM_OK;
}

double hgm_plan_compile_seconds(const hgm_plan *p) { return (p && p->plan.jit) ? p->plan.jit->compile_s : 0.0; }

int hgm_plan_program(const hgm_plan *p, int64_t i, void *rops, int64_t *n_rops, int32_t *first_map,
                     int32_t *last_map, int8_t *local_pos) {
    if (!p || i < 0 || i >= (int64_t)p->plan.passes.size() || !n_rops) { t_err = "bad arguments"; return HGM_E_ARG; }
    const Pass &ps = p->plan.passes[i];
    *n_rops = p->plan.kernel == HGM_KERNEL_REG ? ps.rop_count : 0;
    if (rops && *n_rops)
        std::memcpy(rops, p->plan.rops.data() + ps.rop_begin, (size_t)ps.rop_count * sizeof(RegOp));
    if (first_map) *first_map = ps.mem >= 0 ? p->plan.mems[ps.mem].first_map : -1;
    if (last_map) *last_map = ps.mem >= 0 ? p->plan.mems[ps.mem].last_map : -1;
    if (local_pos) std::memcpy(local_pos, ps.local_pos, sizeof(ps.local_pos));
    return HGM_OK;
}

int hgm_plan_maps(const hgm_plan *p, void *maps, int64_t *n_maps) {
    if (!p || !n_maps) { t_err = "bad arguments"; return HGM_E_ARG; }
    *n_maps = (int64_t)p->plan.maps.size();
    if (maps && *n_maps) std::memcpy(maps, p->plan.maps.data(), p->plan.maps.size() * sizeof(PhaseMap));
    return HGM_OK;
}

int hgm_state_run_plan(hgm_state *s, const hgm_plan *p, hgm_stats *stats) {
    if (!s || !p) { t_err = "null argument"; return HGM_E_ARG; }
    auto t0 = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> g(ctx().mu);
    int rc = init_locked(-1);
    if (!rc) rc = p->graph_on ? run_plan_graph_locked(s, p, stats) : run_plan_locked(s, p->plan, stats, p->plan.plan_ms);
    if (stats) stats->total_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int hgm_batch_create(hgm_state *const *states, const hgm_plan *const *plans, int n, hgm_batch **out) {
    if (!out) { t_err = "null output"; return HGM_E_ARG; }
    *out = nullptr;
    if (!states || !plans || n < 1) { t_err = "bad arguments"; return HGM_E_ARG; }
    std::vector<const void *> seen;
    std::vector<uint64_t> mods;
    for (int i = 0; i < n; ++i) {
        if (!states[i] || !plans[i]) { t_err = "null state or plan"; return HGM_E_ARG; }
        const Plan &p = plans[i]->plan;
        if (p.kernel == HGM_KERNEL_JIT && !p.jit) { t_err = "plan-compiled plan without code"; return HGM_E_ARG; }
        if (states[i]->qubits != p.qubits || states[i]->dtype != p.dtype) {
            t_err = "plan was built for another qubit count or dtype";
            return HGM_E_STATE;
        }
        if (std::find(seen.begin(), seen.end(), states[i]->d) != seen.end()) {
            t_err = "a state appears twice in the batch";
            return HGM_E_ARG;
        }
        seen.push_back(states[i]->d);
        if (p.kernel == HGM_KERNEL_JIT && std::find(mods.begin(), mods.end(), p.jit->hash) == mods.end())
            mods.push_back(p.jit->hash);
    }
    if (mods.size() > kMaxJitModules) {
        t_err = "a batch graph holds at most " + std::to_string(kMaxJitModules) + " distinct compiled modules";
        return HGM_E_ARG;
    }
    auto b = std::make_unique<hgm_batch>();
    b->states.assign(states, states + n);
    b->plans.assign(plans, plans + n);
    // states by bytes moved, largest first, each to the branch with the fewest so far
    b->n_branches = std::min(n, kBatchBranches);
    std::vector<int> order(n);
    std::vector<double> bytes(n), load(b->n_branches, 0.0);
    for (int i = 0; i < n; ++i) {
        const Plan &p = plans[i]->plan;
        int64_t busy = 0;
        for (const Pass &ps : p.passes) busy += ps.op_count > 0;
        bytes[i] = (double)busy * (double)(uint64_t(1) << p.qubits) * (double)amp_bytes(p.dtype);
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return bytes[a] > bytes[c]; });
    b->branch.assign(n, 0);
    for (int i : order) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->branch[i] = k;
        load[k] += bytes[i];
    }
    std::lock_guard<std::mutex> g(ctx().mu);
    int rc = init_locked(-1);
    if (rc) return rc;
    auto release = [&] {
        b->drop_graph();
        if (b->tables) { hipFree(b->tables); b->tables = nullptr; }
        for (hipStream_t &st : b->streams)
            if (st) { hipStreamDestroy(st); st = nullptr; }
        for (hipEvent_t &ev : b->join)
            if (ev) { hipEventDestroy(ev); ev = nullptr; }
        if (b->fork) { hipEventDestroy(b->fork); b->fork = nullptr; }
    };
    hipError_t e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (int k = 0; k < b->n_branches && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&b->streams[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming);
    }
    // one device copy of the tables of each distinct register or LDS plan
    b->table_off.assign(n, -1);
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        const Plan &p = plans[i]->plan;
        if (p.kernel == HGM_KERNEL_JIT) continue;
        int j = 0;
        while (j < i && plans[j] != plans[i]) ++j;
        if (j < i) { b->table_off[i] = b->table_off[j]; continue; }
        b->table_off[i] = (int64_t)total;
        total += align256(plan_table_bytes(p));
        if (e == hipSuccess) e = prepare_pass_kernels(p.dtype);
    }
    if (e == hipSuccess && total) e = hipMalloc(&b->tables, total);
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        const Plan &p = plans[i]->plan;
        if (b->table_off[i] < 0 || (i > 0 && std::find(plans, plans + i, plans[i]) != plans + i)) continue;
        const PlanTables t = batch_tables(b.get(), (size_t)i);
        if (p.kernel == HGM_KERNEL_REG) {
            if (!p.rops.empty())
                e = hipMemcpy(const_cast<void *>(t.rops), p.rops.data(), p.rops.size() * sizeof(RegOp),
                              hipMemcpyHostToDevice);
            if (e == hipSuccess && !p.maps.empty())
                e = hipMemcpy(const_cast<void *>(t.maps), p.maps.data(), p.maps.size() * sizeof(PhaseMap),
                              hipMemcpyHostToDevice);
            if (e == hipSuccess && !p.mems.empty())
                e = hipMemcpy(const_cast<void *>(t.mems), p.mems.data(), p.mems.size() * sizeof(PassMem),
                              hipMemcpyHostToDevice);
        } else if (!p.ops.empty()) {
            e = hipMemcpy(const_cast<DevOp *>(t.dops), p.ops.data(), p.ops.size() * sizeof(DevOp),
                          hipMemcpyHostToDevice);
        }
    }
    if (e != hipSuccess) {
        rc = hip_fail(e, "batch streams and tables");
        release();
        return rc;
    }
    std::vector<JitModule *> jms;
    bool stale = true;
    rc = batch_modules_locked(b.get(), &jms, &stale);
    if (!rc) rc = capture_batch_locked(b.get(), jms);
    if (rc) {
        release();
        return rc;
    }
    *out = b.release();
    return HGM_OK;
}

int hgm_batch_run(hgm_batch *b, hgm_stats *stats) {
    if (!b) { t_err = "null batch"; return HGM_E_ARG; }
    auto t0 = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> g(ctx().mu);
    int rc = init_locked(-1);
    if (rc) return rc;
    Ctx &c = ctx();
    std::vector<JitModule *> jms;
    bool stale = false;
    if ((rc = batch_modules_locked(b, &jms, &stale))) return rc;
    if (stale && (rc = capture_batch_locked(b, jms))) return rc;
    HGM_HIP(hipEventRecord(c.e0, c.stream), "hipEventRecord");
    HGM_HIP(hipGraphLaunch(b->exec, c.stream), "hipGraphLaunch");
    HGM_HIP(hipEventRecord(c.e1, c.stream), "hipEventRecord");
    HGM_HIP(hipEventSynchronize(c.e1), "batch execution");
    const float ms = elapsed(c.e0, c.e1);
    for (size_t i = 0; i < b->plans.size(); ++i) add_stats(stats, b->plans[i]->plan, i == 0 ? ms : 0.0f, 0.0);
    if (stats) stats->total_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return HGM_OK;
}

void hgm_batch_destroy(hgm_batch *b) {
    if (!b) return;
    std::lock_guard<std::mutex> g(ctx().mu);
    if (ctx().stream) hipStreamSynchronize(ctx().stream);
    b->drop_graph();
    if (b->tables) hipFree(b->tables);
    for (hipStream_t st : b->streams)
        if (st) hipStreamDestroy(st);
    for (hipEvent_t ev : b->join)
        if (ev) hipEventDestroy(ev);
    if (b->fork) hipEventDestroy(b->fork);
    delete b;
}

int hgm_plan_set_graph(hgm_plan *p, int enable) {
    if (!p) { t_err = "null plan"; return HGM_E_ARG; }
    if (enable && p->plan.kernel != HGM_KERNEL_JIT) {
        t_err = "graph replay is for plan-compiled plans (HGM_KERNEL_JIT)";
        return HGM_E_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> g(ctx().mu);
    if (!enable) p->drop_graph();
    p->graph_on = enable != 0;
    return HGM_OK;
}

int hgm_execute(void *host_state, int qubits, int dtype, const hgm_op *ops, int64_t n_ops, const hgm_tuning *tuning,
                int iterations, hgm_stats *stats) {
    if (!host_state) { t_err = "null host state"; return HGM_E_ARG; }
    auto t0 = std::chrono::steady_clock::now();
    hgm_tuning t;
    if (tuning) t = *tuning; else hgm_default_tuning(&t);
    std::shared_ptr<const Plan> plan;
    double plan_ms = 0;
    int rc = cached_plan(qubits, dtype, ops, n_ops, iterations, t, &plan, &plan_ms);
    if (rc) return rc;
    const Plan &p = *plan;
    Ctx &c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    rc = init_locked(-1);
    if (rc) return rc;
    const size_t bytes = (size_t(1) << qubits) * amp_bytes(dtype);
    if (bytes > c.cache_cap) {
        if (c.cache) { hipStreamSynchronize(c.stream); hipFree(c.cache); c.cache = nullptr; c.cache_cap = 0; }
        if (hipMalloc(&c.cache, bytes) != hipSuccess) {
            hipGetLastError();
            t_err = "hipMalloc of a " + std::to_string(qubits) + "-qubit state failed";
            return HGM_E_NOMEM;
        }
        c.cache_cap = bytes;
    }
    hgm_state s;
    s.d = c.cache;
    s.qubits = qubits;
    s.dtype = dtype;
    s.owned = false;
    s.bytes = bytes;
    if (host_is_pinned(host_state)) {   // straight to the DMA engines, timed with events
        HGM_HIP(hipEventRecord(c.e2, c.stream), "hipEventRecord");
        HGM_HIP(hipMemcpyAsync(s.d, host_state, bytes, hipMemcpyHostToDevice, c.stream), "H2D state");
        HGM_HIP(hipEventRecord(c.e3, c.stream), "hipEventRecord");
        rc = run_plan_locked(&s, p, stats, plan_ms);
        if (rc) return rc;
        HGM_HIP(hipEventRecord(c.e4, c.stream), "hipEventRecord");
        HGM_HIP(hipMemcpyAsync(host_state, s.d, bytes, hipMemcpyDeviceToHost, c.stream), "D2H state");
        HGM_HIP(hipEventRecord(c.e5, c.stream), "hipEventRecord");
        HGM_HIP(hipEventSynchronize(c.e5), "D2H state");
        if (stats) {
            stats->h2d_ms += elapsed(c.e2, c.e3);
            stats->d2h_ms += elapsed(c.e4, c.e5);
        }
    } else {   // pageable (the reference's calloc'd state): staged, timed on the host clock
        const uint64_t all = (uint64_t(1) << qubits) - 1;
        const size_t amp = amp_bytes(dtype);
        auto w0 = std::chrono::steady_clock::now();
        if ((rc = staged_h2d_locked(static_cast<char *>(s.d), static_cast<const char *>(host_state), bytes, 0, all, amp)))
            return rc;
        auto w1 = std::chrono::steady_clock::now();
        rc = run_plan_locked(&s, p, stats, plan_ms);
        if (rc) return rc;
        HGM_HIP(hipStreamSynchronize(c.stream), "passes");
        auto w2 = std::chrono::steady_clock::now();
        if ((rc = staged_d2h_locked(static_cast<char *>(host_state), static_cast<const char *>(s.d), bytes, 0, all, amp)))
            return rc;
        if (stats) {
            stats->h2d_ms += std::chrono::duration<double, std::milli>(w1 - w0).count();
            stats->d2h_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w2).count();
        }
    }
    if (stats)
        stats->total_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return HGM_OK;
}

}  // extern "C"
