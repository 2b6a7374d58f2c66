// Serving engine: continuous batching of circuit requests on one GPU, and the
// tile-shape autotuner (the hgm_server_* / hgm_autotune part of hgm_gpu.h).
//
// Requests are planned on host threads as they arrive, admitted in arrival
// order while their HBM footprint (state + register programs + readout
// partials) fits the budget and a slot is free, and then advance together:
// every step is one batched register-kernel launch per workgroup shape that
// runs the next pass of every admitted state. States that fit one tile (up to
// 13 qubits) run their whole circuit in one pass and live on a lane (stream)
// of their own, so that long single-tile programs never hold back the passes
// of large states. A state's last pass also reduces its readout (norm and
// argmax, FP64); the step's results come back in one small copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "hgm_dgm.hpp"
#include "hgm_gpu.h"
#include "jit.hpp"
#include "kernels.hpp"
#include "planner.hpp"
#include "runtime.hpp"

namespace hgm {
namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// Message of a failed HIP call; also clears the error, which the next
// launch check (hipGetLastError) would otherwise report as its own.
std::string hip_msg(hipError_t e, const char *what) {
    (void)hipGetLastError();
    return std::string(what) + ": " + hipGetErrorString(e);
}

// First-fit allocator over one device (or pinned host) allocation; freed
// ranges merge with their neighbours.
class Arena {
  public:
    int init(size_t bytes, bool pinned_host, std::string *err) {
        hipError_t e = pinned_host ? hipHostMalloc(&base_, bytes, hipHostMallocDefault) : hipMalloc(&base_, bytes);
        if (e != hipSuccess) {
            *err = hip_msg(e, pinned_host ? "hipHostMalloc(program staging)" : "hipMalloc(serving arena)");
            base_ = nullptr;
            return HGM_E_HIP;
        }
        pinned_ = pinned_host;
        size_ = bytes;
        free_[0] = bytes;
        return HGM_OK;
    }
    void destroy() {
        if (base_) {
            if (pinned_) hipHostFree(base_);
            else hipFree(base_);
        }
        base_ = nullptr;
        free_.clear();
    }
    bool alloc(size_t bytes, size_t *off) {
        bytes = align_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += bytes;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    char *base() const { return base_; }
    size_t size() const { return size_; }
    size_t used() const { return used_; }

  private:
    char *base_ = nullptr;
    bool pinned_ = false;
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;
};

struct Req {
    int64_t id = 0;
    int qubits = 0;
    int iterations = 1;
    uint64_t basis = 0;
    void *state_out = nullptr;
    std::vector<hgm_op> ops;
    double t_submit = 0, t_admit = -1, t_done = -1;
    std::atomic<int> planned{0};   // 0 queued, 1 planned, 2 failed
    int status = HGM_OK;
    std::string err;
    Plan plan;
    std::vector<int> passes;       // non-empty passes in order
    std::vector<int> kernel;       // per entry of passes: compiled serving kernel, or -1 (interpreter)
    size_t next = 0;
    int lane = 0;
    int64_t n_tiles = 1;
    size_t state_bytes = 0, prog_bytes = 0, part_bytes = 0, total_bytes = 0, off = 0, stage_off = 0;
    char *d = nullptr;             // device block: state | programs | maps | pass tables | partials
    ReadoutPartial result{};
};

struct Step {
    std::vector<Req *> finishing;
    int slot = 0;
};

constexpr int kLanes = 2;          // 0: single-tile states (whole circuit per pass), 1: multi-tile states
constexpr int kDepth = 3;          // steps in flight per lane

struct Lane {
    hipStream_t stream = nullptr;
    std::vector<Req *> active;
    std::deque<Step> inflight;
    hipEvent_t ev[kDepth] = {};
    char *d_ring = nullptr;        // kDepth slots: items | tasks | results
    char *h_ring = nullptr;        // pinned mirror
    int next_slot = 0;
};

double now_ms(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Gate lists are fused and row-scaled before planning (same unitary, fewer
// and cheaper gates: Grover's X-H-X runs become one real gate per qubit, and
// each dense update costs 8 multiply-adds per pair, 4 when real).
constexpr int kServingPlanner = HGM_PLAN_GREEDY | HGM_PLAN_FUSE | HGM_PLAN_PIVOT;

// Planner of the qubit counts that have compiled pass kernels: row scaling
// picks each row's pivot by magnitude, which makes a pass's structure (and so
// its kernel) depend on the matrix entries; without it, circuits of one shape
// (a random-layer circuit under any seed, Grover for any marked value) share
// their pass kernels.
constexpr int kCompiledPlanner = HGM_PLAN_GREEDY | HGM_PLAN_FUSE;

// Dynamic LDS a compiled serving kernel may ask for: the CU's 160 KiB less
// the kernel's static readout arrays.
constexpr size_t kServeMaxLds = 160 * 1024 - 1024;

// A compiled serving pass kernel (serving_kernel_source), loaded on the
// server's device.
struct ServeKernel {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    size_t lds = 0;
};

// Amplitudes per thread the batched register kernel is instantiated for.
bool serving_amps_ok(int A) { return A == 2 || A == 4 || A == 8 || A == 16 || A == 32; }

// Tile shape the serving engine uses for a qubit count when no autotuning is
// asked for: 16 amplitudes per thread; states up to 13 qubits are one tile;
// larger ones use tiles of 2^12 amplitudes. The always-local low bits make
// 128-byte runs: 3 bits of 16-byte FP64 amplitudes, 4 bits of 8-byte FP32 ones.
// FP32 states of 21-22 qubits use tiles of 2^13 (512 threads) and of 23-24
// qubits tiles of 2^14 (1024 threads): their n - 4 non-coalesced bits then
// fit two passes' windows of r - 4 bits, so a layer of gates on every qubit
// takes 2 passes instead of 3-4 (Grover with compiled pass kernels: 28-31%
// faster at 21-24 qubits; at 25-26 qubits no tile the register kernels hold
// fits two windows with 128-byte runs).
hgm_tuning serving_default_tuning(int n, int dtype) {
    const int c = dtype == HGM_C64 ? 4 : 3;
    hgm_tuning t;
    hgm_default_tuning(&t);
    t.ops_block = 1 << 30;
    t.kernel = HGM_KERNEL_REG;
    t.planner = kServingPlanner;
    if (n <= 13) {
        const int64_t pairs = int64_t(1) << (n - 1);
        t.repeat_count = (int)std::min<int64_t>(8, pairs);
        t.block_size = (int)(pairs / t.repeat_count);
        t.gpu_region_bits = n;
        t.gpu_coalesced_bits = std::min(c, n);
    } else {
        t.block_size = 256;
        t.repeat_count = 8;
        t.gpu_region_bits = 12;
        t.gpu_coalesced_bits = c;
        if (dtype == HGM_C64 && n - c > 2 * (12 - c))
            for (int r : {13, 14})
                if (n - c <= 2 * (r - c)) {
                    t.block_size = int(1) << (r - 4);
                    t.gpu_region_bits = r;
                    break;
                }
    }
    return t;
}

// FP32 states of 25-26 qubits can also run on tiles of 2^15 amplitudes (1024
// threads x 32 amplitudes, a 140 KiB LDS remap tile in halves), which only the
// compiled pass kernels hold. Their n - 4 non-coalesced bits then fit two
// windows of 11 bits, so Grover takes 2 passes per iteration instead of 4; but
// one such workgroup fills a CU, so a pass streams slower (Grover at 26
// qubits: 0.33 against 0.19 ms per pass on 2^12 tiles, kLargeTilePassCost).
// A request takes this shape only when every pass of its plan has a compiled
// kernel and the plan has fewer than 1/kLargeTilePassCost of the passes.
constexpr double kLargeTilePassCost = 1.7;

bool serving_large_tile(int n, int dtype, hgm_tuning *t) {
    const int c = 4, r = 15;
    if (dtype != HGM_C64 || n - c <= 2 * (14 - c) || n - c > 2 * (r - c)) return false;
    hgm_default_tuning(t);
    t->ops_block = 1 << 30;
    t->kernel = HGM_KERNEL_JIT;   // shape rules of the compiled kernels; planned program-only
    t->planner = kCompiledPlanner;
    t->block_size = 1024;
    t->repeat_count = 16;
    t->gpu_region_bits = r;
    t->gpu_coalesced_bits = c;
    return true;
}

// The autotuner's timing circuit for `n` qubits: the serving mix's expected
// work at that qubit count, one request of each kind (the mix draws the kinds
// uniformly): a Hadamard layer, Grover with 32 iterations (the mean of 1..64),
// the Draper adder on the low 2*floor((n-1)/2)+1 qubits, and random layers of
// depth 10, as one gate list.
std::vector<hgm_op> mix_timing_circuit(int n) {
    DGM d;
    d.exit_on_error = false;
    d.qubits = n;
    auto add = [&](const std::vector<std::string> &steps, int iterations, std::vector<hgm_op> &ops) {
        d.setFunction(steps, iterations);
        std::vector<hgm_op> o = d.ops();
        ops.insert(ops.end(), o.begin(), o.end());
    };
    std::vector<hgm_op> ops;
    add({Hadamard(n, 0, n)}, 1, ops);
    add(GroverIteration(n, (uint64_t(0x5a5a5a5a5a5a5a5aull) & ((uint64_t(1) << n) - 1))), 32, ops);
    const int w = (n - 1) / 2;
    if (w >= 1) {
        DGM da;
        da.exit_on_error = false;
        da.qubits = 2 * w + 1;
        da.setFunction(DraperAdder(w, da.gates));
        std::vector<hgm_op> o = da.ops();
        ops.insert(ops.end(), o.begin(), o.end());
    }
    add(RandomLayers(n, 10, 2026, d.gates), 1, ops);
    return ops;
}

}  // namespace
}  // namespace hgm

using namespace hgm;

struct hgm_server {
    hgm_server_config cfg{};
    int device = 0;
    std::chrono::steady_clock::time_point t0;
    Arena arena;                              // device: states, programs, partials
    Arena staging;                            // pinned host: programs on their way to the device
    Lane lanes[kLanes];
    size_t slot_bytes = 0, tasks_off = 0, results_off = 0;
    std::vector<hgm_tuning> tune;             // per qubit count
    std::vector<hgm_tuning> large;            // per qubit count: the 2^15-tile alternative (block_size 0: none)

    std::mutex mu;                            // submissions, results, plan queue
    std::condition_variable cv_work, cv_done, cv_plan;
    std::deque<Req *> submitted;              // not yet seen by the worker
    std::deque<Req *> plan_q;
    std::deque<hgm_result> done;
    int64_t next_id = 0, n_submitted = 0, n_done = 0;
    bool stop = false;
    std::string fatal;                        // first device error

    // worker-only state
    std::deque<Req *> pending;                // arrival order, awaiting admission
    int active = 0;

    std::thread worker;
    std::vector<std::thread> planners;

    // compiled pass kernels (hgm_server_precompile), by serving_pass_key
    std::mutex kmu;
    std::unordered_map<std::string, int> kernel_of;
    std::deque<ServeKernel> kernels;          // stable addresses while precompile appends
};

namespace {

void finish(hgm_server *s, Req *r) {
    hgm_result res{};
    res.id = r->id;
    res.status = r->status;
    res.passes = (int32_t)r->passes.size();
    res.argmax = r->result.idx;
    res.p_max = r->result.pmax;
    res.norm2 = r->result.sum;
    res.t_submit_ms = r->t_submit;
    res.t_admit_ms = r->t_admit;
    res.t_done_ms = r->t_done;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->done.push_back(res);
        ++s->n_done;
    }
    s->cv_done.notify_all();
    delete r;
}

void fail(hgm_server *s, Req *r, int status, const std::string &msg) {
    r->status = status;
    r->err = msg;
    r->t_done = now_ms(s->t0);
    finish(s, r);
}

// Replans a request on the 2^15-tile alternative when every pass of that plan
// has a compiled kernel and the plan saves enough passes (kLargeTilePassCost);
// otherwise leaves the request as it is.
void try_large_tile(hgm_server *s, Req *r, const hgm_tuning &lt) {
    {
        std::lock_guard<std::mutex> g(s->kmu);
        if (s->kernel_of.empty()) return;
    }
    Plan plan;
    std::string err;
    if (build_plan(r->qubits, s->cfg.dtype, r->ops.data(), (int64_t)r->ops.size(), r->iterations, lt, &plan, &err,
                   true) != HGM_OK)
        return;
    std::vector<int> passes, kernel;
    for (size_t i = 0; i < plan.passes.size(); ++i)
        if (plan.passes[i].op_count > 0) passes.push_back((int)i);
    if (passes.empty() || (double)passes.size() * kLargeTilePassCost >= (double)r->passes.size()) return;
    {
        std::lock_guard<std::mutex> g(s->kmu);
        for (int i : passes) {
            auto it = s->kernel_of.find(serving_pass_key(plan, plan.passes[i]));
            if (it == s->kernel_of.end()) return;
            kernel.push_back(it->second);
        }
    }
    r->plan = std::move(plan);
    r->passes = std::move(passes);
    r->kernel = std::move(kernel);
}

void planner_loop(hgm_server *s) {
    for (;;) {
        Req *r;
        hgm_tuning t, lt;
        {
            std::unique_lock<std::mutex> g(s->mu);
            s->cv_plan.wait(g, [&] { return s->stop || !s->plan_q.empty(); });
            if (s->plan_q.empty()) return;
            r = s->plan_q.front();
            s->plan_q.pop_front();
            t = s->tune[r->qubits];
            lt = s->large[r->qubits];
        }
        std::string err;
        int rc = build_plan(r->qubits, s->cfg.dtype, r->ops.data(), (int64_t)r->ops.size(), r->iterations, t,
                            &r->plan, &err);
        if (rc == HGM_OK && r->plan.kernel != HGM_KERNEL_REG) {
            rc = HGM_E_TUNING;
            err = "serving needs the register kernel for this tile shape";
        }
        if (rc == HGM_OK && !serving_amps_ok(2 * r->plan.shape.R)) {   // fails this request only
            rc = HGM_E_TUNING;
            err = "the serving kernel holds 2, 4, 8, 16 or 32 amplitudes per thread";
        }
        if (rc == HGM_OK) {
            for (size_t i = 0; i < r->plan.passes.size(); ++i)
                if (r->plan.passes[i].op_count > 0) r->passes.push_back((int)i);
            if (r->passes.empty()) { rc = HGM_E_ARG; err = "the request applies no gate"; }
        }
        if (rc == HGM_OK) {   // the compiled kernel of each pass, where one exists
            r->kernel.assign(r->passes.size(), -1);
            std::lock_guard<std::mutex> g(s->kmu);
            if (!s->kernel_of.empty() && r->qubits > r->plan.shape.r)   // one-tile states stay interpreted
                for (size_t i = 0; i < r->passes.size(); ++i) {
                    auto it = s->kernel_of.find(serving_pass_key(r->plan, r->plan.passes[r->passes[i]]));
                    if (it != s->kernel_of.end()) r->kernel[i] = it->second;
                }
        }
        if (rc == HGM_OK && lt.block_size > 0) try_large_tile(s, r, lt);
        if (rc == HGM_OK) {
            const Plan &p = r->plan;
            r->n_tiles = int64_t(1) << (r->qubits - p.shape.r);
            r->lane = r->n_tiles == 1 ? 0 : 1;
            r->state_bytes = align_up((size_t(1) << r->qubits) * amp_bytes(s->cfg.dtype));
            r->prog_bytes = align_up(p.rops.size() * sizeof(RegOp)) + align_up(p.maps.size() * sizeof(PhaseMap)) +
                            align_up(p.mems.size() * sizeof(PassMem));
            r->part_bytes = align_up((size_t)r->n_tiles * sizeof(ReadoutPartial));
            r->total_bytes = r->state_bytes + r->prog_bytes + r->part_bytes;
            if (r->total_bytes > s->arena.size() || r->prog_bytes > s->staging.size()) {
                rc = HGM_E_ARG;
                err = "the request's footprint exceeds the serving HBM budget";
            }
        }
        r->status = rc;
        r->err = err;
        r->planned.store(rc == HGM_OK ? 1 : 2, std::memory_order_release);
        s->cv_work.notify_all();
    }
}

// Returns a request's device and staging blocks (its work has completed).
void release_req(hgm_server *s, Req *r) {
    s->arena.release(r->off, r->total_bytes);
    if (r->prog_bytes) s->staging.release(r->stage_off, r->prog_bytes);
    --s->active;
}

// Admit planned requests in arrival order while memory and slots allow.
bool admit(hgm_server *s) {
    bool any = false;
    for (auto it = s->pending.begin(); it != s->pending.end();) {
        Req *r = *it;
        const int st = r->planned.load(std::memory_order_acquire);
        if (st == 0) { ++it; continue; }
        if (st == 2) {
            it = s->pending.erase(it);
            fail(s, r, r->status, r->err);
            any = true;
            continue;
        }
        if (s->active >= s->cfg.max_active) break;
        size_t off = 0, soff = 0;
        if (!s->arena.alloc(r->total_bytes, &off)) break;   // the oldest waits for memory
        if (r->prog_bytes && !s->staging.alloc(r->prog_bytes, &soff)) {
            s->arena.release(off, r->total_bytes);
            break;
        }
        r->off = off;
        r->stage_off = soff;
        r->d = s->arena.base() + off;
        Lane &L = s->lanes[r->lane];
        const Plan &p = r->plan;
        // programs go through pinned memory, so that the copy is queued on the
        // lane's stream without waiting for the steps already in flight
        char *h = s->staging.base() + soff, *q = h;
        if (!p.rops.empty()) std::memcpy(q, p.rops.data(), p.rops.size() * sizeof(RegOp));
        q += align_up(p.rops.size() * sizeof(RegOp));
        if (!p.maps.empty()) std::memcpy(q, p.maps.data(), p.maps.size() * sizeof(PhaseMap));
        q += align_up(p.maps.size() * sizeof(PhaseMap));
        if (!p.mems.empty()) std::memcpy(q, p.mems.data(), p.mems.size() * sizeof(PassMem));
        hipError_t e = hipMemcpyAsync(r->d + r->state_bytes, h, r->prog_bytes, hipMemcpyHostToDevice, L.stream);
        if (e == hipSuccess)
            e = launch_set_basis(s->cfg.dtype, r->d, uint64_t(1) << r->qubits, r->basis, L.stream);
        it = s->pending.erase(it);
        any = true;
        ++s->active;
        r->t_admit = now_ms(s->t0);
        if (e != hipSuccess) {
            // the copy may still reference the blocks: let the stream drain first
            hipStreamSynchronize(L.stream);
            release_req(s, r);
            fail(s, r, HGM_E_HIP, hip_msg(e, "request upload"));
            continue;
        }
        L.active.push_back(r);
    }
    return any;
}

RegOp *dev_rops(Req *r) { return reinterpret_cast<RegOp *>(r->d + r->state_bytes); }
PhaseMap *dev_maps(Req *r) {
    return reinterpret_cast<PhaseMap *>(r->d + r->state_bytes + align_up(r->plan.rops.size() * sizeof(RegOp)));
}
PassMem *dev_mems(Req *r) {
    return reinterpret_cast<PassMem *>(r->d + r->state_bytes + align_up(r->plan.rops.size() * sizeof(RegOp)) +
                                       align_up(r->plan.maps.size() * sizeof(PhaseMap)));
}
ReadoutPartial *dev_partials(Req *r) {
    return reinterpret_cast<ReadoutPartial *>(r->d + r->state_bytes + r->prog_bytes);
}

// Launches the next pass of every active state of the lane as one step.
bool launch_step(hgm_server *s, Lane &L) {
    if (L.active.empty() || (int)L.inflight.size() >= kDepth) return false;
    const int slot = L.next_slot;
    L.next_slot = (L.next_slot + 1) % kDepth;
    char *h = L.h_ring + (size_t)slot * s->slot_bytes;
    char *d = L.d_ring + (size_t)slot * s->slot_bytes;
    BatchItem *items = reinterpret_cast<BatchItem *>(h);
    ReadoutTask *tasks = reinterpret_cast<ReadoutTask *>(h + s->tasks_off);
    ReadoutPartial *d_results = reinterpret_cast<ReadoutPartial *>(d + s->results_off);

    std::vector<Req *> order(L.active.begin(), L.active.end());
    auto kern = [](const Req *r) { return r->kernel[r->next]; };
    std::stable_sort(order.begin(), order.end(), [&](const Req *a, const Req *b) {
        if (a->plan.shape.R != b->plan.shape.R) return a->plan.shape.R < b->plan.shape.R;
        if (a->plan.shape.B != b->plan.shape.B) return a->plan.shape.B < b->plan.shape.B;
        return kern(a) < kern(b);
    });
    // one launch per workgroup shape and kernel: the interpreter (-1) or a
    // compiled pass kernel shared by the items whose passes have its structure
    struct Group { int A, B, kernel; int first, count; int64_t tiles; size_t lds; };
    std::vector<Group> groups;
    Step step;
    step.slot = slot;
    int n_tasks = 0;
    for (int i = 0; i < (int)order.size(); ++i) {
        Req *r = order[i];
        const int A = 2 * r->plan.shape.R, B = r->plan.shape.B, K = kern(r);
        if (groups.empty() || groups.back().A != A || groups.back().B != B || groups.back().kernel != K)
            groups.push_back({A, B, K, i, 0, 0, 0});
        Group &g = groups.back();
        const Pass &ps = r->plan.passes[r->passes[r->next]];
        const bool last = r->next + 1 == r->passes.size();
        BatchItem &it = items[i];
        char *const arena = s->arena.base();
        it.state_off = (uint64_t)(r->d - arena);
        it.ops_off = (uint64_t)(reinterpret_cast<char *>(dev_rops(r) + ps.rop_begin) - arena);
        it.maps_off = (uint64_t)(reinterpret_cast<char *>(dev_maps(r)) - arena);
        it.pm_off = (uint64_t)(reinterpret_cast<char *>(dev_mems(r) + ps.mem) - arena);
        it.nonlocal_mask = ((uint64_t(1) << r->qubits) - 1) & ~ps.local_mask;
        it.tile_begin = g.tiles;
        it.partials_off = last ? (int64_t)(reinterpret_cast<char *>(dev_partials(r)) - arena) : -1;
        it.n_ops = (int32_t)ps.rop_count;
        it.pad = 0;
        g.count++;
        g.tiles += r->n_tiles;
        if (ps.phases > 1)
            g.lds = std::max(g.lds, (size_t(1) << r->plan.shape.r) * amp_bytes(s->cfg.dtype));
        if (last) {
            tasks[n_tasks].partials = dev_partials(r);
            tasks[n_tasks].n = r->n_tiles;
            tasks[n_tasks].out = d_results + n_tasks;
            ++n_tasks;
            step.finishing.push_back(r);
        }
        r->next++;
    }
    hipError_t e = hipMemcpyAsync(d, h, order.size() * sizeof(BatchItem), hipMemcpyHostToDevice, L.stream);
    if (e == hipSuccess && n_tasks)
        e = hipMemcpyAsync(d + s->tasks_off, h + s->tasks_off, n_tasks * sizeof(ReadoutTask), hipMemcpyHostToDevice,
                           L.stream);
    for (const Group &g : groups) {
        if (e != hipSuccess) break;
        const BatchItem *gi = reinterpret_cast<const BatchItem *>(d) + g.first;
        if (g.kernel < 0) {
            e = launch_batch_reg(s->cfg.dtype, g.A, g.B, gi, g.count, g.tiles, g.lds, s->arena.base(), L.stream);
            continue;
        }
        ServeKernel k;
        {
            std::lock_guard<std::mutex> lk(s->kmu);
            k = s->kernels[(size_t)g.kernel];
        }
        if ((uint64_t)g.tiles * (uint64_t)g.B > 0xffffffffull) { e = hipErrorInvalidValue; break; }
        int count = g.count;
        char *arena = s->arena.base();
        void *args[] = {&gi, &count, &arena, &arena, &arena};
        e = hipModuleLaunchKernel(k.fn, (unsigned)g.tiles, 1, 1, (unsigned)g.B, 1, 1, (unsigned)k.lds, L.stream,
                                  args, nullptr);
    }
    if (e == hipSuccess && n_tasks)
        e = launch_readout_final(reinterpret_cast<const ReadoutTask *>(d + s->tasks_off), n_tasks, L.stream);
    if (e == hipSuccess && n_tasks)
        e = hipMemcpyAsync(h + s->results_off, d + s->results_off, n_tasks * sizeof(ReadoutPartial),
                           hipMemcpyDeviceToHost, L.stream);
    for (Req *r : step.finishing) {
        if (e != hipSuccess) break;
        if (r->state_out)
            e = hipMemcpyAsync(r->state_out, r->d, (size_t(1) << r->qubits) * amp_bytes(s->cfg.dtype),
                               hipMemcpyDeviceToHost, L.stream);
    }
    if (e == hipSuccess) e = hipEventRecord(L.ev[slot], L.stream);
    if (e != hipSuccess) {
        // a failed launch leaves the lane's states undefined: fail all of them
        hipStreamSynchronize(L.stream);
        const std::string msg = hip_msg(e, "serving step");
        {
            std::lock_guard<std::mutex> g(s->mu);
            if (s->fatal.empty()) s->fatal = msg;
        }
        for (Req *r : L.active) {
            release_req(s, r);
            fail(s, r, HGM_E_HIP, msg);
        }
        L.active.clear();
        return true;
    }
    L.active.erase(std::remove_if(L.active.begin(), L.active.end(),
                                  [](const Req *r) { return r->next == r->passes.size(); }),
                   L.active.end());
    L.inflight.push_back(std::move(step));
    return true;
}

bool complete_steps(hgm_server *s, Lane &L) {
    bool any = false;
    while (!L.inflight.empty()) {
        Step &st = L.inflight.front();
        const hipError_t q = hipEventQuery(L.ev[st.slot]);
        if (q == hipErrorNotReady) break;
        const ReadoutPartial *res =
            reinterpret_cast<const ReadoutPartial *>(L.h_ring + (size_t)st.slot * s->slot_bytes + s->results_off);
        const double t = now_ms(s->t0);
        for (size_t k = 0; k < st.finishing.size(); ++k) {
            Req *r = st.finishing[k];
            release_req(s, r);
            if (q != hipSuccess) {
                fail(s, r, HGM_E_HIP, hip_msg(q, "serving step"));
                continue;
            }
            r->result = res[k];
            r->t_done = t;
            finish(s, r);
        }
        L.inflight.pop_front();
        any = true;
    }
    return any;
}

void worker_loop(hgm_server *s) {
    hipSetDevice(s->device);
    for (;;) {
        bool stopping;
        {
            std::lock_guard<std::mutex> g(s->mu);
            while (!s->submitted.empty()) {
                s->pending.push_back(s->submitted.front());
                s->submitted.pop_front();
            }
            stopping = s->stop;
        }
        bool progress = false;
        for (Lane &L : s->lanes) progress |= complete_steps(s, L);
        progress |= admit(s);
        for (Lane &L : s->lanes) progress |= launch_step(s, L);
        bool idle = s->pending.empty();
        for (Lane &L : s->lanes) idle = idle && L.active.empty() && L.inflight.empty();
        if (stopping && idle) return;
        if (progress) continue;
        if (idle) {
            std::unique_lock<std::mutex> g(s->mu);
            s->cv_work.wait_for(g, std::chrono::milliseconds(5),
                                [&] { return s->stop || !s->submitted.empty(); });
        } else {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
}

void release_lanes(hgm_server *s) {
    for (Lane &L : s->lanes) {
        if (L.stream) hipStreamSynchronize(L.stream);
        for (hipEvent_t &e : L.ev)
            if (e) { hipEventDestroy(e); e = nullptr; }
        if (L.d_ring) hipFree(L.d_ring);
        if (L.h_ring) hipHostFree(L.h_ring);
        if (L.stream) hipStreamDestroy(L.stream);
        L.d_ring = L.h_ring = nullptr;
        L.stream = nullptr;
    }
}

}  // namespace

extern "C" {

void hgm_server_default_config(hgm_server_config *c) {
    std::memset(c, 0, sizeof *c);
    c->dtype = HGM_C64;
    c->max_active = 256;
    c->hbm_budget_bytes = uint64_t(32) << 30;
    c->planner_threads = 4;
    c->autotune = 0;
    c->min_qubits = 10;
    c->max_qubits = 26;
}

int hgm_server_create(const hgm_server_config *cfg, hgm_server **out) {
    if (!cfg || !out) { set_error("null argument"); return HGM_E_ARG; }
    if (cfg->dtype != HGM_C64 && cfg->dtype != HGM_C128) { set_error("unknown amplitude dtype"); return HGM_E_ARG; }
    if (cfg->max_active < 1 || cfg->max_active > 65536 || cfg->planner_threads < 1 || cfg->planner_threads > 64 ||
        cfg->hbm_budget_bytes < (uint64_t(1) << 20)) {
        set_error("server config out of range (max_active 1..65536, planner_threads 1..64, budget >= 1 MiB)");
        return HGM_E_ARG;
    }
    if (cfg->autotune && (cfg->min_qubits < 1 || cfg->max_qubits > kMaxQubits || cfg->min_qubits > cfg->max_qubits)) {
        set_error("autotune qubit range out of range");
        return HGM_E_ARG;
    }
    int rc = hgm_init(-1);
    if (rc) return rc;
    auto s = std::make_unique<hgm_server>();
    s->cfg = *cfg;
    hipGetDevice(&s->device);
    s->t0 = std::chrono::steady_clock::now();
    s->tune.resize(kMaxQubits + 1);
    s->large.resize(kMaxQubits + 1);
    for (int n = 1; n <= kMaxQubits; ++n) {
        s->tune[n] = serving_default_tuning(n, cfg->dtype);
        if (!serving_large_tile(n, cfg->dtype, &s->large[n])) s->large[n].block_size = 0;
    }
    if (cfg->autotune) {
        for (int n = std::max(cfg->min_qubits, 14); n <= cfg->max_qubits; ++n) {
            std::vector<hgm_op> ops;
            try {
                ops = mix_timing_circuit(n);
            } catch (const std::exception &e) {
                set_error(std::string("serving autotune: ") + e.what());
                return HGM_E_ARG;
            }
            hgm_tuning best;
            double ms;
            if ((rc = hgm_autotune(n, cfg->dtype, ops.data(), (int64_t)ops.size(), 1, HGM_KERNEL_REG,
                                   kServingPlanner, &best, &ms)))
                return rc;
            best.ops_block = 1 << 30;
            best.planner = kServingPlanner;
            s->tune[n] = best;
        }
    }
    std::string err;
    if ((rc = s->arena.init(cfg->hbm_budget_bytes, false, &err))) { set_error(err); return rc; }
    // pinned staging for programs: 1/16 of the device budget, at least 64 MiB and at most 1 GiB
    const size_t stage = std::min<size_t>(size_t(1) << 30,
                                          std::max<size_t>(size_t(64) << 20, cfg->hbm_budget_bytes / 16));
    if ((rc = s->staging.init(stage, true, &err))) { s->arena.destroy(); set_error(err); return rc; }
    const size_t m = (size_t)cfg->max_active;
    s->tasks_off = align_up(m * sizeof(BatchItem));
    s->results_off = s->tasks_off + align_up(m * sizeof(ReadoutTask));
    s->slot_bytes = s->results_off + align_up(m * sizeof(ReadoutPartial));
    for (Lane &L : s->lanes) {
        hipError_t e = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking);
        for (hipEvent_t &ev : L.ev)
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc(&L.d_ring, kDepth * s->slot_bytes);
        if (e == hipSuccess) e = hipHostMalloc(&L.h_ring, kDepth * s->slot_bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            set_error(hip_msg(e, "serving lane set-up"));
            release_lanes(s.get());
            s->arena.destroy();
            s->staging.destroy();
            return HGM_E_HIP;
        }
    }
    hgm_server *p = s.release();
    for (int i = 0; i < cfg->planner_threads; ++i) p->planners.emplace_back(planner_loop, p);
    p->worker = std::thread(worker_loop, p);
    *out = p;
    return HGM_OK;
}

int hgm_server_submit(hgm_server *s, const hgm_request *rq, int64_t *id) {
    if (!s || !rq) { set_error("null argument"); return HGM_E_ARG; }
    if (rq->qubits < 1 || rq->qubits > kMaxQubits) { set_error("qubits out of range"); return HGM_E_ARG; }
    if (rq->n_ops < 1 || !rq->ops || rq->iterations < 1) {
        set_error("a request needs at least one gate and one iteration");
        return HGM_E_ARG;
    }
    if (rq->basis >> rq->qubits) { set_error("basis index outside the state"); return HGM_E_ARG; }
    Req *r = new Req;
    r->qubits = rq->qubits;
    r->iterations = rq->iterations;
    r->basis = rq->basis;
    r->state_out = rq->state_out;
    r->ops.assign(rq->ops, rq->ops + rq->n_ops);
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) { delete r; set_error("server is shutting down"); return HGM_E_STATE; }
        r->id = s->next_id++;
        r->t_submit = now_ms(s->t0);
        ++s->n_submitted;
        s->plan_q.push_back(r);
        s->submitted.push_back(r);
    }
    if (id) *id = r->id;
    s->cv_plan.notify_one();
    s->cv_work.notify_all();
    return HGM_OK;
}

int hgm_server_poll(hgm_server *s, hgm_result *out, int cap, int *n, int timeout_ms) {
    if (!s || !n || cap < 0 || (cap > 0 && !out)) { set_error("bad argument"); return HGM_E_ARG; }
    std::unique_lock<std::mutex> g(s->mu);
    if (s->done.empty() && timeout_ms > 0)
        s->cv_done.wait_for(g, std::chrono::milliseconds(timeout_ms), [&] { return !s->done.empty(); });
    int k = 0;
    while (k < cap && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    *n = k;
    return HGM_OK;
}

int hgm_server_wait_all(hgm_server *s) {
    if (!s) { set_error("null argument"); return HGM_E_ARG; }
    std::unique_lock<std::mutex> g(s->mu);
    s->cv_done.wait(g, [&] { return s->n_done == s->n_submitted; });
    if (!s->fatal.empty()) { set_error(s->fatal); return HGM_E_HIP; }
    return HGM_OK;
}

int hgm_server_tuning(hgm_server *s, int qubits, hgm_tuning *out) {
    if (!s || !out || qubits < 1 || qubits > kMaxQubits) { set_error("bad argument"); return HGM_E_ARG; }
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->tune[qubits];
    return HGM_OK;
}

int hgm_server_set_tuning(hgm_server *s, int qubits, const hgm_tuning *t) {
    if (!s || !t || qubits < 1 || qubits > kMaxQubits) { set_error("bad argument"); return HGM_E_ARG; }
    if (t->kernel != HGM_KERNEL_REG) { set_error("the serving engine runs the register kernel (HGM_KERNEL_REG)"); return HGM_E_TUNING; }
    TileShape shape;
    std::string err;
    if (int rc = make_shape(qubits, s->cfg.dtype, *t, &shape, &err)) { set_error(err); return rc; }
    const int A = 2 * shape.R;
    if (!serving_amps_ok(A)) { set_error("the serving kernel holds 2, 4, 8, 16 or 32 amplitudes per thread"); return HGM_E_TUNING; }
    if (shape.B > reg_kernel_max_threads(A, amp_bytes(s->cfg.dtype))) { set_error("block_size above the register kernel's limit for this repeat_count"); return HGM_E_TUNING; }
    std::lock_guard<std::mutex> g(s->mu);   // planner threads copy their entry under this lock
    s->tune[qubits] = *t;
    s->large[qubits].block_size = 0;        // the caller's shape replaces the 2^15-tile alternative
    return HGM_OK;
}

int hgm_server_precompile(hgm_server *s, int qubits, const hgm_op *ops, int64_t n_ops, int iterations, int *n_new) {
    if (n_new) *n_new = 0;
    if (!s || !ops || n_ops < 1 || iterations < 1 || qubits < 1 || qubits > kMaxQubits) {
        set_error("bad argument");
        return HGM_E_ARG;
    }
    hgm_tuning t, lt;
    {
        std::lock_guard<std::mutex> g(s->mu);
        t = s->tune[qubits];
        lt = s->large[qubits];
        if (qubits <= t.gpu_region_bits) return HGM_OK;   // one tile: the interpreter runs the whole circuit
        t.planner = kCompiledPlanner;
        s->tune[qubits] = t;   // requests planned from now on find the kernels
    }
    Plan plan, large;
    std::string err;
    int rc = build_plan(qubits, s->cfg.dtype, ops, n_ops, iterations, t, &plan, &err);
    if (rc) { set_error(err); return rc; }
    if (plan.kernel != HGM_KERNEL_REG) { set_error("serving needs the register kernel for this tile shape"); return HGM_E_TUNING; }
    auto busy = [](const Plan &p) {
        size_t k = 0;
        for (const Pass &ps : p.passes) k += ps.op_count > 0;
        return k;
    };
    // the 2^15-tile plan's kernels too, where a request of this shape would take that plan
    bool with_large = false;
    if (lt.block_size > 0) {
        rc = build_plan(qubits, s->cfg.dtype, ops, n_ops, iterations, lt, &large, &err, true);
        if (rc) { set_error(err); return rc; }
        with_large = (double)busy(large) * kLargeTilePassCost < (double)busy(plan);
    }
    struct Job {
        std::string key;
        JitCode code;
        size_t lds = 0;
        int rc = HGM_OK;
        std::string err;
    };
    std::vector<Job> jobs;
    {
        std::lock_guard<std::mutex> g(s->kmu);
        std::unordered_set<std::string> seen;
        for (const Plan *p : {&plan, &large}) {
            if (p == &large && !with_large) continue;
            for (const Pass &ps : p->passes) {
                if (!ps.op_count) continue;
                std::string key = serving_pass_key(*p, ps);
                if (s->kernel_of.count(key) || !seen.insert(key).second) continue;
                Job j;
                j.key = std::move(key);
                j.code.source = serving_kernel_source(*p, ps, "hgm_serve", &j.lds);
                j.code.names = {"hgm_serve"};
                jobs.push_back(std::move(j));
            }
        }
    }
    // hipRTC compiles independent programs concurrently
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nth = std::min<size_t>(jobs.size(), std::min(16u, hw));
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (size_t w = 0; w < nth; ++w)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < jobs.size();) jobs[i].rc = jit_compile(&jobs[i].code, &jobs[i].err);
        });
    for (std::thread &th : pool) th.join();
    int prev = 0;
    hipGetDevice(&prev);
    hipSetDevice(s->device);
    int added = 0;
    for (Job &j : jobs) {
        if (j.rc) { rc = j.rc; set_error(j.err); break; }
        if (j.lds > kServeMaxLds) continue;   // stays interpreted
        ServeKernel k;
        k.lds = j.lds;
        hipError_t e = hipModuleLoadData(&k.mod, j.code.binary.data());
        if (e == hipSuccess) e = hipModuleGetFunction(&k.fn, k.mod, "hgm_serve");
        if (e == hipSuccess && j.lds > 64 * 1024)   // above the default dynamic LDS of a launch
            e = hipFuncSetAttribute(reinterpret_cast<const void *>(k.fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)j.lds);
        if (e != hipSuccess) {
            if (k.mod) hipModuleUnload(k.mod);
            set_error(hip_msg(e, "loading a serving pass kernel"));
            rc = HGM_E_HIP;
            break;
        }
        std::lock_guard<std::mutex> g(s->kmu);
        s->kernel_of[j.key] = (int)s->kernels.size();
        s->kernels.push_back(k);
        ++added;
    }
    hipSetDevice(prev);
    if (n_new) *n_new = added;
    return rc;
}

double hgm_server_clock_ms(const hgm_server *s) { return s ? now_ms(s->t0) : 0.0; }

void hgm_server_destroy(hgm_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_plan.notify_all();
    s->cv_work.notify_all();
    for (std::thread &t : s->planners) t.join();
    s->worker.join();
    release_lanes(s);
    s->arena.destroy();
    s->staging.destroy();
    for (ServeKernel &k : s->kernels)
        if (k.mod) hipModuleUnload(k.mod);
    delete s;
}

int hgm_autotune(int qubits, int dtype, const hgm_op *ops, int64_t n_ops, int iterations, int kernel, int planner,
                 hgm_tuning *best, double *best_ms) {
    if (!best || (n_ops > 0 && !ops) || iterations < 1) { set_error("bad argument"); return HGM_E_ARG; }
    if (kernel < HGM_KERNEL_AUTO || kernel > HGM_KERNEL_JIT) { set_error("unknown kernel family"); return HGM_E_ARG; }
    const bool jit = kernel == HGM_KERNEL_JIT;
    const int r_lo = jit ? 10 : 8, c_run = dtype == HGM_C64 ? 4 : 3;
    hgm_state *st = nullptr;
    int rc = hgm_state_create(qubits, dtype, &st);
    if (rc) return rc;
    double best_t = 1e300;
    bool found = false;
    for (int A : {8, 16, 32}) {
        for (int r = r_lo; r <= 13; ++r) {
            const int64_t B = (int64_t(1) << r) / A;
            if (B < 64 || B > 1024 || (kernel != HGM_KERNEL_LDS && B > reg_kernel_max_threads(A, amp_bytes(dtype)))) continue;
            // plan-compiled candidates: runs of 128 and 256 bytes per wave segment
            for (int c = jit ? c_run : 2; c <= (jit ? c_run + 1 : 4); ++c) {
                hgm_tuning t;
                hgm_default_tuning(&t);
                t.block_size = (int)B;
                t.repeat_count = A / 2;
                t.gpu_region_bits = r;
                t.gpu_coalesced_bits = c;
                t.kernel = kernel;
                t.planner = planner;
                t.ops_block = 1 << 30;
                if (hgm_validate_tuning(qubits, dtype, &t) != HGM_OK) continue;
                hgm_plan *p = nullptr;
                if (hgm_plan_build(qubits, dtype, ops, n_ops, iterations, &t, &p) != HGM_OK) continue;
                double ms = 1e300;
                hgm_state_set_basis(st, 0);
                rc = hgm_state_run_plan(st, p, nullptr);   // warm-up
                for (int rep = 0; rep < 2 && rc == HGM_OK; ++rep) {
                    hgm_stats stats{};
                    rc = hgm_state_run_plan(st, p, &stats);
                    ms = std::min(ms, stats.kernel_ms);
                }
                hgm_plan_destroy(p);
                if (rc) { hgm_state_destroy(st); return rc; }
                if (ms < best_t) { best_t = ms; *best = t; found = true; }
            }
        }
    }
    hgm_state_destroy(st);
    if (!found) { set_error("no valid tile shape for this state"); return HGM_E_TUNING; }
    if (best_ms) *best_ms = best_t;
    return HGM_OK;
}

}  // extern "C"
