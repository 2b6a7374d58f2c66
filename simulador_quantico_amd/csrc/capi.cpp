// extern "C" layer over the executor: the DGM handle, the circuit generators
// by name, and the reference-compatible entry points of include/core/dgm.h:22-32.
#include <chrono>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hgm_dgm.hpp"
#include "hgm_gpu.h"
#include "planner.hpp"
#include "runtime.hpp"

static_assert(offsetof(hgm_ref_pt, matrix) == 8, "PT layout (include/core/common.h:57-90)");
static_assert(offsetof(hgm_ref_pt, target_bit) == 24, "PT layout");
static_assert(offsetof(hgm_ref_pt, control_value) == 32, "PT layout");
static_assert(offsetof(hgm_ref_pt, control_mask) == 40, "PT layout");
static_assert(offsetof(hgm_ref_pt, control_count) == 56, "PT layout");
static_assert(sizeof(hgm_ref_pt) == 80, "PT layout");
static_assert(sizeof(hgm_op) == 88, "hgm_op layout");

struct hgm_dgm {
    int dtype;
    std::unique_ptr<hgm::DGM> d64;
    std::unique_ptr<hgm::DGM32> d32;
    hgm::DGMCore &core() { return d64 ? static_cast<hgm::DGMCore &>(*d64) : static_cast<hgm::DGMCore &>(*d32); }
    const hgm::DGMCore &core() const {
        return d64 ? static_cast<const hgm::DGMCore &>(*d64) : static_cast<const hgm::DGMCore &>(*d32);
    }
};

namespace hgm {

hgm_tuning env_tuning(const hgm_tuning &caller, int qubits, int dtype, const hgm_op *ops, int64_t n_ops,
                      int iterations) {
    const char *e = std::getenv("HGM_TUNING");
    if (!e) return caller;
    const std::string mode(e);
    if (mode != "auto" && mode != "autotune") return caller;
    hgm_tuning t;
    if (hgm_auto_tuning(qubits, dtype, &t)) return caller;
    t.ops_block = caller.ops_block;
    if (mode == "auto") return t;
    // a tile autotuned once per (qubits, dtype), on the first circuit seen
    static std::mutex mu;
    static std::map<std::pair<int, int>, hgm_tuning> tuned;
    std::lock_guard<std::mutex> g(mu);
    auto it = tuned.find({qubits, dtype});
    if (it == tuned.end()) {
        hgm_tuning best;
        double ms = 0;
        if (hgm_autotune(qubits, dtype, ops, n_ops, iterations, t.kernel, t.planner, &best, &ms)) return t;
        best.kernel = t.kernel;
        best.planner = t.planner;
        it = tuned.emplace(std::make_pair(qubits, dtype), best).first;
    }
    hgm_tuning r = it->second;
    r.ops_block = caller.ops_block;
    return r;
}

}  // namespace hgm

namespace {

template <class F>
int guard(F fn) {
    try {
        return fn();
    } catch (const hgm::Error &e) {
        hgm::set_error(e.what());
        return e.code ? e.code : HGM_E_ARG;
    } catch (const std::exception &e) {
        hgm::set_error(e.what());
        return HGM_E_ARG;
    }
}

template <class F>
auto visit(hgm_dgm *d, F fn) {
    return d->d64 ? fn(*d->d64) : fn(*d->d32);
}

long *field(hgm::DGMCore &c, const std::string &f) {
    if (f == "qubits") return &c.qubits;
    if (f == "thread_count") return &c.thread_count;
    if (f == "cpu_region_bits") return &c.cpu_region_bits;
    if (f == "cpu_coalesced_bits") return &c.cpu_coalesced_bits;
    if (f == "gpu_region_bits") return &c.gpu_region_bits;
    if (f == "gpu_coalesced_bits") return &c.gpu_coalesced_bits;
    return nullptr;
}

int *ifield(hgm::DGMCore &c, const std::string &f) {
    if (f == "exec_type") return &c.exec_type;
    if (f == "gpu_count") return &c.gpu_count;
    if (f == "block_size") return &c.block_size;
    if (f == "repeat_count") return &c.repeat_count;
    if (f == "ops_block") return &c.ops_block;
    if (f == "planner") return &c.planner;
    if (f == "kernel") return &c.kernel;
    return nullptr;
}

std::string join(const std::vector<std::string> &v) {
    std::string r;
    for (size_t i = 0; i < v.size(); ++i) { if (i) r += ";"; r += v[i]; }
    return r;
}

}  // namespace

extern "C" {

int hgm_dgm_create(int dtype, hgm_dgm **out) {
    if (!out || (dtype != HGM_C128 && dtype != HGM_C64)) { hgm::set_error("bad dtype"); return HGM_E_ARG; }
    hgm_dgm *h = new hgm_dgm();
    h->dtype = dtype;
    if (dtype == HGM_C128) h->d64.reset(new hgm::DGM());
    else h->d32.reset(new hgm::DGM32());
    h->core().exit_on_error = false;
    *out = h;
    return HGM_OK;
}

void hgm_dgm_destroy(hgm_dgm *d) { delete d; }

int hgm_dgm_set(hgm_dgm *d, const char *name, int64_t value) {
    if (!d || !name) return HGM_E_ARG;
    hgm::DGMCore &c = d->core();
    const std::string f(name);
    if (long *p = field(c, f)) { *p = (long)value; return HGM_OK; }
    if (int *p = ifield(c, f)) { *p = (int)value; return HGM_OK; }
    if (f == "resident") { c.resident = value != 0; return HGM_OK; }
    hgm::set_error("unknown DGM field '" + f + "'");
    return HGM_E_ARG;
}

int hgm_dgm_get(const hgm_dgm *d, const char *name, int64_t *value) {
    if (!d || !name || !value) return HGM_E_ARG;
    hgm::DGMCore &c = const_cast<hgm_dgm *>(d)->core();
    const std::string f(name);
    if (long *p = field(c, f)) { *value = *p; return HGM_OK; }
    if (int *p = ifield(c, f)) { *value = *p; return HGM_OK; }
    if (f == "resident") { *value = c.resident; return HGM_OK; }
    hgm::set_error("unknown DGM field '" + f + "'");
    return HGM_E_ARG;
}

int hgm_dgm_allocate_memory(hgm_dgm *d) {
    return guard([&] { visit(d, [](auto &x) { x.allocateMemory(); return 0; }); return HGM_OK; });
}

int hgm_dgm_free_memory(hgm_dgm *d) {
    return guard([&] { visit(d, [](auto &x) { x.freeMemory(); return 0; }); return HGM_OK; });
}

int hgm_dgm_set_memory_value(hgm_dgm *d, int64_t pos) {
    return guard([&] {
        return visit(d, [&](auto &x) {
            if (!x.state || pos < 0 || pos >= (int64_t(1) << x.qubits)) {
                hgm::set_error("setMemoryValue: position outside the state");
                return (int)HGM_E_ARG;
            }
            x.setMemoryValue((long)pos);
            return (int)HGM_OK;
        });
    });
}

void *hgm_dgm_state(hgm_dgm *d) {
    if (!d) return nullptr;
    void *p = nullptr;
    if (guard([&] { p = visit(d, [](auto &x) { return static_cast<void *>(x.hostState()); }); return HGM_OK; }))
        return nullptr;
    return p;
}

int hgm_dgm_add_gate(hgm_dgm *d, const char *name, const double *m8, int *added) {
    if (!d || !name || !m8) return HGM_E_ARG;
    hgm::cplx m[4] = {{m8[0], m8[1]}, {m8[2], m8[3]}, {m8[4], m8[5]}, {m8[6], m8[7]}};
    bool ok = d->core().gates.addGate(name, m);
    if (added) *added = ok;
    return HGM_OK;
}

int hgm_dgm_get_gate(const hgm_dgm *d, const char *name, double *m8, int *found) {
    if (!d || !name || !m8) return HGM_E_ARG;
    const hgm::cplx *m = d->core().gates.getMatrix(name);
    if (found) *found = m != nullptr;
    if (m)
        for (int k = 0; k < 4; ++k) { m8[2 * k] = m[k].real(); m8[2 * k + 1] = m[k].imag(); }
    return HGM_OK;
}

int hgm_dgm_set_function(hgm_dgm *d, const char *function, int iterations, int reset) {
    if (!d || !function) return HGM_E_ARG;
    return guard([&] { d->core().setFunction(std::string(function), iterations, reset != 0); return HGM_OK; });
}

int hgm_dgm_set_function_steps(hgm_dgm *d, const char *const *steps, int n, int iterations, int reset) {
    if (!d || (n > 0 && !steps)) return HGM_E_ARG;
    std::vector<std::string> v;
    for (int i = 0; i < n; ++i) v.emplace_back(steps[i]);
    return guard([&] { d->core().setFunction(v, iterations, reset != 0); return HGM_OK; });
}

int hgm_dgm_num_pts(const hgm_dgm *d, int64_t *n) {
    if (!d || !n) return HGM_E_ARG;
    const hgm::DGMCore &c = d->core();
    int64_t k = 0;
    while (c.pts && c.pts[k]) ++k;
    *n = k;
    return HGM_OK;
}

int hgm_dgm_get_ops(const hgm_dgm *d, hgm_op *out, int64_t cap) {
    if (!d || (!out && cap)) return HGM_E_ARG;
    std::vector<hgm_op> o = d->core().ops();
    if ((int64_t)o.size() > cap) { hgm::set_error("op buffer too small"); return HGM_E_ARG; }
    std::memcpy(out, o.data(), o.size() * sizeof(hgm_op));
    return HGM_OK;
}

int hgm_dgm_execute(hgm_dgm *d, int iterations) {
    if (!d) return HGM_E_ARG;
    return guard([&] { visit(d, [&](auto &x) { x.execute(iterations); return 0; }); return HGM_OK; });
}

int hgm_dgm_execute_function(hgm_dgm *d, const char *function, int iterations) {
    if (!d || !function) return HGM_E_ARG;
    return guard([&] {
        visit(d, [&](auto &x) { x.executeFunction(std::string(function), iterations); return 0; });
        return HGM_OK;
    });
}

int hgm_dgm_measure(hgm_dgm *d, int qubit, int *outcome) {
    if (!d || !outcome) return HGM_E_ARG;
    return guard([&] { *outcome = visit(d, [&](auto &x) { return x.measure(qubit); }); return HGM_OK; });
}

int hgm_dgm_colapse(hgm_dgm *d, int qubit, int value) {
    if (!d) return HGM_E_ARG;
    return guard([&] {
        if (qubit < 0 || qubit >= (int)d->core().qubits || (value != 0 && value != 1))
            throw hgm::Error(HGM_E_ARG, "colapse: qubit outside the state or value not 0/1");
        visit(d, [&](auto &x) { x.colapse(qubit, value); return 0; });
        return HGM_OK;
    });
}

int hgm_dgm_measure_many(hgm_dgm *d, const int *qubits, int n, int64_t *values, double *probs, int64_t *n_out) {
    if (!d || (!qubits && n) || !n_out) return HGM_E_ARG;
    return guard([&] {
        std::vector<int> q(qubits, qubits + n);
        std::map<long, double> m = visit(d, [&](auto &x) { return x.measure(q); });
        int64_t cap = *n_out, k = 0;
        for (auto &kv : m) {
            if (k < cap && values && probs) { values[k] = kv.first; probs[k] = kv.second; }
            ++k;
        }
        *n_out = k;
        return HGM_OK;
    });
}

int hgm_dgm_sync_host(hgm_dgm *d) {
    if (!d) return HGM_E_ARG;
    return guard([&] { visit(d, [](auto &x) { x.syncHost(); return 0; }); return HGM_OK; });
}

int hgm_dgm_stats(const hgm_dgm *d, hgm_stats *out) {
    if (!d || !out) return HGM_E_ARG;
    *out = d->core().last_stats;
    return HGM_OK;
}

int hgm_dgm_erase(hgm_dgm *d) {
    if (!d) return HGM_E_ARG;
    d->core().erase();
    return HGM_OK;
}

int hgm_dgm_validate_tuning(hgm_dgm *d) {
    if (!d) return HGM_E_ARG;
    return guard([&] { d->core().validateTuning(); return HGM_OK; });
}

int hgm_dgm_count_ops(hgm_dgm *d, int iterations, int64_t counts[7]) {
    if (!d || !counts) return HGM_E_ARG;
    hgm::DGMCore &c = d->core();
    c.CountOps(iterations);
    counts[0] = c.total_op; counts[1] = c.dense; counts[2] = c.main_diag; counts[3] = c.sec_diag;
    counts[4] = c.c_dense; counts[5] = c.c_main_diag; counts[6] = c.c_sec_diag;
    return HGM_OK;
}

int hgm_dgm_print_pts(const hgm_dgm *d) {
    if (!d) return HGM_E_ARG;
    d->core().printPTs();
    return HGM_OK;
}

void hgm_seed_rng(uint64_t seed) { hgm::seed_rng((unsigned long)seed); }

void hgm_frac_approx(int *num, int *den, int width) { hgm::quantum_frac_approx(num, den, width); }
int64_t hgm_mul_inv(int64_t value, int64_t modulus) { return hgm::mul_inv((long)value, (long)modulus); }
int hgm_revert_bits(int value, int bits) { return hgm::revert_bits(value, bits); }

int hgm_shor(hgm_dgm *cfg, int64_t number, int64_t out[4], hgm_stats *stats) {
    if (!cfg || !out) { hgm::set_error("null argument"); return HGM_E_ARG; }
    if (number < 4 || number > (1 << 14)) {
        hgm::set_error("Shor needs 4 <= N <= 16384 (2*bits(N)+3 qubits)");
        return HGM_E_ARG;
    }
    return guard([&] {
        const hgm::DGMCore &c = cfg->core();
        const hgm::ShorResult r = cfg->d64 ? hgm::Shor<double>((long)number, c) : hgm::Shor<float>((long)number, c);
        out[0] = r.factors.size() == 2 ? r.factors[0] : 0;
        out[1] = r.factors.size() == 2 ? r.factors[1] : 0;
        out[2] = r.base;
        out[3] = r.phase_bits;
        if (stats) {   // added, like every other call that takes an hgm_stats*
            stats->gates += r.gates;
            stats->kernel_ms += r.kernel_ms;
        }
        return HGM_OK;
    });
}

int hgm_generate(hgm_dgm *d, const char *which, const int64_t *a, int na, const char *op, char *out, size_t cap,
                 size_t *len) {
    if (!which || !len || (na > 0 && !a)) return HGM_E_ARG;
    return guard([&]() -> int {
        const std::string w(which), o(op ? op : "");
        auto need = [&](int k) {
            if (na < k) throw hgm::Error(HGM_E_ARG, w + " needs " + std::to_string(k) + " integer arguments");
        };
        auto gates = [&]() -> hgm::Gates & {
            if (!d) throw hgm::Error(HGM_E_ARG, w + " registers gates and needs a DGM handle");
            return d->core().gates;
        };
        std::string s;
        if (w == "CNot") { need(4); s = hgm::CNot(a[0], a[1], a[2], a[3]); }
        else if (w == "Toffoli") { need(5); s = hgm::Toffoli(a[0], a[1], a[2], a[3], a[4]); }
        else if (w == "Controlled1") { need(4); s = hgm::Controlled1(a[0], a[1], a[2], o, a[3]); }
        else if (w == "Controlled2") { need(5); s = hgm::Controlled2(a[0], a[1], a[2], a[3], o, a[4]); }
        else if (w == "ControlledN") {   // qubits, target, control_value, ctrls...
            need(3);
            std::vector<int> ctrls(a + 3, a + na);
            s = hgm::ControlledN(a[0], ctrls, a[1], o, a[2]);
        }
        else if (w == "Pauli_X") { need(3); s = hgm::Pauli_X(a[0], a[1], a[2]); }
        else if (w == "Pauli_Z") { need(3); s = hgm::Pauli_Z(a[0], a[1], a[2]); }
        else if (w == "Hadamard") { need(3); s = hgm::Hadamard(a[0], a[1], a[2]); }
        else if (w == "Oracle1") { need(2); s = hgm::Oracle1(a[0], a[1]); }
        else if (w == "ControledZ") { need(1); s = hgm::ControledZ(a[0]); }
        else if (w == "QFT") { need(4); s = join(hgm::QFT(a[0], a[1], a[2], a[3], gates())); }
        else if (w == "RQFT") { need(4); s = join(hgm::RQFT(a[0], a[1], a[2], a[3], gates())); }
        else if (w == "AddF") { need(5); s = hgm::AddF(a[0], a[1], a[2], a[3], a[4], gates()); }
        else if (w == "SubF") { need(5); s = hgm::SubF(a[0], a[1], a[2], a[3], a[4], gates()); }
        else if (w == "AddFc") { need(6); s = join(hgm::AddF(a[0], a[1], a[2], a[3], a[4], a[5] != 0, gates())); }
        else if (w == "SubFc") { need(6); s = join(hgm::SubF(a[0], a[1], a[2], a[3], a[4], a[5] != 0, gates())); }
        else if (w == "CAddF") { need(6); s = hgm::CAddF(a[0], a[1], a[2], a[3], a[4], a[5], gates()); }
        else if (w == "C2AddF") { need(7); s = hgm::C2AddF(a[0], a[1], a[2], a[3], a[4], a[5], a[6], gates()); }
        else if (w == "CSubF") { need(6); s = hgm::CSubF(a[0], a[1], a[2], a[3], a[4], a[5], gates()); }
        else if (w == "C2SubF") { need(7); s = hgm::C2SubF(a[0], a[1], a[2], a[3], a[4], a[5], a[6], gates()); }
        else if (w == "genRot") { need(3); s = hgm::genRot(a[0], a[1], a[2], gates()); }
        else if (w == "CSwapR") { need(5); s = join(hgm::CSwapR(a[0], a[1], a[2], a[3], a[4])); }
        else if (w == "SwapOver") { need(3); s = join(hgm::SwapOver(a[0], a[1], a[2])); }
        else if (w == "GroverIterationRef") { need(2); s = join(hgm::GroverIterationRef(a[0], a[1])); }
        else if (w == "GroverIteration") { need(2); s = join(hgm::GroverIteration(a[0], (uint64_t)a[1])); }
        else if (w == "GroverIterations") { need(1); s = std::to_string(hgm::GroverIterations(a[0])); }
        else if (w == "RandomLayers") { need(3); s = join(hgm::RandomLayers(a[0], a[1], (uint64_t)a[2], gates())); }
        else if (w == "DraperAdder") { need(1); s = join(hgm::DraperAdder(a[0], gates())); }
        else if (w == "C2AddMod") {
            need(9);
            s = join(hgm::C2AddMod(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], gates()));
        } else if (w == "C2SubMod") {
            need(9);
            s = join(hgm::C2SubMod(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], gates()));
        } else if (w == "CMultMod") {
            need(9);
            s = join(hgm::CMultMod(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], gates()));
        } else if (w == "CRMultMod") {
            need(9);
            s = join(hgm::CRMultMod(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], gates()));
        }
        else throw hgm::Error(HGM_E_ARG, "unknown generator '" + w + "'");
        *len = s.size();
        if (out && cap > s.size()) std::memcpy(out, s.c_str(), s.size() + 1);
        return HGM_OK;
    });
}

}  // extern "C"

// ------------------------------------------------------------------ reference ABI

namespace {

[[noreturn]] void ref_die(const std::string &msg) {
    std::printf("error: %s\n", msg.c_str());
    std::exit(1);
}

// One PT as the reference's t_GPU kernel applies it: the pair (pos0, pos1)
// is updated when (pos0 & control_mask) == control_value, pos0 having the
// target bit clear and no bit at or above qubits (src/core/kernel.cu:102-108,
// with the control arguments of src/core/common.cpp:53-65, which zero the
// value when the mask is 0). Returns false for a gate that can never fire.
// Targets outside the state pass through for the planner to reject.
bool ref_op(const hgm_ref_pt &p, int qubits, hgm_op *o) {
    std::memset(o, 0, sizeof(*o));
    o->target_bit = p.target_bit;
    uint64_t mask = (uint64_t)p.control_mask, value = mask ? (uint64_t)p.control_value : 0;
    if (p.target_bit >= 0 && p.target_bit < qubits && qubits < 64) {
        if (value & ~mask) return false;
        const uint64_t clear = (uint64_t(1) << p.target_bit) | ~((uint64_t(1) << qubits) - 1);
        if (value & clear) return false;
        mask &= ~clear;
    }
    o->control_mask = mask;
    o->control_value = value;
    for (int k = 0; k < 4; ++k) {
        o->m[2 * k] = __real__ p.matrix[k];
        o->m[2 * k + 1] = __imag__ p.matrix[k];
    }
    return true;
}

// The region that ProjectState placed on the device, kept for
// GpuExecutionWrapper(NULL, ...) and GetState (the reference's gpu_mem,
// src/core/kernel.cu:321-435).
std::mutex g_proj_mu;
hgm_state *g_proj = nullptr;

// Checks the arguments of ProjectState / GetState and returns the region's
// bits: the lowest region_size set bits of region_mask inside the state. The
// reference copies 2^(region_size - c) portions of 2^c amplitudes, c being the
// low run of region_mask counted inside the state, at the subsets of the mask
// bits above c taken in increasing order (src/core/kernel.cu:329-357), which
// are the same amplitudes.
uint64_t ref_region(const hgm_cfloat *state, int qubits, int region_size, long region_id, long region_mask,
                    int gpu_count, const char *who) {
    const std::string w(who);
    if (gpu_count != 1) ref_die(w + ": gpu_count " + std::to_string(gpu_count) + ": this library drives one GPU per process");
    if (!state || qubits < 1 || qubits > 40 || region_size < 1 || region_size > qubits)
        ref_die(w + ": bad state, qubits or region_size");
    const uint64_t all = (uint64_t(1) << qubits) - 1;
    uint64_t bits = (uint64_t)region_mask & all, mask = 0;
    for (int k = 0; k < region_size && bits; ++k) {
        mask |= bits & (~bits + 1);
        bits &= bits - 1;
    }
    if (__builtin_popcountll(mask) != region_size)
        ref_die(w + ": region_mask has fewer than region_size bits inside the state");
    if (((uint64_t)region_id & ~all) || ((uint64_t)region_id & mask))
        ref_die(w + ": region_id outside the state or on a region bit");
    return mask;
}

}  // namespace

extern "C" {

bool setDevice(int device_id) { return hgm_init(device_id) != HGM_OK; }

int hgm_auto_tuning(int qubits, int dtype, hgm_tuning *out) {
    if (!out || qubits < 1 || qubits > hgm::kMaxQubits || (dtype != HGM_C128 && dtype != HGM_C64)) {
        hgm::set_error("bad argument");
        return HGM_E_ARG;
    }
    hgm_default_tuning(out);
    const bool f32 = dtype == HGM_C64;
    out->block_size = f32 ? 512 : 256;
    out->repeat_count = 8;
    out->gpu_region_bits = f32 ? 13 : 12;
    out->gpu_coalesced_bits = 4;
    out->kernel = HGM_KERNEL_JIT;
    out->planner = HGM_PLAN_GREEDY | HGM_PLAN_FUSE | hgm::kAutoRowFlag | HGM_PLAN_BALANCE;
    return HGM_OK;
}

hgm_cfloat *GpuExecutionWrapper(hgm_cfloat *read_memory, hgm_ref_pt **pts, int qubits, int coalesced_bits,
                                int gpu_region_bits, int gpu_count, int block_size, int repeat_count,
                                int iterations) {
    static const bool debug = std::getenv("HYBRID_DEBUG") != nullptr;
    auto t0 = std::chrono::steady_clock::now();
    if (gpu_count != 1) ref_die("gpu_count " + std::to_string(gpu_count) + ": this library drives one GPU per process");
    std::vector<hgm_op> ops;
    for (size_t i = 0; pts && pts[i]; ++i) {
        hgm_op o;
        if (ref_op(*pts[i], qubits, &o)) ops.push_back(o);
    }
    hgm_tuning t;
    hgm_default_tuning(&t);
    t.block_size = block_size;
    t.repeat_count = repeat_count;
    t.gpu_region_bits = gpu_region_bits;
    t.gpu_coalesced_bits = coalesced_bits;
    if (const char *e = std::getenv("HGM_OPS_BLOCK")) t.ops_block = std::atoi(e);
    t = hgm::env_tuning(t, qubits, HGM_C64, ops.data(), (int64_t)ops.size(), iterations);
    hgm_stats st;
    std::memset(&st, 0, sizeof(st));
    int rc;
    if (read_memory) {
        rc = hgm_execute(read_memory, qubits, HGM_C64, ops.data(), (int64_t)ops.size(), &t, iterations, &st);
    } else {
        std::lock_guard<std::mutex> g(g_proj_mu);
        if (!g_proj) ref_die("GpuExecutionWrapper(NULL, ...): no state was projected (ProjectState)");
        if (hgm_state_qubits(g_proj) != qubits)
            ref_die("GpuExecutionWrapper(NULL, ...): qubits " + std::to_string(qubits) + " but the projected region has " +
                    std::to_string(hgm_state_qubits(g_proj)));
        rc = hgm_state_apply(g_proj, ops.data(), (int64_t)ops.size(), &t, iterations, &st);
    }
    if (rc) {
        std::printf("error: %d - %s\n", rc, hgm_last_error());
        std::exit(1);
    }
    if (debug)
        std::fprintf(stderr, "[GPUEXEC]   qubits=%d elapsed_ms=%.3f passes=%lld kernel_ms=%.3f\n", qubits,
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(),
                     (long long)st.passes, st.kernel_ms);
    return read_memory;
}

bool ProjectState(hgm_cfloat *state, int qubits, int region_size, long region_id, long region_mask, int gpu_count) {
    static const bool debug = std::getenv("HYBRID_DEBUG") != nullptr;
    auto t0 = std::chrono::steady_clock::now();
    const uint64_t mask = ref_region(state, qubits, region_size, region_id, region_mask, gpu_count, "ProjectState");
    std::lock_guard<std::mutex> g(g_proj_mu);
    if (g_proj) {   // a region never collected by GetState: the reference overwrites gpu_mem
        hgm_state_destroy(g_proj);
        g_proj = nullptr;
    }
    int rc = hgm_state_create(region_size, HGM_C64, &g_proj);
    auto t1 = std::chrono::steady_clock::now();
    if (!rc) rc = hgm_state_gather(g_proj, state, qubits, (uint64_t)region_id, mask);
    if (rc) {
        std::printf("error: %d - %s\n", rc, hgm_last_error());
        std::exit(1);
    }
    if (debug) {   // the reference's fields (src/core/kernel.cu:374-377); the portions move in one pipelined copy
        const auto t2 = std::chrono::steady_clock::now();
        const int c = __builtin_ctzll(~mask);
        std::fprintf(stderr,
                     "[PROJSTATE] region_size=%d coalesced_bits=%d mem_portions=%ld portion_size=%ld malloc_ms=%.3f "
                     "memcpy_ms=%.3f elapsed_ms=%.3f\n",
                     region_size, c, 1L << (region_size - c), 1L << c,
                     std::chrono::duration<double, std::milli>(t1 - t0).count(),
                     std::chrono::duration<double, std::milli>(t2 - t1).count(),
                     std::chrono::duration<double, std::milli>(t2 - t0).count());
    }
    return true;
}

bool GetState(hgm_cfloat *state, int qubits, int region_size, long region_id, long region_mask, int gpu_count) {
    static const bool debug = std::getenv("HYBRID_DEBUG") != nullptr;
    auto t0 = std::chrono::steady_clock::now();
    const uint64_t mask = ref_region(state, qubits, region_size, region_id, region_mask, gpu_count, "GetState");
    std::lock_guard<std::mutex> g(g_proj_mu);
    if (!g_proj || hgm_state_qubits(g_proj) != region_size)
        ref_die("GetState: no projected region of " + std::to_string(region_size) + " qubits (ProjectState)");
    int rc = hgm_state_scatter(g_proj, state, qubits, (uint64_t)region_id, mask);
    auto t1 = std::chrono::steady_clock::now();
    hgm_state_destroy(g_proj);
    g_proj = nullptr;
    if (rc) {
        std::printf("error: %d - %s\n", rc, hgm_last_error());
        std::exit(1);
    }
    if (debug) {   // the reference's fields (src/core/kernel.cu:430-433)
        const auto t2 = std::chrono::steady_clock::now();
        const int c = __builtin_ctzll(~mask);
        std::fprintf(stderr,
                     "[GETSTATE]  region_size=%d coalesced_bits=%d mem_portions=%ld portion_size=%ld memcpy_ms=%.3f "
                     "free_ms=%.3f elapsed_ms=%.3f\n",
                     region_size, c, 1L << (region_size - c), 1L << c,
                     std::chrono::duration<double, std::milli>(t1 - t0).count(),
                     std::chrono::duration<double, std::milli>(t2 - t1).count(),
                     std::chrono::duration<double, std::milli>(t2 - t0).count());
    }
    return true;
}

}  // extern "C"
