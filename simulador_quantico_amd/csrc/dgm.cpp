// Executor of the t_GPU path: the reference's DGM API (include/core/dgm.h)
// over the HIP library. Parsing follows src/core/dgm_parser.cpp, the catalog
// src/core/gates.cpp, validation and dispatch src/core/dgm_core.cpp.
#include "hgm_dgm.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "runtime.hpp"

namespace hgm {

std::mt19937 g_rng;
void seed_rng(unsigned long seed) { g_rng.seed(seed); }

// ------------------------------------------------------------------ PT / Gates
long PT::matrixType() const {
    if (matrix[1] == 0.0 && matrix[2] == 0.0) return DIAG_PRI;
    if (matrix[0] == 0.0 && matrix[3] == 0.0) return DIAG_SEC;
    return DENSE;
}

Gates::Gates() { init(); }

void Gates::init() {
    const double s = 1.0 / std::sqrt(2.0);
    addGate("H", s, s, s, -s);
    addGate("X", 0.0, 1.0, 1.0, 0.0);
    addGate("Z", 1.0, 0.0, 0.0, -1.0);
    addGate("Y", 0.0, cplx(0, 1), cplx(0, -1), 0.0);
    // R_k = diag(1, e^(i*pi/2^(k-1))), generated in double
    addGate("R1", 1.0, 0.0, 0.0, std::polar(1.0, M_PI));
    addGate("R2", 1.0, 0.0, 0.0, std::polar(1.0, M_PI / 2.0));
    addGate("R3", 1.0, 0.0, 0.0, std::polar(1.0, M_PI / 4.0));
}

const cplx *Gates::getMatrix(const std::string &gate_name) const {
    auto it = list.find(gate_name);
    return it == list.end() ? nullptr : it->second.data();
}

bool Gates::addGate(const std::string &name, const cplx *m) { return addGate(name, m[0], m[1], m[2], m[3]); }

bool Gates::addGate(const std::string &name, cplx a0, cplx a1, cplx a2, cplx a3) {
    if (list.count(name)) return false;
    list[name] = {a0, a1, a2, a3};
    return true;
}

// ------------------------------------------------------------------ DGMCore
DGMCore::~DGMCore() { erase(); }

void DGMCore::fail(int code, const std::string &msg) const {
    if (exit_on_error) {
        std::printf("%s\n", msg.c_str());
        std::fflush(stdout);
        std::exit(1);
    }
    throw Error(code, msg);
}

void DGMCore::erase() {
    for (PT *p : vec_pts) delete p;
    vec_pts.clear();
    pts = nullptr;
}

void DGMCore::setCpuStructure(long region_bits, long coalesced_bits) {
    cpu_region_bits = region_bits;
    cpu_coalesced_bits = coalesced_bits;
}

// The reference computes block_size = 1 << r/2/R, which C parses as
// 1 << (r/2/R) (src/core/dgm_core.cpp:317); the intent -- and what the tile
// invariant 2*block_size*repeat_count == 2^r needs -- is (1 << r)/2/R.
void DGMCore::setGpuStructure(long coalesced_bits, long region_bits, int rc) {
    gpu_coalesced_bits = coalesced_bits;
    gpu_region_bits = region_bits;
    repeat_count = rc;
    block_size = (int)((1L << region_bits) / 2 / rc);
}

namespace {

std::vector<std::string> tokenize(const std::string &text, const std::string &delims) {
    std::vector<std::string> out;
    std::string::size_type last = text.find_first_not_of(delims, 0);
    std::string::size_type pos = text.find_first_of(delims, last);
    while (pos != std::string::npos || last != std::string::npos) {
        out.push_back(text.substr(last, pos - last));
        last = text.find_first_not_of(delims, pos);
        pos = text.find_first_of(delims, last);
    }
    return out;
}

}  // namespace

void DGMCore::setFunction(const std::string &function, int iterations, bool reset) {
    setFunction(tokenize(function, ";"), iterations, reset);
}

// dgm_parser.cpp:42-66: per step, PTs sorted by target bit, ascending for even
// step indices and descending for odd ones (parity restarts each iteration);
// the whole step list is replicated `iterations` times; reset=false appends.
void DGMCore::setFunction(const std::vector<std::string> &steps, int iterations, bool reset) {
    if (reset) erase();
    else if (!vec_pts.empty() && vec_pts.back() == nullptr) vec_pts.pop_back();
    std::vector<PT *> step_pts;
    for (long it = 0; it < iterations; ++it)
        for (size_t s = 0; s < steps.size(); ++s) {
            std::map<long, Group> groups = genGroups(steps[s]);
            genPTs(groups, step_pts);
            if (s % 2)
                std::sort(step_pts.begin(), step_pts.end(),
                          [](const PT *a, const PT *b) { return a->span_start_bit > b->span_start_bit; });
            else
                std::sort(step_pts.begin(), step_pts.end(),
                          [](const PT *a, const PT *b) { return a->span_start_bit < b->span_start_bit; });
            vec_pts.insert(vec_pts.end(), step_pts.begin(), step_pts.end());
        }
    vec_pts.push_back(nullptr);
    pts = vec_pts.data();
}

// dgm_parser.cpp:70-114. "ControlN(v)" adds a control of value v (any nonzero
// value counts as 1) to group N; "TargetN(g)" adds gate g as a target of group
// N; any other token except "ID" is an uncontrolled gate of group 0. As in the
// reference, the step's token count becomes the executor's qubit count.
std::map<long, Group> DGMCore::genGroups(const std::string &step) {
    std::vector<std::string> toks = tokenize(step, ",");
    qubits = (long)toks.size();
    std::map<long, Group> groups;
    long pos = 0;
    for (const std::string &tok : toks) {
        const size_t ck = tok.find("Control"), tk = tok.find("Target");
        const size_t paren = tok.find("(") + 1;
        char *end = nullptr;
        if (ck != std::string::npos) {
            long g = std::strtol(tok.c_str() + 7, &end, 10);
            long v = std::strtol(tok.c_str() + paren, &end, 10);
            groups[g].ctrl.push_back(v);
            groups[g].pos_ctrl.push_back(pos);
        } else if (tk != std::string::npos) {
            long g = std::strtol(tok.c_str() + 6, &end, 10);
            groups[g].ops.push_back(tok.substr(paren, tok.size() - paren - 1));
            groups[g].pos_ops.push_back(pos);
        } else if (tok != "ID") {
            groups[0].ops.push_back(tok);
            groups[0].pos_ops.push_back(pos);
        }
        ++pos;
    }
    return groups;
}

// dgm_parser.cpp:118-169, with 64-bit masks and an error for an unknown gate
// name (the reference stores a NULL matrix and crashes later, gates.cpp:37).
void DGMCore::genPTs(std::map<long, Group> &groups, std::vector<PT *> &step_pts) {
    step_pts.clear();
    for (auto &kv : groups) {
        Group &g = kv.second;
        long cmask = 0, cval = 0;
        std::vector<long> cbits;
        for (size_t i = 0; i < g.ctrl.size(); ++i) {
            long bit = qubits - g.pos_ctrl[i] - 1;
            cbits.push_back(bit);
            cmask |= 1L << bit;
            if (g.ctrl[i]) cval |= 1L << bit;
        }
        for (size_t i = 0; i < g.ops.size(); ++i) {
            const cplx *m = gates.getMatrix(g.ops[i]);
            if (!m) {
                for (PT *p : step_pts) delete p;
                step_pts.clear();
                fail(HGM_E_PARSE, "unknown gate '" + g.ops[i] + "'");
            }
            PT *t = new PT();
            t->qubits = 1;
            t->span_start_bit = (int)(qubits - g.pos_ops[i]);
            t->target_bit = t->span_start_bit - 1;
            t->matrix_size = 2;
            t->matrix = m;
            t->control_value = cval;
            t->control_mask = cmask;
            t->control_count = (long)g.ctrl.size();
            t->control_bit_positions = cbits;
            step_pts.push_back(t);
        }
    }
}

// Debug printing with the reference's field order and formats
// (src/core/common.cpp:68-85, src/core/dgm_core.cpp:35-39), so the output of
// the two builds can be compared line by line.
void PT::print() const {
    std::printf("qubits: %d\nmatrix_size: %d\nspan_start_bit: %d\ntarget_bit: %d\nAffect: %d\n", qubits, matrix_size,
                span_start_bit, target_bit, (int)affected);
    std::printf("Ctrl-(value: %ld, mask: %ld, count: %ld)\n", control_value, control_mask, control_count);
    for (long i = 0; i < control_count && i < (long)control_bit_positions.size(); ++i)
        std::printf("%d: %ld\n", (int)i, control_bit_positions[(size_t)i]);
}

void PT::printMatrix() const {
    for (int row = 0; row < matrix_size; ++row) {
        for (int col = 0; col < matrix_size; ++col) {
            const cplx v = matrix ? matrix[row * matrix_size + col] : cplx(0, 0);
            // the reference's matrices are float complex: print the same rounding
            std::printf("%d: %.4f, %.4f  \t", row * matrix_size + col, (double)(float)v.real(), (double)(float)v.imag());
        }
        std::printf("\n");
    }
}

void DGMCore::printPTs() const {
    for (size_t i = 0; pts && pts[i]; ++i) pts[i]->print();
    std::fflush(stdout);
}

void DGMCore::CountOps(int iterations) {
    dense = main_diag = sec_diag = c_dense = c_main_diag = c_sec_diag = 0;
    for (size_t i = 0; pts && pts[i]; ++i) {
        const bool ctl = pts[i]->control_mask != 0;
        switch (pts[i]->matrixType()) {
            case DENSE: (ctl ? c_dense : dense)++; break;
            case DIAG_PRI: (ctl ? c_main_diag : main_diag)++; break;
            default: (ctl ? c_sec_diag : sec_diag)++; break;
        }
    }
    dense *= iterations; c_dense *= iterations; main_diag *= iterations;
    c_main_diag *= iterations; sec_diag *= iterations; c_sec_diag *= iterations;
    total_op = dense + c_dense + main_diag + c_main_diag + sec_diag + c_sec_diag;
}

hgm_tuning DGMCore::tuning() const {
    hgm_tuning t;
    hgm_default_tuning(&t);
    t.block_size = block_size;
    t.repeat_count = repeat_count;
    t.gpu_region_bits = (int)gpu_region_bits;
    t.gpu_coalesced_bits = (int)gpu_coalesced_bits;
    t.ops_block = ops_block;
    t.planner = planner;
    t.kernel = kernel;
    return t;
}

// src/core/dgm_core.cpp:177-241. The CPU region clamp is kept; the GPU tile
// invariant is checked on the requested gpu_region_bits before any clamp, and
// states smaller than a tile run as one tile (the reference rejects them with
// the default tuning, since it clamps first). gpu_count must be 1: the
// multi-GPU peer mode is not part of this library.
void DGMCore::validateTuning() {
    if (cpu_region_bits > qubits) cpu_region_bits = qubits;
    if (exec_type != t_GPU && exec_type != t_HYBRID) return;
    if (gpu_count != 1)
        fail(HGM_E_UNSUPPORTED, "Erro de tuning: gpu_count (" + std::to_string(gpu_count) +
                                    ") must be 1 (one GPU per process; the multi-GPU peer mode is not provided)");
    hgm_tuning t = tuning();
    if (int rc = hgm_validate_tuning((int)(qubits > 0 ? qubits : 1), amp_dtype(), &t)) fail(rc, hgm_last_error());
}

std::vector<hgm_op> DGMCore::ops() const {
    std::vector<hgm_op> out;
    for (size_t i = 0; pts && pts[i]; ++i) {
        const PT *p = pts[i];
        hgm_op o;
        std::memset(&o, 0, sizeof(o));
        o.target_bit = p->target_bit;
        o.control_mask = (uint64_t)p->control_mask;
        o.control_value = (uint64_t)p->control_value;
        for (int k = 0; k < 4; ++k) { o.m[2 * k] = p->matrix[k].real(); o.m[2 * k + 1] = p->matrix[k].imag(); }
        out.push_back(o);
    }
    return out;
}

// ------------------------------------------------------------------ BasicDGM
template <class Real>
BasicDGM<Real>::~BasicDGM() {
    freeMemory();
}

template <class Real>
void BasicDGM<Real>::allocateMemory() {
    freeMemory();
    if (qubits < 1 || qubits > QB_LIMIT) fail(HGM_E_ARG, "Invalid qubit count: " + std::to_string(qubits));
    const size_t bytes = (size_t(1) << qubits) * sizeof(amp_t);
    void *p = nullptr;
    int count = 0;
    hgm_device_count(&count);
    if (count > 0 && alloc_pinned(&p, bytes) == HGM_OK) {
        pinned_ = true;
    } else {
        p = std::calloc(size_t(1) << qubits, sizeof(amp_t));
        pinned_ = false;
        if (!p) fail(HGM_E_NOMEM, "cannot allocate a " + std::to_string(qubits) + "-qubit host state");
    }
    state = static_cast<amp_t *>(p);
    state_qubits = qubits;
    host_valid_ = true;
    dev_valid_ = false;
}

template <class Real>
void BasicDGM<Real>::freeMemory() {
    if (dev_) { hgm_state_destroy(dev_); dev_ = nullptr; }
    if (state) {
        if (pinned_) free_pinned(state);
        else std::free(state);
    }
    state = nullptr;
    dev_valid_ = false;
    host_valid_ = true;
}

template <class Real>
void BasicDGM<Real>::setMemoryValue(long pos) {
    syncHost();
    state[pos] = 1;
    dev_valid_ = false;
}

template <class Real>
void BasicDGM<Real>::ensure_device() {
    if (dev_ && hgm_state_qubits(dev_) != state_qubits) { hgm_state_destroy(dev_); dev_ = nullptr; }
    if (!dev_) {
        if (int rc = hgm_state_create((int)state_qubits, dtype(), &dev_)) fail(rc, hgm_last_error());
        dev_valid_ = false;
    }
    if (!dev_valid_) {
        if (int rc = hgm_state_upload(dev_, state)) fail(rc, hgm_last_error());
        dev_valid_ = true;
    }
}

template <class Real>
typename BasicDGM<Real>::amp_t *BasicDGM<Real>::hostState() {
    syncHost();
    // the caller may write through the pointer: the host copy becomes the
    // one the next device use uploads
    if (resident) dev_valid_ = false;
    return state;
}

template <class Real>
void BasicDGM<Real>::syncHost() {
    if (host_valid_ || !dev_ || !dev_valid_) { host_valid_ = true; return; }
    if (int rc = hgm_state_download(dev_, state)) fail(rc, hgm_last_error());
    host_valid_ = true;
}

template <class Real>
typename BasicDGM<Real>::amp_t *BasicDGM<Real>::execute(int iterations) {
    validateTuning();
    if (exec_type < t_CPU || exec_type > t_HYBRID) fail(HGM_E_ARG, "Erro exec type");
    if (exec_type == t_CPU || exec_type == t_PAR_CPU)
        fail(HGM_E_UNSUPPORTED, "exec_type " + std::to_string(exec_type) +
                                    " is not provided by this library: only t_GPU (2) and t_HYBRID (3) run here");
    // t_HYBRID: the GPU takes every region, the reference's own schedule when
    // thread_count is 1, and the PT list runs once whatever iterations says,
    // as DGM::execute calls HybridExecution(pts) (src/core/dgm_core.cpp:262-264)
    if (exec_type == t_HYBRID) iterations = 1;
    if (!state) fail(HGM_E_STATE, "execute() before allocateMemory()");
    if (qubits != state_qubits)
        fail(HGM_E_STATE, "the circuit has " + std::to_string(qubits) + " qubits but the state was allocated for " +
                              std::to_string(state_qubits));
    std::vector<hgm_op> o = ops();
    const hgm_tuning t = env_tuning(tuning(), (int)qubits, dtype(), o.data(), (int64_t)o.size(), iterations);
    std::memset(&last_stats, 0, sizeof(last_stats));
    if (resident) {
        ensure_device();
        if (int rc = hgm_state_apply(dev_, o.data(), (int64_t)o.size(), &t, iterations, &last_stats))
            fail(rc, hgm_last_error());
        host_valid_ = false;
        return nullptr;   // the result stays on the device: syncHost() or hostState() reads it
    } else {
        if (dev_) { hgm_state_destroy(dev_); dev_ = nullptr; dev_valid_ = false; }
        if (int rc = hgm_execute(state, (int)qubits, dtype(), o.data(), (int64_t)o.size(), &t, iterations, &last_stats))
            fail(rc, hgm_last_error());
    }
    return state;
}

template <class Real>
void BasicDGM<Real>::executeFunction(const std::string &function, int iterations) {
    if (function.empty()) return;
    setFunction(function);
    execute(iterations);
}

template <class Real>
void BasicDGM<Real>::executeFunction(const std::vector<std::string> &steps, int iterations) {
    setFunction(steps);
    execute(iterations);
}

// dgm_core.cpp:69-125: probabilities of the two halves, one uniform draw in
// [0,1) from g_rng, outcome 0 iff P(0) > draw, then the kept half moves to
// the bit=0 slot divided by sqrt(P) and the bit=1 slot is zeroed (the qubit is
// reset to |0>). As in the reference (:75-99), P(0) and P(1) are kept in the
// amplitude type, each |a|^2 formed in double and added in index order, and
// the double draw is rounded to float before the test `zero > draw`; sqrt(P)
// and the division are taken in the amplitude type (:106-109,118). The FP32
// executor's host path therefore repeats the reference's arithmetic. On a
// device-resident state both halves are summed on the device in FP64 (each
// directly, neither as the total minus the other) and then rounded to the
// amplitude type: the summation order differs from the reference's.
template <class Real>
int BasicDGM<Real>::measure(int qubit_pos) {
    if (qubit_pos < 0 || qubit_pos >= qubits) fail(HGM_E_ARG, "measure: qubit out of range");
    Real zero = 0, one = 0;
    const bool on_device = resident && dev_ && dev_valid_ && !host_valid_;
    const int shift = (int)(qubits - 1 - qubit_pos);
    if (on_device) {
        double z = 0, o = 0;
        if (int rc = hgm_state_prob_halves(dev_, qubit_pos, &z, &o)) fail(rc, hgm_last_error());
        zero = (Real)z;
        one = (Real)o;
    } else {
        syncHost();
        const size_t n = size_t(1) << qubits;
        for (size_t i = 0; i < n; ++i) {
            const double re = state[i].real(), im = state[i].imag();
            const double p = re * re + im * im;
            if ((i >> shift) & 1) one = (Real)((double)one + p); else zero = (Real)((double)zero + p);
        }
    }
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    const float r = (float)dist(g_rng);
    const int bit = (zero > (Real)r) ? 0 : 1;
    const Real kept = bit ? one : zero;
    measure_value = (double)kept;
    const Real norm = (Real)std::sqrt((double)kept);
    if (on_device) {
        // the kernel divides in double by sqrt(p). For float amplitudes
        // p = norm^2 is exact in double, so it divides by `norm` itself, and
        // a float divided in double and rounded to float is the float
        // quotient; for double amplitudes sqrt(kept) is `norm`
        const double p = sizeof(Real) == 4 ? (double)norm * (double)norm : (double)kept;
        if (int rc = hgm_state_collapse_reset(dev_, qubit_pos, bit, p)) fail(rc, hgm_last_error());
    } else {
        const size_t half = size_t(1) << (qubits - 1), low = (size_t(1) << shift) - 1;
        for (size_t i = 0; i < half; ++i) {
            const size_t pos = (i << 1) - (i & low);
            const amp_t v = state[pos | (size_t(bit) << shift)];
            state[pos] = amp_t(v.real() / norm, v.imag() / norm);
            state[pos | (size_t(1) << shift)] = amp_t(0);
        }
        dev_valid_ = false;
    }
    return bit;
}

// dgm_core.cpp:150-162: marginal probabilities keyed by the masked index bits.
template <class Real>
std::map<long, double> BasicDGM<Real>::measure(const std::vector<int> &qubit_positions) {
    syncHost();
    long mask = 0;
    for (int q : qubit_positions) mask |= 1L << (qubits - 1 - q);
    // sums kept in the amplitude type, each |a|^2 formed and added in double,
    // as the reference's map<long, float> does (src/core/dgm_core.cpp:158): the
    // FP32 executor returns the reference's values exactly
    std::map<long, Real> sums;
    const size_t n = size_t(1) << qubits;
    for (size_t i = 0; i < n; ++i) {
        const double re = state[i].real(), im = state[i].imag();
        Real &acc = sums[(long)i & mask];
        acc = (Real)((double)acc + (re * re + im * im));
    }
    std::map<long, double> probs;
    for (const auto &kv : sums) probs[kv.first] = (double)kv.second;
    return probs;
}

// dgm_core.cpp:129-146: force qubit_pos to value and renormalise.
template <class Real>
void BasicDGM<Real>::colapse(int qubit_pos, int value) {
    syncHost();
    const int shift = (int)(qubits - 1 - qubit_pos);
    const size_t n = size_t(1) << qubits;
    // The sum is kept in the amplitude type, each |a|^2 formed in double and
    // added in double, and printed, as src/core/dgm_core.cpp:133-139 does for
    // its float state; the kept amplitudes are divided by the square root in
    // that type (:141-145), so the FP32 executor reproduces the reference.
    Real sum = 0;
    for (size_t i = 0; i < n; ++i)
        if ((int)((i >> shift) & 1) == value) {
            const double re = state[i].real(), im = state[i].imag();
            sum = (Real)((double)sum + (re * re + im * im));
        }
    std::printf("%g\n", (double)(float)sum);
    std::fflush(stdout);
    const Real norm = (Real)std::sqrt((double)sum);
    for (size_t i = 0; i < n; ++i)
        state[i] = ((int)((i >> shift) & 1) == value) ? amp_t(state[i].real() / norm, state[i].imag() / norm) : amp_t(0);
    dev_valid_ = false;
}

template class BasicDGM<double>;
template class BasicDGM<float>;

}  // namespace hgm
