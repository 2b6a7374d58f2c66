// ORACLE -- test infrastructure only. Nothing in the shipped library links,
// loads or calls this file; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it, and only as the checker.
//
// A CPU restatement of HybriD-GM's circuit semantics, written from the
// reference's behaviour (file:line citations below), templated on the
// amplitude precision (double or float):
//   * circuit language: Tokenize / genGroups / genPTs / per-step sort /
//     iteration replication / reset=false append      (src/core/dgm_parser.cpp)
//   * gate catalog with first-writer-wins addGate      (src/core/gates.cpp:23-60)
//   * t_CPU engine: dense / main-diagonal / anti-diagonal loops with the
//     controlled subset walk (pos+inc)&mask            (src/core/dgm_cpu_exec.cpp)
//   * t_PAR_CPU engine: region planner + dynamically claimed regions +
//     per-region op loops                               (src/core/dgm_par_exec.cpp:38-253)
//   * measure(qubit) with collapse-and-reset-to-|0>, measure(vector) marginals
//                                                       (src/core/dgm_core.cpp:69-162)
//
// Pinning (tests/test_oracle.py): the reference's test_gates truth tables,
// the closed forms of the Hadamard layer and of Grover's iterate, a numpy
// Kronecker-product brute force for n <= 8, and golden vectors written by the
// reference itself (tests/golden/, produced through oracle/_ref).
//
// Two catalog modes: CAT_DOUBLE builds R1..R3 as exp(i*pi/2^(k-1)) in double
// (the precision contract of the HIP library); CAT_REFERENCE builds them the
// way gates.cpp:31-33 does, with cpowf in single precision, so that an FP32
// oracle run can be compared bit for bit with the reference's own FP32 output.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex.h>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <vector>
#include <omp.h>

namespace {

enum { KIND_DENSE = 0, KIND_DIAG = 1, KIND_ANTI = 2 };   // common.h:48-52
enum { CAT_DOUBLE = 0, CAT_REFERENCE = 1 };
enum { ENGINE_TCPU = 0, ENGINE_TPARCPU = 1 };

using cd = std::complex<double>;

struct Gate {            // the reference's PT (include/core/common.h:57-90), the fields that matter
    int target_bit;
    int64_t control_mask;
    int64_t control_value;
    int64_t control_count;
    std::string name;    // resolved against the catalog at execute time
};

// Tokenize (dgm_parser.cpp:10-26): split on any delimiter, skipping runs of them.
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

struct GroupAcc {
    std::vector<std::string> ops;
    std::vector<long> op_pos;
    std::vector<long> ctrl_val;
    std::vector<long> ctrl_pos;
};

template <class R>
struct Oracle {
    using C = std::complex<R>;
    int qubits = 0;
    int state_qubits = 0;        // qubits the state was allocated with
    std::vector<C> state;
    std::map<std::string, std::array<cd, 4>> catalog;
    std::vector<Gate> gates;     // the compiled PT list (without the NULL terminator)
    std::mt19937 rng;            // common.cpp:10 (default seed 5489 until seeded)
    std::string error;

    explicit Oracle(int q, int catalog_mode) : qubits(q), state_qubits(q) {
        state.assign(size_t(1) << q, C(0));
        init_catalog(catalog_mode);
    }

    bool add_gate(const std::string &name, const cd m[4]) {      // gates.cpp:49-60
        if (catalog.count(name)) return false;
        catalog[name] = {m[0], m[1], m[2], m[3]};
        return true;
    }

    void init_catalog(int mode) {                                   // gates.cpp:23-34
        const double s = 1.0 / std::sqrt(2.0);
        cd h[4] = {s, s, s, -s};
        cd x[4] = {0.0, 1.0, 1.0, 0.0};
        cd z[4] = {1.0, 0.0, 0.0, -1.0};
        cd y[4] = {0.0, cd(0, 1), cd(0, -1), 0.0};
        add_gate("H", h); add_gate("X", x); add_gate("Z", z); add_gate("Y", y);
        // gates.cpp:31-33 call cpowf with literal arguments, which the compiler
        // folds at build time (even at -O0); the same literal calls here give
        // the same single-precision values.
        float _Complex ref_ph[3] = {cpowf(M_E, M_PI * I), cpowf(M_E, M_PI * I / 2.0), cpowf(M_E, M_PI * I / 4.0)};
        for (int k = 1; k <= 3; ++k) {
            cd ph;
            if (mode == CAT_REFERENCE)
                ph = cd(crealf(ref_ph[k - 1]), cimagf(ref_ph[k - 1]));
            else
                ph = std::polar(1.0, M_PI / std::pow(2.0, k - 1));
            cd r[4] = {1.0, 0.0, 0.0, ph};
            add_gate("R" + std::to_string(k), r);
        }
    }

    // genGroups (dgm_parser.cpp:70-114). Note the side effect: qubits := token count.
    std::map<long, GroupAcc> gen_groups(const std::string &step) {
        std::vector<std::string> toks = tokenize(step, ",");
        qubits = (int)toks.size();
        std::map<long, GroupAcc> groups;
        long qpos = 0;
        for (const std::string &tok : toks) {
            size_t ck = tok.find("Control"), tk = tok.find("Target");
            size_t paren = tok.find("(") + 1;
            char *end = nullptr;
            if (ck != std::string::npos) {
                long gnum = strtol(tok.c_str() + 7, &end, 10);
                long val = strtol(tok.c_str() + paren, &end, 10);
                groups[gnum].ctrl_val.push_back(val);
                groups[gnum].ctrl_pos.push_back(qpos);
            } else if (tk != std::string::npos) {
                long gnum = strtol(tok.c_str() + 6, &end, 10);
                std::string name = tok.substr(paren, tok.size() - paren - 1);
                groups[gnum].ops.push_back(name);
                groups[gnum].op_pos.push_back(qpos);
            } else if (tok != "ID") {
                groups[0].ops.push_back(tok);
                groups[0].op_pos.push_back(qpos);
            }
            ++qpos;
        }
        return groups;
    }

    // genPTs (dgm_parser.cpp:118-169): one gate per target, sharing the group's control.
    void gen_gates(std::map<long, GroupAcc> &groups, std::vector<Gate> &out) {
        out.clear();
        for (auto &kv : groups) {
            GroupAcc &g = kv.second;
            int64_t cmask = 0, cval = 0;
            for (size_t i = 0; i < g.ctrl_pos.size(); ++i) {
                int bit = qubits - (int)g.ctrl_pos[i] - 1;
                cmask += int64_t(1) << bit;
                if (g.ctrl_val[i]) cval += int64_t(1) << bit;
            }
            for (size_t i = 0; i < g.ops.size(); ++i) {
                Gate gt;
                gt.target_bit = qubits - (int)g.op_pos[i] - 1;
                gt.control_mask = cmask;
                gt.control_value = cval;
                gt.control_count = (int64_t)g.ctrl_pos.size();
                gt.name = g.ops[i];
                out.push_back(gt);
            }
        }
    }

    // setFunction(vector) (dgm_parser.cpp:42-66): odd step index -> descending
    // target bit (the comparator named "increasing", common.cpp:166-171),
    // even step index -> ascending ("decreasing", common.cpp:173-178). The
    // parity restarts for every iteration.
    void set_function(const std::vector<std::string> &steps, int iterations, bool reset) {
        if (reset) gates.clear();
        std::vector<Gate> step_gates;
        for (int it = 0; it < iterations; ++it)
            for (size_t s = 0; s < steps.size(); ++s) {
                auto groups = gen_groups(steps[s]);
                gen_gates(groups, step_gates);
                if (s % 2)
                    std::sort(step_gates.begin(), step_gates.end(),
                              [](const Gate &a, const Gate &b) { return a.target_bit > b.target_bit; });
                else
                    std::sort(step_gates.begin(), step_gates.end(),
                              [](const Gate &a, const Gate &b) { return a.target_bit < b.target_bit; });
                gates.insert(gates.end(), step_gates.begin(), step_gates.end());
            }
    }

    bool resolve(const Gate &g, C m[4]) {
        auto it = catalog.find(g.name);
        if (it == catalog.end()) { error = "unknown gate '" + g.name + "'"; return false; }
        for (int i = 0; i < 4; ++i) m[i] = C((R)it->second[i].real(), (R)it->second[i].imag());
        return true;
    }

    static int kind_of(const C m[4]) {                 // PT::matrixType, common.cpp:33-38
        if (m[1] == C(0) && m[2] == C(0)) return KIND_DIAG;
        if (m[0] == C(0) && m[3] == C(0)) return KIND_ANTI;
        return KIND_DENSE;
    }

    // ---------------- t_CPU (dgm_cpu_exec.cpp:11-126) ----------------
    void tcpu_gate(const Gate &g, const C m[4]) {
        const int64_t size = int64_t(1) << qubits;
        const int64_t tmask = int64_t(1) << g.target_bit;
        C *st = state.data();
        int kind = kind_of(m);
        if (kind == KIND_DIAG) {
            if (!g.control_count) {
                for (int64_t p = 0; p < size; ++p) st[p] = m[((p >> g.target_bit) & 1) * 3] * st[p];
            } else {
                int64_t mask = ~g.control_mask, inc = (~mask) + 1;
                for (int64_t p = 0; p < size; p = (p + inc) & mask) {
                    int64_t p0 = p | g.control_value;
                    st[p0] = m[((p0 >> g.target_bit) & 1) * 3] * st[p0];
                }
            }
            return;
        }
        auto pair = [&](int64_t p0) {
            int64_t p1 = p0 | tmask;
            if (kind == KIND_DENSE) {
                C t = m[0] * st[p0] + m[1] * st[p1];
                st[p1] = m[2] * st[p0] + m[3] * st[p1];
                st[p0] = t;
            } else {
                C t = m[1] * st[p1];
                st[p1] = m[2] * st[p0];
                st[p0] = t;
            }
        };
        if (!g.control_count) {
            for (int64_t p = 0; p < size / 2; ++p) pair((p * 2) - (p & (tmask - 1)));
        } else {
            int64_t mask = ~(g.control_mask | tmask), inc = (~mask) + 1;
            for (int64_t p = 0; p < size; p = (p + inc) & mask) pair(p | g.control_value);
        }
    }

    bool run_tcpu(int iterations) {
        std::vector<std::array<C, 4>> mats(gates.size());
        for (size_t i = 0; i < gates.size(); ++i)
            if (!resolve(gates[i], mats[i].data())) return false;
        for (int it = 0; it < iterations; ++it)
            for (size_t i = 0; i < gates.size(); ++i) tcpu_gate(gates[i], mats[i].data());
        return true;
    }

    // ---------------- t_PAR_CPU (dgm_par_exec.cpp:38-253) ----------------
    struct Region { int64_t mask; int64_t bits; int64_t count; int64_t pos_count; size_t end; };

    // compute_region (dgm_par_exec.cpp:38-77) for a free-standing region.
    Region plan_region(size_t start, int64_t coalesced, int64_t rbits) {
        int64_t nbits = coalesced;
        int64_t mask = coalesced ? ((int64_t(1) << coalesced) - 1) : 0;
        size_t i = start;
        while (nbits < rbits && i < gates.size()) {
            if (!((mask >> gates[i].target_bit) & 1)) ++nbits;
            if (nbits <= rbits) mask |= int64_t(1) << gates[i].target_bit;
            ++i;
        }
        while (i < gates.size() && ((mask >> gates[i].target_bit) & 1)) ++i;
        for (int64_t b = 1; nbits < rbits; b <<= 1)
            if (b & ~mask) { mask |= b; ++nbits; }
        if (nbits < rbits) rbits = nbits;
        Region r;
        r.mask = mask; r.bits = rbits;
        r.count = int64_t(1) << (qubits - rbits);
        r.pos_count = int64_t(1) << (rbits - 1);
        r.end = i;
        return r;
    }

    // PCpuExecution1_0 (dgm_par_exec.cpp:138-253): ops [a,b) restricted to one region.
    void region_ops(size_t a, size_t b, int64_t pos_count, int64_t rid, int64_t rmask,
                    const std::vector<std::array<C, 4>> &mats) {
        C *st = state.data();
        for (size_t oi = a; oi < b; ++oi) {
            const Gate &g = gates[oi];
            const C *m = mats[oi].data();
            int64_t tmask = int64_t(1) << g.target_bit;
            int kind = kind_of(m);
            int64_t base = rid, cmask_region = rmask, count = pos_count;
            if (g.control_count) {
                if ((g.control_mask & rid & ~rmask) != (g.control_value & ~rmask)) continue;
                base = rid | g.control_value;
                for (int q = 0; q < qubits; ++q) {
                    int64_t bm = int64_t(1) << q;
                    if (bm & rmask & g.control_mask) { cmask_region ^= bm; count /= 2; }
                }
            }
            int64_t pmask = cmask_region & ~tmask, inc = ~pmask + 1, pos = 0;
            for (int64_t k = 0; k < count; ++k) {
                int64_t p0 = pos | base, p1 = p0 | tmask;
                pos = (pos + inc) & pmask;
                if (kind == KIND_DENSE) {
                    C t = m[2] * st[p0] + m[3] * st[p1];
                    st[p0] = m[0] * st[p0] + m[1] * st[p1];
                    st[p1] = t;
                } else if (kind == KIND_DIAG) {
                    C t = m[3] * st[p1];
                    st[p0] *= m[0];
                    st[p1] = t;
                } else {
                    C t = m[2] * st[p0];
                    st[p0] = m[1] * st[p1];
                    st[p1] = t;
                }
            }
        }
    }

    // PCpuExecution1 (dgm_par_exec.cpp:82-132). Like the reference it does
    // not loop over "iterations" (t_PAR_CPU applies the PT list once while
    // t_GPU replicates it); threads claim the next region id from a shared
    // counter.
    bool run_tparcpu(int threads, int coalesced, int region_bits) {
        std::vector<std::array<C, 4>> mats(gates.size());
        for (size_t i = 0; i < gates.size(); ++i)
            if (!resolve(gates[i], mats[i].data())) return false;
        if (region_bits > qubits) region_bits = qubits;           // validateTuning clamp (dgm_core.cpp:185)
        size_t i = 0;
        int64_t rbits = region_bits;
        while (i < gates.size()) {
            Region r = plan_region(i, coalesced, rbits);
            rbits = r.bits;
            std::atomic<int64_t> claimed{0};
            int64_t rmask = r.mask;
            int64_t nreg = r.count;
            size_t a = i, b = r.end;
            #pragma omp parallel num_threads(threads)
            {
                for (;;) {
                    int64_t k = claimed.fetch_add(1);
                    if (k >= nreg) break;
                    // k-th combination of the bits outside the region mask
                    int64_t rid = 0, src = k;
                    for (int q = 0; q < qubits && src; ++q)
                        if (!((rmask >> q) & 1)) { rid |= (src & 1) << q; src >>= 1; }
                    region_ops(a, b, r.pos_count, rid, rmask, mats);
                }
            }
            i = r.end;
        }
        return true;
    }

    // ---------------- readout (dgm_core.cpp:69-162) ----------------
    // measure(qubit) follows dgm_core.cpp:75-122: P(0) and P(1) are summed in
    // index order into accumulators of the amplitude type (the reference's
    // `float zero, one`, :75), each term formed in double (pow(float, 2.0),
    // :81,83); the double draw is rounded to float before the test
    // `zero > random_sample` (:98-99); the kept half is divided by sqrt(P)
    // taken in the amplitude type (:106-109,118). With float amplitudes every
    // step is the reference's own arithmetic.
    int measure(int qubit) {
        const int64_t size = int64_t(1) << qubits;
        const int shift = qubits - 1 - qubit;
        R one = 0, zero = 0;
        for (int64_t p = 0; p < size; ++p) {
            const double re = state[p].real(), im = state[p].imag();
            const double pr = re * re + im * im;
            if ((p >> shift) & 1) one = R(double(one) + pr); else zero = R(double(zero) + pr);
        }
        std::uniform_real_distribution<double> dist(0.0, 1.0);
        const float r = float(dist(rng));
        const int bit = (zero > R(r)) ? 0 : 1;
        const R norm = R(std::sqrt(double(bit ? one : zero)));
        const int64_t low = (int64_t(1) << shift) - 1;
        for (int64_t p = 0; p < size / 2; ++p) {
            const int64_t pos = (p << 1) - (p & low);
            const C v = state[pos | (int64_t(bit) << shift)];
            state[pos] = C(v.real() / norm, v.imag() / norm);
            state[pos | (int64_t(1) << shift)] = C(0);
        }
        return bit;
    }
};

struct Handle {
    int precision;     // 64 or 32
    Oracle<double> *d = nullptr;
    Oracle<float> *f = nullptr;
    ~Handle() { delete d; delete f; }
};

template <class F> auto with(Handle *h, F fn) {
    return h->precision == 64 ? fn(*h->d) : fn(*h->f);
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

void *orc_new(int qubits, int precision, int catalog_mode) {
    if (qubits < 1 || qubits > 34 || (precision != 64 && precision != 32)) return nullptr;
    Handle *h = new Handle();
    h->precision = precision;
    if (precision == 64) h->d = new Oracle<double>(qubits, catalog_mode);
    else h->f = new Oracle<float>(qubits, catalog_mode);
    return h;
}

void orc_free(void *p) { delete static_cast<Handle *>(p); }

int orc_add_gate(void *p, const char *name, const double *m8) {
    cd m[4] = {{m8[0], m8[1]}, {m8[2], m8[3]}, {m8[4], m8[5]}, {m8[6], m8[7]}};
    return with(static_cast<Handle *>(p), [&](auto &o) { return o.add_gate(name, m) ? 1 : 0; });
}

int orc_get_gate(void *p, const char *name, double *m8) {
    return with(static_cast<Handle *>(p), [&](auto &o) {
        auto it = o.catalog.find(name);
        if (it == o.catalog.end()) return 0;
        for (int i = 0; i < 4; ++i) { m8[2 * i] = it->second[i].real(); m8[2 * i + 1] = it->second[i].imag(); }
        return 1;
    });
}

// function: steps separated by ';' (setFunction(string), dgm_parser.cpp:29-35)
int orc_set_function(void *p, const char *function, int iterations, int reset) {
    std::vector<std::string> steps = tokenize(function, ";");
    with(static_cast<Handle *>(p), [&](auto &o) { o.set_function(steps, iterations, reset != 0); return 0; });
    return 0;
}

int orc_set_function_steps(void *p, const char *const *steps, int n, int iterations, int reset) {
    std::vector<std::string> v(steps, steps + n);
    with(static_cast<Handle *>(p), [&](auto &o) { o.set_function(v, iterations, reset != 0); return 0; });
    return 0;
}

int orc_qubits(void *p) { return with(static_cast<Handle *>(p), [](auto &o) { return o.qubits; }); }

int64_t orc_num_gates(void *p) {
    return with(static_cast<Handle *>(p), [](auto &o) { return (int64_t)o.gates.size(); });
}

// parsed gate list: target bit, control mask, control value, matrix (8 doubles)
int orc_get_gates(void *p, int32_t *tbits, int64_t *cmask, int64_t *cval, double *m8) {
    return with(static_cast<Handle *>(p), [&](auto &o) {
        for (size_t i = 0; i < o.gates.size(); ++i) {
            tbits[i] = o.gates[i].target_bit;
            cmask[i] = o.gates[i].control_mask;
            cval[i] = o.gates[i].control_value;
            auto it = o.catalog.find(o.gates[i].name);
            for (int k = 0; k < 8; ++k) m8[8 * i + k] = NAN;
            if (it != o.catalog.end())
                for (int k = 0; k < 4; ++k) { m8[8 * i + 2 * k] = it->second[k].real(); m8[8 * i + 2 * k + 1] = it->second[k].imag(); }
        }
        return 0;
    });
}

// Replace the compiled gate list directly (used to replay a reordered schedule).
int orc_set_gates(void *p, int64_t n, const int32_t *tbits, const int64_t *cmask, const int64_t *cval,
                  const char *const *names) {
    return with(static_cast<Handle *>(p), [&](auto &o) {
        o.gates.clear();
        for (int64_t i = 0; i < n; ++i) {
            Gate g;
            g.target_bit = tbits[i]; g.control_mask = cmask[i]; g.control_value = cval[i];
            g.control_count = __builtin_popcountll((unsigned long long)cmask[i]);
            g.name = names[i];
            o.gates.push_back(g);
        }
        return 0;
    });
}

int orc_set_basis(void *p, int64_t index) {
    return with(static_cast<Handle *>(p), [&](auto &o) {
        using C = typename std::decay_t<decltype(o)>::C;
        std::fill(o.state.begin(), o.state.end(), C(0));
        o.state[index] = C(1);
        return 0;
    });
}

int orc_set_state(void *p, const void *amps) {
    return with(static_cast<Handle *>(p), [&](auto &o) {
        std::memcpy(o.state.data(), amps, o.state.size() * sizeof(o.state[0]));
        return 0;
    });
}

int orc_get_state(void *p, void *out) {
    return with(static_cast<Handle *>(p), [&](auto &o) {
        std::memcpy(out, o.state.data(), o.state.size() * sizeof(o.state[0]));
        return 0;
    });
}

// engine 0 = t_CPU (honours iterations), 1 = t_PAR_CPU (ignores them, as the reference does)
int orc_execute(void *p, int engine, int threads, int coalesced_bits, int region_bits, int iterations) {
    return with(static_cast<Handle *>(p), [&](auto &o) {
        if (o.qubits != o.state_qubits) {
            g_err = "circuit qubit count differs from the state's";
            return 2;
        }
        bool ok = engine == ENGINE_TCPU ? o.run_tcpu(iterations)
                                        : o.run_tparcpu(threads < 1 ? 1 : threads, coalesced_bits, region_bits);
        if (!ok) { g_err = o.error; return 1; }
        return 0;
    });
}

void orc_seed(void *p, unsigned long seed) {
    with(static_cast<Handle *>(p), [&](auto &o) { o.rng.seed(seed); return 0; });
}

int orc_measure(void *p, int qubit) {
    return with(static_cast<Handle *>(p), [&](auto &o) { return o.measure(qubit); });
}

// Fixture search (tests/golden/make_golden.py): the smallest seed >= start,
// below start + count, whose first uniform_real_distribution<double>(0, 1)
// draw from a std::mt19937 seeded with it (src/core/common.cpp:10-14, the
// draw of src/core/dgm_core.cpp:97) lies in [lo, hi); -1 if none. Searched on
// `threads` OpenMP threads.
long orc_find_draw_seed(double lo, double hi, long start, long count, int threads) {
    std::atomic<long> best{-1};
    const long chunk = 1 << 16;
    std::atomic<long> next{0};
    #pragma omp parallel num_threads(threads < 1 ? 1 : threads)
    {
        std::mt19937 g;
        std::uniform_real_distribution<double> dist(0.0, 1.0);
        for (;;) {
            const long c0 = next.fetch_add(chunk);
            const long b = best.load();
            if (c0 >= count || (b >= 0 && start + c0 > b)) break;
            const long c1 = std::min(count, c0 + chunk);
            for (long k = c0; k < c1; ++k) {
                g.seed((unsigned long)(start + k));
                dist.reset();
                const double r = dist(g);
                if (r >= lo && r < hi) {
                    long cur = best.load();
                    while ((cur < 0 || start + k < cur) && !best.compare_exchange_weak(cur, start + k)) {}
                    break;
                }
            }
        }
    }
    return best.load();
}

}  // extern "C"
