// Pass planner (see planner.hpp).
#include "planner.hpp"

#include <complex>

#include <chrono>
#include <cstring>

#include "jit.hpp"

namespace hgm {

namespace {

constexpr size_t kMaxLdsBytes = 160 * 1024;   // gfx950 LDS per workgroup
constexpr size_t kLdsReserve = 2 * 1024;      // bit-deposit tables of the LDS kernel

bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }
int log2i(int64_t v) { int k = 0; while ((int64_t(1) << k) < v) ++k; return k; }

struct GateInfo {
    int kind;
    int target;
    uint64_t need;    // bits that must be tile-local (target of a non-diagonal gate)
    uint64_t writes;  // bits whose amplitudes the gate mixes
    uint64_t reads;   // bits the gate only inspects (controls, diagonal targets)
};

void fill_pass_ops(const std::vector<GateInfo> &info, const hgm_op *ops, int64_t n_ops,
                   const std::vector<int64_t> &chosen, Pass &pass, Plan &plan) {
    int tile_of[64];
    for (int b = 0; b < 64; ++b) tile_of[b] = -1;
    for (int j = 0; j < plan.shape.r; ++j) tile_of[(int)pass.local_pos[j]] = j;
    pass.op_begin = (int64_t)plan.ops.size();
    for (int64_t gi : chosen) {
        const hgm_op &o = ops[gi % n_ops];
        const GateInfo &g = info[gi % n_ops];
        DevOp d;
        std::memset(&d, 0, sizeof(d));
        std::memcpy(d.m, o.m, sizeof(d.m));
        if (g.kind == K_DIAG && tile_of[g.target] < 0) {
            d.kind = K_DIAG_GLOBAL;
            d.tbit = g.target;
        } else {
            d.kind = g.kind;
            d.tbit = tile_of[g.target];
        }
        for (int b = 0; b < 64; ++b) {
            if (!((o.control_mask >> b) & 1)) continue;
            uint64_t v = (o.control_value >> b) & 1;
            if (tile_of[b] >= 0) {
                d.lmask |= 1u << tile_of[b];
                d.lval |= uint32_t(v) << tile_of[b];
            } else {
                d.gmask |= uint64_t(1) << b;
                d.gval |= v << b;
            }
        }
        plan.ops.push_back(d);
        plan.src_index.push_back(gi);
        plan.pre_phase.push_back(g.kind == K_DIAG && g.need ? 1 : 0);
    }
    pass.op_count = (int64_t)plan.ops.size() - pass.op_begin;
}

void finish_pass(uint64_t local, const Plan &plan, Pass &pass) {
    const int n = plan.qubits, c = plan.shape.c, r = plan.shape.r;
    // top up to r local bits with the lowest free bits
    int have = __builtin_popcountll(local);
    for (int b = 0; b < n && have < r; ++b)
        if (!((local >> b) & 1)) { local |= uint64_t(1) << b; ++have; }
    pass.local_mask = local;
    std::memset(pass.local_pos, -1, sizeof(pass.local_pos));
    int j = 0;
    for (int b = 0; b < c; ++b) pass.local_pos[j++] = (int8_t)b;
    for (int b = c; b < n; ++b)
        if ((local >> b) & 1) pass.local_pos[j++] = (int8_t)b;
}

// Arithmetic of one gate per amplitude pair as the generated pass code does
// it: per output component, a term with coefficient +-1 is an addend and
// every other nonzero term one multiply or FMA; controls leave 2^-k of the
// pairs to update. Used to balance passes, not as a flop count.
double gate_cost(const hgm_op &o) {
    auto unit = [](double re, double im) { return im == 0.0 && (re == 1.0 || re == -1.0); };
    double ops = 0;
    for (int row = 0; row < 2; ++row) {
        const double *a = o.m + 4 * row, *b = o.m + 4 * row + 2;
        // both components of the row's output use the same coefficients
        int terms = 0, units = 0;
        for (const double *e : {a, b}) {
            const int nz = (e[0] != 0.0) + (e[1] != 0.0);
            if (unit(e[0], e[1])) { ++terms; ++units; }
            else terms += nz;
        }
        if (terms == 1 && units == 1) continue;                  // a copy or a sign
        ops += 2.0 * (terms - (units ? 1 : 0));
    }
    return ops / double(uint64_t(1) << __builtin_popcountll(o.control_mask));
}

// Arithmetic (gate_cost units: multiply-adds per amplitude pair) that one pass
// over the state hides under its memory sweep on MI355X: a wave64 FP64 FMA
// issues every 4 cycles on each of the 1024 SIMDs (about 3e13 pair-ops/s at
// the loaded clock), and a sweep moves 64 bytes per FP64 pair at about
// 6.1 TB/s (1e11 pairs/s); FP32 halves both. Measured on the 30-qubit
// random-layer circuit, the per-pass VALU floors and memory floors sum to
// about the same time at ~300 per pass.
constexpr double kSweepCost = 300.0;

// Moves gates between passes up to four apart, from the one with more
// arithmetic to the one with less, while that lowers the sum of the two
// passes' times (max of sweep and arithmetic) or, at equal sums, the larger of
// them: a gate moves to the front of a later pass when it commutes with every
// gate after it in its own pass and with every gate of the passes in between,
// or to the end of an earlier pass likewise, and only where its target is
// tile-local there; a
// phase that rides with the rotation after it (HGM_PLAN_ROTATE) moves with
// that rotation. The pass count and the set of gates are unchanged.
void balance_passes(const std::vector<GateInfo> &info, const hgm_op *ops, int64_t n_ops, int64_t cap,
                    const std::vector<uint64_t> &local, std::vector<std::vector<int64_t>> &lists) {
    const size_t P = lists.size();
    if (P < 2) return;
    // a phase tagged to ride with the rotation after it moves with that rotation
    auto tagged = [&](int64_t k) { return ops[k % n_ops].reserved == kPrePhaseTag && info[k % n_ops].need; };
    auto unit_begin = [&](const std::vector<int64_t> &l, int64_t j) {
        return (j > 0 && tagged(l[j - 1]) && ops[l[j - 1] % n_ops].target_bit == ops[l[j] % n_ops].target_bit) ? j - 1
                                                                                                                    : j;
    };
    auto unit_end = [&](const std::vector<int64_t> &l, int64_t j) {   // one past the unit starting at j
        return (tagged(l[j]) && j + 1 < (int64_t)l.size() &&
                ops[l[j + 1] % n_ops].target_bit == ops[l[j] % n_ops].target_bit) ? j + 2 : j + 1;
    };
    auto cost_of = [&](const std::vector<int64_t> &l, int64_t a, int64_t b) {
        double c = 0;
        for (int64_t k = a; k < b; ++k) c += gate_cost(ops[l[k] % n_ops]);
        return c;
    };
    std::vector<double> cost(P, 0.0);
    for (size_t i = 0; i < P; ++i)
        for (int64_t k : lists[i]) cost[i] += gate_cost(ops[k % n_ops]);
    // A pass takes about max(memory sweep, arithmetic): arithmetic below the
    // sweep is hidden. Moving cost c from pass a to pass b helps when it
    // lowers the sum of the two, or, at equal sums, the larger of the two.
    auto eff = [&](double x) { return std::max(x, kSweepCost); };
    auto helps = [&](double from, double to, double c) {
        const double gain = eff(from) + eff(to) - eff(from - c) - eff(to + c);
        return c > 0 && (gain > 1e-9 || (gain > -1e-9 && c < from - to));
    };
    // the writes and reads of all the gates of a pass
    std::vector<uint64_t> pw(P, 0), pr(P, 0);
    auto refresh = [&](size_t i) {
        pw[i] = pr[i] = 0;
        for (int64_t k : lists[i]) { pw[i] |= info[k % n_ops].writes; pr[i] |= info[k % n_ops].reads; }
    };
    for (size_t i = 0; i < P; ++i) refresh(i);
    constexpr size_t kHops = 4;   // a gate moves at most this many passes
    for (int sweep = 0; sweep < 16; ++sweep) {
        bool moved = false;
        for (size_t i = 0; i < P; ++i) {
            // forward: a unit at the tail of pass i (commuting with the gates after it there) to the
            // front of a later pass j, when it commutes with every gate of the passes in between
            for (bool again = true; again;) {
                again = false;
                uint64_t bw = 0, br = 0;
                for (int64_t jj = (int64_t)lists[i].size() - 1; jj >= 0 && !again; --jj) {
                    const int64_t a = unit_begin(lists[i], jj);
                    uint64_t uw = 0, ur = 0, need = 0;
                    for (int64_t k = a; k <= jj; ++k) {
                        const GateInfo &g = info[lists[i][k] % n_ops];
                        uw |= g.writes;
                        ur |= g.reads;
                        need |= g.need;
                    }
                    const double c = cost_of(lists[i], a, jj + 1);
                    if (!(uw & (bw | br)) && !(ur & bw)) {
                        uint64_t mw = 0, mr = 0;
                        for (size_t j = i + 1; j < P && j <= i + kHops; ++j) {
                            if ((int64_t)lists[j].size() < cap && !(need & ~local[j]) && helps(cost[i], cost[j], c)) {
                                lists[j].insert(lists[j].begin(), lists[i].begin() + a, lists[i].begin() + jj + 1);
                                lists[i].erase(lists[i].begin() + a, lists[i].begin() + jj + 1);
                                cost[i] -= c;
                                cost[j] += c;
                                refresh(i);
                                refresh(j);
                                moved = again = true;
                                break;
                            }
                            mw |= pw[j];
                            mr |= pr[j];
                            if ((uw & (mw | mr)) || (ur & mw)) break;   // blocked by the gates of pass j
                        }
                    }
                    bw |= uw;
                    br |= ur;
                    jj = a;
                }
            }
            // backward: a unit at the head of pass i to the end of an earlier pass j
            for (bool again = true; again;) {
                again = false;
                uint64_t bw = 0, br = 0;
                for (int64_t jj = 0; jj < (int64_t)lists[i].size() && !again;) {
                    const int64_t e = unit_end(lists[i], jj);
                    uint64_t uw = 0, ur = 0, need = 0;
                    for (int64_t k = jj; k < e; ++k) {
                        const GateInfo &g = info[lists[i][k] % n_ops];
                        uw |= g.writes;
                        ur |= g.reads;
                        need |= g.need;
                    }
                    const double c = cost_of(lists[i], jj, e);
                    if (!(uw & (bw | br)) && !(ur & bw)) {
                        uint64_t mw = 0, mr = 0;
                        for (size_t d = 1; d <= kHops && d <= i; ++d) {
                            const size_t j = i - d;
                            if ((int64_t)lists[j].size() < cap && !(need & ~local[j]) && helps(cost[i], cost[j], c)) {
                                lists[j].insert(lists[j].end(), lists[i].begin() + jj, lists[i].begin() + e);
                                lists[i].erase(lists[i].begin() + jj, lists[i].begin() + e);
                                cost[i] -= c;
                                cost[j] += c;
                                refresh(i);
                                refresh(j);
                                moved = again = true;
                                break;
                            }
                            mw |= pw[j];
                            mr |= pr[j];
                            if ((uw & (mw | mr)) || (ur & mw)) break;
                        }
                    }
                    bw |= uw;
                    br |= ur;
                    jj = e;
                }
            }
        }
        if (!moved) break;
    }
}

}  // namespace

int classify(const double m[8]) {
    auto zero = [&](int k) { return m[2 * k] == 0.0 && m[2 * k + 1] == 0.0; };
    if (zero(1) && zero(2)) return K_DIAG;
    if (zero(0) && zero(3)) return K_ANTI;
    return K_DENSE;
}

size_t amp_bytes(int dtype) { return dtype == HGM_C64 ? 8 : 16; }

int make_shape(int qubits, int dtype, const hgm_tuning &t, TileShape *out, std::string *err) {
    if (qubits < 1 || qubits > kMaxQubits) {
        *err = "qubits must be in [1, " + std::to_string(kMaxQubits) + "], got " + std::to_string(qubits);
        return HGM_E_ARG;
    }
    if (dtype != HGM_C128 && dtype != HGM_C64) { *err = "unknown amplitude dtype"; return HGM_E_ARG; }
    const int B = t.block_size, R = t.repeat_count, r = t.gpu_region_bits, c = t.gpu_coalesced_bits;
    if (B < 1 || R < 1 || r < 1 || r > kMaxTileBits) {
        *err = "Erro de tuning: block_size, repeat_count and gpu_region_bits must be positive (gpu_region_bits <= " +
               std::to_string(kMaxTileBits) + ")";
        return HGM_E_TUNING;
    }
    // src/core/dgm_core.cpp:207-212
    if (2LL * B * R != (int64_t(1) << r)) {
        *err = "Erro de tuning: 2*block_size*repeat_count (" + std::to_string(2LL * B * R) +
               ") precisa ser igual a 2^gpu_region_bits (" + std::to_string(int64_t(1) << r) + ")";
        return HGM_E_TUNING;
    }
    // src/core/dgm_core.cpp:218-223
    if (c < 0 || r < c) {
        *err = "Erro de tuning: gpu_region_bits (" + std::to_string(r) + ") precisa ser >= gpu_coalesced_bits (" +
               std::to_string(c) + ")";
        return HGM_E_TUNING;
    }
    if (B > 1024) { *err = "Erro de tuning: block_size must be <= 1024 on gfx950"; return HGM_E_TUNING; }
    if (!is_pow2(B) || !is_pow2(R)) { *err = "Erro de tuning: block_size and repeat_count must be powers of two"; return HGM_E_TUNING; }
    if (t.ops_block < 1) { *err = "Erro de tuning: ops_block must be >= 1"; return HGM_E_TUNING; }
    const int base_planner = t.planner & ~(HGM_PLAN_FOLD_PHASES | HGM_PLAN_FUSE | HGM_PLAN_PIVOT | HGM_PLAN_BALANCE |
                                           HGM_PLAN_LOOKAHEAD | HGM_PLAN_ROTATE);
    if (base_planner != HGM_PLAN_GREEDY && base_planner != HGM_PLAN_WINDOW) {
        *err = "unknown planner";
        return HGM_E_TUNING;
    }
    if (t.kernel < HGM_KERNEL_AUTO || t.kernel > HGM_KERNEL_JIT) { *err = "unknown kernel family"; return HGM_E_TUNING; }
    // the plan-compiled kernels move a tile above 32 KiB through LDS in halves (real parts, then imaginary
    // parts), with one padding slot per 16 and per 256 entries: FP32 tiles of 2^15 fit
    const size_t tile_bytes = (size_t(1) << r) * amp_bytes(dtype);
    const size_t lds_need = t.kernel == HGM_KERNEL_JIT && tile_bytes > 32 * 1024
                                ? (tile_bytes + (tile_bytes >> 4) + (tile_bytes >> 8)) / 2
                                : tile_bytes;
    if (lds_need > kMaxLdsBytes - kLdsReserve) {
        *err = "Erro de tuning: a tile of 2^" + std::to_string(r) + " amplitudes exceeds the 160 KiB LDS of a gfx950 workgroup";
        return HGM_E_TUNING;
    }
    TileShape s;
    if (qubits <= r) {
        // the whole state is one tile; shrink the workgroup to the state
        s.r = qubits;
        s.c = c < qubits ? c : qubits;
        int64_t pairs = int64_t(1) << (qubits - 1);
        s.B = (int)(B < pairs ? B : pairs);
        s.R = (int)(pairs / s.B);
    } else {
        // a pass must have at least one selectable bit beyond the coalesced
        // ones, or a gate on a high bit could never be scheduled (the
        // reference loops forever there, src/core/kernel.cu:208,281)
        if (c >= r) {
            *err = "Erro de tuning: gpu_region_bits (" + std::to_string(r) + ") must exceed gpu_coalesced_bits (" +
                   std::to_string(c) + ") when the state has more qubits than a tile";
            return HGM_E_TUNING;
        }
        s.r = r; s.c = c; s.B = B; s.R = R;
    }
    *out = s;
    return HGM_OK;
}

void fuse_gates(int qubits, const hgm_op *ops, int64_t n_ops, std::vector<hgm_op> *out) {
    using C = std::complex<double>;
    struct M2 { C a, b, c, d; };   // row-major 2x2
    const M2 I{C(1, 0), C(0, 0), C(0, 0), C(1, 0)};
    std::vector<M2> P((size_t)qubits, I);
    std::vector<char> has((size_t)qubits, 0);
    auto mat = [](const hgm_op &o) {
        return M2{C(o.m[0], o.m[1]), C(o.m[2], o.m[3]), C(o.m[4], o.m[5]), C(o.m[6], o.m[7])};
    };
    auto mul = [](const M2 &x, const M2 &y) {   // x * y
        return M2{x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
    };
    auto diag = [](const M2 &m) { return m.b == C(0, 0) && m.c == C(0, 0); };
    auto emit = [&](int q) {
        if (!has[q]) return;
        has[q] = 0;
        const M2 m = P[q];
        P[q] = I;
        if (m.a == C(1, 0) && m.b == C(0, 0) && m.c == C(0, 0) && m.d == C(1, 0)) return;   // cancelled
        hgm_op g{};
        g.target_bit = q;
        const C v[4] = {m.a, m.b, m.c, m.d};
        for (int k = 0; k < 4; ++k) { g.m[2 * k] = v[k].real(); g.m[2 * k + 1] = v[k].imag(); }
        out->push_back(g);
    };
    out->clear();
    out->reserve((size_t)n_ops + (size_t)qubits);
    for (int64_t i = 0; i < n_ops; ++i) {
        const hgm_op &o = ops[i];
        const int q = o.target_bit;
        if (!o.control_mask) {            // accumulate: the gate acts after what is pending on q
            P[q] = has[q] ? mul(mat(o), P[q]) : mat(o);
            has[q] = 1;
            continue;
        }
        // a controlled gate: pending factors it does not commute with go first
        const bool d = classify(o.m) == K_DIAG;
        if (has[q] && !(d && diag(P[q]))) emit(q);
        for (uint64_t cm = o.control_mask; cm; cm &= cm - 1) {
            const int c = __builtin_ctzll(cm);
            if (has[c] && !diag(P[c])) emit(c);
        }
        out->push_back(o);
    }
    for (int q = 0; q < qubits; ++q) emit(q);
}

void fold_phases(int qubits, const hgm_op *ops, int64_t n_ops, std::vector<hgm_op> *out) {
    using C = std::complex<double>;
    std::vector<C> d0((size_t)qubits, C(1, 0)), d1((size_t)qubits, C(1, 0));
    auto get = [](const hgm_op &o, int k) { return C(o.m[2 * k], o.m[2 * k + 1]); };
    auto set = [](hgm_op &o, int k, C v) { o.m[2 * k] = v.real(); o.m[2 * k + 1] = v.imag(); };
    auto pending = [&](int b) { return d0[b] != C(1, 0) || d1[b] != C(1, 0); };
    auto flush = [&](int b) {
        if (!pending(b)) return;
        hgm_op g{};
        g.target_bit = b;
        set(g, 0, d0[b]);
        set(g, 3, d1[b]);
        out->push_back(g);
        d0[b] = d1[b] = C(1, 0);
    };
    out->clear();
    out->reserve((size_t)n_ops + (size_t)qubits);
    for (int64_t i = 0; i < n_ops; ++i) {
        hgm_op o = ops[i];
        const int b = o.target_bit;
        const int kind = classify(o.m);
        if (kind == K_DIAG) {            // commutes with every pending diagonal factor
            out->push_back(o);
            continue;
        }
        if (o.control_mask) {            // the factor on b must act on both control halves first
            flush(b);
            out->push_back(o);
            continue;
        }
        // U' = U * diag(d0, d1)
        const C u00 = get(o, 0) * d0[b], u01 = get(o, 1) * d1[b], u10 = get(o, 2) * d0[b], u11 = get(o, 3) * d1[b];
        if (kind == K_ANTI) {            // stays anti-diagonal: absorb and emit
            set(o, 1, u01);
            set(o, 2, u10);
            d0[b] = d1[b] = C(1, 0);
            out->push_back(o);
            continue;
        }
        // U' = diag(p0, p1) * M with the first column of M real and non-negative
        const double a00 = std::abs(u00), a10 = std::abs(u10);
        const C p0 = a00 > 0 ? u00 / a00 : C(1, 0), p1 = a10 > 0 ? u10 / a10 : C(1, 0);
        o.m[0] = a00;
        o.m[1] = 0.0;
        set(o, 1, std::conj(p0) * u01);
        o.m[4] = a10;
        o.m[5] = 0.0;
        set(o, 3, std::conj(p1) * u11);
        d0[b] = p0;
        d1[b] = p1;
        out->push_back(o);
    }
    for (int b = 0; b < qubits; ++b) flush(b);
}

void pivot_rows(int qubits, const hgm_op *ops, int64_t n_ops, std::vector<hgm_op> *out) {
    using C = std::complex<double>;
    // What is still to be applied after the emitted list: g * prod_b diag(1, r[b]) on bit b.
    std::vector<C> r((size_t)qubits, C(1, 0));
    C g(1, 0);
    auto get = [](const hgm_op &o, int k) { return C(o.m[2 * k], o.m[2 * k + 1]); };
    auto set = [](hgm_op &o, int k, C v) { o.m[2 * k] = v.real(); o.m[2 * k + 1] = v.imag(); };
    auto diag_op = [&](int b, C a, C d) {
        hgm_op o{};
        o.target_bit = b;
        set(o, 0, a);
        set(o, 3, d);
        out->push_back(o);
    };
    auto flush = [&](int b) {
        if (r[b] == C(1, 0)) return;
        diag_op(b, C(1, 0), r[b]);
        r[b] = C(1, 0);
    };
    // the emitted amplitudes are the true ones divided by g: keep |g| near 1
    // so that FP32 states neither overflow nor lose small amplitudes
    auto rescale = [&]() {
        const double s = std::abs(g);
        if (s >= 0x1p-24 && s <= 0x1p24) return;
        diag_op(0, C(s, 0), C(s, 0));
        g /= s;
    };
    out->clear();
    out->reserve((size_t)n_ops + (size_t)qubits + 1);
    for (int64_t i = 0; i < n_ops; ++i) {
        hgm_op o = ops[i];
        const int b = o.target_bit;
        const int kind = classify(o.m);
        const C u00 = get(o, 0), u11 = get(o, 3);
        if (kind == K_DIAG) {
            if (o.control_mask || u00 == C(0, 0)) {   // commutes with every carried factor
                out->push_back(o);
                continue;
            }
            g *= u00;                                  // diag(u00, u11) = u00 * diag(1, u11 / u00)
            r[b] *= u11 / u00;
            rescale();
            continue;
        }
        if (o.control_mask) {            // the factor on b must act on both control halves first
            flush(b);
            out->push_back(o);
            continue;
        }
        // U' = U * diag(1, r): each row scaled by its larger entry, which becomes exactly 1
        const C v[2][2] = {{u00, get(o, 1) * r[b]}, {get(o, 2), u11 * r[b]}};
        int p[2];
        C d[2];
        for (int row = 0; row < 2; ++row) {
            p[row] = std::abs(v[row][1]) > std::abs(v[row][0]) ? 1 : 0;
            d[row] = v[row][p[row]];
        }
        if (d[0] == C(0, 0) || d[1] == C(0, 0)) {   // a zero row: emit U' itself
            set(o, 0, v[0][0]);
            set(o, 1, v[0][1]);
            set(o, 2, v[1][0]);
            set(o, 3, v[1][1]);
            r[b] = C(1, 0);
            out->push_back(o);
            continue;
        }
        for (int row = 0; row < 2; ++row) {
            set(o, 2 * row + p[row], C(1, 0));
            set(o, 2 * row + 1 - p[row], v[row][1 - p[row]] / d[row]);
        }
        out->push_back(o);
        g *= d[0];                       // U' = diag(d0, d1) * M = d0 * diag(1, d1 / d0) * M
        r[b] = d[1] / d[0];
        rescale();
    }
    for (int b = 0; b < qubits; ++b) {
        if (r[b] == C(1, 0)) continue;
        diag_op(b, g, g * r[b]);         // the global factor rides on the first remaining diagonal
        g = C(1, 0);
        r[b] = C(1, 0);
    }
    if (g != C(1, 0)) diag_op(0, g, g);
}

// Phases x0, x1, y of modulus 1 with v[i][j] = x_i * o[i][j] * y_j (y_0 = 1)
// and every o[i][j] real, for the nonzero entries of v; false when they do
// not exist up to rounding (more than 16 units in the last place of the
// entry), as for a matrix that is no unitary times a scale. The nonzero
// entries link rows to columns; fixing column 0's phase at 1 and walking the
// links determines the others, and the remaining link (all four entries
// nonzero) is the check.
bool real_middle(const std::complex<double> v[2][2], std::complex<double> x[2], std::complex<double> *y,
                 double o[2][2]) {
    using C = std::complex<double>;
    auto unit = [](C z) { return z / std::abs(z); };
    C row[2], col[2];
    bool have_r[2] = {false, false}, have_c[2] = {true, false};
    col[0] = C(1, 0);
    for (int pass = 0; pass < 3; ++pass)
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                if (v[i][j] == C(0, 0)) continue;
                if (have_c[j] && !have_r[i]) { row[i] = unit(v[i][j]) / col[j]; have_r[i] = true; }
                else if (have_r[i] && !have_c[j]) { col[j] = unit(v[i][j]) / row[i]; have_c[j] = true; }
            }
    for (int i = 0; i < 2; ++i) if (!have_r[i]) row[i] = C(1, 0);
    if (!have_c[1]) col[1] = C(1, 0);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const C w = v[i][j] / (row[i] * col[j]);
            if (std::abs(w.imag()) > 16 * 0x1p-52 * std::abs(w)) return false;
            o[i][j] = w.real();
        }
    x[0] = row[0];
    x[1] = row[1];
    *y = col[1];
    return true;
}

void rotate_rows(int qubits, const hgm_op *ops, int64_t n_ops, std::vector<hgm_op> *out) {
    using C = std::complex<double>;
    // Still to be applied after the emitted list: g * prod_b diag(1, r[b]) on bit b (as in pivot_rows).
    std::vector<C> r((size_t)qubits, C(1, 0));
    C g(1, 0);
    auto get = [](const hgm_op &o, int k) { return C(o.m[2 * k], o.m[2 * k + 1]); };
    auto set = [](hgm_op &o, int k, C v) { o.m[2 * k] = v.real(); o.m[2 * k + 1] = v.imag(); };
    auto diag_op = [&](int b, C a, C d, int32_t tag) {
        hgm_op o{};
        o.target_bit = b;
        o.reserved = tag;
        set(o, 0, a);
        set(o, 3, d);
        out->push_back(o);
    };
    auto flush = [&](int b) {
        if (r[b] == C(1, 0)) return;
        diag_op(b, C(1, 0), r[b], 0);
        r[b] = C(1, 0);
    };
    auto rescale = [&]() {
        const double s = std::abs(g);
        if (s >= 0x1p-24 && s <= 0x1p24) return;
        diag_op(0, C(s, 0), C(s, 0), 0);
        g /= s;
    };
    out->clear();
    out->reserve(2 * (size_t)n_ops + (size_t)qubits + 1);
    for (int64_t i = 0; i < n_ops; ++i) {
        hgm_op o = ops[i];
        o.reserved = 0;
        const int b = o.target_bit;
        const int kind = classify(o.m);
        const C u00 = get(o, 0), u11 = get(o, 3);
        if (kind == K_DIAG) {
            if (o.control_mask || u00 == C(0, 0)) {   // commutes with every carried factor
                out->push_back(o);
                continue;
            }
            g *= u00;
            r[b] *= u11 / u00;
            rescale();
            continue;
        }
        if (o.control_mask) {            // the factor on b must act on both control halves first
            flush(b);
            out->push_back(o);
            continue;
        }
        // V = U * diag(1, r) = diag(x0, x1) * O * diag(1, y), O real
        const C v[2][2] = {{u00, get(o, 1) * r[b]}, {get(o, 2), u11 * r[b]}};
        C x[2], y;
        double om[2][2];
        const bool zero_row = (v[0][0] == C(0, 0) && v[0][1] == C(0, 0)) || (v[1][0] == C(0, 0) && v[1][1] == C(0, 0));
        if (kind == K_ANTI || zero_row || !real_middle(v, x, &y, om)) {
            // the row-scaled complex form of pivot_rows
            int p[2];
            C d[2];
            for (int row = 0; row < 2; ++row) {
                p[row] = std::abs(v[row][1]) > std::abs(v[row][0]) ? 1 : 0;
                d[row] = v[row][p[row]];
            }
            if (d[0] == C(0, 0) || d[1] == C(0, 0)) {
                set(o, 0, v[0][0]);
                set(o, 1, v[0][1]);
                set(o, 2, v[1][0]);
                set(o, 3, v[1][1]);
                r[b] = C(1, 0);
                out->push_back(o);
                continue;
            }
            for (int row = 0; row < 2; ++row) {
                set(o, 2 * row + p[row], C(1, 0));
                set(o, 2 * row + 1 - p[row], v[row][1 - p[row]] / d[row]);
            }
            out->push_back(o);
            g *= d[0];
            r[b] = d[1] / d[0];
            rescale();
            continue;
        }
        // the phase y rides in front of the real gate; each row of O is
        // scaled by its larger entry (exactly 1 after scaling), and the row
        // factors join x as the factor carried forward
        if (y != C(1, 0)) diag_op(b, C(1, 0), y, kPrePhaseTag);
        double d[2];
        for (int row = 0; row < 2; ++row) {
            const int pc = std::abs(om[row][1]) > std::abs(om[row][0]) ? 1 : 0;
            d[row] = om[row][pc];
            o.m[4 * row + 2 * pc] = 1.0;
            o.m[4 * row + 2 * pc + 1] = 0.0;
            o.m[4 * row + 2 * (1 - pc)] = om[row][1 - pc] / d[row];
            o.m[4 * row + 2 * (1 - pc) + 1] = 0.0;
        }
        out->push_back(o);
        const C l0 = x[0] * d[0], l1 = x[1] * d[1];
        g *= l0;
        r[b] = l1 / l0;
        rescale();
    }
    for (int b = 0; b < qubits; ++b) {
        if (r[b] == C(1, 0)) continue;
        diag_op(b, g, g * r[b], 0);
        g = C(1, 0);
        r[b] = C(1, 0);
    }
    if (g != C(1, 0)) diag_op(0, g, g, 0);
}

int build_plan(int qubits, int dtype, const hgm_op *ops, int64_t n_ops, int iterations, const hgm_tuning &t,
               Plan *out, std::string *err, bool program_only) {
    auto t0 = std::chrono::steady_clock::now();
    Plan plan;
    plan.qubits = qubits;
    plan.dtype = dtype;
    int st = make_shape(qubits, dtype, t, &plan.shape, err);
    if (st) return st;
    std::vector<hgm_op> expanded, fused, folded;
    const bool rotate = (t.planner & HGM_PLAN_ROTATE) != 0;
    if ((t.planner & (HGM_PLAN_FOLD_PHASES | HGM_PLAN_FUSE | HGM_PLAN_PIVOT | HGM_PLAN_ROTATE)) && n_ops > 0 && ops) {
        for (int64_t i = 0; i < n_ops; ++i)
            if (ops[i].target_bit < 0 || ops[i].target_bit >= qubits || (ops[i].control_mask >> ops[i].target_bit) & 1) {
                *err = "gate " + std::to_string(i) + ": target bit " + std::to_string(ops[i].target_bit) +
                       " outside the state or among its controls";
                return HGM_E_ARG;
            }
        if (t.planner & HGM_PLAN_FUSE) {
            if (iterations > 1) {   // fuse across iteration boundaries too
                expanded.reserve((size_t)(n_ops * iterations));
                for (int it = 0; it < iterations; ++it) expanded.insert(expanded.end(), ops, ops + n_ops);
                ops = expanded.data();
                n_ops = (int64_t)expanded.size();
                iterations = 1;
            }
            fuse_gates(qubits, ops, n_ops, &fused);
            ops = fused.data();
            n_ops = (int64_t)fused.size();
        }
        if (rotate) {                       // subsumes row scaling
            rotate_rows(qubits, ops, n_ops, &folded);
            ops = folded.data();
            n_ops = (int64_t)folded.size();
        } else if (t.planner & HGM_PLAN_PIVOT) {   // subsumes phase folding
            pivot_rows(qubits, ops, n_ops, &folded);
            ops = folded.data();
            n_ops = (int64_t)folded.size();
        } else if (t.planner & HGM_PLAN_FOLD_PHASES) {
            fold_phases(qubits, ops, n_ops, &folded);
            ops = folded.data();
            n_ops = (int64_t)folded.size();
        }
        if (n_ops == 0) {   // the whole list cancelled: one identity keeps the plan well formed
            hgm_op g{};
            g.m[0] = g.m[6] = 1.0;
            folded.assign(1, g);
            ops = folded.data();
            n_ops = 1;
        }
    }
    if (iterations < 0) { *err = "iterations must be >= 0"; return HGM_E_ARG; }
    if (n_ops < 0 || (n_ops > 0 && !ops)) { *err = "bad op array"; return HGM_E_ARG; }
    plan.kernel = t.kernel == HGM_KERNEL_AUTO ? HGM_KERNEL_LDS : t.kernel;

    const uint64_t all = qubits >= 64 ? ~uint64_t(0) : ((uint64_t(1) << qubits) - 1);
    std::vector<GateInfo> info((size_t)n_ops);
    for (int64_t i = 0; i < n_ops; ++i) {
        const hgm_op &o = ops[i];
        if (o.target_bit < 0 || o.target_bit >= qubits) {
            *err = "gate " + std::to_string(i) + ": target bit " + std::to_string(o.target_bit) + " outside the state";
            return HGM_E_ARG;
        }
        uint64_t tb = uint64_t(1) << o.target_bit;
        if ((o.control_mask & ~all) || (o.control_mask & tb) || (o.control_value & ~o.control_mask)) {
            *err = "gate " + std::to_string(i) + ": control mask/value inconsistent with the target or the state";
            return HGM_E_ARG;
        }
        GateInfo g;
        g.kind = classify(o.m);
        g.target = o.target_bit;
        if (g.kind == K_DIAG) {
            // a phase in front of a rotation stays with it: its target must be
            // local, and it is ordered like the rotation against gates on its qubit
            const bool pre = rotate && o.reserved == kPrePhaseTag;
            g.need = pre ? tb : 0;
            g.writes = pre ? tb : 0;
            g.reads = tb | o.control_mask;
        }
        else { g.need = tb; g.writes = tb; g.reads = o.control_mask; }
        info[i] = g;
    }

    const int64_t N = n_ops * (int64_t)iterations;   // expanded list: gate k is ops[k % n_ops]
    plan.gates = N;
    const int r = plan.shape.r, c = plan.shape.c;
    const uint64_t low = c >= 64 ? ~uint64_t(0) : ((uint64_t(1) << c) - 1);
    const int64_t window = t.lookahead > 0 ? t.lookahead : 4096;
    const int64_t cap = t.ops_block;

    std::vector<int64_t> chosen;
    if (qubits <= r) {
        // whole state in one tile: only the ops_block cap splits passes
        for (int64_t k = 0; k < N; k += cap) {
            chosen.clear();
            for (int64_t j = k; j < N && j < k + cap; ++j) chosen.push_back(j);
            Pass p;
            finish_pass(all, plan, p);
            fill_pass_ops(info, ops, n_ops, chosen, p, plan);
            plan.passes.push_back(p);
        }
    } else if ((t.planner & 0xF) == HGM_PLAN_WINDOW) {
        // In order; targets >= c must fit one window of e = r-c contiguous bits.
        const int e = r - c;
        int64_t k = 0;
        while (k < N) {
            chosen.clear();
            bool have = false;
            int lo = 0, hi = 0;
            while (k < N && (int64_t)chosen.size() < cap) {
                const GateInfo &g = info[k % n_ops];
                if (g.need && g.target >= c) {
                    int b = g.target;
                    if (!have) { lo = hi = b; have = true; }
                    else if (b >= lo && b - lo < e) { if (b > hi) hi = b; }
                    else if (b <= hi && hi - b < e) { if (b < lo) lo = b; }
                    else break;
                }
                chosen.push_back(k++);
            }
            int rs = have ? (hi - e + 1 > c ? hi - e + 1 : c) : c;
            if (rs + e > qubits) rs = qubits - e;
            uint64_t local = low | (((uint64_t(1) << e) - 1) << rs);
            Pass p;
            finish_pass(local, plan, p);
            fill_pass_ops(info, ops, n_ops, chosen, p, plan);
            plan.passes.push_back(p);
        }
    } else {
        // Greedy: scan ahead from the oldest unscheduled gate; take every gate
        // that commutes with all the gates skipped before it and whose target
        // is local. The local set is either grown on demand (the first r-c
        // new targets met) or fixed to a window of r-c contiguous bits above
        // the coalesced ones; every window is tried and the set that admits
        // the most gates wins (on layered circuits the windows tile the
        // light cone better than first-come targets do).
        std::vector<char> done((size_t)N, 0);
        std::vector<int64_t> cand;
        int64_t head = 0;
        auto scan_from = [&](int64_t from, uint64_t fixed, bool grow, uint64_t *local_out, std::vector<int64_t> &out) {
            out.clear();
            uint64_t local = grow ? low : fixed;
            int nlocal = __builtin_popcountll(local);
            uint64_t bw = 0, br = 0;   // writes / reads of the skipped gates
            int64_t scanned = 0;
            for (int64_t k = from; k < N && (int64_t)out.size() < cap && scanned < window; ++k) {
                if (done[k]) continue;
                ++scanned;
                const GateInfo &g = info[k % n_ops];
                bool ok = !(g.writes & (bw | br)) && !(g.reads & bw);
                if (ok && (g.need & ~local)) {
                    if (grow && nlocal < r) { local |= g.need; ++nlocal; }
                    else ok = false;
                }
                if (ok) {
                    out.push_back(k);
                } else {
                    bw |= g.writes;
                    br |= g.reads;
                    if (bw == all) break;
                }
            }
            *local_out = local;
        };
        const int e = r - c;
        // the candidate local sets of a pass: grown on demand (index 0), or one window per index
        const int n_cands = 1 + std::max(0, qubits - e - c + 1);
        auto cand_scan = [&](int64_t from, int i, uint64_t *local_out, std::vector<int64_t> &out) {
            if (i == 0) scan_from(from, 0, true, local_out, out);
            else scan_from(from, low | (((uint64_t(1) << e) - 1) << (c + i - 1)), false, local_out, out);
        };
        const bool lookahead = (t.planner & HGM_PLAN_LOOKAHEAD) != 0;
        std::vector<int64_t> next;
        // the most gates one pass could take once `taken` is done (marked in place, then unmarked)
        auto best_next = [&](const std::vector<int64_t> &taken) -> size_t {
            for (int64_t k : taken) done[k] = 1;
            int64_t h = head;
            while (h < N && done[h]) ++h;
            size_t best = 0;
            uint64_t l;
            for (int i = 0; h < N && i < n_cands; ++i) {
                cand_scan(h, i, &l, next);
                best = std::max(best, next.size());
            }
            for (int64_t k : taken) done[k] = 0;
            return best;
        };
        std::vector<std::vector<int64_t>> lists;
        std::vector<Pass> shells;
        std::vector<uint64_t> locals;
        while (head < N) {
            uint64_t local = 0;
            size_t best_value = 0;
            chosen.clear();
            for (int i = 0; i < n_cands; ++i) {
                uint64_t L;
                cand_scan(head, i, &L, cand);
                // plain greedy: the most gates now; lookahead: the most over this pass and the next
                const size_t value = lookahead ? cand.size() + best_next(cand) : cand.size();
                if (i == 0 || value > best_value || (value == best_value && cand.size() > chosen.size())) {
                    best_value = value;
                    chosen.swap(cand);
                    local = L;
                }
            }
            for (int64_t k : chosen) done[k] = 1;
            while (head < N && done[head]) ++head;
            Pass p;
            finish_pass(local, plan, p);
            shells.push_back(p);
            locals.push_back(p.local_mask);
            lists.push_back(chosen);
        }
        if (t.planner & HGM_PLAN_BALANCE) balance_passes(info, ops, n_ops, cap, locals, lists);
        for (size_t i = 0; i < shells.size(); ++i) {
            fill_pass_ops(info, ops, n_ops, lists[i], shells[i], plan);
            plan.passes.push_back(shells[i]);
        }
    }
    const int A = 2 * plan.shape.R;
    const bool reg_ok = A <= kMaxRegAmps && plan.shape.B <= reg_kernel_max_threads(A, amp_bytes(dtype));
    const int max_B = t.kernel == HGM_KERNEL_JIT ? compiled_kernel_max_threads(A, amp_bytes(dtype))
                                                 : reg_kernel_max_threads(A, amp_bytes(dtype));
    if ((t.kernel == HGM_KERNEL_REG || t.kernel == HGM_KERNEL_JIT) && (A > kMaxRegAmps || plan.shape.B > max_B)) {
        *err = "Erro de tuning: the register kernel holds at most " + std::to_string(kMaxRegAmps) +
               " amplitudes per thread and at most " + std::to_string(max_B) + " threads for 2*repeat_count = " +
               std::to_string(A);
        return HGM_E_TUNING;
    }
    if (program_only && t.kernel != HGM_KERNEL_JIT) {
        *err = "a program-only plan is for plan-compiled tile shapes (HGM_KERNEL_JIT)";
        return HGM_E_ARG;
    }
    if (t.kernel == HGM_KERNEL_REG || t.kernel == HGM_KERNEL_JIT || (t.kernel == HGM_KERNEL_AUTO && reg_ok)) {
        plan.kernel = t.kernel == HGM_KERNEL_JIT ? HGM_KERNEL_JIT : HGM_KERNEL_REG;
        build_register_program(plan);
        if (plan.kernel == HGM_KERNEL_JIT) {
            // one workgroup per tile: a dispatch holds at most 2^32 - 1 work-items
            const uint64_t items = (uint64_t(1) << (qubits - plan.shape.r)) * (uint64_t)plan.shape.B;
            if (items > 0xffffffffull) {
                *err = "Erro de tuning: a plan-compiled pass launches 2^(qubits-1)/repeat_count work-items, above "
                       "the 2^32 - 1 of one dispatch; raise repeat_count";
                return HGM_E_TUNING;
            }
            if (program_only) {
                plan.kernel = HGM_KERNEL_REG;
            } else {
                plan.jit = std::make_shared<JitCode>();
                jit_generate(plan, plan.jit.get());
                if (int rc = jit_compile(plan.jit.get(), err)) return rc;
            }
        }
    }
    plan.plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = std::move(plan);
    return HGM_OK;
}

namespace {

int popc(uint64_t v) { return __builtin_popcountll(v); }

// Splits a tile-coordinate condition into register-index and thread-index parts.
void split_cond(uint32_t mask, uint32_t val, uint32_t S, uint32_t all, RegOp &o) {
    for (int b = 0; b < 32; ++b) {
        if (!((mask >> b) & 1)) continue;
        const uint32_t v = (val >> b) & 1;
        const uint32_t below = (1u << b) - 1;
        if ((S >> b) & 1) {
            const int i = popc(S & below);
            o.rmask |= 1u << i;
            o.rval |= v << i;
        } else {
            const int j = popc(~S & all & below);
            o.tmask |= 1u << j;
            o.tval |= v << j;
        }
    }
}

// LDS slot of tile index x: the upper nibbles are XOR-folded into the low
// one, so that lanes walking any 4 consecutive tile bits hit distinct 16-byte
// slots. The map is linear over GF(2) and its own inverse.
uint32_t lds_swizzle(uint32_t x) { return x ^ ((x >> 4) & 15u) ^ ((x >> 8) & 15u) ^ ((x >> 12) & 15u); }

// Tile indices of a phase assignment (unswizzled).
PhaseMap make_map(uint32_t S, int r) {
    PhaseMap pm;
    std::memset(&pm, 0, sizeof(pm));
    int sb[kMaxTileBits], tbits[kMaxTileBits];
    int na = 0, nt = 0;
    for (int b = 0; b < r; ++b) {
        if ((S >> b) & 1) sb[na++] = b;
        else tbits[nt++] = b;
    }
    for (int k = 0; k < 64; ++k) {
        uint32_t v = 0;
        for (int i = 0; i < na; ++i)
            if ((k >> i) & 1) v |= 1u << sb[i];
        pm.reg[k] = (k < (1 << na)) ? (uint16_t)v : 0;
    }
    for (int h = 0; h < 2; ++h)
        for (int x = 0; x < 32; ++x) {
            uint32_t v = 0;
            for (int j = 0; j < 5; ++j)
                if (((x >> j) & 1) && 5 * h + j < nt) v |= 1u << tbits[5 * h + j];
            pm.thr[h][x] = (uint16_t)v;
        }
    return pm;
}

PhaseMap swizzled(const PhaseMap &m) {
    PhaseMap s;
    for (int k = 0; k < 64; ++k) s.reg[k] = (uint16_t)lds_swizzle(m.reg[k]);
    for (int h = 0; h < 2; ++h)
        for (int x = 0; x < 32; ++x) s.thr[h][x] = (uint16_t)lds_swizzle(m.thr[h][x]);
    return s;
}

uint64_t to_state(uint32_t tile_mask, const Pass &ps) {
    uint64_t g = 0;
    for (int b = 0; b < kMaxTileBits; ++b)
        if ((tile_mask >> b) & 1) g |= uint64_t(1) << ps.local_pos[b];
    return g;
}

bool is_one(double re, double im) { return re == 1.0 && im == 0.0; }

// Column of row `row` of the row-major 2x2 matrix m that holds an exact 1
// (the first such column).
bool unit_pivot(const double m[8], int row, int *col) {
    for (int c = 0; c < 2; ++c)
        if (is_one(m[4 * row + 2 * c], m[4 * row + 2 * c + 1])) { *col = c; return true; }
    return false;
}

}  // namespace

// Second-level schedule: inside each pass, the same commutation-aware greedy
// as the pass planner, with the tile as the "state" and the log2(A) register
// bits as the "tile". The first and the last phase keep the coalesced tile
// bits on the thread index so that HBM loads and stores stay contiguous
// across lanes (an empty phase is added when a gate needs them earlier or the
// schedule ends with them in registers).
void build_register_program(Plan &plan) {
    const int r = plan.shape.r, c = plan.shape.c;
    const int A = 2 * plan.shape.R;
    int a = 0;
    while ((1 << a) < A) ++a;
    const uint32_t all = (uint32_t)((1u << r) - 1);
    const uint32_t forbid = (r - c >= a) ? (uint32_t)((1u << c) - 1) : 0u;
    plan.rops.clear();
    plan.maps.clear();
    plan.mems.clear();
    plan.max_phases = 0;

    for (Pass &ps : plan.passes) {
        const DevOp *ops = plan.ops.data() + ps.op_begin;
        const int64_t n = ps.op_count;
        std::vector<uint32_t> need(n), wr(n), rd(n);
        for (int64_t i = 0; i < n; ++i) {
            const DevOp &o = ops[i];
            if (o.kind == K_DENSE || o.kind == K_ANTI) {
                need[i] = wr[i] = 1u << o.tbit;
                rd[i] = o.lmask;
            } else if (o.kind == K_DIAG) {
                // a phase in front of a rotation goes to the rotation's register phase
                need[i] = wr[i] = plan.pre_phase[ps.op_begin + i] ? (1u << o.tbit) : 0;
                rd[i] = (1u << o.tbit) | o.lmask;
            } else {
                need[i] = wr[i] = 0;
                rd[i] = o.lmask;
            }
        }
        std::vector<uint32_t> sets;
        std::vector<std::vector<int64_t>> lists;
        std::vector<char> done((size_t)n, 0);
        int64_t head = 0;
        while (head < n) {
            const bool first = sets.empty();
            uint32_t S = 0, bw = 0, br = 0;
            int nS = 0;
            std::vector<int64_t> list;
            for (int64_t k = head; k < n; ++k) {
                if (done[k]) continue;
                bool ok = !(wr[k] & (bw | br)) && !(rd[k] & bw);
                if (ok && (need[k] & ~S)) {
                    if (nS < a && !(first && (need[k] & forbid))) { S |= need[k]; ++nS; }
                    else ok = false;
                }
                if (ok) { list.push_back(k); done[k] = 1; }
                else { bw |= wr[k]; br |= rd[k]; if (bw == all) break; }
            }
            while (head < n && done[head]) ++head;
            sets.push_back(S);
            lists.push_back(std::move(list));
        }
        if (sets.empty() || (sets.back() & forbid)) { sets.push_back(0); lists.emplace_back(); }
        // complete every assignment to a register bits, highest free bits first
        // (never the coalesced ones in the first and last phase)
        for (size_t i = 0; i < sets.size(); ++i) {
            const bool edge = (i == 0 || i + 1 == sets.size());
            for (int b = r - 1; b >= 0 && popc(sets[i]) < a; --b)
                if (!((sets[i] >> b) & 1) && !(edge && ((forbid >> b) & 1))) sets[i] |= 1u << b;
            for (int b = r - 1; b >= 0 && popc(sets[i]) < a; --b) sets[i] |= 1u << b;
        }
        const int32_t map0 = (int32_t)plan.maps.size();
        const PhaseMap first_map = make_map(sets.front(), r), last_map = make_map(sets.back(), r);
        for (uint32_t S : sets) plan.maps.push_back(swizzled(make_map(S, r)));

        ps.rop_begin = (int64_t)plan.rops.size();
        for (size_t i = 0; i < sets.size(); ++i) {
            const uint32_t S = sets[i];
            if (i > 0) {
                RegOp o;
                std::memset(&o, 0, sizeof(o));
                o.kind = R_REMAP;
                o.rbit = map0 + (int32_t)i;
                plan.rops.push_back(o);
            }
            for (int64_t k : lists[i]) {
                const DevOp &d = ops[k];
                if (d.kind == K_DENSE || d.kind == K_ANTI) {
                    RegOp o;
                    std::memset(&o, 0, sizeof(o));
                    o.kind = d.kind == K_DENSE ? R_DENSE : R_ANTI;
                    if (d.kind == K_ANTI && is_one(d.m[2], d.m[3]) && is_one(d.m[4], d.m[5])) o.kind = R_SWAP;
                    o.rbit = popc(S & ((1u << d.tbit) - 1));
                    split_cond(d.lmask, d.lval, S, all, o);
                    // conditioned anti-diagonal gates run as dense ones: the
                    // kernel resolves conditions by selecting matrix or identity
                    if (o.kind != R_DENSE && (o.rmask || o.tmask)) o.kind = R_DENSE;
                    if (o.kind == R_DENSE && d.m[1] == 0.0 && d.m[3] == 0.0 && d.m[5] == 0.0 && d.m[7] == 0.0)
                        o.kind = R_DENSE_REAL;
                    o.gmask = d.gmask;
                    o.gval = d.gval;
                    std::memcpy(o.m, d.m, sizeof(o.m));
                    int p0, p1;
                    if (d.kind == K_DENSE && !o.rmask && !o.tmask && unit_pivot(d.m, 0, &p0) &&
                        unit_pivot(d.m, 1, &p1)) {
                        const double *t0 = d.m + 2 * (1 - p0), *t1 = d.m + 4 + 2 * (1 - p1);
                        o.kind = (t0[1] == 0.0 && t1[1] == 0.0) ? R_DENSE_PIV_REAL : R_DENSE_PIV;
                        o.tval = (uint32_t)(p0 | (p1 << 1));
                        const double m4[4] = {t0[0], t0[1], t1[0], t1[1]};
                        std::memset(o.m, 0, sizeof(o.m));
                        std::memcpy(o.m, m4, sizeof(m4));
                    }
                    plan.rops.push_back(o);
                    continue;
                }
                for (int v = 0; v < 2; ++v) {   // diagonal: one conditional phase per target value
                    const double fr = d.m[6 * v], fi = d.m[6 * v + 1];
                    if (is_one(fr, fi)) continue;
                    RegOp o;
                    std::memset(&o, 0, sizeof(o));
                    o.kind = fi == 0.0 ? R_PHASE_REAL : R_PHASE;
                    uint32_t lm = d.lmask, lv = d.lval;
                    o.gmask = d.gmask;
                    o.gval = d.gval;
                    if (d.kind == K_DIAG) { lm |= 1u << d.tbit; lv |= uint32_t(v) << d.tbit; }
                    else { o.gmask |= uint64_t(1) << d.tbit; o.gval |= uint64_t(v) << d.tbit; }
                    split_cond(lm, lv, S, all, o);
                    o.m[0] = fr;
                    o.m[1] = fi;
                    plan.rops.push_back(o);
                }
            }
        }
        ps.rop_count = (int64_t)plan.rops.size() - ps.rop_begin;
        ps.phases = (int32_t)sets.size();
        if (ps.phases > plan.max_phases) plan.max_phases = ps.phases;

        PassMem pm;
        std::memset(&pm, 0, sizeof(pm));
        const PhaseMap &f = first_map, &l = last_map;
        for (int h = 0; h < 2; ++h)
            for (int x = 0; x < 32; ++x) {
                pm.lthr[h][x] = to_state(f.thr[h][x], ps);
                pm.sthr[h][x] = to_state(l.thr[h][x], ps);
            }
        for (int k = 0; k < 64; ++k) {
            pm.lreg[k] = to_state(f.reg[k], ps);
            pm.sreg[k] = to_state(l.reg[k], ps);
        }
        pm.first_map = map0;
        pm.last_map = (int32_t)plan.maps.size() - 1;
        ps.mem = (int64_t)plan.mems.size();
        plan.mems.push_back(pm);
    }
}

}  // namespace hgm
