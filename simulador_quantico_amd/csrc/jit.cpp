// Plan-time specialisation of the register kernel (see jit.hpp).
#include "jit.hpp"

#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>

namespace hgm {

namespace {

uint64_t fnv1a(const char *p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) { h ^= (unsigned char)p[i]; h *= 1099511628211ull; }
    return h;
}
uint64_t fnv1a(const std::string &s) { return fnv1a(s.data(), s.size()); }

std::string hex64(uint64_t v) {
    char b[32];
    std::snprintf(b, sizeof b, "0x%llxull", (unsigned long long)v);
    return b;
}

std::string hex32(uint32_t v) {
    char b[24];
    std::snprintf(b, sizeof b, "0x%xu", v);
    return b;
}

struct Gen {
    bool f32;
    bool split;   // remaps move real parts, then imaginary parts, through an LDS tile of scalars
    std::string V, T;

    double round(double v) const { return f32 ? (double)(float)v : v; }

    std::string lit(double v) const {
        char b[64];
        if (f32) std::snprintf(b, sizeof b, "%af", (double)(float)v);
        else std::snprintf(b, sizeof b, "%a", v);
        return b;
    }

    // sum of coef * var, skipping zero coefficients and folding +-1. Terms
    // with coefficient +-1 come first: the sum is evaluated left to right, so
    // u + a * v + b * w contracts to two FMAs, while a * v + b * w + u needs a
    // multiply, an FMA and an add.
    std::string lin(const std::vector<std::pair<double, std::string>> &terms) const {
        std::string s;
        for (int unit = 1; unit >= 0; --unit)
            for (const auto &tv : terms) {
                const double c = round(tv.first);
                if (c == 0.0 || (c == 1.0 || c == -1.0) != (unit == 1)) continue;
                if (c == 1.0) s += (s.empty() ? "" : " + ") + tv.second;
                else if (c == -1.0) s += (s.empty() ? "-" : " - ") + tv.second;
                else s += (s.empty() ? "" : " + ") + lit(c) + " * " + tv.second;
            }
        return s.empty() ? lit(0.0) : s;
    }
};

// Expression of a thread-dependent quantity built from single-bit
// contributions: OR (state offsets) of ((t >> j) & 1) << bit_j, grouping runs.
std::string thread_or(const uint64_t *contrib, int nbits) {
    std::string s;
    int j = 0;
    while (j < nbits) {
        if (contrib[j] == 0) { ++j; continue; }
        int g = __builtin_ctzll(contrib[j]);
        int len = 1;
        while (j + len < nbits && contrib[j + len] == (uint64_t(1) << (g + len))) ++len;
        char b[128];
        std::snprintf(b, sizeof b, "(((unsigned long long)(t >> %d) & 0x%llxull) << %d)", j,
                      (unsigned long long)((1ull << len) - 1), g);
        s += (s.empty() ? "" : " | ") + std::string(b);
        j += len;
    }
    return s.empty() ? "0ull" : s;
}

// Sum of ((t >> j) & 1) * contrib_j.
std::string thread_sum(const uint32_t *contrib, int nbits) {
    std::string s;
    for (int j = 0; j < nbits; ++j) {
        if (!contrib[j]) continue;
        char b[96];
        std::snprintf(b, sizeof b, "(((t >> %d) & 1u) * %uu)", j, contrib[j]);
        s += (s.empty() ? "" : " + ") + std::string(b);
    }
    return s.empty() ? "0u" : s;
}

// The interpreter's phase maps hold XOR-swizzled tile indices; the swizzle is
// its own inverse.
uint32_t unswizzle(uint32_t x) { return x ^ ((x >> 4) & 15u) ^ ((x >> 8) & 15u) ^ ((x >> 12) & 15u); }

// LDS slot of tile index x in the generated kernels: one padding slot per 16
// and per 256 entries. Additive over disjoint bit sets, so a slot is a
// per-thread part plus a per-register constant (an immediate offset).
uint32_t padded(uint32_t x) { return x + (x >> 4) + (x >> 8); }

// Amplitude types of the generated kernels and their state accessors. A pass
// of the plan-compiled kernels reads and writes every amplitude once and
// nothing re-reads it before the next pass: their loads and stores carry the
// nontemporal (streaming) hint (random30 FP64 99.4-99.9 -> 96.7-97.1 ms per
// step, random32 FP32 221 -> 215 ms, on one box). The serving kernels do not:
// with the hint the mix served 1086-1089 against 1093 circuits/s, and with it
// on states above 256 MiB only (26 qubits) 1066-1068 against 1088-1090.
const char *const kAmpTypes =
    "struct __attribute__((aligned(16))) hgm_d2 { double x, y; };\n"
    "struct __attribute__((aligned(8))) hgm_f2 { float x, y; };\n"
    "typedef double hgm_v2d __attribute__((ext_vector_type(2)));\n"
    "typedef float hgm_v2f __attribute__((ext_vector_type(2)));\n"
    "__device__ inline hgm_d2 hgm_ld(const hgm_d2 *p) {\n"
    "    hgm_v2d v = __builtin_nontemporal_load((const hgm_v2d *)p); return {v.x, v.y}; }\n"
    "__device__ inline hgm_f2 hgm_ld(const hgm_f2 *p) {\n"
    "    hgm_v2f v = __builtin_nontemporal_load((const hgm_v2f *)p); return {v.x, v.y}; }\n"
    "__device__ inline void hgm_st(hgm_d2 *p, hgm_d2 a) {\n"
    "    hgm_v2d v = {a.x, a.y}; __builtin_nontemporal_store(v, (hgm_v2d *)p); }\n"
    "__device__ inline void hgm_st(hgm_f2 *p, hgm_f2 a) {\n"
    "    hgm_v2f v = {a.x, a.y}; __builtin_nontemporal_store(v, (hgm_v2f *)p); }\n";

// base = deposit of the tile id into the non-local state bits, by runs
std::string base_expr(uint64_t nonlocal) {
    std::string s;
    int q = 0;
    for (int p = 0; p < 64;) {
        if (!((nonlocal >> p) & 1)) { ++p; continue; }
        int len = 0;
        while (p + len < 64 && ((nonlocal >> (p + len)) & 1)) ++len;
        char b[128];
        std::snprintf(b, sizeof b, "(((unsigned long long)tid >> %d) & 0x%llxull) << %d", q,
                      (unsigned long long)(len >= 64 ? ~0ull : ((1ull << len) - 1)), p);
        s += (s.empty() ? "" : " | ") + std::string("(") + b + ")";
        q += len;
        p += len;
    }
    return s.empty() ? "0ull" : s;
}

std::string reg(int k) { return "a" + std::to_string(k); }

// The 2x2 matrix an op applies to its register pair (row-major, re/im), with
// the exact 1s of the row-scaled kinds written back in.
void op_matrix(const RegOp &op, double m[8]) {
    if (op.kind == R_DENSE_PIV || op.kind == R_DENSE_PIV_REAL) {
        std::fill(m, m + 8, 0.0);
        for (int row = 0; row < 2; ++row) {
            const int p = (int)((op.tval >> row) & 1u);
            m[4 * row + 2 * p] = 1.0;
            m[4 * row + 2 * (1 - p)] = op.m[2 * row];
            m[4 * row + 2 * (1 - p) + 1] = op.m[2 * row + 1];
        }
        return;
    }
    std::copy(op.m, op.m + 8, m);
}

// One op on the register pair (k0, k1). The generated code keeps each
// register as its true amplitude divided by a real scale known at generation
// time (scale[k], 1 unless a gathered phase left one): the op's matrix is
// applied to the true values, M * diag(s0, s1), and each output row is
// divided by the scale that makes its largest-weight exact +-1 entry stay
// exactly 1 (or by 1), which becomes the output register's scale. An op the
// code applies under a run-time condition keeps each register's scale, so
// that the registers mean the same whichever branch a thread takes.
void emit_pair(const Gen &g, std::string &o, const RegOp &op, int k0, int k1, double *scale, bool runtime_cond) {
    double m[8];
    op_matrix(op, m);
    const double s[2] = {scale[k0], scale[k1]};
    o += "        { const " + g.V + " x = " + reg(k0) + ", y = " + reg(k1) + ";\n";
    if (op.kind == R_SWAP && !runtime_cond) {   // a renaming; the scales follow the registers
        o += "          " + reg(k0) + " = y; " + reg(k1) + " = x; }\n";
        std::swap(scale[k0], scale[k1]);
        return;
    }
    double u[2];
    for (int row = 0; row < 2; ++row) {
        if (runtime_cond) { u[row] = s[row]; continue; }
        u[row] = 1.0;
        for (int col = 0; col < 2; ++col) {
            const double re = m[4 * row + 2 * col], im = m[4 * row + 2 * col + 1];
            if (im == 0.0 && (re == 1.0 || re == -1.0)) { u[row] = re * s[col]; break; }
        }
    }
    double w[8];
    for (int row = 0; row < 2; ++row)
        for (int col = 0; col < 2; ++col)
            for (int c = 0; c < 2; ++c) w[4 * row + 2 * col + c] = m[4 * row + 2 * col + c] * s[col] / u[row];
    // exact units stay exact (the division above can leave 1 - 2^-53)
    for (int row = 0; row < 2; ++row) {
        for (int col = 0; col < 2; ++col) {
            const double re = m[4 * row + 2 * col], im = m[4 * row + 2 * col + 1];
            if (im == 0.0 && (re == 1.0 || re == -1.0) && re * s[col] == u[row]) {
                w[4 * row + 2 * col] = 1.0;
                w[4 * row + 2 * col + 1] = 0.0;
            }
        }
    }
    scale[k0] = u[0];
    scale[k1] = u[1];
    o += "          " + reg(k0) + ".x = " +
         g.lin({{w[0], "x.x"}, {-w[1], "x.y"}, {w[2], "y.x"}, {-w[3], "y.y"}}) + ";\n";
    o += "          " + reg(k0) + ".y = " + g.lin({{w[0], "x.y"}, {w[1], "x.x"}, {w[2], "y.y"}, {w[3], "y.x"}}) + ";\n";
    o += "          " + reg(k1) + ".x = " +
         g.lin({{w[4], "x.x"}, {-w[5], "x.y"}, {w[6], "y.x"}, {-w[7], "y.y"}}) + ";\n";
    o += "          " + reg(k1) + ".y = " + g.lin({{w[4], "x.y"}, {w[5], "x.x"}, {w[6], "y.y"}, {w[7], "y.x"}}) +
         "; }\n";
}

void emit_factor(const Gen &g, std::string &o, double fr, double fi, int k) {
    fr = g.round(fr);
    fi = g.round(fi);
    const std::string a = reg(k);
    if (fi == 0.0) {
        if (fr == 1.0) return;
        if (fr == -1.0) { o += "        " + a + ".x = -" + a + ".x; " + a + ".y = -" + a + ".y;\n"; return; }
        o += "        " + a + ".x *= " + g.lit(fr) + "; " + a + ".y *= " + g.lit(fr) + ";\n";
        return;
    }
    o += "        { const " + g.V + " x = " + a + ";\n";
    o += "          " + a + ".x = " + g.lin({{fr, "x.x"}, {-fi, "x.y"}}) + ";\n";
    o += "          " + a + ".y = " + g.lin({{fr, "x.y"}, {fi, "x.x"}}) + "; }\n";
}

void emit_phase(const Gen &g, std::string &o, const RegOp &op, int k) { emit_factor(g, o, op.m[0], op.m[1], k); }

bool is_phase(const RegOp &o) { return o.kind == R_PHASE || o.kind == R_PHASE_REAL; }

// A phase that tests register bits only: its factor per register slot is a
// constant of the generated code.
bool slot_phase(const RegOp &o) { return is_phase(o) && !o.tmask && !o.gmask; }

// Whether two register-program ops commute: phases always do; a phase and a
// gate on register bit b do when the phase does not depend on b; two gates do
// when their register bits differ and neither tests the other's bit.
bool commute(const RegOp &a, const RegOp &b) {
    const bool pa = is_phase(a), pb = is_phase(b);
    if (pa && pb) return true;
    if (pa) return !((a.rmask >> b.rbit) & 1);
    if (pb) return !((b.rmask >> a.rbit) & 1);
    return a.rbit != b.rbit && !((a.rmask >> b.rbit) & 1) && !((b.rmask >> a.rbit) & 1);
}

// Reorders the ops between two remaps into waves: every phase whose
// predecessors (the earlier ops it does not commute with) are done, then
// every gate whose predecessors are done, and again. The phases in front of
// a wave of rotations then form one run, and a run costs one multiply per
// register slot. The order stays a topological order of the non-commuting
// pairs, so the product is unchanged.
void schedule_waves(std::vector<RegOp> &ops) {
    size_t a = 0;
    while (a < ops.size()) {
        size_t b = a;
        while (b < ops.size() && ops[b].kind != R_REMAP) ++b;
        const size_t n = b - a;
        if (n > 2) {
            std::vector<std::vector<uint32_t>> succ(n);
            std::vector<uint32_t> preds(n, 0);
            for (size_t i = 0; i < n; ++i)
                for (size_t j = i + 1; j < n; ++j)
                    if (!commute(ops[a + i], ops[a + j])) { succ[i].push_back((uint32_t)j); ++preds[j]; }
            std::vector<RegOp> out;
            out.reserve(n);
            std::vector<char> done(n, 0);
            size_t emitted = 0;
            while (emitted < n) {
                for (int want_phase = 1; want_phase >= 0; --want_phase) {
                    bool again = true;
                    while (again) {
                        again = false;
                        for (size_t i = 0; i < n; ++i) {
                            if (done[i] || preds[i] || is_phase(ops[a + i]) != (want_phase == 1)) continue;
                            done[i] = 1;
                            ++emitted;
                            out.push_back(ops[a + i]);
                            for (uint32_t j : succ[i]) if (--preds[j] == 0 && j < i) again = true;
                        }
                    }
                }
            }
            std::copy(out.begin(), out.end(), ops.begin() + a);
        }
        a = b + 1;
    }
}

std::string kernel_body(const Gen &g, const Plan &plan, const Pass &ps, const std::string &name) {
    const int r = plan.shape.r, B = plan.shape.B, A = 2 * plan.shape.R;
    int tb = 0;
    while ((1 << tb) < B) ++tb;
    const PassMem &pm = plan.mems[ps.mem];
    const uint64_t all = (plan.qubits >= 64) ? ~0ull : ((1ull << plan.qubits) - 1);
    const uint64_t nonlocal = all & ~ps.local_mask;
    const int64_t n_tiles = int64_t(1) << (plan.qubits - r);

    uint64_t lt[kMaxThreadBits] = {0}, st[kMaxThreadBits] = {0};
    for (int j = 0; j < tb; ++j) {
        lt[j] = pm.lthr[j / 5][1 << (j % 5)];
        st[j] = pm.sthr[j / 5][1 << (j % 5)];
    }
    std::string o;
    o += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(B) + ") " + name + "(" + g.V +
         " *__restrict__ s) {\n";
    if (ps.phases > 1) o += "    extern __shared__ " + (g.split ? g.T : g.V) + " lds[];\n";
    o += "    const unsigned t = threadIdx.x;\n";
    o += "    const unsigned long long lt = " + thread_or(lt, tb) + ";\n";
    o += "    const unsigned long long st = " + thread_or(st, tb) + ";\n";
    // thread parts of the LDS slot of every phase map of this pass
    const int m_first = pm.first_map, m_last = pm.last_map;
    if (ps.phases > 1)
        for (int mi = m_first; mi <= m_last; ++mi) {
            uint32_t c[kMaxThreadBits] = {0};
            for (int j = 0; j < tb; ++j) c[j] = padded(unswizzle(plan.maps[mi].thr[j / 5][1 << (j % 5)]));
            o += "    const unsigned f" + std::to_string(mi - m_first) + " = " + thread_sum(c, tb) + ";\n";
        }
    // one tile per workgroup: no loop, so the matrix literals are not kept
    // live across tiles
    (void)n_tiles;
    o += "    {\n";
    o += "        const unsigned long long tid = blockIdx.x;\n";
    o += "        const unsigned long long base = " + base_expr(nonlocal) + ";\n";
    for (int k = 0; k < A; ++k)
        o += "        " + g.V + " " + reg(k) + " = hgm_ld(&s[(base | " + hex64(pm.lreg[k]) + ") + lt]);\n";
    int cur = m_first;
    std::vector<RegOp> ops(plan.rops.begin() + ps.rop_begin, plan.rops.begin() + ps.rop_begin + ps.rop_count);
    schedule_waves(ops);
    std::vector<std::complex<double>> factor((size_t)A, std::complex<double>(1, 0));
    std::vector<double> scale((size_t)A, 1.0);   // register k holds (true amplitude) / scale[k]
    bool pending = false;
    // The product z of a run of slot phases, once per register slot, as
    // z = a * (1 + i t) with |t| <= 1 when |Re z| >= |Im z|, else
    // z = b * (t + i) with b = Im z: two multiply-adds per register, and the
    // real a or b joins the register's scale.
    auto flush_factors = [&]() {
        if (!pending) return;
        for (int k = 0; k < A; ++k) {
            const std::complex<double> z = factor[k];
            if (z == std::complex<double>(1, 0)) continue;
            const std::string a = reg(k);
            if (std::abs(z.real()) >= std::abs(z.imag())) {
                scale[k] *= z.real();
                const double t = z.imag() / z.real();
                if (g.round(t) == 0.0) continue;
                o += "        { const " + g.V + " x = " + a + ";\n";
                o += "          " + a + ".x = " + g.lin({{1.0, "x.x"}, {-t, "x.y"}}) + ";\n";
                o += "          " + a + ".y = " + g.lin({{1.0, "x.y"}, {t, "x.x"}}) + "; }\n";
            } else {
                scale[k] *= z.imag();
                const double t = z.real() / z.imag();
                o += "        { const " + g.V + " x = " + a + ";\n";
                o += "          " + a + ".x = " + g.lin({{-1.0, "x.y"}, {t, "x.x"}}) + ";\n";
                o += "          " + a + ".y = " + g.lin({{1.0, "x.x"}, {t, "x.y"}}) + "; }\n";
            }
        }
        std::fill(factor.begin(), factor.end(), std::complex<double>(1, 0));
        pending = false;
    };
    // registers back to their true amplitudes (before LDS exchanges and the store)
    auto flush_scales = [&]() {
        for (int k = 0; k < A; ++k) {
            const double sc = g.round(scale[k]);
            scale[k] = 1.0;
            if (sc == 1.0) continue;
            emit_factor(g, o, sc, 0.0, k);
        }
    };
    for (int64_t i = 0; i < ps.rop_count; ++i) {
        const RegOp &op = ops[i];
        if (slot_phase(op)) {
            for (int k = 0; k < A; ++k)
                if (((uint32_t)k & op.rmask) == op.rval) factor[k] *= std::complex<double>(op.m[0], op.m[1]);
            pending = true;
            continue;
        }
        if (!is_phase(op)) flush_factors();   // other phases commute with the pending factors
        if (op.kind == R_REMAP) {
            flush_scales();
            const int nxt = op.rbit;
            const char *parts[2][2] = {{"", nullptr}, {".x", ".y"}};
            for (int h = 0; h < 2 && parts[g.split][h]; ++h) {
                const std::string part = parts[g.split][h];
                o += "        __syncthreads();\n";
                for (int k = 0; k < A; ++k)
                    o += "        lds[f" + std::to_string(cur - m_first) + " + " +
                         std::to_string(padded(unswizzle(plan.maps[cur].reg[k]))) + "u] = " + reg(k) + part + ";\n";
                o += "        __syncthreads();\n";
                for (int k = 0; k < A; ++k)
                    o += "        " + reg(k) + part + " = lds[f" + std::to_string(nxt - m_first) + " + " +
                         std::to_string(padded(unswizzle(plan.maps[nxt].reg[k]))) + "u];\n";
            }
            cur = nxt;
            continue;
        }
        std::string body;
        if (op.kind == R_PHASE || op.kind == R_PHASE_REAL) {
            for (int k = 0; k < A; ++k)
                if (((uint32_t)k & op.rmask) == op.rval) emit_phase(g, body, op, k);
        } else {
            const bool runtime_cond = op.tmask || op.gmask;
            for (int k0 = 0; k0 < A; ++k0) {
                if ((k0 >> op.rbit) & 1) continue;
                if (((uint32_t)k0 & op.rmask) != op.rval) continue;
                emit_pair(g, body, op, k0, k0 | (1 << op.rbit), scale.data(), runtime_cond);
            }
        }
        if (body.empty()) continue;
        std::string cond;
        if (op.gmask) cond += "(base & " + hex64(op.gmask) + ") == " + hex64(op.gval);
        if (op.tmask) cond += std::string(cond.empty() ? "" : " && ") + "(t & " + hex32(op.tmask) + ") == " + hex32(op.tval);
        if (cond.empty()) o += body;
        else o += "        if (" + cond + ") {\n" + body + "        }\n";
    }
    flush_factors();
    flush_scales();
    for (int k = 0; k < A; ++k)
        o += "        hgm_st(&s[(base | " + hex64(pm.sreg[k]) + ") + st], " + reg(k) + ");\n";
    o += "    }\n}\n";
    return o;
}

// Everything kernel_body reads for one pass, as bytes: passes with equal keys
// (the repeated iterations of a circuit) have equal bodies, so only the first
// is generated. Map indices are made relative to the pass's first map.
std::string pass_key(const Plan &plan, const Pass &ps) {
    const PassMem &pm = plan.mems[ps.mem];
    std::string k;
    auto put = [&](const void *p, size_t n) { k.append(static_cast<const char *>(p), n); };
    put(&ps.local_mask, sizeof(ps.local_mask));
    put(&ps.phases, sizeof(ps.phases));
    PassMem m = pm;
    m.first_map = 0;
    m.last_map = pm.last_map - pm.first_map;
    put(&m, sizeof(m));
    for (int32_t mi = pm.first_map; mi <= pm.last_map; ++mi) put(&plan.maps[mi], sizeof(PhaseMap));
    for (int64_t i = 0; i < ps.rop_count; ++i) {
        RegOp o = plan.rops[ps.rop_begin + i];
        if (o.kind == R_REMAP) o.rbit -= pm.first_map;
        put(&o, sizeof(o));
    }
    return k;
}

struct Compiled {
    std::string source;          // compared on lookup: the hash only narrows the search
    std::vector<char> binary;
    uint64_t binary_hash;
    double seconds;
    uint64_t last_use;
};

// Compiled code objects by source hash, least recently used first out beyond
// kMaxCompiled (an autotune sweep or a stream of distinct circuits would
// otherwise keep every source and binary for the life of the process).
constexpr size_t kMaxCompiled = 64;
std::mutex g_cache_mu;
std::unordered_map<uint64_t, std::shared_ptr<Compiled>> g_cache;
uint64_t g_cache_tick = 0;

}  // namespace

void jit_generate(const Plan &plan, JitCode *out) {
    Gen g;
    g.f32 = plan.dtype == HGM_C64;
    g.V = g.f32 ? "hgm_f2" : "hgm_d2";
    g.T = g.f32 ? "float" : "double";
    // LDS per workgroup caps the workgroups a CU holds (160 KiB each): a tile
    // above 32 KiB moves through LDS in two halves of scalars instead, which
    // doubles the resident workgroups for two more barriers per remap
    const uint64_t last = (uint64_t(1) << plan.shape.r) - 1;
    const size_t entries = (size_t)padded((uint32_t)last) + 1;
    g.split = entries * amp_bytes(plan.dtype) > 32 * 1024;
    out->lds_bytes = entries * amp_bytes(plan.dtype) / (g.split ? 2 : 1);
    out->source = std::string("// generated pass kernels\n") + kAmpTypes;
    out->pass_kernel.clear();
    out->names.clear();
    std::map<std::string, int> seen;                 // body (with a placeholder name) -> kernel index
    std::unordered_map<std::string, int> seen_key;   // pass_key -> kernel index
    for (const Pass &ps : plan.passes) {
        if (ps.op_count == 0) { out->pass_kernel.push_back(-1); continue; }
        std::string key = pass_key(plan, ps);
        auto kt = seen_key.find(key);
        if (kt != seen_key.end()) { out->pass_kernel.push_back(kt->second); continue; }
        std::string body = kernel_body(g, plan, ps, "@NAME@");
        auto it = seen.find(body);
        if (it != seen.end()) {
            seen_key.emplace(std::move(key), it->second);
            out->pass_kernel.push_back(it->second);
            continue;
        }
        const int idx = (int)out->names.size();
        const std::string name = "hgm_pass_" + std::to_string(idx);
        seen[body] = idx;
        seen_key.emplace(std::move(key), idx);
        out->names.push_back(name);
        out->pass_kernel.push_back(idx);
        body.replace(body.find("@NAME@"), 6, name);
        out->source += body;
    }
    out->hash = fnv1a(out->source);
}

int jit_compile(JitCode *out, std::string *err) {
    if (!out->hash) out->hash = fnv1a(out->source);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_cache.find(out->hash);
        if (it != g_cache.end() && it->second->source == out->source) {
            out->binary = it->second->binary;
            out->binary_hash = it->second->binary_hash;
            out->compile_s = 0;
            it->second->last_use = ++g_cache_tick;
            return HGM_OK;
        }
    }
    auto t0 = std::chrono::steady_clock::now();
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, out->source.c_str(), "hgm_passes.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        *err = "hiprtcCreateProgram failed";
        return HGM_E_HIP;
    }
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=fast",
                          // keep the compiler from emitting v_pk_{add,mul,fma}_f32
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        *err = "hipRTC compilation of the pass kernels failed: " + log.substr(0, 2000);
        hiprtcDestroyProgram(&prog);
        return HGM_E_HIP;
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    out->binary.assign(n, 0);
    hiprtcGetCode(prog, out->binary.data());
    hiprtcDestroyProgram(&prog);
    out->binary_hash = fnv1a(out->binary.data(), out->binary.size());
    out->compile_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    auto c = std::make_shared<Compiled>();
    c->source = out->source;
    c->binary = out->binary;
    c->binary_hash = out->binary_hash;
    c->seconds = out->compile_s;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    c->last_use = ++g_cache_tick;
    g_cache[out->hash] = c;
    while (g_cache.size() > kMaxCompiled) {
        auto lru = g_cache.begin();
        for (auto it = g_cache.begin(); it != g_cache.end(); ++it)
            if (it->second->last_use < lru->second->last_use) lru = it;
        g_cache.erase(lru);
    }
    return HGM_OK;
}

size_t jit_compiled_count() {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    return g_cache.size();
}

// ---- serving pass kernels ---------------------------------------------------

std::string serving_pass_key(const Plan &plan, const Pass &ps) {
    const PassMem &pm = plan.mems[ps.mem];
    std::string k;
    auto put = [&](const void *p, size_t n) { k.append(static_cast<const char *>(p), n); };
    const int32_t hdr[6] = {plan.qubits, plan.dtype, plan.shape.r, plan.shape.B, plan.shape.R, plan.shape.c};
    put(hdr, sizeof hdr);
    put(&ps.local_mask, sizeof(ps.local_mask));
    put(&ps.phases, sizeof(ps.phases));
    PassMem m = pm;
    m.first_map = 0;
    m.last_map = pm.last_map - pm.first_map;
    put(&m, sizeof(m));
    for (int32_t mi = pm.first_map; mi <= pm.last_map; ++mi) put(&plan.maps[mi], sizeof(PhaseMap));
    for (int64_t i = 0; i < ps.rop_count; ++i) {
        RegOp o = plan.rops[ps.rop_begin + i];
        std::memset(o.m, 0, sizeof(o.m));
        o.rval = 0;
        o.gval = 0;
        if (o.kind != R_DENSE_PIV && o.kind != R_DENSE_PIV_REAL) o.tval = 0;   // the pivot columns are code
        if (o.kind == R_REMAP) o.rbit -= pm.first_map;
        put(&o, sizeof(o));
    }
    return k;
}

namespace {

// Code of one op on the register pair (k0, k1) of a serving kernel; the
// coefficients are the op's locals m0..m7.
std::string serving_pair(const Gen &g, const RegOp &op, int k0, int k1) {
    const std::string a0 = reg(k0), a1 = reg(k1);
    std::string o = "        { const " + g.V + " x = " + a0 + ", y = " + a1 + ";\n";
    switch (op.kind) {
        case R_DENSE:
            o += "          " + a0 + ".x = m0 * x.x - m1 * x.y + m2 * y.x - m3 * y.y;\n";
            o += "          " + a0 + ".y = m0 * x.y + m1 * x.x + m2 * y.y + m3 * y.x;\n";
            o += "          " + a1 + ".x = m4 * x.x - m5 * x.y + m6 * y.x - m7 * y.y;\n";
            o += "          " + a1 + ".y = m4 * x.y + m5 * x.x + m6 * y.y + m7 * y.x; }\n";
            return o;
        case R_DENSE_REAL:
            o += "          " + a0 + ".x = m0 * x.x + m2 * y.x; " + a0 + ".y = m0 * x.y + m2 * y.y;\n";
            o += "          " + a1 + ".x = m4 * x.x + m6 * y.x; " + a1 + ".y = m4 * x.y + m6 * y.y; }\n";
            return o;
        case R_ANTI:
            o += "          " + a0 + ".x = m2 * y.x - m3 * y.y; " + a0 + ".y = m2 * y.y + m3 * y.x;\n";
            o += "          " + a1 + ".x = m4 * x.x - m5 * x.y; " + a1 + ".y = m4 * x.y + m5 * x.x; }\n";
            return o;
        case R_SWAP:
            o += "          " + a0 + " = y; " + a1 + " = x; }\n";
            return o;
        default: {   // R_DENSE_PIV, R_DENSE_PIV_REAL: out_i = v[p_i] + t_i * v[1 - p_i]
            const char *v[2] = {"x", "y"};
            const int p0 = (int)(op.tval & 1u), p1 = (int)(op.tval >> 1);
            const std::string u0 = v[p0], w0 = v[1 - p0], u1 = v[p1], w1 = v[1 - p1];
            if (op.kind == R_DENSE_PIV_REAL) {
                o += "          " + a0 + ".x = " + u0 + ".x + m0 * " + w0 + ".x; " + a0 + ".y = " + u0 + ".y + m0 * " + w0 +
                     ".y;\n";
                o += "          " + a1 + ".x = " + u1 + ".x + m2 * " + w1 + ".x; " + a1 + ".y = " + u1 + ".y + m2 * " + w1 +
                     ".y; }\n";
            } else {
                o += "          " + a0 + ".x = " + u0 + ".x + m0 * " + w0 + ".x - m1 * " + w0 + ".y;\n";
                o += "          " + a0 + ".y = " + u0 + ".y + m0 * " + w0 + ".y + m1 * " + w0 + ".x;\n";
                o += "          " + a1 + ".x = " + u1 + ".x + m2 * " + w1 + ".x - m3 * " + w1 + ".y;\n";
                o += "          " + a1 + ".y = " + u1 + ".y + m2 * " + w1 + ".y + m3 * " + w1 + ".x; }\n";
            }
            return o;
        }
    }
}

}  // namespace

std::string serving_kernel_source(const Plan &plan, const Pass &ps, const std::string &name, size_t *lds_bytes) {
    Gen g;
    g.f32 = plan.dtype == HGM_C64;
    g.V = g.f32 ? "hgm_f2" : "hgm_d2";
    g.T = g.f32 ? "float" : "double";
    const uint64_t last = (uint64_t(1) << plan.shape.r) - 1;
    const size_t entries = (size_t)padded((uint32_t)last) + 1;
    g.split = entries * amp_bytes(plan.dtype) > 32 * 1024;
    *lds_bytes = ps.phases > 1 ? entries * amp_bytes(plan.dtype) / (g.split ? 2 : 1) : 0;
    const int B = plan.shape.B, A = 2 * plan.shape.R;
    int tb = 0;
    while ((1 << tb) < B) ++tb;
    const PassMem &pm = plan.mems[ps.mem];
    const uint64_t all = (plan.qubits >= 64) ? ~0ull : ((1ull << plan.qubits) - 1);
    const uint64_t nonlocal = all & ~ps.local_mask;
    uint64_t lt[kMaxThreadBits] = {0}, st[kMaxThreadBits] = {0};
    for (int j = 0; j < tb; ++j) {
        lt[j] = pm.lthr[j / 5][1 << (j % 5)];
        st[j] = pm.sthr[j / 5][1 << (j % 5)];
    }
    std::string o;
    o += std::string("// generated serving pass kernel\n") + kAmpTypes +
         "struct hgm_item { unsigned long long state_off, ops_off, maps_off, pm_off, nonlocal_mask;\n"
         "                  long long tile_begin, partials_off; int n_ops, pad; };\n"
         "struct hgm_rop { int kind, rbit; unsigned rmask, rval, tmask, tval; unsigned long long gmask, gval;\n"
         "                 double m[8]; };\n"
         "struct hgm_part { double sum, pmax; unsigned long long idx, pad; };\n"
         "static_assert(sizeof(hgm_item) == 64 && sizeof(hgm_rop) == 104 && sizeof(hgm_part) == 32, \"layouts\");\n";
    o += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(B) + ") " + name +
         "(const hgm_item *__restrict__ items, int n_items, const char *__restrict__ progs,\n"
         "        char *__restrict__ states, char *__restrict__ parts) {\n";
    if (ps.phases > 1) o += "    extern __shared__ " + (g.split ? g.T : g.V) + " lds[];\n";
    o += "    __shared__ double red_s[16], red_m[16];\n"
         "    __shared__ unsigned long long red_i[16];\n"
         "    const unsigned t = threadIdx.x;\n"
         "    const long long b = blockIdx.x;\n"
         "    int lo = 0, hi = n_items - 1;\n"
         "    while (lo < hi) {\n"
         "        const int mid = (lo + hi + 1) >> 1;\n"
         "        if (items[mid].tile_begin <= b) lo = mid; else hi = mid - 1;\n"
         "    }\n"
         "    const hgm_item &it = items[lo];\n";
    o += "    " + g.V + " *__restrict__ s = (" + g.V + " *)(states + it.state_off);\n";
    o += "    const hgm_rop *__restrict__ ops = (const hgm_rop *)(progs + it.ops_off);\n"
         "    const unsigned long long tid = (unsigned long long)(b - it.tile_begin);\n";
    o += "    const unsigned long long base = " + base_expr(nonlocal) + ";\n";
    o += "    const unsigned long long lt = " + thread_or(lt, tb) + ";\n";
    o += "    const unsigned long long st = " + thread_or(st, tb) + ";\n";
    const int m_first = pm.first_map, m_last = pm.last_map;
    if (ps.phases > 1)
        for (int mi = m_first; mi <= m_last; ++mi) {
            uint32_t c[kMaxThreadBits] = {0};
            for (int j = 0; j < tb; ++j) c[j] = padded(unswizzle(plan.maps[mi].thr[j / 5][1 << (j % 5)]));
            o += "    const unsigned f" + std::to_string(mi - m_first) + " = " + thread_sum(c, tb) + ";\n";
        }
    for (int k = 0; k < A; ++k)
        o += "    " + g.V + " " + reg(k) + " = s[(base | " + hex64(pm.lreg[k]) + ") + lt];\n";
    int cur = m_first;
    for (int64_t i = 0; i < ps.rop_count; ++i) {
        const RegOp &op = plan.rops[ps.rop_begin + i];
        const std::string oi = "ops[" + std::to_string(i) + "]";
        if (op.kind == R_REMAP) {
            const int nxt = op.rbit;
            const char *parts[2][2] = {{"", nullptr}, {".x", ".y"}};
            for (int h = 0; h < 2 && parts[g.split][h]; ++h) {
                const std::string part = parts[g.split][h];
                o += "    __syncthreads();\n";
                for (int k = 0; k < A; ++k)
                    o += "    lds[f" + std::to_string(cur - m_first) + " + " +
                         std::to_string(padded(unswizzle(plan.maps[cur].reg[k]))) + "u] = " + reg(k) + part + ";\n";
                o += "    __syncthreads();\n";
                for (int k = 0; k < A; ++k)
                    o += "    " + reg(k) + part + " = lds[f" + std::to_string(nxt - m_first) + " + " +
                         std::to_string(padded(unswizzle(plan.maps[nxt].reg[k]))) + "u];\n";
            }
            cur = nxt;
            continue;
        }
        // the op's coefficients and control values are read from the item's
        // register program; its kind, register bit, masks and pivots are code
        std::string body = "    {\n";
        int nm = 0;
        switch (op.kind) {
            case R_DENSE: nm = 8; break;
            case R_DENSE_REAL: nm = 8; break;
            case R_ANTI: nm = 6; break;
            case R_PHASE: nm = 2; break;
            case R_PHASE_REAL: nm = 1; break;
            case R_DENSE_PIV: nm = 4; break;
            case R_DENSE_PIV_REAL: nm = 3; break;
            default: nm = 0;
        }
        for (int e = 0; e < nm; ++e)
            body += "        const " + g.T + " m" + std::to_string(e) + " = (" + g.T + ")" + oi + ".m[" + std::to_string(e) +
                    "];\n";
        if (op.rmask) body += "        const unsigned rv = " + oi + ".rval;\n";
        const bool phase = op.kind == R_PHASE || op.kind == R_PHASE_REAL;
        for (int k0 = 0; k0 < A; ++k0) {
            std::string one;
            if (phase) {
                const std::string a = reg(k0);
                if (op.kind == R_PHASE_REAL) one = "        " + a + ".x *= m0; " + a + ".y *= m0;\n";
                else
                    one = "        { const " + g.V + " x = " + a + "; " + a + ".x = m0 * x.x - m1 * x.y; " + a +
                          ".y = m0 * x.y + m1 * x.x; }\n";
            } else {
                if ((k0 >> op.rbit) & 1) continue;
                one = serving_pair(g, op, k0, k0 | (1 << op.rbit));
            }
            if (op.rmask) body += "        if ((" + std::to_string(k0) + "u & " + hex32(op.rmask) + ") == rv) {\n" + one + "        }\n";
            else body += one;
        }
        std::string cond;
        if (op.gmask) cond += "(base & " + hex64(op.gmask) + ") == " + oi + ".gval";
        if (op.tmask) cond += std::string(cond.empty() ? "" : " && ") + "(t & " + hex32(op.tmask) + ") == " + oi + ".tval";
        if (!cond.empty()) body = "    if (" + cond + ")\n" + body;
        o += body + "    }\n";
    }
    for (int k = 0; k < A; ++k)
        o += "    s[(base | " + hex64(pm.sreg[k]) + ") + st] = " + reg(k) + ";\n";
    // the last pass of a state reduces its readout per tile: FP64 sum of |a|^2
    // and the largest |a|^2 with the lowest index
    const int width = B < 64 ? B : 64, waves = (B + 63) / 64;
    o += "    if (it.partials_off < 0) return;\n"
         "    double sum = 0.0, best = -1.0;\n"
         "    unsigned long long bi = ~0ull;\n";
    for (int k = 0; k < A; ++k)
        o += "    { const double p = (double)" + reg(k) + ".x * (double)" + reg(k) + ".x + (double)" + reg(k) +
             ".y * (double)" + reg(k) + ".y;\n      const unsigned long long idx = (base | " + hex64(pm.sreg[k]) +
             ") | st;\n      sum += p;\n      if (p > best || (p == best && idx < bi)) { best = p; bi = idx; } }\n";
    o += "    for (int off = " + std::to_string(width >> 1) + "; off > 0; off >>= 1) {\n"
         "        sum += __shfl_xor(sum, off, " + std::to_string(width) + ");\n"
         "        const double ob = __shfl_xor(best, off, " + std::to_string(width) + ");\n"
         "        const unsigned long long oi = __shfl_xor(bi, off, " + std::to_string(width) + ");\n"
         "        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }\n"
         "    }\n"
         "    if ((t & 63u) == 0) { red_s[t >> 6] = sum; red_m[t >> 6] = best; red_i[t >> 6] = bi; }\n"
         "    __syncthreads();\n"
         "    if (t == 0) {\n"
         "        for (int w = 1; w < " + std::to_string(waves) + "; ++w) {\n"
         "            sum += red_s[w];\n"
         "            if (red_m[w] > best || (red_m[w] == best && red_i[w] < bi)) { best = red_m[w]; bi = red_i[w]; }\n"
         "        }\n"
         "        hgm_part r;\n"
         "        r.sum = sum; r.pmax = best; r.idx = bi; r.pad = 0;\n"
         "        ((hgm_part *)(parts + it.partials_off))[tid] = r;\n"
         "    }\n"
         "}\n";
    return o;
}

}  // namespace hgm
