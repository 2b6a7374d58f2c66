// Shor's algorithm on the t_GPU path: the modular-arithmetic circuit builders
// (src/algorithms/lib_shor_circuits.cpp:47-166), the classical number theory
// (src/algorithms/lib_shor_number_theory.cpp, src/core/common.cpp:209-219)
// and the semiclassical driver (src/algorithms/lib_shor.cpp:17-163).
//
// The builders emit the same steps in the same order as the reference and
// call the sub-builders in the same order, so the gate catalog (first writer
// wins) ends up with the same entries. The driver keeps the state resident on
// the device across its 2*width rounds; each round's measurement reads the
// probability on the device and collapses there.
#include <cmath>
#include <random>

#include "hgm_dgm.hpp"

namespace hgm {

namespace {

void append(std::vector<std::string> &dst, const std::vector<std::string> &src) {
    dst.insert(dst.end(), src.begin(), src.end());
}

}  // namespace

// |x> -> |(x + a) mod N> on the Fourier-basis register `reg` (with overflow
// qubit `over`), when both controls are 1; over_bool is the ancilla that
// records the overflow test (src/algorithms/lib_shor_circuits.cpp:100-130).
std::vector<std::string> C2AddMod(int qubits, int ctrl1, int ctrl2, int reg, int over, int over_bool, int width,
                                  long a, long N, Gates &g) {
    const std::vector<std::string> qft = QFT(qubits, reg, over, width, g);
    const std::vector<std::string> rqft = RQFT(qubits, reg, over, width, g);
    const std::string add_a = C2AddF(qubits, ctrl1, ctrl2, reg, over, a, width, g);
    const std::string sub_a = C2SubF(qubits, ctrl1, ctrl2, reg, over, a, width, g);
    const std::string sub_n = SubF(qubits, reg, over, N, width, g);
    const std::string cadd_n = CAddF(qubits, over_bool, reg, over, N, width, g);
    const std::string x_over = Pauli_X(qubits, over, 1);
    const std::string copy_over = CNot(qubits, over, over_bool);

    std::vector<std::string> s{add_a, sub_n};       // x + a - N
    append(s, rqft);
    s.push_back(copy_over);                         // sign bit -> ancilla
    append(s, qft);
    s.push_back(cadd_n);                            // add N back when negative
    s.push_back(sub_a);                             // uncompute the ancilla
    append(s, rqft);
    s.push_back(x_over);
    s.push_back(copy_over);
    s.push_back(x_over);
    append(s, qft);
    s.push_back(add_a);
    return s;
}

// The inverse of C2AddMod (src/algorithms/lib_shor_circuits.cpp:132-166).
std::vector<std::string> C2SubMod(int qubits, int ctrl1, int ctrl2, int reg, int over, int over_bool, int width,
                                  long a, long N, Gates &g) {
    const std::vector<std::string> qft = QFT(qubits, reg, over, width, g);
    const std::vector<std::string> rqft = RQFT(qubits, reg, over, width, g);
    const std::string add_a = C2AddF(qubits, ctrl1, ctrl2, reg, over, a, width, g);
    const std::string sub_a = C2SubF(qubits, ctrl1, ctrl2, reg, over, a, width, g);
    const std::string add_n = AddF(qubits, reg, over, N, width, g);
    (void)CAddF(qubits, over_bool, reg, over, N, width, g);   // registers its gates, as the reference does
    const std::string csub_n = CSubF(qubits, over_bool, reg, over, N, width, g);
    const std::string x_over = Pauli_X(qubits, over, 1);
    const std::string copy_over = CNot(qubits, over, over_bool);

    std::vector<std::string> s{sub_a};
    append(s, rqft);
    s.push_back(x_over);
    s.push_back(copy_over);
    s.push_back(x_over);
    append(s, qft);
    s.push_back(add_a);
    s.push_back(csub_n);
    append(s, rqft);
    s.push_back(copy_over);
    append(s, qft);
    s.push_back(add_n);
    s.push_back(sub_a);
    return s;
}

// Controlled |x>|0> -> |x>|a*x mod N>: the second register starts at 0, so its
// QFT is a Hadamard layer (src/algorithms/lib_shor_circuits.cpp:47-74).
std::vector<std::string> CMultMod(int qubits, int ctrl, int reg1, int reg2, int over, int over_bool, int width,
                                  long a, long N, Gates &g) {
    (void)QFT(qubits, reg2, over, width, g);
    const std::string h_reg2 = Hadamard(qubits, reg2, width);
    const std::vector<std::string> rqft = RQFT(qubits, reg2, over, width, g);
    std::vector<std::string> s{Hadamard(qubits, over, 1), h_reg2};
    const int top = reg1 + width - 1;
    for (int bit = 0; bit < width; ++bit) {
        append(s, C2AddMod(qubits, ctrl, top - bit, reg2, over, over_bool, width, a, N, g));
        a = (a * 2) % N;
    }
    append(s, rqft);
    return s;
}

// The inverse multiplication, built back to front
// (src/algorithms/lib_shor_circuits.cpp:76-98).
std::vector<std::string> CRMultMod(int qubits, int ctrl, int reg1, int reg2, int over, int over_bool, int width,
                                   long a, long N, Gates &g) {
    const std::vector<std::string> qft = QFT(qubits, reg2, over, width, g);
    const std::vector<std::string> rqft = RQFT(qubits, reg2, over, width, g);
    std::vector<std::vector<std::string>> blocks;
    const int top = reg1 + width - 1;
    for (int bit = 0; bit < width; ++bit) {
        blocks.push_back(C2SubMod(qubits, ctrl, top - bit, reg2, over, over_bool, width, a, N, g));
        a = (a * 2) % N;
    }
    std::vector<std::string> s(qft);
    for (auto it = blocks.rbegin(); it != blocks.rend(); ++it) append(s, *it);
    append(s, rqft);
    return s;
}

// ---- number theory ---------------------------------------------------------

int revert_bits(int value, int bits) {   // lib_shor_number_theory.cpp:14-22
    int r = 0;
    for (int k = 0; k < bits; ++k, value >>= 1) r = (r << 1) | (value & 1);
    return r;
}

int quantum_gcd(int a, int b) {          // lib_shor_number_theory.cpp:27-38
    while (b) {
        const int t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// Continued-fraction approximation of num/den with denominator below 2^width
// (lib_shor_number_theory.cpp:43-71). The reference keeps the running
// remainder in single precision and adds its 5e-6 guards in double; the
// same mixed arithmetic is kept here because the period candidate depends on it.
void quantum_frac_approx(int *num, int *den, int width) {
    const float target = (float)*num / (float)*den;
    float rest = target;
    int p2 = 0, q2 = 1, p1 = 1, q1 = 0, p = 0, q = 0;
    do {
        const int term = (int)((double)rest + 0.000005);
        rest = (float)((double)rest - ((double)term - 0.000005));
        rest = (float)(1.0 / (double)rest);
        if (term * q1 + q2 > (1 << width)) break;
        p = term * p1 + p2;
        q = term * q1 + q2;
        p2 = p1;
        q2 = q1;
        p1 = p;
        q1 = q;
    } while (std::fabs((double)p / q - (double)target) > 1.0 / (2 * (1 << width)));
    *num = p;
    *den = q;
}

long mul_inv(long value, long modulus) { // extended Euclid, lib_shor_number_theory.cpp:77-90
    const long m0 = modulus;
    long x0 = 0, x1 = 1;
    if (modulus == 1) return 1;
    while (value > 1) {
        const long q = value / modulus;
        long t = modulus;
        modulus = value % modulus;
        value = t;
        t = x0;
        x0 = x1 - q * x0;
        x1 = t;
    }
    return x1 < 0 ? x1 + m0 : x1;
}

long modular_pow(long base, long exponent, long modulus) {   // src/core/common.cpp:209-219
    long r = 1;
    base %= modulus;
    for (; exponent; exponent >>= 1) {
        if (exponent & 1) r = (r * base) % modulus;
        base = (base * base) % modulus;
    }
    return r;
}

// ---- driver ----------------------------------------------------------------

template <class Real>
ShorResult Shor(long N, const DGMCore &cfg) {
    ShorResult out;
    BasicDGM<Real> dgm;
    dgm.exec_type = cfg.exec_type;
    dgm.gpu_count = cfg.gpu_count;
    dgm.thread_count = cfg.thread_count;
    dgm.cpu_region_bits = cfg.cpu_region_bits;
    dgm.cpu_coalesced_bits = cfg.cpu_coalesced_bits;
    dgm.gpu_region_bits = cfg.gpu_region_bits;
    dgm.gpu_coalesced_bits = cfg.gpu_coalesced_bits;
    dgm.block_size = cfg.block_size;
    dgm.repeat_count = cfg.repeat_count;
    dgm.ops_block = cfg.ops_block;
    dgm.planner = cfg.planner;
    dgm.kernel = cfg.kernel;
    dgm.resident = cfg.resident;
    dgm.exit_on_error = cfg.exit_on_error;

    int width = 0;
    for (long v = N; v; v >>= 1) ++width;
    const int qubits = 2 * width + 3;
    dgm.qubits = qubits;
    dgm.allocateMemory();
    dgm.setMemoryValue(1L << (width + 2));

    const int phase_qb = 0, reg1 = 1, reg2 = width + 2, over = width + 1, over_bool = qubits - 1;
    long a = 0;
    std::uniform_int_distribution<long> pick(0, N - 1);
    while (quantum_gcd((int)N, (int)a) > 1 || a < 2) a = pick(g_rng);
    out.base = a;
    out.qubits = qubits;

    const std::string h_phase = Hadamard(qubits, phase_qb, 1);
    long phase_bits = 0;
    for (int round = 2 * width - 1; round >= 0; --round) {
        const long a_pow = modular_pow(a, 1L << round, N);
        std::vector<std::string> steps{h_phase};
        append(steps, CMultMod(qubits, phase_qb, reg1, reg2, over, over_bool, width, a_pow, N, dgm.gates));
        append(steps, CSwapR(qubits, phase_qb, reg1, reg2, width));
        append(steps, CRMultMod(qubits, phase_qb, reg1, reg2, over, over_bool, width, a_pow, N, dgm.gates));
        steps.push_back(h_phase);
        if (phase_bits) steps.push_back(genRot(qubits, phase_qb, phase_bits, dgm.gates));
        dgm.executeFunction(steps);
        out.gates += dgm.last_stats.gates;
        out.kernel_ms += dgm.last_stats.kernel_ms;
        phase_bits = (phase_bits << 1) | dgm.measure(phase_qb);
    }
    out.phase_bits = phase_bits;

    // period from the measured phase (src/algorithms/lib_shor.cpp:100-160)
    int num = revert_bits((int)phase_bits, 2 * width);
    if (num == 0) return out;
    int den = 1 << (2 * width);
    quantum_frac_approx(&num, &den, width);
    const int cf_den = den;
    for (int k = 1; cf_den * k < (1 << width); ++k)
        if (modular_pow(a, (long)cf_den * k, N) == 1) { den = cf_den * k; break; }
    if (den >= N) den = cf_den;

    auto try_period = [&](int r) {
        const int h = (int)modular_pow(a, r / 2, N);
        const int g1 = quantum_gcd((int)N, h + 1), g2 = quantum_gcd((int)N, h - 1);
        const int f = g1 > g2 ? g1 : g2;
        if (f < N && f > 1) {
            out.factors = {f, (int)(N / f)};
            return true;
        }
        return false;
    };
    if (try_period(den)) return out;
    if (cf_den != den) try_period(cf_den);
    return out;
}

template ShorResult Shor<double>(long, const DGMCore &);
template ShorResult Shor<float>(long, const DGMCore &);

}  // namespace hgm
