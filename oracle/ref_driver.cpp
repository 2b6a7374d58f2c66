// ORACLE -- test infrastructure only (see oracle/oracle.cpp).
//
// A thin extern "C" driver around the reference's own DGM, compiled together
// with the reference's CPU sources (read in place under the reference tree by
// oracle/Makefile, never copied) into oracle/_ref/libhgmref.so. It lets the
// tests and the bench's cpu_baseline leg run the reference's real t_CPU and
// t_PAR_CPU engines and its real circuit generators, in FP32 as the reference
// computes. The t_GPU path (kernel.cu) is not built: the stub that the
// reference's own "GPU=stub" build uses stands in for it in libhgmref.so, and
// this repository's libhgm.so in libhgmref_hgm.so.

#include "core/dgm.h"
#include "core/gates.h"
#include "algorithms/lib_general.h"
#include "algorithms/lib_grover.h"
#include "algorithms/lib_shor.h"

#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

int put(const std::string &s, char *out, size_t cap) {
    if (out && cap > s.size()) std::memcpy(out, s.c_str(), s.size() + 1);
    return (int)s.size();
}

std::string join(const std::vector<std::string> &v) {
    std::string r;
    for (size_t i = 0; i < v.size(); ++i) { if (i) r += ";"; r += v[i]; }
    return r;
}

}  // namespace

extern "C" {

void *ref_new(int qubits) {
    DGM *d = new DGM();
    d->qubits = qubits;
    d->allocateMemory();
    d->thread_count = 1;
    d->cpu_region_bits = 14;        // include/cli/cli_common.h:10-11
    d->cpu_coalesced_bits = 11;
    d->gpu_region_bits = 8;
    d->gpu_coalesced_bits = 4;
    d->block_size = 64;
    d->repeat_count = 2;
    return d;
}

void ref_free(void *p) { delete static_cast<DGM *>(p); }

// GPU tuning of t_GPU / t_HYBRID (validated by DGM::validateTuning,
// src/core/dgm_core.cpp:177-241, which exits on a bad tuple).
void ref_set_gpu_tuning(void *p, int gpu_region_bits, int gpu_coalesced_bits, int block_size, int repeat_count) {
    DGM *d = static_cast<DGM *>(p);
    d->gpu_region_bits = gpu_region_bits;
    d->gpu_coalesced_bits = gpu_coalesced_bits;
    d->block_size = block_size;
    d->repeat_count = repeat_count;
}

int ref_add_gate(void *p, const char *name, const float *m8) {
    DGM *d = static_cast<DGM *>(p);
    float complex a[4];
    for (int i = 0; i < 4; ++i) { __real__ a[i] = m8[2 * i]; __imag__ a[i] = m8[2 * i + 1]; }
    return d->gates.addGate(name, a[0], a[1], a[2], a[3]) ? 1 : 0;
}

int ref_get_gate(void *p, const char *name, float *m8) {
    DGM *d = static_cast<DGM *>(p);
    float complex *m = d->gates.getMatrix(name);
    if (!m) return 0;
    for (int i = 0; i < 4; ++i) { m8[2 * i] = crealf(m[i]); m8[2 * i + 1] = cimagf(m[i]); }
    return 1;
}

int ref_set_function_steps(void *p, const char *const *steps, int n, int iterations, int reset) {
    DGM *d = static_cast<DGM *>(p);
    std::vector<std::string> v(steps, steps + n);
    d->setFunction(v, iterations, reset != 0);
    return 0;
}

int ref_qubits(void *p) { return (int)static_cast<DGM *>(p)->qubits; }

// measure(vector<int>) (src/core/dgm_core.cpp:150-162): the map's entries in
// key order; returns their number (at most cap are written).
long ref_measure_many(void *p, const int *qubits, int n, long *keys, float *probs, long cap) {
    std::map<long, float> m = static_cast<DGM *>(p)->measure(std::vector<int>(qubits, qubits + n));
    long k = 0;
    for (const auto &kv : m) {
        if (k < cap) { keys[k] = kv.first; probs[k] = kv.second; }
        ++k;
    }
    return k;
}

void ref_colapse(void *p, int qubit, int value) {
    static_cast<DGM *>(p)->colapse(qubit, value);
    std::fflush(stdout);
}

void ref_print_pts(void *p) {
    static_cast<DGM *>(p)->printPTs();
    std::fflush(stdout);
}

long ref_num_pts(void *p) {
    DGM *d = static_cast<DGM *>(p);
    long n = 0;
    while (d->pts && d->pts[n]) ++n;
    return n;
}

int ref_get_pts(void *p, int *tbits, long *cmask, long *cval, float *m8) {
    DGM *d = static_cast<DGM *>(p);
    for (long i = 0; d->pts && d->pts[i]; ++i) {
        PT *t = d->pts[i];
        tbits[i] = t->target_bit;
        cmask[i] = t->control_mask;
        cval[i] = t->control_value;
        for (int k = 0; k < 4; ++k) {
            m8[8 * i + 2 * k] = t->matrix ? crealf(t->matrix[k]) : 0.f;
            m8[8 * i + 2 * k + 1] = t->matrix ? cimagf(t->matrix[k]) : 0.f;
        }
    }
    return 0;
}

int ref_set_basis(void *p, long index) {
    DGM *d = static_cast<DGM *>(p);
    std::memset(d->state, 0, sizeof(float complex) * (size_t(1) << d->qubits));
    d->setMemoryValue((int)index);
    return 0;
}

int ref_set_state(void *p, const float *amps) {
    DGM *d = static_cast<DGM *>(p);
    std::memcpy(d->state, amps, sizeof(float complex) * (size_t(1) << d->qubits));
    return 0;
}

int ref_get_state(void *p, float *out) {
    DGM *d = static_cast<DGM *>(p);
    std::memcpy(out, d->state, sizeof(float complex) * (size_t(1) << d->qubits));
    return 0;
}

// exec_type 0 = t_CPU, 1 = t_PAR_CPU; 2 = t_GPU and 3 = t_HYBRID reach the GPU
// stub in libhgmref.so and libhgm.so in libhgmref_hgm.so.
int ref_execute(void *p, int exec_type, int threads, int cpu_coalesced_bits, int cpu_region_bits, int iterations) {
    DGM *d = static_cast<DGM *>(p);
    d->exec_type = exec_type;
    d->thread_count = threads;
    d->cpu_coalesced_bits = cpu_coalesced_bits;
    d->cpu_region_bits = cpu_region_bits;
    d->execute(iterations);
    return 0;
}

void ref_seed(unsigned long seed) { seed_rng(seed); }

int ref_measure(void *p, int qubit) { return static_cast<DGM *>(p)->measure(qubit); }

// Step generators of src/core/gates.cpp:69-163 and the Shor circuit builders
// (src/algorithms/lib_shor_circuits.cpp), by name. Multi-step results are
// joined with ';'. Returns the length of the text written to out.
int ref_gen(void *p, const char *which, const long *a, int na, const char *op, char *out, size_t cap) {
    std::string w(which);
    DGM *d = static_cast<DGM *>(p);
    if (w == "CNot") return put(CNot(a[0], a[1], a[2], a[3]), out, cap);
    if (w == "Toffoli") return put(Toffoli(a[0], a[1], a[2], a[3], a[4]), out, cap);
    if (w == "Controlled1") return put(Controlled1(a[0], a[1], a[2], op, a[3]), out, cap);
    if (w == "Controlled2") return put(Controlled2(a[0], a[1], a[2], a[3], op, a[4]), out, cap);
    if (w == "ControlledN") {
        std::vector<int> ctrls(a + 3, a + na);
        return put(ControlledN(a[0], ctrls, a[1], op, a[2]), out, cap);
    }
    if (w == "Pauli_X") return put(Pauli_X(a[0], a[1], a[2]), out, cap);
    if (w == "Pauli_Z") return put(Pauli_Z(a[0], a[1], a[2]), out, cap);
    if (w == "Hadamard") return put(Hadamard(a[0], a[1], a[2]), out, cap);
    if (w == "Oracle1") return put(Oracle1(a[0], a[1]), out, cap);
    if (w == "ControledZ") return put(ControledZ(a[0]), out, cap);
    if (w == "QFT") return put(join(QFT(a[0], a[1], a[2], a[3], d->gates)), out, cap);
    if (w == "RQFT") return put(join(RQFT(a[0], a[1], a[2], a[3], d->gates)), out, cap);
    if (w == "AddF") return put(AddF(a[0], a[1], a[2], a[3], a[4], d->gates), out, cap);
    if (w == "SubF") return put(SubF(a[0], a[1], a[2], a[3], a[4], d->gates), out, cap);
    if (w == "AddFc") return put(join(AddF(a[0], a[1], a[2], a[3], a[4], a[5] != 0, d->gates)), out, cap);
    if (w == "SubFc") return put(join(SubF(a[0], a[1], a[2], a[3], a[4], a[5] != 0, d->gates)), out, cap);
    if (w == "CAddF") return put(CAddF(a[0], a[1], a[2], a[3], a[4], a[5], d->gates), out, cap);
    if (w == "genRot") return put(genRot(a[0], a[1], a[2], d->gates), out, cap);
    if (w == "CSwapR") return put(join(CSwapR(a[0], a[1], a[2], a[3], a[4])), out, cap);
    if (w == "C2AddMod") return put(join(C2AddMod(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], d->gates)), out, cap);
    if (w == "C2SubMod") return put(join(C2SubMod(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], d->gates)), out, cap);
    if (w == "CMultMod") return put(join(CMultMod(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], d->gates)), out, cap);
    if (w == "CRMultMod") return put(join(CRMultMod(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], d->gates)), out, cap);
    return -1;
}

// The reference's classical helpers of Shor's algorithm.
void ref_frac_approx(int *num, int *den, int width) { quantum_frac_approx(num, den, width); }
long ref_mul_inv(long v, long m) { return mul_inv(v, m); }
int ref_revert_bits(int v, int bits) { return revert_bits(v, bits); }

// One attempt of the reference's Shor() with its t_CPU engine; out = {f, N/f}
// or {0, 0}.
void ref_shor(long number, int *out) {
    std::vector<int> f = Shor(number, t_CPU, 1, 14, 11, 1, 8, 4, 64, 2);
    out[0] = f.size() == 2 ? f[0] : 0;
    out[1] = f.size() == 2 ? f[1] : 0;
}

// The reference's application drivers with any exec_type and the CLI's tuning
// defaults (include/cli/cli_common.h:4-13): general.out's HadamardNQubits
// (src/algorithms/lib_general.cpp:15-47, 3 iterations, src/cli/general.cpp:43)
// prints amplitude 0 to stdout; Grover (src/algorithms/lib_grover.cpp:15-91)
// returns the measured value; Shor (src/algorithms/lib_shor.cpp:18-163) the
// factors, out = {f, N/f} or {0, 0}. The seed goes to the measurement RNG.
void ref_general_hadamard(long qubits, int exec_type, int threads) {
    HadamardNQubits(qubits, 3, exec_type, threads, 14, 11, 1, 8, 4, 64, 2);
    std::fflush(stdout);
}

long ref_grover_run(long qubits, long search_value, int exec_type, int threads, unsigned long seed) {
    seed_rng(seed);
    return Grover(qubits, search_value, exec_type, threads, 14, 11, 1, 8, 4, 64, 2);
}

void ref_shor_run(long number, int exec_type, int threads, unsigned long seed, int *out) {
    seed_rng(seed);
    std::vector<int> f = Shor(number, exec_type, threads, 14, 11, 1, 8, 4, 64, 2);
    out[0] = f.size() == 2 ? f[0] : 0;
    out[1] = f.size() == 2 ? f[1] : 0;
}

// The step list Grover() (src/algorithms/lib_grover.cpp:15-66) executes after
// its initial Hadamard layer, for one iteration, joined with ';'.
int ref_grover_iteration(long qubits, long search_value, char *out, size_t cap) {
    std::vector<std::string> steps;
    steps.push_back(Oracle1(qubits, search_value));
    for (int q = 1; q < qubits; ++q) { steps.push_back(Hadamard(qubits, q, 1)); steps.push_back(Pauli_X(qubits, q, 1)); }
    steps.push_back(ControledZ(qubits));
    for (int q = qubits - 1; q >= 1; --q) { steps.push_back(Pauli_X(qubits, q, 1)); steps.push_back(Hadamard(qubits, q, 1)); }
    return put(join(steps), out, cap);
}

}  // extern "C"
