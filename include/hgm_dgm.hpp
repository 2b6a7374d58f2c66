// hgm_dgm.hpp -- C++ executor API of the MI355X t_GPU library.
//
// Mirrors the reference's DGM executor (include/core/dgm.h:65-163), its gate
// catalog (include/core/gates.h:40-54), its PT record (include/core/common.h:57-90)
// and its step generators (include/core/gates.h:15-31), so that code written
// against the reference keeps its shape:
//
//     hgm::DGM dgm;                       // complex<double> amplitudes
//     dgm.qubits = 20;
//     dgm.exec_type = hgm::t_GPU;
//     dgm.allocateMemory();
//     dgm.setMemoryValue(0);
//     dgm.setFunction(hgm::Hadamard(20, 0, 20));
//     dgm.execute(1);                     // host state updated in place
//     auto a = dgm.state[0];
//
// hgm::DGM32 is the same executor on complex<float> amplitudes (the
// reference's precision). Gate matrices are generated and stored in double in
// both cases; the FP32 executor rounds them when it uploads a pass.
//
// Only exec_type == t_GPU executes; t_CPU, t_PAR_CPU and t_HYBRID are
// rejected with an error (the CPU engines stay with the reference).
#pragma once

#include <array>
#include <complex>
#include <map>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "hgm_gpu.h"

namespace hgm {

using cplx = std::complex<double>;

enum { t_CPU, t_PAR_CPU, t_GPU, t_HYBRID };   // include/core/dgm.h:42-47
enum { DENSE, DIAG_PRI, DIAG_SEC };           // include/core/common.h:48-52

constexpr int QB_LIMIT = 40;                  // 64-bit indexing throughout (the reference stops at 30)

// Shared generator for measurement draws (src/core/common.cpp:10-14).
extern std::mt19937 g_rng;
void seed_rng(unsigned long seed);

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// One compiled single-qubit gate (include/core/common.h:57-90).
struct PT {
    int qubits = 1;
    const cplx *matrix = nullptr;   // borrowed from the DGM's catalog
    int matrix_size = 2;
    int span_start_bit = 0, target_bit = 0;
    bool affected = false;
    long control_value = 0, control_mask = 0;
    std::vector<long> control_bit_positions;
    long control_count = 0;

    long matrixType() const;        // DENSE / DIAG_PRI / DIAG_SEC (src/core/common.cpp:33-38)
    void print() const;             // fields to stdout, as src/core/common.cpp:68-74 prints them
    void printMatrix() const;       // the 2x2 matrix, as src/core/common.cpp:77-85
};

// Gate catalog: name -> row-major 2x2 matrix; addGate never overwrites
// (src/core/gates.cpp:42-60).
class Gates {
public:
    std::map<std::string, std::array<cplx, 4>> list;
    Gates();
    void init();                    // H, X, Z, Y, R1, R2, R3 (src/core/gates.cpp:23-34)
    const cplx *getMatrix(const std::string &gate_name) const;
    bool addGate(const std::string &name, const cplx *matrix);
    bool addGate(const std::string &name, cplx a0, cplx a1, cplx a2, cplx a3);
};

// Controls and targets of one group of a step (include/core/dgm.h:51-61).
class Group {
public:
    std::vector<std::string> ops;
    std::vector<long> pos_ops;
    std::vector<long> ctrl;
    std::vector<long> pos_ctrl;
};

// Everything of the executor that does not depend on the amplitude type.
class DGMCore {
public:
    long total_op = 0, dense = 0, main_diag = 0, sec_diag = 0, c_dense = 0, c_main_diag = 0, c_sec_diag = 0;
    Gates gates;
    int exec_type = t_GPU;
    long thread_count = 1, cpu_coalesced_bits = 11, cpu_region_bits = 14;
    int gpu_count = 1;
    long gpu_coalesced_bits = 4, gpu_region_bits = 8;
    int block_size = 64, repeat_count = 2;
    int ops_block = 300;            // OPS_BLOCK (makefile:7), a runtime value here
    int planner = HGM_PLAN_GREEDY;
    int kernel = HGM_KERNEL_AUTO;
    bool resident = false;          // keep the state on the device between execute() calls
    bool exit_on_error = true;      // print + exit(1) like the reference; false: throw hgm::Error
    std::vector<PT *> vec_pts;      // NULL-terminated like the reference's
    PT **pts = nullptr;
    long qubits = 0;
    double measure_value = 0;
    hgm_stats last_stats{};

    DGMCore() = default;
    DGMCore(const DGMCore &) = delete;
    DGMCore &operator=(const DGMCore &) = delete;
    virtual ~DGMCore();

    void erase();
    void setExecType(int type) { exec_type = type; }
    void setCpuStructure(long cpu_region_bits, long cpu_coalesced_bits);
    void setGpuStructure(long gpu_coalesced_bits, long gpu_region_bits, int repeat_count = 1);

    void setFunction(const std::string &function, int iterations = 1, bool reset = true);
    void setFunction(const std::vector<std::string> &steps, int iterations = 1, bool reset = true);
    std::map<long, Group> genGroups(const std::string &step);
    void genPTs(std::map<long, Group> &groups, std::vector<PT *> &step_pts);
    void CountOps(int iterations = 1);
    void printPTs() const;          // print() of every PT (src/core/dgm_core.cpp:35-39)
    void validateTuning();

    hgm_tuning tuning() const;
    std::vector<hgm_op> ops() const;    // the PT list as POD ops
    virtual int amp_dtype() const { return HGM_C128; }

protected:
    [[noreturn]] void fail(int code, const std::string &msg) const;
    long state_qubits = 0;              // qubits the state was allocated with
};

template <class Real>
class BasicDGM : public DGMCore {
public:
    using amp_t = std::complex<Real>;
    amp_t *state = nullptr;

    BasicDGM() = default;
    ~BasicDGM() override;

    static constexpr int dtype() { return sizeof(Real) == 8 ? HGM_C128 : HGM_C64; }
    int amp_dtype() const override { return dtype(); }

    void allocateMemory();              // zeroed 2^qubits amplitudes (pinned when a GPU is present)
    void freeMemory();
    void setMemoryValue(long pos);      // state[pos] = 1 (src/core/dgm_core.cpp:63-65)

    int measure(int qubit_pos);         // src/core/dgm_core.cpp:69-125
    std::map<long, double> measure(const std::vector<int> &qubit_positions);   // :150-162
    void colapse(int qubit_pos, int value);   // :129-146

    void executeFunction(const std::string &function, int iterations = 1);
    void executeFunction(const std::vector<std::string> &steps, int iterations = 1);
    // src/core/dgm_core.cpp:243-267: t_GPU, and t_HYBRID with every region on the GPU (the PT list
    // once, as HybridExecution runs it); t_CPU / t_PAR_CPU fail with HGM_E_UNSUPPORTED. Returns the
    // updated host state; in resident mode the result stays on the device and nullptr is returned
    // (syncHost() or hostState() brings the host copy up to date)
    amp_t *execute(int iterations);

    void syncHost();                    // resident mode: bring the host copy up to date
    // syncHost(), then hand out the host copy for reading and writing: in resident mode the
    // device copy is treated as stale, so host writes reach the next execute()
    amp_t *hostState();
    hgm_state *deviceState() { return dev_; }

private:
    bool pinned_ = false;
    hgm_state *dev_ = nullptr;
    bool host_valid_ = true;
    bool dev_valid_ = false;
    void ensure_device();
};

using DGM = BasicDGM<double>;
using DGM32 = BasicDGM<float>;

// ---- step generators (src/core/gates.cpp:69-163) ----
std::string CNot(int qubits, int ctrl, int target, int control_value = 1);
std::string Toffoli(int qubits, int ctrl1, int ctrl2, int target, int control_value = 3);
std::string Controlled1(int qubits, int ctrl, int target, const std::string &op, int control_value = 1);
std::string Controlled2(int qubits, int ctrl1, int ctrl2, int target, const std::string &op, int control_value = 3);
std::string ControlledN(int qubits, const std::vector<int> &ctrls, int target, const std::string &op,
                        int control_value = -1);
std::string Pauli_X(int qubits, int reg, int width = 1);
std::string Pauli_Z(int qubits, int reg, int width = 1);
std::string Hadamard(int qubits, int reg, int width = 1);
std::string concatena(const std::vector<std::string> &step_ops, int qubits, bool reverse = false);

// ---- circuit builders (src/algorithms/lib_grover.cpp, lib_shor_circuits.cpp) ----
std::string Oracle1(long qubits, long search_value);
std::string ControledZ(int qubits);
std::vector<std::string> QFT(int qubits, int reg, int over, int width, Gates &g);
std::vector<std::string> RQFT(int qubits, int reg, int over, int width, Gates &g);
std::vector<std::string> AddF(int qubits, int reg, int over, long value, int width, bool controlled, Gates &g);
std::vector<std::string> SubF(int qubits, int reg, int over, long value, int width, bool controlled, Gates &g);
std::string AddF(int qubits, int reg, int over, long value, int width, Gates &g);
std::string SubF(int qubits, int reg, int over, long value, int width, Gates &g);
std::string CAddF(int qubits, int ctrl1, int reg, int over, long value, int width, Gates &g);
std::string C2AddF(int qubits, int ctrl1, int ctrl2, int reg, int over, long value, int width, Gates &g);
std::string CSubF(int qubits, int ctrl1, int reg, int over, long value, int width, Gates &g);
std::string C2SubF(int qubits, int ctrl1, int ctrl2, int reg, int over, long value, int width, Gates &g);
std::string genRot(int qubits, int reg, long phase_bits, Gates &g);
std::vector<std::string> CSwapR(int qubits, int ctrl, int reg1, int reg2, int width);
std::vector<std::string> SwapOver(int qubits, int reg, int width);

// Modular arithmetic on Fourier-basis registers (src/algorithms/lib_shor_circuits.cpp:47-166).
std::vector<std::string> C2AddMod(int qubits, int ctrl1, int ctrl2, int reg, int over, int over_bool, int width,
                                  long a, long N, Gates &g);
std::vector<std::string> C2SubMod(int qubits, int ctrl1, int ctrl2, int reg, int over, int over_bool, int width,
                                  long a, long N, Gates &g);
std::vector<std::string> CMultMod(int qubits, int ctrl, int reg1, int reg2, int over, int over_bool, int width,
                                  long a, long N, Gates &g);
std::vector<std::string> CRMultMod(int qubits, int ctrl, int reg1, int reg2, int over, int over_bool, int width,
                                   long a, long N, Gates &g);

// Classical helpers of Shor's algorithm (src/algorithms/lib_shor_number_theory.cpp,
// src/core/common.cpp:209-219).
int revert_bits(int value, int bits);
int quantum_gcd(int a, int b);
void quantum_frac_approx(int *num, int *den, int width);
long mul_inv(long value, long modulus);
long modular_pow(long base, long exponent, long modulus);

struct ShorResult {
    std::vector<int> factors;   // {f, N/f}, or empty when this attempt failed
    long base = 0;              // the random base drawn from g_rng
    long phase_bits = 0;        // measured phase bits, first round in the highest bit
    int qubits = 0;
    int64_t gates = 0;          // gates executed over all rounds
    double kernel_ms = 0;       // pass-kernel time over all rounds
};

// One attempt of Shor's algorithm with semiclassical phase estimation
// (src/algorithms/lib_shor.cpp:17-163): 2*width+3 qubits, 2*width rounds on
// one executor configured from cfg (exec_type, tuning, resident, ...).
template <class Real>
ShorResult Shor(long N, const DGMCore &cfg);

// The step list Grover() runs after its Hadamard layer, one iteration
// (ancilla on qubit 0; src/algorithms/lib_grover.cpp:34-51).
std::vector<std::string> GroverIterationRef(long qubits, long search_value);
// Grover without ancilla: oracle = X on the zero bits of `marked`, n-controlled
// Z, X; diffusion = H layer, X layer, n-controlled Z, X layer, H layer.
std::vector<std::string> GroverIteration(int qubits, uint64_t marked);
// floor(pi/4 * sqrt(2^bits)) (src/algorithms/lib_grover.cpp:53).
long GroverIterations(int search_bits);

// Random-layer circuit: `depth` layers of U3(theta, phi, lambda) on every qubit
// (angles uniform in [0, 2pi) from std::mt19937_64(seed), drawn per layer, per
// qubit, in the order theta, phi, lambda), each followed by a CZ brickwork on
// pairs (0,1),(2,3),... for even layers and (1,2),(3,4),... for odd ones.
// Registers the U3 matrices in g as "U3_<layer>_<qubit>".
std::vector<std::string> RandomLayers(int qubits, int depth, uint64_t seed, Gates &g);
std::array<cplx, 4> U3(double theta, double phi, double lambda);

// Quantum-quantum Draper adder on |a>_width |0 b>_(width+1): QFT on the
// (carry, b) register, controlled phase rotations from the bits of a, inverse
// QFT. Qubit layout: a = qubits [0, width), carry = qubit width, b = qubits
// [width+1, 2*width+1). Leaves |a, a+b> with a+b read from qubits
// [width, 2*width+1) (carry is the most significant bit).
std::vector<std::string> DraperAdder(int width, Gates &g);

}  // namespace hgm
