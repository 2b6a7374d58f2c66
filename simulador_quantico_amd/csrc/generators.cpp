// Circuit step generators and benchmark circuit builders.
//
// The step generators produce the reference's circuit language byte for byte
// (src/core/gates.cpp:69-163, src/algorithms/lib_grover.cpp:68-91,
// src/algorithms/lib_shor_circuits.cpp:11-359): one token per qubit, joined
// with ','. Phases are generated in double (the reference uses cpowf in
// single precision); gate names and the decisions of which steps to emit are
// the same.
#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>

#include "hgm_dgm.hpp"

namespace hgm {

namespace {

std::string istr(long v) { return std::to_string((int)v); }   // int2str takes an int

std::string ctl(int value) { return value ? "Control1(1)" : "Control1(0)"; }

}  // namespace

std::string concatena(const std::vector<std::string> &ops, int qubits, bool reverse) {
    std::string s;
    if (!reverse) {
        s = ops[0];
        for (int q = 1; q < qubits; ++q) s += "," + ops[q];
    } else {
        s = ops[qubits - 1];
        for (int q = qubits - 2; q >= 0; --q) s += "," + ops[q];
    }
    return s;
}

std::string CNot(int qubits, int ctrl, int target, int control_value) {
    std::vector<std::string> ops(qubits, "ID");
    ops[ctrl] = ctl(control_value);
    ops[target] = "Target1(X)";
    return concatena(ops, qubits);
}

std::string Toffoli(int qubits, int ctrl1, int ctrl2, int target, int control_value) {
    std::vector<std::string> ops(qubits, "ID");
    ops[ctrl1] = ctl(control_value >> 1);
    ops[ctrl2] = ctl(control_value & 1);
    ops[target] = "Target1(X)";
    return concatena(ops, qubits);
}

std::string Controlled1(int qubits, int ctrl, int target, const std::string &op, int control_value) {
    std::vector<std::string> ops(qubits, "ID");
    ops[ctrl] = ctl(control_value);
    ops[target] = "Target1(" + op + ")";
    return concatena(ops, qubits);
}

std::string Controlled2(int qubits, int ctrl1, int ctrl2, int target, const std::string &op, int control_value) {
    std::vector<std::string> ops(qubits, "ID");
    ops[ctrl1] = ctl(control_value & 2);
    ops[ctrl2] = ctl(control_value & 1);
    ops[target] = "Target1(" + op + ")";
    return concatena(ops, qubits);
}

// The last control takes the least significant bit of control_value; -1 means
// "all ones" (src/core/gates.cpp:110-123).
std::string ControlledN(int qubits, const std::vector<int> &ctrls, int target, const std::string &op,
                        int control_value) {
    if (control_value == -1) control_value = (int)((1L << ctrls.size()) - 1);
    std::vector<std::string> ops(qubits, "ID");
    for (int i = (int)ctrls.size() - 1; i >= 0; --i) {
        ops[ctrls[i]] = ctl(control_value & 1);
        control_value >>= 1;
    }
    ops[target] = "Target1(" + op + ")";
    return concatena(ops, qubits);
}

static std::string layer(int qubits, int reg, int width, const char *g) {
    std::vector<std::string> ops(qubits, "ID");
    for (int q = 0; q < width; ++q) ops[q + reg] = g;
    return concatena(ops, qubits);
}

std::string Pauli_X(int qubits, int reg, int width) { return layer(qubits, reg, width, "X"); }
std::string Pauli_Z(int qubits, int reg, int width) { return layer(qubits, reg, width, "Z"); }
std::string Hadamard(int qubits, int reg, int width) { return layer(qubits, reg, width, "H"); }

// Oracle of the ancilla form: qubits 1..n-1 are controls carrying the bits of
// search_value (MSB first), qubit 0 is the X target (src/algorithms/lib_grover.cpp:68-82).
std::string Oracle1(long qubits, long search_value) {
    std::vector<std::string> ops((size_t)qubits);
    for (long q = qubits - 1; q >= 1; --q) {
        ops[q] = ctl((int)(search_value & 1));
        search_value >>= 1;
    }
    ops[0] = "Target1(X)";
    return concatena(ops, (int)qubits);
}

std::string ControledZ(int qubits) {
    std::vector<std::string> ops;
    ops.push_back("ID");
    for (int q = 0; q < qubits - 2; ++q) ops.push_back("Control1(1)");
    ops.push_back("Target1(Z)");
    return concatena(ops, qubits);
}

std::vector<std::string> GroverIterationRef(long qubits, long search_value) {
    std::vector<std::string> steps;
    steps.push_back(Oracle1(qubits, search_value));
    for (int q = 1; q < qubits; ++q) {
        steps.push_back(Hadamard((int)qubits, q, 1));
        steps.push_back(Pauli_X((int)qubits, q, 1));
    }
    steps.push_back(ControledZ((int)qubits));
    for (int q = (int)qubits - 1; q >= 1; --q) {
        steps.push_back(Pauli_X((int)qubits, q, 1));
        steps.push_back(Hadamard((int)qubits, q, 1));
    }
    return steps;
}

long GroverIterations(int search_bits) { return (long)(M_PI / 4.0 * std::sqrt((double)(1L << search_bits))); }

std::vector<std::string> GroverIteration(int qubits, uint64_t marked) {
    std::vector<std::string> xs(qubits, "ID");
    for (int q = 0; q < qubits; ++q)
        if (!((marked >> (qubits - 1 - q)) & 1)) xs[q] = "X";
    const std::string oracle_x = concatena(xs, qubits);
    std::vector<int> ctrls;
    for (int q = 0; q < qubits - 1; ++q) ctrls.push_back(q);
    const std::string mcz = ControlledN(qubits, ctrls, qubits - 1, "Z", -1);
    const std::string h = Hadamard(qubits, 0, qubits), x = Pauli_X(qubits, 0, qubits);
    return {oracle_x, mcz, oracle_x, h, x, mcz, x, h};
}

// src/algorithms/lib_shor_circuits.cpp:268-313: QFT on [over, reg .. reg+width)
// with rotations R<k> = diag(1, e^(+-2 pi i / 2^k)) registered as "R<k>" or
// "R'<k>"; no final swaps; the inverse is the reversed step list.
static std::vector<std::string> qft_impl(int qubits, int reg, int over, int width, bool inverse, Gates &g) {
    const std::string prefix = inverse ? "R'" : "R";
    const double sign = inverse ? -1.0 : 1.0;
    for (int level = 1; level <= width + 1; ++level)
        g.addGate(prefix + istr(level), 1.0, 0.0, 0.0, std::polar(1.0, sign * 2.0 * M_PI / std::pow(2.0, level)));
    std::vector<std::string> steps;
    std::vector<std::string> ops(qubits, "ID");
    steps.push_back(Hadamard(qubits, over, 1));
    for (int c = 0; c < width; ++c) {
        ops[c + reg] = "Control1(1)";
        ops[over] = "Target1(" + prefix + istr(c + 2) + ")";
        steps.push_back(concatena(ops, qubits));
        ops[c + reg] = "ID";
    }
    ops[over] = "ID";
    for (int t = 0; t < width; ++t) {
        steps.push_back(Hadamard(qubits, t + reg, 1));
        for (int c = t + 1; c < width; ++c) {
            ops[c + reg] = "Control1(1)";
            ops[t + reg] = "Target1(" + prefix + istr(c - t + 1) + ")";
            steps.push_back(concatena(ops, qubits));
            ops[c + reg] = "ID";
        }
        ops[t + reg] = "ID";
    }
    if (inverse) std::reverse(steps.begin(), steps.end());
    return steps;
}

std::vector<std::string> QFT(int qubits, int reg, int over, int width, Gates &g) {
    return qft_impl(qubits, reg, over, width, false, g);
}
std::vector<std::string> RQFT(int qubits, int reg, int over, int width, Gates &g) {
    return qft_impl(qubits, reg, over, width, true, g);
}

// src/algorithms/lib_shor_circuits.cpp:190-227: Draper constant adder. Digit d
// (d = 0 least significant, width+1 digits) gets the phase
// sum over set bits i <= d of value of +-2 pi / 2^(d-i+1); digits with a
// nonzero phase get a gate "ADD_<value>_<d>" / "SUB_<value>_<d>" on qubit
// reg+width-1-d, and the top digit is moved from qubit reg-1 to `over`.
static std::vector<std::string> addsub_impl(int qubits, int reg, int over, long value, int width, bool controlled,
                                            bool subtract, Gates &g) {
    if (reg < 1) throw Error(HGM_E_ARG, "AddF/SubF need reg >= 1 (the top digit sits on qubit reg-1 first)");
    const int digits = width + 1;
    const double sign = subtract ? -1.0 : 1.0;
    std::vector<double> angle((size_t)digits, 0.0);
    long rest = value;
    for (int d = 0; d < digits; ++d) {
        if (rest & 1)
            for (int h = d; h < digits; ++h) angle[h] += sign * 2.0 * M_PI / std::pow(2.0, h - d + 1);
        rest >>= 1;
    }
    std::vector<std::string> ops(qubits, "ID");
    const std::string prefix = subtract ? "SUB_" : "ADD_";
    const int msb = reg + width - 1;
    for (int d = 0; d < digits; ++d) {
        if (angle[d] == 0.0) continue;
        std::string name = prefix + istr(value) + "_" + istr(d);
        g.addGate(name, 1.0, 0.0, 0.0, std::polar(1.0, angle[d]));
        if (controlled) name = "Target1(" + name + ")";
        ops[msb - d] = name;
    }
    std::string top = ops[reg - 1];
    ops[reg - 1] = "ID";
    ops[over] = top;
    return ops;
}

std::vector<std::string> AddF(int qubits, int reg, int over, long value, int width, bool controlled, Gates &g) {
    return addsub_impl(qubits, reg, over, value, width, controlled, false, g);
}
std::vector<std::string> SubF(int qubits, int reg, int over, long value, int width, bool controlled, Gates &g) {
    return addsub_impl(qubits, reg, over, value, width, controlled, true, g);
}
std::string AddF(int qubits, int reg, int over, long value, int width, Gates &g) {
    return concatena(AddF(qubits, reg, over, value, width, false, g), qubits);
}
std::string SubF(int qubits, int reg, int over, long value, int width, Gates &g) {
    return concatena(SubF(qubits, reg, over, value, width, false, g), qubits);
}
std::string CAddF(int qubits, int ctrl1, int reg, int over, long value, int width, Gates &g) {
    std::vector<std::string> ops = AddF(qubits, reg, over, value, width, true, g);
    ops[ctrl1] = "Control1(1)";
    return concatena(ops, qubits);
}
std::string C2AddF(int qubits, int ctrl1, int ctrl2, int reg, int over, long value, int width, Gates &g) {
    std::vector<std::string> ops = AddF(qubits, reg, over, value, width, true, g);
    ops[ctrl1] = "Control1(1)";
    ops[ctrl2] = "Control1(1)";
    return concatena(ops, qubits);
}
std::string CSubF(int qubits, int ctrl1, int reg, int over, long value, int width, Gates &g) {
    std::vector<std::string> ops = SubF(qubits, reg, over, value, width, true, g);
    ops[ctrl1] = "Control1(1)";
    return concatena(ops, qubits);
}
std::string C2SubF(int qubits, int ctrl1, int ctrl2, int reg, int over, long value, int width, Gates &g) {
    std::vector<std::string> ops = SubF(qubits, reg, over, value, width, true, g);
    ops[ctrl1] = "Control1(1)";
    ops[ctrl2] = "Control1(1)";
    return concatena(ops, qubits);
}

// src/algorithms/lib_shor_circuits.cpp:11-45: semiclassical phase correction,
// bit b of phase_bits contributing -2 pi / 2^(b+2).
std::string genRot(int qubits, int reg, long phase_bits, Gates &g) {
    std::vector<std::string> ops(qubits, "ID");
    double ang = 0;
    long bits = phase_bits;
    for (int k = 2; bits; ++k, bits >>= 1)
        if (bits & 1) ang += -2.0 * M_PI / std::pow(2.0, k);
    if (ang != 0.0) {
        std::string name = "Rot_" + istr(phase_bits);
        g.addGate(name, 1.0, 0.0, 0.0, std::polar(1.0, ang));
        ops[reg] = name;
    }
    return concatena(ops, qubits);
}

std::vector<std::string> CSwapR(int qubits, int ctrl, int reg1, int reg2, int width) {
    std::vector<std::string> steps, ops(qubits, "ID");
    for (int q = 0; q < width; ++q) {
        ops[ctrl] = "Control1(1)";
        ops[q + reg1] = "Target1(X)";
        ops[q + reg2] = "Control1(1)";
        const std::string s1 = concatena(ops, qubits);
        ops[ctrl] = "ID";
        ops[q + reg1] = "Control1(1)";
        ops[q + reg2] = "Target1(X)";
        const std::string s2 = concatena(ops, qubits);
        ops[q + reg1] = ops[q + reg2] = "ID";
        steps.push_back(s2);
        steps.push_back(s1);
        steps.push_back(s2);
    }
    return steps;
}

std::vector<std::string> SwapOver(int qubits, int reg, int width) {
    std::vector<std::string> steps;
    for (int q = 0; q < width / 2; ++q) {
        steps.push_back(CNot(qubits, reg + width - q - 1, reg + q));
        steps.push_back(CNot(qubits, reg + q, reg + width - q - 1));
        steps.push_back(CNot(qubits, reg + width - q - 1, reg + q));
    }
    return steps;
}

std::array<cplx, 4> U3(double theta, double phi, double lambda) {
    const double c = std::cos(theta / 2), s = std::sin(theta / 2);
    return {cplx(c, 0), -std::polar(s, lambda), std::polar(s, phi), std::polar(c, phi + lambda)};
}

std::vector<std::string> RandomLayers(int qubits, int depth, uint64_t seed, Gates &g) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> ang(0.0, 2.0 * M_PI);
    std::vector<std::string> steps;
    for (int l = 0; l < depth; ++l) {
        std::vector<std::string> ops(qubits);
        for (int q = 0; q < qubits; ++q) {
            const double th = ang(rng), ph = ang(rng), la = ang(rng);
            const std::string name = "U3_" + std::to_string(l) + "_" + std::to_string(q);
            std::array<cplx, 4> m = U3(th, ph, la);
            g.addGate(name, m.data());
            ops[q] = name;
        }
        steps.push_back(concatena(ops, qubits));
        std::vector<std::string> cz(qubits, "ID");
        int group = 1;
        for (int q = l % 2; q + 1 < qubits; q += 2, ++group) {
            cz[q] = "Control" + std::to_string(group) + "(1)";
            cz[q + 1] = "Target" + std::to_string(group) + "(Z)";
        }
        if (group > 1) steps.push_back(concatena(cz, qubits));
    }
    return steps;
}

std::vector<std::string> DraperAdder(int width, Gates &g) {
    const int qubits = 2 * width + 1, carry = width, reg = width + 1;
    std::vector<std::string> steps = QFT(qubits, reg, carry, width, g);
    // a-bit i (weight 2^(width-1-i)) rotates register position j (qubit
    // carry + j) by 2 pi / 2^(2 + i - j) whenever that is not a full turn
    for (int i = 0; i < width; ++i)
        for (int j = 0; j <= i + 1 && j <= width; ++j)
            steps.push_back(Controlled1(qubits, i, carry + j, "R" + std::to_string(2 + i - j)));
    std::vector<std::string> inv = RQFT(qubits, reg, carry, width, g);
    steps.insert(steps.end(), inv.begin(), inv.end());
    return steps;
}

}  // namespace hgm
