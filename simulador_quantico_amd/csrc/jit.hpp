// Plan-time specialisation of the register kernel.
//
// The register program of every pass (RegOp list, phase maps, load/store
// assignments) is translated into straight-line HIP source: amplitudes in
// named registers, gate matrices as exact hexadecimal literals (zero, +-1 and
// real entries fold away), register- and global-bit conditions resolved at
// generation time. The source is compiled for gfx950 with hipRTC when the plan
// is built (no GPU needed) and loaded into a module on first use.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "planner.hpp"

namespace hgm {

struct JitCode {
    std::string source;          // generated HIP source of all pass kernels
    std::vector<char> binary;    // gfx950 code object
    uint64_t binary_hash = 0;    // of the code object
    std::vector<int32_t> pass_kernel;   // kernel index of each pass (passes with identical code share one)
    std::vector<std::string> names;     // kernel symbol names
    size_t lds_bytes = 0;        // dynamic LDS of a pass kernel with remaps
    uint64_t hash = 0;           // of the source
    double compile_s = 0;
};

// Generates the source of every pass kernel of a register-kernel plan.
void jit_generate(const Plan &plan, JitCode *out);

// Compiles out->source with hipRTC for gfx950 (cached by source hash within
// the process, the 64 most recently used). Returns an HGM_* status.
int jit_compile(JitCode *out, std::string *err);

// Number of code objects in the compile cache.
size_t jit_compiled_count();

// Serving pass kernels. A register-kernel pass's kernel for the serving engine
// is a function of its structure only: op kinds, register bits, condition
// masks, pivot columns, phase maps and load/store tables, the qubit count and
// tile shape. Matrix entries and control values are read at run time from
// the item's register program (as the interpreter reads them), so requests
// whose passes have equal keys share one kernel. The kernel takes the batched
// interpreter's arguments (items, n_items, programs, states, partials) and
// also reduces the readout of an item that asks for it.
std::string serving_pass_key(const Plan &plan, const Pass &ps);
std::string serving_kernel_source(const Plan &plan, const Pass &ps, const std::string &name, size_t *lds_bytes);

}  // namespace hgm
