// sv_jit.h -- the pass compiler: straight-line HIP kernels generated per register-resident pass.
//
// The interpreter (k_pass_reg) dispatches on op kinds at run time; at every merge of its control
// flow the compiler keeps the 16-amplitude register array in one assignment by copying it, which
// costs about as many instructions as the gate arithmetic itself.  A generated kernel has the op
// sequence, register bits, masks and LDS offsets of one pass as constants: CNOTs become register
// renaming, sign flips fold into the arithmetic, and nothing is copied.  Gate coefficients stay
// run-time data (read from the pass's RegOp array), so one kernel serves every pass with the same
// structure, whatever the angles or the tile's qubits.  Kernels are compiled with hipRTC and
// cached in the process by their source text.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "sv_program.h"

namespace llsv {

enum JitMode : int { kJitOff = 0, kJitAuto = 1, kJitAlways = 2 };

// Automatic mode compiles passes of states with at least this many qubits (below it a pass
// takes microseconds and the compile, about a second per distinct pass, does not pay); it
// compiles in a background thread and runs the interpreter until a pass's kernel is ready.
// Mode kJitAlways compiles every eligible pass before its first launch.
constexpr int kJitMinQubits = 20;
// Passes with more ops than this keep the interpreter (compile time grows with the length).
constexpr int kJitMaxOps = 1024;

int jit_mode();
int set_jit_mode(int mode);

// Whether a pass of nops register ops (dense ops excluded) on an n-qubit state is compiled.
bool jit_wanted(int n, int nops, bool dense);

// Source of the generated kernel for one pass (the kernel is named LLSV_KERNEL; a compile
// batch renames it and prepends the shared prelude).
std::string jit_source(const RegOp* ops, int nops);

// Source of a batched pass of a parameterised circuit (states A + b * size for b = blockIdx.y,
// one-qubit coefficients from the table T + b * tstride).  backward: lambda goes through every
// op too, and the RO_GRAD ops' per-tile sums land in partials[(b * nslot + slot) * ntiles + tile].
std::string jit_adjoint_source(const RegOp* ops, int nops, bool backward);

struct JitEntry;

// The cache entry for a kernel source (created on first sight, not compiled).
JitEntry* jit_entry(const std::string& body);

// jit_entry(jit_source(ops, nops)), memoized on the ops' structure: a pass seen before (the same
// circuit applied again, whatever its angles) finds its kernel without generating the source.
JitEntry* jit_program_entry(const RegOp* ops, int nops);
bool jit_ready(const JitEntry* e);

// Asks for the entries to be compiled, all in one hipRTC program.  now: in the calling thread,
// returning when every entry is compiled or has failed; otherwise by a background thread
// while the caller goes on (an entry that failed once is not compiled again).
void jit_request(const std::vector<JitEntry*>& entries, bool now);

// Blocks until the background thread has nothing left to compile.
void jit_wait();

// Launches a ready entry's kernel on the current device: grid workgroups of 256 threads.
hipError_t jit_launch(JitEntry* e, unsigned grid, hipStream_t st, v2d* psi, const PassDesc& d, const RegOp* dops);

// Launches a ready batched entry (jit_adjoint_source) on grid (tiles, batch).
hipError_t jit_launch_batched(JitEntry* e, unsigned tiles, unsigned batch, hipStream_t st, v2d* A, v2d* Lm,
                              uint64_t size, const PassDesc& d, const double* T, uint64_t tstride, double* partials,
                              int nslot, int ntiles);

void jit_stats(uint64_t* compiled, uint64_t* launches, uint64_t* failures);

}  // namespace llsv
