// sv_program.h -- the gate-program path: whole circuits applied in LDS-staged HBM passes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/llsv.h"
#include "sv_common.h"

namespace llsv {

// Tile geometry of one pass.  The tile holds 2^b amplitudes: local bit j is global qubit
// tq[j] (tq[0..L) = 0..L-1, so every tile is made of contiguous runs of 2^L amplitudes);
// the other n - b bits are fixed per workgroup.  pos[] = tq[] sorted ascending.
constexpr int kMaxTileBits = 12;
struct PassDesc {
    int n, b, L, nops;
    int tq[kMaxTileBits];
    int pos[kMaxTileBits];
};

// One op of a pass, in tile-local coordinates.
enum PassOpKind : int {
    P_1Q = 0,      // a = local bit, m = 2x2
    P_DIAG = 1,    // a = local bit, m[0..1] = d0, m[6..7] = d1 (diagonal one-qubit gate)
    P_SCALE = 2,   // global qubit a outside the tile, d0/d1 as P_DIAG: a per-tile scalar
    P_CNOT = 3,    // a = local control, b = local target
    P_MCZ = 4,     // negate where (l & lmask) == lmask, if (tile base & omask) == omask
    P_DENSE = 5,   // a = k, q[0..k) local bits, moff = offset (in v2d) of the matrix
    // register-resident variant (b = 12) only:
    P_LAYOUT = 6,  // q[0..4) = the 4 local bits held in registers from now on (ascending)
    R_1Q = 7,      // a = register bit (0..3), m = 2x2
    R_CNOT = 8,    // a = local control bit, b = register bit of the target
};

// Register-resident tiles: 2^12 amplitudes per workgroup of 256 threads, 16 per thread.
constexpr int kRegTileBits = 12;
constexpr int kRegBits = 4;

struct PassOp {
    int kind;
    int a, b;
    int moff;
    int q[8];
    uint64_t lmask, omask;
    double m[8];
    uint32_t dep[32];  // P_DENSE: local index of operator column c, precomputed on the host
};

// Per-state host/device buffers of the program path.
struct ProgramCache {
    unsigned char* dev = nullptr;   // pass ops + dense matrices, device
    unsigned char* host = nullptr;  // pinned staging for the upload
    size_t cap = 0;
    hipEvent_t uploaded = nullptr;  // staging may be rewritten once this has completed
    int max_tile_bits = 0;          // from the device's LDS size, filled on first use
    ~ProgramCache();
};

int run_program(llsv_state* s, const llsv_op* ops, uint64_t n_ops);
int device_tile_bits(llsv_state* s);

// The scheduler alone (host only): fills order[] with the op indices in execution order and
// pass_end[p] with the end of pass p in order[] (either may be null); returns the number
// of passes, or a negative error code.
int plan_order(int n, const llsv_op* ops, uint64_t n_ops, int b, int L, uint64_t* order, uint64_t* pass_end);

}  // namespace llsv
