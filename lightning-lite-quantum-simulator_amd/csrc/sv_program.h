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
    int rq[4];  // register-resident passes: register bits of the final layout (ascending)
};

// One op of a pass, in tile-local coordinates.
enum PassOpKind : int {
    P_1Q = 0,      // a = local bit, m = 2x2
    P_DIAG = 1,    // a = local bit, m[0..1] = d0, m[6..7] = d1 (diagonal one-qubit gate)
    P_SCALE = 2,   // global qubit a outside the tile, d0/d1 as P_DIAG: a per-tile scalar
    P_CNOT = 3,    // a = local control, b = local target
    P_MCZ = 4,     // negate where (l & lmask) == lmask, if (tile base & omask) == omask
    P_DENSE = 5,   // a = k, q[0..k) local bits, moff = offset (in v2d) of the matrix
    // batched parameterised circuits (register-resident passes only):
    P_1QT = 6,     // a = local bit, moff = index of the 2x2 in the per-batch coefficient table,
                   // b = its form: 0 general, 1 [[p, iq], [iq, p]] (RX), 2 [[p, -q], [q, p]] (RY),
                   // 3 diagonal (RZ; needs no register residency)
    P_GRAD = 7,    // a = local bit, b = generator (0 X, 1 Y, 2 Z), moff = gradient slot:
                   // accumulate Im <lambda| G |psi> over the tile (Z needs no register residency)
    // register-resident passes with carried phases only:
    P_1QN = 8,     // a = local bit, a one-qubit matrix without its row scales (r_nu0 / r_nu1):
                   // m[0..3] = b, c; form b = 0 for (x + b y, c x + y), 1 for (b x + y, x + c y)
    P_RSCALE = 9,  // every amplitude of the tile times the real m[0] (row scales taken out of
                   // P_1QN ops before it)
};

struct PassOp {
    int kind;
    int a, b;
    int moff;
    int q[8];
    uint64_t lmask, omask;
    double m[8];
    uint32_t dep[32];  // P_DENSE: local index of operator column c, precomputed on the host
};

// Register-resident tiles: 2^12 amplitudes per workgroup of 256 threads, 16 per thread.
// Register j of a lane holds local index lt | R(j), where lt is the lane's share of the
// local index and R(j) deposits the 4 bits of j at the current register bits.
constexpr int kRegTileBits = 12;
constexpr int kRegBits = 4;

// One op of a register-resident pass.  Everything that depends on the current register
// layout (which registers an op touches, which lanes, the LDS slots of a layout change) is
// resolved on the host, so the kernel does only the arithmetic and the data movement.
enum RegOpKind : int {
    RO_1Q = 0,      // a = register bit of the target, m = 2x2
    RO_CNOT = 1,    // a = register bit of the target; registers j (bit a clear) in rmask swap
                    // with j | 2^a on lanes whose local index bits include tmask
    RO_DIAG = 2,    // d1 = m[6..7] on registers in rmask (tmask = 0) or on lanes holding tmask,
                    // d0 = m[0..1] elsewhere; a = 1 when d0 == 1 (those amplitudes are left alone)
    RO_MCZ = 3,     // negate registers in rmask on lanes holding tmask, if (base & omask) == omask
    RO_SCALE = 4,   // per-tile scalar: d0/d1 selected by global qubit a (outside the tile)
    RO_LAYOUT = 5,  // q[0..4) = new register bits (ascending); off0[j] / off1[j] = LDS byte
                    // slot of register j before / after (the lane's part is XORed in)
    RO_DENSE = 6,   // a = k, q[0..k) = local target bits, tmask = their mask, moff = matrix;
                    // off0[j] = LDS byte slot of register j, off1[j] = R(j), dep[c] = column c
    RO_1QT = 7,     // a = register bit of the target, moff = coefficient table index, q[0] = form
                    // (as P_1QT); form 3 (diagonal) as RO_DIAG: rmask, or tmask when the bit is
                    // a lane bit (compiled passes only)
    RO_GRAD = 8,    // a = register bit, q[0] = generator (0 X, 1 Y, 2 Z), moff = gradient slot;
                    // Z on a lane bit: tmask = that bit
    RO_1QN = 9,     // a = register bit of the target, q[0] = form, m as P_1QN
    RO_RSCALE = 10, // every register times the real m[0]
};

struct RegOp {
    int32_t kind, a;
    uint32_t rmask, tmask;
    uint64_t omask;
    int32_t moff, q[5];
    double m[8];
    uint32_t off0[16], off1[16];
    uint16_t dep[32];
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

// A pass's ops (tile-local coordinates) as a register-resident program; rq receives the
// register bits of the final layout.
std::vector<RegOp> to_register_program(const std::vector<PassOp>& in, int L, int rq[4]);
int jit_precompile(int n, const llsv_op* ops, uint64_t n_ops);
int device_tile_bits(llsv_state* s);

// The scheduler alone (host only): fills order[] with the op indices in execution order and
// pass_end[p] with the end of pass p in order[] (either may be null); returns the number
// of passes, or a negative error code.
int plan_order(int n, const llsv_op* ops, uint64_t n_ops, int b, int L, uint64_t* order, uint64_t* pass_end);

}  // namespace llsv
