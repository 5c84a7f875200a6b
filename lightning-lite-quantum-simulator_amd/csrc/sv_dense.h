// sv_dense.h -- fused dense blocks on FP64 MFMA (internal to libllsv).
#pragma once

#include "sv_common.h"
#include "sv_program.h"

namespace llsv {

constexpr int kDenseTileBits = 12;  // largest tile of the dense kernel (index bits)

// Operator bit j of the block is tile-local bit tloc[j]; gbits[] are the other local bits,
// ascending (they index the tile's groups of 2^k amplitudes).
struct DenseTile {
    int k;
    int tloc[5];
    int gbits[kDenseTileBits];
};

// Tile geometry for a k-qubit block on n qubits (k = 4 or 5; n >= 11 for k = 5, n >= 12 for
// k = 4); false if not applicable.
bool make_dense_tile(int n, int k, const int* targets, PassDesc* d, DenseTile* dt);
hipError_t launch_dense_mfma(hipStream_t st, v2d* psi, const PassDesc& d, const DenseTile& dt, const v2d* U_dev);

}  // namespace llsv
