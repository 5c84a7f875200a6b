// sv_kernels.h -- launchers of the gfx950 kernels (internal to libllsv).
#pragma once

#include "sv_common.h"

namespace llsv {

// Sorted bit positions (ascending) plus their mask; m <= 40.
struct BitList {
    int m;
    int pos[40];
    uint64_t mask;
};

// Grid size used by the two-stage reductions; partial buffers hold 2 * kRedBlocks doubles.
constexpr int kRedBlocks = 2048;

hipError_t launch_copy(hipStream_t st, v2d* dst, const v2d* src, int n);
hipError_t launch_fill_basis(hipStream_t st, v2d* psi, int n, uint64_t index);
hipError_t launch_apply_1q(hipStream_t st, v2d* psi, int n, int q, const Gate2& g);
hipError_t launch_cnot(hipStream_t st, v2d* psi, int n, int c, int t);
hipError_t launch_mcz(hipStream_t st, v2d* psi, int n, const BitList& bits);
hipError_t launch_dense(hipStream_t st, v2d* psi, int n, int k, const BitList& sorted_targets, const v2d* U_dev,
                        const uint64_t* off_dev);

// mode 0: sum |a|^2 ; mode 1: sum (-1)^{bit q} |a|^2.  Result (one double) in out_dev.
hipError_t launch_reduce_prob(hipStream_t st, const v2d* psi, int n, int mode, int q, double* partials,
                              double* out_dev);
// <a|b> into out_dev[0..1].
hipError_t launch_inner(hipStream_t st, const v2d* a, const v2d* b, int n, double* partials, double* out_dev);
// |psi_i|^2 for i < size into out_dev[i].
hipError_t launch_probabilities(hipStream_t st, const v2d* psi, uint64_t size, double* out_dev);

}  // namespace llsv
