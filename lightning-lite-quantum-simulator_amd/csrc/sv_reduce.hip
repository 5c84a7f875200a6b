// sv_reduce.hip -- readout kernels: norm, <Z_q>, <a|b>, probabilities.
//
// Two-stage, fixed-shape reductions: stage 1 runs a fixed grid (kRedBlocks blocks of
// 256 threads) in which each thread walks a fixed grid-stride sequence, then sums its
// wavefront with xor-butterfly shuffles and the 4 wavefronts of the block through LDS;
// stage 2 is one block that folds the partials in a fixed tree.  The summation order
// therefore depends only on the state size, so results are bitwise reproducible run to
// run (no atomics anywhere).
#include "sv_common.h"
#include "sv_kernels.h"

namespace llsv {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the 256 threads of a block; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds4) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) s = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    __syncthreads();
    return s;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_reduce_prob(const v2d* __restrict__ psi, uint64_t size, int q,
                                                     double* __restrict__ partials) {
    __shared__ double lds4[4];
    const uint64_t stride = uint64_t(gridDim.x) * 256;
    double acc0 = 0.0, acc1 = 0.0;
    uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    for (; i + stride < size; i += 2 * stride) {
        const v2d a = ld_nt(psi + i), b = ld_nt(psi + i + stride);
        double pa = a.x * a.x + a.y * a.y, pb = b.x * b.x + b.y * b.y;
        if (MODE == 1) {
            if ((i >> q) & 1) pa = -pa;
            if (((i + stride) >> q) & 1) pb = -pb;
        }
        acc0 += pa;
        acc1 += pb;
    }
    if (i < size) {
        const v2d a = ld_nt(psi + i);
        double pa = a.x * a.x + a.y * a.y;
        if (MODE == 1 && ((i >> q) & 1)) pa = -pa;
        acc0 += pa;
    }
    const double s = block_sum(acc0 + acc1, lds4);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_inner(const v2d* __restrict__ a, const v2d* __restrict__ b, uint64_t size,
                                               double* __restrict__ partials) {
    __shared__ double lds4[4];
    const uint64_t stride = uint64_t(gridDim.x) * 256;
    double re = 0.0, im = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < size; i += stride) {
        const v2d x = a[i], y = b[i];
        re += x.x * y.x + x.y * y.y;
        im += x.x * y.y - x.y * y.x;
    }
    const double sr = block_sum(re, lds4);
    const double si = block_sum(im, lds4);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sr;
        partials[2 * blockIdx.x + 1] = si;
    }
}

// One block: out[c] = sum_b partials[b * ncomp + c] for c < ncomp (ncomp = 1 or 2).
__global__ __launch_bounds__(256) void k_reduce_final(const double* __restrict__ partials, int nparts, int ncomp,
                                                      double* __restrict__ out) {
    __shared__ double lds4[4];
    for (int c = 0; c < ncomp; ++c) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nparts; b += 256) acc += partials[b * ncomp + c];
        const double s = block_sum(acc, lds4);
        if (threadIdx.x == 0) out[c] = s;
    }
}

static int red_grid(uint64_t size) {
    const uint64_t want = (size + 255) / 256;
    return int(want < uint64_t(kRedBlocks) ? want : uint64_t(kRedBlocks));
}

hipError_t launch_reduce_prob(hipStream_t st, const v2d* psi, int n, int mode, int q, double* partials,
                              double* out_dev) {
    const uint64_t size = uint64_t(1) << n;
    const int grid = red_grid(size);
    if (mode == 0)
        hipLaunchKernelGGL(k_reduce_prob<0>, dim3(grid), dim3(256), 0, st, psi, size, q, partials);
    else
        hipLaunchKernelGGL(k_reduce_prob<1>, dim3(grid), dim3(256), 0, st, psi, size, q, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, partials, grid, 1, out_dev);
    return hipGetLastError();
}

hipError_t launch_inner(hipStream_t st, const v2d* a, const v2d* b, int n, double* partials, double* out_dev) {
    const uint64_t size = uint64_t(1) << n;
    const int grid = red_grid(size);
    hipLaunchKernelGGL(k_inner, dim3(grid), dim3(256), 0, st, a, b, size, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, partials, grid, 2, out_dev);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_probabilities(const v2d* __restrict__ psi, uint64_t size,
                                                       double* __restrict__ out) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < size; i += uint64_t(gridDim.x) * 256) {
        const v2d a = ld_nt(psi + i);
        out[i] = a.x * a.x + a.y * a.y;
    }
}

hipError_t launch_probabilities(hipStream_t st, const v2d* psi, uint64_t size, double* out_dev) {
    hipLaunchKernelGGL(k_probabilities, dim3(grid_for((size + 255) / 256)), dim3(256), 0, st, psi, size, out_dev);
    return hipGetLastError();
}

}  // namespace llsv
