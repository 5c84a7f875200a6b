// sv_traj.hip -- batched gate-level noisy trajectories in one persistent launch.
//
// Workgroups (256 threads) take items from a device work counter until it runs past the
// batch, so a mixed batch of 10..16-qubit circuits balances itself (items are handed out in
// decreasing size).  Each workgroup owns a state slot of 2^max_n amplitudes in the workspace
// (256 KiB at 14 qubits, resident in L2 / Infinity Cache while the item runs).  Per gate:
//   pass 1: apply the gate and accumulate the joint probabilities of its qubits;
//   thread 0: draw Philox4x32-10 uniforms (key = seed, counter = (trajectory, gate, site)),
//             decide AD then PD on each touched qubit from the joint weights, updated in closed
//             form after each outcome, and form the product operator (normalised);
//   pass 2: apply that operator.
// Every wave reaches the exit test of the work loop; nothing waits on another workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/llsv.h"
#include "sv_common.h"

namespace llsv {

namespace {

struct TG {
    int kind, q0, q1, pad;
    double m[8];
};
static_assert(sizeof(TG) == sizeof(llsv_tgate), "llsv_tgate layout");

struct TItem {
    int n, first_gate, n_gates, first_rate;
    uint64_t traj;
};
static_assert(sizeof(TItem) == sizeof(llsv_traj_item), "llsv_traj_item layout");

__device__ __host__ inline void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c[3] ^ k1;
        c[1] = uint32_t(p1);
        c[3] = uint32_t(p0);
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ inline double uniform(uint64_t seed, uint64_t traj, uint32_t gate, uint32_t site) {
    uint32_t c[4] = {uint32_t(traj), uint32_t(traj >> 32), gate, site};
    philox(c, uint32_t(seed), uint32_t(seed >> 32));
    const uint64_t x = uint64_t(c[0]) | (uint64_t(c[1]) << 32);
    return double(x >> 11) * (1.0 / 9007199254740992.0);
}

__device__ inline void mm2(const double* a, const double* b, double* c) {
    double r[8];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < 2; ++k) {
                const double ar = a[2 * (2 * i + k)], ai = a[2 * (2 * i + k) + 1];
                const double br = b[2 * (2 * k + j)], bi = b[2 * (2 * k + j) + 1];
                re = __dadd_rn(re, __dsub_rn(__dmul_rn(ar, br), __dmul_rn(ai, bi)));
                im = __dadd_rn(im, __dadd_rn(__dmul_rn(ar, bi), __dmul_rn(ai, br)));
            }
            r[2 * (2 * i + j)] = re;
            r[2 * (2 * i + j) + 1] = im;
        }
    for (int i = 0; i < 8; ++i) c[i] = r[i];
}

// AD then PD on one qubit; P[a][b]: joint weights (a = this qubit's bit).  Returns jumps.
__device__ inline int decide(double P[2][2], double g, double l, uint64_t seed, uint64_t traj, uint32_t gate,
                             uint32_t site0, double* k, uint64_t* hash) {
    int jumps = 0;
    double ad[8] = {0, 0, 0, 0, 0, 0, 0, 0}, pd[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const double tot = (P[0][0] + P[0][1]) + (P[1][0] + P[1][1]);
    const double p1 = (P[1][0] + P[1][1]) / tot;
    const double u = uniform(seed, traj, gate, site0);
    if (u >= __dsub_rn(1.0, __dmul_rn(g, p1))) {
        ad[2] = sqrt(g);
        P[0][0] = g * P[1][0];
        P[0][1] = g * P[1][1];
        P[1][0] = P[1][1] = 0.0;
        ++jumps;
        *hash = (*hash ^ (uint64_t(gate) * 4 + site0)) * 0x100000001B3ull;
    } else {
        ad[0] = 1.0;
        ad[6] = sqrt(1.0 - g);
        P[1][0] *= 1.0 - g;
        P[1][1] *= 1.0 - g;
    }
    const double tot2 = (P[0][0] + P[0][1]) + (P[1][0] + P[1][1]);
    const double q1 = (P[1][0] + P[1][1]) / tot2;
    const double v = uniform(seed, traj, gate, site0 + 1);
    if (v >= __dsub_rn(1.0, __dmul_rn(l, q1))) {
        pd[6] = sqrt(l);
        P[0][0] = P[0][1] = 0.0;
        P[1][0] *= l;
        P[1][1] *= l;
        ++jumps;
        *hash = (*hash ^ (uint64_t(gate) * 4 + site0 + 1)) * 0x100000001B3ull;
    } else {
        pd[0] = 1.0;
        pd[6] = sqrt(1.0 - l);
        P[1][0] *= 1.0 - l;
        P[1][1] *= 1.0 - l;
    }
    mm2(pd, ad, k);
    return jumps;
}

__device__ __forceinline__ v2d ap(const double* m, v2d a0, v2d a1) {
    v2d o;
    o.x = m[0] * a0.x - m[1] * a0.y + m[2] * a1.x - m[3] * a1.y;
    o.y = m[0] * a0.y + m[1] * a0.x + m[2] * a1.y + m[3] * a1.x;
    return o;
}

__device__ __forceinline__ double nrm2(v2d a) { return a.x * a.x + a.y * a.y; }

// Sums NV values over the 256 threads of the block; the totals land in out[] (all threads).
template <int NV>
__device__ inline void block_sums(double* v, double (*lds)[4], double* out) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[k][w] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = (lds[k][0] + lds[k][1]) + (lds[k][2] + lds[k][3]);
    __syncthreads();
}

__global__ void k_traj_reset(unsigned* counter) { *counter = 0u; }

__global__ __launch_bounds__(256) void k_traj(const TItem* __restrict__ items, int n_items,
                                              const int* __restrict__ order, const TG* __restrict__ gates,
                                              const double* __restrict__ gamma, const double* __restrict__ lam,
                                              uint64_t seed, v2d* __restrict__ ws, uint64_t slot_size,
                                              unsigned* __restrict__ counter, double* __restrict__ zout,
                                              unsigned* __restrict__ jout, uint64_t* __restrict__ hout) {
    __shared__ int s_item;
    __shared__ double s_k[2][8];
    __shared__ double lds[20][4];
    v2d* s = ws + uint64_t(blockIdx.x) * slot_size;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) {
            const unsigned k = atomicAdd(counter, 1u);
            s_item = k < unsigned(n_items) ? order[k] : -1;
        }
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it < 0) return;
        const TItem item = items[it];
        const int n = item.n;
        const uint64_t N = uint64_t(1) << n;
        for (uint64_t i = tid; i < N; i += 256) {
            v2d z;
            z.x = (i == 0) ? 1.0 : 0.0;
            z.y = 0.0;
            s[i] = z;
        }
        __syncthreads();
        unsigned jumps = 0;
        uint64_t hash = 0xCBF29CE484222325ull;
        for (int gi = 0; gi < item.n_gates; ++gi) {
            const TG g = gates[item.first_gate + gi];
            if (g.kind == 0) {
                const int q = g.q0;
                const uint64_t lomask = (uint64_t(1) << q) - 1, st = uint64_t(1) << q;
                double v[2] = {0.0, 0.0};
                for (uint64_t p = tid; p < N / 2; p += 256) {
                    const uint64_t i0 = ((p & ~lomask) << 1) | (p & lomask);
                    const v2d a0 = s[i0], a1 = s[i0 + st];
                    const v2d b0 = ap(g.m, a0, a1), b1 = ap(g.m + 4, a0, a1);
                    s[i0] = b0;
                    s[i0 + st] = b1;
                    v[0] += nrm2(b0);
                    v[1] += nrm2(b1);
                }
                double P2[2];
                block_sums<2>(v, lds, P2);
                if (tid == 0) {
                    double P[2][2] = {{P2[0], 0.0}, {P2[1], 0.0}};
                    double k[8];
                    jumps += decide(P, gamma[item.first_rate + q], lam[item.first_rate + q], seed, item.traj,
                                    uint32_t(gi), 0, k, &hash);
                    const double nr = 1.0 / sqrt(P[0][0] + P[1][0]);
                    for (int j = 0; j < 8; ++j) s_k[0][j] = k[j] * nr;
                }
                __syncthreads();
                double k[8];
                for (int j = 0; j < 8; ++j) k[j] = s_k[0][j];
                for (uint64_t p = tid; p < N / 2; p += 256) {
                    const uint64_t i0 = ((p & ~lomask) << 1) | (p & lomask);
                    const v2d a0 = s[i0], a1 = s[i0 + st];
                    s[i0] = ap(k, a0, a1);
                    s[i0 + st] = ap(k + 4, a0, a1);
                }
                __syncthreads();
            } else {
                const int c = g.q0, t = g.q1;
                const int lo = c < t ? c : t, hi = c < t ? t : c;
                const uint64_t cb = uint64_t(1) << c, tb = uint64_t(1) << t;
                double v[4] = {0.0, 0.0, 0.0, 0.0};  // weights (c, t) = 00, 01, 10, 11
                for (uint64_t p = tid; p < N / 4; p += 256) {
                    const uint64_t b = insert0(insert0(p, lo), hi);
                    const v2d a00 = s[b], a01 = s[b | tb], a10 = s[b | cb], a11 = s[b | cb | tb];
                    s[b | cb] = a11;  // control set: swap target
                    s[b | cb | tb] = a10;
                    v[0] += nrm2(a00);
                    v[1] += nrm2(a01);
                    v[2] += nrm2(a11);
                    v[3] += nrm2(a10);
                }
                double P4[4];
                block_sums<4>(v, lds, P4);
                if (tid == 0) {
                    double P[2][2] = {{P4[0], P4[1]}, {P4[2], P4[3]}};
                    double kc[8], kt[8];
                    jumps += decide(P, gamma[item.first_rate + c], lam[item.first_rate + c], seed, item.traj,
                                    uint32_t(gi), 0, kc, &hash);
                    double Q[2][2] = {{P[0][0], P[1][0]}, {P[0][1], P[1][1]}};
                    jumps += decide(Q, gamma[item.first_rate + t], lam[item.first_rate + t], seed, item.traj,
                                    uint32_t(gi), 2, kt, &hash);
                    const double nr = 1.0 / sqrt((Q[0][0] + Q[0][1]) + (Q[1][0] + Q[1][1]));
                    for (int j = 0; j < 8; ++j) {
                        s_k[0][j] = kc[j];
                        s_k[1][j] = kt[j] * nr;
                    }
                }
                __syncthreads();
                double kc[8], kt[8];
                for (int j = 0; j < 8; ++j) {
                    kc[j] = s_k[0][j];
                    kt[j] = s_k[1][j];
                }
                for (uint64_t p = tid; p < N / 4; p += 256) {
                    const uint64_t b = insert0(insert0(p, lo), hi);
                    v2d a00 = s[b], a01 = s[b | tb], a10 = s[b | cb], a11 = s[b | cb | tb];
                    // control qubit first, then target (the operators act on different qubits)
                    const v2d c00 = ap(kc, a00, a10), c10 = ap(kc + 4, a00, a10);
                    const v2d c01 = ap(kc, a01, a11), c11 = ap(kc + 4, a01, a11);
                    s[b] = ap(kt, c00, c01);
                    s[b | tb] = ap(kt + 4, c00, c01);
                    s[b | cb] = ap(kt, c10, c11);
                    s[b | cb | tb] = ap(kt + 4, c10, c11);
                }
                __syncthreads();
            }
        }
        double zq[20];
        for (int q = 0; q < 20; ++q) zq[q] = 0.0;
        for (uint64_t i = tid; i < N; i += 256) {
            const double p = nrm2(s[i]);
            for (int q = 0; q < 20; ++q)
                if (q < n) zq[q] += ((i >> q) & 1) ? -p : p;
        }
        double Z[20];
        block_sums<20>(zq, lds, Z);
        if (tid == 0) {
            for (int q = 0; q < n; ++q) zout[size_t(it) * LLSV_TRAJ_ZSTRIDE + q] = Z[q];
            jout[it] = jumps;
            hout[it] = hash;
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace llsv

extern "C" int llsv_traj_run(const llsv_traj_item* items, int n_items, const llsv_tgate* gates, int n_gates,
                             const double* gamma, const double* lambda, int n_rates, uint64_t seed, double* z,
                             uint32_t* jumps, uint64_t* hashes, double* state) {
    using namespace llsv;
    if (!items || n_items < 1 || !gates || n_gates < 1 || !gamma || !lambda || n_rates < 1 || !z)
        return fail(LLSV_ERR_ARG, "llsv_traj_run: empty or null argument");
    int maxn = 0;
    for (int i = 0; i < n_items; ++i) {
        const llsv_traj_item& it = items[i];
        if (it.n_qubits < 1 || it.n_qubits > 20) return fail(LLSV_ERR_ARG, "llsv_traj_run: item %d: n_qubits must be in [1, 20]", i);
        if (it.first_gate < 0 || it.n_gates < 0 || it.first_gate + it.n_gates > n_gates)
            return fail(LLSV_ERR_ARG, "llsv_traj_run: item %d: gate range outside the gate list", i);
        if (it.first_rate < 0 || it.first_rate + it.n_qubits > n_rates)
            return fail(LLSV_ERR_ARG, "llsv_traj_run: item %d: rate range outside the rate lists", i);
        for (int k = 0; k < it.n_gates; ++k) {
            const llsv_tgate& g = gates[it.first_gate + k];
            const bool ok = (g.kind == 0 && g.q0 >= 0 && g.q0 < it.n_qubits) ||
                            (g.kind == 1 && g.q0 >= 0 && g.q0 < it.n_qubits && g.q1 >= 0 && g.q1 < it.n_qubits && g.q0 != g.q1);
            if (!ok) return fail(LLSV_ERR_ARG, "llsv_traj_run: item %d gate %d: bad kind or qubits", i, k);
        }
        for (int q = 0; q < it.n_qubits; ++q) {
            const double g = gamma[it.first_rate + q], l = lambda[it.first_rate + q];
            if (!(g >= 0.0 && g <= 1.0 && l >= 0.0 && l <= 1.0))
                return fail(LLSV_ERR_ARG, "llsv_traj_run: item %d: rates must be in [0, 1]", i);
        }
        maxn = std::max(maxn, it.n_qubits);
    }
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int blocks = std::min(n_items, 4 * cus);
    const uint64_t slot = uint64_t(1) << maxn;
    // largest items first
    std::vector<int> order(n_items);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return items[a].n_qubits > items[b].n_qubits; });

    const size_t b_items = sizeof(llsv_traj_item) * n_items, b_order = sizeof(int) * n_items;
    const size_t b_gates = sizeof(llsv_tgate) * n_gates, b_rates = sizeof(double) * n_rates;
    const size_t b_z = sizeof(double) * LLSV_TRAJ_ZSTRIDE * n_items, b_j = sizeof(unsigned) * n_items;
    const size_t b_h = sizeof(uint64_t) * n_items, b_ws = slot * 16 * size_t(blocks);
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t total = al(b_items) + al(b_order) + al(b_gates) + 2 * al(b_rates) + al(b_z) + al(b_j) + al(b_h) + 256 + al(b_ws);
    unsigned char* buf = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&buf), total) != hipSuccess) {
        (void)hipGetLastError();
        return fail(LLSV_ERR_NOMEM, "llsv_traj_run: cannot allocate %zu bytes", total);
    }
    size_t off = 0;
    auto take = [&](size_t b) { unsigned char* p = buf + off; off += al(b); return p; };
    TItem* d_items = reinterpret_cast<TItem*>(take(b_items));
    int* d_order = reinterpret_cast<int*>(take(b_order));
    TG* d_gates = reinterpret_cast<TG*>(take(b_gates));
    double* d_gamma = reinterpret_cast<double*>(take(b_rates));
    double* d_lam = reinterpret_cast<double*>(take(b_rates));
    double* d_z = reinterpret_cast<double*>(take(b_z));
    unsigned* d_j = reinterpret_cast<unsigned*>(take(b_j));
    uint64_t* d_h = reinterpret_cast<uint64_t*>(take(b_h));
    unsigned* d_counter = reinterpret_cast<unsigned*>(take(256));
    v2d* d_ws = reinterpret_cast<v2d*>(take(b_ws));
    hipStream_t st = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemcpyAsync(d_items, items, b_items, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_order, order.data(), b_order, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_gates, gates, b_gates, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_gamma, gamma, b_rates, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_lam, lambda, b_rates, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_traj_reset, dim3(1), dim3(1), 0, st, d_counter);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_traj, dim3(blocks), dim3(256), 0, st, d_items, n_items, d_order, d_gates, d_gamma, d_lam,
                           seed, d_ws, slot, d_counter, d_z, d_j, d_h);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(z, d_z, b_z, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && jumps) e = hipMemcpyAsync(jumps, d_j, b_j, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && hashes) e = hipMemcpyAsync(hashes, d_h, b_h, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && state) {
        // item 0's final state is in the slot of whichever workgroup ran it last: rerun it alone.
        e = hipStreamSynchronize(st);
        if (e == hipSuccess) {
            const int zero = 0;
            e = hipMemcpyAsync(d_order, &zero, sizeof(int), hipMemcpyHostToDevice, st);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_traj_reset, dim3(1), dim3(1), 0, st, d_counter);
                hipLaunchKernelGGL(k_traj, dim3(1), dim3(256), 0, st, d_items, 1, d_order, d_gates, d_gamma, d_lam,
                                   seed, d_ws, slot, d_counter, d_z, d_j, d_h);
                e = hipGetLastError();
            }
            if (e == hipSuccess)
                e = hipMemcpyAsync(state, d_ws, sizeof(v2d) << items[0].n_qubits, hipMemcpyDeviceToHost, st);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_traj_run: %s", hipGetErrorString(e));
    return LLSV_OK;
}
