// sv_traj.hip -- batched gate-level noisy trajectories in one persistent launch.
//
// Workgroups take items from a device work counter until it runs past the batch, so a mixed
// batch of 10..16-qubit circuits balances itself (items are handed out in decreasing size).  Each
// workgroup owns a state slot of 2^max_n amplitudes in the workspace (256 KiB at 14 qubits,
// resident in L2 / Infinity Cache while the item runs).
//
// Every noise site's Philox4x32-10 uniform (key = seed, counter = (trajectory, gate, site)) is
// drawn before its gate runs.  A site whose uniform lies below 1 - rate cannot jump, whatever the
// state; there the operator is the fixed no-jump product, which joins the pending 2x2 operator of
// its qubit instead of touching the state.  Pending operators reach the state inside the pass of
// the next CNOT on their qubit, or when a site that may jump needs the joint weights of its
// qubits: then every pending operator is applied (two qubits per pass), the weights are reduced
// over the workgroup, and each thread takes the same decisions from them (AD then PD on each
// touched qubit, weights updated in closed form after each outcome).  The state is renormalised
// only at those decisions; <Z_q> divides by the final norm.
// Every wave reaches the exit test of the work loop; nothing waits on another workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <memory>
#include <mutex>
#include <numeric>
#include <tuple>
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

// Sums NV values over the WG threads of the block; the totals land in out[] (all threads).
template <int WG, int NV>
__device__ inline void block_sums(double* v, double (*lds)[WG / 64], double* out) {
    constexpr int W = WG / 64;
#pragma unroll
    for (int k = 0; k < NV; ++k)
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[k][w] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < W; i += 2) t += lds[k][i] + lds[k][i + 1];
        out[k] = t;
    }
    __syncthreads();
}

__global__ void k_traj_reset(unsigned* counter) { *counter = 0u; }
__global__ void k_traj_set_item(int* slot, int v) { *slot = v; }

// Joint weights of bits (sa, sb) of amplitude i: v[2 * bit_sa + bit_sb] += p (selects, so v stays
// in registers).  A one-qubit statistic uses sa = sb: v[0] / v[3] get the weight of bit 0 / 1.
__device__ __forceinline__ void acc4(double* v, uint64_t i, int sa, int sb, double p) {
    const int k = int(((i >> sa) & 1) * 2 + ((i >> sb) & 1));
    v[0] += k == 0 ? p : 0.0;
    v[1] += k == 1 ? p : 0.0;
    v[2] += k == 2 ? p : 0.0;
    v[3] += k == 3 ? p : 0.0;
}

// Index bits of local index l of a tile over qubit set T: bit j of l at the j-th lowest qubit of T.
__device__ __forceinline__ uint64_t tile_dep(uint32_t l, uint32_t T) {
    uint64_t o = 0;
    for (uint32_t m = T; m && l; m &= m - 1u, l >>= 1)
        if (l & 1u) o |= uint64_t(m & (~m + 1u));
    return o;
}

// Operator ma on qubit a, one pass over the pairs.
template <int WG, bool STATS>
__device__ inline void pass_one(v2d* s, uint64_t N, int a, const double* ma, int sa, int sb, double* v) {
    const uint64_t lomask = (uint64_t(1) << a) - 1, ab = uint64_t(1) << a;
    for (uint64_t p = threadIdx.x; p < N / 2; p += WG) {
        const uint64_t i = ((p & ~lomask) << 1) | (p & lomask);
        const v2d x0 = s[i], x1 = s[i | ab];
        const v2d y0 = ap(ma, x0, x1), y1 = ap(ma + 4, x0, x1);
        s[i] = y0;
        s[i | ab] = y1;
        if (STATS) {
            acc4(v, i, sa, sb, nrm2(y0));
            acc4(v, i | ab, sa, sb, nrm2(y1));
        }
    }
}

// ma on qubit a and mb on qubit b (a != b; they commute), then, with CNOT, the CNOT with control
// a and target b; one pass over groups of 4 amplitudes.
template <int WG, bool STATS, bool CNOT>
__device__ inline void pass_two(v2d* s, uint64_t N, int a, const double* ma, int b, const double* mb, int sa, int sb,
                                double* v) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    const uint64_t ab = uint64_t(1) << a, bb = uint64_t(1) << b;
    for (uint64_t p = threadIdx.x; p < N / 4; p += WG) {
        const uint64_t i = insert0(insert0(p, lo), hi);
        const v2d x00 = s[i], x10 = s[i | ab], x01 = s[i | bb], x11 = s[i | ab | bb];  // x(a bit)(b bit)
        const v2d y00 = ap(ma, x00, x10), y10 = ap(ma + 4, x00, x10);
        const v2d y01 = ap(ma, x01, x11), y11 = ap(ma + 4, x01, x11);
        const v2d z00 = ap(mb, y00, y01), z01 = ap(mb + 4, y00, y01);
        const v2d z10 = ap(mb, y10, y11), z11 = ap(mb + 4, y10, y11);
        const v2d f10 = CNOT ? z11 : z10, f11 = CNOT ? z10 : z11;  // control a set: target b flipped
        s[i] = z00;
        s[i | ab] = f10;
        s[i | bb] = z01;
        s[i | ab | bb] = f11;
        if (STATS) {
            acc4(v, i, sa, sb, nrm2(z00));
            acc4(v, i | ab, sa, sb, nrm2(f10));
            acc4(v, i | bb, sa, sb, nrm2(z01));
            acc4(v, i | ab | bb, sa, sb, nrm2(f11));
        }
    }
}

// Two CNOTs on disjoint qubit pairs, (c0, t0) then (c1, t1), in one pass over groups of 16
// amplitudes: for each pair in turn, the pending operators of its control and target, then the
// CNOT -- the arithmetic of two pass_two calls, with one read and write of the state.
template <int WG>
__device__ inline void pass_four(v2d* s, uint64_t N, int c0, const double* mc0, int t0, const double* mt0, int c1,
                                 const double* mc1, int t1, const double* mt1) {
    int pos[4] = {c0, t0, c1, t1};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3 - i; ++j)
            if (pos[j] > pos[j + 1]) {
                const int x = pos[j];
                pos[j] = pos[j + 1];
                pos[j + 1] = x;
            }
    // amplitude j of a group: bits c0, t0, c1, t1 of the index = bits 0, 1, 2, 3 of j
    const uint64_t bit[4] = {uint64_t(1) << c0, uint64_t(1) << t0, uint64_t(1) << c1, uint64_t(1) << t1};
    const double* op[4] = {mc0, mt0, mc1, mt1};
    for (uint64_t p = threadIdx.x; p < N / 16; p += WG) {
        const uint64_t i = insert0(insert0(insert0(insert0(p, pos[0]), pos[1]), pos[2]), pos[3]);
        v2d x[16];
#pragma unroll
        for (int j = 0; j < 16; ++j)
            x[j] = s[i | ((j & 1) ? bit[0] : 0) | ((j & 2) ? bit[1] : 0) | ((j & 4) ? bit[2] : 0) | ((j & 8) ? bit[3] : 0)];
#pragma unroll
        for (int pr = 0; pr < 2; ++pr) {
            const int K0 = 1 << (2 * pr), K1 = 2 << (2 * pr);  // control bit, target bit of j
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (!(j & K0)) {
                    const v2d y0 = ap(op[2 * pr], x[j], x[j | K0]), y1 = ap(op[2 * pr] + 4, x[j], x[j | K0]);
                    x[j] = y0;
                    x[j | K0] = y1;
                }
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (!(j & K1)) {
                    const v2d y0 = ap(op[2 * pr + 1], x[j], x[j | K1]), y1 = ap(op[2 * pr + 1] + 4, x[j], x[j | K1]);
                    x[j] = y0;
                    x[j | K1] = y1;
                }
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if ((j & K0) && !(j & K1)) {  // control set: target flipped
                    const v2d y = x[j];
                    x[j] = x[j | K1];
                    x[j | K1] = y;
                }
        }
#pragma unroll
        for (int j = 0; j < 16; ++j)
            s[i | ((j & 1) ? bit[0] : 0) | ((j & 2) ? bit[1] : 0) | ((j & 4) ? bit[2] : 0) | ((j & 8) ? bit[3] : 0)] = x[j];
    }
}

// Held CNOTs of a tile pass: at most kMaxHeld, their qubits within one tile of at most kTileQubits.
constexpr int kMaxHeld = 16, kTileQubits = 12;

// One pass of the state of n qubits through LDS tiles: the held CNOTs h < nh in order (each with
// its control's and its target's operators before it, hm[h][0] and hm[h][1]: the arithmetic of
// pass_two), then the operator pend[q] on every qubit q of m1 (two qubits per sweep of the tile),
// and with `stats` the joint weights of bits (sa, sb) of the result into v (acc4).  The tile set T
// holds those qubits (at most min(n, kTileQubits), the caller's to ensure) plus the lowest other
// qubits up to that many bits (local bit j = the j-th lowest qubit of T); each of the
// 2^(n - |T|) tiles is loaded into `tile` once (full tiles: the next one's loads in flight while
// the current one is worked on) and goes back once.
// LDS slot of tile element l: l ^ (bits 4..7 of l), an XOR-linear bijection on any tile of 16 or
// more elements that spreads the strided accesses of sweeps over low bits across the banks.
__device__ __forceinline__ uint32_t tsw(uint32_t l) { return l ^ ((l >> 4) & 15u); }

template <int WG>
__device__ __forceinline__ void tile_ops(v2d* tile, uint32_t ts, uint32_t T, int nh, const int (*hq)[2],
                                const double (*hm)[2][8], uint32_t m1, const double (*pend)[8]) {
    const uint32_t tid = threadIdx.x;
    for (int h = 0; h < nh; ++h) {
        const int a = __builtin_popcount(T & ((1u << hq[h][0]) - 1u));  // local control
        const int b = __builtin_popcount(T & ((1u << hq[h][1]) - 1u));  // local target
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        const uint32_t ab = 1u << a, bb = 1u << b;
        double ma[8], mb[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            ma[j] = hm[h][0][j];
            mb[j] = hm[h][1][j];
        }
        const uint32_t sa = tsw(ab), sb = tsw(bb), sab = sa ^ sb;
        for (uint32_t p = tid; p < ts / 4; p += WG) {
            const uint32_t i = tsw(uint32_t(insert0(insert0(p, lo), hi)));
            const v2d x00 = tile[i], x10 = tile[i ^ sa], x01 = tile[i ^ sb], x11 = tile[i ^ sab];
            const v2d y00 = ap(ma, x00, x10), y10 = ap(ma + 4, x00, x10);
            const v2d y01 = ap(ma, x01, x11), y11 = ap(ma + 4, x01, x11);
            const v2d z00 = ap(mb, y00, y01), z01 = ap(mb + 4, y00, y01);
            const v2d z10 = ap(mb, y10, y11), z11 = ap(mb + 4, y10, y11);
            tile[i] = z00;
            tile[i ^ sa] = z11;  // control set: target flipped
            tile[i ^ sb] = z01;
            tile[i ^ sab] = z10;
        }
        __syncthreads();
    }
    while (m1) {
        const int qa = __builtin_ctz(m1);
        m1 &= m1 - 1u;
        const int a = __builtin_popcount(T & ((1u << qa) - 1u));
        double ma[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) ma[j] = pend[qa][j];
        if (m1) {  // two operators per sweep (they commute)
            const int qb = __builtin_ctz(m1);
            m1 &= m1 - 1u;
            const int b = __builtin_popcount(T & ((1u << qb) - 1u));
            double mb[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) mb[j] = pend[qb][j];
            const int lo = a < b ? a : b, hi = a < b ? b : a;
            const uint32_t ab = 1u << a, bb = 1u << b;
            const uint32_t sa = tsw(ab), sb = tsw(bb), sab = sa ^ sb;
            for (uint32_t p = tid; p < ts / 4; p += WG) {
                const uint32_t i = tsw(uint32_t(insert0(insert0(p, lo), hi)));
                const v2d x00 = tile[i], x10 = tile[i ^ sa], x01 = tile[i ^ sb], x11 = tile[i ^ sab];
                const v2d y00 = ap(ma, x00, x10), y10 = ap(ma + 4, x00, x10);
                const v2d y01 = ap(ma, x01, x11), y11 = ap(ma + 4, x01, x11);
                tile[i] = ap(mb, y00, y01);
                tile[i ^ sb] = ap(mb + 4, y00, y01);
                tile[i ^ sa] = ap(mb, y10, y11);
                tile[i ^ sab] = ap(mb + 4, y10, y11);
            }
        } else {
            const uint32_t ab = 1u << a;
            const uint32_t sa = tsw(ab);
            for (uint32_t p = tid; p < ts / 2; p += WG) {
                const uint32_t i = tsw(uint32_t(insert0(p, a)));
                const v2d x0 = tile[i], x1 = tile[i ^ sa];
                tile[i] = ap(ma, x0, x1);
                tile[i ^ sa] = ap(ma + 4, x0, x1);
            }
        }
        __syncthreads();
    }
}

template <int WG>
__device__ __forceinline__ void pass_tile(v2d* s, int n, uint32_t hmask, int nh, const int (*hq)[2], const double (*hm)[2][8],
                                 uint32_t m1, const double (*pend)[8], bool stats, int sa, int sb, double* v,
                                 v2d* tile) {
    const int tb = n < kTileQubits ? n : kTileQubits;
    uint32_t T = hmask | m1;
    for (int q = 0; q < n && __builtin_popcount(T) < tb; ++q) T |= 1u << q;
    const uint32_t ts = 1u << tb;
    const uint64_t ntiles = uint64_t(1) << (n - tb);
    const uint32_t tid = threadIdx.x;
    auto tile_base = [&](uint64_t w) {
        uint64_t base = w;
        for (int q = 0; q < n; ++q)
            if ((T >> q) & 1u) base = insert0(base, q);
        return base;
    };
    if (ts < (1u << kTileQubits)) {  // a state of fewer qubits than a full tile: one tile, the whole state
        for (uint32_t l = tid; l < ts; l += WG) tile[tsw(l)] = s[tile_dep(l, T)];
        __syncthreads();
        tile_ops<WG>(tile, ts, T, nh, hq, hm, m1, pend);
        for (uint32_t l = tid; l < ts; l += WG) {
            const uint64_t i = tile_dep(l, T);
            const v2d x = tile[tsw(l)];
            s[i] = x;
            if (stats) acc4(v, i, sa, sb, nrm2(x));
        }
        __syncthreads();
        return;
    }
    // full tiles: element tid + k * WG of a tile is at base | dt | dk[k] (disjoint bits)
    constexpr int R = (1 << kTileQubits) / WG;
    const uint64_t dt = tile_dep(tid, T);
    uint64_t dk[R];
#pragma unroll
    for (int k = 0; k < R; ++k) dk[k] = tile_dep(uint32_t(k) * WG, T);
    v2d nxt[R];
    {
        const uint64_t b0 = tile_base(0) | dt;
#pragma unroll
        for (int k = 0; k < R; ++k) nxt[k] = s[b0 | dk[k]];
    }
    for (uint64_t w = 0; w < ntiles; ++w) {
        const uint64_t base = tile_base(w) | dt;
#pragma unroll
        for (int k = 0; k < R; ++k) tile[tsw(tid + uint32_t(k) * WG)] = nxt[k];
        __syncthreads();
        if (w + 1 < ntiles) {
            const uint64_t bn = tile_base(w + 1) | dt;
#pragma unroll
            for (int k = 0; k < R; ++k) nxt[k] = s[bn | dk[k]];
        }
        tile_ops<WG>(tile, ts, T, nh, hq, hm, m1, pend);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const v2d x = tile[tsw(tid + uint32_t(k) * WG)];
            s[base | dk[k]] = x;
            if (stats) acc4(v, base | dk[k], sa, sb, nrm2(x));
        }
        __syncthreads();
    }
}

// Bits of the possible-jump mask of a gate: site s may jump only if u_s >= 1 - rate (a jump needs
// u >= 1 - rate * p with p <= 1, both sides rounded the way decide() rounds them).
__device__ inline unsigned possible_sites(const TG& g, const double* gamma, const double* lam, int first_rate,
                                          uint64_t seed, uint64_t traj, uint32_t gi) {
    unsigned m = 0;
    const int nq = g.kind == 0 ? 1 : 2;
    for (int k = 0; k < nq; ++k) {
        const int q = k == 0 ? g.q0 : g.q1;
        const double r0 = gamma[first_rate + q], r1 = lam[first_rate + q];
        if (uniform(seed, traj, gi, uint32_t(2 * k)) >= __dsub_rn(1.0, r0)) m |= 1u << (2 * k);
        if (uniform(seed, traj, gi, uint32_t(2 * k + 1)) >= __dsub_rn(1.0, r1)) m |= 1u << (2 * k + 1);
    }
    return m;
}

// The no-jump product PD0 AD0 = diag(1, sqrt(1 - l) sqrt(1 - g)) (unnormalised).
__device__ inline void no_jump_op(double g, double l, double* k) {
    for (int j = 0; j < 8; ++j) k[j] = 0.0;
    k[0] = 1.0;
    k[6] = sqrt(1.0 - l) * sqrt(1.0 - g);
}

__device__ inline void copy8(const double* a, double* b) {
    for (int j = 0; j < 8; ++j) b[j] = a[j];
}

constexpr int TRAJ_CHUNK = 256;  // gates whose possible-jump masks are drawn together

// WG threads per trajectory.  Decisions are the reference's (src/TrajectorySolver.h:425, with the
// joint-weight updates of decide()); what changes is when the state is touched:
//   * every site's uniform is drawn ahead (TRAJ_CHUNK gates at a time, one gate per thread); a
//     site with u < 1 - rate cannot jump whatever the state, so its operator is the fixed no-jump
//     product and no probability is needed;
//   * the state is not renormalised between decisions (a decision only uses ratios of weights),
//     so a gate with no possible jump is a 2x2 product on the host-free side: thread 0 multiplies
//     gate and no-jump operator into the pending operator of the gate's qubit (s_pend);
//   * DEFER: pending operators stay pending until a CNOT touches their qubit (applied in the
//     CNOT's pass) or a possible jump needs weights (then every pending operator is applied, two
//     qubits per pass, the last pass accumulating the weights); !DEFER: each gate's operators are
//     applied right after it, one pass per gate (plus one for a CNOT's noise).
//   * DEFER, WG <= 512 (1024-thread workgroups have half the registers per thread and would
//     spill the 16-amplitude groups): a CNOT with no possible jump is held, with a copy of the
//     pending operators of its two qubits; when the next such CNOT is on two other qubits, both
//     run in one pass (pass_four), else the held one runs alone first.  Operators pushed on a
//     held CNOT's qubits after it stay pending, so they apply after it; anything that reads or
//     needs the state runs it first.
//   * TILE (with DEFER): CNOTs with no possible jump are held in a list instead (s_hq, s_hm: their
//     qubits and the pending operators of both before each), up to kMaxHeld of them on at most
//     kTileQubits qubits; the list runs in one pass through LDS tiles of the state (pass_tile)
//     when the next one would not fit, or before anything reads or needs the state.
// At a possible jump every thread runs decide() on the same block sums, so the outcome needs no
// broadcast; the chosen operator, scaled to renormalise, becomes the qubit's pending operator.
// <Z_q> divides by the final norm.
template <int WG, bool DEFER, bool TILE = false>
__global__ __launch_bounds__(WG, TILE && WG == 512 ? 4 : 1) void k_traj(const TItem* __restrict__ items, int n_items,
                                             const int* __restrict__ order, const TG* __restrict__ gates,
                                             const double* __restrict__ gamma, const double* __restrict__ lam,
                                             const uint64_t* __restrict__ seedp, v2d* __restrict__ ws,
                                             uint64_t slot_size, unsigned* __restrict__ counter,
                                             double* __restrict__ zout, unsigned* __restrict__ jout,
                                             uint64_t* __restrict__ hout, int norm_state) {
    constexpr bool kPair = DEFER && WG <= 512 && !TILE;  // held CNOTs run in pairs (pass_four)
    static_assert(!TILE || DEFER, "tile passes hold deferred CNOTs");
    __shared__ int s_item;
    __shared__ int s_hq[TILE ? kMaxHeld : 1][2];
    __shared__ double s_hm[TILE ? kMaxHeld : 1][2][8];
    extern __shared__ v2d s_tile[];  // TILE: 2^kTileQubits amplitudes (dynamic)
    __shared__ double s_pend[20][8];
    __shared__ unsigned char s_pm[TRAJ_CHUNK];
    __shared__ double lds[8][WG / 64];
    v2d* s = ws + uint64_t(blockIdx.x) * slot_size;
    const int tid = threadIdx.x;
    const uint64_t seed = *seedp;
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
        for (uint64_t i = tid; i < N; i += WG) {
            v2d z;
            z.x = (i == 0) ? 1.0 : 0.0;
            z.y = 0.0;
            s[i] = z;
        }
        unsigned jumps = 0;
        uint64_t hash = 0xCBF29CE484222325ull;
        uint32_t pmask = 0;  // qubits with a pending operator in s_pend (the same in every thread)

        // Applies the pending operators of the qubits in `mask`, two per pass; the last pass
        // accumulates the joint weights of bits (sa, sb) into v when STATS.
        auto flush = [&](uint32_t mask, bool stats, int sa, int sb, double* v) {
            __syncthreads();
            while (mask) {
                const int a = __builtin_ctz(mask);
                mask &= mask - 1;
                double ma[8], mb[8];
                copy8(s_pend[a], ma);
                if (mask) {
                    const int b = __builtin_ctz(mask);
                    mask &= mask - 1;
                    copy8(s_pend[b], mb);
                    if (stats && !mask)
                        pass_two<WG, true, false>(s, N, a, ma, b, mb, sa, sb, v);
                    else
                        pass_two<WG, false, false>(s, N, a, ma, b, mb, sa, sb, v);
                } else if (stats) {
                    pass_one<WG, true>(s, N, a, ma, sa, sb, v);
                } else {
                    pass_one<WG, false>(s, N, a, ma, sa, sb, v);
                }
                __syncthreads();
            }
        };
        int nh = 0;          // TILE: held CNOTs in s_hq / s_hm
        uint32_t hmask = 0;  // TILE: their qubits
        // TILE: the held CNOTs, then the pending operators of the qubits in mask (and with stats the
        // weights of bits (sa, sb) of the result), in as few tile passes as the tile's qubits allow;
        // clears mask from pmask.
        auto flush_tile = [&](uint32_t mask, bool stats, int sa, int sb, double* v) __attribute__((always_inline)) {
            mask &= pmask;
            pmask &= ~mask;
            if (!mask && !nh && !stats) return;
            const int tb = n < kTileQubits ? n : kTileQubits;
            do {
                uint32_t part = 0;
                for (uint32_t r = mask; r; r &= r - 1u) {
                    const uint32_t bit = r & (~r + 1u);
                    if (__builtin_popcount(hmask | part | bit) <= tb) part |= bit;
                }
                const bool last = part == mask;
                __syncthreads();
                pass_tile<WG>(s, n, hmask, nh, s_hq, s_hm, part, s_pend, stats && last, sa, sb, v, s_tile);
                nh = 0;
                hmask = 0;
                mask &= ~part;
            } while (mask);
        };
        // the held CNOT (DEFER), if any: control hc, target ht, their operators before it
        bool hold = false;
        int hc = 0, ht = 0;
        double hmc[8], hmt[8];
        auto run_held = [&] {
            if (TILE) {
                flush_tile(0u, false, 0, 0, nullptr);
                return;
            }
            if (!hold) return;
            __syncthreads();
            pass_two<WG, false, true>(s, N, hc, hmc, ht, hmt, hc, ht, nullptr);
            __syncthreads();
            hold = false;
        };
        // TILE: appends CNOT (c, t) with the pending operators of its qubits to the held list
        // (running the list first when it would not fit).
        auto hold_cnot = [&](int c, int t) {
            const uint32_t ct = (1u << c) | (1u << t);
            if (nh == kMaxHeld || __builtin_popcount(hmask | ct) > (n < kTileQubits ? n : kTileQubits)) run_held();
            __syncthreads();  // pushes to s_pend visible, the last pass done with s_hm
            if (tid == 0) {
                s_hq[nh][0] = c;
                s_hq[nh][1] = t;
                for (int e = 0; e < 2; ++e) {
                    const int q = e ? t : c;
                    if (pmask & (1u << q)) {
                        copy8(s_pend[q], s_hm[nh][e]);
                    } else {
                        for (int j = 0; j < 8; ++j) s_hm[nh][e][j] = (j == 0 || j == 6) ? 1.0 : 0.0;
                    }
                }
            }
            ++nh;
            hmask |= ct;
            pmask &= ~ct;
        };
        // s_pend[q] <- a * s_pend[q] (or a when q has nothing pending); thread 0 only.
        auto push = [&](int q, const double* a) {
            if (tid == 0) {
                if (pmask & (1u << q))
                    mm2(a, s_pend[q], s_pend[q]);
                else
                    copy8(a, s_pend[q]);
            }
            pmask |= 1u << q;
        };

        for (int gi = 0; gi < item.n_gates; ++gi) {
            if (gi % TRAJ_CHUNK == 0) {
                __syncthreads();
                for (int k = tid; k < TRAJ_CHUNK && gi + k < item.n_gates; k += WG)
                    s_pm[k] = (unsigned char)possible_sites(gates[item.first_gate + gi + k], gamma, lam, item.first_rate,
                                                            seed, item.traj, uint32_t(gi + k));
                __syncthreads();
            }
            const unsigned pm = s_pm[gi % TRAJ_CHUNK];
            const TG g = gates[item.first_gate + gi];
            if (g.kind == 0) {
                const int q = g.q0;
                const double gq = gamma[item.first_rate + q], lq = lam[item.first_rate + q];
                if (!pm) {
                    double k[8], kg[8];
                    no_jump_op(gq, lq, k);
                    mm2(k, g.m, kg);
                    push(q, kg);
                } else {
                    push(q, g.m);
                    double v[4] = {0.0, 0.0, 0.0, 0.0}, P4[4];
                    if (TILE) {
                        flush_tile(pmask, true, q, q, v);
                    } else {
                        run_held();
                        flush(pmask, true, q, q, v);
                    }
                    pmask = 0;
                    block_sums<WG, 4>(v, lds, P4);
                    double P[2][2] = {{P4[0], 0.0}, {P4[3], 0.0}};
                    double k[8];
                    jumps += decide(P, gq, lq, seed, item.traj, uint32_t(gi), 0, k, &hash);
                    const double nr = 1.0 / sqrt(P[0][0] + P[1][0]);
                    for (int j = 0; j < 8; ++j) k[j] *= nr;
                    push(q, k);
                }
                if (!DEFER) {
                    flush(pmask, false, 0, 0, nullptr);
                    pmask = 0;
                }
            } else {
                const int c = g.q0, t = g.q1;
                const uint32_t ct = (1u << c) | (1u << t);
                if (TILE && !pm) {
                    // no possible jump: joins the held list (which runs first if this one does not fit)
                    hold_cnot(c, t);
                    double kc[8], kt[8];
                    no_jump_op(gamma[item.first_rate + c], lam[item.first_rate + c], kc);
                    no_jump_op(gamma[item.first_rate + t], lam[item.first_rate + t], kt);
                    push(c, kc);  // thread 0 read s_pend[c], s_pend[t] above; it alone writes them
                    push(t, kt);
                    continue;
                }
                if (kPair && !pm) {
                    // no possible jump: held for one pass with the next CNOT on two other qubits
                    __syncthreads();
                    double mc[8] = {1, 0, 0, 0, 0, 0, 1, 0}, mt[8] = {1, 0, 0, 0, 0, 0, 1, 0};
                    if (pmask & (1u << c)) copy8(s_pend[c], mc);
                    if (pmask & (1u << t)) copy8(s_pend[t], mt);
                    if (hold && hc != c && hc != t && ht != c && ht != t) {
                        pass_four<WG>(s, N, hc, hmc, ht, hmt, c, mc, t, mt);
                        __syncthreads();
                        hold = false;
                    } else {
                        run_held();
                        hold = true;
                        hc = c;
                        ht = t;
                        copy8(mc, hmc);
                        copy8(mt, hmt);
                    }
                    pmask &= ~ct;
                    double kc[8], kt[8];
                    no_jump_op(gamma[item.first_rate + c], lam[item.first_rate + c], kc);
                    no_jump_op(gamma[item.first_rate + t], lam[item.first_rate + t], kt);
                    __syncthreads();  // every thread has read s_pend[c], s_pend[t] before thread 0 rewrites them
                    push(c, kc);
                    push(t, kt);
                    continue;
                }
                double v[4] = {0.0, 0.0, 0.0, 0.0};
                if (TILE) {
                    // (a possible jump here) the CNOT joins the held list, and one pass runs the
                    // list, every other pending operator and the weights of (c, t)
                    hold_cnot(c, t);
                    flush_tile(pmask, true, c, t, v);
                } else {
                    run_held();
                    if (pm) {  // weights after the CNOT need every other pending operator applied first
                        flush(pmask & ~ct, false, 0, 0, nullptr);
                        pmask &= ct;
                    }
                    __syncthreads();
                    double mc[8] = {1, 0, 0, 0, 0, 0, 1, 0}, mt[8] = {1, 0, 0, 0, 0, 0, 1, 0};
                    if (pmask & (1u << c)) copy8(s_pend[c], mc);
                    if (pmask & (1u << t)) copy8(s_pend[t], mt);
                    if (pm)
                        pass_two<WG, true, true>(s, N, c, mc, t, mt, c, t, v);
                    else
                        pass_two<WG, false, true>(s, N, c, mc, t, mt, c, t, v);
                    __syncthreads();
                    pmask &= ~ct;
                }
                const double gc = gamma[item.first_rate + c], lc = lam[item.first_rate + c];
                const double gt = gamma[item.first_rate + t], lt = lam[item.first_rate + t];
                double kc[8], kt[8];
                if (!pm) {
                    no_jump_op(gc, lc, kc);
                    no_jump_op(gt, lt, kt);
                } else {
                    double P4[4];
                    block_sums<WG, 4>(v, lds, P4);
                    double P[2][2] = {{P4[0], P4[1]}, {P4[2], P4[3]}};
                    jumps += decide(P, gc, lc, seed, item.traj, uint32_t(gi), 0, kc, &hash);
                    double Q[2][2] = {{P[0][0], P[1][0]}, {P[0][1], P[1][1]}};
                    jumps += decide(Q, gt, lt, seed, item.traj, uint32_t(gi), 2, kt, &hash);
                    const double nr = 1.0 / sqrt((Q[0][0] + Q[0][1]) + (Q[1][0] + Q[1][1]));
                    for (int j = 0; j < 8; ++j) kt[j] *= nr;
                }
                push(c, kc);
                push(t, kt);
                if (!DEFER) {
                    flush(pmask, false, 0, 0, nullptr);
                    pmask = 0;
                }
            }
        }
        if (TILE) {
            flush_tile(pmask, false, 0, 0, nullptr);
        } else {
            run_held();
            flush(pmask, false, 0, 0, nullptr);
        }
        pmask = 0;
        // <Z_q> = sum_i p_i (1 - 2 bit_q(i)) / sum_i p_i, seven qubits per read of the state (few
        // accumulators, so the loop stays in registers); the norm comes with the first seven.
        double norm = 1.0;
        for (int q0 = 0; q0 < n; q0 += 7) {
            double zq[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) zq[j] = 0.0;
            for (uint64_t i = tid; i < N; i += WG) {
                const double p = nrm2(s[i]);
                zq[7] += p;
#pragma unroll
                for (int j = 0; j < 7; ++j) zq[j] += ((i >> (q0 + j)) & 1) ? -p : p;
            }
            double Z[8];
            block_sums<WG, 8>(zq, lds, Z);
            if (q0 == 0) norm = Z[7];
            if (tid == 0) {
#pragma unroll
                for (int j = 0; j < 7; ++j)
                    if (q0 + j < n) zout[size_t(it) * LLSV_TRAJ_ZSTRIDE + q0 + j] = Z[j] / norm;
            }
        }
        if (norm_state) {  // the caller reads this state back: normalise it
            const double r = 1.0 / sqrt(norm);
            for (uint64_t i = tid; i < N; i += WG) s[i] = s[i] * r;
        }
        if (tid == 0) {
            jout[it] = jumps;
            hout[it] = hash;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- host: buffers, launch, tuning
struct TrajConfig {
    int wg = 256;       // threads per trajectory (workgroup size)
    int per_cu = 4;     // persistent workgroups per CU
    int fuse = 1;       // 0: one pass per gate; 1: deferred operators, CNOTs held in pairs;
                        // 2: deferred operators, CNOTs held in lists run through LDS tiles
};

int device_cus() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return cus;
}

// Workgroups of a launch: at most n_items, and at most what a 4 GiB workspace holds.
int traj_blocks(const TrajConfig& c, int n_items, int maxn, int cus) {
    const uint64_t slot_bytes = uint64_t(16) << maxn;
    const uint64_t cap = std::max<uint64_t>(1, (uint64_t(4) << 30) / slot_bytes);
    return int(std::min<uint64_t>({uint64_t(n_items), uint64_t(c.per_cu) * uint64_t(cus), cap}));
}

hipError_t launch_traj(const TrajConfig& c, int blocks, hipStream_t st, const TItem* items, int n_items,
                       const int* order, const TG* gates, const double* gamma, const double* lam, const uint64_t* seed,
                       v2d* ws, uint64_t slot, unsigned* counter, double* z, unsigned* j, uint64_t* h,
                       int norm_state = 0) {
    hipLaunchKernelGGL(k_traj_reset, dim3(1), dim3(1), 0, st, counter);
#define LLSV_TRAJ_LAUNCH(WG, F, T)                                                                             \
    hipLaunchKernelGGL((k_traj<WG, F, T>), dim3(blocks), dim3(WG), T ? kTileBytes : 0, st, items, n_items, order, \
                       gates, gamma, lam, seed, ws, slot, counter, z, j, h, norm_state)
    constexpr size_t kTileBytes = size_t(16) << kTileQubits;
    if (c.wg == 1024) {
        if (c.fuse == 2) LLSV_TRAJ_LAUNCH(1024, true, true);
        else if (c.fuse) LLSV_TRAJ_LAUNCH(1024, true, false);
        else LLSV_TRAJ_LAUNCH(1024, false, false);
    } else if (c.wg == 512) {
        if (c.fuse == 2) LLSV_TRAJ_LAUNCH(512, true, true);
        else if (c.fuse) LLSV_TRAJ_LAUNCH(512, true, false);
        else LLSV_TRAJ_LAUNCH(512, false, false);
    } else {
        if (c.fuse == 2) LLSV_TRAJ_LAUNCH(256, true, true);
        else if (c.fuse) LLSV_TRAJ_LAUNCH(256, true, false);
        else LLSV_TRAJ_LAUNCH(256, false, false);
    }
#undef LLSV_TRAJ_LAUNCH
    return hipGetLastError();
}

// Shape of a batch for the tuning cache: qubit range, log2 of the item count and of the mean
// gate count per item.
struct TrajShape {
    int maxn, minn, lg_items, lg_gates;
    bool operator<(const TrajShape& o) const {
        return std::tie(maxn, minn, lg_items, lg_gates) < std::tie(o.maxn, o.minn, o.lg_items, o.lg_gates);
    }
};

bool valid_config(const TrajConfig& c) {
    return (c.wg == 256 || c.wg == 512 || c.wg == 1024) && (c.per_cu == 1 || c.per_cu == 2 || c.per_cu == 4) &&
           (c.fuse == 0 || c.fuse == 1 || c.fuse == 2) && c.wg * c.per_cu <= 2048;
}

// ---- tuning cache --------------------------------------------------------------------------
// Choices of the tuner, per batch shape, for the process; with LLSV_TUNE_CACHE=<file> also kept
// on disk as JSON (read on first use, rewritten after every new choice), so later processes skip
// the tuning runs.  The file is a list of flat objects:
//   {"format": "llsv-traj-tune-1", "entries": [
//     {"maxn": 14, "minn": 14, "lg_items": 12, "lg_gates": 8, "threads": 512, "per_cu": 2, "fused": 1}, ...]}
// LLSV_TRAJ_CONFIG=<threads>,<per_cu>,<fused> (e.g. 512,2,1) sets the default configuration and
// takes the tuner's place: plans use it instead of tuning.
std::mutex g_tune_mu;
TrajConfig g_default;  // used when no tuned choice is cached (guarded by g_tune_mu)
bool g_forced = false;  // LLSV_TRAJ_CONFIG given: no tuning
std::map<TrajShape, TrajConfig>& tune_cache_raw() {
    static auto* m = new std::map<TrajShape, TrajConfig>();
    return *m;
}

// The integer after "key": in [b, e), or fails.
bool json_int(const char* b, const char* e, const char* key, int* out) {
    const std::string k = std::string("\"") + key + "\"";
    const char* p = std::search(b, e, k.begin(), k.end());
    if (p == e) return false;
    p += k.size();
    while (p < e && (*p == ' ' || *p == ':' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    char* end = nullptr;
    const long v = std::strtol(p, &end, 10);
    if (end == p || end > e) return false;
    *out = int(v);
    return true;
}

// Parses a cache file's text; false (and nothing added) on any malformed entry.
bool parse_tune_json(const std::string& text, std::map<TrajShape, TrajConfig>* into) {
    if (text.find("\"llsv-traj-tune-1\"") == std::string::npos) return false;
    const size_t ent = text.find("\"entries\"");
    if (ent == std::string::npos) return false;
    std::map<TrajShape, TrajConfig> got;
    size_t pos = text.find('[', ent);
    if (pos == std::string::npos) return false;
    for (;;) {
        const size_t ob = text.find('{', pos), close = text.find(']', pos);
        if (ob == std::string::npos || (close != std::string::npos && close < ob)) break;
        const size_t cb = text.find('}', ob);
        if (cb == std::string::npos) return false;
        const char *b = text.data() + ob, *e = text.data() + cb;
        TrajShape sh{};
        TrajConfig c{};
        if (!json_int(b, e, "maxn", &sh.maxn) || !json_int(b, e, "minn", &sh.minn) ||
            !json_int(b, e, "lg_items", &sh.lg_items) || !json_int(b, e, "lg_gates", &sh.lg_gates) ||
            !json_int(b, e, "threads", &c.wg) || !json_int(b, e, "per_cu", &c.per_cu) || !json_int(b, e, "fused", &c.fuse))
            return false;
        if (!valid_config(c) || sh.minn < 1 || sh.maxn > 20 || sh.minn > sh.maxn) return false;
        got[sh] = c;
        pos = cb + 1;
    }
    for (const auto& kv : got) (*into)[kv.first] = kv.second;
    return true;
}

std::string tune_json(const std::map<TrajShape, TrajConfig>& m) {
    std::string s = "{\"format\": \"llsv-traj-tune-1\", \"entries\": [";
    bool first = true;
    for (const auto& kv : m) {
        char buf[256];
        snprintf(buf, sizeof buf,
                 "%s\n  {\"maxn\": %d, \"minn\": %d, \"lg_items\": %d, \"lg_gates\": %d, \"threads\": %d, "
                 "\"per_cu\": %d, \"fused\": %d}",
                 first ? "" : ",", kv.first.maxn, kv.first.minn, kv.first.lg_items, kv.first.lg_gates, kv.second.wg,
                 kv.second.per_cu, kv.second.fuse);
        s += buf;
        first = false;
    }
    s += "\n]}\n";
    return s;
}

bool read_file(const char* path, std::string* out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char buf[4096];
    size_t k;
    out->clear();
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) out->append(buf, k);
    std::fclose(f);
    return true;
}

// Writes through a temporary file and a rename, so a reader never sees half a file.
bool write_file(const char* path, const std::string& text) {
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    if (std::fclose(f) != 0 || !ok) return false;
    return std::rename(tmp.c_str(), path) == 0;
}

const char* cache_path() {
    static const char* p = [] {
        const char* e = std::getenv("LLSV_TUNE_CACHE");
        return (e && *e) ? e : nullptr;
    }();
    return p;
}

// The cache (caller holds g_tune_mu); the environment is read on first use.
std::map<TrajShape, TrajConfig>& tune_cache() {
    static bool init = false;
    if (!init) {
        init = true;
        std::string text;
        if (cache_path() && read_file(cache_path(), &text)) (void)parse_tune_json(text, &tune_cache_raw());
        if (const char* e = std::getenv("LLSV_TRAJ_CONFIG")) {
            TrajConfig c{};
            if (std::sscanf(e, "%d,%d,%d", &c.wg, &c.per_cu, &c.fuse) == 3 && valid_config(c)) {
                g_default = c;
                g_forced = true;
            }
        }
    }
    return tune_cache_raw();
}

// Records a new choice; rewrites the cache file (caller holds g_tune_mu).
void tune_store(const TrajShape& sh, const TrajConfig& c) {
    tune_cache()[sh] = c;
    if (cache_path()) (void)write_file(cache_path(), tune_json(tune_cache()));
}

int ilog2(uint64_t x) { return x ? 63 - __builtin_clzll(x) : 0; }

// Validates a batch; fills the largest-first order, the qubit range and the shape.
int check_batch(const llsv_traj_item* items, int n_items, const llsv_tgate* gates, int n_gates, const double* gamma,
                const double* lambda, int n_rates, const char* who, std::vector<int>* order, TrajShape* shape) {
    if (!items || n_items < 1 || !gates || n_gates < 1 || !gamma || !lambda || n_rates < 1)
        return fail(LLSV_ERR_ARG, "%s: empty or null argument", who);
    int maxn = 0, minn = 99;
    uint64_t gsum = 0;
    for (int i = 0; i < n_items; ++i) {
        const llsv_traj_item& it = items[i];
        if (it.n_qubits < 1 || it.n_qubits > 20) return fail(LLSV_ERR_ARG, "%s: item %d: n_qubits must be in [1, 20]", who, i);
        if (it.first_gate < 0 || it.n_gates < 0 || it.first_gate + it.n_gates > n_gates)
            return fail(LLSV_ERR_ARG, "%s: item %d: gate range outside the gate list", who, i);
        if (it.first_rate < 0 || it.first_rate + it.n_qubits > n_rates)
            return fail(LLSV_ERR_ARG, "%s: item %d: rate range outside the rate lists", who, i);
        for (int k = 0; k < it.n_gates; ++k) {
            const llsv_tgate& g = gates[it.first_gate + k];
            const bool ok = (g.kind == 0 && g.q0 >= 0 && g.q0 < it.n_qubits) ||
                            (g.kind == 1 && g.q0 >= 0 && g.q0 < it.n_qubits && g.q1 >= 0 && g.q1 < it.n_qubits && g.q0 != g.q1);
            if (!ok) return fail(LLSV_ERR_ARG, "%s: item %d gate %d: bad kind or qubits", who, i, k);
        }
        for (int q = 0; q < it.n_qubits; ++q) {
            const double g = gamma[it.first_rate + q], l = lambda[it.first_rate + q];
            if (!(g >= 0.0 && g <= 1.0 && l >= 0.0 && l <= 1.0))
                return fail(LLSV_ERR_ARG, "%s: item %d: rates must be in [0, 1]", who, i);
        }
        maxn = std::max(maxn, it.n_qubits);
        minn = std::min(minn, it.n_qubits);
        gsum += uint64_t(it.n_gates);
    }
    order->resize(n_items);
    std::iota(order->begin(), order->end(), 0);
    std::stable_sort(order->begin(), order->end(), [&](int a, int b) { return items[a].n_qubits > items[b].n_qubits; });
    *shape = TrajShape{maxn, minn, ilog2(uint64_t(n_items)), ilog2(gsum / uint64_t(n_items) + 1)};
    return LLSV_OK;
}

// Device copy of one batch: inputs, results, work counter, seed and the per-workgroup state slots.
struct TrajDev {
    unsigned char* buf = nullptr;
    TItem* items = nullptr;
    int* order = nullptr;
    int* sample = nullptr;  // tuning sample order
    TG* gates = nullptr;
    double *gamma = nullptr, *lam = nullptr, *z = nullptr;
    unsigned *j = nullptr, *counter = nullptr;
    uint64_t *h = nullptr, *seed = nullptr;
    v2d* ws = nullptr;
    uint64_t slot = 0;
    int max_blocks = 0;
    ~TrajDev() {
        if (buf) (void)hipFree(buf);
    }
};

int traj_upload(TrajDev* d, const llsv_traj_item* items, int n_items, const std::vector<int>& order,
                const llsv_tgate* gates, int n_gates, const double* gamma, const double* lambda, int n_rates,
                int maxn, int max_blocks, hipStream_t st, const char* who) {
    const size_t b_items = sizeof(llsv_traj_item) * n_items, b_order = sizeof(int) * n_items;
    const size_t b_gates = sizeof(llsv_tgate) * n_gates, b_rates = sizeof(double) * n_rates;
    const size_t b_z = sizeof(double) * LLSV_TRAJ_ZSTRIDE * n_items, b_j = sizeof(unsigned) * n_items;
    const size_t b_h = sizeof(uint64_t) * n_items;
    d->slot = uint64_t(1) << maxn;
    d->max_blocks = max_blocks;
    const size_t b_ws = d->slot * 16 * size_t(max_blocks);
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t total = al(b_items) + 2 * al(b_order) + al(b_gates) + 2 * al(b_rates) + al(b_z) + al(b_j) + al(b_h) +
                         512 + al(b_ws);
    if (dev_alloc(reinterpret_cast<void**>(&d->buf), total) != hipSuccess) {
        (void)hipGetLastError();
        d->buf = nullptr;
        return fail(LLSV_ERR_NOMEM, "%s: cannot allocate %zu bytes", who, total);
    }
    size_t off = 0;
    auto take = [&](size_t b) { unsigned char* p = d->buf + off; off += al(b); return p; };
    d->items = reinterpret_cast<TItem*>(take(b_items));
    d->order = reinterpret_cast<int*>(take(b_order));
    d->sample = reinterpret_cast<int*>(take(b_order));
    d->gates = reinterpret_cast<TG*>(take(b_gates));
    d->gamma = reinterpret_cast<double*>(take(b_rates));
    d->lam = reinterpret_cast<double*>(take(b_rates));
    d->z = reinterpret_cast<double*>(take(b_z));
    d->j = reinterpret_cast<unsigned*>(take(b_j));
    d->h = reinterpret_cast<uint64_t*>(take(b_h));
    d->counter = reinterpret_cast<unsigned*>(take(256));
    d->seed = reinterpret_cast<uint64_t*>(take(256));
    d->ws = reinterpret_cast<v2d*>(take(b_ws));
    hipError_t e = hipMemcpyAsync(d->items, items, b_items, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d->order, order.data(), b_order, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d->gates, gates, b_gates, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d->gamma, gamma, b_rates, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d->lam, lambda, b_rates, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(LLSV_ERR_DEVICE, "%s: upload: %s", who, hipGetErrorString(e));
    return LLSV_OK;
}

// The autotuner: times every candidate configuration (threads per trajectory 256 / 512 / 1024,
// 1 / 2 / 4 persistent workgroups per CU, one pass per gate / deferred operators with CNOTs held
// in pairs / in LDS tile lists) on a sample of the batch -- every k-th item, about 2048 in all,
// largest first -- and returns the fastest.
int traj_tune(TrajDev& d, const llsv_traj_item* items, int n_items, const std::vector<int>& order, int maxn,
              hipStream_t st, TrajConfig* best, double* tune_ms) {
    const int cus = device_cus();
    std::vector<int> sample;
    const int step = std::max(1, n_items / 2048);
    for (int i = 0; i < n_items; i += step) sample.push_back(order[i]);  // keeps the largest-first order
    hipError_t e = hipMemcpyAsync(d.sample, sample.data(), sizeof(int) * sample.size(), hipMemcpyHostToDevice, st);
    const uint64_t seed_probe = 0x5EEDull;
    if (e == hipSuccess) e = hipMemcpyAsync(d.seed, &seed_probe, sizeof seed_probe, hipMemcpyHostToDevice, st);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto t_start = std::chrono::steady_clock::now();
    float best_ms = 1e30f;
    for (int wg : {256, 512, 1024})
        for (int per_cu : {1, 2, 4})
            for (int fuse : {0, 1, 2}) {
                if (e != hipSuccess) break;
                if (wg * per_cu > 2048) continue;  // more threads than one CU holds at once
                if (fuse == 2 && per_cu > 2) continue;  // 64 KiB LDS tiles: at most two workgroups per CU
                const TrajConfig c{wg, per_cu, fuse};
                const int blocks = std::min(traj_blocks(c, int(sample.size()), maxn, cus), d.max_blocks);
                float ms = 0.0f;
                for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {  // first run warms up
                    e = hipEventRecord(e0, st);
                    if (e == hipSuccess)
                        e = launch_traj(c, blocks, st, d.items, int(sample.size()), d.sample, d.gates, d.gamma, d.lam,
                                        d.seed, d.ws, d.slot, d.counter, d.z, d.j, d.h);
                    if (e == hipSuccess) e = hipEventRecord(e1, st);
                    if (e == hipSuccess) e = hipEventSynchronize(e1);
                    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                }
                if (e == hipSuccess && ms < best_ms) {
                    best_ms = ms;
                    *best = c;
                }
            }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    *tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    if (e != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_traj_plan_create: tuning: %s", hipGetErrorString(e));
    (void)items;
    return LLSV_OK;
}

}  // namespace
}  // namespace llsv

// A prepared batch: device data, the chosen configuration and the captured run.
struct llsv_traj_plan {
    llsv::TrajDev dev;
    llsv::TrajConfig cfg;
    int n_items = 0, blocks = 0, maxn = 0;
    double tune_ms = 0.0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t* h_seed = nullptr;  // pinned
    double* h_z = nullptr;       // pinned results
    unsigned* h_j = nullptr;
    uint64_t* h_h = nullptr;
    ~llsv_traj_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamSynchronize(stream);
        if (h_seed) (void)hipHostFree(h_seed);
        if (h_z) (void)hipHostFree(h_z);
        if (h_j) (void)hipHostFree(h_j);
        if (h_h) (void)hipHostFree(h_h);
        if (stream) (void)hipStreamDestroy(stream);
    }
};



extern "C" int llsv_traj_run(const llsv_traj_item* items, int n_items, const llsv_tgate* gates, int n_gates,
                             const double* gamma, const double* lambda, int n_rates, uint64_t seed, double* z,
                             uint32_t* jumps, uint64_t* hashes, double* state) {
    using namespace llsv;
    if (!z) return fail(LLSV_ERR_ARG, "llsv_traj_run: z is null");
    std::vector<int> order;
    TrajShape shape{};
    if (int rc = check_batch(items, n_items, gates, n_gates, gamma, lambda, n_rates, "llsv_traj_run", &order, &shape))
        return rc;
    TrajConfig cfg;  // a configuration the tuner chose for this shape (this process or the cache file), else the default
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = tune_cache().find(shape);
        cfg = (it != tune_cache().end() && !g_forced) ? it->second : g_default;
    }
    const int cus = device_cus();
    const int blocks = traj_blocks(cfg, n_items, shape.maxn, cus);
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        return fail(LLSV_ERR_DEVICE, "llsv_traj_run: stream creation failed");
    int rc = LLSV_OK;
    {
        TrajDev d;
        rc = traj_upload(&d, items, n_items, order, gates, n_gates, gamma, lambda, n_rates, shape.maxn, blocks, st,
                         "llsv_traj_run");
        hipError_t e = hipSuccess;
        if (rc == LLSV_OK) {
            e = hipMemcpyAsync(d.seed, &seed, sizeof seed, hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = launch_traj(cfg, blocks, st, d.items, n_items, d.order, d.gates, d.gamma, d.lam, d.seed, d.ws,
                                d.slot, d.counter, d.z, d.j, d.h);
            if (e == hipSuccess)
                e = hipMemcpyAsync(z, d.z, sizeof(double) * LLSV_TRAJ_ZSTRIDE * n_items, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess && jumps) e = hipMemcpyAsync(jumps, d.j, sizeof(unsigned) * n_items, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess && hashes) e = hipMemcpyAsync(hashes, d.h, sizeof(uint64_t) * n_items, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess && state) {
                // item 0's final state is in the slot of whichever workgroup ran it last: rerun it alone.
                e = hipStreamSynchronize(st);
                if (e == hipSuccess) {
                    hipLaunchKernelGGL(k_traj_set_item, dim3(1), dim3(1), 0, st, d.sample, 0);
                    e = hipGetLastError();
                }
                if (e == hipSuccess)
                    e = launch_traj(cfg, 1, st, d.items, 1, d.sample, d.gates, d.gamma, d.lam, d.seed, d.ws, d.slot,
                                    d.counter, d.z, d.j, d.h, 1);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(state, d.ws, sizeof(v2d) << items[0].n_qubits, hipMemcpyDeviceToHost, st);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) rc = fail(LLSV_ERR_DEVICE, "llsv_traj_run: %s", hipGetErrorString(e));
        }
    }
    (void)hipStreamDestroy(st);
    return rc;
}

extern "C" int llsv_traj_plan_create(const llsv_traj_item* items, int n_items, const llsv_tgate* gates, int n_gates,
                                     const double* gamma, const double* lambda, int n_rates, int tune,
                                     llsv_traj_plan** out) {
    using namespace llsv;
    if (!out) return fail(LLSV_ERR_ARG, "llsv_traj_plan_create: out is null");
    *out = nullptr;
    std::vector<int> order;
    TrajShape shape{};
    if (int rc = check_batch(items, n_items, gates, n_gates, gamma, lambda, n_rates, "llsv_traj_plan_create", &order,
                             &shape))
        return rc;
    std::unique_ptr<llsv_traj_plan> p(new llsv_traj_plan());
    p->n_items = n_items;
    p->maxn = shape.maxn;
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(LLSV_ERR_DEVICE, "llsv_traj_plan_create: stream creation failed");
    const int cus = device_cus();
    // workspace for the largest candidate grid
    const int max_blocks = traj_blocks(TrajConfig{256, 4, 1}, n_items, shape.maxn, cus);
    if (int rc = traj_upload(&p->dev, items, n_items, order, gates, n_gates, gamma, lambda, n_rates, shape.maxn,
                             max_blocks, p->stream, "llsv_traj_plan_create"))
        return rc;
    bool cached = false;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = tune_cache().find(shape);
        p->cfg = g_default;
        if (g_forced) {
            cached = true;  // LLSV_TRAJ_CONFIG: the given configuration, no tuning
        } else if (tune >= 0 && it != tune_cache().end()) {
            p->cfg = it->second;
            cached = true;
        }
    }
    if (tune == 1 && !cached) {
        if (int rc = traj_tune(p->dev, items, n_items, order, shape.maxn, p->stream, &p->cfg, &p->tune_ms)) return rc;
        std::lock_guard<std::mutex> lock(g_tune_mu);
        tune_store(shape, p->cfg);
    }
    p->blocks = std::min(traj_blocks(p->cfg, n_items, shape.maxn, cus), max_blocks);
    if (dev_host_alloc(reinterpret_cast<void**>(&p->h_seed), 256) != hipSuccess ||
        dev_host_alloc(reinterpret_cast<void**>(&p->h_z), sizeof(double) * LLSV_TRAJ_ZSTRIDE * n_items) != hipSuccess ||
        dev_host_alloc(reinterpret_cast<void**>(&p->h_j), sizeof(unsigned) * n_items) != hipSuccess ||
        dev_host_alloc(reinterpret_cast<void**>(&p->h_h), sizeof(uint64_t) * n_items) != hipSuccess) {
        (void)hipGetLastError();
        return fail(LLSV_ERR_NOMEM, "llsv_traj_plan_create: cannot allocate pinned result buffers");
    }
    // Capture one run: seed upload, counter reset (a kernel, not a memset), the persistent kernel,
    // result copies into the pinned buffers.
    TrajDev& d = p->dev;
    hipError_t e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t ec = hipMemcpyAsync(d.seed, p->h_seed, sizeof(uint64_t), hipMemcpyHostToDevice, p->stream);
        if (ec == hipSuccess)
            ec = launch_traj(p->cfg, p->blocks, p->stream, d.items, n_items, d.order, d.gates, d.gamma, d.lam, d.seed,
                             d.ws, d.slot, d.counter, d.z, d.j, d.h);
        if (ec == hipSuccess)
            ec = hipMemcpyAsync(p->h_z, d.z, sizeof(double) * LLSV_TRAJ_ZSTRIDE * n_items, hipMemcpyDeviceToHost, p->stream);
        if (ec == hipSuccess) ec = hipMemcpyAsync(p->h_j, d.j, sizeof(unsigned) * n_items, hipMemcpyDeviceToHost, p->stream);
        if (ec == hipSuccess) ec = hipMemcpyAsync(p->h_h, d.h, sizeof(uint64_t) * n_items, hipMemcpyDeviceToHost, p->stream);
        e = hipStreamEndCapture(p->stream, &p->graph);
        if (ec != hipSuccess) e = ec;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(LLSV_ERR_DEVICE, "llsv_traj_plan_create: graph capture: %s", hipGetErrorString(e));
    }
    *out = p.release();
    return LLSV_OK;
}

extern "C" int llsv_traj_default_config(int threads, int per_cu, int fused) {
    using namespace llsv;
    if (!valid_config(TrajConfig{threads, per_cu, fused}))
        return fail(LLSV_ERR_ARG, "llsv_traj_default_config: threads in {256, 512, 1024}, per_cu in {1, 2, 4}, "
                                  "threads * per_cu <= 2048, fused in {0, 1, 2}");
    std::lock_guard<std::mutex> lock(g_tune_mu);
    (void)tune_cache();  // the environment's settings first, so this call overrides them
    g_default = TrajConfig{threads, per_cu, fused};
    return LLSV_OK;
}

extern "C" int llsv_traj_tune_load(const char* path) {
    using namespace llsv;
    if (!path) return fail(LLSV_ERR_ARG, "llsv_traj_tune_load: path is null");
    std::string text;
    if (!read_file(path, &text)) return fail(LLSV_ERR_ARG, "llsv_traj_tune_load: cannot read %s", path);
    std::lock_guard<std::mutex> lock(g_tune_mu);
    if (!parse_tune_json(text, &tune_cache()))
        return fail(LLSV_ERR_ARG, "llsv_traj_tune_load: %s is not a tuning cache (llsv-traj-tune-1)", path);
    return LLSV_OK;
}

extern "C" int llsv_traj_tune_save(const char* path) {
    using namespace llsv;
    if (!path) return fail(LLSV_ERR_ARG, "llsv_traj_tune_save: path is null");
    std::lock_guard<std::mutex> lock(g_tune_mu);
    if (!write_file(path, tune_json(tune_cache()))) return fail(LLSV_ERR_ARG, "llsv_traj_tune_save: cannot write %s", path);
    return LLSV_OK;
}

extern "C" int llsv_traj_tune_entries(void) {
    std::lock_guard<std::mutex> lock(llsv::g_tune_mu);
    return int(llsv::tune_cache().size());
}

extern "C" int llsv_traj_plan_run(llsv_traj_plan* p, uint64_t seed, double* z, uint32_t* jumps, uint64_t* hashes) {
    using namespace llsv;
    if (!p) return fail(LLSV_ERR_HANDLE, "llsv_traj_plan_run: null plan");
    if (!z) return fail(LLSV_ERR_ARG, "llsv_traj_plan_run: z is null");
    *p->h_seed = seed;
    hipError_t e = hipGraphLaunch(p->exec, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_traj_plan_run: %s", hipGetErrorString(e));
    const size_t k = size_t(p->n_items);
    std::memcpy(z, p->h_z, sizeof(double) * LLSV_TRAJ_ZSTRIDE * k);
    if (jumps) std::memcpy(jumps, p->h_j, sizeof(unsigned) * k);
    if (hashes) std::memcpy(hashes, p->h_h, sizeof(uint64_t) * k);
    return LLSV_OK;
}

extern "C" int llsv_traj_plan_config(const llsv_traj_plan* p, int* threads, int* workgroups, int* fused, double* tune_ms) {
    if (!p) return llsv::fail(LLSV_ERR_HANDLE, "llsv_traj_plan_config: null plan");
    if (threads) *threads = p->cfg.wg;
    if (workgroups) *workgroups = p->blocks;
    if (fused) *fused = p->cfg.fuse;
    if (tune_ms) *tune_ms = p->tune_ms;
    return LLSV_OK;
}

extern "C" int llsv_traj_plan_destroy(llsv_traj_plan* p) {
    delete p;
    return LLSV_OK;
}
