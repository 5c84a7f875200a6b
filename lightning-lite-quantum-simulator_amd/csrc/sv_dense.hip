// sv_dense.hip -- fused dense k-qubit blocks (k = 4, 5) on FP64 matrix cores.
//
// A dense block U (D x D complex, D = 2^k, operator bit j <-> target j) applied to every group
// of D amplitudes is the complex product Y = U X over all 2^(n-k) groups.  Its arithmetic
// intensity is 8 * 2^k flop per 32 bytes (8 flop/B at k = 5), close to the FP64 ridge, so
// it runs on v_mfma_f64_16x16x4f64 with the real embedding
//     Yr = Ur Xr - Ui Xi,   Yi = Ui Xr + Ur Xi        (four MFMAs per 4-column K step).
// A workgroup stages one tile of 2^11 (k = 5) or 2^12 (k = 4) amplitudes in LDS: the k targets plus
// the lowest other qubits (always qubits 0..2, so the HBM accesses are whole 128-byte lines).  The
// tile's groups are cut into 16-group column blocks; each wavefront takes column blocks
// in turn, reads its B operands (one amplitude per lane) from LDS, keeps U's A operands in
// registers for the whole tile, and writes the 16 x 16 result back to the same LDS slots
// (no other wavefront reads them).  Then the tile goes back to HBM.
//
// MFMA lane layout (checked on gfx950): A[i][k] is lane i + 16 k, B[k][j] is lane j + 16 k,
// D[i][j] is lane j + 16 (i % 4), element i / 4.
#include <algorithm>

#include "sv_common.h"
#include "sv_dense.h"
#include "sv_program.h"

namespace llsv {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t dswz(uint32_t l) { return l ^ ((l >> 4) & 15u); }

__device__ __forceinline__ uint64_t dense_tile_offset(const PassDesc& d, uint32_t l) {
    uint64_t o = l & ((1u << d.L) - 1);
    for (int j = d.L; j < d.b; ++j)
        if ((l >> j) & 1u) o |= uint64_t(1) << d.tq[j];
    return o;
}

template <int K>
__device__ __forceinline__ uint32_t tdep(const DenseTile& t, uint32_t c) {
    uint32_t o = 0;
#pragma unroll
    for (int j = 0; j < K; ++j)
        if ((c >> j) & 1u) o |= 1u << t.tloc[j];
    return o;
}

// Tile bits for a K-qubit block: 5-qubit blocks take tiles of 2^11 amplitudes (32 KiB of LDS, four
// workgroups per CU), 4-qubit blocks 2^12 (two per CU).  Measured on the 28-qubit layered circuit:
// 1.67 against 1.79 ms per 5-qubit block, 1.70 against 1.76 ms per 4-qubit block.
constexpr int dense_tile_bits(int k) { return k == 5 ? 11 : 12; }

template <int K>
__device__ __forceinline__ uint32_t gdep(const DenseTile& t, uint32_t g) {
    constexpr int kTB = dense_tile_bits(K);
    uint32_t o = 0;
#pragma unroll
    for (int j = 0; j < kTB - K; ++j)
        if ((g >> j) & 1u) o |= 1u << t.gbits[j];
    return o;
}

template <int K>
__global__ __launch_bounds__(256) void k_dense_mfma(v2d* __restrict__ psi, PassDesc d, DenseTile dt,
                                                    const v2d* __restrict__ U, uint32_t ntiles) {
    constexpr int D = 1 << K;
    constexpr int kTB = dense_tile_bits(K);
    constexpr int RT = D / 16;  // 16-row blocks of U
    constexpr int KS = D / 4;   // 4-column K steps
    constexpr int NCB = (1 << kTB) / D / 16;  // 16-group column blocks per tile
    constexpr int R = (1 << kTB) / 256;       // amplitudes per thread
    extern __shared__ v2d tile[];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // Persistent: this workgroup walks tiles blockIdx.x, + gridDim.x, ...  The next tile is
    // loaded into registers while the current one is on the matrix cores.
    uint64_t goff[R];  // global offset of element tid + r * 256 within a tile
    const uint64_t off_t = dense_tile_offset(d, tid);
#pragma unroll
    for (int r = 0; r < R; ++r) goff[r] = off_t | dense_tile_offset(d, uint32_t(r) * 256);
    auto tile_base = [&](uint32_t t) {
        uint64_t base = t;
        for (int j = 0; j < kTB; ++j) base = insert0(base, d.pos[j]);
        return base;
    };

    double ar[RT][KS], ai[RT][KS], an[RT][KS];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const v2d u = U[(rt * 16 + (lane & 15)) * D + ks * 4 + (lane >> 4)];
            ar[rt][ks] = u.x;
            ai[rt][ks] = u.y;
            an[rt][ks] = -u.y;
        }
    const uint32_t d_in = tdep<K>(dt, lane >> 4);   // column bits 0..1 of a K step
    const uint32_t d_out = tdep<K>(dt, lane >> 4);  // row bits 0..1 of an output block
    const uint32_t g_lane = gdep<K>(dt, lane & 15);

    v2d pf[R];
    uint32_t t = blockIdx.x;
    if (t < ntiles) {
        const uint64_t base = tile_base(t);
#pragma unroll
        for (int r = 0; r < R; ++r) pf[r] = ld_nt(psi + (base | goff[r]));
    }
    for (; t < ntiles; t += gridDim.x) {
        const uint64_t base = tile_base(t);
#pragma unroll
        for (int r = 0; r < R; ++r) tile[dswz(tid + r * 256)] = pf[r];
        __syncthreads();
        const uint32_t tn = t + gridDim.x;
        if (tn < ntiles) {
            const uint64_t nb = tile_base(tn);
#pragma unroll
            for (int r = 0; r < R; ++r) pf[r] = ld_nt(psi + (nb | goff[r]));
        }
        for (int cb = wave; cb < NCB; cb += 4) {
            const uint32_t go = g_lane | gdep<K>(dt, uint32_t(cb) << 4);
            d4 yr[RT], yi[RT];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                yr[rt] = d4{0.0, 0.0, 0.0, 0.0};
                yi[rt] = d4{0.0, 0.0, 0.0, 0.0};
            }
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const v2d x = tile[dswz(go | d_in | tdep<K>(dt, uint32_t(ks) << 2))];
                // 2 * RT independent accumulator chains, interleaved so back-to-back MFMAs never
                // wait on each other's result
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) {
                    yr[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[rt][ks], x.x, yr[rt], 0, 0, 0);
                    yi[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[rt][ks], x.x, yi[rt], 0, 0, 0);
                }
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) {
                    yr[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(an[rt][ks], x.y, yr[rt], 0, 0, 0);
                    yi[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[rt][ks], x.y, yi[rt], 0, 0, 0);
                }
            }
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    v2d y;
                    y.x = yr[rt][v];
                    y.y = yi[rt][v];
                    tile[dswz(go | d_out | tdep<K>(dt, (uint32_t(v) << 2) | (uint32_t(rt) << 4)))] = y;
                }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) st_nt(psi + (base | goff[r]), tile[dswz(tid + r * 256)]);
        __syncthreads();
    }
}

bool make_dense_tile(int n, int k, const int* targets, PassDesc* d, DenseTile* dt) {
    if (k < 4 || k > 5 || n < dense_tile_bits(k)) return false;
    const int b = dense_tile_bits(k), L = 3;
    uint64_t T = (uint64_t(1) << L) - 1;
    for (int j = 0; j < k; ++j) T |= uint64_t(1) << targets[j];
    for (int q = 0; q < n && __builtin_popcountll(T) < b; ++q) T |= uint64_t(1) << q;
    if (__builtin_popcountll(T) != b) return false;
    d->n = n;
    d->b = b;
    d->L = L;
    d->nops = 0;
    int local_of[64];
    int j = 0;
    for (int q = 0; q < L; ++q) d->tq[j++] = q;
    for (int q = L; q < n; ++q)
        if ((T >> q) & 1) d->tq[j++] = q;
    for (int t = 0; t < b; ++t) {
        local_of[d->tq[t]] = t;
        d->pos[t] = d->tq[t];
    }
    // pos ascending: tq is ascending already (0..L-1, then the rest in order)
    dt->k = k;
    uint32_t tl = 0;
    for (int t = 0; t < k; ++t) {
        dt->tloc[t] = local_of[targets[t]];
        tl |= 1u << dt->tloc[t];
    }
    int g = 0;
    for (int t = 0; t < b; ++t)
        if (!((tl >> t) & 1)) dt->gbits[g++] = t;
    return true;
}

hipError_t launch_dense_mfma(hipStream_t st, v2d* psi, const PassDesc& d, const DenseTile& dt, const v2d* U_dev) {
    const int tb = dense_tile_bits(dt.k);
    const uint64_t tiles = uint64_t(1) << (d.n - tb);
    const size_t lds = size_t(16) << tb;
    static int cus = 0;
    if (!cus) {
        int dev = 0, c = 256;
        if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
        cus = c;
    }
    // persistent: as many workgroups as fit the CUs' LDS (64 KiB tiles: two per CU; 32 KiB: four)
    const uint64_t persistent = uint64_t(cus) * (tb == 11 ? 4 : 2);
    const unsigned grid = unsigned(std::min<uint64_t>(tiles, persistent));
    if (dt.k == 5)
        hipLaunchKernelGGL(k_dense_mfma<5>, dim3(grid), dim3(256), lds, st, psi, d, dt, U_dev, uint32_t(tiles));
    else
        hipLaunchKernelGGL(k_dense_mfma<4>, dim3(grid), dim3(256), lds, st, psi, d, dt, U_dev, uint32_t(tiles));
    return hipGetLastError();
}

}  // namespace llsv
