// sv_gates.hip -- streaming gfx950 kernels: one gate = one pass over the state in HBM.
//
// Every kernel here reads each amplitude it changes once and writes it once, with
// 128-bit accesses; consecutive lanes of a 64-wide wavefront touch consecutive
// amplitudes wherever the gate's index map allows, so each wavefront access is a
// 1 KiB contiguous burst (runs of 2^q amplitudes when the gate's lowest qubit q < 6).
// Grids are sized for >> 256 CUs.
//
// Split-chunk order (the *_split kernels, used from 2^13 amplitudes up): a workgroup takes
// two runs of 256 units (pairs, swaps or amplitudes) whose addresses lie 32-64 KiB apart,
// instead of one contiguous run.  Measured on MI355X at 30 qubits, in-place streams in that
// order run 8-12% faster (6.2-6.6 TB/s against 5.6-6.0 for the contiguous order), which
// points at DRAM bank parallelism within each workgroup's accesses.
#include <algorithm>
#include <vector>

#include "sv_common.h"
#include "sv_kernels.h"

namespace llsv {

// ---------------------------------------------------------------- basis fill
__global__ __launch_bounds__(256) void k_fill_basis(v2d* __restrict__ psi, uint64_t size, uint64_t index) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < size; i += uint64_t(gridDim.x) * 256) {
        v2d v;
        v.x = (i == index) ? 1.0 : 0.0;
        v.y = 0.0;
        st_nt(psi + i, v);
    }
}

hipError_t launch_fill_basis(hipStream_t st, v2d* psi, int n, uint64_t index) {
    const uint64_t size = uint64_t(1) << n;
    hipLaunchKernelGGL(k_fill_basis, dim3(grid_for((size + 255) / 256)), dim3(256), 0, st, psi, size, index);
    return hipGetLastError();
}

// ---------------------------------------------------------------- copy
// Plain device copy, 4 amplitudes per thread in flight.  Also the calibration kernel for
// HBM traffic counters: it moves exactly 32 bytes per amplitude.
__global__ __launch_bounds__(256) void k_copy(v2d* __restrict__ dst, const v2d* __restrict__ src, uint64_t size, uint64_t nchunks) {
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t i0 = chunk * 1024 + threadIdx.x;
    v2d v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (i0 + u * 256 < size) v[u] = ld_nt(src + i0 + u * 256);
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (i0 + u * 256 < size) st_nt(dst + i0 + u * 256, v[u]);
  }
}

hipError_t launch_copy(hipStream_t st, v2d* dst, const v2d* src, int n) {
    const uint64_t size = uint64_t(1) << n;
    const uint64_t nch = (size + 1023) / 1024;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(nch)), dim3(256), 0, st, dst, src, size, nch);
    return hipGetLastError();
}

// ---------------------------------------------------------------- dense 2x2
// Pair p (bit q removed) -> i0 = insert0(p, q), i1 = i0 + 2^q.  Each thread owns PPT pairs
// spaced one block apart, so a wavefront still covers 64 consecutive pairs per access
// while every thread keeps 2 * PPT independent 16-byte loads in flight.
//
// XCD_MAP: workgroups are dealt round-robin to the 8 XCDs; with the remap, XCD k walks the
// k-th contiguous eighth of the pairs instead of every 8th block.  Measured on MI355X at 30
// qubits, that order streams ~4% faster for 6 <= q < 20 and ~4% slower for q >= 20.
__device__ __forceinline__ uint64_t xcd_block(uint64_t b, uint64_t nb) { return (b % 8) * (nb / 8) + b / 8; }

template <int PPT, bool XCD_MAP>
__global__ __launch_bounds__(256) void k_apply_1q(v2d* __restrict__ psi, uint64_t npairs, int q, Gate2 g, uint64_t nchunks) {
    const uint64_t lomask = (uint64_t(1) << q) - 1;
    const uint64_t stride = uint64_t(1) << q;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t blk = XCD_MAP ? xcd_block(chunk, nchunks) : chunk;
    const uint64_t p0 = blk * (256 * PPT) + threadIdx.x;
    v2d a0[PPT], a1[PPT];
    uint64_t i0[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t p = p0 + uint64_t(u) * 256;
        i0[u] = ((p & ~lomask) << 1) | (p & lomask);
        if (p < npairs) {
            a0[u] = ld_nt(psi + i0[u]);
            a1[u] = ld_nt(psi + i0[u] + stride);
        }
    }
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t p = p0 + uint64_t(u) * 256;
        if (p < npairs) {
            st_nt(psi + i0[u], row2(g.r, a0[u], a1[u]));
            st_nt(psi + i0[u] + stride, row2(g.r + 4, a0[u], a1[u]));
        }
    }
  }
}

// Target q < 6: the partner of amplitude i (i ^ 2^q) sits in the same wavefront when each
// lane loads one of 64 consecutive amplitudes, so it is fetched with a cross-lane shuffle and
// every load and store instruction covers 1 KiB of contiguous memory.  (Pairing by index
// instead would make each store write every other 16 bytes of a line, which costs up to 66%
// extra HBM write traffic for q <= 2.)
template <int PPT>
__global__ __launch_bounds__(256) void k_apply_1q_lane(v2d* __restrict__ psi, uint64_t size, int q, Gate2 g, uint64_t nchunks) {
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t i0 = chunk * (256 * PPT) + threadIdx.x;
    const int m = 1 << q;
    const bool hi = (threadIdx.x >> q) & 1;
    const double* r = hi ? g.r + 4 : g.r;
    v2d a[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u)
        if (i0 + u * 256 < size) a[u] = ld_nt(psi + i0 + u * 256);
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        v2d p;
        p.x = __shfl_xor(a[u].x, m, 64);
        p.y = __shfl_xor(a[u].y, m, 64);
        if (i0 + u * 256 < size) st_nt(psi + i0 + u * 256, hi ? row2(r, p, a[u]) : row2(r, a[u], p));
    }
  }
}

// Diagonal gate diag(1, d1): only the bit-q = 1 half changes (Z, S, T, phase gates).
template <int PPT>
__global__ __launch_bounds__(256) void k_phase_upper(v2d* __restrict__ psi, uint64_t npairs, int q, double dr,
                                                     double di, uint64_t nchunks) {
    const uint64_t lomask = (uint64_t(1) << q) - 1;
    const uint64_t stride = uint64_t(1) << q;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t p0 = chunk * (256 * PPT) + threadIdx.x;
    v2d a[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t p = p0 + uint64_t(u) * 256;
        if (p < npairs) a[u] = ld_nt(psi + (((p & ~lomask) << 1) | (p & lomask)) + stride);
    }
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t p = p0 + uint64_t(u) * 256;
        if (p < npairs) st_nt(psi + (((p & ~lomask) << 1) | (p & lomask)) + stride, cmul(dr, di, a[u]));
    }
  }
}

// Diagonal gate diag(d0, d1) with d0 != 1: every amplitude is scaled, one linear stream.
template <int PPT>
__global__ __launch_bounds__(256) void k_diag_1q(v2d* __restrict__ psi, uint64_t size, int q, Gate2 g, uint64_t nchunks) {
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t i0 = chunk * (256 * PPT) + threadIdx.x;
    v2d a[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t i = i0 + uint64_t(u) * 256;
        if (i < size) a[u] = ld_nt(psi + i);
    }
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t i = i0 + uint64_t(u) * 256;
        if (i < size) {
            const bool hi = (i >> q) & 1;
            st_nt(psi + i, hi ? cmul(g.r[6], g.r[7], a[u]) : cmul(g.r[0], g.r[1], a[u]));
        }
    }
  }
}

// Unit of the split-chunk order for thread t of chunk `chunk`, half h (0, 1): chunks come in
// groups of 2^(logD - 8) that share a superchunk of 2^(logD + 1) units; the two halves of a
// chunk are 2^logD units apart.  The unit count must be a multiple of 2^(logD + 1).
__device__ __forceinline__ uint64_t split_unit(uint64_t chunk, uint32_t t, int h, int logD) {
    const uint64_t per = uint64_t(1) << (logD - 8);
    return ((chunk >> (logD - 8)) << (logD + 1)) + ((chunk & (per - 1)) << 8) + t + (h ? (uint64_t(1) << logD) : 0);
}

// Offset exponent (in units) of the split for a unit -> amplitude map that inserts zero bits
// at the ascending positions ins[0..m): the unit bit that lands on amplitude bit 12 (64 KiB),
// else 11, else 13.  -1 if none (then the contiguous kernels run).
static int split_bit(const int* ins, int m) {
    for (int want : {12, 11, 13})
        for (int d = 8; d < 40; ++d) {
            int x = d;
            for (int j = 0; j < m; ++j)
                if (x >= ins[j]) ++x;
            if (x == want) return d;
            if (x > want) break;
        }
    return -1;
}

constexpr int kSplitPairs = 11;  // pair-index split of the one-qubit kernels (2^11 pairs apart)

// R rounds per chunk: a thread has 2R pairs (4R loads) in flight, the R split units of its
// chunk one after another.  Measured on MI355X at 30 qubits, targets 6..29: R = 4 streams 3%
// faster than R = 1 (the lane-shuffle kernel for q < 6 is fastest at R = 1).
constexpr int kRounds1q = 4;

template <bool XCD_MAP, int R>
__global__ __launch_bounds__(256) void k_apply_1q_split(v2d* __restrict__ psi, int q, Gate2 g, uint64_t nchunks, int logD) {
    const uint64_t lomask = (uint64_t(1) << q) - 1;
    const uint64_t stride = uint64_t(1) << q;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t blk = XCD_MAP ? xcd_block(chunk, nchunks) : chunk;
    v2d a0[2 * R], a1[2 * R];
    uint64_t i0[2 * R];
#pragma unroll
    for (int u = 0; u < 2 * R; ++u) {
        const uint64_t p = split_unit(R * blk + (u >> 1), threadIdx.x, u & 1, logD);
        i0[u] = ((p & ~lomask) << 1) | (p & lomask);
        a0[u] = ld_nt(psi + i0[u]);
        a1[u] = ld_nt(psi + i0[u] + stride);
    }
#pragma unroll
    for (int u = 0; u < 2 * R; ++u) {
        st_nt(psi + i0[u], row2(g.r, a0[u], a1[u]));
        st_nt(psi + i0[u] + stride, row2(g.r + 4, a0[u], a1[u]));
    }
  }
}

__global__ __launch_bounds__(256) void k_apply_1q_lane_split(v2d* __restrict__ psi, int q, Gate2 g, uint64_t nchunks,
                                                             int logD) {
    const int m = 1 << q;
    const bool hi = (threadIdx.x >> q) & 1;
    const double* r = hi ? g.r + 4 : g.r;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    v2d a[2];
    uint64_t ii[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        ii[u] = split_unit(chunk, threadIdx.x, u, logD);
        a[u] = ld_nt(psi + ii[u]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        v2d p;
        p.x = __shfl_xor(a[u].x, m, 64);
        p.y = __shfl_xor(a[u].y, m, 64);
        st_nt(psi + ii[u], hi ? row2(r, p, a[u]) : row2(r, a[u], p));
    }
  }
}

__global__ __launch_bounds__(256) void k_phase_upper_split(v2d* __restrict__ psi, int q, double dr, double di,
                                                           uint64_t nchunks, int logD) {
    const uint64_t lomask = (uint64_t(1) << q) - 1;
    const uint64_t stride = uint64_t(1) << q;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    v2d a[2];
    uint64_t i[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const uint64_t p = split_unit(chunk, threadIdx.x, u, logD);
        i[u] = (((p & ~lomask) << 1) | (p & lomask)) + stride;
        a[u] = ld_nt(psi + i[u]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) st_nt(psi + i[u], cmul(dr, di, a[u]));
  }
}

__global__ __launch_bounds__(256) void k_diag_1q_split(v2d* __restrict__ psi, int q, Gate2 g, uint64_t nchunks, int logD) {
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    v2d a[2];
    uint64_t ii[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        ii[u] = split_unit(chunk, threadIdx.x, u, logD);
        a[u] = ld_nt(psi + ii[u]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const bool hi = (ii[u] >> q) & 1;
        st_nt(psi + ii[u], hi ? cmul(g.r[6], g.r[7], a[u]) : cmul(g.r[0], g.r[1], a[u]));
    }
  }
}

constexpr int kPPT = 4;
// Lowest qubit whose 2^q-amplitude runs (16 * 2^q bytes) fill whole 128-byte lines.
// Tile order (states of 18 qubits and more): workgroup w takes a tile of 128 runs of 512 bytes
// (index bits 0..4) 32 KiB apart (bits 11..17), or of bits 0..4, the target and 11..16 for targets
// 5..9, or of bits 0..4, 11..17 and the target for target 10 and targets >= 18 (2^13 amplitudes, 32
// per thread); the tile's other index bits are w's, 12-bit tiles dealt to XCDs in contiguous
// eighths.  A target among the register bits is the top one: each thread applies the gate to its
// own pairs (registers j and j + R/2); targets 0..4 and 11 are thread bits 0..5, whose pairs meet by
// a cross-lane exchange.  Measured on MI355X at 30 qubits with a plain in-place pass over tiles of
// bits 0..4 and 11..17: 6.7 TB/s, against 5.9 for contiguous 64 KiB tiles.
struct TileMap {
    int nb;       // tile bits: 12 or 13
    int pos[13];  // the tile's index bits, ascending
    int lane[8];  // index bit of thread bit k
    int reg[5];   // index bit of register bit k (the last = the gate's target)
};

constexpr int kTileMinQubits = 18;

// Thread bit of the target when it is one (then the gate's pairs meet by a cross-lane exchange
// within the wavefront), else -1.
int tile_lane_bit(int q) { return q < 5 ? q : q == 11 ? 5 : -1; }

TileMap tile_map_1q(int q) {
    std::vector<int> lanes = {0, 1, 2, 3, 4, 11, 12, 13}, regs;
    if (tile_lane_bit(q) >= 0) {
        regs = {14, 15, 16, 17};  // bits 0..4 and 11..17
    } else if (q == 12 || q == 13) {  // 17 takes the target's place among the lanes
        for (int& l : lanes)
            if (l == q) l = 17;
        regs = {14, 15, 16, q};
    } else if (q >= 14 && q <= 17) {
        for (int b = 14; b <= 17; ++b)
            if (b != q) regs.push_back(b);
        regs.push_back(q);
    } else if (q < 10) {  // 0..4, q, 11..16
        regs = {14, 15, 16, q};
    } else {  // 10 or >= 18: 0..4, 11..17, q (13 bits)
        regs = {14, 15, 16, 17, q};
    }
    // a register target is the top register bit: a thread loads its pairs' lower halves first
    TileMap m{};
    std::vector<int> all(lanes);
    all.insert(all.end(), regs.begin(), regs.end());
    std::sort(all.begin(), all.end());
    m.nb = int(all.size());
    for (int k = 0; k < m.nb; ++k) m.pos[k] = all[k];
    for (int k = 0; k < 8; ++k) m.lane[k] = lanes[k];
    for (size_t k = 0; k < regs.size(); ++k) m.reg[k] = regs[k];
    return m;
}

// LANE: the target is thread bit qb (< 6, within the wavefront) instead of a register bit; each lane
// takes its partner's amplitude with a cross-lane exchange.
template <int RB, bool XCD, bool LANE>
__global__ __launch_bounds__(256, 2) void k_apply_1q_tile(v2d* __restrict__ psi, Gate2 g, TileMap tm, int qb) {
    constexpr int R = 1 << RB;
    uint64_t base = XCD ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
#pragma unroll
    for (int k = 0; k < 8 + RB; ++k) base = insert0(base, tm.pos[k]);
    uint64_t gt = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if ((threadIdx.x >> k) & 1u) gt |= uint64_t(1) << tm.lane[k];
    uint64_t rb[RB];
#pragma unroll
    for (int k = 0; k < RB; ++k) rb[k] = uint64_t(1) << tm.reg[k];
    v2d* const p = psi + (base | gt);
    auto off = [&](int j) {
        uint64_t o = 0;
#pragma unroll
        for (int k = 0; k < RB; ++k)
            if ((j >> k) & 1) o |= rb[k];
        return o;
    };
    v2d a[R];
#pragma unroll
    for (int j = 0; j < R; ++j) a[j] = ld_nt(p + off(j));
    if (LANE) {
        const int m = 1 << qb;
        const bool hi = (threadIdx.x >> qb) & 1;
        const double* r = hi ? g.r + 4 : g.r;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            v2d o;
            o.x = __shfl_xor(a[j].x, m, 64);
            o.y = __shfl_xor(a[j].y, m, 64);
            a[j] = hi ? row2(r, o, a[j]) : row2(r, a[j], o);
        }
    } else {
#pragma unroll
        for (int j = 0; j < R / 2; ++j) {
            const v2d x0 = a[j], x1 = a[j + R / 2];
            a[j] = row2(g.r, x0, x1);
            a[j + R / 2] = row2(g.r + 4, x0, x1);
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) st_nt(p + off(j), a[j]);
}

constexpr int kLineBits = 3;

hipError_t launch_apply_1q(hipStream_t st, v2d* psi, int n, int q, const Gate2& g) {
    const uint64_t npairs = uint64_t(1) << (n - 1);
    const bool diag = g.r[2] == 0.0 && g.r[3] == 0.0 && g.r[4] == 0.0 && g.r[5] == 0.0;
    if (diag && g.r[0] == 1.0 && g.r[1] == 0.0 && g.r[6] == 1.0 && g.r[7] == 0.0) return hipSuccess;  // identity
    if (npairs >= (uint64_t(1) << (kSplitPairs + 1))) {
        // split-chunk order: pairs (or amplitudes) in chunks of 512, halves 2^11 pairs apart
        const int ld = kSplitPairs;
        if (diag && g.r[0] == 1.0 && g.r[1] == 0.0 && q >= kLineBits) {
            const uint64_t nch = npairs / 512;
            hipLaunchKernelGGL(k_phase_upper_split, dim3(grid_for(nch)), dim3(256), 0, st, psi, q, g.r[6], g.r[7], nch, ld);
        } else if (diag) {
            const uint64_t nch = 2 * npairs / 512;
            hipLaunchKernelGGL(k_diag_1q_split, dim3(grid_for(nch)), dim3(256), 0, st, psi, q, g, nch, ld);
        } else if (n >= kTileMinQubits) {
            const TileMap tm = tile_map_1q(q);
            const dim3 grid(unsigned(uint64_t(1) << (n - tm.nb)));
            const int qb = tile_lane_bit(q);
            if (qb >= 0)
                hipLaunchKernelGGL((k_apply_1q_tile<4, true, true>), grid, dim3(256), 0, st, psi, g, tm, qb);
            else if (tm.nb == 13)
                hipLaunchKernelGGL((k_apply_1q_tile<5, false, false>), grid, dim3(256), 0, st, psi, g, tm, 0);
            else
                hipLaunchKernelGGL((k_apply_1q_tile<4, true, false>), grid, dim3(256), 0, st, psi, g, tm, 0);
        } else if (q < 6) {
            const uint64_t nch = 2 * npairs / 512;
            hipLaunchKernelGGL(k_apply_1q_lane_split, dim3(grid_for(nch)), dim3(256), 0, st, psi, q, g, nch, ld);
        } else if (q < 9 && (npairs / 512 / kRounds1q) % 8 == 0) {
            // XCD-contiguous chunk order (xcd_block needs a multiple of 8 chunks): another 2% at
            // these targets
            const uint64_t nch = npairs / 512 / kRounds1q;
            hipLaunchKernelGGL((k_apply_1q_split<true, kRounds1q>), dim3(grid_for(nch)), dim3(256), 0, st, psi, q, g, nch, ld);
        } else {
            const uint64_t nch = npairs / 512 / kRounds1q;
            hipLaunchKernelGGL((k_apply_1q_split<false, kRounds1q>), dim3(grid_for(nch)), dim3(256), 0, st, psi, q, g, nch,
                               ld);
        }
        return hipGetLastError();
    }
    if (diag && g.r[0] == 1.0 && g.r[1] == 0.0 && q >= kLineBits) {
        const uint64_t blocks = (npairs + 256 * kPPT - 1) / (256 * kPPT);
        hipLaunchKernelGGL(k_phase_upper<kPPT>, dim3(grid_for(blocks)), dim3(256), 0, st, psi, npairs, q, g.r[6],
                           g.r[7], blocks);
    } else if (diag) {
        const uint64_t size = uint64_t(1) << n;
        const uint64_t blocks = (size + 256 * kPPT - 1) / (256 * kPPT);
        hipLaunchKernelGGL(k_diag_1q<kPPT>, dim3(grid_for(blocks)), dim3(256), 0, st, psi, size, q, g, blocks);
    } else if (q < 6 && n >= 6) {
        const uint64_t size = uint64_t(1) << n;
        const uint64_t blocks = (size + 256 * kPPT - 1) / (256 * kPPT);
        hipLaunchKernelGGL(k_apply_1q_lane<kPPT>, dim3(grid_for(blocks)), dim3(256), 0, st, psi, size, q, g, blocks);
    } else if (q >= 20 || npairs % (256 * 2 * 8) != 0) {
        const uint64_t blocks = (npairs + 256 * 4 - 1) / (256 * 4);
        hipLaunchKernelGGL((k_apply_1q<4, false>), dim3(grid_for(blocks)), dim3(256), 0, st, psi, npairs, q, g, blocks);
    } else {
        const uint64_t blocks = npairs / (256 * 2);  // a multiple of 8, as the remap needs
        hipLaunchKernelGGL((k_apply_1q<2, true>), dim3(grid_for(blocks)), dim3(256), 0, st, psi, npairs, q, g, blocks);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- CNOT
// Control c >= 3: only the control = 1 half moves, 2^(n-2) swaps: index p with the two gate
// bits removed; i = insert(p) | control bit, partner i | target bit.  Every touched run is
// >= 8 amplitudes (a whole 128-byte line).  Non-temporal accesses when the target is >= 3
// too; with a low target the two halves of each swap share lines, and cached stores let L2
// merge them.
template <int PPT, bool NT>
__global__ __launch_bounds__(256) void k_cnot(v2d* __restrict__ psi, uint64_t nswaps, int lo, int hi, uint64_t cb,
                                              uint64_t tb, uint64_t nchunks) {
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t p0 = chunk * (256 * PPT) + threadIdx.x;
    v2d a[PPT], b[PPT];
    uint64_t idx[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t p = p0 + uint64_t(u) * 256;
        idx[u] = insert0(insert0(p, lo), hi) | cb;
        if (p < nswaps) {
            a[u] = NT ? ld_nt(psi + idx[u]) : psi[idx[u]];
            b[u] = NT ? ld_nt(psi + (idx[u] | tb)) : psi[idx[u] | tb];
        }
    }
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t p = p0 + uint64_t(u) * 256;
        if (p < nswaps) {
            if (NT) {
                st_nt(psi + idx[u], b[u]);
                st_nt(psi + (idx[u] | tb), a[u]);
            } else {
                psi[idx[u]] = b[u];
                psi[idx[u] | tb] = a[u];
            }
        }
    }
  }
}

// Control c < 3: the control = 1 amplitudes are interleaved with the others at 16-byte
// granularity, so every line is touched anyway.  One full pass over whole lines instead:
// target t < 6 pairs through a cross-lane shuffle, t >= 6 through the index pairing of the
// one-qubit kernel; amplitudes whose control bit is 0 are written back unchanged.
template <int PPT>
__global__ __launch_bounds__(256) void k_cnot_lane(v2d* __restrict__ psi, uint64_t size, int c, int t, uint64_t nchunks) {
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t i0 = chunk * (256 * PPT) + threadIdx.x;
    const int m = 1 << t;
    v2d a[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u)
        if (i0 + u * 256 < size) a[u] = ld_nt(psi + i0 + u * 256);
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        v2d p;
        p.x = __shfl_xor(a[u].x, m, 64);
        p.y = __shfl_xor(a[u].y, m, 64);
        const uint64_t i = i0 + u * 256;
        if (i < size) st_nt(psi + i, ((i >> c) & 1) ? p : a[u]);
    }
  }
}

template <int PPT>
__global__ __launch_bounds__(256) void k_cnot_pairs(v2d* __restrict__ psi, uint64_t npairs, int c, int t, uint64_t nchunks) {
    const uint64_t lomask = (uint64_t(1) << t) - 1;
    const uint64_t stride = uint64_t(1) << t;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t p0 = chunk * (256 * PPT) + threadIdx.x;
    v2d a0[PPT], a1[PPT];
    uint64_t i0[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t p = p0 + uint64_t(u) * 256;
        i0[u] = ((p & ~lomask) << 1) | (p & lomask);
        if (p < npairs) {
            a0[u] = ld_nt(psi + i0[u]);
            a1[u] = ld_nt(psi + i0[u] + stride);
        }
    }
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t p = p0 + uint64_t(u) * 256;
        if (p < npairs) {
            const bool sw = (i0[u] >> c) & 1;
            st_nt(psi + i0[u], sw ? a1[u] : a0[u]);
            st_nt(psi + i0[u] + stride, sw ? a0[u] : a1[u]);
        }
    }
  }
}

// Split-chunk versions of the half-state swap and of the lane kernel.
// Two rounds per chunk (four swaps, eight loads in flight per thread): measured on MI355X at 30
// qubits over controls 3..28, 3% faster than one, and no control bit above 2.9 ms (3.4 before).
constexpr int kRoundsCnot = 2;

__global__ __launch_bounds__(256) void k_cnot_split(v2d* __restrict__ psi, int lo, int hi, uint64_t cb, uint64_t tb,
                                                    uint64_t nchunks, int logD) {
  constexpr int R = kRoundsCnot;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    v2d a[2 * R], b[2 * R];
    uint64_t idx[2 * R];
#pragma unroll
    for (int u = 0; u < 2 * R; ++u) {
        idx[u] = insert0(insert0(split_unit(R * chunk + (u >> 1), threadIdx.x, u & 1, logD), lo), hi) | cb;
        a[u] = ld_nt(psi + idx[u]);
        b[u] = ld_nt(psi + (idx[u] | tb));
    }
#pragma unroll
    for (int u = 0; u < 2 * R; ++u) {
        st_nt(psi + idx[u], b[u]);
        st_nt(psi + (idx[u] | tb), a[u]);
    }
  }
}

__global__ __launch_bounds__(256) void k_cnot_lane_split(v2d* __restrict__ psi, int c, int t, uint64_t nchunks, int logD) {
    const int m = 1 << t;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    v2d a[2];
    uint64_t ii[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        ii[u] = split_unit(chunk, threadIdx.x, u, logD);
        a[u] = ld_nt(psi + ii[u]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        v2d p;
        p.x = __shfl_xor(a[u].x, m, 64);
        p.y = __shfl_xor(a[u].y, m, 64);
        st_nt(psi + ii[u], ((ii[u] >> c) & 1) ? p : a[u]);
    }
  }
}

// Tile order for the half-state swap: the tile of k_apply_1q_tile with the target and the control
// as the two top register bits; a thread loads the control = 1 half of its registers (16
// amplitudes) and stores each at its partner across the target bit.  Used where control and target
// are >= 5 and neither is bit 10 or 11 (then thread bit 5 is index bit 11, as in the one-qubit
// tiles): measured on MI355X at 30 qubits over the CNOT ladder, 2.68-3.10 ms per swap there against
// 2.69-3.34 for k_cnot_split, while controls 3-4 and 10-11 took 3.1-3.7 ms this way.
TileMap tile_map_cnot(int c, int t) {
    constexpr int RB = 5;
    std::vector<int> lanes, regs;
    for (int b = 0; b < 5; ++b)
        if (b != c && b != t) lanes.push_back(b);
    int b = 11;
    auto next = [&] {
        while (b == c || b == t) ++b;
        return b++;
    };
    while (lanes.size() < 8) lanes.push_back(next());
    while (regs.size() < RB - 2) regs.push_back(next());
    regs.push_back(t);
    regs.push_back(c);
    TileMap m{};
    std::vector<int> all(lanes);
    all.insert(all.end(), regs.begin(), regs.end());
    std::sort(all.begin(), all.end());
    m.nb = int(all.size());
    for (int k = 0; k < m.nb; ++k) m.pos[k] = all[k];
    for (int k = 0; k < 8; ++k) m.lane[k] = lanes[k];
    for (size_t k = 0; k < regs.size(); ++k) m.reg[k] = regs[k];
    return m;
}

__global__ __launch_bounds__(256, 2) void k_cnot_tile(v2d* __restrict__ psi, TileMap tm) {
    constexpr int RB = 5, H = 1 << (RB - 1), T = 1 << (RB - 2);
    uint64_t base = blockIdx.x;
#pragma unroll
    for (int k = 0; k < 8 + RB; ++k) base = insert0(base, tm.pos[k]);
    uint64_t gt = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if ((threadIdx.x >> k) & 1u) gt |= uint64_t(1) << tm.lane[k];
    uint64_t rb[RB - 1];
#pragma unroll
    for (int k = 0; k < RB - 1; ++k) rb[k] = uint64_t(1) << tm.reg[k];
    v2d* const p = psi + (base | gt | (uint64_t(1) << tm.reg[RB - 1]));  // control = 1
    auto off = [&](int j) {
        uint64_t o = 0;
#pragma unroll
        for (int k = 0; k < RB - 1; ++k)
            if ((j >> k) & 1) o |= rb[k];
        return o;
    };
    v2d a[H];
#pragma unroll
    for (int j = 0; j < H; ++j) a[j] = ld_nt(p + off(j));
#pragma unroll
    for (int j = 0; j < H; ++j) st_nt(p + off(j ^ T), a[j]);
}

hipError_t launch_cnot(hipStream_t st, v2d* psi, int n, int c, int t) {
    const int lo = c < t ? c : t, hi = c < t ? t : c;
    if (c >= 5 && t >= 5 && c != 10 && c != 11 && t != 10 && t != 11 && n >= kTileMinQubits) {
        const TileMap tm = tile_map_cnot(c, t);
        if (tm.pos[tm.nb - 1] < n) {
            hipLaunchKernelGGL(k_cnot_tile, dim3(unsigned(uint64_t(1) << (n - tm.nb))), dim3(256), 0, st, psi, tm);
            return hipGetLastError();
        }
    }
    if (c >= kLineBits && t >= kLineBits) {
        const int ins[2] = {lo, hi};
        const int d = split_bit(ins, 2);
        const uint64_t nswaps = uint64_t(1) << (n - 2);
        if (d > 0 && nswaps >= (uint64_t(1) << (d + 1)) * kRoundsCnot) {
            const uint64_t nch = nswaps / 512 / kRoundsCnot;
            hipLaunchKernelGGL(k_cnot_split, dim3(grid_for(nch)), dim3(256), 0, st, psi, lo, hi, uint64_t(1) << c,
                               uint64_t(1) << t, nch, d);
            return hipGetLastError();
        }
    } else if (c < kLineBits && t < 6 && n >= kSplitPairs + 2) {
        const uint64_t nch = (uint64_t(1) << n) / 512;
        hipLaunchKernelGGL(k_cnot_lane_split, dim3(grid_for(nch)), dim3(256), 0, st, psi, c, t, nch, kSplitPairs);
        return hipGetLastError();
    }
    if (c >= kLineBits || n < 6) {
        const uint64_t nswaps = uint64_t(1) << (n - 2);
        const uint64_t blocks = (nswaps + 256 * kPPT - 1) / (256 * kPPT);
        if (t >= kLineBits)
            hipLaunchKernelGGL((k_cnot<kPPT, true>), dim3(grid_for(blocks)), dim3(256), 0, st, psi, nswaps, lo, hi,
                               uint64_t(1) << c, uint64_t(1) << t, blocks);
        else
            hipLaunchKernelGGL((k_cnot<kPPT, false>), dim3(grid_for(blocks)), dim3(256), 0, st, psi, nswaps, lo, hi,
                               uint64_t(1) << c, uint64_t(1) << t, blocks);
    } else if (t < 6) {
        const uint64_t size = uint64_t(1) << n;
        const uint64_t blocks = (size + 1023) / 1024;
        hipLaunchKernelGGL(k_cnot_lane<kPPT>, dim3(grid_for(blocks)), dim3(256), 0, st, psi, size, c, t, blocks);
    } else {
        const uint64_t npairs = uint64_t(1) << (n - 1);
        const uint64_t blocks = (npairs + 1023) / 1024;
        hipLaunchKernelGGL(k_cnot_pairs<kPPT>, dim3(grid_for(blocks)), dim3(256), 0, st, psi, npairs, c, t, blocks);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- MCZ / CZ
// Only the 2^(n-m) amplitudes with all m listed bits set are touched: index p with the
// m bits removed, bits re-inserted in ascending position order, then OR the mask.
template <int PPT>
__global__ __launch_bounds__(256) void k_mcz(v2d* __restrict__ psi, uint64_t count, BitList bits, uint64_t nchunks) {
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t p0 = chunk * (256 * PPT) + threadIdx.x;
    v2d a[PPT];
    uint64_t idx[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t p = p0 + uint64_t(u) * 256;
        uint64_t i = p;
        for (int j = 0; j < bits.m; ++j) i = insert0(i, bits.pos[j]);
        idx[u] = i | bits.mask;
        if (p < count) a[u] = ld_nt(psi + idx[u]);
    }
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t p = p0 + uint64_t(u) * 256;
        if (p < count) st_nt(psi + idx[u], -a[u]);
    }
  }
}

// Lowest masked qubit < 3: the touched amplitudes are spread over every line, so stream
// the whole state once (negating the matching ones) rather than write partial lines.
template <int PPT>
__global__ __launch_bounds__(256) void k_mcz_full(v2d* __restrict__ psi, uint64_t size, uint64_t mask, uint64_t nchunks) {
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t i0 = chunk * (256 * PPT) + threadIdx.x;
    v2d a[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u)
        if (i0 + u * 256 < size) a[u] = ld_nt(psi + i0 + u * 256);
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const uint64_t i = i0 + u * 256;
        if (i < size) st_nt(psi + i, ((i & mask) == mask) ? -a[u] : a[u]);
    }
  }
}

__global__ __launch_bounds__(256) void k_mcz_split(v2d* __restrict__ psi, BitList bits, uint64_t nchunks, int logD) {
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    v2d a[2];
    uint64_t idx[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        uint64_t i = split_unit(chunk, threadIdx.x, u, logD);
        for (int j = 0; j < bits.m; ++j) i = insert0(i, bits.pos[j]);
        idx[u] = i | bits.mask;
        a[u] = ld_nt(psi + idx[u]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) st_nt(psi + idx[u], -a[u]);
  }
}

__global__ __launch_bounds__(256) void k_mcz_full_split(v2d* __restrict__ psi, uint64_t mask, uint64_t nchunks, int logD) {
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    v2d a[2];
    uint64_t ii[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        ii[u] = split_unit(chunk, threadIdx.x, u, logD);
        a[u] = ld_nt(psi + ii[u]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) st_nt(psi + ii[u], ((ii[u] & mask) == mask) ? -a[u] : a[u]);
  }
}

hipError_t launch_mcz(hipStream_t st, v2d* psi, int n, const BitList& bits) {
    if (bits.pos[0] < kLineBits && n >= kSplitPairs + 2) {
        const uint64_t nch = (uint64_t(1) << n) / 512;
        hipLaunchKernelGGL(k_mcz_full_split, dim3(grid_for(nch)), dim3(256), 0, st, psi, bits.mask, nch, kSplitPairs);
        return hipGetLastError();
    }
    if (bits.pos[0] >= kLineBits) {
        const int d = split_bit(bits.pos, bits.m);
        const uint64_t count = uint64_t(1) << (n - bits.m);
        if (d > 0 && count >= (uint64_t(1) << (d + 1))) {
            const uint64_t nch = count / 512;
            hipLaunchKernelGGL(k_mcz_split, dim3(grid_for(nch)), dim3(256), 0, st, psi, bits, nch, d);
            return hipGetLastError();
        }
    }
    if (bits.pos[0] < kLineBits && n >= 6) {
        const uint64_t size = uint64_t(1) << n;
        const uint64_t blocks = (size + 1023) / 1024;
        hipLaunchKernelGGL(k_mcz_full<kPPT>, dim3(grid_for(blocks)), dim3(256), 0, st, psi, size, bits.mask, blocks);
        return hipGetLastError();
    }
    const uint64_t count = uint64_t(1) << (n - bits.m);
    const uint64_t blocks = (count + 256 * kPPT - 1) / (256 * kPPT);
    hipLaunchKernelGGL(k_mcz<kPPT>, dim3(grid_for(blocks)), dim3(256), 0, st, psi, count, bits, blocks);
    return hipGetLastError();
}

// ---------------------------------------------------------------- dense k-qubit (VALU)
// Block of 256 threads = 256 / D groups of D = 2^k amplitudes.  Thread (g, r) gathers
// amplitude r of its group into LDS, then computes row r: y = sum_c U[r][c] x[c] with U
// staged once per block in LDS.  Each block walks `iters` consecutive group batches.
template <int K>
__global__ __launch_bounds__(256) void k_dense(v2d* __restrict__ psi, uint64_t ngroups, BitList tb,
                                               const v2d* __restrict__ U, const uint64_t* __restrict__ off,
                                               int iters) {
    constexpr int D = 1 << K;
    constexpr int GPB = 256 / D;
    __shared__ v2d sU[D * D];
    __shared__ v2d sx[256];
    for (int i = threadIdx.x; i < D * D; i += 256) sU[i] = U[i];
    const int r = threadIdx.x % D;
    const int gl = threadIdx.x / D;
    const uint64_t my_off = off[r];
    __syncthreads();
    for (int it = 0; it < iters; ++it) {
        const uint64_t g = (uint64_t(blockIdx.x) * iters + it) * GPB + gl;
        uint64_t base = g;
        for (int j = 0; j < tb.m; ++j) base = insert0(base, tb.pos[j]);
        const bool live = g < ngroups;
        if (live) sx[threadIdx.x] = ld_nt(psi + (base | my_off));
        __syncthreads();
        if (live) {
            v2d acc = {0.0, 0.0};
            const v2d* xr = sx + gl * D;
            const v2d* ur = sU + r * D;
#pragma unroll 8
            for (int c = 0; c < D; ++c) {
                const v2d u = ur[c], x = xr[c];
                acc.x += u.x * x.x - u.y * x.y;
                acc.y += u.x * x.y + u.y * x.x;
            }
            st_nt(psi + (base | my_off), acc);
        }
        __syncthreads();
    }
}

hipError_t launch_dense(hipStream_t st, v2d* psi, int n, int k, const BitList& sorted_targets, const v2d* U_dev,
                        const uint64_t* off_dev) {
    const uint64_t ngroups = uint64_t(1) << (n - k);
    const int gpb = 256 >> k;
    const uint64_t batches = (ngroups + gpb - 1) / gpb;
    const int iters = batches >= 4096 ? 8 : 1;
    const uint64_t blocks = (batches + iters - 1) / iters;
#define LLSV_DENSE_CASE(KK)                                                                                \
    case KK:                                                                                               \
        hipLaunchKernelGGL(k_dense<KK>, dim3(unsigned(blocks)), dim3(256), 0, st, psi, ngroups, sorted_targets, \
                           U_dev, off_dev, iters);                                                         \
        break;
    switch (k) {
        LLSV_DENSE_CASE(1)
        LLSV_DENSE_CASE(2)
        LLSV_DENSE_CASE(3)
        LLSV_DENSE_CASE(4)
        LLSV_DENSE_CASE(5)
        default:
            return hipErrorInvalidValue;
    }
#undef LLSV_DENSE_CASE
    return hipGetLastError();
}

}  // namespace llsv
