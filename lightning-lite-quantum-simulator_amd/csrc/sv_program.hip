// sv_program.hip -- gate programs applied in LDS-staged HBM passes.
//
// A pass reads every amplitude of the state once and writes it once.  Workgroup w owns one
// tile of 2^b amplitudes: the amplitudes whose index bits outside a chosen qubit set T
// (|T| = b) equal w.  T always contains the L lowest qubits, so a tile is 2^(b-L) runs of
// 2^L contiguous amplitudes and every wavefront access is a full 128-byte line or more.
// The tile sits in LDS (64 KiB at b = 12) while every op of the pass is applied to it,
// with one workgroup barrier between ops; then it is written back.  Ops that act on
// qubits of T are applied on the tile; diagonal ops (CZ, MCZ, phase gates) may touch
// any qubit, because the tile's fixed index bits are known to the workgroup.
//
// The host scheduler walks the circuit in order and grows a pass greedily: an op joins
// when its qubits fit T (|T| <= b) and it shares no qubit with an earlier op that was
// left for a later pass (so joining never reorders two ops that do not commute).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "sv_dense.h"
#include "sv_kernels.h"
#include "sv_jit.h"
#include "sv_program.h"
#include "sv_regops.h"
#include "sv_state.h"

namespace llsv {

ProgramCache::~ProgramCache() {
    if (dev) (void)hipFree(dev);
    if (host) (void)hipHostFree(host);
    if (uploaded) (void)hipEventDestroy(uploaded);
}

// ---------------------------------------------------------------- device
__device__ __forceinline__ uint32_t ins0(uint32_t x, int pos) {
    const uint32_t lo = x & ((1u << pos) - 1);
    return ((x >> pos) << (pos + 1)) | lo;
}

__device__ __forceinline__ uint64_t tile_offset(const PassDesc& d, uint32_t l) {
    uint64_t o = l & ((1u << d.L) - 1);
    for (int j = d.L; j < d.b; ++j)
        if ((l >> j) & 1u) o |= uint64_t(1) << d.tq[j];
    return o;
}

// The fields of a PassOp that every kind but P_DENSE uses.
struct OpHead {
    int kind, a, b;
    uint64_t lmask, omask;
    double m[8];
};

__device__ __forceinline__ OpHead op_head(const PassOp& op) {
    OpHead h;
    h.kind = op.kind;
    h.a = op.a;
    h.b = op.b;
    h.lmask = op.lmask;
    h.omask = op.omask;
#pragma unroll
    for (int j = 0; j < 8; ++j) h.m[j] = op.m[j];
    return h;
}

// DENSE selects the variant that can run dense blocks (it holds a column of outputs in
// registers, so the plain variant stays lean for passes without one).
template <int NT, bool DENSE>
__global__ __launch_bounds__(NT) void k_pass(v2d* __restrict__ psi, PassDesc d, const PassOp* __restrict__ ops,
                                             const v2d* __restrict__ mats) {
    extern __shared__ v2d tile[];
    const int b = d.b;
    const uint32_t tsize = 1u << b;
    const uint32_t tid = threadIdx.x;

    uint64_t base = xcd_tile(blockIdx.x, gridDim.x);
    for (int j = 0; j < b; ++j) base = insert0(base, d.pos[j]);

    // l = tid + k * NT: the offset splits into a per-thread and a per-iteration part.
    const uint64_t off_t = tile_offset(d, tid < tsize ? tid : 0);
    for (uint32_t k = 0; k * NT < tsize; ++k) {
        const uint32_t l = tid + k * NT;
        if (l < tsize) tile[l] = ld_nt(psi + (base | off_t | tile_offset(d, k * NT)));
    }
    __syncthreads();

    double sr = 1.0, si = 0.0;  // accumulated per-tile scalar (phases of qubits outside T)
    // The op's fields are read while the previous op runs: a one-tile state runs its whole
    // circuit in this loop, and a read of the op list on the critical path costs as much as the op.
    OpHead next = d.nops > 0 ? op_head(ops[0]) : OpHead{};
    for (int i = 0; i < d.nops; ++i) {
        const OpHead op = next;
        if (i + 1 < d.nops) next = op_head(ops[i + 1]);
        const int kind = op.kind;
        if (kind == P_SCALE) {
            const bool hi = (base >> op.a) & 1;
            const double dr = hi ? op.m[6] : op.m[0], di = hi ? op.m[7] : op.m[1];
            const double nr = sr * dr - si * di;
            si = sr * di + si * dr;
            sr = nr;
            continue;
        }
        if (kind == P_1Q) {
            const uint32_t bit = 1u << op.a;
            for (uint32_t p = tid; p < tsize / 2; p += NT) {
                const uint32_t l0 = ins0(p, op.a);
                const v2d a0 = tile[l0], a1 = tile[l0 | bit];
                tile[l0] = row2(op.m, a0, a1);
                tile[l0 | bit] = row2(op.m + 4, a0, a1);
            }
        } else if (kind == P_DIAG) {
            for (uint32_t l = tid; l < tsize; l += NT) {
                const bool hi = (l >> op.a) & 1u;
                tile[l] = hi ? cmul(op.m[6], op.m[7], tile[l]) : cmul(op.m[0], op.m[1], tile[l]);
            }
        } else if (kind == P_CNOT) {
            const int lo = op.a < op.b ? op.a : op.b, hi = op.a < op.b ? op.b : op.a;
            const uint32_t cb = 1u << op.a, tb = 1u << op.b;
            for (uint32_t p = tid; p < tsize / 4; p += NT) {
                const uint32_t l = ins0(ins0(p, lo), hi) | cb;
                const v2d x = tile[l];
                tile[l] = tile[l | tb];
                tile[l | tb] = x;
            }
        } else if (kind == P_MCZ) {
            if ((base & op.omask) == op.omask) {
                const uint32_t lm = uint32_t(op.lmask);
                for (uint32_t l = tid; l < tsize; l += NT)
                    if ((l & lm) == lm) tile[l] = -tile[l];
            }
        } else if (DENSE && kind == P_DENSE) {
            const PassOp& dop = ops[i];
            const int k = dop.a, D = 1 << k;
            uint32_t tmask = 0;
            for (int j = 0; j < k; ++j) tmask |= 1u << dop.q[j];
            const v2d* U = mats + dop.moff;
            constexpr int kMaxPer = (1 << kMaxTileBits) / NT;
            v2d out[kMaxPer];
#pragma unroll
            for (int r = 0; r < kMaxPer; ++r) {
                const uint32_t o = tid + r * NT;
                if (o < tsize) {
                    uint32_t row = 0;
                    for (int j = 0; j < k; ++j) row |= ((o >> dop.q[j]) & 1u) << j;
                    const uint32_t gb = o & ~tmask;
                    const v2d* ur = U + row * D;
                    v2d acc = {0.0, 0.0};
                    for (int c = 0; c < D; ++c) {
                        const v2d u = ur[c], x = tile[gb | dop.dep[c]];
                        acc.x += u.x * x.x - u.y * x.y;
                        acc.y += u.x * x.y + u.y * x.x;
                    }
                    out[r] = acc;
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kMaxPer; ++r) {
                const uint32_t o = tid + r * NT;
                if (o < tsize) tile[o] = out[r];
            }
        }
        __syncthreads();
    }

    const bool unit = (sr == 1.0 && si == 0.0);
    for (uint32_t k = 0; k * NT < tsize; ++k) {
        const uint32_t l = tid + k * NT;
        if (l < tsize) {
            const v2d v = tile[l];
            st_nt(psi + (base | off_t | tile_offset(d, k * NT)), unit ? v : cmul(sr, si, v));
        }
    }
}

constexpr int kPassThreads = 256;
constexpr int kSingleThreads = 1024;  // a one-tile state: 16 waves hide the LDS latency of each op

// ---------------------------------------------------------------- register-resident tiles
// The tile (2^12 amplitudes) lives in registers, 16 per thread: 4 "register bits" of the
// local index select the register, the other 8 local bits are the thread id (ascending bit
// order both ways).  A one-qubit gate or a CNOT whose target is a register bit is a pure
// register update; diagonal ops (phases, CZ, MCZ) need no data movement in any layout, since
// each thread knows the index of every amplitude it holds.  LDS is used only to change the
// register bits (write all, barrier, read all) and to feed dense blocks.  The host resolves
// every layout-dependent quantity (see RegOp), so per op the kernel reads a few uniform
// words and runs straight-line register code.
//
// LDS slots are XOR-swizzled, slot(l) = l ^ ((l >> 4) & 15): a layout change whose register
// bits include bits 4..7 would otherwise put all 64 lanes of a wavefront on one bank.  The
// swizzle is XOR-linear, so slot(lt | R(j)) = slot(lt) ^ slot(R(j)): one XOR per access.
LLSV_REGOPS_SOURCE

__device__ __forceinline__ uint64_t local_to_global(const PassDesc& d, int bit) {
    return bit < d.L ? (uint64_t(1) << bit) : (uint64_t(1) << d.tq[bit]);
}

template <int K>
__device__ __forceinline__ void reg_1q(v2d* a, const double* m) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (!(j & (1 << K))) r_pu(a[j], a[j | (1 << K)], m);
}

// CNOT with the control on a register bit: registers in rmask swap with their partner.
// Written as selects on a uniform condition, so the register array keeps one assignment
// across all op kinds (a branchy swap makes the compiler copy the whole array every op).
template <int K>
__device__ __forceinline__ void reg_swap(v2d* a, uint32_t rmask) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (!(j & (1 << K))) {
            const bool c = (rmask >> j) & 1u;
            const v2d x = a[j], y = a[j | (1 << K)];
            a[j] = c ? y : x;
            a[j | (1 << K)] = c ? x : y;
        }
}

// CNOT with the control on a thread bit: the swap happens on lanes with tok set.
template <int K>
__device__ __forceinline__ void reg_cswap(v2d* a, bool tok) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (!(j & (1 << K))) {
            const v2d x = a[j], y = a[j | (1 << K)];
            a[j] = tok ? y : x;
            a[j | (1 << K)] = tok ? x : y;
        }
}

// RO_1QN on register bit K in form F.
template <int K, int F>
__device__ __forceinline__ void reg_1qn(v2d* a, const double* m) {
    const double br = m[0], bi = m[1], cr = m[2], ci = m[3];
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (!(j & (1 << K))) {
            if (F) r_nu1(a[j], a[j | (1 << K)], br, bi, cr, ci);
            else r_nu0(a[j], a[j | (1 << K)], br, bi, cr, ci);
        }
}

template <bool DENSE>
__global__ __launch_bounds__(256) void k_pass_reg(v2d* __restrict__ psi, PassDesc d, const RegOp* __restrict__ ops,
                                                  const v2d* __restrict__ mats) {
    extern __shared__ v2d tile[];
    char* const lds = reinterpret_cast<char*>(tile);
    const uint32_t tid = threadIdx.x;
    uint64_t base = xcd_tile(blockIdx.x, gridDim.x);
    for (int j = 0; j < kRegTileBits; ++j) base = insert0(base, d.pos[j]);

    // initial layout: register bits = local bits 8..11, thread bits = 0..7 (coalesced runs of 2^L)
    uint32_t lt = tid;             // this lane's share of the local index
    uint32_t sl = r_swz(tid) << 4;  // its LDS byte slot
    v2d a[16];
    {
        uint64_t gt = 0;
        for (int j = 0; j < 8; ++j)
            if ((tid >> j) & 1u) gt |= local_to_global(d, j);
        uint64_t gs[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) gs[i] = local_to_global(d, 8 + i);
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint64_t go = ((j & 1) ? gs[0] : 0) | ((j & 2) ? gs[1] : 0) | ((j & 4) ? gs[2] : 0) | ((j & 8) ? gs[3] : 0);
            a[j] = ld_nt(psi + (base | gt | go));
        }
    }

    // Every op kind (and every target register bit) is its own if-then region with no else:
    // the register array then merges through simple two-way joins, which the compiler keeps
    // in one register assignment (an if-else chain or a switch makes it copy the array).
    double sr = 1.0, si = 0.0;
    for (int i = 0; i < d.nops; ++i) {
        const RegOp& op = ops[i];
        const int kind = op.kind;
        if (kind == RO_1Q) {
            const int k = op.a;
            if (k == 0) reg_1q<0>(a, op.m);
            if (k == 1) reg_1q<1>(a, op.m);
            if (k == 2) reg_1q<2>(a, op.m);
            if (k == 3) reg_1q<3>(a, op.m);
        }
        if (kind == RO_1QN) {
            const int k = op.a + 4 * op.q[0];
            if (k == 0) reg_1qn<0, 0>(a, op.m);
            if (k == 1) reg_1qn<1, 0>(a, op.m);
            if (k == 2) reg_1qn<2, 0>(a, op.m);
            if (k == 3) reg_1qn<3, 0>(a, op.m);
            if (k == 4) reg_1qn<0, 1>(a, op.m);
            if (k == 5) reg_1qn<1, 1>(a, op.m);
            if (k == 6) reg_1qn<2, 1>(a, op.m);
            if (k == 7) reg_1qn<3, 1>(a, op.m);
        }
        if (kind == RO_RSCALE) {
            const double f = op.m[0];
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] *= f;
        }
        if (kind == RO_MCZ && (base & op.omask) == op.omask) {
            const uint32_t sgn = ((lt & op.tmask) == op.tmask) ? 0x80000000u : 0u;
            const uint32_t rm = op.rmask;
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = r_ng(a[j], ((rm >> j) & 1u) ? sgn : 0u);
        }
        if (kind == RO_LAYOUT) {
#pragma unroll
            for (int j = 0; j < 16; ++j) *reinterpret_cast<v2d*>(lds + (sl ^ op.off0[j])) = a[j];
            __syncthreads();
            uint32_t x = tid;
#pragma unroll
            for (int t = 0; t < 4; ++t) x = r_ins0(x, op.q[t]);
            lt = x;
            sl = r_swz(x) << 4;
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = *reinterpret_cast<const v2d*>(lds + (sl ^ op.off1[j]));
            __syncthreads();
        }
        if (kind == RO_CNOT) {
            // control on a register bit (tmask = 0): registers in rmask swap on every lane;
            // on a thread bit: every pair swaps on the lanes holding that bit
            const int k = op.a;
            const uint32_t rm = op.rmask;
            if (op.tmask == 0) {
                if (k == 0) reg_swap<0>(a, rm);
                if (k == 1) reg_swap<1>(a, rm);
                if (k == 2) reg_swap<2>(a, rm);
                if (k == 3) reg_swap<3>(a, rm);
            } else {
                const bool tok = (lt & op.tmask) != 0;
                if (k == 0) reg_cswap<0>(a, tok);
                if (k == 1) reg_cswap<1>(a, tok);
                if (k == 2) reg_cswap<2>(a, tok);
                if (k == 3) reg_cswap<3>(a, tok);
            }
        }
        if (kind == RO_DIAG) {
            const uint32_t rm = op.rmask;
            const bool lane_hi = (lt & op.tmask) != 0;  // tmask = 0 when the bit is a register bit
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const bool hi = lane_hi || ((rm >> j) & 1u);
                r_sc(a[j], hi ? op.m[6] : op.m[0], hi ? op.m[7] : op.m[1]);
            }
        }
        if (kind == RO_SCALE) {
            const bool hi = (base >> op.a) & 1;
            r_phase(sr, si, hi ? op.m[6] : op.m[0], hi ? op.m[7] : op.m[1]);
        }
        if (DENSE && kind == RO_DENSE) {
#pragma unroll
            for (int j = 0; j < 16; ++j) *reinterpret_cast<v2d*>(lds + (sl ^ op.off0[j])) = a[j];
            __syncthreads();
            const int k = op.a, D = 1 << k;
            const v2d* U = mats + op.moff;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const uint32_t o = lt | op.off1[j];
                uint32_t row = 0;
                for (int t = 0; t < k; ++t) row |= ((o >> op.q[t]) & 1u) << t;
                const uint32_t gb = o & ~op.tmask;
                const v2d* ur = U + row * D;
                v2d acc = {0.0, 0.0};
                for (int c = 0; c < D; ++c) {
                    const v2d u = ur[c], x = tile[r_swz(gb | op.dep[c])];
                    acc.x += u.x * x.x - u.y * x.y;
                    acc.y += u.x * x.y + u.y * x.x;
                }
                a[j] = acc;
            }
            __syncthreads();
        }
    }

    // the planner ends every pass in a layout whose thread bits include local bits 0..L-1
    uint64_t gt = 0;
    for (int j = 0; j < kRegTileBits; ++j)
        if ((lt >> j) & 1u) gt |= local_to_global(d, j);
    uint64_t gs[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) gs[t] = local_to_global(d, d.rq[t]);
    if (si != 0.0) {
#pragma unroll
        for (int j = 0; j < 16; ++j) r_sc(a[j], sr, si);
    } else if (sr != 1.0) {  // a real factor: two products per amplitude
#pragma unroll
        for (int j = 0; j < 16; ++j) a[j] *= sr;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint64_t go = ((j & 1) ? gs[0] : 0) | ((j & 2) ? gs[1] : 0) | ((j & 4) ? gs[2] : 0) | ((j & 8) ? gs[3] : 0);
        st_nt(psi + (base | gt | go), a[j]);
    }
}

// ---------------------------------------------------------------- host: validation
namespace {

struct Checked {
    uint64_t mask;  // qubits touched
    bool diag;      // diagonal in the computational basis (commutes with other diagonal ops)
    bool single;    // can run as one streamed pass by itself
};

int check_op(int n, const llsv_op& op, uint64_t i, Checked* c) {
    auto bad = [&](const char* what) {
        return fail(LLSV_ERR_ARG, "llsv_apply_program: op %llu: %s", (unsigned long long)i, what);
    };
    auto qubit_ok = [&](int q) { return q >= 0 && q < n; };
    c->mask = 0;
    c->diag = false;
    c->single = true;
    switch (op.kind) {
        case LLSV_OP_1Q:
            if (!qubit_ok(op.q[0])) return bad("qubit outside the register");
            c->mask = uint64_t(1) << op.q[0];
            c->diag = op.m[2] == 0.0 && op.m[3] == 0.0 && op.m[4] == 0.0 && op.m[5] == 0.0;
            return LLSV_OK;
        case LLSV_OP_CNOT:
            if (!qubit_ok(op.q[0]) || !qubit_ok(op.q[1])) return bad("qubit outside the register");
            if (op.q[0] == op.q[1]) return bad("control == target");
            c->mask = (uint64_t(1) << op.q[0]) | (uint64_t(1) << op.q[1]);
            return LLSV_OK;
        case LLSV_OP_CZ:
        case LLSV_OP_MCZ:
        case LLSV_OP_DENSE: {
            const int nq = op.kind == LLSV_OP_CZ ? 2 : op.nq;
            if (op.kind == LLSV_OP_MCZ && (nq < 1 || nq > 8)) return bad("mcz needs 1..8 qubits");
            if (op.kind == LLSV_OP_DENSE && (nq < 1 || nq > LLSV_MAX_DENSE_K)) return bad("dense block needs 1..5 qubits");
            if (op.kind == LLSV_OP_DENSE && !op.dense) return bad("dense matrix is null");
            for (int j = 0; j < nq; ++j) {
                if (!qubit_ok(op.q[j])) return bad("qubit outside the register");
                const uint64_t bit = uint64_t(1) << op.q[j];
                if (c->mask & bit) return bad("duplicate qubit");
                c->mask |= bit;
            }
            c->diag = op.kind != LLSV_OP_DENSE;
            c->single = op.kind != LLSV_OP_DENSE;
            return LLSV_OK;
        }
        default:
            return bad("unknown op kind");
    }
}

struct PlannedPass {
    std::vector<uint64_t> ops;
    uint64_t tile = 0;
};

int popc(uint64_t x) { return __builtin_popcountll(x); }

// The pass schedule.  A pass takes a set T of b qubits (always the L lowest, so a tile is made of
// whole 128-byte lines) and every pending op that can run inside it: an op whose non-diagonal
// qubits lie in T (diagonal ops run in any tile: each tile knows its fixed bits) and whose
// predecessors on each of its qubits have run (earlier passes, or earlier in this pass) -- so ops
// move only past ops they share no qubit with.  T grows as ops join, so which ops join first
// decides it: the planner tries a scan seeded at the first pending op of every qubit (and at the
// first pending op overall, which keeps circuit order when everything fits), each swept to a
// fixpoint over a window of pending ops, once with T holding only the low bits and once with the
// seed op's qubits in T from the start, and keeps the candidate that schedules the most ops.  On a
// 20-qubit hardware-efficient ansatz (10 layers) this gives 10 passes where a single in-order scan
// gives 17; on the 28-qubit layered circuit (20 layers) 11 instead of 16.
std::vector<PlannedPass> plan(int n, const llsv_op* ops, const std::vector<Checked>& info, int b, int L) {
    const uint64_t n_ops = info.size();
    const uint64_t all = n >= 64 ? ~uint64_t(0) : ((uint64_t(1) << n) - 1);
    std::vector<PlannedPass> passes;
    if (b > n) b = n;
    if (L > b) L = b;
    // predecessors: for every op, the previous op on each of its qubits (CSR)
    std::vector<uint64_t> dep_start(n_ops + 1, 0), deps;
    {
        std::vector<int64_t> last(size_t(std::max(n, 1)), -1);
        deps.reserve(n_ops * 2);
        for (uint64_t i = 0; i < n_ops; ++i) {
            dep_start[i] = deps.size();
            const uint64_t m = info[i].mask;
            for (int q = 0; q < n; ++q)
                if ((m >> q) & 1u) {
                    if (last[size_t(q)] >= 0 &&
                        std::find(deps.begin() + int64_t(dep_start[i]), deps.end(), uint64_t(last[size_t(q)])) == deps.end())
                        deps.push_back(uint64_t(last[size_t(q)]));
                    last[size_t(q)] = int64_t(i);
                }
        }
        dep_start[n_ops] = deps.size();
    }
    constexpr uint64_t kWindow = 4096;  // pending ops a pass considers
    constexpr int kSweeps = 8;
    std::vector<char> done(n_ops, 0);
    std::vector<uint32_t> mark(n_ops, 0);  // == stamp: taken by the candidate being built
    uint32_t stamp = 0;
    std::vector<uint64_t> cand, best;
    uint64_t first = 0;
    while (first < n_ops) {
        // the window of pending ops and the candidate seeds
        std::vector<uint64_t> win;
        std::vector<uint64_t> seeds{first};
        uint64_t seen = 0;
        for (uint64_t i = first; i < n_ops && win.size() < kWindow; ++i) {
            if (done[i]) continue;
            win.push_back(i);
            const uint64_t fresh = info[i].mask & ~seen;
            if (fresh) {
                if (i != first) seeds.push_back(uint64_t(win.size() - 1));  // position in win
                seen |= info[i].mask;
            }
        }
        seeds[0] = 0;  // position of `first` in win
        uint64_t bestT = 0;
        best.clear();
        for (size_t si = 0; si < 2 * seeds.size(); ++si) {
            const uint64_t sp = seeds[si / 2];
            ++stamp;
            cand.clear();
            uint64_t T = (b >= n) ? all : ((uint64_t(1) << L) - 1);
            // odd candidates start with the seed op's qubits in T, so that ops which are ready
            // earlier in the window cannot fill the tile before the seed's own chain runs
            if (si & 1) {
                const Checked& c = info[win[sp]];
                if (c.diag || popc(T | c.mask) > b || (c.mask & ~T) == 0) continue;
                T |= c.mask;
            }
            for (int sweep = 0; sweep < kSweeps; ++sweep) {
                bool changed = false;
                for (uint64_t k = 0; k < win.size(); ++k) {
                    const uint64_t i = win[(sp + k) % win.size()];
                    if (mark[i] == stamp) continue;
                    bool ready = true;
                    for (uint64_t d = dep_start[i]; d < dep_start[i + 1] && ready; ++d)
                        ready = done[deps[d]] || mark[deps[d]] == stamp;
                    if (!ready) continue;
                    const Checked& c = info[i];
                    if (!c.diag) {
                        const uint64_t need = c.mask & ~T;
                        if (popc(T) + popc(need) > b) continue;
                        T |= need;
                    }
                    mark[i] = stamp;
                    cand.push_back(i);
                    changed = true;
                }
                if (!changed) break;
            }
            if (cand.size() > best.size()) {
                best.swap(cand);
                bestT = T;
            }
        }
        PlannedPass p;
        p.ops = best;
        for (uint64_t i : best) done[i] = 1;
        uint64_t T = bestT;
        for (int q = 0; q < n && popc(T) < b; ++q) T |= uint64_t(1) << q;  // longer contiguous runs
        p.tile = T;
        passes.push_back(std::move(p));
        while (first < n_ops && done[first]) ++first;
    }
    (void)ops;
    return passes;
}

}  // namespace

int device_tile_bits(llsv_state* s) {
    if (s->program->max_tile_bits) return s->program->max_tile_bits;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device) != hipSuccess) lds = 65536;
    int b = 1;
    while (b < kMaxTileBits && (size_t(16) << (b + 1)) <= size_t(lds)) ++b;
    s->program->max_tile_bits = b;
    return b;
}

namespace {

int ensure_capacity(llsv_state* s, size_t bytes) {
    ProgramCache& pc = *s->program;
    if (!pc.uploaded && hipEventCreateWithFlags(&pc.uploaded, hipEventDisableTiming) != hipSuccess)
        return fail(LLSV_ERR_DEVICE, "llsv_apply_program: event creation failed");
    if (bytes <= pc.cap) {
        // the staging buffer is rewritten only after the previous upload has finished
        if (hipEventSynchronize(pc.uploaded) != hipSuccess)
            return fail(LLSV_ERR_DEVICE, "llsv_apply_program: upload wait failed");
        return LLSV_OK;
    }
    size_t cap = 1 << 16;
    while (cap < bytes) cap *= 2;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_apply_program: sync failed");
    if (pc.dev) (void)hipFree(pc.dev);
    if (pc.host) (void)hipHostFree(pc.host);
    pc.dev = nullptr;
    pc.host = nullptr;
    pc.cap = 0;
    if (dev_alloc(reinterpret_cast<void**>(&pc.dev), cap) != hipSuccess ||
        dev_host_alloc(reinterpret_cast<void**>(&pc.host), cap) != hipSuccess) {
        (void)hipGetLastError();
        return fail(LLSV_ERR_NOMEM, "llsv_apply_program: cannot allocate %zu bytes of program buffers", cap);
    }
    pc.cap = cap;
    return LLSV_OK;
}

}  // namespace

int plan_order(int n, const llsv_op* ops, uint64_t n_ops, int b, int L, uint64_t* order, uint64_t* pass_end) {
    std::vector<Checked> info(n_ops);
    for (uint64_t i = 0; i < n_ops; ++i)
        if (int rc = check_op(n, ops[i], i, &info[i])) return rc;
    const std::vector<PlannedPass> passes = plan(n, ops, info, b, L);
    uint64_t k = 0;
    for (size_t p = 0; p < passes.size(); ++p) {
        for (uint64_t i : passes[p].ops) {
            if (order) order[k] = i;
            ++k;
        }
        if (pass_end) pass_end[p] = k;
    }
    return int(passes.size());
}

static constexpr int kLowBits = 3;

namespace {

// ---- plan cache --------------------------------------------------------------------------------
// The schedule depends only on the circuit's structure (op kinds, qubits, which one-qubit matrices
// are diagonal), not on the matrix values, so a circuit applied again with new angles reuses it.
// A register-tile schedule also picks its tile's low bits: tiles are made of contiguous runs of
// 2^L amplitudes (16 << L bytes), and runs of 256 or 512 bytes stream faster than 128-byte ones
// (a 30-qubit one-qubit + CNOT-ladder pass: 5.5 against 4.2 TB/s), but every low bit is a tile
// bit that no gate can use.  The schedule keeps the largest L in [3, 5] that needs no more passes
// than L = 3 does.
struct PlanEntry {
    int L = 0;
    std::vector<PlannedPass> passes;
};

constexpr size_t kPlanCacheEntries = 16;
std::mutex g_plan_mu;
std::list<std::pair<std::string, std::shared_ptr<const PlanEntry>>>& plan_cache() {
    static auto* c = new std::list<std::pair<std::string, std::shared_ptr<const PlanEntry>>>();  // most recent first
    return *c;
}

std::string plan_key(int n, int b, bool reg_tiles, const llsv_op* ops, const std::vector<Checked>& info) {
    std::string k;
    k.reserve(8 + info.size() * 6);
    k += char(n);
    k += char(b);
    k += char(reg_tiles);
    for (size_t i = 0; i < info.size(); ++i) {
        const llsv_op& op = ops[i];
        const int nq = op.kind == LLSV_OP_1Q ? 1 : (op.kind == LLSV_OP_CNOT || op.kind == LLSV_OP_CZ) ? 2 : op.nq;
        k += char(op.kind);
        k += char(nq | (info[i].diag ? 0x40 : 0));
        for (int t = 0; t < nq; ++t) k += char(op.q[t]);
    }
    return k;
}

std::shared_ptr<const PlanEntry> cached_plan(int n, int b, bool reg_tiles, const llsv_op* ops,
                                             const std::vector<Checked>& info) {
    const std::string key = plan_key(n, b, reg_tiles, ops, info);
    {
        std::lock_guard<std::mutex> lock(g_plan_mu);
        auto& c = plan_cache();
        for (auto it = c.begin(); it != c.end(); ++it)
            if (it->first == key) {
                c.splice(c.begin(), c, it);
                return c.front().second;
            }
    }
    auto e = std::make_shared<PlanEntry>();
    e->L = std::min(kLowBits, b);
    e->passes = plan(n, ops, info, b, e->L);
    if (reg_tiles && n > b) {
        for (int L = kLowBits + 1; L <= 5; ++L) {
            std::vector<PlannedPass> alt = plan(n, ops, info, b, L);
            if (alt.size() > e->passes.size()) break;
            e->L = L;
            e->passes.swap(alt);
        }
    }
    std::lock_guard<std::mutex> lock(g_plan_mu);
    auto& c = plan_cache();
    c.emplace_front(key, e);
    if (c.size() > kPlanCacheEntries) c.pop_back();
    return e;
}

}  // namespace

namespace {

// prev := next . prev for two one-qubit ops on the same tile bit (next applied after prev).
// The product is diagonal when both are.
void merge_1q(PassOp* prev, const PassOp& next) {
    const double* a = next.m;
    double* b = prev->m;
    double r[8];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < 2; ++k) {
                const double xr = a[2 * (2 * i + k)], xi = a[2 * (2 * i + k) + 1];
                const double yr = b[2 * (2 * k + j)], yi = b[2 * (2 * k + j) + 1];
                re += xr * yr - xi * yi;
                im += xr * yi + xi * yr;
            }
            r[2 * (2 * i + j)] = re;
            r[2 * (2 * i + j) + 1] = im;
        }
    std::memcpy(b, r, sizeof r);
    prev->kind = (prev->kind == P_DIAG && next.kind == P_DIAG) ? P_DIAG : P_1Q;
}

// ---- one-qubit matrices without their row scales ---------------------------------------------
// A merged one-qubit matrix U = [[a, b], [c, d]] costs 16 FP64 operations per amplitude pair as a
// general 2x2.  Written U = diag(a, d) [[1, b/a], [c/d, 1]] (or diag(b, c) [[a/b, 1], [1, d/c]]
// when |b c| > |a d|), the right-hand factor costs 8, one FMA per product (r_nu0 / r_nu1) -- and
// the diagonal need not be applied where it stands: its modulus |a| (|b|) is a real factor of the
// whole tile, multiplied up over the pass and applied once at its end (P_RSCALE), and the rest,
// diag(a, d) / |a|, commutes with every diagonal op (CZ, MCZ, phases) and with a CNOT it controls,
// so it is carried forward on its qubit, into the next matrix on that qubit (in this pass or a
// later one), as a global factor and a relative phase d / a.  What is left at the end of the
// program becomes per-tile scalars (qubits outside the last tile) or diagonal ops; a CNOT target
// or a dense block applies what its qubits carry first.  Both register-tile kernels (interpreted
// and compiled) see the rewritten ops, so they still agree bit for bit.
using cx = std::complex<double>;

bool unit_modulus(cx z) { return std::abs(std::norm(z) - 1.0) <= 1e-12; }

// Phases carried between the passes of one program: a global factor and, per qubit, the relative
// phase of its |1> (the diagonal diag(1, rel[q])).
struct CarriedPhases {
    cx glob = 1.0;
    std::vector<cx> rel;
};

PassOp diag_op(int bit, cx d0, cx d1, bool resident) {
    PassOp po{};
    po.kind = P_DIAG;
    po.a = bit;
    po.b = resident ? 1 : 0;
    po.m[0] = d0.real();
    po.m[1] = d0.imag();
    po.m[6] = d1.real();
    po.m[7] = d1.imag();
    return po;
}

PassOp scale_op(int qubit, cx d0, cx d1) {
    PassOp po = diag_op(0, d0, d1, false);
    po.kind = P_SCALE;
    po.a = qubit;
    return po;
}

// Rewrites a register-tile pass's ops (after merging): unitaries split as above where the phase
// they leave can move on, diagonal unit-modulus ops carried.  tq: global qubit of each tile bit.
// last: the program's final pass, which applies whatever is still carried.
void carry_phases(std::vector<PassOp>& pops, const int* tq, int b, int n, bool last, CarriedPhases* cp) {
    std::vector<PassOp> out;
    out.reserve(pops.size() + 8);
    double rscale = 1.0;  // product of the row scales taken out of P_1QN ops not yet applied
    auto emit_rscale = [&] {
        PassOp rs{};
        rs.kind = P_RSCALE;
        rs.m[0] = rscale;
        out.push_back(rs);
        rscale = 1.0;
    };
    auto flush = [&](int bit, bool resident) {
        cx& r = cp->rel[size_t(tq[bit])];
        if (r != 1.0) out.push_back(diag_op(bit, 1.0, r, resident));
        r = 1.0;
    };
    for (size_t i = 0; i < pops.size(); ++i) {
        PassOp po = pops[i];
        switch (po.kind) {
            case P_1Q: {
                const int t = po.a;
                cx& r = cp->rel[size_t(tq[t])];
                // U diag(1, r): the carried phase goes into U's second column
                const cx a(po.m[0], po.m[1]), u10(po.m[4], po.m[5]);
                const cx b = cx(po.m[2], po.m[3]) * r, d = cx(po.m[6], po.m[7]) * r;
                // U = diag(a, d) [[1, b/a], [c/d, 1]] (form 0) or diag(b, c) [[a/b, 1], [1, d/c]]
                // (form 1).  The divisions lose no accuracy (the quotient's rounding error, scaled
                // back, is relative to the term it multiplies); the bound only keeps the quotients
                // and the pass's scale in range.  Form 0 whenever its divisors are not below 1e-6
                // of the other pair, so that angles drifting between calls almost never change a
                // pass's kernel; the row factors leave as |a| (|b|) into the tile's real scale and a
                // diagonal carried on the qubit like any other.
                const double na = std::abs(a), nb = std::abs(b), nc = std::abs(u10), nd = std::abs(d);
                const double kRatio = 1e-6;
                int form = -1;
                if (na > 0.0 && nd > 0.0 && std::min(na, nd) >= kRatio * std::max(nb, nc)) form = 0;
                else if (nb > 0.0 && nc > 0.0 && std::min(nb, nc) >= kRatio * std::max(na, nd)) form = 1;
                if (form < 0) {
                    po.m[2] = b.real();
                    po.m[3] = b.imag();
                    po.m[6] = d.real();
                    po.m[7] = d.imag();
                    r = 1.0;
                    out.push_back(po);
                    break;
                }
                const cx p0 = form ? b : a, p1 = form ? u10 : d;
                const cx beta = form ? a / b : b / a, gamma = form ? d / u10 : u10 / d;
                const double sig = std::abs(p0);
                PassOp nu{};
                nu.kind = P_1QN;
                nu.a = t;
                nu.b = form;
                nu.m[0] = beta.real();
                nu.m[1] = beta.imag();
                nu.m[2] = gamma.real();
                nu.m[3] = gamma.imag();
                out.push_back(nu);
                rscale *= sig;
                if (rscale < 1e-100 || rscale > 1e100) emit_rscale();
                cp->glob *= p0 / sig;
                r = p1 / p0;
                break;
            }
            case P_DIAG:
            case P_SCALE: {
                const cx d0(po.m[0], po.m[1]), d1(po.m[6], po.m[7]);
                if (!unit_modulus(d0) || !unit_modulus(d1)) {
                    out.push_back(po);
                    break;
                }
                const int q = po.kind == P_DIAG ? tq[po.a] : po.a;
                cp->glob *= d0;
                cp->rel[size_t(q)] *= d1 / d0;
                break;
            }
            case P_CNOT:
                flush(po.b, true);  // the target's phase does not commute with it
                out.push_back(po);
                break;
            case P_DENSE:
                for (int q = 0; q < po.a; ++q) flush(po.q[q], false);
                out.push_back(po);
                break;
            default:  // P_MCZ: diagonal
                out.push_back(po);
                break;
        }
    }
    if (rscale != 1.0) emit_rscale();
    if (last) {
        std::vector<char> in_tile(size_t(n), 0);
        for (int t = 0; t < b; ++t) in_tile[size_t(tq[t])] = 1;
        int outside = -1;
        for (int q = 0; q < n; ++q) {
            if (!in_tile[size_t(q)]) outside = q;
            cx& r = cp->rel[size_t(q)];
            if (r == 1.0 || in_tile[size_t(q)]) continue;
            out.push_back(scale_op(q, 1.0, r));
            r = 1.0;
        }
        for (int t = 0; t < b; ++t) flush(t, false);
        if (cp->glob != 1.0) {
            out.push_back(outside >= 0 ? scale_op(outside, cp->glob, cp->glob) : diag_op(0, cp->glob, cp->glob, false));
            cp->glob = 1.0;
        }
    }
    pops.swap(out);
}

// Local index bits of register j when the register bits are S (4 bits of a 12-bit mask):
// bit t of j lands on the t-th lowest bit of S.
uint32_t reg_local(uint32_t S, int j) {
    uint32_t o = 0;
    int t = 0;
    for (int bit = 0; bit < kRegTileBits; ++bit)
        if ((S >> bit) & 1u) {
            if ((j >> t) & 1) o |= 1u << bit;
            ++t;
        }
    return o;
}

// LDS byte slot of the register part of a local index (the swizzle of the kernel, x 16 B).
uint32_t reg_slot(uint32_t S, int j) {
    const uint32_t l = reg_local(S, j);
    return (l ^ ((l >> 4) & 15u)) << 4;
}

// Registers whose register part of the local index contains every bit of m (m within S).
uint32_t regs_with(uint32_t S, uint32_t m) {
    uint32_t r = 0;
    for (int j = 0; j < 16; ++j)
        if ((reg_local(S, j) & m) == m) r |= 1u << j;
    return r;
}

}  // namespace

// Turns a pass's ops (tile-local coordinates) into the register-resident program: a layout
// change is inserted before every one-qubit gate or CNOT whose target bit is not held in
// registers; the new register set takes that bit, then the target bits of the next ops (so
// one change serves several gates), then bits of the current set.  The pass ends in a layout
// whose thread bits include local bits 0..L-1, so the stores write whole lines; its register
// bits go to rq.
std::vector<RegOp> to_register_program(const std::vector<PassOp>& in, int L, int rq[4]) {
    constexpr uint32_t kStart = 0xF00u;  // local bits 8..11
    const uint32_t low = (1u << L) - 1;
    auto need_of = [](const PassOp& po) -> uint32_t {
        if (po.kind == P_DIAG && po.b) return 1u << po.a;  // wants its bit in registers (half the work)
        if (po.kind == P_1QT && po.b == 3) return 0u;   // diagonal: any residency
        if (po.kind == P_GRAD && po.b == 2) return 0u;  // Z generator: any residency
        if (po.kind == P_1Q || po.kind == P_1QN || po.kind == P_1QT || po.kind == P_GRAD) return 1u << po.a;
        if (po.kind == P_CNOT) return 1u << po.b;
        return 0u;
    };
    auto rank = [](uint32_t set, int bit) { return __builtin_popcount(set & ((1u << bit) - 1)); };
    auto layout = [](uint32_t from, uint32_t to) {
        RegOp lay{};
        lay.kind = RO_LAYOUT;
        int t = 0;
        for (int bit = 0; bit < kRegTileBits; ++bit)
            if ((to >> bit) & 1u) lay.q[t++] = bit;
        for (int j = 0; j < 16; ++j) {
            lay.off0[j] = reg_slot(from, j);
            lay.off1[j] = reg_slot(to, j);
        }
        return lay;
    };
    // Ops run in an order of their own: an op may move past ops that touch none of its tile bits
    // (they commute).  Ready ops whose bits are in registers run first, in input order; when none
    // is left, the new register set takes the bits that ready ops need (then ops after them).  An
    // op sequence that goes layer by layer over the tile's qubits then continues on the qubits
    // already in registers instead of starting each layer over (the light-cone pass of a layered
    // circuit: 21 layout changes in input order, 11 this way).
    const size_t N = in.size();
    auto bits_of = [](const PassOp& po) -> uint32_t {
        switch (po.kind) {
            case P_CNOT:
                return (1u << po.a) | (1u << po.b);
            case P_MCZ:
                return uint32_t(po.lmask);
            case P_SCALE:
                return 0u;  // a per-tile scalar commutes with every op of the tile
            case P_RSCALE:
                // keeps its place: it undoes the growth of the P_1QN ops before it (range only)
                return (1u << kMaxTileBits) - 1u;
            case P_DENSE: {
                uint32_t m = 0;
                for (int t = 0; t < po.a; ++t) m |= 1u << po.q[t];
                return m;
            }
            default:
                return 1u << po.a;
        }
    };
    std::vector<uint32_t> need(N);
    std::vector<int> npred(N, 0);
    std::vector<std::vector<int>> succ(N);
    {
        int last_on[kMaxTileBits];
        for (int t = 0; t < kMaxTileBits; ++t) last_on[t] = -1;
        for (size_t i = 0; i < N; ++i) {
            need[i] = need_of(in[i]);
            const uint32_t m = bits_of(in[i]);
            for (int t = 0; t < kMaxTileBits; ++t) {
                if (!((m >> t) & 1u)) continue;
                const int p = last_on[t];
                if (p >= 0 && (succ[size_t(p)].empty() || succ[size_t(p)].back() != int(i))) {
                    succ[size_t(p)].push_back(int(i));
                    ++npred[i];
                }
                last_on[t] = int(i);
            }
        }
    }
    std::vector<int> ready;  // ascending
    for (size_t i = 0; i < N; ++i)
        if (!npred[i]) ready.push_back(int(i));
    std::vector<char> done(N, 0);
    std::vector<int> work, undo;  // scratch of the lookahead
    std::vector<RegOp> out;
    out.reserve(N + 8);
    uint32_t S = kStart;
    for (size_t emitted = 0; emitted < N;) {
        int pick = -1;
        for (int i : ready)
            if (!(need[size_t(i)] & ~S)) {
                pick = i;
                break;
            }
        if (pick < 0) {
            // A CNOT whose control is a register bit too is a renaming of registers (no
            // instructions); with the control on a lane bit it is a select per register word.  So
            // after the bits ops need, a CNOT's control joins the set where there is room.
            auto control_of = [&](size_t k, uint32_t set) -> uint32_t {
                const PassOp& po = in[k];
                return (po.kind == P_CNOT && ((set >> po.b) & 1u)) ? 1u << po.a : 0u;
            };
            uint32_t ns = 0;
            for (int i : ready) {
                const uint32_t nb = need[size_t(i)];
                if (__builtin_popcount(ns | nb) <= kRegBits) ns |= nb;
                if (__builtin_popcount(ns) == kRegBits) break;
            }
            for (int i : ready) {
                const uint32_t cb = control_of(size_t(i), ns);
                if (__builtin_popcount(ns | cb) <= kRegBits) ns |= cb;
            }
            for (size_t k = 0; k < N && __builtin_popcount(ns) < kRegBits; ++k) {
                if (done[k]) continue;
                const uint32_t nb = need[k];
                if (__builtin_popcount(ns | nb) <= kRegBits) ns |= nb;
                const uint32_t cb = control_of(k, ns);
                if (__builtin_popcount(ns | cb) <= kRegBits) ns |= cb;
            }
            for (int bit = kRegTileBits - 1; bit >= 0 && __builtin_popcount(ns) < kRegBits; --bit)
                if ((S >> bit) & 1u) ns |= 1u << bit;
            for (int bit = kRegTileBits - 1; bit >= 0 && __builtin_popcount(ns) < kRegBits; --bit) ns |= 1u << bit;
            // Lookahead: every 4-bit set is scored by the ops it lets run before the next change
            // (a CNOT whose control stays on a lane bit counts half, ops that need no register bit
            // a quarter); the set built above is the default, another must score higher.
            auto score = [&](uint32_t cand) {
                double sc = 0.0;
                work.assign(ready.begin(), ready.end());
                undo.clear();
                for (size_t w = 0; w < work.size(); ++w) {
                    const size_t i = size_t(work[w]);
                    if (need[i] & ~cand) continue;
                    const PassOp& po = in[i];
                    sc += !need[i] ? 0.25 : (po.kind == P_CNOT && !((cand >> po.a) & 1u)) ? 0.5 : 1.0;
                    for (int sx : succ[i]) {
                        undo.push_back(sx);
                        if (--npred[size_t(sx)] == 0) work.push_back(sx);
                    }
                }
                for (int sx : undo) ++npred[size_t(sx)];
                return sc;
            };
            double best = score(ns);
            for (uint32_t c = 0xFu; c < (1u << kRegTileBits);) {
                const double sc = score(c);
                if (sc > best) {
                    best = sc;
                    ns = c;
                }
                const uint32_t lo = c & (~c + 1u), r = c + lo;  // next mask with 4 bits set
                c = (((r ^ c) >> 2) / lo) | r;
            }
            out.push_back(layout(S, ns));
            S = ns;
            continue;
        }
        ready.erase(std::find(ready.begin(), ready.end(), pick));
        done[size_t(pick)] = 1;
        ++emitted;
        for (int sx : succ[size_t(pick)])
            if (--npred[size_t(sx)] == 0) ready.insert(std::lower_bound(ready.begin(), ready.end(), sx), sx);
        const PassOp& po = in[size_t(pick)];
        RegOp ro{};
        std::memcpy(ro.m, po.m, sizeof ro.m);
        switch (po.kind) {
            case P_1Q:
                ro.kind = RO_1Q;
                ro.a = rank(S, po.a);
                break;
            case P_1QN:
                ro.kind = RO_1QN;
                ro.a = rank(S, po.a);
                ro.q[0] = po.b;
                break;
            case P_RSCALE:
                ro.kind = RO_RSCALE;
                break;
            case P_CNOT: {
                ro.kind = RO_CNOT;
                ro.a = rank(S, po.b);
                const uint32_t cb = 1u << po.a;
                if (S & cb) {
                    ro.rmask = regs_with(S, cb);
                } else {
                    ro.rmask = 0xFFFFu;
                    ro.tmask = cb;
                }
                break;
            }
            case P_DIAG: {
                ro.kind = RO_DIAG;
                const uint32_t bit = 1u << po.a;
                if (S & bit) {
                    ro.rmask = regs_with(S, bit);
                    ro.a = (po.m[0] == 1.0 && po.m[1] == 0.0) ? 1 : 0;
                } else {
                    ro.tmask = bit;
                }
                break;
            }
            case P_MCZ: {
                ro.kind = RO_MCZ;
                const uint32_t lm = uint32_t(po.lmask);
                ro.rmask = regs_with(S, lm & S);
                ro.tmask = lm & ~S;
                ro.omask = po.omask;
                break;
            }
            case P_SCALE:
                ro.kind = RO_SCALE;
                ro.a = po.a;
                break;
            case P_1QT:
            case P_GRAD: {
                ro.kind = po.kind == P_1QT ? RO_1QT : RO_GRAD;
                ro.q[0] = po.b;
                ro.moff = po.moff;
                const uint32_t bit = 1u << po.a;
                if (S & bit) {
                    ro.a = rank(S, po.a);
                    ro.rmask = regs_with(S, bit);
                } else {
                    ro.tmask = bit;  // diagonal forms only (need_of)
                }
                break;
            }
            case P_DENSE: {
                ro.kind = RO_DENSE;
                ro.a = po.a;
                ro.moff = po.moff;
                for (int t = 0; t < po.a; ++t) {
                    ro.q[t] = po.q[t];
                    ro.tmask |= 1u << po.q[t];
                }
                for (int c = 0; c < (1 << po.a); ++c) ro.dep[c] = uint16_t(po.dep[c]);
                for (int j = 0; j < 16; ++j) {
                    ro.off0[j] = reg_slot(S, j);
                    ro.off1[j] = reg_local(S, j);
                }
                break;
            }
        }
        out.push_back(ro);
    }
    if (S & low) {
        out.push_back(layout(S, kStart));
        S = kStart;
    }
    int t = 0;
    for (int bit = 0; bit < kRegTileBits; ++bit)
        if ((S >> bit) & 1u) rq[t++] = bit;
    return out;
}

namespace {

// One kernel launch of a program: a streamed single op, a lone dense block on the matrix cores,
// or a pass (LDS tile or register tile, interpreted or compiled).
struct Launch {
    bool streamed = false;
    bool dense = false;
    bool mfma = false;  // a lone dense block of 4-5 qubits: the FP64-MFMA kernel
    uint64_t op = 0;
    PassDesc d{};
    DenseTile dt{};
    size_t first = 0;  // first op of the pass in ops (LDS tiles) or rops (register tiles)
    size_t moff = 0;
};

struct BuiltProgram {
    int b = 0;
    bool reg_tiles = false;
    std::vector<Launch> launches;
    std::vector<PassOp> ops;
    std::vector<RegOp> rops;
    std::vector<const llsv_op*> dense_src;  // dense ops in upload order
    size_t mat_v2d = 0;
};

// Host half of llsv_apply_program: validation, the pass schedule and every pass's op list.
// single_lds: a state that is one tile (n <= tile bits) runs its pass LDS-resident (k_pass with
// kSingleThreads threads) instead of in register tiles.
int build_program(int n, const llsv_op* ops, uint64_t n_ops, int tile_bits, BuiltProgram* out,
                  bool single_lds = false) {
    std::vector<Checked> info(n_ops);
    for (uint64_t i = 0; i < n_ops; ++i)
        if (int rc = check_op(n, ops[i], i, &info[i])) return rc;

    const int b = std::min(n, tile_bits);
    const bool reg_tiles = (b == kRegTileBits) && !(single_lds && n == b);
    out->b = b;
    out->reg_tiles = reg_tiles;
    const std::shared_ptr<const PlanEntry> pe = cached_plan(n, b, reg_tiles, ops, info);
    const int L = pe->L;
    const std::vector<PlannedPass>& passes = pe->passes;
    out->launches.reserve(passes.size());
    // phases are carried between passes only when every pass runs in register tiles
    bool carry = reg_tiles;
    for (const PlannedPass& p : passes)
        carry = carry && !(p.ops.size() == 1 && (info[p.ops[0]].single || ops[p.ops[0]].kind == LLSV_OP_DENSE) && n > b);
    CarriedPhases cp;
    cp.rel.assign(size_t(n), cx(1.0));
    for (const PlannedPass& p : passes) {
        Launch la{};
        if (p.ops.size() == 1 && info[p.ops[0]].single && n > b) {
            la.streamed = true;
            la.op = p.ops[0];
            out->launches.push_back(la);
            continue;
        }
        if (p.ops.size() == 1 && ops[p.ops[0]].kind == LLSV_OP_DENSE &&
            make_dense_tile(n, ops[p.ops[0]].nq, ops[p.ops[0]].q, &la.d, &la.dt)) {
            la.mfma = true;
            la.moff = out->mat_v2d;
            out->dense_src.push_back(&ops[p.ops[0]]);
            out->mat_v2d += size_t(1) << (2 * ops[p.ops[0]].nq);
            out->launches.push_back(la);
            continue;
        }
        PassDesc& d = la.d;
        d.n = n;
        d.b = b;
        d.L = L;
        int local_of[64];
        for (int q = 0; q < 64; ++q) local_of[q] = -1;
        int j = 0;
        for (int q = 0; q < L; ++q) d.tq[j++] = q;
        for (int q = L; q < n; ++q)
            if ((p.tile >> q) & 1) d.tq[j++] = q;
        for (int t = 0; t < b; ++t) {
            local_of[d.tq[t]] = t;
            d.pos[t] = d.tq[t];
        }
        std::sort(d.pos, d.pos + b);
        std::vector<PassOp> pops;
        pops.reserve(p.ops.size());
        // last_1q[bit]: index in pops of the latest one-qubit op on tile bit `bit` that no later op
        // touching that bit follows; a new one-qubit op on the bit is multiplied into it.
        int last_1q[kMaxTileBits];
        for (int t = 0; t < kMaxTileBits; ++t) last_1q[t] = -1;
        for (uint64_t i : p.ops) {
            const llsv_op& op = ops[i];
            PassOp po{};
            switch (op.kind) {
                case LLSV_OP_1Q:
                    std::memcpy(po.m, op.m, sizeof po.m);
                    if (info[i].diag) {
                        if (local_of[op.q[0]] >= 0) {
                            po.kind = P_DIAG;
                            po.a = local_of[op.q[0]];
                        } else {
                            po.kind = P_SCALE;
                            po.a = op.q[0];
                        }
                    } else {
                        po.kind = P_1Q;
                        po.a = local_of[op.q[0]];
                    }
                    break;
                case LLSV_OP_CNOT:
                    po.kind = P_CNOT;
                    po.a = local_of[op.q[0]];
                    po.b = local_of[op.q[1]];
                    break;
                case LLSV_OP_CZ:
                case LLSV_OP_MCZ: {
                    po.kind = P_MCZ;
                    const int nq = op.kind == LLSV_OP_CZ ? 2 : op.nq;
                    for (int t = 0; t < nq; ++t) {
                        const int q = op.q[t];
                        if (local_of[q] >= 0) po.lmask |= uint64_t(1) << local_of[q];
                        else po.omask |= uint64_t(1) << q;
                    }
                    break;
                }
                case LLSV_OP_DENSE: {
                    la.dense = true;
                    po.kind = P_DENSE;
                    po.a = op.nq;
                    for (int t = 0; t < op.nq; ++t) po.q[t] = local_of[op.q[t]];
                    po.moff = int(out->mat_v2d);
                    for (int c = 0; c < (1 << op.nq); ++c) {
                        uint32_t o = 0;
                        for (int t = 0; t < op.nq; ++t)
                            if ((c >> t) & 1) o |= 1u << po.q[t];
                        po.dep[c] = o;
                    }
                    out->dense_src.push_back(&op);
                    out->mat_v2d += size_t(1) << (2 * op.nq);
                    break;
                }
            }
            if (po.kind == P_1Q || po.kind == P_DIAG) {
                const int bit = po.a;
                if (last_1q[bit] >= 0) {
                    PassOp& prev = pops[last_1q[bit]];
                    merge_1q(&prev, po);
                    continue;
                }
                last_1q[bit] = int(pops.size());
            } else if (po.kind == P_CNOT) {
                last_1q[po.a] = last_1q[po.b] = -1;
            } else if (po.kind == P_MCZ) {
                for (int t = 0; t < b; ++t)
                    if ((po.lmask >> t) & 1) last_1q[t] = -1;
            } else if (po.kind == P_DENSE) {
                for (int t = 0; t < po.a; ++t) last_1q[po.q[t]] = -1;
            }
            pops.push_back(po);
        }
        if (carry) carry_phases(pops, d.tq, b, n, &p == &passes.back(), &cp);
        if (reg_tiles) {
            const std::vector<RegOp> rops = to_register_program(pops, L, d.rq);
            d.nops = int(rops.size());
            la.first = out->rops.size();
            out->rops.insert(out->rops.end(), rops.begin(), rops.end());
        } else {
            d.nops = int(pops.size());
            la.first = out->ops.size();
            out->ops.insert(out->ops.end(), pops.begin(), pops.end());
        }
        out->launches.push_back(la);
    }

    return LLSV_OK;
}

}  // namespace

int run_program(llsv_state* s, const llsv_op* ops, uint64_t n_ops) {
    const int n = s->n;
    BuiltProgram bp;
    // A state that fits one tile is latency-bound (one workgroup does all the work): there the
    // LDS-resident pass with 1024 threads, one barrier per op, beats register tiles, whose ops on
    // lane bits each need a full exchange through LDS.  In the automatic mode only: mode 0 runs the
    // register-tile interpreter and mode 2 compiles register tiles (the two agree bit for bit).
    if (int rc = build_program(n, ops, n_ops, device_tile_bits(s), &bp, jit_mode() == kJitAuto)) return rc;
    const int b = bp.b;
    const bool reg_tiles = bp.reg_tiles;

    const size_t ops_bytes = bp.ops.size() * sizeof(PassOp);
    const size_t rops_off_bytes = (ops_bytes + 255) & ~size_t(255);
    const size_t rops_bytes = bp.rops.size() * sizeof(RegOp);
    const size_t mat_off_bytes = (rops_off_bytes + rops_bytes + 255) & ~size_t(255);
    const size_t total = mat_off_bytes + bp.mat_v2d * sizeof(v2d);
    if (bp.ops.size() || bp.rops.size() || bp.mat_v2d) {
        if (int rc = ensure_capacity(s, total)) return rc;
        std::memcpy(s->program->host, bp.ops.data(), ops_bytes);
        std::memcpy(s->program->host + rops_off_bytes, bp.rops.data(), rops_bytes);
        v2d* hmats = reinterpret_cast<v2d*>(s->program->host + mat_off_bytes);
        size_t mi = 0;
        for (const llsv_op* op : bp.dense_src) {
            const size_t cnt = size_t(1) << (2 * op->nq);
            std::memcpy(hmats + mi, op->dense, cnt * sizeof(v2d));
            mi += cnt;
        }
        hipError_t e = hipMemcpyAsync(s->program->dev, s->program->host, total, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipEventRecord(s->program->uploaded, s->stream);
        if (e != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_apply_program: upload: %s", hipGetErrorString(e));
    }
    const PassOp* dops = reinterpret_cast<const PassOp*>(s->program->dev);
    const RegOp* drops = reinterpret_cast<const RegOp*>(s->program->dev + rops_off_bytes);
    const v2d* dmats = reinterpret_cast<const v2d*>(s->program->dev + mat_off_bytes);
    const size_t lds = size_t(16) << b;

    // compiled pass kernels: look up (and request) every eligible pass first, so the uncompiled
    // ones go to hipRTC as one batch
    std::vector<JitEntry*> jit(bp.launches.size(), nullptr);
    if (reg_tiles) {
        std::vector<JitEntry*> want;
        for (size_t i = 0; i < bp.launches.size(); ++i) {
            const Launch& la = bp.launches[i];
            if (la.streamed || la.mfma || !jit_wanted(n, la.d.nops, la.dense)) continue;
            jit[i] = jit_program_entry(bp.rops.data() + la.first, la.d.nops);
            if (!jit_ready(jit[i])) want.push_back(jit[i]);
        }
        if (!want.empty()) jit_request(want, jit_mode() == kJitAlways);
    }
    for (size_t li = 0; li < bp.launches.size(); ++li) {
        const Launch& la = bp.launches[li];
        hipError_t e = hipSuccess;
        if (la.mfma) {
            e = launch_dense_mfma(s->stream, s->amps, la.d, la.dt, dmats + la.moff);
        } else if (la.streamed) {
            const llsv_op& op = ops[la.op];
            if (op.kind == LLSV_OP_1Q) {
                Gate2 g;
                std::memcpy(g.r, op.m, sizeof g.r);
                e = launch_apply_1q(s->stream, s->amps, n, op.q[0], g);
            } else if (op.kind == LLSV_OP_CNOT) {
                e = launch_cnot(s->stream, s->amps, n, op.q[0], op.q[1]);
            } else {
                BitList bl{};
                bl.m = op.kind == LLSV_OP_CZ ? 2 : op.nq;
                for (int t = 0; t < bl.m; ++t) {
                    bl.pos[t] = op.q[t];
                    bl.mask |= uint64_t(1) << op.q[t];
                }
                std::sort(bl.pos, bl.pos + bl.m);
                e = launch_mcz(s->stream, s->amps, n, bl);
            }
        } else {
            const uint64_t tiles = uint64_t(1) << (n - b);
            const dim3 grid{unsigned(tiles), 1, 1}, block{unsigned(kPassThreads), 1, 1};
            const bool single = !reg_tiles && n == b;
            if (jit_ready(jit[li])) {
                // a pass whose kernel is not ready (compiling, or failed to compile) is interpreted
                e = jit_launch(jit[li], unsigned(tiles), s->stream, s->amps, la.d, drops + la.first);
            } else if (reg_tiles) {
                if (la.dense)
                    hipLaunchKernelGGL((k_pass_reg<true>), grid, block, lds, s->stream, s->amps, la.d, drops + la.first, dmats);
                else
                    hipLaunchKernelGGL((k_pass_reg<false>), grid, block, lds, s->stream, s->amps, la.d, drops + la.first, dmats);
                e = hipGetLastError();
            } else if (single) {
                const dim3 sblock{unsigned(kSingleThreads), 1, 1};
                if (la.dense)
                    hipLaunchKernelGGL((k_pass<kSingleThreads, true>), grid, sblock, lds, s->stream, s->amps, la.d,
                                       dops + la.first, dmats);
                else
                    hipLaunchKernelGGL((k_pass<kSingleThreads, false>), grid, sblock, lds, s->stream, s->amps, la.d,
                                       dops + la.first, dmats);
                e = hipGetLastError();
            } else if (la.dense) {
                hipLaunchKernelGGL((k_pass<kPassThreads, true>), grid, block, lds, s->stream, s->amps, la.d,
                                   dops + la.first, dmats);
                e = hipGetLastError();
            } else {
                hipLaunchKernelGGL((k_pass<kPassThreads, false>), grid, block, lds, s->stream, s->amps, la.d,
                                   dops + la.first, dmats);
                e = hipGetLastError();
            }
        }
        if (e != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_apply_program: launch: %s", hipGetErrorString(e));
    }
    return LLSV_OK;
}

// Compiles the pass kernels a program would use on an n-qubit state with 12-bit register tiles
// (host only); returns how many passes have a compiled kernel.
int jit_precompile(int n, const llsv_op* ops, uint64_t n_ops) {
    BuiltProgram bp;
    if (int rc = build_program(n, ops, n_ops, kRegTileBits, &bp)) return rc;
    if (!bp.reg_tiles) return 0;
    std::vector<JitEntry*> want;
    for (const Launch& la : bp.launches) {
        if (la.streamed || la.mfma || !jit_wanted(n, la.d.nops, la.dense)) continue;
        want.push_back(jit_program_entry(bp.rops.data() + la.first, la.d.nops));
    }
    jit_request(want, true);
    int count = 0;
    for (JitEntry* e : want) count += jit_ready(e) ? 1 : 0;
    return count;
}

}  // namespace llsv

extern "C" int llsv_set_program_jit(int mode) {
    if (mode < llsv::kJitOff || mode > llsv::kJitAlways)
        return llsv::fail(LLSV_ERR_ARG, "llsv_set_program_jit: mode must be 0, 1 or 2");
    return llsv::set_jit_mode(mode);
}

extern "C" void llsv_program_jit_stats(uint64_t* compiled, uint64_t* launches, uint64_t* failures) {
    llsv::jit_stats(compiled, launches, failures);
}

extern "C" int llsv_program_jit_precompile(int n_qubits, const llsv_op* ops, uint64_t n_ops) {
    using namespace llsv;
    if (n_qubits < 1 || n_qubits > LLSV_MAX_QUBITS) return fail(LLSV_ERR_ARG, "llsv_program_jit_precompile: bad qubit count");
    if (n_ops && !ops) return fail(LLSV_ERR_ARG, "llsv_program_jit_precompile: ops is null");
    return jit_precompile(n_qubits, ops, n_ops);
}

extern "C" void llsv_program_jit_wait(void) { llsv::jit_wait(); }

extern "C" int llsv_program_tile_bits(llsv_state* s) {
    if (!s) return llsv::fail(LLSV_ERR_HANDLE, "llsv_program_tile_bits: null state handle");
    return llsv::device_tile_bits(s);
}

extern "C" int llsv_plan_passes(int n_qubits, const llsv_op* ops, uint64_t n_ops, int tile_bits, uint64_t* n_passes,
                                uint64_t* order, uint64_t* pass_end) {
    using namespace llsv;
    if (!n_passes) return fail(LLSV_ERR_ARG, "llsv_plan_passes: n_passes is null");
    if (n_qubits < 1 || n_qubits > LLSV_MAX_QUBITS) return fail(LLSV_ERR_ARG, "llsv_plan_passes: bad qubit count");
    if (tile_bits < 1 || tile_bits > kMaxTileBits)
        return fail(LLSV_ERR_ARG, "llsv_plan_passes: tile_bits must be in [1, %d]", kMaxTileBits);
    if (n_ops && !ops) return fail(LLSV_ERR_ARG, "llsv_plan_passes: ops is null");
    const int b = std::min(n_qubits, tile_bits);
    std::vector<Checked> info(n_ops);
    for (uint64_t i = 0; i < n_ops; ++i)
        if (int rc = check_op(n_qubits, ops[i], i, &info[i])) return rc;
    // the schedule llsv_apply_program makes: register tiles when the tile has 12 bits
    const std::shared_ptr<const PlanEntry> pe = cached_plan(n_qubits, b, b == kRegTileBits, ops, info);
    uint64_t k = 0;
    for (size_t p = 0; p < pe->passes.size(); ++p) {
        for (uint64_t i : pe->passes[p].ops) {
            if (order) order[k] = i;
            ++k;
        }
        if (pass_end) pass_end[p] = k;
    }
    *n_passes = uint64_t(pe->passes.size());
    return LLSV_OK;
}
