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
#include <cstring>
#include <vector>

#include "sv_dense.h"
#include "sv_kernels.h"
#include "sv_program.h"
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

// DENSE selects the variant that can run dense blocks (it holds a column of outputs in
// registers, so the plain variant stays lean for passes without one).
template <int NT, bool DENSE>
__global__ __launch_bounds__(NT) void k_pass(v2d* __restrict__ psi, PassDesc d, const PassOp* __restrict__ ops,
                                             const v2d* __restrict__ mats) {
    extern __shared__ v2d tile[];
    const int b = d.b;
    const uint32_t tsize = 1u << b;
    const uint32_t tid = threadIdx.x;

    uint64_t base = blockIdx.x;
    for (int j = 0; j < b; ++j) base = insert0(base, d.pos[j]);

    // l = tid + k * NT: the offset splits into a per-thread and a per-iteration part.
    const uint64_t off_t = tile_offset(d, tid < tsize ? tid : 0);
    for (uint32_t k = 0; k * NT < tsize; ++k) {
        const uint32_t l = tid + k * NT;
        if (l < tsize) tile[l] = ld_nt(psi + (base | off_t | tile_offset(d, k * NT)));
    }
    __syncthreads();

    double sr = 1.0, si = 0.0;  // accumulated per-tile scalar (phases of qubits outside T)
    for (int i = 0; i < d.nops; ++i) {
        const PassOp& op = ops[i];
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
            const int k = op.a, D = 1 << k;
            uint32_t tmask = 0;
            for (int j = 0; j < k; ++j) tmask |= 1u << op.q[j];
            const v2d* U = mats + op.moff;
            constexpr int kMaxPer = (1 << kMaxTileBits) / NT;
            v2d out[kMaxPer];
#pragma unroll
            for (int r = 0; r < kMaxPer; ++r) {
                const uint32_t o = tid + r * NT;
                if (o < tsize) {
                    uint32_t row = 0;
                    for (int j = 0; j < k; ++j) row |= ((o >> op.q[j]) & 1u) << j;
                    const uint32_t gb = o & ~tmask;
                    const v2d* ur = U + row * D;
                    v2d acc = {0.0, 0.0};
                    for (int c = 0; c < D; ++c) {
                        const v2d u = ur[c], x = tile[gb | op.dep[c]];
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

// ---------------------------------------------------------------- register-resident tiles
// The tile (2^12 amplitudes) lives in registers, 16 per thread: 4 "register bits" S of the
// local index select the register, the other 8 local bits are the thread id (ascending bit
// order both ways).  A one-qubit gate or a CNOT whose target is in S is a pure register
// update; diagonal ops (phases, CZ, MCZ) need no data movement in any layout, since each
// thread knows the index of every amplitude it holds.  LDS is used only to change S (write
// all, barrier, read all) and to feed dense blocks.  LDS addresses are XOR-swizzled
// (bits 0..3 ^= bits 4..7) so that layouts whose thread bits start above bit 3 do not
// serialise on one bank.
__device__ __forceinline__ uint32_t swz(uint32_t l) { return l ^ ((l >> 4) & 15u); }

__device__ __forceinline__ uint32_t reg_dep(int j, const uint32_t* sb) {
    return ((j & 1) ? sb[0] : 0u) | ((j & 2) ? sb[1] : 0u) | ((j & 4) ? sb[2] : 0u) | ((j & 8) ? sb[3] : 0u);
}

__device__ __forceinline__ uint32_t thread_base(uint32_t tid, uint32_t smask) {
    uint32_t out = 0;
    int k = 0;
#pragma unroll
    for (int j = 0; j < kRegTileBits; ++j)
        if (!((smask >> j) & 1u)) {
            out |= ((tid >> k) & 1u) << j;
            ++k;
        }
    return out;
}

__device__ __forceinline__ uint64_t local_to_global(const PassDesc& d, int bit) {
    return bit < d.L ? (uint64_t(1) << bit) : (uint64_t(1) << d.tq[bit]);
}

template <int K>
__device__ __forceinline__ void reg_1q(v2d* a, const double* m) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (!(j & (1 << K))) {
            const v2d x = a[j], y = a[j | (1 << K)];
            a[j] = row2(m, x, y);
            a[j | (1 << K)] = row2(m + 4, x, y);
        }
}

template <int K>
__device__ __forceinline__ void reg_cnot(v2d* a, uint32_t lt, const uint32_t* sb, uint32_t cbit) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (!(j & (1 << K))) {
            if ((lt | reg_dep(j, sb)) & cbit) {
                const v2d x = a[j];
                a[j] = a[j | (1 << K)];
                a[j | (1 << K)] = x;
            }
        }
}

template <bool DENSE>
__global__ __launch_bounds__(256) void k_pass_reg(v2d* __restrict__ psi, PassDesc d, const PassOp* __restrict__ ops,
                                                  const v2d* __restrict__ mats) {
    extern __shared__ v2d tile[];
    const uint32_t tid = threadIdx.x;
    uint64_t base = blockIdx.x;
    for (int j = 0; j < kRegTileBits; ++j) base = insert0(base, d.pos[j]);

    // initial layout: S = local bits 8..11, thread bits 0..7 (coalesced runs of 2^L)
    uint32_t sb[4] = {1u << 8, 1u << 9, 1u << 10, 1u << 11};
    uint32_t lt = tid;
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

    double sr = 1.0, si = 0.0;
    for (int i = 0; i < d.nops; ++i) {
        const PassOp& op = ops[i];
        const int kind = op.kind;
        if (kind == R_1Q) {
            switch (op.a) {
                case 0: reg_1q<0>(a, op.m); break;
                case 1: reg_1q<1>(a, op.m); break;
                case 2: reg_1q<2>(a, op.m); break;
                default: reg_1q<3>(a, op.m); break;
            }
        } else if (kind == R_CNOT) {
            const uint32_t cbit = 1u << op.a;
            switch (op.b) {
                case 0: reg_cnot<0>(a, lt, sb, cbit); break;
                case 1: reg_cnot<1>(a, lt, sb, cbit); break;
                case 2: reg_cnot<2>(a, lt, sb, cbit); break;
                default: reg_cnot<3>(a, lt, sb, cbit); break;
            }
        } else if (kind == P_DIAG) {
            const uint32_t bit = 1u << op.a;
#pragma unroll
            for (int j = 0; j < 16; ++j)
                a[j] = ((lt | reg_dep(j, sb)) & bit) ? cmul(op.m[6], op.m[7], a[j]) : cmul(op.m[0], op.m[1], a[j]);
        } else if (kind == P_MCZ) {
            if ((base & op.omask) == op.omask) {
                const uint32_t lm = uint32_t(op.lmask);
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    if (((lt | reg_dep(j, sb)) & lm) == lm) a[j] = -a[j];
            }
        } else if (kind == P_SCALE) {
            const bool hi = (base >> op.a) & 1;
            const double dr = hi ? op.m[6] : op.m[0], di = hi ? op.m[7] : op.m[1];
            const double nr = sr * dr - si * di;
            si = sr * di + si * dr;
            sr = nr;
        } else if (kind == P_LAYOUT) {
#pragma unroll
            for (int j = 0; j < 16; ++j) tile[swz(lt | reg_dep(j, sb))] = a[j];
            __syncthreads();
            uint32_t smask = 0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                sb[t] = 1u << op.q[t];
                smask |= sb[t];
            }
            lt = thread_base(tid, smask);
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = tile[swz(lt | reg_dep(j, sb))];
            __syncthreads();
        } else if (DENSE && kind == P_DENSE) {
#pragma unroll
            for (int j = 0; j < 16; ++j) tile[swz(lt | reg_dep(j, sb))] = a[j];
            __syncthreads();
            const int k = op.a, D = 1 << k;
            uint32_t tmask = 0;
            for (int t = 0; t < k; ++t) tmask |= 1u << op.q[t];
            const v2d* U = mats + op.moff;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const uint32_t o = lt | reg_dep(j, sb);
                uint32_t row = 0;
                for (int t = 0; t < k; ++t) row |= ((o >> op.q[t]) & 1u) << t;
                const uint32_t gb = o & ~tmask;
                const v2d* ur = U + row * D;
                v2d acc = {0.0, 0.0};
                for (int c = 0; c < D; ++c) {
                    const v2d u = ur[c], x = tile[swz(gb | op.dep[c])];
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
    for (int t = 0; t < 4; ++t) gs[t] = local_to_global(d, 31 - __builtin_clz(sb[t]));
    const bool unit = (sr == 1.0 && si == 0.0);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint64_t go = ((j & 1) ? gs[0] : 0) | ((j & 2) ? gs[1] : 0) | ((j & 4) ? gs[2] : 0) | ((j & 8) ? gs[3] : 0);
        st_nt(psi + (base | gt | go), unit ? a[j] : cmul(sr, si, a[j]));
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

std::vector<PlannedPass> plan(int n, const llsv_op* ops, const std::vector<Checked>& info, int b, int L) {
    const uint64_t n_ops = info.size();
    const uint64_t all = n >= 64 ? ~uint64_t(0) : ((uint64_t(1) << n) - 1);
    std::vector<PlannedPass> passes;
    std::vector<char> done(n_ops, 0);
    uint64_t first = 0;
    if (b > n) b = n;
    if (L > b) L = b;
    while (first < n_ops) {
        PlannedPass p;
        uint64_t T = (b >= n) ? all : ((uint64_t(1) << L) - 1);
        uint64_t blocked = 0;
        for (uint64_t i = first; i < n_ops && blocked != all; ++i) {
            if (done[i]) continue;
            const Checked& c = info[i];
            if (c.mask & blocked) {
                blocked |= c.mask;
                continue;
            }
            if (!c.diag) {
                const uint64_t need = c.mask & ~T;
                if (popc(T) + popc(need) > b) {
                    blocked |= c.mask;
                    continue;
                }
                T |= need;
            }
            p.ops.push_back(i);
            done[i] = 1;
        }
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
    if (hipMalloc(reinterpret_cast<void**>(&pc.dev), cap) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&pc.host), cap) != hipSuccess) {
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

// Turns a pass's ops (tile-local coordinates) into the register-resident program: a layout
// change is inserted before every one-qubit gate or CNOT whose target bit is not held in
// registers; the new register set takes that bit, then the target bits of the next ops (so
// one change serves several gates), then bits of the current set.  The pass ends in a layout
// whose thread bits include local bits 0..L-1, so the stores write whole lines.
std::vector<PassOp> to_register_program(const std::vector<PassOp>& in, int L) {
    constexpr uint32_t kStart = 0xF00u;  // local bits 8..11
    const uint32_t low = (1u << L) - 1;
    auto need_of = [](const PassOp& po) -> uint32_t {
        if (po.kind == P_1Q) return 1u << po.a;
        if (po.kind == P_CNOT) return 1u << po.b;
        return 0u;
    };
    auto rank = [](uint32_t set, int bit) { return __builtin_popcount(set & ((1u << bit) - 1)); };
    auto layout = [](uint32_t set) {
        PassOp lay{};
        lay.kind = P_LAYOUT;
        int t = 0;
        for (int bit = 0; bit < kRegTileBits; ++bit)
            if ((set >> bit) & 1u) lay.q[t++] = bit;
        return lay;
    };
    std::vector<PassOp> out;
    out.reserve(in.size() + 8);
    uint32_t S = kStart;
    for (size_t i = 0; i < in.size(); ++i) {
        const uint32_t need = need_of(in[i]);
        if (need & ~S) {
            uint32_t ns = need;
            for (size_t k = i + 1; k < in.size() && __builtin_popcount(ns) < kRegBits; ++k) {
                const uint32_t nk = need_of(in[k]);
                if (__builtin_popcount(ns | nk) <= kRegBits) ns |= nk;
            }
            for (int bit = kRegTileBits - 1; bit >= 0 && __builtin_popcount(ns) < kRegBits; --bit)
                if ((S >> bit) & 1u) ns |= 1u << bit;
            for (int bit = kRegTileBits - 1; bit >= 0 && __builtin_popcount(ns) < kRegBits; --bit) ns |= 1u << bit;
            out.push_back(layout(ns));
            S = ns;
        }
        PassOp po = in[i];
        if (po.kind == P_1Q) {
            po.kind = R_1Q;
            po.a = rank(S, in[i].a);
        } else if (po.kind == P_CNOT) {
            po.kind = R_CNOT;
            po.a = in[i].a;
            po.b = rank(S, in[i].b);
        }
        out.push_back(po);
    }
    if (S & low) out.push_back(layout(kStart));
    return out;
}

}  // namespace

int run_program(llsv_state* s, const llsv_op* ops, uint64_t n_ops) {
    const int n = s->n;
    std::vector<Checked> info(n_ops);
    for (uint64_t i = 0; i < n_ops; ++i)
        if (int rc = check_op(n, ops[i], i, &info[i])) return rc;

    const int b = std::min(n, device_tile_bits(s));
    const int L = std::min(kLowBits, b);
    const bool reg_tiles = (b == kRegTileBits);
    const std::vector<PlannedPass> passes = plan(n, ops, info, b, L);

    struct Launch {
        bool streamed;
        bool dense;
        bool mfma;  // a lone dense block of 4-5 qubits: the FP64-MFMA kernel
        uint64_t op;
        PassDesc d;
        DenseTile dt;
        size_t first;
        size_t moff;
    };
    std::vector<Launch> launches;
    launches.reserve(passes.size());
    std::vector<PassOp> all_ops;
    std::vector<const llsv_op*> dense_src;  // dense ops in upload order
    size_t mat_v2d = 0;
    for (const PlannedPass& p : passes) {
        Launch la{};
        if (p.ops.size() == 1 && info[p.ops[0]].single && n > b) {
            la.streamed = true;
            la.op = p.ops[0];
            launches.push_back(la);
            continue;
        }
        if (p.ops.size() == 1 && ops[p.ops[0]].kind == LLSV_OP_DENSE &&
            make_dense_tile(n, ops[p.ops[0]].nq, ops[p.ops[0]].q, &la.d, &la.dt)) {
            la.mfma = true;
            la.moff = mat_v2d;
            dense_src.push_back(&ops[p.ops[0]]);
            mat_v2d += size_t(1) << (2 * ops[p.ops[0]].nq);
            launches.push_back(la);
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
                    po.moff = int(mat_v2d);
                    for (int c = 0; c < (1 << op.nq); ++c) {
                        uint32_t o = 0;
                        for (int t = 0; t < op.nq; ++t)
                            if ((c >> t) & 1) o |= 1u << po.q[t];
                        po.dep[c] = o;
                    }
                    dense_src.push_back(&op);
                    mat_v2d += size_t(1) << (2 * op.nq);
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
        if (reg_tiles) pops = to_register_program(pops, L);
        d.nops = int(pops.size());
        la.first = all_ops.size();
        all_ops.insert(all_ops.end(), pops.begin(), pops.end());
        launches.push_back(la);
    }

    const size_t ops_bytes = all_ops.size() * sizeof(PassOp);
    const size_t mat_off_bytes = (ops_bytes + 255) & ~size_t(255);
    const size_t total = mat_off_bytes + mat_v2d * sizeof(v2d);
    if (all_ops.size() || mat_v2d) {
        if (int rc = ensure_capacity(s, total)) return rc;
        std::memcpy(s->program->host, all_ops.data(), ops_bytes);
        v2d* hmats = reinterpret_cast<v2d*>(s->program->host + mat_off_bytes);
        size_t mi = 0;
        for (const llsv_op* op : dense_src) {
            const size_t cnt = size_t(1) << (2 * op->nq);
            std::memcpy(hmats + mi, op->dense, cnt * sizeof(v2d));
            mi += cnt;
        }
        hipError_t e = hipMemcpyAsync(s->program->dev, s->program->host, total, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipEventRecord(s->program->uploaded, s->stream);
        if (e != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_apply_program: upload: %s", hipGetErrorString(e));
    }
    const PassOp* dops = reinterpret_cast<const PassOp*>(s->program->dev);
    const v2d* dmats = reinterpret_cast<const v2d*>(s->program->dev + mat_off_bytes);
    const size_t lds = size_t(16) << b;
    for (const Launch& la : launches) {
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
            if (reg_tiles) {
                if (la.dense)
                    hipLaunchKernelGGL((k_pass_reg<true>), grid, block, lds, s->stream, s->amps, la.d, dops + la.first, dmats);
                else
                    hipLaunchKernelGGL((k_pass_reg<false>), grid, block, lds, s->stream, s->amps, la.d, dops + la.first, dmats);
            } else if (la.dense) {
                hipLaunchKernelGGL((k_pass<kPassThreads, true>), grid, block, lds, s->stream, s->amps, la.d,
                                   dops + la.first, dmats);
            } else {
                hipLaunchKernelGGL((k_pass<kPassThreads, false>), grid, block, lds, s->stream, s->amps, la.d,
                                   dops + la.first, dmats);
            }
            e = hipGetLastError();
        }
        if (e != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_apply_program: launch: %s", hipGetErrorString(e));
    }
    return LLSV_OK;
}

}  // namespace llsv

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
    const int c = plan_order(n_qubits, ops, n_ops, b, std::min(kLowBits, b), order, pass_end);
    if (c < 0) return c;
    *n_passes = uint64_t(c);
    return LLSV_OK;
}
