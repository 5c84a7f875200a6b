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

int run_program(llsv_state* s, const llsv_op* ops, uint64_t n_ops) {
    const int n = s->n;
    std::vector<Checked> info(n_ops);
    for (uint64_t i = 0; i < n_ops; ++i)
        if (int rc = check_op(n, ops[i], i, &info[i])) return rc;

    const int b = std::min(n, device_tile_bits(s));
    const int L = std::min(kLowBits, b);
    const std::vector<PlannedPass> passes = plan(n, ops, info, b, L);

    // Pack every tiled pass's ops and the dense matrices into one upload.
    size_t op_count = 0, mat_v2d = 0;
    for (const PlannedPass& p : passes) {
        const bool streamed = p.ops.size() == 1 && info[p.ops[0]].single && n > b;
        if (streamed) continue;
        op_count += p.ops.size();
        for (uint64_t i : p.ops)
            if (ops[i].kind == LLSV_OP_DENSE) mat_v2d += size_t(1) << (2 * ops[i].nq);
    }
    const size_t ops_bytes = op_count * sizeof(PassOp);
    const size_t mat_off_bytes = (ops_bytes + 255) & ~size_t(255);
    const size_t total = mat_off_bytes + mat_v2d * sizeof(v2d);
    if (total) {
        if (int rc = ensure_capacity(s, total)) return rc;
    }
    PassOp* hops = reinterpret_cast<PassOp*>(s->program->host);
    v2d* hmats = total ? reinterpret_cast<v2d*>(s->program->host + mat_off_bytes) : nullptr;

    struct Launch {
        bool streamed;
        bool dense;
        uint64_t op;
        PassDesc d;
        size_t first;
    };
    std::vector<Launch> launches;
    launches.reserve(passes.size());
    size_t oi = 0, mi = 0;
    for (const PlannedPass& p : passes) {
        Launch la{};
        if (p.ops.size() == 1 && info[p.ops[0]].single && n > b) {
            la.streamed = true;
            la.op = p.ops[0];
            launches.push_back(la);
            continue;
        }
        PassDesc& d = la.d;
        d.n = n;
        d.b = b;
        d.L = L;
        d.nops = int(p.ops.size());
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
        la.first = oi;
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
                    po.moff = int(mi);
                    for (int c = 0; c < (1 << op.nq); ++c) {
                        uint32_t o = 0;
                        for (int t = 0; t < op.nq; ++t)
                            if ((c >> t) & 1) o |= 1u << po.q[t];
                        po.dep[c] = o;
                    }
                    const size_t cnt = size_t(1) << (2 * op.nq);
                    std::memcpy(hmats + mi, op.dense, cnt * sizeof(v2d));
                    mi += cnt;
                    break;
                }
            }
            hops[oi++] = po;
        }
        launches.push_back(la);
    }

    if (total) {
        hipError_t e = hipMemcpyAsync(s->program->dev, s->program->host, total, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipEventRecord(s->program->uploaded, s->stream);
        if (e != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_apply_program: upload: %s", hipGetErrorString(e));
    }
    const PassOp* dops = reinterpret_cast<const PassOp*>(s->program->dev);
    const v2d* dmats = reinterpret_cast<const v2d*>(s->program->dev + mat_off_bytes);
    const size_t lds = size_t(16) << b;
    for (const Launch& la : launches) {
        hipError_t e = hipSuccess;
        if (la.streamed) {
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
            if (la.dense)
                hipLaunchKernelGGL((k_pass<kPassThreads, true>), dim3(unsigned(tiles)), dim3(kPassThreads), lds,
                                   s->stream, s->amps, la.d, dops + la.first, dmats);
            else
                hipLaunchKernelGGL((k_pass<kPassThreads, false>), dim3(unsigned(tiles)), dim3(kPassThreads), lds,
                                   s->stream, s->amps, la.d, dops + la.first, dmats);
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
