// sv_adjoint.hip -- batched parameterised circuits: energies and adjoint gradients.
//
// A batch of B parameter vectors shares one circuit.  Every kernel handles one gate for all
// B states at once (blockIdx.y = batch index, states B x 2^n contiguous in the workspace),
// so a circuit of G gates is G launches regardless of B.  Rotation angles are read from the
// parameter matrix on the device and turned into (cos, sin) in the kernel, so the same
// launches serve any parameter values (torch tensors, shifted copies for parameter shift).
//
// Adjoint sweep (for U_k = exp(-i theta_k G_k / 2), G in {X, Y, Z}):
//   psi = U_G ... U_1 |0>;  lambda = H psi;  E = Re <psi|lambda>
//   for k = G..1:  g_k = Im <lambda|G_k|psi>;  psi <- U_k^dag psi;  lambda <- U_k^dag lambda
// The inner product of step k and both inverse gates are one pass over (psi, lambda).
// Reductions use a fixed grid and fixed trees: gradients are bitwise reproducible.
//
// Tiled sweeps (n >= 12, when the pass compiler is on): the forward and backward sweeps are
// grouped into register-resident passes of 12-qubit tiles (the program path's scheduler and
// layouts), each compiled once per circuit structure.  A forward pass applies its gates to all
// B states (grid.y = batch); a backward pass holds psi and lambda of a tile in registers and,
// for each gate, reduces Im <lambda|G|psi> over the tile (one slot per gate) before applying
// the inverse gate to both.  One-qubit coefficients (forward and inverse) come from a per-batch
// table computed on the device from the parameters.  A 20-qubit, 590-gate circuit then costs
// a few dozen passes per sweep instead of one per gate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/llsv.h"
#include "sv_common.h"
#include "sv_jit.h"
#include "sv_program.h"

namespace llsv {

namespace {

constexpr int kRedGrid = 128;  // blocks per state for kernels that reduce

struct PG {
    int kind, q0, q1, param;
    double m[8];
};
static_assert(sizeof(PG) == sizeof(llsv_pgate), "llsv_pgate layout");

struct PauliDev {
    double coeff;
    uint64_t x, z;
};

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double bsum(double v, double* lds4) {
    v = wsum(v);
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) s = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    __syncthreads();
    return s;
}

// The 2x2 of gate g for batch element b (or its adjoint).
__device__ __forceinline__ void gate_matrix(const PG& g, const double* __restrict__ params, int P, int b,
                                            bool dagger, double* m) {
    if (g.kind == LLSV_PG_1Q) {
        if (!dagger) {
            for (int i = 0; i < 8; ++i) m[i] = g.m[i];
        } else {  // conjugate transpose
            m[0] = g.m[0], m[1] = -g.m[1];
            m[2] = g.m[4], m[3] = -g.m[5];
            m[4] = g.m[2], m[5] = -g.m[3];
            m[6] = g.m[6], m[7] = -g.m[7];
        }
        return;
    }
    const double th = params[size_t(b) * P + g.param];
    double s, c;
    sincos(0.5 * th, &s, &c);
    if (dagger) s = -s;  // U(theta)^dagger = U(-theta) for these rotations
    if (g.kind == LLSV_PG_RX) {
        m[0] = c, m[1] = 0, m[2] = 0, m[3] = -s, m[4] = 0, m[5] = -s, m[6] = c, m[7] = 0;
    } else if (g.kind == LLSV_PG_RY) {
        m[0] = c, m[1] = 0, m[2] = -s, m[3] = 0, m[4] = s, m[5] = 0, m[6] = c, m[7] = 0;
    } else {  // RZ
        m[0] = c, m[1] = -s, m[2] = 0, m[3] = 0, m[4] = 0, m[5] = 0, m[6] = c, m[7] = s;
    }
}

__device__ __forceinline__ v2d r2(const double* r, v2d a0, v2d a1) {
    v2d o;
    o.x = r[0] * a0.x - r[1] * a0.y + r[2] * a1.x - r[3] * a1.y;
    o.y = r[0] * a0.y + r[1] * a0.x + r[2] * a1.y + r[3] * a1.x;
    return o;
}

__global__ __launch_bounds__(256) void k_pc_init(v2d* __restrict__ A, uint64_t size) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    v2d* s = A + uint64_t(blockIdx.y) * size;
    if (i < size) {
        v2d v;
        v.x = (i == 0) ? 1.0 : 0.0;
        v.y = 0.0;
        s[i] = v;
    }
}

// One-qubit gate on every state of the batch (A, and Lm when TWO).  With RED, first adds
// Im <lm|G psi> for the gate's generator G into partials[(b * nslot + slot) * gridDim.x + blk].
template <bool TWO, bool RED>
__global__ __launch_bounds__(256) void k_pc_1q(v2d* __restrict__ A, v2d* __restrict__ Lm, uint64_t size, PG g,
                                               const double* __restrict__ params, int P, int dagger,
                                               double* __restrict__ partials, int nslot, int slot) {
    __shared__ double lds4[4];
    const int b = blockIdx.y;
    const int q = g.q0;
    double m[8];
    gate_matrix(g, params, P, b, dagger != 0, m);
    v2d* psi = A + uint64_t(b) * size;
    v2d* lam = TWO ? Lm + uint64_t(b) * size : nullptr;
    const uint64_t npairs = size >> 1;
    const uint64_t lomask = (uint64_t(1) << q) - 1, stride = uint64_t(1) << q;
    double acc = 0.0;
    for (uint64_t p = uint64_t(blockIdx.x) * 256 + threadIdx.x; p < npairs; p += uint64_t(gridDim.x) * 256) {
        const uint64_t i0 = ((p & ~lomask) << 1) | (p & lomask), i1 = i0 + stride;
        const v2d a0 = psi[i0], a1 = psi[i1];
        psi[i0] = r2(m, a0, a1);
        psi[i1] = r2(m + 4, a0, a1);
        if (TWO) {
            const v2d l0 = lam[i0], l1 = lam[i1];
            if (RED) {
                // Im of conj(l0) (G a)_0 + conj(l1) (G a)_1
                if (g.kind == LLSV_PG_RZ) {         // G = Z: (a0, -a1)
                    acc += (l0.x * a0.y - l0.y * a0.x) - (l1.x * a1.y - l1.y * a1.x);
                } else if (g.kind == LLSV_PG_RX) {  // G = X: (a1, a0)
                    acc += (l0.x * a1.y - l0.y * a1.x) + (l1.x * a0.y - l1.y * a0.x);
                } else {                            // G = Y: (-i a1, i a0)
                    // Im(conj(l) * (-i a)) = -Re(conj(l) a); Im(conj(l) * (i a)) = Re(conj(l) a)
                    acc += -(l0.x * a1.x + l0.y * a1.y) + (l1.x * a0.x + l1.y * a0.y);
                }
            }
            lam[i0] = r2(m, l0, l1);
            lam[i1] = r2(m + 4, l0, l1);
        }
    }
    if (RED) {
        const double s = bsum(acc, lds4);
        if (threadIdx.x == 0) partials[(size_t(b) * nslot + slot) * gridDim.x + blockIdx.x] = s;
    }
}

// CNOT (kind 4) or CZ (kind 5) on every state of the batch (A, and Lm when TWO).
template <bool TWO>
__global__ __launch_bounds__(256) void k_pc_2q(v2d* __restrict__ A, v2d* __restrict__ Lm, uint64_t size, int kind,
                                               int q0, int q1) {
    const int b = blockIdx.y;
    const int lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
    const uint64_t b0 = uint64_t(1) << q0, b1 = uint64_t(1) << q1;
    const uint64_t quads = size >> 2;
    const uint64_t p = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= quads) return;
    const uint64_t i = insert0(insert0(p, lo), hi) | b0;  // control (or first qubit) set
    for (int arr = 0; arr < (TWO ? 2 : 1); ++arr) {
        v2d* s = (arr == 0 ? A : Lm) + uint64_t(b) * size;
        if (kind == LLSV_PG_CNOT) {
            const v2d x = s[i];
            s[i] = s[i | b1];
            s[i | b1] = x;
        } else {
            s[i | b1] = -s[i | b1];
        }
    }
}

// lam = H psi (when STORE) and partials of Re <psi|H psi> per batch element.
template <bool STORE>
__global__ __launch_bounds__(256) void k_pc_hpsi(const v2d* __restrict__ A, v2d* __restrict__ Lm, uint64_t size,
                                                 const PauliDev* __restrict__ terms, int nterms,
                                                 double* __restrict__ partials) {
    __shared__ double lds4[4];
    const int b = blockIdx.y;
    const v2d* psi = A + uint64_t(b) * size;
    double acc = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < size; i += uint64_t(gridDim.x) * 256) {
        double lr = 0.0, li = 0.0;
        for (int t = 0; t < nterms; ++t) {
            const PauliDev T = terms[t];
            const uint64_t src = i ^ T.x;
            const v2d a = psi[src];
            // phase = i^{nY} (-1)^{popcount(src & z)}, nY = popcount(x & z)
            const int ny = __popcll(T.x & T.z) & 3;
            const double sg = (__popcll(src & T.z) & 1) ? -T.coeff : T.coeff;
            double pr = a.x, pi = a.y;
            if (ny == 1) { const double t0 = pr; pr = -pi; pi = t0; }
            else if (ny == 2) { pr = -pr; pi = -pi; }
            else if (ny == 3) { const double t0 = pr; pr = pi; pi = -t0; }
            lr += sg * pr;
            li += sg * pi;
        }
        const v2d p = psi[i];
        acc += p.x * lr + p.y * li;
        if (STORE) {
            v2d l;
            l.x = lr;
            l.y = li;
            Lm[uint64_t(b) * size + i] = l;
        }
    }
    const double s = bsum(acc, lds4);
    if (threadIdx.x == 0) partials[size_t(b) * gridDim.x + blockIdx.x] = s;
}

// ---- H psi over 12-qubit tiles -----------------------------------------------------------
// The terms are grouped into passes whose X masks fit one tile (qubits 0..2 plus at most nine
// more); a workgroup loads its tile of psi into LDS once and forms every term of the pass from
// there, so psi is read once per pass instead of once per term.  Pass contributions add:
// lambda += H_pass psi (stored in the first pass, accumulated after), and
// partials[b][pass][tile] = Re <psi|H_pass psi> over the tile.
struct HTile {
    int tq[12];  // tile qubits, ascending: local bit k <-> qubit tq[k] (tq[0..2] = 0, 1, 2)
};
struct HTerm {
    double cr, ci;  // c i^ny, ny = popcount(x & z) mod 4 (the Y factors)
    uint64_t x, z;
    uint32_t xl;  // x as local bits
    uint32_t pad;
};

template <bool STORE>
__global__ __launch_bounds__(256) void k_pc_hpsi_tile(const v2d* __restrict__ A, v2d* __restrict__ Lm, uint64_t size,
                                                      HTile tl, const HTerm* __restrict__ terms, int nterms, int first,
                                                      double* __restrict__ partials, int pass, int npass) {
    __shared__ v2d tile[4096];
    __shared__ double lds4[4];
    const int b = blockIdx.y;
    const uint32_t tid = threadIdx.x;
    const v2d* psi = A + uint64_t(b) * size;
    v2d* lam = STORE ? Lm + uint64_t(b) * size : nullptr;
    const uint64_t tix = xcd_tile(blockIdx.x, gridDim.x);  // this workgroup's tile
    uint64_t base = tix;
    for (int j = 0; j < 12; ++j) base = insert0(base, tl.tq[j]);
    uint32_t glane = 0, greg[16];  // states here have at most 30 qubits: 32-bit indices
    for (int k = 0; k < 8; ++k)
        if ((tid >> k) & 1u) glane |= 1u << tl.tq[k];
    v2d own[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        uint32_t gr = uint32_t(base) | glane;
        for (int k = 0; k < 4; ++k)
            if ((r >> k) & 1) gr |= 1u << tl.tq[8 + k];
        greg[r] = gr;
        own[r] = psi[gr];
        tile[tid + 256u * uint32_t(r)] = own[r];
    }
    __syncthreads();
    double lr[16], li[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) lr[r] = li[r] = 0.0;
    // one term at a time over the thread's 16 amplitudes, branch-free (no control flow merges the
    // accumulators): (P psi)_g = c i^ny (-1)^popc(src & z) psi_src, src = g ^ x
    for (int t = 0; t < nterms; ++t) {
        const HTerm T = terms[first + t];
        const uint32_t x = uint32_t(T.x), z = uint32_t(T.z), xl = T.xl;
        const double cr = T.cr, ci = T.ci;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const v2d a = tile[(tid + 256u * uint32_t(r)) ^ xl];
            const bool neg = __popc((greg[r] ^ x) & z) & 1;
            const double sr = neg ? -cr : cr, si = neg ? -ci : ci;
            lr[r] = __builtin_fma(sr, a.x, __builtin_fma(-si, a.y, lr[r]));
            li[r] = __builtin_fma(sr, a.y, __builtin_fma(si, a.x, li[r]));
        }
    }
    double e = 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        e += own[r].x * lr[r] + own[r].y * li[r];
        if (STORE) {
            v2d o;
            o.x = lr[r];
            o.y = li[r];
            if (pass > 0) o = o + lam[greg[r]];
            lam[greg[r]] = o;
        }
    }
    const double sum = bsum(e, lds4);
    if (tid == 0) partials[(size_t(b) * npass + pass) * gridDim.x + tix] = sum;
}

// Groups the terms into at most four tile passes (diagonal terms in the first); false when some
// X mask cannot be placed (the per-term gather kernel runs instead).
bool plan_hpsi(int n, const llsv_pauli* terms, int T, std::vector<HTile>& tiles, std::vector<int>& first,
               std::vector<int>& count, std::vector<HTerm>& out) {
    if (n < 12) return false;
    std::vector<char> done(size_t(T), 0);
    int left = T;
    while (left > 0) {
        if (tiles.size() == 4) return false;
        uint32_t bits = 0x7u;  // qubits 0..2
        const size_t start = out.size();
        for (int t = 0; t < T; ++t) {
            if (done[t]) continue;
            const uint64_t x = terms[t].xmask;
            if (x >> 32) return false;
            const uint32_t nb = bits | uint32_t(x);
            if (__builtin_popcount(nb) > 12) continue;
            bits = nb;
        }
        for (int q = 0; q < n && __builtin_popcount(bits) < 12; ++q) bits |= 1u << q;
        HTile tl{};
        int k = 0;
        for (int q = 0; q < n; ++q)
            if ((bits >> q) & 1u) tl.tq[k++] = q;
        for (int t = 0; t < T; ++t) {
            if (done[t] || (terms[t].xmask & ~uint64_t(bits))) continue;
            HTerm h{};
            h.x = terms[t].xmask;
            h.z = terms[t].zmask;
            for (int j = 0; j < 12; ++j)
                if ((h.x >> tl.tq[j]) & 1u) h.xl |= 1u << j;
            const int ny = __builtin_popcountll(h.x & h.z) & 3;
            const double c = terms[t].coeff;
            h.cr = ny == 0 ? c : ny == 2 ? -c : 0.0;  // c i^ny
            h.ci = ny == 1 ? c : ny == 3 ? -c : 0.0;
            out.push_back(h);
            done[t] = 1;
            --left;
        }
        if (out.size() == start) return false;
        tiles.push_back(tl);
        first.push_back(int(start));
        count.push_back(int(out.size() - start));
    }
    return true;
}

// out[b * nslot + s] = sum over nblk partials (fixed order); grid (nslot, B).
__global__ __launch_bounds__(256) void k_pc_reduce(const double* __restrict__ partials, int nblk, int nslot,
                                                   double* __restrict__ out) {
    __shared__ double lds4[4];
    const size_t row = size_t(blockIdx.y) * nslot + blockIdx.x;
    double acc = 0.0;
    for (int k = threadIdx.x; k < nblk; k += 256) acc += partials[row * nblk + k];
    const double s = bsum(acc, lds4);
    if (threadIdx.x == 0) out[row] = s;
}

// grads[b * P + param(s)] += slot_vals[b * nslot + s], slots in order; one thread per b.
__global__ void k_pc_scatter(const double* __restrict__ slot_vals, const int* __restrict__ slot_param, int nslot,
                             int P, int B, double* __restrict__ grads) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    for (int p = 0; p < P; ++p) grads[size_t(b) * P + p] = 0.0;
    for (int s = 0; s < nslot; ++s) grads[size_t(b) * P + slot_param[s]] += slot_vals[size_t(b) * nslot + s];
}

struct Layout {
    uint64_t size;
    size_t states, partials, slots, terms, slotmap, table, total;
    int nblk, nslot;
};

// Tiles of a tiled sweep (12-qubit register tiles), 0 when the state is smaller than a tile.
uint64_t tiled_tiles(int n) { return n >= kRegTileBits ? uint64_t(1) << (n - kRegTileBits) : 0; }

Layout layout_of(int n, int B, int n_gates, int nslot, int nterms) {
    Layout L{};
    L.size = uint64_t(1) << n;
    const uint64_t npairs = L.size >> 1;
    L.nblk = int(std::min<uint64_t>(kRedGrid, (npairs + 255) / 256));
    L.nslot = nslot;
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    L.states = al(L.size * 16 * size_t(B));
    const uint64_t per_slot = std::max<uint64_t>(kRedGrid, tiled_tiles(n));
    L.partials = al(sizeof(double) * size_t(B) * (size_t(n_gates) + 1) * per_slot);
    L.slots = al(sizeof(double) * size_t(B) * (size_t(n_gates) + 1));
    L.terms = al(std::max(sizeof(PauliDev), sizeof(HTerm)) * size_t(nterms > 0 ? nterms : 1));
    L.slotmap = al(sizeof(int) * size_t(n_gates + 1));
    L.table = al(sizeof(double) * 16 * size_t(B) * size_t(n_gates + 1));  // forward + inverse 2x2 per one-qubit gate
    L.total = 2 * L.states + L.partials + L.slots + L.terms + L.slotmap + L.table;
    return L;
}

int validate(int n, int B, const llsv_pgate* gates, int G, int P, const double* params, const llsv_pauli* terms,
             int T, const char* who) {
    if (n < 1 || n > 30) return fail(LLSV_ERR_ARG, "%s: n_qubits must be in [1, 30]", who);
    if (B < 1 || B > 65535) return fail(LLSV_ERR_ARG, "%s: batch must be in [1, 65535]", who);
    if (G < 0 || (G > 0 && !gates)) return fail(LLSV_ERR_ARG, "%s: bad gate list", who);
    if (P < 0 || (P > 0 && !params)) return fail(LLSV_ERR_ARG, "%s: bad parameter matrix", who);
    if (T < 1 || T > 4096 || !terms) return fail(LLSV_ERR_ARG, "%s: need 1..4096 Hamiltonian terms", who);
    const uint64_t all = (uint64_t(1) << n) - 1;
    for (int t = 0; t < T; ++t)
        if ((terms[t].xmask | terms[t].zmask) & ~all) return fail(LLSV_ERR_ARG, "%s: term %d acts outside the register", who, t);
    for (int k = 0; k < G; ++k) {
        const llsv_pgate& g = gates[k];
        if (g.kind < LLSV_PG_RX || g.kind > LLSV_PG_CZ) return fail(LLSV_ERR_ARG, "%s: gate %d: unknown kind", who, k);
        if (g.q0 < 0 || g.q0 >= n) return fail(LLSV_ERR_ARG, "%s: gate %d: qubit outside the register", who, k);
        if (g.kind >= LLSV_PG_CNOT && (g.q1 < 0 || g.q1 >= n || g.q1 == g.q0))
            return fail(LLSV_ERR_ARG, "%s: gate %d: bad second qubit", who, k);
        if (g.kind <= LLSV_PG_RZ && (g.param < 0 || g.param >= P))
            return fail(LLSV_ERR_ARG, "%s: gate %d: parameter index outside [0, %d)", who, k, P);
    }
    return LLSV_OK;
}

hipError_t apply_gate(hipStream_t st, v2d* A, v2d* Lm, bool two, const Layout& L, int B, const llsv_pgate& g,
                      const double* params, int P, bool dagger, double* partials, int slot) {
    const uint64_t npairs = L.size >> 1;
    if (g.kind >= LLSV_PG_CNOT) {
        const uint64_t quads = L.size >> 2;
        const dim3 grid(unsigned((quads + 255) / 256), unsigned(B));
        if (two)
            hipLaunchKernelGGL(k_pc_2q<true>, grid, dim3(256), 0, st, A, Lm, L.size, g.kind, g.q0, g.q1);
        else
            hipLaunchKernelGGL(k_pc_2q<false>, grid, dim3(256), 0, st, A, Lm, L.size, g.kind, g.q0, g.q1);
        return hipGetLastError();
    }
    PG pg;
    std::memcpy(&pg, &g, sizeof pg);
    const bool red = slot >= 0;
    const unsigned gx = red ? unsigned(L.nblk) : unsigned(std::min<uint64_t>((npairs + 255) / 256, 4096));
    const dim3 grid(gx, unsigned(B));
    if (!two)
        hipLaunchKernelGGL((k_pc_1q<false, false>), grid, dim3(256), 0, st, A, Lm, L.size, pg, params, P, int(dagger),
                           partials, L.nslot, 0);
    else if (!red)
        hipLaunchKernelGGL((k_pc_1q<true, false>), grid, dim3(256), 0, st, A, Lm, L.size, pg, params, P, int(dagger),
                           partials, L.nslot, 0);
    else
        hipLaunchKernelGGL((k_pc_1q<true, true>), grid, dim3(256), 0, st, A, Lm, L.size, pg, params, P, int(dagger),
                           partials, L.nslot, slot);
    return hipGetLastError();
}

// ---------------------------------------------------------------- tiled sweeps
// An RX / RY entry of the table holds, in place of its first two coefficients, the shears
// (a, b) that the compiled passes apply (r_shx / r_shy): the rotation by f = theta / 2 is -1
// times the rotation by f - pi, so f is folded into [-pi/2, pi/2] first, where
// a = -tan(f/2) = -sin f / (1 + cos f) and b = +-sin f are at most 1 in modulus.  The dropped
// sign multiplies a whole state; psi and lambda carry the same sign, so energies and gradients
// do not see it.
__device__ __forceinline__ void rotation_shears(const PG& g, double* m) {
    double c = m[0], s = g.kind == LLSV_PG_RX ? -m[5] : m[4];
    if (c < 0.0) {
        c = -c;
        s = -s;
    }
    m[0] = -s / (1.0 + c);
    m[1] = g.kind == LLSV_PG_RX ? -s : s;
}

// table[b * tstride + 8 * j] = forward 2x2 of one-qubit gate coef_gate[j] for batch element b,
// table[b * tstride + 8 * (ncoef + j)] = its inverse (RX / RY: with their shears).
__global__ __launch_bounds__(256) void k_pc_table(const PG* __restrict__ gates, const int* __restrict__ coef_gate,
                                                  int ncoef, const double* __restrict__ params, int P, int B,
                                                  double* __restrict__ table, uint64_t tstride) {
    const int j = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y;
    if (j >= ncoef || b >= B) return;
    const PG g = gates[coef_gate[j]];
    double* out = table + uint64_t(b) * tstride;
    gate_matrix(g, params, P, b, false, out + 8 * j);
    gate_matrix(g, params, P, b, true, out + 8 * (ncoef + j));
    if (g.kind == LLSV_PG_RX || g.kind == LLSV_PG_RY) {
        rotation_shears(g, out + 8 * j);
        rotation_shears(g, out + 8 * (ncoef + j));
    }
}

struct TiledPass {
    PassDesc d;
    JitEntry* e = nullptr;
};

struct TiledPlan {
    int ncoef = 0;
    std::vector<int> coef_gate;  // gate index of coefficient j
    std::vector<TiledPass> fwd, bwd;
    std::vector<int> slot_of;  // gradient slot of each rotation (backward-sweep order), else -1
};

// Schedule of the gate list as register-tile passes: ops in execution order; a one-qubit gate
// becomes P_1QT on its coefficient (the inverse one in the backward sweep, preceded there by
// P_GRAD for a rotation), CNOT and CZ as in the program path.
// order / pass_end: the schedule (planned here when empty; the backward sweep gets the forward
// schedule reversed, which is valid for the reversed gate list and has as few passes).
std::vector<TiledPass> tiled_passes(int n, const std::vector<const llsv_pgate*>& seq, const std::vector<int>& coef_of,
                                    std::vector<int>* slot_of, const std::vector<int>& gate_index, int ncoef,
                                    bool backward, std::vector<uint64_t>* order_io, std::vector<uint64_t>* end_io) {
    const int G = int(seq.size());
    std::vector<llsv_op> ops(G);
    for (int k = 0; k < G; ++k) {
        const llsv_pgate& g = *seq[k];
        llsv_op& o = ops[k];
        o = llsv_op{};
        if (g.kind >= LLSV_PG_CNOT) {
            o.kind = g.kind == LLSV_PG_CNOT ? LLSV_OP_CNOT : LLSV_OP_CZ;
            o.nq = 2;
            o.q[0] = g.q0;
            o.q[1] = g.q1;
        } else {
            o.kind = LLSV_OP_1Q;  // scheduled as a dense 2x2 (the values live in the table)
            o.nq = 1;
            o.q[0] = g.q0;
            for (int i = 0; i < 8; ++i) o.m[i] = 1.0;
        }
    }
    const int b = kRegTileBits, L = 3;
    std::vector<uint64_t>& order = *order_io;
    std::vector<uint64_t>& pass_end = *end_io;
    if (order.empty()) {
        order.assign(static_cast<size_t>(G), 0);
        pass_end.assign(static_cast<size_t>(G) + 1, 0);
        const int np = plan_order(n, ops.data(), uint64_t(G), b, L, order.data(), pass_end.data());
        pass_end.resize(size_t(std::max(np, 0)));
    }
    const int np = int(pass_end.size());
    int next_slot = 0;
    for (int v : *slot_of) next_slot += v >= 0 ? 1 : 0;
    std::vector<TiledPass> out;
    uint64_t first = 0;
    for (int p = 0; p < np; ++p) {
        const uint64_t last = pass_end[p];
        uint64_t T = (uint64_t(1) << L) - 1;
        for (uint64_t i = first; i < last; ++i) {
            const llsv_op& o = ops[order[i]];
            if (o.kind == LLSV_OP_1Q) T |= uint64_t(1) << o.q[0];
            if (o.kind == LLSV_OP_CNOT) T |= (uint64_t(1) << o.q[0]) | (uint64_t(1) << o.q[1]);
        }
        for (int q = 0; q < n && __builtin_popcountll(T) < b; ++q) T |= uint64_t(1) << q;
        TiledPass tp;
        PassDesc& d = tp.d;
        d.n = n;
        d.b = b;
        d.L = L;
        int local_of[64];
        for (int q = 0; q < 64; ++q) local_of[q] = -1;
        int j = 0;
        for (int q = 0; q < L; ++q) d.tq[j++] = q;
        for (int q = L; q < n; ++q)
            if ((T >> q) & 1) d.tq[j++] = q;
        for (int t = 0; t < b; ++t) {
            local_of[d.tq[t]] = t;
            d.pos[t] = d.tq[t];
        }
        std::sort(d.pos, d.pos + b);
        std::vector<PassOp> pops;
        for (uint64_t i = first; i < last; ++i) {
            const int k = int(order[i]);
            const llsv_pgate& g = *seq[k];
            PassOp po{};
            if (g.kind == LLSV_PG_CNOT) {
                po.kind = P_CNOT;
                po.a = local_of[g.q0];
                po.b = local_of[g.q1];
                pops.push_back(po);
            } else if (g.kind == LLSV_PG_CZ) {
                po.kind = P_MCZ;
                for (int q : {g.q0, g.q1}) {
                    if (local_of[q] >= 0) po.lmask |= uint64_t(1) << local_of[q];
                    else po.omask |= uint64_t(1) << q;
                }
                pops.push_back(po);
            } else {
                const int gi = gate_index[k];
                if (backward && g.kind <= LLSV_PG_RZ) {
                    PassOp gr{};
                    gr.kind = P_GRAD;
                    gr.a = local_of[g.q0];
                    gr.b = g.kind == LLSV_PG_RX ? 0 : g.kind == LLSV_PG_RY ? 1 : 2;
                    // slots are numbered in the order the backward sweep meets them, so the
                    // slots of one pass are a contiguous range
                    int& sl = (*slot_of)[gi];
                    if (sl < 0) sl = next_slot++;
                    gr.moff = sl;
                    pops.push_back(gr);
                }
                po.kind = P_1QT;
                po.a = local_of[g.q0];
                po.b = g.kind == LLSV_PG_RX ? 1 : g.kind == LLSV_PG_RY ? 2 : g.kind == LLSV_PG_RZ ? 3 : 0;
                po.moff = coef_of[gi] + (backward ? ncoef : 0);
                pops.push_back(po);
            }
        }
        const std::vector<RegOp> rops = to_register_program(pops, L, d.rq);
        d.nops = int(rops.size());
        tp.e = jit_entry(jit_adjoint_source(rops.data(), int(rops.size()), backward));
        out.push_back(tp);
        first = last;
    }
    return out;
}

std::mutex g_tiled_mu;
std::map<std::string, std::unique_ptr<TiledPlan>>& tiled_cache() {
    static auto* m = new std::map<std::string, std::unique_ptr<TiledPlan>>();  // lives to process end
    return *m;
}

// The tiled plan of a circuit (built and cached on first sight), or nullptr when the sweeps run
// gate by gate: state below one tile, pass compiler off, or kernels not compiled (yet).
TiledPlan* tiled_plan(int n, const llsv_pgate* gates, int G) {
    if (n < kRegTileBits || G == 0 || jit_mode() == kJitOff) return nullptr;
    if (!(jit_mode() == kJitAlways || n >= kJitMinQubits)) return nullptr;
    std::string key = std::to_string(n) + ":";
    for (int k = 0; k < G; ++k) {
        const llsv_pgate& g = gates[k];
        key += std::to_string(g.kind) + "," + std::to_string(g.q0) + "," + std::to_string(g.kind >= LLSV_PG_CNOT ? g.q1 : -1) + ";";
    }
    TiledPlan* tp = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_tiled_mu);
        auto it = tiled_cache().find(key);
        if (it == tiled_cache().end()) {
            auto plan = std::make_unique<TiledPlan>();
            std::vector<int> coef_of(G, -1), idx(G);
            for (int k = 0; k < G; ++k)
                if (gates[k].kind <= LLSV_PG_1Q) {
                    coef_of[k] = plan->ncoef++;
                    plan->coef_gate.push_back(k);
                }
            std::vector<const llsv_pgate*> seq(G);
            for (int k = 0; k < G; ++k) {
                seq[k] = &gates[k];
                idx[k] = k;
            }
            std::vector<uint64_t> forder, fend;
            plan->slot_of.assign(size_t(G), -1);
            plan->fwd = tiled_passes(n, seq, coef_of, &plan->slot_of, idx, plan->ncoef, false, &forder, &fend);
            // the backward sweep: the forward passes last to first, each one's gates last to first
            // (positions in the reversed gate list)
            std::vector<uint64_t> border, bend;
            for (size_t p = fend.size(); p-- > 0;) {
                const uint64_t start = p ? fend[p - 1] : 0;
                for (uint64_t i = fend[p]; i-- > start;) border.push_back(uint64_t(G - 1) - forder[i]);
                bend.push_back(border.size());
            }
            std::reverse(seq.begin(), seq.end());
            std::reverse(idx.begin(), idx.end());
            plan->bwd = tiled_passes(n, seq, coef_of, &plan->slot_of, idx, plan->ncoef, true, &border, &bend);
            it = tiled_cache().emplace(key, std::move(plan)).first;
        }
        tp = it->second.get();
    }
    std::vector<JitEntry*> want;
    for (const auto* v : {&tp->fwd, &tp->bwd})
        for (const TiledPass& p : *v)
            if (!jit_ready(p.e)) want.push_back(p.e);
    if (!want.empty()) jit_request(want, jit_mode() == kJitAlways);
    for (const auto* v : {&tp->fwd, &tp->bwd})
        for (const TiledPass& p : *v)
            if (!jit_ready(p.e)) return nullptr;
    return tp;
}

int run(int n, int B, const llsv_pgate* gates, int G, int P, const double* params, const llsv_pauli* terms, int T,
        double* energies, double* grads, void* ws, uint64_t ws_bytes, void* stream, const char* who) {
    if (int rc = validate(n, B, gates, G, P, params, terms, T, who)) return rc;
    if (!energies) return fail(LLSV_ERR_ARG, "%s: energies is null", who);
    const bool adj = grads != nullptr;
    std::vector<int> slot_param;
    std::vector<int> slot_of(G, -1);
    for (int k = 0; k < G; ++k)
        if (gates[k].kind <= LLSV_PG_RZ) {
            slot_of[k] = int(slot_param.size());
            slot_param.push_back(gates[k].param);
        }
    const Layout L = layout_of(n, B, G, int(slot_param.size()), T);
    const uint64_t need = adj ? L.total : L.total - L.states;
    if (!ws || ws_bytes < need)
        return fail(LLSV_ERR_ARG, "%s: workspace of %llu bytes needed, %llu given", who, (unsigned long long)need,
                    (unsigned long long)ws_bytes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned char* base = static_cast<unsigned char*>(ws);
    v2d* A = reinterpret_cast<v2d*>(base);
    size_t off = L.states;
    v2d* Lm = adj ? reinterpret_cast<v2d*>(base + off) : nullptr;
    if (adj) off += L.states;
    double* partials = reinterpret_cast<double*>(base + off);
    off += L.partials;
    double* slots = reinterpret_cast<double*>(base + off);
    off += L.slots;
    PauliDev* dterms = reinterpret_cast<PauliDev*>(base + off);
    off += L.terms;
    int* dslotmap = reinterpret_cast<int*>(base + off);

#define CK(expr)                                                                                       \
    do {                                                                                               \
        const hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(hip_code(e_), "%s: %s: %s", who, #expr, hipGetErrorString(e_)); \
    } while (0)

    std::vector<PauliDev> ht(T);
    for (int t = 0; t < T; ++t) ht[t] = PauliDev{terms[t].coeff, terms[t].xmask, terms[t].zmask};
    CK(hipMemcpyAsync(dterms, ht.data(), sizeof(PauliDev) * T, hipMemcpyHostToDevice, st));
    TiledPlan* tiled = tiled_plan(n, gates, G);
    if (tiled) {
        // the compiled sweeps number the slots themselves
        slot_of = tiled->slot_of;
        for (int k = 0; k < G; ++k)
            if (slot_of[k] >= 0) slot_param[size_t(slot_of[k])] = gates[k].param;
    }
    if (adj && !slot_param.empty())
        CK(hipMemcpyAsync(dslotmap, slot_param.data(), sizeof(int) * slot_param.size(), hipMemcpyHostToDevice, st));

    off += L.slotmap;
    double* table = reinterpret_cast<double*>(base + off);
    const uint64_t tstride = 16 * uint64_t(tiled ? tiled->ncoef : 0);
    const int ntiles = int(tiled_tiles(n));
    if (tiled && tiled->ncoef) {
        std::vector<PG> hg(G);
        std::memcpy(hg.data(), gates, sizeof(PG) * size_t(G));
        // gates and the coefficient map go through the (not yet used) partials area
        PG* dgates = reinterpret_cast<PG*>(partials);
        int* dcoef = reinterpret_cast<int*>(dgates + G);
        CK(hipMemcpyAsync(dgates, hg.data(), sizeof(PG) * size_t(G), hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(dcoef, tiled->coef_gate.data(), sizeof(int) * tiled->coef_gate.size(), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_pc_table, dim3(unsigned((tiled->ncoef + 255) / 256), unsigned(B)), dim3(256), 0, st, dgates,
                           dcoef, tiled->ncoef, params, P, B, table, tstride);
        CK(hipGetLastError());
    }

    hipLaunchKernelGGL(k_pc_init, dim3(unsigned((L.size + 255) / 256), unsigned(B)), dim3(256), 0, st, A, L.size);
    CK(hipGetLastError());
    if (tiled) {
        for (const TiledPass& p : tiled->fwd)
            CK(jit_launch_batched(p.e, unsigned(ntiles), unsigned(B), st, A, nullptr, L.size, p.d, table, tstride, nullptr,
                                  0, ntiles));
    } else {
        for (int k = 0; k < G; ++k) CK(apply_gate(st, A, nullptr, false, L, B, gates[k], params, P, false, partials, -1));
    }

    std::vector<HTile> htiles;
    std::vector<int> hfirst, hcount;
    std::vector<HTerm> hterms;
    // (the partials area holds B x (G + 1) x ntiles values: at most G + 1 passes)
    if (tiled && plan_hpsi(n, terms, T, htiles, hfirst, hcount, hterms) && int(htiles.size()) <= G + 1) {
        // lambda = H psi and the energies, tile pass by tile pass
        HTerm* dht = reinterpret_cast<HTerm*>(dterms);
        CK(hipMemcpyAsync(dht, hterms.data(), sizeof(HTerm) * hterms.size(), hipMemcpyHostToDevice, st));
        const int np = int(htiles.size());
        for (int ps = 0; ps < np; ++ps) {
            if (adj)
                hipLaunchKernelGGL(k_pc_hpsi_tile<true>, dim3(unsigned(ntiles), unsigned(B)), dim3(256), 0, st, A, Lm,
                                   L.size, htiles[ps], dht, hcount[ps], hfirst[ps], partials, ps, np);
            else
                hipLaunchKernelGGL(k_pc_hpsi_tile<false>, dim3(unsigned(ntiles), unsigned(B)), dim3(256), 0, st, A, Lm,
                                   L.size, htiles[ps], dht, hcount[ps], hfirst[ps], partials, ps, np);
            CK(hipGetLastError());
        }
        hipLaunchKernelGGL(k_pc_reduce, dim3(1, unsigned(B)), dim3(256), 0, st, partials, np * ntiles, 1, energies);
        CK(hipGetLastError());
    } else {
        const dim3 rgrid(unsigned(L.nblk), unsigned(B));
        if (adj)
            hipLaunchKernelGGL(k_pc_hpsi<true>, rgrid, dim3(256), 0, st, A, Lm, L.size, dterms, T, partials);
        else
            hipLaunchKernelGGL(k_pc_hpsi<false>, rgrid, dim3(256), 0, st, A, Lm, L.size, dterms, T, partials);
        CK(hipGetLastError());
        hipLaunchKernelGGL(k_pc_reduce, dim3(1, unsigned(B)), dim3(256), 0, st, partials, L.nblk, 1, energies);
        CK(hipGetLastError());
    }
    if (!adj) return LLSV_OK;

    int per_slot = L.nblk;
    if (tiled) {
        for (const TiledPass& p : tiled->bwd)
            CK(jit_launch_batched(p.e, unsigned(ntiles), unsigned(B), st, A, Lm, L.size, p.d, table, tstride, partials,
                                  L.nslot, ntiles));
        per_slot = ntiles;
    } else {
        for (int k = G - 1; k >= 0; --k)
            CK(apply_gate(st, A, Lm, true, L, B, gates[k], params, P, true, partials, slot_of[k]));
    }
    if (!slot_param.empty()) {
        hipLaunchKernelGGL(k_pc_reduce, dim3(unsigned(L.nslot), unsigned(B)), dim3(256), 0, st, partials, per_slot,
                           L.nslot, slots);
        CK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_pc_scatter, dim3(unsigned((B + 63) / 64)), dim3(64), 0, st, slots, dslotmap, L.nslot, P, B,
                       grads);
    CK(hipGetLastError());
#undef CK
    return LLSV_OK;
}

}  // namespace
}  // namespace llsv

extern "C" {

uint64_t llsv_pc_workspace_bytes(int n_qubits, int batch, int n_gates) {
    if (n_qubits < 1 || n_qubits > 30 || batch < 1 || n_gates < 0) return 0;
    return llsv::layout_of(n_qubits, batch, n_gates, n_gates, 4096).total;
}

int llsv_pc_energy(int n, int batch, const llsv_pgate* gates, int n_gates, int n_params, const double* params,
                   const llsv_pauli* terms, int n_terms, double* energies, void* ws, uint64_t ws_bytes, void* stream) {
    return llsv::run(n, batch, gates, n_gates, n_params, params, terms, n_terms, energies, nullptr, ws, ws_bytes,
                     stream, "llsv_pc_energy");
}

int llsv_pc_precompile(int n, const llsv_pgate* gates, int n_gates, int* forward_passes) {
    using namespace llsv;
    if (n < 1 || n > 30 || n_gates < 0 || (n_gates && !gates)) return fail(LLSV_ERR_ARG, "llsv_pc_precompile: bad arguments");
    for (int k = 0; k < n_gates; ++k) {
        if (gates[k].kind < LLSV_PG_RX || gates[k].kind > LLSV_PG_CZ || gates[k].q0 < 0 || gates[k].q0 >= n ||
            (gates[k].kind >= LLSV_PG_CNOT && (gates[k].q1 < 0 || gates[k].q1 >= n || gates[k].q1 == gates[k].q0)))
            return fail(LLSV_ERR_ARG, "llsv_pc_precompile: gate %d: bad kind or qubits", k);
    }
    if (forward_passes) *forward_passes = 0;
    const int prev = jit_mode();
    if (prev == kJitOff || (prev == kJitAuto && n < kJitMinQubits)) return 0;
    set_jit_mode(kJitAlways);  // compile now, in this thread
    TiledPlan* tp = tiled_plan(n, gates, n_gates);
    set_jit_mode(prev);
    if (!tp) return 0;
    if (forward_passes) *forward_passes = int(tp->fwd.size());
    return int(tp->fwd.size() + tp->bwd.size());
}

int llsv_pc_adjoint(int n, int batch, const llsv_pgate* gates, int n_gates, int n_params, const double* params,
                    const llsv_pauli* terms, int n_terms, double* energies, double* grads, void* ws,
                    uint64_t ws_bytes, void* stream) {
    if (!grads) return llsv::fail(LLSV_ERR_ARG, "llsv_pc_adjoint: grads is null");
    return llsv::run(n, batch, gates, n_gates, n_params, params, terms, n_terms, energies, grads, ws, ws_bytes,
                     stream, "llsv_pc_adjoint");
}

}  // extern "C"
