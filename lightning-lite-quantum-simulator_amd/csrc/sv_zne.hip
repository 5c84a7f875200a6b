// sv_zne.hip -- zero-noise-extrapolation trajectories: folded circuits under sampled depolarizing
// Pauli noise, many trajectories per launch.
//
// Pauli noise does not depend on the state, so a trajectory's noise is fixed before it runs: the
// Pauli inserted after each gate application (if any) comes either from a list the host draws
// from the reference's sequential stream (std::mt19937_64, the draws ll::NoisyBackend makes), or
// from counter-based Philox draws made on the device.  Nothing is normalised and nothing waits on
// a probability, so a trajectory is a plain gate sweep: each one lives in LDS (2^n x 16 B, n <= 12)
// and owns L = min(512, 2^(n-1)) threads of a 512-thread workgroup (512 / L trajectories per
// workgroup); above 12 qubits (up to 20) one trajectory per 1024-thread workgroup, its state in the
// workgroup's own slot in device memory.  Per gate application: one pass over the gate's
// amplitude pairs (one-qubit gates), quads (cnot, two-qubit mcz) or single amplitudes (wider
// mcz); the sampled Pauli acts on the same qubits and is applied in the same pass (X^x Z^z on the
// pair or quad), a workgroup barrier ends the pass.  A wider mcz takes a second pass for its
// per-qubit Paulis.  In a device-memory slot, a run of consecutive applications on at most 12
// qubits is applied through 64 KiB LDS tiles instead, one read and one write of the slot per run.
// At the end each group reduces sum_i w_i |psi_i|^2 for its trajectory and may copy its state out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "../../include/llsv.h"
#include "sv_common.h"

namespace llsv {

namespace {

constexpr int kZThreads = 512;
constexpr int kZSlotThreads = 1024;  // above 12 qubits: one trajectory per workgroup
constexpr int kZTileQubits = 12;     // device-memory states: runs of applications on at most 12 qubits
constexpr int kZRunMax = 256;        // go through LDS tiles of 2^12 amplitudes, at most this many per run
constexpr int64_t kZMaxApps = int64_t(1) << 22;  // gate applications of one folded circuit
constexpr size_t kZKeepBytes = size_t(64) << 20;    // per-application tables kept between calls
constexpr int kZLdsQubits = 12;  // up to 2^12 amplitudes (64 KiB) per state: 512-thread workgroups

// One gate application of the folded circuit.
struct ZApp {
    int32_t kind;  // 0 one-qubit matrix, 1 cnot, 2 mcz
    int32_t nq;    // qubits the noise acts on
    int32_t q0, q1;
    uint32_t mask;  // mcz: all its qubits
    uint32_t chan;  // 1: one 1-qubit channel, 2: one 2-qubit channel, 3: nq 1-qubit channels
    double m[8];    // one-qubit matrix, row-major (re, im)
    int8_t ql[16];  // noise slot i acts on qubit ql[i]
};

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

// Two uniforms in [0, 1) (53 bits each) for (trajectory, site, channel slot).
__device__ __host__ inline void philox_u2(uint64_t seed, uint64_t traj, uint32_t site, uint32_t slot, double* u,
                                          double* v) {
    uint32_t c[4] = {uint32_t(traj), uint32_t(traj >> 32), site, 0x5A4E0000u | slot};
    philox(c, uint32_t(seed), uint32_t(seed >> 32));
    const uint64_t x = uint64_t(c[0]) | (uint64_t(c[1]) << 32), y = uint64_t(c[2]) | (uint64_t(c[3]) << 32);
    *u = double(x >> 11) * (1.0 / 9007199254740992.0);
    *v = double(y >> 11) * (1.0 / 9007199254740992.0);
}

// Noise code of one site (two bits per channel slot) from Philox draws.
__device__ __host__ inline uint32_t philox_code(const ZApp& a, double p, uint64_t seed, uint64_t traj, uint32_t site) {
    double u, v;
    if (a.chan == 1) {
        philox_u2(seed, traj, site, 0, &u, &v);
        return u < p ? uint32_t(1 + int(v * 3.0)) : 0u;
    }
    if (a.chan == 2) {
        philox_u2(seed, traj, site, 0, &u, &v);
        return u < p ? uint32_t(1 + int(v * 15.0)) : 0u;
    }
    uint32_t code = 0;
    for (int i = 0; i < a.nq; ++i) {
        philox_u2(seed, traj, site, uint32_t(1 + i), &u, &v);
        if (u < p) code |= uint32_t(1 + int(v * 3.0)) << (2 * i);
    }
    return code;
}

// X and Z masks (over the state's qubits) of a noise code.  Pauli 1 = X, 2 = Y (Z then X), 3 = Z.
__device__ __forceinline__ void code_masks(const ZApp& a, uint32_t code, uint32_t& xm, uint32_t& zm) {
    xm = 0;
    zm = 0;
    for (int i = 0; i < a.nq && code; ++i, code >>= 2) {
        const uint32_t w = code & 3u, b = 1u << a.ql[i];
        if (w == 1u || w == 2u) xm |= b;
        if (w == 2u || w == 3u) zm |= b;
    }
}

__device__ __forceinline__ v2d neg_if(v2d v, bool neg) { return neg ? -v : v; }

// Group-wide sum of L consecutive lanes (L a power of two, L / 64 <= entries of red).
__device__ double group_sum(double v, int L, double* red) {
    const int w = int(threadIdx.x) >> 6;
    for (int o = 1; o < L && o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
    if (L <= 64) return v;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    const int first = (int(threadIdx.x) / L) * (L / 64);
    double s = 0.0;
    for (int k = 0; k < L / 64; ++k) s += red[first + k];
    __syncthreads();
    return s;
}

// The bits of v deposited, lowest first, at the set bits of m.
__device__ __forceinline__ uint32_t pdep(uint32_t v, uint32_t m) {
    uint32_t out = 0;
    for (uint32_t b = 1; m; b <<= 1, m &= m - 1)
        if (v & b) out |= m & (0u - m);
    return out;
}

// One gate application with its sampled Pauli (code) on a state of dim amplitudes at s (LDS or
// device memory), by thread r of the L threads that work on it.  Every thread of the workgroup
// calls it with the same application (a wider mcz has a barrier between its two passes); the
// caller ends the pass with a barrier.
__device__ __forceinline__ void apply_app(v2d* s, uint32_t dim, const ZApp& a, uint32_t code, uint32_t r,
                                          uint32_t L) {
    uint32_t xm, zm;
    code_masks(a, code, xm, zm);
    if (a.kind == 0 || (a.kind == 2 && a.nq == 1)) {
        // pairs along q0: gate (or Z for a one-qubit mcz), then X^x Z^z on the pair
        const int q = a.q0;
        const bool zq = (zm >> q) & 1u, xq = (xm >> q) & 1u;
        for (uint32_t pi = r; pi < dim / 2; pi += L) {
            const uint32_t i0 = uint32_t(insert0(pi, q)), i1 = i0 | (1u << q);
            v2d b0 = s[i0], b1 = s[i1];
            if (a.kind == 0) {
                const v2d c0 = row2(a.m, b0, b1), c1 = row2(a.m + 4, b0, b1);
                b0 = c0;
                b1 = c1;
            } else {
                b1 = -b1;
            }
            b1 = neg_if(b1, zq);
            s[i0] = xq ? b1 : b0;
            s[i1] = xq ? b0 : b1;
        }
    } else if (a.kind == 1 || (a.kind == 2 && a.nq == 2)) {
        // quads over (q0, q1): cnot or cz, then X^x Z^z on the quad; k = bit q0 + 2 bit q1
        const int qa = a.q0, qb = a.q1, lo = qa < qb ? qa : qb, hi = qa < qb ? qb : qa;
        const uint32_t xk = ((xm >> qa) & 1u) | (((xm >> qb) & 1u) << 1);
        const uint32_t zk = ((zm >> qa) & 1u) | (((zm >> qb) & 1u) << 1);
        for (uint32_t qi = r; qi < dim / 4; qi += L) {
            const uint32_t base = uint32_t(insert0(insert0(qi, lo), hi));
            uint32_t idx[4];
            v2d v[4];
            for (int k = 0; k < 4; ++k) {
                idx[k] = base | ((k & 1) ? (1u << qa) : 0u) | ((k & 2) ? (1u << qb) : 0u);
                v[k] = s[idx[k]];
            }
            if (a.kind == 1) {  // control qa, target qb: |1,0> <-> |1,1>
                const v2d tmp = v[1];
                v[1] = v[3];
                v[3] = tmp;
            } else {
                v[3] = -v[3];
            }
            for (int k = 0; k < 4; ++k) v[k] = neg_if(v[k], __builtin_popcount(uint32_t(k) & zk) & 1);
            for (int k = 0; k < 4; ++k) s[idx[k]] = v[uint32_t(k) ^ xk];
        }
    } else {
        // wider mcz: signs, then the per-qubit Paulis in a second pass
        for (uint32_t i = r; i < dim; i += L)
            if ((i & a.mask) == a.mask) s[i] = -s[i];
        __syncthreads();
        if (code) {
            if (xm == 0) {
                for (uint32_t i = r; i < dim; i += L) s[i] = neg_if(s[i], __builtin_popcount(i & zm) & 1);
            } else {
                const int hb = 31 - __builtin_clz(xm);
                for (uint32_t pi = r; pi < dim / 2; pi += L) {
                    const uint32_t i = uint32_t(insert0(pi, hb)), k = i ^ xm;
                    const v2d vi = s[i], vk = s[k];
                    s[i] = neg_if(vk, __builtin_popcount(k & zm) & 1);
                    s[k] = neg_if(vi, __builtin_popcount(i & zm) & 1);
                }
            }
        }
    }
}

// NT threads per workgroup, 2^lgL of them per trajectory.  SLOT: the state is slot blockIdx.x of
// `slots` (2^n amplitudes each) instead of LDS (then NT = 1024, lgL = 10: one trajectory per
// workgroup), and run[j].x > 1 marks a run of applications j .. j + run[j].x - 1 whose qubits all lie
// in the kZTileQubits-qubit set run[j].y: the run is applied tile by tile (the 2^kZTileQubits
// amplitudes that share their bits outside the set, staged in the dynamic LDS; tapps holds the run's
// applications with qubits renumbered by rank in the set), so the slot is read and written once per
// run instead of once per application.
template <bool PHILOX, bool SLOT, int NT>
__global__ __launch_bounds__(NT) void k_zne(const ZApp* __restrict__ apps, int S, int n, int lgL,
                                            uint64_t n_traj, const uint2* __restrict__ ev,
                                            const uint64_t* __restrict__ ev_off, double p, uint64_t seed,
                                            uint64_t traj_base, const double* __restrict__ w,
                                            double* __restrict__ vals, v2d* __restrict__ amps,
                                            v2d* __restrict__ slots, const uint2* __restrict__ run,
                                            const ZApp* __restrict__ tapps) {
    extern __shared__ v2d lds_states[];
    __shared__ double red[NT / 64];
    __shared__ uint32_t run_codes[SLOT ? kZRunMax : 1];
    const int L = 1 << lgL, G = NT >> lgL;
    const int g = int(threadIdx.x) >> lgL, r = int(threadIdx.x) & (L - 1);
    const uint32_t dim = 1u << n;
    v2d* s = SLOT ? slots + size_t(blockIdx.x) * dim : lds_states + size_t(g) * dim;
    for (uint64_t tb = uint64_t(blockIdx.x) * G; tb < n_traj; tb += uint64_t(gridDim.x) * G) {
        const uint64_t t = tb + uint64_t(g);
        const bool live = t < n_traj;
        for (uint32_t i = uint32_t(r); i < dim; i += uint32_t(L)) s[i] = v2d{i == 0 ? 1.0 : 0.0, 0.0};
        uint64_t cur = 0, end = 0;
        if (!PHILOX && live) {
            cur = ev_off[t];
            end = ev_off[t + 1];
        }
        uint32_t next_site = cur < end ? ev[cur].x : 0xFFFFFFFFu;
        __syncthreads();
        for (int j = 0; j < S;) {
            if (SLOT && run[j].x > 1) {
                const int R = int(run[j].x);  // 2 .. kZRunMax
                // the run's noise codes, by application
                if (r < R)
                    run_codes[r] = (PHILOX && live && p > 0.0)
                                       ? philox_code(apps[j + r], p, seed, traj_base + t, uint32_t(j + r))
                                       : 0u;
                __syncthreads();
                if (!PHILOX) {
                    // every thread walks the event list (the same cursor everywhere); thread 0 writes
                    while (live && next_site < uint32_t(j + R)) {
                        if (r == 0) run_codes[next_site - uint32_t(j)] = ev[cur].y;
                        ++cur;
                        next_site = cur < end ? ev[cur].x : 0xFFFFFFFFu;
                    }
                    __syncthreads();
                }
                // tile element e = r + i L sits at slot offset pdep(e, T) + pdep(k, O) in tile k
                constexpr uint32_t kTile = 1u << kZTileQubits;
                constexpr int kPer = int(kTile) / kZSlotThreads;  // L = kZSlotThreads here
                const uint32_t T = run[j].y, O = (dim - 1u) & ~T;
                uint32_t off[kPer];
                for (int i = 0; i < kPer; ++i) off[i] = pdep(uint32_t(r) + uint32_t(i) * uint32_t(L), T);
                for (uint32_t k = 0; k < (dim >> kZTileQubits); ++k) {
                    v2d* gt = s + pdep(k, O);
                    for (int i = 0; i < kPer; ++i) lds_states[uint32_t(r) + uint32_t(i) * uint32_t(L)] = gt[off[i]];
                    __syncthreads();
                    for (int u = 0; u < R; ++u) {
                        apply_app(lds_states, kTile, tapps[j + u], run_codes[u], uint32_t(r), uint32_t(L));
                        __syncthreads();
                    }
                    // each thread stores the amplitudes it loaded, so the next tile's loads need no barrier
                    for (int i = 0; i < kPer; ++i) gt[off[i]] = lds_states[uint32_t(r) + uint32_t(i) * uint32_t(L)];
                }
                __syncthreads();
                j += R;
                continue;
            }
            const ZApp& a = apps[j];
            uint32_t code = 0;
            if (live) {
                if (PHILOX) {
                    if (p > 0.0) code = philox_code(a, p, seed, traj_base + t, uint32_t(j));
                } else if (next_site == uint32_t(j)) {
                    code = ev[cur].y;
                    ++cur;
                    next_site = cur < end ? ev[cur].x : 0xFFFFFFFFu;
                }
            }
            apply_app(s, dim, a, code, uint32_t(r), uint32_t(L));
            __syncthreads();
            ++j;
        }
        if (w && vals) {
            double acc = 0.0;
            for (uint32_t i = uint32_t(r); i < dim; i += uint32_t(L)) {
                const v2d v = s[i];
                acc += w[i] * (v.x * v.x + v.y * v.y);
            }
            acc = group_sum(acc, L, red);
            if (live && r == 0) vals[t] = acc;
        }
        if (amps && live)
            for (uint32_t i = uint32_t(r); i < dim; i += uint32_t(L)) amps[t * dim + i] = s[i];
        __syncthreads();
    }
}

// Arguments of one k_zne launch.
struct ZArgs {
    const ZApp* apps;
    int S, n, lgL;
    uint64_t n_traj;
    const uint2* ev;
    const uint64_t* ev_off;
    double p;
    uint64_t seed, traj_base;
    const double* w;
    double* vals;
    v2d* amps;
    v2d* slots;
    const uint2* run;
    const ZApp* tapps;
};

template <bool PHILOX, bool SLOT, int NT>
hipError_t launch_zne(unsigned grid, size_t lds, const ZArgs& a) {
    hipLaunchKernelGGL((k_zne<PHILOX, SLOT, NT>), dim3(grid), dim3(NT), lds, nullptr, a.apps, a.S, a.n, a.lgL,
                       a.n_traj, a.ev, a.ev_off, a.p, a.seed, a.traj_base, a.w, a.vals, a.amps, a.slots,
                       a.run, a.tapps);
    return hipGetLastError();
}

}  // namespace
}  // namespace llsv

struct llsv_noisy {
    double p = 0.0;
    int rng = 0;
    uint64_t seed = 1;
    std::mt19937_64 mt;
    uint64_t philox_next = 0;  // trajectories drawn so far (rng 1)
};

namespace llsv {
namespace {

double u01(std::mt19937_64& mt) { return double(mt() >> 11) * (1.0 / 9007199254740992.0); }

// One noise site from the reference stream: NoisyBackend::noise / depolarize1 / depolarize2.
uint32_t mt_site(std::mt19937_64& mt, double p, int nq) {
    if (nq == 1) {
        if (u01(mt) >= p) return 0;
        return uint32_t(1 + int(u01(mt) * 3.0));
    }
    if (nq == 2) {
        if (u01(mt) >= p) return 0;
        return uint32_t(1 + int(u01(mt) * 15.0));
    }
    uint32_t code = 0;
    for (int i = 0; i < nq; ++i) {
        if (u01(mt) >= p) continue;
        code |= uint32_t(1 + int(u01(mt) * 3.0)) << (2 * i);
    }
    return code;
}

void mat_of(const llsv_zgate& g, bool inverse, double* m) {
    const double t = inverse ? -g.theta : g.theta;
    const double c = std::cos(0.5 * t), s = std::sin(0.5 * t);
    const double r = 1.0 / std::sqrt(2.0);
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    switch (g.kind) {
        case LLSV_ZG_H: a[0] = r, a[2] = r, a[4] = r, a[6] = -r; break;
        case LLSV_ZG_X: a[2] = 1.0, a[4] = 1.0; break;
        case LLSV_ZG_RX: a[0] = c, a[3] = -s, a[5] = -s, a[6] = c; break;
        case LLSV_ZG_RY: a[0] = c, a[2] = -s, a[4] = s, a[6] = c; break;
        default: a[0] = c, a[1] = -s, a[6] = c, a[7] = s; break;  // RZ
    }
    std::memcpy(m, a, sizeof a);
}

int check_gates(int n, const llsv_zgate* gates, int G) {
    for (int k = 0; k < G; ++k) {
        const llsv_zgate& g = gates[k];
        if (g.kind < LLSV_ZG_H || g.kind > LLSV_ZG_MCZ) return fail(LLSV_ERR_ARG, "llsv_noisy_run: gate %d: unknown kind", k);
        const int want = g.kind == LLSV_ZG_CNOT ? 2 : g.kind == LLSV_ZG_MCZ ? g.nq : 1;
        if (g.kind == LLSV_ZG_MCZ && (g.nq < 1 || g.nq > 16))
            return fail(LLSV_ERR_ARG, "llsv_noisy_run: gate %d: mcz on %d qubits", k, g.nq);
        if (g.nq != want) return fail(LLSV_ERR_ARG, "llsv_noisy_run: gate %d: nq = %d, expected %d", k, g.nq, want);
        uint32_t seen = 0;
        for (int i = 0; i < want; ++i) {
            const int q = g.q[i];
            if (q < 0 || q >= n) return fail(LLSV_ERR_ARG, "llsv_noisy_run: gate %d: qubit %d outside [0, %d)", k, q, n);
            if (seen & (1u << q)) return fail(LLSV_ERR_ARG, "llsv_noisy_run: gate %d: duplicate qubit %d", k, q);
            seen |= 1u << q;
        }
        if (g.kind >= LLSV_ZG_RX && g.kind <= LLSV_ZG_RZ && !std::isfinite(g.theta))
            return fail(LLSV_ERR_ARG, "llsv_noisy_run: gate %d: non-finite angle", k);
    }
    return LLSV_OK;
}

// The folded application list: U (U^+ U)^k per logical gate.
std::vector<ZApp> fold(const llsv_zgate* gates, int G, int folds) {
    std::vector<ZApp> out;
    out.reserve(size_t(G) * size_t(2 * folds + 1));
    for (int k = 0; k < G; ++k) {
        const llsv_zgate& g = gates[k];
        ZApp a{};
        a.nq = g.kind == LLSV_ZG_MCZ ? g.nq : g.kind == LLSV_ZG_CNOT ? 2 : 1;
        for (int i = 0; i < a.nq; ++i) {
            a.ql[i] = g.q[i];
            a.mask |= 1u << g.q[i];
        }
        a.q0 = g.q[0];
        a.q1 = a.nq > 1 ? g.q[1] : 0;
        a.chan = a.nq == 1 ? 1u : a.nq == 2 ? 2u : 3u;
        a.kind = g.kind == LLSV_ZG_CNOT ? 1 : g.kind == LLSV_ZG_MCZ ? 2 : 0;
        ZApp inv = a;
        if (a.kind == 0) {
            mat_of(g, false, a.m);
            mat_of(g, true, inv.m);
        }
        out.push_back(a);
        for (int f = 0; f < folds; ++f) {
            out.push_back(inv);
            out.push_back(a);
        }
    }
    return out;
}

// Device buffers, grown on demand and kept for the process (one set; calls are serialised).
struct ZBuf {
    void* p = nullptr;
    size_t cap = 0;
    void release_above(size_t keep) {
        if (cap > keep) {
            (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
    }
    hipError_t need(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const hipError_t e = dev_alloc(&p, bytes);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
};

// Pinned host staging, grown on demand and kept for the process.
struct PinBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t need(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const hipError_t e = dev_host_alloc(&p, bytes);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
};

}  // namespace
}  // namespace llsv

extern "C" {

int llsv_noisy_create(double p, uint64_t seed, int rng, llsv_noisy** out) {
    if (!out) return llsv::fail(LLSV_ERR_ARG, "llsv_noisy_create: out is null");
    *out = nullptr;
    if (!(p >= 0.0 && p <= 1.0)) return llsv::fail(LLSV_ERR_ARG, "NoisyBackend: p must be in [0, 1]");
    if (rng != 0 && rng != 1) return llsv::fail(LLSV_ERR_ARG, "llsv_noisy_create: rng must be 0 or 1");
    auto* z = new llsv_noisy();
    z->p = p;
    z->rng = rng;
    z->seed = seed;
    z->mt.seed(seed);
    *out = z;
    return LLSV_OK;
}

void llsv_noisy_destroy(llsv_noisy* z) { delete z; }

int llsv_noisy_reseed(llsv_noisy* z, uint64_t seed) {
    if (!z) return llsv::fail(LLSV_ERR_HANDLE, "llsv_noisy_reseed: null handle");
    z->seed = seed;
    z->mt.seed(seed);
    z->philox_next = 0;
    return LLSV_OK;
}

int llsv_noisy_site(llsv_noisy* z, int nq, uint32_t* code) {
    if (!z) return llsv::fail(LLSV_ERR_HANDLE, "llsv_noisy_site: null handle");
    if (!code || nq < 1 || nq > 16) return llsv::fail(LLSV_ERR_ARG, "llsv_noisy_site: bad arguments");
    *code = z->p == 0.0 ? 0u : llsv::mt_site(z->mt, z->p, nq);
    return LLSV_OK;
}

int llsv_noisy_run(llsv_noisy* z, int n, const llsv_zgate* gates, int n_gates, int scale, int64_t n_traj,
                   const double* weights, double* values, double* amps) {
    using namespace llsv;
    if (!z) return fail(LLSV_ERR_HANDLE, "llsv_noisy_run: null handle");
    if (n < 1 || n > LLSV_ZNE_MAX_QUBITS)
        return fail(LLSV_ERR_ARG, "llsv_noisy_run: n_qubits must be in [1, %d]", LLSV_ZNE_MAX_QUBITS);
    if (scale < 1 || scale > 99 || scale % 2 == 0)
        return fail(LLSV_ERR_ARG, "llsv_noisy_run: scale must be an odd integer in [1, 99]");
    if (n_gates < 1 || !gates) return fail(LLSV_ERR_ARG, "llsv_noisy_run: empty circuit");
    if (n_traj < 0) return fail(LLSV_ERR_ARG, "llsv_noisy_run: n_traj < 0");
    if ((weights == nullptr) != (values == nullptr))
        return fail(LLSV_ERR_ARG, "llsv_noisy_run: weights and values go together");
    if (int64_t(n_gates) * scale > kZMaxApps)
        return fail(LLSV_ERR_ARG, "llsv_noisy_run: %d gates at scale %d exceed %lld gate applications", n_gates, scale,
                    static_cast<long long>(kZMaxApps));
    if (int rc = check_gates(n, gates, n_gates)) return rc;
    if (n_traj == 0) return LLSV_OK;

    const std::vector<ZApp> apps = fold(gates, n_gates, (scale - 1) / 2);
    const int S = int(apps.size());
    const uint64_t N = uint64_t(n_traj), dim = uint64_t(1) << n;

    static std::mutex mu;  // the device buffers below are shared by every handle
    std::lock_guard<std::mutex> lock(mu);
    static ZBuf b_apps, b_w, b_vals, b_amps, b_slots, b_run, b_tapps;
    static ZBuf b_ev[2], b_off[2];
    static hipEvent_t done[2] = {nullptr, nullptr};
    static int owner = -1;  // the device the cached buffers and events live on
    {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_noisy_run: no device");
        if (owner != dev) {  // another device is current: the buffers are allocated again there
            for (ZBuf* b : {&b_apps, &b_w, &b_vals, &b_amps, &b_slots, &b_run, &b_tapps, &b_ev[0], &b_ev[1],
                            &b_off[0], &b_off[1]}) {
                if (b->p) (void)hipFree(b->p);
                b->p = nullptr;
                b->cap = 0;
            }
            for (hipEvent_t& e : done) {
                if (e) (void)hipEventDestroy(e);
                e = nullptr;
            }
            // 64 KiB of dynamic LDS: one 12-qubit state, or one tile of a device-memory state
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_zne<false, false, kZThreads>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 65536);
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_zne<true, false, kZThreads>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 65536);
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_zne<false, true, kZSlotThreads>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 65536);
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_zne<true, true, kZSlotThreads>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 65536);
            owner = dev;
        }
    }
    // Launch shape.  Up to 12 qubits: 512-thread workgroups, L = min(512, 2^(n-1)) threads per
    // trajectory, 512 / L states in LDS.  Above: one trajectory per 1024-thread workgroup, its state in
    // a device-memory slot with LDS-tiled runs (MI355X, 13 qubits: 107 gate applications per µs against
    // 85 with the whole 128 KiB state in LDS, one workgroup per CU).
    const bool in_slot = n > kZLdsQubits;
    const int lgL = in_slot ? 10 : std::min(9, std::max(0, n - 1));
    const int G = in_slot ? 1 : kZThreads >> lgL;
    const size_t lds = in_slot ? size_t(16) << kZTileQubits : size_t(G) * dim * 16;
    using Launch = hipError_t (*)(unsigned, size_t, const ZArgs&);
    const bool philox = z->rng != 0;
    const Launch launch = in_slot ? (philox ? launch_zne<true, true, kZSlotThreads> : launch_zne<false, true, kZSlotThreads>)
                                  : (philox ? launch_zne<true, false, kZThreads> : launch_zne<false, false, kZThreads>);
#define ZK(expr)                                                                                         \
    do {                                                                                                 \
        const hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(hip_code(e_), "llsv_noisy_run: %s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
    ZK(b_apps.need(sizeof(ZApp) * apps.size()));
    ZK(hipMemcpy(b_apps.p, apps.data(), sizeof(ZApp) * apps.size(), hipMemcpyHostToDevice));
    if (weights) {
        ZK(b_w.need(sizeof(double) * dim));
        ZK(hipMemcpy(b_w.p, weights, sizeof(double) * dim, hipMemcpyHostToDevice));
        ZK(b_vals.need(sizeof(double) * N));
    }
    // Trajectories go in chunks: final states at most 256 MiB at a time, and with the reference
    // stream at most 2048 trajectories, so the host draws chunk c + 1 while the device runs chunk c
    // (two event buffers, pinned host staging, an event per buffer).
    uint64_t chunk = N;
    if (amps) chunk = std::max<uint64_t>(1, (uint64_t(256) << 20) / (dim * 16));
    if (z->rng == 0) chunk = std::min<uint64_t>(chunk, 2048);
    if (amps) ZK(b_amps.need(size_t(std::min(chunk, N) * dim * 16)));
    // device-memory states: one slot per workgroup, one workgroup per CU, at most 4 GiB of slots
    // (MI355X, with the LDS-tiled runs: 128 slots read 12.8 TB/s of state-equivalent traffic at 16
    // qubits, 256 slots 19.2, 512 slots 18.3)
    unsigned slot_blocks = 0;
    if (in_slot) {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        const uint64_t slot_bytes = dim * 16;
        const uint64_t cap = std::max<uint64_t>(1, (uint64_t(4) << 30) / slot_bytes);
        slot_blocks = unsigned(std::min<uint64_t>({uint64_t(std::max(1, cus)), cap, std::min(chunk, N)}));
        ZK(b_slots.need(size_t(slot_blocks) * slot_bytes));
        // runs: greedily the longest stretch of applications whose qubits together number at most
        // kZTileQubits (and at most kZRunMax applications), if it holds two or more; the tile's qubit
        // set is theirs, filled up with the lowest other qubits (whole 16-byte amplitudes, and runs of
        // consecutive ones, per load)
        std::vector<uint2> runs(size_t(S), make_uint2(1u, 0u));
        std::vector<ZApp> tapps(apps);
        for (int j = 0; j < S;) {
            uint32_t u = 0;
            int k = j;
            while (k < S && k - j < kZRunMax && __builtin_popcount(u | apps[size_t(k)].mask) <= kZTileQubits)
                u |= apps[size_t(k++)].mask;
            if (k - j < 2) {
                ++j;
                continue;
            }
            uint32_t T = u;
            for (int b = 0; __builtin_popcount(T) < kZTileQubits; ++b) T |= 1u << b;
            runs[size_t(j)] = make_uint2(uint32_t(k - j), T);
            const auto rank = [T](int q) { return int8_t(__builtin_popcount(T & ((1u << q) - 1u))); };
            for (int x = j; x < k; ++x) {
                ZApp& b = tapps[size_t(x)];
                b.q0 = rank(b.q0);
                b.q1 = rank(b.q1);
                b.mask = 0;
                for (int i = 0; i < b.nq; ++i) {
                    b.ql[i] = rank(b.ql[i]);
                    b.mask |= 1u << b.ql[i];
                }
            }
            j = k;
        }
        ZK(b_run.need(sizeof(uint2) * runs.size()));
        ZK(hipMemcpy(b_run.p, runs.data(), sizeof(uint2) * runs.size(), hipMemcpyHostToDevice));
        ZK(b_tapps.need(sizeof(ZApp) * tapps.size()));
        ZK(hipMemcpy(b_tapps.p, tapps.data(), sizeof(ZApp) * tapps.size(), hipMemcpyHostToDevice));
    }
    static PinBuf h_ev[2], h_off[2];
    for (int k = 0; k < 2; ++k)
        if (!done[k]) ZK(hipEventCreateWithFlags(&done[k], hipEventDisableTiming));
    std::vector<uint2> ev;
    std::vector<uint64_t> off;
    int c = 0;
    for (uint64_t t0 = 0; t0 < N; t0 += chunk, ++c) {
        const uint64_t nt = std::min(chunk, N - t0);
        const int ev_slot = c & 1;
        if (z->rng == 0) {
            if (c >= 2) ZK(hipEventSynchronize(done[ev_slot]));  // chunk c - 2 no longer reads these buffers
            // the reference stream, trajectory by trajectory, site by site
            ev.clear();
            off.assign(nt + 1, 0);
            for (uint64_t t = 0; t < nt; ++t) {
                if (z->p > 0.0)
                    for (int j = 0; j < S; ++j) {
                        const uint32_t code = mt_site(z->mt, z->p, apps[size_t(j)].nq);
                        if (code) ev.push_back(make_uint2(uint32_t(j), code));
                    }
                off[t + 1] = ev.size();
            }
            const size_t evb = sizeof(uint2) * std::max<size_t>(1, ev.size()), offb = sizeof(uint64_t) * off.size();
            ZK(h_ev[ev_slot].need(evb));
            ZK(h_off[ev_slot].need(offb));
            if (!ev.empty()) std::memcpy(h_ev[ev_slot].p, ev.data(), sizeof(uint2) * ev.size());
            std::memcpy(h_off[ev_slot].p, off.data(), offb);
            ZK(b_ev[ev_slot].need(evb));
            ZK(b_off[ev_slot].need(offb));
            if (!ev.empty()) ZK(hipMemcpyAsync(b_ev[ev_slot].p, h_ev[ev_slot].p, sizeof(uint2) * ev.size(), hipMemcpyHostToDevice, nullptr));
            ZK(hipMemcpyAsync(b_off[ev_slot].p, h_off[ev_slot].p, offb, hipMemcpyHostToDevice, nullptr));
        }
        const unsigned grid = unsigned(std::min<uint64_t>((nt + G - 1) / G, in_slot ? slot_blocks : 16384));
        // w is null unless weights were given (the cached buffer may hold an earlier call's)
        const double* dw = weights ? static_cast<const double*>(b_w.p) : nullptr;
        double* dvals = weights ? static_cast<double*>(b_vals.p) + t0 : nullptr;
        v2d* damps = amps ? static_cast<v2d*>(b_amps.p) : nullptr;
        ZArgs args{static_cast<const ZApp*>(b_apps.p), S, n, lgL, nt, nullptr, nullptr, z->p, z->seed, 0,
                   dw, dvals, damps, static_cast<v2d*>(b_slots.p),
                   in_slot ? static_cast<const uint2*>(b_run.p) : nullptr,
                   in_slot ? static_cast<const ZApp*>(b_tapps.p) : nullptr};
        if (philox) {
            args.traj_base = z->philox_next + t0;
        } else {
            args.ev = static_cast<const uint2*>(b_ev[ev_slot].p);
            args.ev_off = static_cast<const uint64_t*>(b_off[ev_slot].p);
        }
        ZK(launch(grid, lds, args));
        if (z->rng == 0) ZK(hipEventRecord(done[ev_slot], nullptr));
        if (amps) ZK(hipMemcpy(amps + 2 * t0 * dim, damps, size_t(nt * dim * 16), hipMemcpyDeviceToHost));
    }
    if (weights) ZK(hipMemcpy(values, b_vals.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    ZK(hipDeviceSynchronize());
#undef ZK
    // the per-application tables grow with the folded circuit (104 bytes per application with
    // the renumbered copy): keep them for the next call up to kZKeepBytes, release larger ones
    for (ZBuf* b : {&b_apps, &b_tapps, &b_run}) b->release_above(kZKeepBytes);
    if (z->rng == 1) z->philox_next += N;
    return LLSV_OK;
}

}  // extern "C"
