// sv_jit.hip -- generation, compilation (hipRTC) and launch of straight-line pass kernels.
#include "sv_jit.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "sv_regops.h"

namespace llsv {

static_assert(sizeof(PassDesc) == 4 * (4 + 2 * kMaxTileBits + 4), "generated kernels restate PassDesc");
static_assert(offsetof(RegOp, m) % sizeof(double) == 0 && sizeof(RegOp) % sizeof(double) == 0,
              "generated kernels address RegOp coefficients as doubles");

namespace {

std::atomic<int> g_mode{-1};
std::atomic<uint64_t> g_compiled{0}, g_launches{0}, g_failures{0};

const char kPrelude[] =
    "typedef double v2d __attribute__((ext_vector_type(2)));\n"
    "struct PassDesc { int n, b, L, nops; int tq[12]; int pos[12]; int rq[4]; };\n"
    "__device__ __forceinline__ unsigned int r_swzm(unsigned int l, unsigned int r0, unsigned int r1,"
    " unsigned int r2, unsigned int r3) { return l ^ ((__builtin_popcount(l & r0) & 1u) |"
    " ((__builtin_popcount(l & r1) & 1u) << 1) | ((__builtin_popcount(l & r2) & 1u) << 2) |"
    " ((__builtin_popcount(l & r3) & 1u) << 3)); }\n"
    // r_wsum: the sum over the wavefront's 64 lanes in a fixed order, by DPP moves (no LDS round
    // trips): pairs, quads, eights, rows of 16, then rows 0+1 into row 1 and rows 0..3 into lane 63
    // (r_dpp: every lane reads a valid source lane; r_dppz: rows outside RM read 0)
    "template <int C> __device__ __forceinline__ double r_dpp(double v) {"
    " const unsigned long long u = __builtin_bit_cast(unsigned long long, v);"
    " const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, C, 0xf, 0xf, false);"
    " const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), C, 0xf, 0xf, false);"
    " return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo); }\n"
    "template <int C, int RM> __device__ __forceinline__ double r_dppz(double v) {"
    " const unsigned long long u = __builtin_bit_cast(unsigned long long, v);"
    " const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, C, RM, 0xf, false);"
    " const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), C, RM, 0xf, false);"
    " return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo); }\n"
    "__device__ __forceinline__ double r_wsum(double v) {"
    " v += r_dpp<0xB1>(v); v += r_dpp<0x4E>(v); v += r_dpp<0x141>(v); v += r_dpp<0x140>(v);"
    " v += r_dppz<0x142, 0xA>(v); v += r_dppz<0x143, 0xC>(v);"
    " const unsigned long long u = __builtin_bit_cast(unsigned long long, v);"
    " const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, 63);"
    " const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), 63);"
    " return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo); }\n"
    LLSV_STRINGIZE(LLSV_REGOPS_SOURCE) "\n";

// ---- LDS slot swizzle of a generated pass ----------------------------------------------------
// A layout change writes every amplitude of the tile into LDS and reads it back in the new
// layout; amplitude L (12-bit local index) sits in 16-byte slot f(L), f(L) = L ^ (M L) with M a
// 4 x 12 matrix over GF(2) into bits 0..3 (rows[b] = the bits whose parity flips slot bit b).  f
// is linear, so a lane's slot is f(its lane part) ^ f(the register part).  A 128-bit LDS access
// serves 8 (or 16) consecutive lanes per cycle, so it is free of bank conflicts when the lowest 3
// (or 4) lane positions of the layout map to independent slot bits mod 8 (or 16).  The fixed
// swizzle of the interpreter (fold bits 4..7 onto 0..3) fails that for some register sets; a
// generated pass picks its own M: the one, among the fixed fold and a deterministic family of
// candidates, with the fewest conflict cycles over every layout of the pass.
struct Swz {
    uint32_t row[4];
};

uint32_t swz_apply(const Swz& m, uint32_t l) {
    uint32_t o = 0;
    for (int b = 0; b < 4; ++b) o |= uint32_t(__builtin_popcount(l & m.row[b]) & 1) << b;
    return l ^ o;
}

// Local index bits of register j in register set S (deposit the 4 bits of j at S's positions).
uint32_t reg_bits(uint32_t S, int j) {
    uint32_t o = 0;
    int t = 0;
    for (int bit = 0; bit < kRegTileBits; ++bit)
        if ((S >> bit) & 1u) {
            if ((j >> t) & 1) o |= 1u << bit;
            ++t;
        }
    return o;
}

int rank_gf2(uint32_t* v, int n) {
    int r = 0;
    for (int bit = 3; bit >= 0; --bit) {
        int piv = -1;
        for (int i = r; i < n; ++i)
            if ((v[i] >> bit) & 1u) {
                piv = i;
                break;
            }
        if (piv < 0) continue;
        std::swap(v[r], v[piv]);
        for (int i = 0; i < n; ++i)
            if (i != r && ((v[i] >> bit) & 1u)) v[i] ^= v[r];
        ++r;
    }
    return r;
}

// Extra cycles of the 128-bit accesses made with register set S (lanes deposit tid into the other
// positions, ascending), for 8- and 16-lane service groups.
int swz_cost(const Swz& m, uint32_t S) {
    int pos[8], k = 0;
    for (int bit = 0; bit < kRegTileBits && k < 8; ++bit)
        if (!((S >> bit) & 1u)) pos[k++] = bit;
    uint32_t v16[4], v8[3];
    for (int i = 0; i < 4; ++i) v16[i] = swz_apply(m, 1u << pos[i]) & 15u;
    for (int i = 0; i < 3; ++i) v8[i] = swz_apply(m, 1u << pos[i]) & 7u;
    return ((1 << (4 - rank_gf2(v16, 4))) - 1) + ((1 << (3 - rank_gf2(v8, 3))) - 1);
}

bool swz_invertible(const Swz& m) {
    uint32_t v[4];
    for (int c = 0; c < 4; ++c) v[c] = swz_apply(m, 1u << c) & 15u;
    return rank_gf2(v, 4) == 4;
}

Swz choose_swizzle(const RegOp* ops, int nops) {
    std::vector<uint32_t> sets;  // register sets of every LDS access (both sides of each layout change)
    uint32_t S = 0xF00u;
    for (int i = 0; i < nops; ++i)
        if (ops[i].kind == RO_LAYOUT) {
            sets.push_back(S);
            S = (1u << ops[i].q[0]) | (1u << ops[i].q[1]) | (1u << ops[i].q[2]) | (1u << ops[i].q[3]);
            sets.push_back(S);
        }
    Swz best{{0x10u, 0x20u, 0x40u, 0x80u}};  // the interpreter's fold
    auto cost = [&](const Swz& m) {
        int c = 0;
        for (uint32_t x : sets) c += swz_cost(m, x);
        return c;
    };
    int best_cost = cost(best);
    uint64_t st = 0x9E3779B97F4A7C15ull;  // fixed seed: the same pass always gets the same swizzle
    for (int it = 0; it < 2048 && best_cost > 0; ++it) {
        Swz c;
        for (int b = 0; b < 4; ++b) {
            st ^= st << 13;
            st ^= st >> 7;
            st ^= st << 17;
            c.row[b] = uint32_t(st) & 0xFFFu;
        }
        if (!swz_invertible(c)) continue;
        const int k = cost(c);
        if (k < best_cost) {
            best = c;
            best_cost = k;
        }
    }
    return best;
}

struct Src {
    std::string s;
    // Appends one formatted line of any length (a pass's list of gradient slots can run to
    // thousands of characters).
    void line(const char* fmt, ...) __attribute__((format(printf, 2, 3))) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        va_list ap2;
        va_copy(ap2, ap);
        const int len = vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        if (len >= int(sizeof buf)) {
            std::string big(size_t(len) + 1, '\0');
            vsnprintf(&big[0], big.size(), fmt, ap2);
            big.resize(size_t(len));
            s += big;
        } else if (len > 0) {
            s += buf;
        }
        va_end(ap2);
        s += '\n';
    }
};

}  // namespace

int jit_mode() {
    int m = g_mode.load();
    if (m < 0) {
        // LLSV_PROGRAM_JIT = 0 (never), 1 (automatic, the default) or 2 (always)
        const char* e = std::getenv("LLSV_PROGRAM_JIT");
        m = (e && *e >= '0' && *e <= '2' && !e[1]) ? (*e - '0') : int(kJitAuto);
        int expect = -1;
        g_mode.compare_exchange_strong(expect, m);
        m = g_mode.load();
    }
    return m;
}

int set_jit_mode(int mode) {
    const int prev = jit_mode();
    g_mode.store(mode);
    return prev;
}

bool jit_wanted(int n, int nops, bool dense) {
    const int m = jit_mode();
    if (m == kJitOff || dense || nops > kJitMaxOps) return false;
    return m == kJitAlways || n >= kJitMinQubits;
}

namespace {

// The arithmetic form of an RO_1Q matrix: 2 for the RY pattern [[p, -q], [q, p]] (p, q real), 1 for
// the RX pattern [[p, iq], [iq, p]], 0 for a general 2x2.
int rotation_form(const double* m) {
    if (m[1] == 0.0 && m[3] == 0.0 && m[5] == 0.0 && m[7] == 0.0 && m[0] == m[6] && m[2] == -m[4]) return 2;
    if (m[1] == 0.0 && m[2] == 0.0 && m[4] == 0.0 && m[7] == 0.0 && m[0] == m[6] && m[3] == m[5]) return 1;
    return 0;
}

// What a generated kernel looks like around the op sequence.
struct Shape {
    bool batched = false;  // blockIdx.y = batch element: states A + b * size, coefficients T + b * tstride
    bool two = false;      // a second state (lambda) goes through every op; RO_GRAD ops reduce into partials
};

// Every generated kernel asks for two waves per SIMD (__launch_bounds__(256, 2)): with the 64 KiB
// exchange buffer two workgroups fit a CU, and a kernel allowed more than 256 registers per lane
// (the two-state backward passes reach ~300) would run one.
std::string emit(const RegOp* ops, int nops, const Shape& sh) {
    const size_t stride = sizeof(RegOp) / sizeof(double), moff = offsetof(RegOp, m) / sizeof(double);
    const char* arrs[2] = {"a", "l"};
    const int na = sh.two ? 2 : 1;
    std::vector<int> gslot;  // gradient slots of this pass, in order of appearance
    for (int i = 0; i < nops; ++i)
        if (ops[i].kind == RO_GRAD) gslot.push_back(ops[i].moff);
    Src o;
    o.s.reserve(4096 + size_t(nops) * 600 * na);
    if (!sh.batched) {
        o.line("extern \"C\" __global__ __launch_bounds__(256, 2) void LLSV_KERNEL(v2d* __restrict__ psi, PassDesc d, "
               "const double* __restrict__ M) {");
    } else {
        o.line("extern \"C\" __global__ __launch_bounds__(256, 2) void LLSV_KERNEL(v2d* __restrict__ A, v2d* __restrict__ Lm, "
               "unsigned long long size, PassDesc d, const double* __restrict__ T, unsigned long long tstride, "
               "double* __restrict__ partials, int nslot, int ntiles) {");
        o.line("v2d* const psi = A + (unsigned long long)blockIdx.y * size;");
        if (sh.two) o.line("v2d* const lam = Lm + (unsigned long long)blockIdx.y * size;");
        o.line("const double* const M = T + (unsigned long long)blockIdx.y * tstride;");
    }
    o.line("__shared__ v2d tile[4096];");
    if (!gslot.empty()) o.line("__shared__ double gs[%zu][4];", gslot.size());
    o.line("char* const lds = (char*)tile;");
    o.line("const unsigned int tid = threadIdx.x;");
    // tile of this workgroup: with a multiple of 8 tiles, XCD k (workgroups k, k + 8, ...) takes
    // the k-th eighth of them in order; gradient partials are stored per tile
    o.line("const unsigned long long tb = (gridDim.x & 7u) ? blockIdx.x : (blockIdx.x & 7u) * (gridDim.x >> 3) + "
           "(blockIdx.x >> 3);");
    o.line("unsigned long long base = tb;");
    o.line("for (int j = 0; j < 12; ++j) base = ((base >> d.pos[j]) << (d.pos[j] + 1)) | (base & ((1ull << d.pos[j]) - 1ull));");
    o.line("unsigned long long lg[12];");
    o.line("#pragma unroll");
    o.line("for (int j = 0; j < 12; ++j) lg[j] = j < d.L ? (1ull << j) : (1ull << d.tq[j]);");
    for (int k = 0; k < na; ++k) o.line("v2d %s[16];", arrs[k]);
    o.line("{ unsigned long long gt = 0;");
    o.line("#pragma unroll");
    o.line("for (int j = 0; j < 8; ++j) if ((tid >> j) & 1u) gt |= lg[j];");
    o.line("#pragma unroll");
    o.line("for (int j = 0; j < 16; ++j) { const unsigned long long o = base | gt | ((j & 1) ? lg[8] : 0ull) | "
           "((j & 2) ? lg[9] : 0ull) | ((j & 4) ? lg[10] : 0ull) | ((j & 8) ? lg[11] : 0ull);");
    o.line("a[j] = __builtin_nontemporal_load(psi + o);%s } }", sh.two ? " l[j] = __builtin_nontemporal_load(lam + o);" : "");
    const Swz sw = choose_swizzle(ops, nops);
    o.line("unsigned int lt = tid, sl = r_swzm(tid, %uu, %uu, %uu, %uu) << 4;", sw.row[0], sw.row[1], sw.row[2], sw.row[3]);
    o.line("double sr = 1.0, si = 0.0;");

    uint32_t S = 0xF00u;  // register bits of the current layout (the kernel starts with 8..11)
    bool scaled = false;
    // RO_1QT diagonals (RZ) defer what they can into the lane's factor (sr, si), which every
    // amplitude of the lane shares: on a lane bit the whole diagonal entry, on a register bit d0
    // (the registers with the bit set take d1 / d0).  A factor that differs between lanes goes onto the
    // amplitudes before they change lanes (lane_factor marks it); the pass's last step applies the
    // rest.  Inner products of the gradients see |factor|^2 = 1 (unit-modulus entries, the same
    // factor on psi and lambda), so they read the deferred state as it is.
    bool lane_factor = false;
    int gi = 0;
    for (int i = 0; i < nops; ++i) {
        const RegOp& op = ops[i];
        const size_t c = size_t(i) * stride + moff;
        switch (op.kind) {
            case RO_1Q:
            case RO_1QT: {
                const int K = 1 << op.a;
                // RO_1QT carries its form; an RO_1Q matrix of the RX or RY pattern (zeros and equal
                // entries where the rotation has them) runs the 8-operation form, which rounds exactly
                // like the general update on such a matrix (the dropped terms are products with zero)
                const int form = op.kind == RO_1QT ? op.q[0] : rotation_form(op.m);
                const size_t t8 = size_t(8) * size_t(op.moff);
                if (op.kind == RO_1Q)
                    o.line("{ const double* m = M + %zu;", c);
                else
                    o.line("{ const double* m = M + %zu;", t8);
                if (form == 3 && op.kind == RO_1QT) {
                    scaled = true;
                    if (op.tmask) {
                        o.line("const bool hi = (lt & %uu) != 0u;", op.tmask);
                        o.line("r_phase(sr, si, hi ? m[6] : m[0], hi ? m[7] : m[1]);");
                        lane_factor = true;
                    } else {
                        // the form-3 coefficients are RZ's, |d0| = 1: d1 / d0 = d1 conj(d0)
                        o.line("const double rr = __builtin_fma(m[6], m[0], m[7] * m[1]), "
                               "ri = __builtin_fma(m[7], m[0], -(m[6] * m[1]));");
                        for (int k = 0; k < na; ++k)
                            for (int j = 0; j < 16; ++j)
                                if ((op.rmask >> j) & 1u) o.line("r_sc(%s[%d], rr, ri);", arrs[k], j);
                        o.line("r_phase(sr, si, m[0], m[1]);");
                    }
                } else if (form == 3) {
                    if (op.tmask) {
                        o.line("const bool hi = (lt & %uu) != 0u;", op.tmask);
                        o.line("const double dr = hi ? m[6] : m[0], di = hi ? m[7] : m[1];");
                        for (int k = 0; k < na; ++k)
                            for (int j = 0; j < 16; ++j) o.line("r_sc(%s[%d], dr, di);", arrs[k], j);
                    } else {
                        for (int k = 0; k < na; ++k)
                            for (int j = 0; j < 16; ++j)
                                o.line("r_sc(%s[%d], m[%d], m[%d]);", arrs[k], j, (op.rmask >> j) & 1u ? 6 : 0,
                                       (op.rmask >> j) & 1u ? 7 : 1);
                    }
                } else if ((form == 1 || form == 2) && op.kind == RO_1QT) {
                    // table coefficients of RX / RY are the shears of the angle folded into
                    // [-pi/2, pi/2] (the sign this drops is the same on psi and lambda)
                    o.line("const double sa = m[0], sb = m[1];");
                    for (int k = 0; k < na; ++k)
                        for (int j = 0; j < 16; ++j)
                            if (!(j & K))
                                o.line("%s(%s[%d], %s[%d], sa, sb);", form == 1 ? "r_shx" : "r_shy", arrs[k], j, arrs[k],
                                       j | K);
                } else if (form == 1 || form == 2) {
                    o.line("const double p = m[0], q = m[%d];", form == 1 ? 3 : 4);
                    for (int k = 0; k < na; ++k)
                        for (int j = 0; j < 16; ++j)
                            if (!(j & K))
                                o.line("%s(%s[%d], %s[%d], p, q);", form == 1 ? "r_rx" : "r_ry", arrs[k], j, arrs[k],
                                       j | K);
                } else {
                    for (int k = 0; k < na; ++k)
                        for (int j = 0; j < 16; ++j)
                            if (!(j & K)) o.line("r_pu(%s[%d], %s[%d], m);", arrs[k], j, arrs[k], j | K);
                }
                o.line("}");
                break;
            }
            case RO_1QN: {
                const int K = 1 << op.a;
                o.line("{ const double* m = M + %zu; const double br = m[0], bi = m[1], cr = m[2], ci = m[3];", c);
                for (int k = 0; k < na; ++k)
                    for (int j = 0; j < 16; ++j)
                        if (!(j & K)) o.line("r_nu%d(%s[%d], %s[%d], br, bi, cr, ci);", op.q[0], arrs[k], j, arrs[k], j | K);
                o.line("}");
                break;
            }
            case RO_RSCALE:
                o.line("{ const double f = M[%zu];", c);
                for (int k = 0; k < na; ++k)
                    for (int j = 0; j < 16; ++j) o.line("%s[%d] *= f;", arrs[k], j);
                o.line("}");
                break;
            case RO_GRAD: {
                // Im <lambda| G |psi> over this lane's pairs along register bit a (G = X, Y or Z on it)
                // Every product goes into a chain of FMAs (four per amplitude pair), alternating
                // between two accumulators so that consecutive FMAs do not wait on each other.
                const int K = 1 << op.a;
                o.line("{ double g0 = 0.0, g1 = 0.0;");
                auto fma4 = [&](int acc, int p, char cp, int q, char cq, bool neg0, int r, char cr, int s, char cs,
                                bool neg1, int t, char ct, int v, char cv, bool neg2, int w, char cw, int x, char cx,
                                bool neg3) {
                    const char* g = acc ? "g1" : "g0";
                    o.line("%s = __builtin_fma(%sl[%d].%c, a[%d].%c, %s); %s = __builtin_fma(%sl[%d].%c, a[%d].%c, %s); "
                           "%s = __builtin_fma(%sl[%d].%c, a[%d].%c, %s); %s = __builtin_fma(%sl[%d].%c, a[%d].%c, %s);",
                           g, neg0 ? "-" : "", p, cp, q, cq, g, g, neg1 ? "-" : "", r, cr, s, cs, g, g, neg2 ? "-" : "", t,
                           ct, v, cv, g, g, neg3 ? "-" : "", w, cw, x, cx, g);
                };
                if (op.q[0] == 2 && op.tmask) {
                    // Z on a lane bit: Im conj(l) a over all 16, the sign is the lane's
                    for (int j = 0; j < 16; j += 2)
                        fma4((j >> 1) & 1, j, 'x', j, 'y', false, j, 'y', j, 'x', true, j + 1, 'x', j + 1, 'y', false, j + 1,
                             'y', j + 1, 'x', true);
                    o.line("double g = g0 + g1;");
                    o.line("if ((lt & %uu) != 0u) g = -g;", op.tmask);
                } else {
                    int pair = 0;
                    for (int j = 0; j < 16; ++j) {
                        if (j & K) continue;
                        const int u = j | K, acc = pair++ & 1;
                        if (op.q[0] == 2)  // Im conj(l_j) a_j - Im conj(l_u) a_u
                            fma4(acc, j, 'x', j, 'y', false, j, 'y', j, 'x', true, u, 'x', u, 'y', true, u, 'y', u, 'x', false);
                        else if (op.q[0] == 0)  // Im conj(l_j) a_u + Im conj(l_u) a_j
                            fma4(acc, j, 'x', u, 'y', false, j, 'y', u, 'x', true, u, 'x', j, 'y', false, u, 'y', j, 'x', true);
                        else  // Re conj(l_u) a_j - Re conj(l_j) a_u
                            fma4(acc, u, 'x', j, 'x', false, u, 'y', j, 'y', false, j, 'x', u, 'x', true, j, 'y', u, 'y', true);
                    }
                    o.line("double g = g0 + g1;");
                }
                o.line("g = r_wsum(g);");
                o.line("if ((tid & 63u) == 0u) gs[%d][tid >> 6] = g; }", gi);
                ++gi;
                break;
            }
            case RO_CNOT: {
                const int K = 1 << op.a;
                if (op.tmask == 0) {
                    for (int k = 0; k < na; ++k)
                        for (int j = 0; j < 16; ++j)
                            if (!(j & K) && ((op.rmask >> j) & 1u))
                                o.line("{ const v2d t = %s[%d]; %s[%d] = %s[%d]; %s[%d] = t; }", arrs[k], j, arrs[k], j,
                                       arrs[k], j | K, arrs[k], j | K);
                } else {
                    o.line("{ const bool tk = (lt & %uu) != 0u;", op.tmask);
                    for (int k = 0; k < na; ++k)
                        for (int j = 0; j < 16; ++j)
                            if (!(j & K))
                                o.line("{ const v2d x = %s[%d], y = %s[%d]; %s[%d] = tk ? y : x; %s[%d] = tk ? x : y; }",
                                       arrs[k], j, arrs[k], j | K, arrs[k], j, arrs[k], j | K);
                    o.line("}");
                }
                break;
            }
            case RO_DIAG: {
                if (op.tmask == 0) {
                    for (int k = 0; k < na; ++k)
                        for (int j = 0; j < 16; ++j) {
                            if ((op.rmask >> j) & 1u)
                                o.line("r_sc(%s[%d], M[%zu], M[%zu]);", arrs[k], j, c + 6, c + 7);
                            else if (!op.a)
                                o.line("r_sc(%s[%d], M[%zu], M[%zu]);", arrs[k], j, c + 0, c + 1);
                        }
                } else {
                    o.line("{ const bool hi = (lt & %uu) != 0u;", op.tmask);
                    o.line("const double dr = hi ? M[%zu] : M[%zu], di = hi ? M[%zu] : M[%zu];", c + 6, c, c + 7, c + 1);
                    o.line("#pragma unroll");
                    o.line("for (int j = 0; j < 16; ++j) { r_sc(a[j], dr, di);%s } }", sh.two ? " r_sc(l[j], dr, di);" : "");
                }
                break;
            }
            case RO_MCZ: {
                if (!op.rmask) break;
                if (op.omask == 0 && op.tmask == 0) {
                    for (int k = 0; k < na; ++k)
                        for (int j = 0; j < 16; ++j)
                            if ((op.rmask >> j) & 1u) o.line("%s[%d] = -%s[%d];", arrs[k], j, arrs[k], j);
                    break;
                }
                std::string cond;
                char buf[128];
                if (op.omask) {
                    snprintf(buf, sizeof buf, "(base & %lluull) == %lluull", (unsigned long long)op.omask,
                             (unsigned long long)op.omask);
                    cond += buf;
                }
                if (op.tmask) {
                    snprintf(buf, sizeof buf, "%s(lt & %uu) == %uu", cond.empty() ? "" : " && ", op.tmask, op.tmask);
                    cond += buf;
                }
                o.line("{ const unsigned int sg = (%s) ? 0x80000000u : 0u;", cond.c_str());
                for (int k = 0; k < na; ++k)
                    for (int j = 0; j < 16; ++j)
                        if ((op.rmask >> j) & 1u) o.line("%s[%d] = r_ng(%s[%d], sg);", arrs[k], j, arrs[k], j);
                o.line("}");
                break;
            }
            case RO_SCALE:
                scaled = true;
                o.line("{ const bool hi = (base >> %d) & 1ull; r_phase(sr, si, hi ? M[%zu] : M[%zu], hi ? M[%zu] : M[%zu]); }",
                       op.a, c + 6, c, c + 7, c + 1);
                break;
            case RO_LAYOUT: {
                if (lane_factor) {
                    // the lanes' factors go onto their amplitudes before these change lanes
                    o.line("#pragma unroll");
                    o.line("for (int j = 0; j < 16; ++j) { r_sc(a[j], sr, si);%s }", sh.two ? " r_sc(l[j], sr, si);" : "");
                    o.line("sr = 1.0; si = 0.0;");
                    lane_factor = false;
                }
                // one state at a time through the 64 KiB exchange buffer
                o.line("{ const unsigned int nl = r_ins0(r_ins0(r_ins0(r_ins0(tid, %d), %d), %d), %d);", op.q[0], op.q[1],
                       op.q[2], op.q[3]);
                o.line("const unsigned int nsl = r_swzm(nl, %uu, %uu, %uu, %uu) << 4;", sw.row[0], sw.row[1], sw.row[2],
                       sw.row[3]);
                const uint32_t Snew = (1u << op.q[0]) | (1u << op.q[1]) | (1u << op.q[2]) | (1u << op.q[3]);
                for (int k = 0; k < na; ++k) {
                    if (k) o.line("__syncthreads();");
                    for (int j = 0; j < 16; ++j)
                        o.line("*(v2d*)(lds + (sl ^ %uu)) = %s[%d];", swz_apply(sw, reg_bits(S, j)) << 4, arrs[k], j);
                    o.line("__syncthreads();");
                    for (int j = 0; j < 16; ++j)
                        o.line("%s[%d] = *(const v2d*)(lds + (nsl ^ %uu));", arrs[k], j, swz_apply(sw, reg_bits(Snew, j)) << 4);
                }
                o.line("lt = nl; sl = nsl;");
                o.line("__syncthreads(); }");
                S = (1u << op.q[0]) | (1u << op.q[1]) | (1u << op.q[2]) | (1u << op.q[3]);
                break;
            }
            default:
                break;  // dense ops never reach the generator (jit_wanted)
        }
    }

    int rq[4], t = 0;
    for (int bit = 0; bit < kRegTileBits; ++bit)
        if ((S >> bit) & 1u) rq[t++] = bit;
    // When the pass ends in its starting layout, the store addresses equal the load addresses and
    // the compiler would keep all of them (64-bit, one per register) live through the pass; the
    // empty asm makes it compute them again here.
    o.line("__asm__ volatile(\"\" : \"+v\"(lt));");
    o.line("{ unsigned long long gt = 0;");
    o.line("#pragma unroll");
    o.line("for (int j = 0; j < 12; ++j) if ((lt >> j) & 1u) gt |= lg[j];");
    // one state at a time (scale, then store): scaling both before the stores keeps two versions of
    // all 32 amplitudes of a two-state pass live across the branch, past the 256-register budget
    const char* dst[2] = {"psi", "lam"};
    for (int k = 0; k < na; ++k) {
        if (scaled) {
            o.line("if (si != 0.0) {");
            o.line("#pragma unroll");
            o.line("for (int j = 0; j < 16; ++j) r_sc(%s[j], sr, si);", arrs[k]);
            o.line("} else if (sr != 1.0) {");
            o.line("#pragma unroll");
            o.line("for (int j = 0; j < 16; ++j) %s[j] *= sr; }", arrs[k]);
        }
        o.line("#pragma unroll");
        o.line("for (int j = 0; j < 16; ++j) { const unsigned long long o = base | gt | ((j & 1) ? lg[%d] : 0ull) | "
               "((j & 2) ? lg[%d] : 0ull) | ((j & 4) ? lg[%d] : 0ull) | ((j & 8) ? lg[%d] : 0ull);",
               rq[0], rq[1], rq[2], rq[3]);
        o.line("__builtin_nontemporal_store(%s[j], %s + o); }", arrs[k], dst[k]);
    }
    o.line("}");
    if (!gslot.empty()) {
        bool contiguous = true;
        for (size_t k = 0; k < gslot.size(); ++k) contiguous = contiguous && gslot[k] == gslot[0] + int(k);
        if (contiguous) {
            // the usual case (slots numbered in sweep order): no table, which a thread would index
            // at run time and the compiler would keep in scratch memory
            o.line("__syncthreads();");
            o.line("for (unsigned int k = tid; k < %zuu; k += 256u)", gslot.size());
            o.line("partials[((unsigned long long)blockIdx.y * nslot + (%du + k)) * ntiles + tb] = "
                   "(gs[k][0] + gs[k][1]) + (gs[k][2] + gs[k][3]);", gslot[0]);
        } else {
            std::string list;
            for (size_t k = 0; k < gslot.size(); ++k) list += (k ? ", " : "") + std::to_string(gslot[k]);
            o.line("{ const int gslot[%zu] = {%s};", gslot.size(), list.c_str());
            o.line("__syncthreads();");
            o.line("for (unsigned int k = tid; k < %zuu; k += 256u)", gslot.size());
            o.line("partials[((unsigned long long)blockIdx.y * nslot + gslot[k]) * ntiles + tb] = "
                   "(gs[k][0] + gs[k][1]) + (gs[k][2] + gs[k][3]); }");
        }
    }
    o.line("}");
    return o.s;
}

}  // namespace

std::string jit_source(const RegOp* ops, int nops) { return emit(ops, nops, Shape{}); }

namespace {
std::mutex g_sig_mu;
std::unordered_map<std::string, JitEntry*>& sig_cache() {
    static auto* c = new std::unordered_map<std::string, JitEntry*>();  // lives to process end
    return *c;
}
}  // namespace

JitEntry* jit_program_entry(const RegOp* ops, int nops) {
    // everything the generated source depends on: per op its kind, bits, masks, coefficient slot
    // and (RO_1Q) arithmetic form; the coefficients themselves are run-time data
    std::string sig;
    sig.reserve(size_t(nops) * 44);
    auto put = [&](const void* p, size_t n) { sig.append(static_cast<const char*>(p), n); };
    for (int i = 0; i < nops; ++i) {
        const RegOp& op = ops[i];
        const int form = op.kind == RO_1Q ? rotation_form(op.m) : -1;
        put(&op.kind, sizeof op.kind);
        put(&op.a, sizeof op.a);
        put(&op.rmask, sizeof op.rmask);
        put(&op.tmask, sizeof op.tmask);
        put(&op.omask, sizeof op.omask);
        put(&op.moff, sizeof op.moff);
        put(op.q, sizeof op.q);
        put(&form, sizeof form);
    }
    {
        std::lock_guard<std::mutex> lock(g_sig_mu);
        auto it = sig_cache().find(sig);
        if (it != sig_cache().end()) return it->second;
    }
    JitEntry* e = jit_entry(jit_source(ops, nops));
    std::lock_guard<std::mutex> lock(g_sig_mu);
    sig_cache().emplace(std::move(sig), e);
    return e;
}

std::string jit_adjoint_source(const RegOp* ops, int nops, bool backward) {
    Shape sh;
    sh.batched = true;
    sh.two = backward;
    return emit(ops, nops, sh);
}

// A code object holds the kernels of one compile batch; modules are loaded per device on first launch.
struct CodeObject {
    std::vector<char> code;
    std::map<int, hipModule_t> mod;  // kept until the process ends
};

struct JitEntry {
    enum : int { kNew, kQueued, kBusy, kReady, kFailed };
    std::atomic<int> state{kNew};
    std::string body;  // kernel source, named LLSV_KERNEL
    std::string name;  // symbol in obj
    std::shared_ptr<CodeObject> obj;
    std::string error;
    std::map<int, hipFunction_t> fn;  // per device
};

namespace {

std::mutex g_mu;
std::condition_variable g_cv;
std::deque<JitEntry*> g_queue;
int g_busy = 0;
bool g_stop = false;
std::thread* g_worker = nullptr;
uint64_t g_uid = 0;

std::unordered_map<std::string, std::unique_ptr<JitEntry>>& cache() {
    static auto* c = new std::unordered_map<std::string, std::unique_ptr<JitEntry>>();  // lives to process end
    return *c;
}

// The process keeps its own cache of compiled passes; keep the code-object manager from writing
// one to disk.  Set once when the library is loaded, before it starts any thread (setenv must not
// race with getenv on another thread, and the compiler worker reads the environment).
// LLSV_JIT_DUMP=<directory>: every compile batch's source and code object are also written
// there, as <first kernel name>.hip and .co, for reading the generated passes.
std::string* g_dump_dir = nullptr;
__attribute__((constructor)) void jit_env_init() {
    setenv("AMD_COMGR_CACHE", "0", 0);
    const char* d = std::getenv("LLSV_JIT_DUMP");
    if (d && *d) g_dump_dir = new std::string(d);
}

void dump_file(const std::string& name, const char* data, size_t size) {
    if (!g_dump_dir) return;
    const std::string path = *g_dump_dir + "/" + name;
    if (FILE* f = std::fopen(path.c_str(), "wb")) {
        std::fwrite(data, 1, size, f);
        std::fclose(f);
    }
}

// Compiles the entries (state kBusy, set by the caller) as one hipRTC program, without the lock.
void compile_batch(const std::vector<JitEntry*>& batch) {
    std::string src = kPrelude;
    std::vector<std::string> names;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        for (size_t i = 0; i < batch.size(); ++i) names.push_back("llsv_pass_" + std::to_string(g_uid++));
    }
    for (size_t i = 0; i < batch.size(); ++i) {
        std::string body = batch[i]->body;
        body.replace(body.find("LLSV_KERNEL"), 11, names[i]);
        src += body;
    }
    dump_file(names.front() + ".hip", src.data(), src.size());
    auto obj = std::make_shared<CodeObject>();
    std::string error;
    hiprtcProgram prog = nullptr;
    hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "llsv_pass.hip", 0, nullptr, nullptr);
    if (r == HIPRTC_SUCCESS) {
        // CodeGenPrepare and the SLP vectorizer take half the compile time of these long
        // straight-line kernels and change nothing in the code they produce.
        const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-fno-slp-vectorize", "-mllvm", "-disable-cgp",
                              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
        r = hiprtcCompileProgram(prog, int(sizeof opts / sizeof opts[0]), opts);
        size_t n = 0;
        if (r == HIPRTC_SUCCESS && hiprtcGetCodeSize(prog, &n) == HIPRTC_SUCCESS && n) {
            obj->code.resize(n);
            r = hiprtcGetCode(prog, obj->code.data());
        } else if (r == HIPRTC_SUCCESS) {
            r = HIPRTC_ERROR_COMPILATION;
        }
        if (r != HIPRTC_SUCCESS) {
            size_t ls = 0;
            if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
                std::string log(ls, '\0');
                if (hiprtcGetProgramLog(prog, &log[0]) == HIPRTC_SUCCESS) error = log.substr(0, 4000);
            }
        }
        (void)hiprtcDestroyProgram(&prog);
    }
    if (r != HIPRTC_SUCCESS) error = std::string("hipRTC: ") + hiprtcGetErrorString(r) + (error.empty() ? "" : ": " + error);
    else dump_file(names.front() + ".co", obj->code.data(), obj->code.size());
    std::lock_guard<std::mutex> lock(g_mu);
    for (size_t i = 0; i < batch.size(); ++i) {
        JitEntry* e = batch[i];
        if (r == HIPRTC_SUCCESS) {
            e->obj = obj;
            e->name = names[i];
            e->state = JitEntry::kReady;
            ++g_compiled;
        } else {
            e->error = error;
            e->state = JitEntry::kFailed;
            ++g_failures;
        }
    }
    g_busy -= int(batch.size());
    g_cv.notify_all();
}

void stop_worker() {
    {
        std::lock_guard<std::mutex> lock(g_mu);
        g_stop = true;
    }
    g_cv.notify_all();
    if (g_worker && g_worker->joinable()) g_worker->join();
}

void worker_loop() {
    for (;;) {
        std::vector<JitEntry*> batch;
        {
            std::unique_lock<std::mutex> lock(g_mu);
            g_cv.wait(lock, [] { return g_stop || !g_queue.empty(); });
            if (g_stop) {
                for (JitEntry* e : g_queue) e->state = JitEntry::kNew;
                g_queue.clear();
                g_cv.notify_all();
                return;
            }
            while (!g_queue.empty() && batch.size() < 64) {
                JitEntry* e = g_queue.front();
                g_queue.pop_front();
                e->state = JitEntry::kBusy;
                batch.push_back(e);
            }
            g_busy += int(batch.size());
        }
        compile_batch(batch);
    }
}

}  // namespace

JitEntry* jit_entry(const std::string& body) {
    std::lock_guard<std::mutex> lock(g_mu);
    auto& c = cache();
    auto it = c.find(body);
    if (it == c.end()) {
        auto e = std::make_unique<JitEntry>();
        e->body = body;
        it = c.emplace(body, std::move(e)).first;
    }
    return it->second.get();
}

bool jit_ready(const JitEntry* e) { return e && e->state.load() == JitEntry::kReady; }

void jit_request(const std::vector<JitEntry*>& entries, bool now) {
    std::vector<JitEntry*> mine;
    {
        std::unique_lock<std::mutex> lock(g_mu);
        for (JitEntry* e : entries) {
            const int st = e->state.load();
            if (st == JitEntry::kNew || (now && st == JitEntry::kQueued)) {
                if (now) {
                    if (st == JitEntry::kQueued) g_queue.erase(std::find(g_queue.begin(), g_queue.end(), e));
                    e->state = JitEntry::kBusy;
                    mine.push_back(e);
                } else if (!g_stop) {
                    e->state = JitEntry::kQueued;
                    g_queue.push_back(e);
                }
            }
        }
        g_busy += int(mine.size());
        if (!now && !g_queue.empty() && !g_worker && !g_stop) {
            g_worker = new std::thread(worker_loop);
            std::atexit(stop_worker);
        }
    }
    g_cv.notify_all();
    if (!now) return;
    if (!mine.empty()) compile_batch(mine);
    // entries the worker is compiling right now
    std::unique_lock<std::mutex> lock(g_mu);
    g_cv.wait(lock, [&] {
        for (JitEntry* e : entries)
            if (e->state.load() == JitEntry::kBusy) return false;
        return true;
    });
}

void jit_wait() {
    std::unique_lock<std::mutex> lock(g_mu);
    g_cv.wait(lock, [] { return g_queue.empty() && g_busy == 0; });
}

namespace {

// The entry's kernel on the current device (its module is loaded on first use).
hipError_t entry_function(JitEntry* e, hipFunction_t* out) {
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = e->fn.find(dev);
    if (it != e->fn.end()) {
        *out = it->second;
        return hipSuccess;
    }
    CodeObject& obj = *e->obj;
    auto mi = obj.mod.find(dev);
    if (mi == obj.mod.end()) {
        hipModule_t mod = nullptr;
        err = hipModuleLoadData(&mod, obj.code.data());
        if (err != hipSuccess) return err;
        mi = obj.mod.emplace(dev, mod).first;
    }
    hipFunction_t fn = nullptr;
    err = hipModuleGetFunction(&fn, mi->second, e->name.c_str());
    if (err != hipSuccess) return err;
    e->fn[dev] = fn;
    *out = fn;
    return hipSuccess;
}

}  // namespace

hipError_t jit_launch(JitEntry* e, unsigned grid, hipStream_t st, v2d* psi, const PassDesc& d, const RegOp* dops) {
    hipFunction_t fn = nullptr;
    hipError_t err = entry_function(e, &fn);
    if (err != hipSuccess) return err;
    PassDesc dd = d;
    const double* M = reinterpret_cast<const double*>(dops);
    void* args[] = {&psi, &dd, &M};
    err = hipModuleLaunchKernel(fn, grid, 1, 1, 256, 1, 1, 0, st, args, nullptr);
    if (err == hipSuccess) ++g_launches;
    return err;
}

hipError_t jit_launch_batched(JitEntry* e, unsigned tiles, unsigned batch, hipStream_t st, v2d* A, v2d* Lm,
                              uint64_t size, const PassDesc& d, const double* T, uint64_t tstride, double* partials,
                              int nslot, int ntiles) {
    hipFunction_t fn = nullptr;
    hipError_t err = entry_function(e, &fn);
    if (err != hipSuccess) return err;
    PassDesc dd = d;
    unsigned long long sz = size, ts = tstride;
    void* args[] = {&A, &Lm, &sz, &dd, &T, &ts, &partials, &nslot, &ntiles};
    err = hipModuleLaunchKernel(fn, tiles, batch, 1, 256, 1, 1, 0, st, args, nullptr);
    if (err == hipSuccess) ++g_launches;
    return err;
}

void jit_stats(uint64_t* compiled, uint64_t* launches, uint64_t* failures) {
    if (compiled) *compiled = g_compiled.load();
    if (launches) *launches = g_launches.load();
    if (failures) *failures = g_failures.load();
}

}  // namespace llsv
