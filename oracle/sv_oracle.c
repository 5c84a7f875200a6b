/*
 * sv_oracle.c -- CPU restatement of the reference state-vector semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * entry and the CPU-baseline leg of the benchmark may load it.  The product
 * (libllsv, quantum_sim_v3/v2) never links or calls it.
 *
 * Layout: a state of n qubits is 2^n complex128 amplitudes stored as
 * interleaved doubles (re, im); qubit q is bit q of the basis index
 * (little endian, src/Backend.h:3-4).
 *
 * Build: gcc -O3 -std=c11 -ffp-contract=off -fopenmp -shared -fPIC.
 * -ffp-contract=off matters: the reference's own build line (README:126-133) has
 * no -march, so on x86-64 nothing is fused into FMAs; with contraction off the
 * restated products below round exactly like that build of
 * std::complex<double> operator* for finite operands, on any host.
 *
 * Every function returns 0 on success and a negative code on a bad argument
 * (the reference throws std::invalid_argument in the same places).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define ORC_EINVAL (-1)

typedef struct { double re, im; } cplx;

/* (a + ib)(c + id) = (ac - bd) + i(ad + bc): the finite-operand branch of
 * GCC's complex multiply used by std::complex<double>::operator*. */
static inline cplx cmul(cplx a, cplx b) {
    cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}
static inline cplx cadd(cplx a, cplx b) {
    cplx r;
    r.re = a.re + b.re;
    r.im = a.im + b.im;
    return r;
}

/* (cos x, sin x) as the reference's compiled gate constructors get them.  RX/RY/RZ
 * evaluate std::cos and std::sin of the same argument (src/Backend.h:75, :81, :87 via
 * std::polar); GCC at -O2/-O3 folds such a pair into one sincos() call, whose results
 * differ from separate cos()/sin() calls in the last bit for some arguments. */
void orc_sincos(double x, double* out) {
    double s, c;
    sincos(x, &s, &c);
    out[0] = c;
    out[1] = s;
}

/* |0...0>: src/StateVectorOptimized.cpp:17-24. */
int orc_init_zero(double* psi, int n) {
    if (n < 0 || n > 40) return ORC_EINVAL;
    const uint64_t size = (uint64_t)1 << n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)(2 * size); ++i) psi[i] = 0.0;
    psi[0] = 1.0;
    return 0;
}

/* One-qubit gate, restating StateVectorOptimized::apply_gate_cache_optimized
 * (src/StateVectorOptimized.cpp:56-78): blocks of 2^(k+1) amplitudes, each
 * block walked contiguously over its 2^k pairs, OpenMP static over blocks
 * (so a target of n-1 runs on one thread, as in the reference).
 * a0' = g0 a0 + g1 a1 ; a1' = g2 a0 + g3 a1, gate row-major
 * g = {g00re, g00im, g01re, g01im, g10re, g10im, g11re, g11im}.
 * The reference does not check target; a target outside [0, n) is rejected here. */
int orc_apply_1q(double* psi, int n, int target, const double* g) {
    if (target < 0 || target >= n) return ORC_EINVAL;
    cplx* s = (cplx*)psi;
    const cplx g0 = {g[0], g[1]}, g1 = {g[2], g[3]}, g2 = {g[4], g[5]}, g3 = {g[6], g[7]};
    const uint64_t size = (uint64_t)1 << n;
    const uint64_t stride = (uint64_t)1 << target;
    const uint64_t num_blocks = size >> (target + 1);
    const uint64_t block_size = stride << 1;
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < (int64_t)num_blocks; ++b) {
        const uint64_t base = (uint64_t)b * block_size;
        for (uint64_t j = 0; j < stride; ++j) {
            const uint64_t i0 = base + j, i1 = i0 + stride;
            const cplx a0 = s[i0], a1 = s[i1];
            s[i0] = cadd(cmul(g0, a0), cmul(g1, a1));
            s[i1] = cadd(cmul(g2, a0), cmul(g3, a1));
        }
    }
    return 0;
}

/* Same update over a bit-mask loop (src/StateVectorOptimized.cpp:81-97 and the
 * v2 apply_gate_optimized, src/bindings_v2.cpp:56-72).  Identical results. */
int orc_apply_1q_standard(double* psi, int n, int target, const double* g) {
    if (target < 0 || target >= n) return ORC_EINVAL;
    cplx* s = (cplx*)psi;
    const cplx g0 = {g[0], g[1]}, g1 = {g[2], g[3]}, g2 = {g[4], g[5]}, g3 = {g[6], g[7]};
    const uint64_t size = (uint64_t)1 << n, mask = (uint64_t)1 << target;
#pragma omp parallel for
    for (int64_t ii = 0; ii < (int64_t)size; ++ii) {
        const uint64_t i = (uint64_t)ii;
        if ((i & mask) == 0) {
            const cplx a0 = s[i], a1 = s[i | mask];
            s[i] = cadd(cmul(g0, a0), cmul(g1, a1));
            s[i | mask] = cadd(cmul(g2, a0), cmul(g3, a1));
        }
    }
    return 0;
}

/* Insert a zero bit at position `pos` of x. */
static inline uint64_t insert0(uint64_t x, int pos) {
    const uint64_t lo = x & (((uint64_t)1 << pos) - 1);
    return ((x >> pos) << (pos + 1)) | lo;
}

/* CNOT: swap psi[i] <-> psi[i | t] for every i with control bit 1 and target
 * bit 0 (src/Backend.h:90-97).  Work split as in StateVectorBackend::cnot
 * (src/StateVectorBackend.h:109-135): 2^(n-2) swaps, OpenMP over them. */
int orc_apply_cnot(double* psi, int n, int control, int target) {
    if (control < 0 || control >= n || target < 0 || target >= n || control == target) return ORC_EINVAL;
    cplx* s = (cplx*)psi;
    const int lo = control < target ? control : target;
    const int hi = control < target ? target : control;
    const uint64_t cb = (uint64_t)1 << control, tb = (uint64_t)1 << target;
    const int64_t npairs = (int64_t)(((uint64_t)1 << n) >> 2);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npairs; ++p) {
        const uint64_t i = insert0(insert0((uint64_t)p, lo), hi) | cb;
        const cplx t = s[i];
        s[i] = s[i | tb];
        s[i | tb] = t;
    }
    return 0;
}

/* Multi-controlled Z: negate every amplitude whose bits are all 1 on `qs`
 * (src/Backend.h:99-110).  CZ(a, b) = mcz({a, b}) (src/Backend.h:7).
 * Empty list, duplicate or out-of-range qubit -> error. */
int orc_apply_mcz(double* psi, int n, const int* qs, int nq) {
    if (nq <= 0) return ORC_EINVAL;
    uint64_t mask = 0;
    for (int k = 0; k < nq; ++k) {
        if (qs[k] < 0 || qs[k] >= n) return ORC_EINVAL;
        const uint64_t b = (uint64_t)1 << qs[k];
        if (mask & b) return ORC_EINVAL;
        mask |= b;
    }
    cplx* s = (cplx*)psi;
    const int64_t size = (int64_t)((uint64_t)1 << n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < size; ++i)
        if (((uint64_t)i & mask) == mask) {
            s[i].re = -s[i].re;
            s[i].im = -s[i].im;
        }
    return 0;
}

/* Dense k-qubit operator U (2^k x 2^k complex, row-major, interleaved re/im).
 * Operator bit j <-> targets[j] (the convention of src/Noise.h:58-65).
 * For every setting of the other n-k bits: y[r] = sum_{c=0}^{2^k-1} U[r][c] x[c],
 * accumulated in increasing c.  New work (the reference fuses only 2x2). */
int orc_apply_dense(double* psi, int n, int k, const int* targets, const double* m) {
    if (k < 1 || k > 10 || k > n) return ORC_EINVAL;
    uint64_t tmask = 0;
    for (int j = 0; j < k; ++j) {
        if (targets[j] < 0 || targets[j] >= n) return ORC_EINVAL;
        const uint64_t b = (uint64_t)1 << targets[j];
        if (tmask & b) return ORC_EINVAL;
        tmask |= b;
    }
    int sorted[10];
    for (int j = 0; j < k; ++j) sorted[j] = targets[j];
    for (int a = 1; a < k; ++a)
        for (int b = a; b > 0 && sorted[b - 1] > sorted[b]; --b) {
            const int t = sorted[b];
            sorted[b] = sorted[b - 1];
            sorted[b - 1] = t;
        }
    const int D = 1 << k;
    uint64_t off[1024];
    for (int c = 0; c < D; ++c) {
        uint64_t o = 0;
        for (int j = 0; j < k; ++j)
            if ((c >> j) & 1) o |= (uint64_t)1 << targets[j];
        off[c] = o;
    }
    cplx* s = (cplx*)psi;
    const cplx* U = (const cplx*)m;
    const int64_t ngroups = (int64_t)(((uint64_t)1 << n) >> k);
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < ngroups; ++g) {
        uint64_t base = (uint64_t)g;
        for (int j = 0; j < k; ++j) base = insert0(base, sorted[j]);
        cplx x[1024], y;
        for (int c = 0; c < D; ++c) x[c] = s[base | off[c]];
        for (int r = 0; r < D; ++r) {
            y.re = 0.0;
            y.im = 0.0;
            for (int c = 0; c < D; ++c) y = cadd(y, cmul(U[(size_t)r * D + c], x[c]));
            s[base | off[r]] = y;
        }
    }
    return 0;
}

/* sum_i |psi_i|^2 with |a|^2 = re^2 + im^2 (src/StateVectorBackend.h:159). */
double orc_norm2(const double* psi, int n) {
    const int64_t size = (int64_t)((uint64_t)1 << n);
    double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
    for (int64_t i = 0; i < size; ++i) acc += psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
    return acc;
}

/* Probabilities |psi_i|^2 (src/StateVectorBackend.h:153-161). */
int orc_probabilities(const double* psi, int n, double* p) {
    const int64_t size = (int64_t)((uint64_t)1 << n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < size; ++i) p[i] = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
    return 0;
}

/* <Z_q> = P(bit q = 0) - P(bit q = 1). */
double orc_expect_z(const double* psi, int n, int q) {
    const int64_t size = (int64_t)((uint64_t)1 << n);
    const uint64_t b = (uint64_t)1 << q;
    double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
    for (int64_t i = 0; i < size; ++i) {
        const double p = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        acc += ((uint64_t)i & b) ? -p : p;
    }
    return acc;
}

/* <a|b> = sum_i conj(a_i) b_i, written to out[0..1]. */
void orc_inner(const double* a, const double* b, int n, double* out) {
    const int64_t size = (int64_t)((uint64_t)1 << n);
    double re = 0.0, im = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : re, im)
    for (int64_t i = 0; i < size; ++i) {
        const double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
        re += ar * br + ai * bi;
        im += ar * bi - ai * br;
    }
    out[0] = re;
    out[1] = im;
}

/* out += c * P psi for the Pauli string P = prod_{q in xmask} X_q * prod_{q in zmask} Z_q
 * (X applied after Z: P|i> = (-1)^{popcount(i & zmask)} |i ^ xmask>), c a real coefficient. */
void orc_axpy_pauli(const double* psi, int n, uint64_t xmask, uint64_t zmask, double c, double* out) {
    const int64_t size = (int64_t)((uint64_t)1 << n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < size; ++i) {
        const uint64_t src = (uint64_t)i ^ xmask; /* out[i] gets (P psi)[i] = sign(src) psi[src] */
        const double sg = (__builtin_popcountll(src & zmask) & 1) ? -c : c;
        out[2 * i] += sg * psi[2 * src];
        out[2 * i + 1] += sg * psi[2 * src + 1];
    }
}

/* STREAM copy (McCalpin's "Copy" kernel, a[i] = b[i]) over n doubles, OpenMP static schedule,
 * for the host-bandwidth figure reported beside the CPU baseline: 16 bytes moved per element. */
void orc_stream_copy(double* dst, const double* src, uint64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) dst[i] = src[i];
}

/* Parallel first touch of n doubles (zeros), so pages land near the threads that use them. */
void orc_fill_zero(double* dst, uint64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) dst[i] = 0.0;
}
