/*
 * dm_oracle.c -- CPU density-matrix reference for noisy circuits. TEST INFRASTRUCTURE ONLY.
 *
 * Restates KrausChannel::apply (src/Noise.cpp:174-258): a channel with Kraus operators K_k on
 * m target qubits (operator bit j <-> targets[j], src/Noise.h:58-65) updates rho block by block.
 * Write a register index as (rest, t): t = the target bits in operator order, rest = every
 * other bit.  For each pair (r1, r2) of rest values the entries rho[(r1, .), (r2, .)] form a
 * 2^m x 2^m block B, and the update is B -> sum_k K_k B K_k^dagger.  Blocks partition rho,
 * so the update is done in place (the reference writes a second matrix and swaps; every entry
 * is written once either way).  rest_base[r] deposits the bits of r at the non-target
 * positions in ascending order; tgt_off[t] deposits bit j of t at targets[j]
 * (src/Noise.cpp:196-218).  The reference caps n at MAX_DM_QUBITS = 10 (src/Noise.h:19) for
 * its Eigen matrices; the algorithm has no such limit and this restatement runs to n = 14
 * (rho = 4^14 x 16 B = 4 GiB).
 *
 * A unitary gate is the one-operator channel {U}, so rho <- U rho U^dagger uses the same code
 * (CNOT as the 4 x 4 operator on targets {control, target}).
 *
 * rho is row-major dim x dim complex128, interleaved (re, im), dim = 2^n.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cx;

static inline cx cmul(cx a, cx b) { return (cx){a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
/* a * conj(b) */
static inline cx cmulc(cx a, cx b) { return (cx){a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im}; }

#define DM_MAX_M 4

/* B <- sum_k K_k B K_k^dagger for one block (blk x blk, row-major). */
static inline __attribute__((always_inline)) void block_update(int blk, cx* B, int n_ops, const cx* ops) {
    cx acc[(1 << DM_MAX_M) * (1 << DM_MAX_M)], tmp[(1 << DM_MAX_M) * (1 << DM_MAX_M)];
    memset(acc, 0, sizeof(cx) * blk * blk);
    for (int k = 0; k < n_ops; ++k) {
        const cx* K = ops + (size_t)k * blk * blk;
        /* tmp = K B */
        for (int a = 0; a < blk; ++a)
            for (int b = 0; b < blk; ++b) {
                cx s = {0.0, 0.0};
                for (int c = 0; c < blk; ++c) {
                    const cx p = cmul(K[a * blk + c], B[c * blk + b]);
                    s.re += p.re;
                    s.im += p.im;
                }
                tmp[a * blk + b] = s;
            }
        /* acc += tmp K^dagger: (tmp K^dagger)[a][b] = sum_c tmp[a][c] conj(K[b][c]) */
        for (int a = 0; a < blk; ++a)
            for (int b = 0; b < blk; ++b) {
                cx s = {0.0, 0.0};
                for (int c = 0; c < blk; ++c) {
                    const cx p = cmulc(tmp[a * blk + c], K[b * blk + c]);
                    s.re += p.re;
                    s.im += p.im;
                }
                acc[a * blk + b].re += s.re;
                acc[a * blk + b].im += s.im;
            }
    }
    memcpy(B, acc, sizeof(cx) * blk * blk);
}

/* Channels 0..n_ch-1, all on the same targets, applied in order: channel c has n_ops[c]
 * operators, stored one after another in ops.  Every channel maps each block to itself, so the
 * sequence is applied block by block in one sweep over rho; each block sees exactly the
 * arithmetic of n_ch separate sweeps, so the result is bit-identical to them.
 * Returns 0, or -1 for a bad argument (m outside [1, 4], a target outside [0, n) or repeated). */
int orc_dm_apply_channel_seq(int n, double* rho_d, int m, const int* targets, int n_ch, const int* n_ops,
                             const double* ops_d) {
    if (n < 1 || n > 15 || m < 1 || m > DM_MAX_M || m > n || n_ch < 1) return -1;
    for (int c = 0; c < n_ch; ++c)
        if (n_ops[c] < 1) return -1;
    for (int j = 0; j < m; ++j) {
        if (targets[j] < 0 || targets[j] >= n) return -1;
        for (int i = 0; i < j; ++i)
            if (targets[i] == targets[j]) return -1;
    }
    cx* rho = (cx*)rho_d;
    const cx* ops = (const cx*)ops_d;
    const int blk = 1 << m;
    const uint64_t dim = (uint64_t)1 << n;
    const uint64_t n_rest = (uint64_t)1 << (n - m);

    int rest_pos[64];
    int nr = 0;
    for (int q = 0; q < n; ++q) {
        int is_t = 0;
        for (int j = 0; j < m; ++j) is_t |= targets[j] == q;
        if (!is_t) rest_pos[nr++] = q;
    }
    uint64_t* rest_base = (uint64_t*)malloc(sizeof(uint64_t) * n_rest);
    if (!rest_base) return -2;
    for (uint64_t r = 0; r < n_rest; ++r) {
        uint64_t base = 0;
        for (int b = 0; b < nr; ++b)
            if ((r >> b) & 1) base |= (uint64_t)1 << rest_pos[b];
        rest_base[r] = base;
    }
    uint64_t tgt_off[1 << DM_MAX_M];
    for (int t = 0; t < blk; ++t) {
        uint64_t off = 0;
        for (int j = 0; j < m; ++j)
            if ((t >> j) & 1) off |= (uint64_t)1 << targets[j];
        tgt_off[t] = off;
    }

#pragma omp parallel for schedule(static)
    for (int64_t r1 = 0; r1 < (int64_t)n_rest; ++r1) {
        cx B[(1 << DM_MAX_M) * (1 << DM_MAX_M)];
        const uint64_t row0 = rest_base[r1];
        for (uint64_t r2 = 0; r2 < n_rest; ++r2) {
            const uint64_t col0 = rest_base[r2];
            for (int a = 0; a < blk; ++a)
                for (int b = 0; b < blk; ++b) B[a * blk + b] = rho[(row0 | tgt_off[a]) * dim + (col0 | tgt_off[b])];
            const cx* K = ops;
            for (int c = 0; c < n_ch; ++c) {
                /* constant block sizes, so each update is unrolled (same operation order) */
                if (blk == 2)
                    block_update(2, B, n_ops[c], K);
                else if (blk == 4)
                    block_update(4, B, n_ops[c], K);
                else
                    block_update(blk, B, n_ops[c], K);
                K += (size_t)n_ops[c] * blk * blk;
            }
            for (int a = 0; a < blk; ++a)
                for (int b = 0; b < blk; ++b) rho[(row0 | tgt_off[a]) * dim + (col0 | tgt_off[b])] = B[a * blk + b];
        }
    }
    free(rest_base);
    return 0;
}

/* One channel: KrausChannel::apply. */
int orc_dm_apply_channel(int n, double* rho_d, int m, const int* targets, int n_ops, const double* ops_d) {
    return orc_dm_apply_channel_seq(n, rho_d, m, targets, 1, &n_ops, ops_d);
}

/* |0...0><0...0| */
void orc_dm_init_zero(int n, double* rho_d) {
    const uint64_t dim = (uint64_t)1 << n;
    cx* rho = (cx*)rho_d;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)dim; ++r) memset(rho + (uint64_t)r * dim, 0, sizeof(cx) * dim);
    rho[0].re = 1.0;
}

/* z[q] = sum_i rho[i][i] (1 - 2 bit_q(i)) for q < n; *trace = sum_i rho[i][i] (real parts). */
void orc_dm_expect_z(int n, const double* rho_d, double* z, double* trace) {
    const uint64_t dim = (uint64_t)1 << n;
    const cx* rho = (const cx*)rho_d;
    double tr = 0.0;
    for (int q = 0; q < n; ++q) z[q] = 0.0;
    for (uint64_t i = 0; i < dim; ++i) {
        const double p = rho[i * dim + i].re;
        tr += p;
        for (int q = 0; q < n; ++q) z[q] += ((i >> q) & 1) ? -p : p;
    }
    *trace = tr;
}
