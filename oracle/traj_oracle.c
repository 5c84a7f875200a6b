/*
 * traj_oracle.c -- CPU reference for gate-level noisy trajectories. TEST INFRASTRUCTURE ONLY.
 *
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11): 10 rounds of (hi, lo) = mulhi/mullo with multipliers 0xD2511F53, 0xCD9E8D57, key
 * bumped by the Weyl constants 0x9E3779B9, 0xBB67AE85 between rounds.
 *
 * A trajectory is an n-qubit state vector driven through a gate list; after every gate each
 * qubit the gate touched goes through amplitude damping AD(gamma_q) then phase damping
 * PD(lambda_q), unravelled by Kraus sampling with the reference's jump rule
 * (src/TrajectorySolver.h:425): jump iff u >= survive = ||K0 psi||^2 (state normalised), then
 * psi <- K psi / ||K psi||.  Kraus operators (src/Noise.h:114-122):
 *   AD: K0 = diag(1, sqrt(1-g)),  K1 = sqrt(g) |0><1|
 *   PD: K0 = diag(1, sqrt(1-l)),  K1 = diag(0, sqrt(l))
 * Uniforms: u = (x >> 11) * 2^-53 from a 64-bit draw (src/Mitigation.h:206-207); the draw for
 * (trajectory, gate g, site s) is Philox(key = seed, counter = (traj_lo, traj_hi, g, s)),
 * x = out0 | out1 << 32.  Sites: 0, 1 = AD, PD on the gate's first qubit; 2, 3 on the second.
 *
 * The channel decisions of one gate are taken from the joint probabilities of the gate's
 * qubits, updated in closed form after each outcome (AD's jump moves the |1> weight of the
 * damped qubit to |0> scaled by gamma, a no-jump scales it by 1 - gamma; PD likewise), then
 * the product of the chosen operators is applied once and the state renormalised.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        if (r < 9) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

double orc_traj_uniform(uint64_t seed, uint64_t traj, uint32_t gate, uint32_t site) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)traj, (uint32_t)(traj >> 32), gate, site};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    const uint64_t x = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct { double re, im; } cx;

typedef struct {
    int32_t kind; /* 0 = one-qubit gate m, 1 = CNOT(q0 control, q1 target) */
    int32_t q0, q1, pad;
    double m[8];
} orc_tgate;

/* 2x2 complex product c = a b (row-major, interleaved). */
static void mm2(const double* a, const double* b, double* c) {
    double r[8];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            double re = 0, im = 0;
            for (int k = 0; k < 2; ++k) {
                const double ar = a[2 * (2 * i + k)], ai = a[2 * (2 * i + k) + 1];
                const double br = b[2 * (2 * k + j)], bi = b[2 * (2 * k + j) + 1];
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
            }
            r[2 * (2 * i + j)] = re;
            r[2 * (2 * i + j) + 1] = im;
        }
    memcpy(c, r, sizeof r);
}

static void apply2(cx* s, int n, int q, const double* m) {
    const uint64_t N = (uint64_t)1 << n, b = (uint64_t)1 << q;
    for (uint64_t i = 0; i < N; ++i)
        if (!(i & b)) {
            const cx a0 = s[i], a1 = s[i | b];
            s[i].re = m[0] * a0.re - m[1] * a0.im + m[2] * a1.re - m[3] * a1.im;
            s[i].im = m[0] * a0.im + m[1] * a0.re + m[2] * a1.im + m[3] * a1.re;
            s[i | b].re = m[4] * a0.re - m[5] * a0.im + m[6] * a1.re - m[7] * a1.im;
            s[i | b].im = m[4] * a0.im + m[5] * a0.re + m[6] * a1.im + m[7] * a1.re;
        }
}

/* One qubit's AD then PD decisions.  P[a][b]: joint weights with a = this qubit's bit, b = the
 * other qubit's bit (b = 0 only when there is no other qubit).  Updates P in place, writes the
 * unnormalised operator K = PD_k . AD_k into k, returns the number of jumps (0..2) and
 * accumulates the decision hash. */
static int decide_qubit(double P[2][2], double g, double l, uint64_t seed, uint64_t traj, uint32_t gate,
                        uint32_t site0, double* k, uint64_t* hash) {
    int jumps = 0;
    double ad[8] = {0}, pd[8] = {0};
    const double tot = P[0][0] + P[0][1] + P[1][0] + P[1][1];
    const double p1 = (P[1][0] + P[1][1]) / tot;
    const double u = orc_traj_uniform(seed, traj, gate, site0);
    if (u >= 1.0 - g * p1) { /* AD jump: sqrt(g) |0><1| */
        ad[2] = sqrt(g);
        P[0][0] = g * P[1][0];
        P[0][1] = g * P[1][1];
        P[1][0] = P[1][1] = 0.0;
        ++jumps;
        *hash = (*hash ^ ((uint64_t)gate * 4 + site0)) * 0x100000001B3ull;
    } else {
        ad[0] = 1.0;
        ad[6] = sqrt(1.0 - g);
        P[1][0] *= 1.0 - g;
        P[1][1] *= 1.0 - g;
    }
    const double tot2 = P[0][0] + P[0][1] + P[1][0] + P[1][1];
    const double q1 = (P[1][0] + P[1][1]) / tot2;
    const double v = orc_traj_uniform(seed, traj, gate, site0 + 1);
    if (v >= 1.0 - l * q1) { /* PD jump: diag(0, sqrt(l)) */
        pd[6] = sqrt(l);
        P[0][0] = P[0][1] = 0.0;
        P[1][0] *= l;
        P[1][1] *= l;
        ++jumps;
        *hash = (*hash ^ ((uint64_t)gate * 4 + site0 + 1)) * 0x100000001B3ull;
    } else {
        pd[0] = 1.0;
        pd[6] = sqrt(1.0 - l);
        P[1][0] *= 1.0 - l;
        P[1][1] *= 1.0 - l;
    }
    mm2(pd, ad, k);
    return jumps;
}

/* Runs one trajectory.  z_out[q] = <Z_q> of the final state; returns the number of jumps and
 * writes the decision hash.  state_out (2^n complex) receives the final state if not NULL. */
int orc_traj_run(int n, const orc_tgate* gates, int n_gates, const double* gamma, const double* lam, uint64_t seed,
                 uint64_t traj, double* z_out, uint64_t* hash_out, double* state_out) {
    const uint64_t N = (uint64_t)1 << n;
    cx* s = (cx*)calloc(N, sizeof(cx));
    if (!s) return -1;
    s[0].re = 1.0;
    int jumps = 0;
    uint64_t hash = 0xCBF29CE484222325ull;
    for (int gi = 0; gi < n_gates; ++gi) {
        const orc_tgate* g = &gates[gi];
        double P[2][2] = {{0, 0}, {0, 0}};
        if (g->kind == 0) {
            apply2(s, n, g->q0, g->m);
            const uint64_t b = (uint64_t)1 << g->q0;
            for (uint64_t i = 0; i < N; ++i) P[(i & b) ? 1 : 0][0] += s[i].re * s[i].re + s[i].im * s[i].im;
            double k[8];
            jumps += decide_qubit(P, gamma[g->q0], lam[g->q0], seed, traj, (uint32_t)gi, 0, k, &hash);
            const double nrm = 1.0 / sqrt(P[0][0] + P[1][0]);
            for (int j = 0; j < 8; ++j) k[j] *= nrm;
            apply2(s, n, g->q0, k);
        } else {
            const uint64_t cb = (uint64_t)1 << g->q0, tb = (uint64_t)1 << g->q1;
            for (uint64_t i = 0; i < N; ++i)
                if ((i & cb) && !(i & tb)) {
                    const cx t = s[i];
                    s[i] = s[i | tb];
                    s[i | tb] = t;
                }
            for (uint64_t i = 0; i < N; ++i)
                P[(i & cb) ? 1 : 0][(i & tb) ? 1 : 0] += s[i].re * s[i].re + s[i].im * s[i].im;
            double kc[8], kt[8];
            jumps += decide_qubit(P, gamma[g->q0], lam[g->q0], seed, traj, (uint32_t)gi, 0, kc, &hash);
            double Q[2][2] = {{P[0][0], P[1][0]}, {P[0][1], P[1][1]}}; /* transpose: target bit first */
            jumps += decide_qubit(Q, gamma[g->q1], lam[g->q1], seed, traj, (uint32_t)gi, 2, kt, &hash);
            const double nrm = 1.0 / sqrt(Q[0][0] + Q[0][1] + Q[1][0] + Q[1][1]);
            for (int j = 0; j < 8; ++j) kt[j] *= nrm;
            apply2(s, n, g->q0, kc);
            apply2(s, n, g->q1, kt);
        }
    }
    for (int q = 0; q < n; ++q) {
        double z = 0.0;
        const uint64_t b = (uint64_t)1 << q;
        for (uint64_t i = 0; i < N; ++i) {
            const double p = s[i].re * s[i].re + s[i].im * s[i].im;
            z += (i & b) ? -p : p;
        }
        z_out[q] = z;
    }
    if (state_out) memcpy(state_out, s, N * sizeof(cx));
    if (hash_out) *hash_out = hash;
    free(s);
    return jumps;
}
