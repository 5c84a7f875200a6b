"""CPU oracle for zero-noise extrapolation -- TEST INFRASTRUCTURE ONLY.

Restates, over a NumPy state vector / density matrix:
  * the sampled noise model of ll::NoisyBackend (src/Mitigation.h:196-303): gate folding
    U (U^+ U)^k (src/Mitigation.h:226-239), a depolarizing site after every application
    (src/Mitigation.h:254-262: one 1-qubit channel, one 2-qubit channel for a 2-qubit gate,
    independent 1-qubit channels for a wider mcz), draws u01 = (x >> 11) 2^-53 from one
    std::mt19937_64 stream (src/Mitigation.h:265), Pauli choice 1 + int(u 3) or 1 + int(u 15)
    with the low two bits on the first qubit (src/Mitigation.h:267-278), Y applied as Z then X
    (src/Mitigation.h:281-299);
  * the ZNE driver (src/Mitigation.h:340-390: Welford mean and variance per scale, weighted
    standard error) and Richardson weights (src/Mitigation.h:300-314);
  * the exact density-matrix counterpart (src/Mitigation.cpp:54-205, DensityNoiseBackend, and
    zne_exact, src/Mitigation.cpp:213-246) with the depolarizing channel of
    src/Noise.cpp:316-333: rho -> (1 - p) rho + p / (4^m - 1) sum_{P != I} P rho P.

Pinned by (tests/test_zne_oracle.py): the mt19937_64 known answer of the C++ standard (10000th
output for the default seed 5489 = 9981545732273789042); the compiled reference's own
NoisyBackend / ZNE (oracle/_ref/libref_mitigation.so) trajectory by trajectory when built; the
closed forms of the reference's tests (1 qubit: <Z>(lambda) = z0 r1^(g lambda), r1 = 1 - 4p/3;
Bell: <ZZ>(lambda) = r2^lambda, r2 = 1 - 16p/15); hand Richardson weights.
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1


class MT19937_64:
    """std::mt19937_64 ([rand.predef]: w 64, n 312, m 156, r 31, a 0xB5026F5AA96619E9,
    u 29, d 0x5555555555555555, s 17, b 0x71D67FFFEDA60000, t 37, c 0xFFF7EEE000000000, l 43,
    f 6364136223846793005)."""

    N, M = 312, 156
    A = 0xB5026F5AA96619E9
    UPPER, LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed: int = 5489):
        self.seed(seed)

    def seed(self, seed: int):
        mt = [0] * self.N
        mt[0] = seed & MASK64
        for i in range(1, self.N):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & MASK64
        self.mt, self.i = mt, self.N

    def _twist(self):
        mt, N, M = self.mt, self.N, self.M
        for k in range(N):
            y = (mt[k] & self.UPPER) | (mt[(k + 1) % N] & self.LOWER)
            v = mt[(k + M) % N] ^ (y >> 1)
            if y & 1:
                v ^= self.A
            mt[k] = v
        self.i = 0

    def __call__(self) -> int:
        if self.i >= self.N:
            self._twist()
        x = self.mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & MASK64

    def u01(self) -> float:
        return float(self() >> 11) * (1.0 / 9007199254740992.0)


# ---- gates (qubit q = bit q of the index), src/Backend.h:66-110 ----------------------------------

def gate_matrix(kind: str, theta: float = 0.0) -> np.ndarray:
    c, s = math.cos(0.5 * theta), math.sin(0.5 * theta)
    r = 1.0 / math.sqrt(2.0)
    return {
        "h": np.array([[r, r], [r, -r]], dtype=complex),
        "x": np.array([[0, 1], [1, 0]], dtype=complex),
        "rx": np.array([[c, -1j * s], [-1j * s, c]]),
        "ry": np.array([[c, -s], [s, c]], dtype=complex),
        "rz": np.array([[complex(c, -s), 0], [0, complex(c, s)]]),
    }[kind]


def gate_qubits(g) -> list:
    if g[0] == "cnot":
        return [int(g[1]), int(g[2])]
    if g[0] == "mcz":
        return [int(q) for q in g[1]]
    return [int(g[1])]


def apply_1q(psi: np.ndarray, n: int, q: int, m: np.ndarray) -> np.ndarray:
    v = psi.reshape(2 ** (n - q - 1), 2, 2 ** q)
    a0, a1 = v[:, 0, :].copy(), v[:, 1, :].copy()
    v[:, 0, :] = m[0, 0] * a0 + m[0, 1] * a1
    v[:, 1, :] = m[1, 0] * a0 + m[1, 1] * a1
    return psi


def apply_gate(psi: np.ndarray, n: int, g, inverse: bool = False) -> np.ndarray:
    k = g[0]
    idx = np.arange(2 ** n)
    if k == "cnot":
        c, t = int(g[1]), int(g[2])
        sel = ((idx >> c) & 1 == 1) & ((idx >> t) & 1 == 0)
        i0 = idx[sel]
        i1 = i0 | (1 << t)
        psi[i0], psi[i1] = psi[i1].copy(), psi[i0].copy()
        return psi
    if k == "mcz":
        mask = 0
        for q in g[1]:
            mask |= 1 << int(q)
        psi[(idx & mask) == mask] *= -1
        return psi
    theta = float(g[2]) if k in ("rx", "ry", "rz") else 0.0
    return apply_1q(psi, n, int(g[1]), gate_matrix(k, -theta if inverse else theta))


def apply_pauli(psi: np.ndarray, n: int, q: int, which: int) -> np.ndarray:
    """NoisyBackend::pauli: 1 X, 2 Y as Z then X, 3 Z."""
    if which in (2, 3):
        apply_gate(psi, n, ("mcz", [q]))
    if which in (1, 2):
        apply_gate(psi, n, ("x", q))
    return psi


# ---- sampled noise: NoisyBackend ------------------------------------------------------------------

class NoisyOracle:
    """NoisyBackend over a NumPy state vector: folding, sampled depolarizing sites, one stream."""

    def __init__(self, n: int, p: float, seed: int = 1):
        if not (0.0 <= p <= 1.0):
            raise ValueError("NoisyBackend: p must be in [0, 1]")
        self.n, self.p, self.rng, self.folds = n, p, MT19937_64(seed), 0
        self.events = []  # (site, code) of the current trajectory
        self.reset()

    def reset(self):
        self.psi = np.zeros(2 ** self.n, dtype=complex)
        self.psi[0] = 1.0
        self.site = 0
        self.events = []

    def set_scale(self, scale: int):
        if scale < 1 or scale > 99 or scale % 2 == 0:
            raise ValueError("scale must be an odd integer in [1, 99]")
        self.folds = (scale - 1) // 2

    def _noise(self, qs):
        if self.p == 0.0:
            self.site += 1
            return
        code = 0
        u = self.rng.u01
        if len(qs) == 1:
            if u() < self.p:
                code = 1 + int(u() * 3.0)
                apply_pauli(self.psi, self.n, qs[0], code)
        elif len(qs) == 2:
            if u() < self.p:
                code = 1 + int(u() * 15.0)
                apply_pauli(self.psi, self.n, qs[0], code & 3)
                apply_pauli(self.psi, self.n, qs[1], code >> 2)
        else:
            for i, q in enumerate(qs):
                if u() < self.p:
                    w = 1 + int(u() * 3.0)
                    code |= w << (2 * i)
                    apply_pauli(self.psi, self.n, q, w)
        if code:
            self.events.append((self.site, code))
        self.site += 1

    def gate(self, g):
        qs = gate_qubits(g)
        apply_gate(self.psi, self.n, g)
        self._noise(qs)
        for _ in range(self.folds):
            apply_gate(self.psi, self.n, g, inverse=True)
            self._noise(qs)
            apply_gate(self.psi, self.n, g)
            self._noise(qs)

    def probabilities(self) -> np.ndarray:
        return np.abs(self.psi) ** 2


def noisy_values(n, gates, p, seed, scale, n_traj, weights, oracle: NoisyOracle | None = None):
    """Per-trajectory sum_i w_i p_i at one scale; the stream continues in `oracle` when given."""
    nb = oracle or NoisyOracle(n, p, seed)
    nb.set_scale(scale)
    w = np.asarray(weights, dtype=float)
    out = np.empty(n_traj)
    for t in range(n_traj):
        nb.reset()
        for g in gates:
            nb.gate(g)
        out[t] = float(np.dot(w, nb.probabilities()))
    return out


# ---- ZNE driver and Richardson --------------------------------------------------------------------

def richardson_weights(scales) -> list:
    """c_k = prod_{j != k} lambda_j / (lambda_j - lambda_k)."""
    n = len(scales)
    c = [1.0] * n
    for k in range(n):
        for j in range(n):
            if j != k:
                c[k] *= scales[j] / (scales[j] - scales[k])
    return c


def welford(values):
    mean, m2 = 0.0, 0.0
    for t, v in enumerate(values):
        d = v - mean
        mean += d / (t + 1)
        m2 += d * (v - mean)
    n = len(values)
    se = math.sqrt(max(0.0, m2 / (n - 1)) / n) if n > 1 else 0.0
    return mean, se


def extrapolate(n, gates, p, seed, trajectories, scales, weights):
    """ZNE(trajectories).extrapolate(NoisyBackend(p, seed), ...) with a diagonal observable."""
    nb = NoisyOracle(n, p, seed)
    ntr = trajectories if p > 0.0 else 1
    vals, ses = [], []
    for s in scales:
        m, se = welford(noisy_values(n, gates, p, seed, int(s), ntr, weights, oracle=nb))
        if ntr == 1 and p > 0.0:
            se = float("nan")
        vals.append(m)
        ses.append(se)
    c = richardson_weights(scales)
    value = sum(ck * v for ck, v in zip(c, vals))
    std = math.sqrt(sum(ck * ck * se * se for ck, se in zip(c, ses)))
    return dict(value=value, std_error=std, values=vals, std_errors=ses, weights=c, trajectories=ntr)


# ---- exact: DensityNoiseBackend -------------------------------------------------------------------

_PAULI = [np.eye(2, dtype=complex), np.array([[0, 1], [1, 0]], dtype=complex),
          np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]], dtype=complex)]


def _op_on(rho: np.ndarray, n: int, qs, m: np.ndarray) -> np.ndarray:
    """M on qubits qs (operator index bit j = qs[j]) applied from the left: M rho."""
    k = len(qs)
    t = rho.reshape([2] * n + [2 ** n])
    axes = [n - 1 - q for q in reversed(qs)]  # tensor axis of qubit q is n - 1 - q
    mt = m.reshape([2] * (2 * k))
    t = np.tensordot(mt, t, axes=(list(range(k, 2 * k)), axes))
    t = np.moveaxis(t, list(range(k)), axes)
    return t.reshape(2 ** n, 2 ** n)


def _conj_apply(rho, n, qs, m):
    """m rho m^+."""
    r = _op_on(rho, n, qs, m)
    return _op_on(r.conj().T, n, qs, m).conj().T


def _unitary(g, inverse=False) -> tuple:
    k = g[0]
    if k == "cnot":
        u = np.zeros((4, 4), dtype=complex)  # operator bit 0 = control, bit 1 = target
        u[0, 0] = u[2, 2] = u[3, 1] = u[1, 3] = 1.0
        return [int(g[1]), int(g[2])], u
    if k == "mcz":
        qs = [int(q) for q in g[1]]
        u = np.eye(2 ** len(qs), dtype=complex)
        u[-1, -1] = -1.0
        return qs, u
    theta = float(g[2]) if k in ("rx", "ry", "rz") else 0.0
    return [int(g[1])], gate_matrix(k, -theta if inverse else theta)


def _depolarize(rho, n, qs, p):
    if p == 0.0:
        return rho
    if len(qs) > 2:
        for q in qs:
            rho = _depolarize(rho, n, [q], p)
        return rho
    m = len(qs)
    out = (1.0 - p) * rho
    w = p / (4 ** m - 1)
    for code in range(1, 4 ** m):
        P = np.array([[1.0]], dtype=complex)
        for j in range(m - 1, -1, -1):
            P = np.kron(P, _PAULI[(code >> (2 * j)) & 3])
        out = out + w * _conj_apply(rho, n, qs, P)
    return out


def density_run(n, gates, p, scale) -> np.ndarray:
    rho = np.zeros((2 ** n, 2 ** n), dtype=complex)
    rho[0, 0] = 1.0
    folds = (scale - 1) // 2
    for g in gates:
        qs, u = _unitary(g)
        _, ui = _unitary(g, inverse=True)
        rho = _depolarize(_conj_apply(rho, n, qs, u), n, qs, p)
        for _ in range(folds):
            rho = _depolarize(_conj_apply(rho, n, qs, ui), n, qs, p)
            rho = _depolarize(_conj_apply(rho, n, qs, u), n, qs, p)
    return rho


def zne_exact(n, gates, p, scales, weights):
    vals = [float(np.dot(weights, np.real(np.diag(density_run(n, gates, p, int(s)))))) for s in scales]
    c = richardson_weights(scales)
    return dict(value=sum(ck * v for ck, v in zip(c, vals)), values=vals, weights=c)
