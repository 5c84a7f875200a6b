"""CPU oracle for the dense Lindblad trajectory solver -- TEST INFRASTRUCTURE ONLY.

Restates src/TrajectorySolver.h over NumPy:
  * traj_detail::splitmix64 (:86-91) and make_rng (:311-313): trajectory i draws from
    std::mt19937_64(splitmix64(seed ^ splitmix64(i + 1)));
  * std::uniform_real_distribution<double>(0, 1) and std::normal_distribution<double>(0, s) as
    libstdc++ implements them (generate_canonical<double, 53>: one 64-bit draw x -> x / 2^64,
    clamped below 1; normal: Marsaglia's polar method on 2u - 1 pairs, the second value cached);
  * traj_detail::expm (:97-111): scaling and squaring, 18-term Taylor series;
  * pauli_like (:316-326), make_plan (:268-290), the jump algorithm (:379-428) and the diffusive
    Euler-Maruyama step (:432-460).
Exact reference for ensembles: the Lindblad equation integrated through its Liouvillian
(scipy.linalg.expm of the 4^n x 4^n generator).

Pinned by (tests/test_lindblad_oracle.py): splitmix64 known answers; libstdc++'s own uniform and
normal draws (tests/golden/lindblad_rng.json, written by a g++ program of this repository,
tests/golden/make_lindblad_rng.cpp); scipy's expm; the closed forms of the reference's tests
(T1: <P1> = exp(-t/T1); T2: <X> = exp(-t/T2); the dephasing variance ratio 1 + exp(4 gamma t));
ensemble rho against the exact Liouvillian solution.
"""
from __future__ import annotations

import math

import numpy as np

from oracle.zne import MT19937_64

M64 = (1 << 64) - 1
STANDARD, PROJECTOR, ANALOG = 0, 1, 2


def splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


class Rng:
    """std::mt19937_64 with libstdc++'s uniform_real_distribution(0, 1) and normal_distribution."""

    def __init__(self, seed: int):
        self.mt = MT19937_64(seed & M64)
        self.saved = None

    def uniform(self) -> float:
        r = float(self.mt()) / 18446744073709551616.0
        return r if r < 1.0 else math.nextafter(1.0, 0.0)

    def normal(self, stddev: float) -> float:
        if self.saved is not None:
            v, self.saved = self.saved, None
            return v * stddev
        while True:
            x = 2.0 * self.uniform() - 1.0
            y = 2.0 * self.uniform() - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        self.saved = x * mult
        return y * mult * stddev


def make_rng(seed: int, index: int) -> Rng:
    return Rng(splitmix64((seed ^ splitmix64((index + 1) & M64)) & M64))


def expm(a: np.ndarray) -> np.ndarray:
    nrm = float(np.max(np.sum(np.abs(a), axis=1)))
    s = 0
    if nrm > 0.5:
        s = max(0, int(math.ceil(math.log2(nrm / 0.5))))
    b = a / (2.0 ** s)
    term = np.eye(a.shape[0], dtype=complex)
    result = term.copy()
    for k in range(1, 19):
        term = (term @ b) / float(k)
        result = result + term
    for _ in range(s):
        result = result @ result
    return result


def pauli_like(l: np.ndarray):
    if np.linalg.norm(l - l.conj().T) > 1e-9 * max(1.0, np.linalg.norm(l)):
        return False, None, 0.0
    l2 = l @ l
    c = float(l2[0, 0].real)
    if not c > 1e-12:
        return False, None, 0.0
    if np.linalg.norm(l2 - c * np.eye(l.shape[0])) > 1e-9 * max(1.0, np.linalg.norm(l2)):
        return False, None, 0.0
    return True, l / math.sqrt(c), c


def make_plan(H, channels, kind, duration, dt):
    """H: matrix or callable t -> matrix; channels: list of (L, gamma)."""
    dim = channels[0][0].shape[0] if channels else (H(0.0) if callable(H) else H).shape[0]
    chs = [(np.asarray(l, dtype=complex), float(g)) for l, g in channels if float(g) > 0.0]
    if not chs:
        kind = STANDARD
    steps = max(1, int(math.ceil(duration / dt - 1e-9)))
    h = duration / steps
    eye = np.eye(dim, dtype=complex)
    ops = []
    for l, g in chs:
        ok, p, c = pauli_like(l)
        if kind == PROJECTOR and ok:
            ops.append(math.sqrt(c * g) * (eye - p))
        else:
            ops.append(math.sqrt(g) * l)
    op_dop = [j.conj().T @ j for j in ops]
    jump = kind != ANALOG
    hnh = np.zeros((dim, dim), dtype=complex)
    if jump:
        for m in op_dop:
            hnh = hnh - 0.5j * m
    plan = dict(kind=kind, steps=steps, h=h, op=ops, op_dop=op_dop, constant=not callable(H))
    if not callable(H):
        plan["prop"] = expm((-1j) * (np.asarray(H, dtype=complex) + hnh) * h)
    else:
        n_nodes = 2 * steps + 1 if jump else steps
        plan["nodes"] = [(-1j) * (np.asarray(H(0.5 * h * j if jump else h * j), dtype=complex) + hnh)
                         for j in range(n_nodes)]
    return plan


def jumps(psi, plan, rng: Rng):
    psi = psi.copy()
    h, ops = plan["h"], plan["op"]
    nch = len(ops)
    for s in range(plan["steps"]):
        if plan["constant"]:
            tmp = plan["prop"] @ psi
        else:
            a0, a1, a2 = plan["nodes"][2 * s], plan["nodes"][2 * s + 1], plan["nodes"][2 * s + 2]
            k1 = a0 @ psi
            w = psi + (0.5 * h) * k1
            k2 = a1 @ w
            w = psi + (0.5 * h) * k2
            k3 = a1 @ w
            w = psi + h * k3
            k4 = a2 @ w
            tmp = psi + (h / 6.0) * (((k1 + 2.0 * k2) + 2.0 * k3) + k4)
        survive = float(np.real(np.vdot(tmp, tmp)))
        if not (survive > 0.0) or not math.isfinite(survive):
            raise RuntimeError("TrajectorySolver: state norm vanished; reduce dt")
        if nch > 0 and rng.uniform() >= survive:
            ws = [float(np.real(np.vdot(o @ tmp, o @ tmp))) for o in ops]
            total = sum(ws)
            if total > 0.0:
                r = rng.uniform() * total
                pick, acc = nch, 0.0
                for k in range(nch):
                    acc += ws[k]
                    if r < acc and ws[k] > 0.0:
                        pick = k
                        break
                if pick == nch:
                    for k in range(nch - 1, -1, -1):
                        if ws[k] > 0.0:
                            pick = k
                            break
                psi = (ops[pick] @ tmp) / math.sqrt(ws[pick])
                continue
        psi = tmp / math.sqrt(survive)
    return psi


def diffusive(psi, plan, rng: Rng):
    psi = psi.copy()
    h, ops, dops = plan["h"], plan["op"], plan["op_dop"]
    sq = math.sqrt(h)
    for s in range(plan["steps"]):
        if plan["constant"]:
            psi = plan["prop"] @ psi
            drift = np.zeros_like(psi)
        else:
            drift = plan["nodes"][s] @ psi
        noise = np.zeros_like(psi)
        for k in range(len(ops)):
            cp = ops[k] @ psi
            x = 2.0 * float(np.real(np.vdot(psi, cp)))
            t2 = dops[k] @ psi
            drift = drift + (-0.5) * ((t2 - x * cp) + (0.25 * x * x) * psi)
            noise = noise + (cp - (0.5 * x) * psi) * rng.normal(sq)
        psi = psi + (h * drift + noise)
        nrm = float(np.linalg.norm(psi))
        if not (nrm > 0.0) or not math.isfinite(nrm):
            raise RuntimeError("TrajectorySolver: diffusion state diverged; reduce dt")
        psi = psi / nrm
    return psi


def trajectory(psi0, plan, seed, index):
    rng = make_rng(seed, index)
    psi = np.asarray(psi0, dtype=complex) / np.linalg.norm(psi0)
    return diffusive(psi, plan, rng) if plan["kind"] == ANALOG else jumps(psi, plan, rng)


def ensemble_states(psi0, plan, seed, n):
    return np.array([trajectory(psi0, plan, seed, i) for i in range(n)])


def lindblad_exact(H, channels, rho0, t):
    """rho(t) through the Liouvillian (time-independent H), vec index r + dim c (column stacking)."""
    from scipy.linalg import expm as sexpm

    H = np.asarray(H, dtype=complex)
    d = H.shape[0]
    eye = np.eye(d)
    # vec(A X B) = (B^T kron A) vec(X) with column stacking
    Lv = -1j * (np.kron(eye, H) - np.kron(H.T, eye))
    for l, g in channels:
        l = np.asarray(l, dtype=complex)
        ld = l.conj().T @ l
        Lv = Lv + g * (np.kron(l.conj(), l) - 0.5 * np.kron(eye, ld) - 0.5 * np.kron(ld.T, eye))
    v = sexpm(Lv * t) @ np.asarray(rho0, dtype=complex).reshape(-1, order="F")
    return v.reshape(d, d, order="F")
