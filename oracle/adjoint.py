"""CPU oracle for energies and gradients of parameterised circuits -- TEST INFRASTRUCTURE ONLY.

* energy: psi = U(theta)|0> through the C oracle, E = Re <psi|H|psi> with H applied by
  pauli_apply (NumPy, phase i^{nY} (-1)^{popcount(src & z)}).
* param_shift: the reference's rule, src/Algorithms.h:277-292:
  dE/dtheta_k = [E(theta + pi/2 e_k) - E(theta - pi/2 e_k)] / 2 (exact for RX/RY/RZ).
* finite_diff: central differences.
* adjoint: the adjoint sweep (Jones & Gacon, arXiv:2009.02823) restated with the C oracle:
  lambda = H psi; for gates in reverse: g_k = Im <lambda|G_k psi>; psi, lambda <- U_k^dag.

Circuits: ('rx'|'ry'|'rz', q, p_index) ('u', q, g4) ('cnot', c, t) ('cz', a, b).
Hamiltonians: [(coeff, xmask, zmask)].
"""
from __future__ import annotations

import math

import numpy as np

from . import oracle as O


def pauli_apply(psi, x, z):
    n = int(psi.size).bit_length() - 1
    idx = np.arange(psi.size, dtype=np.uint64)
    src = idx ^ np.uint64(x)
    ny = bin(int(x) & int(z)).count("1") % 4
    sign = 1.0 - 2.0 * (np.bitwise_count(src & np.uint64(z)) & 1)
    return (1j ** ny) * sign * psi[src.astype(np.int64)]


def h_apply(psi, terms):
    out = np.zeros_like(psi)
    for c, x, z in terms:
        out += c * pauli_apply(psi, x, z)
    return out


def _ops(gates, theta):
    ops = []
    for g in gates:
        if g[0] in ("rx", "ry", "rz"):
            ops.append((g[0], g[1], float(theta[g[2]])))
        else:
            ops.append(g)
    return ops


def state(n, gates, theta):
    return O.run(n, _ops(gates, theta))


def energy(n, gates, theta, terms):
    psi = state(n, gates, theta)
    return float(np.vdot(psi, h_apply(psi, terms)).real)


def param_shift(n, gates, theta, terms, which=None):
    theta = np.asarray(theta, dtype=np.float64)
    ks = range(theta.size) if which is None else which
    g = np.zeros(len(list(ks)) if which is not None else theta.size)
    for j, k in enumerate(ks):
        tp = theta.copy()
        tp[k] += math.pi / 2
        tm = theta.copy()
        tm[k] -= math.pi / 2
        g[j] = 0.5 * (energy(n, gates, tp, terms) - energy(n, gates, tm, terms))
    return g


def finite_diff(n, gates, theta, terms, h=1e-6, which=None):
    theta = np.asarray(theta, dtype=np.float64)
    ks = list(range(theta.size)) if which is None else list(which)
    g = np.zeros(len(ks))
    for j, k in enumerate(ks):
        tp = theta.copy()
        tp[k] += h
        tm = theta.copy()
        tm[k] -= h
        g[j] = (energy(n, gates, tp, terms) - energy(n, gates, tm, terms)) / (2 * h)
    return g


def _dagger(g4):
    m = np.asarray(g4, dtype=np.complex128).reshape(2, 2)
    return m.conj().T.reshape(4).tolist()


def adjoint(n, gates, theta, terms):
    theta = np.asarray(theta, dtype=np.float64)
    ops = _ops(gates, theta)
    psi = O.run(n, ops)
    lam = h_apply(psi, terms)
    e = float(np.vdot(psi, lam).real)
    grad = np.zeros(theta.size)
    gen = {"rx": (1, 0), "ry": (1, 1), "rz": (0, 1)}  # (x bit, z bit) of the generator on the qubit
    for g, op in zip(reversed(gates), reversed(ops)):
        if g[0] in gen:
            xb, zb = gen[g[0]]
            b = 1 << g[1]
            grad[g[2]] += float(np.vdot(lam, pauli_apply(psi, b * xb, b * zb)).imag)
        if op[0] in ("rx", "ry", "rz", "u"):
            inv = _dagger(O.gate4(op))
            O.apply_1q(psi, op[1], inv)
            O.apply_1q(lam, op[1], inv)
        else:
            O.apply_op(psi, op)
            O.apply_op(lam, op)
    return e, grad
