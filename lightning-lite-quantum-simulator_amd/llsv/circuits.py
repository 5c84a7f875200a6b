"""Seeded circuit builders for the benchmark and parity workloads.

Every builder returns a list of op tuples:
  ('h', q) ('x', q) ('rx', q, t) ('ry', q, t) ('rz', q, t) ('u', q, [g00, g01, g10, g11])
  ('cnot', c, t) ('cz', a, b) ('mcz', [q...]) ('dense', [t0, ...], U)
Qubit q is bit q of the basis index (src/Backend.h:3-4).
"""
from __future__ import annotations

import math

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def random_h_cnot(seed: int, n: int = 12, n_ops: int = 1000):
    """1000 mixed ops: with p = 0.5 an H on a uniform qubit, else a CNOT on a uniform
    ordered distinct (control, target) pair -- the published mixed H + CNOT workload
    (README:19-27).  Draw order per op: u = random(); H: q = integers(n);
    CNOT: c = integers(n), t = integers(n - 1), t += (t >= c)."""
    rng = _rng(seed)
    ops = []
    for _ in range(n_ops):
        if rng.random() < 0.5:
            ops.append(("h", int(rng.integers(0, n))))
        else:
            c = int(rng.integers(0, n))
            t = int(rng.integers(0, n - 1))
            t += int(t >= c)
            ops.append(("cnot", c, t))
    return ops


def layered_rotations(seed: int, n: int, layers: int):
    """Per layer: RX, RY, RZ on every qubit (angles uniform in [0, 2pi)), then a brick of
    CZ(q, q+1) over even q on even layers and odd q on odd layers."""
    rng = _rng(seed)
    ops = []
    for layer in range(layers):
        for q in range(n):
            a = rng.uniform(0.0, 2.0 * math.pi, size=3)
            ops.append(("rx", q, float(a[0])))
            ops.append(("ry", q, float(a[1])))
            ops.append(("rz", q, float(a[2])))
        for q in range(layer % 2, n - 1, 2):
            ops.append(("cz", q, q + 1))
    return ops


def haar_2x2(rng: np.random.Generator) -> list:
    """Haar-random U(2): QR of a complex Ginibre matrix with the R-diagonal phase fix
    (Mezzadri, Notices AMS 54 (2007) 592)."""
    z = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))) / math.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    q = q * (d / np.abs(d))
    return [complex(v) for v in q.reshape(4)]


def haar_cnot_sweep(seed: int, n: int, sweeps: int = 1):
    """Bandwidth sweep: a Haar 2x2 on every qubit 0..n-1, then a CNOT ladder (q, q+1),
    q = 0..n-2; repeated `sweeps` times."""
    rng = _rng(seed)
    ops = []
    for _ in range(sweeps):
        for q in range(n):
            ops.append(("u", q, haar_2x2(rng)))
        for q in range(n - 1):
            ops.append(("cnot", q, q + 1))
    return ops


def hea_ansatz(n: int, layers: int, params):
    """Hardware-efficient ansatz: per layer RY(theta) then RZ(phi) on every qubit, then
    CNOT(q, q+1) for q = 0..n-2.  params has 2 * n * layers entries, ordered
    [layer][qubit][ry, rz]."""
    p = np.asarray(params, dtype=np.float64).reshape(layers, n, 2)
    ops = []
    for layer in range(layers):
        for q in range(n):
            ops.append(("ry", q, float(p[layer, q, 0])))
            ops.append(("rz", q, float(p[layer, q, 1])))
        for q in range(n - 1):
            ops.append(("cnot", q, q + 1))
    return ops


def zz_x_hamiltonian(seed: int, n: int):
    """H = sum_{i=0}^{n-2} a_i Z_i Z_{i+1} + sum_{i=0}^{n-1} b_i X_i with a, b ~ N(0, 1).
    Returned as a list of (coeff, xmask, zmask) Pauli terms."""
    rng = _rng(seed)
    a = rng.standard_normal(n - 1)
    b = rng.standard_normal(n)
    terms = [(float(a[i]), 0, (1 << i) | (1 << (i + 1))) for i in range(n - 1)]
    terms += [(float(b[i]), 1 << i, 0) for i in range(n)]
    return terms


def noisy_brick(seed: int, n: int = 14, layers: int = 8):
    """Trajectory workload: per layer H on every qubit, RZ(theta ~ U[0, 2pi)) on every qubit,
    then CNOT(q, q+1) over even q on even layers and odd q on odd layers.  Noise rates
    gamma_q ~ U[1e-3, 1e-2] per qubit, lambda_q = gamma_q / 2.  Returns (gates, gamma, lam)."""
    rng = _rng(seed)
    gamma = rng.uniform(1e-3, 1e-2, size=n)
    ops = []
    for layer in range(layers):
        ops += [("h", q) for q in range(n)]
        ops += [("rz", q, float(rng.uniform(0.0, 2.0 * math.pi))) for q in range(n)]
        ops += [("cnot", q, q + 1) for q in range(layer % 2, n - 1, 2)]
    return ops, gamma, gamma / 2.0


def mixed_batch(seed: int, count: int = 2048, n_min: int = 10, n_max: int = 16, layers: int = 8):
    """`count` noisy brick circuits with n ~ U{n_min..n_max}, each with its own angles and rates.
    Returns a list of (n, gates, gamma, lam, traj_id)."""
    rng = _rng(seed)
    sizes = rng.integers(n_min, n_max + 1, size=count)
    out = []
    for i, n in enumerate(sizes):
        g, gm, lm = noisy_brick(seed * 100003 + i, int(n), layers)
        out.append((int(n), g, gm, lm, i))
    return out
