"""Shared helpers for the tests (fixture decoding, random circuits)."""
import math
import os

import numpy as np

from conftest import GOLDEN


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def decode_h_cnot(arr):
    ops = []
    for k, a, b in arr.tolist():
        ops.append(("h", a) if k == 0 else ("cnot", a, b))
    return ops


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.float64), np.ascontiguousarray(b).view(np.float64))


def random_unitary(rng, D):
    z = (rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))) / math.sqrt(2.0)
    q, r = np.linalg.qr(z)
    return q * (np.diagonal(r) / np.abs(np.diagonal(r)))


def random_ops(rng, n, count, kinds=("h", "x", "rx", "ry", "rz", "u", "cnot", "cz", "mcz", "dense"), max_k=5):
    ops = []
    for _ in range(count):
        kind = kinds[int(rng.integers(0, len(kinds)))]
        if kind in ("cnot", "cz") and n < 2:
            kind = "u"
        q = int(rng.integers(0, n))
        if kind in ("h", "x"):
            ops.append((kind, q))
        elif kind in ("rx", "ry", "rz"):
            ops.append((kind, q, float(rng.uniform(-3.2, 3.2))))
        elif kind == "u":
            ops.append(("u", q, random_unitary(rng, 2).reshape(4).tolist()))
        elif kind in ("cnot", "cz"):
            t = int(rng.integers(0, n - 1))
            t += int(t >= q)
            ops.append((kind, q, t))
        elif kind == "mcz":
            m = int(rng.integers(1, min(n, 4) + 1))
            ops.append(("mcz", [int(v) for v in rng.permutation(n)[:m]]))
        else:
            k = int(rng.integers(1, min(n, max_k) + 1))
            targets = [int(v) for v in rng.permutation(n)[:k]]
            ops.append(("dense", targets, random_unitary(rng, 1 << k)))
    return ops
