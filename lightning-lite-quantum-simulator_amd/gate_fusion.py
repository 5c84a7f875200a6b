"""Gate-fusion pre-pass.

Keeps the reference's ``GateFuser`` interface (python/gate_fusion.py:3-51) so its scripts
run unchanged:
  * ``multiply_gates(g1, g2)`` -> row-major 4-list of ``g2 @ g1`` ("g1 first, then g2",
    python/gate_fusion.py:7-12);
  * ``fuse_circuit(gates)`` merges runs of *adjacent* gates on the same target
    (python/gate_fusion.py:15-37);
  * ``analyze_circuit(gates)`` prints the counts and returns the fused list
    (python/gate_fusion.py:40-51);
  * ``GATES`` holds X, H, S, T (python/gate_fusion.py:54-60).

New here: ``fuse_dense(ops, max_k=5)`` turns a circuit of op tuples (see llsv.circuits)
into dense blocks of at most ``max_k`` qubits, ``('dense', targets, U)`` with U of shape
2^k x 2^k and operator bit j <-> targets[j] (the bit convention of src/Noise.h:58-65),
or ``('u', q, g4)`` for one-qubit blocks.  Blocks are grown greedily along the circuit:
an op joins the open block while the union of qubits stays within ``max_k``; otherwise
the block is emitted and a new one opens.  The emitted circuit equals the input one up to
rounding.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["GateFuser", "GATES", "fuse_dense", "op_matrix"]


class GateFuser:
    """Pre-processor that merges consecutive gates acting on the same qubit."""

    @staticmethod
    def multiply_gates(gate1, gate2):
        a = np.asarray(gate1, dtype=np.complex128).reshape(2, 2)
        b = np.asarray(gate2, dtype=np.complex128).reshape(2, 2)
        return (b @ a).reshape(4).tolist()

    @staticmethod
    def fuse_circuit(gates):
        out = []
        for target, gate in gates:
            if out and out[-1][0] == target:
                out[-1] = (target, GateFuser.multiply_gates(out[-1][1], gate))
            else:
                out.append((target, gate))
        return out

    @staticmethod
    def analyze_circuit(gates):
        fused = GateFuser.fuse_circuit(gates)
        before, after = len(gates), len(fused)
        print(f"Original gates: {before}")
        print(f"After fusion: {after}")
        print(f"Reduction: {(1 - after / before) * 100:.1f}%")
        return fused

    @staticmethod
    def fuse_dense(ops, max_k: int = 5):
        return fuse_dense(ops, max_k)


_R2 = 1.0 / np.sqrt(2.0)
GATES = {
    "X": [0j, 1 + 0j, 1 + 0j, 0j],
    "H": [_R2 + 0j, _R2 + 0j, _R2 + 0j, -_R2 + 0j],
    "S": [1 + 0j, 0j, 0j, 1j],
    "T": [1 + 0j, 0j, 0j, np.exp(1j * np.pi / 4)],
}


def _gate4(op):
    kind = op[0]
    if kind == "u":
        return np.asarray(op[2], dtype=np.complex128).reshape(2, 2)
    if kind == "h":
        r = 1.0 / math.sqrt(2.0)
        return np.array([[r, r], [r, -r]], dtype=np.complex128)
    if kind == "x":
        return np.array([[0, 1], [1, 0]], dtype=np.complex128)
    t = float(op[2])
    c, s = math.cos(0.5 * t), math.sin(0.5 * t)
    if kind == "rx":
        return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)
    if kind == "ry":
        return np.array([[c, -s], [s, c]], dtype=np.complex128)
    if kind == "rz":
        return np.array([[complex(math.cos(-0.5 * t), math.sin(-0.5 * t)), 0],
                         [0, complex(math.cos(0.5 * t), math.sin(0.5 * t))]], dtype=np.complex128)
    raise KeyError(kind)


def op_qubits(op):
    kind = op[0]
    if kind in ("cnot", "cz"):
        return [int(op[1]), int(op[2])]
    if kind in ("mcz", "dense"):
        return [int(q) for q in op[1]]
    return [int(op[1])]


def op_matrix(op, targets):
    """Matrix of `op` on the block qubits `targets` (op's qubits must be among them),
    2^k x 2^k with local bit j <-> targets[j]."""
    k = len(targets)
    D = 1 << k
    pos = {q: j for j, q in enumerate(targets)}
    idx = np.arange(D)
    M = np.zeros((D, D), dtype=np.complex128)
    kind = op[0]
    if kind in ("cnot",):
        c, t = pos[int(op[1])], pos[int(op[2])]
        dst = np.where((idx >> c) & 1, idx ^ (1 << t), idx)
        M[dst, idx] = 1.0
        return M
    if kind in ("cz", "mcz"):
        mask = sum(1 << pos[q] for q in op_qubits(op))
        return np.diag(np.where((idx & mask) == mask, -1.0, 1.0).astype(np.complex128))
    if kind == "dense":
        sub = [pos[int(q)] for q in op[1]]
        U = np.asarray(op[2], dtype=np.complex128)
        smask = sum(1 << j for j in sub)
        for col in range(D):
            c_loc = sum(((col >> j) & 1) << i for i, j in enumerate(sub))
            rest = col & ~smask
            for r_loc in range(1 << len(sub)):
                row = rest | sum(((r_loc >> i) & 1) << j for i, j in enumerate(sub))
                M[row, col] = U[r_loc, c_loc]
        return M
    g = _gate4(op)
    j = pos[int(op[1])]
    bit = (idx >> j) & 1
    base = idx & ~(1 << j)
    for rb in (0, 1):
        M[base | (rb << j), idx] = g[rb, bit]
    return M


def _widen(M, old, new):
    """Re-express block matrix M on qubits `old` as a matrix on `new` (old is a prefix of new)."""
    extra = len(new) - len(old)
    if extra == 0:
        return M
    # New local bits beyond len(old) are untouched: M (x) I in the new high bits.
    return np.kron(np.eye(1 << extra, dtype=np.complex128), M)


def fuse_dense(ops, max_k: int = 5):
    if max_k < 1:
        raise ValueError("max_k must be >= 1")
    out = []
    targets: list = []
    M = None

    def flush():
        nonlocal targets, M
        if M is None:
            return
        if len(targets) == 1:
            out.append(("u", targets[0], M.reshape(4).tolist()))
        else:
            out.append(("dense", list(targets), M))
        targets, M = [], None

    for op in ops:
        qs = op_qubits(op)
        if len(set(qs)) > max_k:
            flush()
            out.append(op)
            continue
        union = targets + [q for q in qs if q not in targets]
        if M is not None and len(union) > max_k:
            flush()
            union = list(dict.fromkeys(qs))
        if M is None:
            targets = union
            M = np.eye(1 << len(targets), dtype=np.complex128)
        else:
            M = _widen(M, targets, union)
            targets = union
        M = op_matrix(op, targets) @ M
    flush()
    return out
