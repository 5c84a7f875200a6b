"""CPU oracle for the state-vector hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  The product
(libllsv, the ``quantum_sim_v3`` / ``quantum_sim_v2`` modules, the ``llsv``
package) never imports it.

Contents
  * ctypes wrapper around ``liborc.so`` (sv_oracle.c): the C restatement of the
    reference one-qubit update (src/StateVectorOptimized.cpp:56-78), CNOT
    (src/Backend.h:90-97), MCZ/CZ (src/Backend.h:99-110), dense k-qubit apply,
    norm, probabilities, <Z_q>, inner product and Pauli-string axpy.
  * gate matrices restated from src/Backend.h:66-88 and python/gate_fusion.py:54-60.
  * ``run(n, ops)``: drives a circuit (the op-tuple format below) through the C oracle.
  * pure-NumPy brute force (``dense_unitary``, ``run_bruteforce``) for tiny n, an
    independent second oracle that builds the full 2^n x 2^n operator with np.kron.

Op tuples (shared with the product's circuit builders):
  ('h', q) ('x', q) ('rx', q, t) ('ry', q, t) ('rz', q, t) ('u', q, [g00, g01, g10, g11])
  ('cnot', c, t) ('cz', a, b) ('mcz', [q...]) ('dense', [t0, t1, ...], U)   U: 2^k x 2^k, bit j <-> t_j
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", _HERE, "liborc.so"], check=True, capture_output=True)
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        U64 = ctypes.c_uint64
        lib.orc_init_zero.argtypes = [P, I]
        lib.orc_apply_1q.argtypes = [P, I, I, P]
        lib.orc_apply_1q_standard.argtypes = [P, I, I, P]
        lib.orc_apply_cnot.argtypes = [P, I, I, I]
        lib.orc_apply_mcz.argtypes = [P, I, P, I]
        lib.orc_apply_dense.argtypes = [P, I, I, P, P]
        lib.orc_norm2.argtypes = [P, I]
        lib.orc_norm2.restype = D
        lib.orc_probabilities.argtypes = [P, I, P]
        lib.orc_expect_z.argtypes = [P, I, I]
        lib.orc_expect_z.restype = D
        lib.orc_inner.argtypes = [P, P, I, P]
        lib.orc_inner.restype = None
        lib.orc_axpy_pauli.argtypes = [P, I, U64, U64, D, P]
        lib.orc_axpy_pauli.restype = None
        lib.orc_sincos.argtypes = [D, P]
        lib.orc_sincos.restype = None
        _LIB = lib
    return _LIB


class OracleError(ValueError):
    pass


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _nq(psi: np.ndarray) -> int:
    n = int(psi.size).bit_length() - 1
    assert psi.dtype == np.complex128 and psi.flags.c_contiguous and (1 << n) == psi.size
    return n


def _check(rc: int, what: str):
    if rc != 0:
        raise OracleError(f"oracle {what}: invalid argument")


# ---------------------------------------------------------------- gate matrices
R2 = 1.0 / math.sqrt(2.0)


def _cs(x):
    """(cos x, sin x) through libm sincos, as GCC compiles the reference's paired
    std::cos/std::sin calls (see orc_sincos in sv_oracle.c)."""
    out = np.zeros(2, dtype=np.float64)
    _lib().orc_sincos(float(x), _ptr(out))
    return float(out[0]), float(out[1])


def mat_h():
    # src/Backend.h:66-69: r = 1/sqrt(2), [r, r; r, -r]
    return [complex(R2), complex(R2), complex(R2), complex(-R2)]


def mat_x():
    return [0j, 1 + 0j, 1 + 0j, 0j]  # src/Backend.h:71


def mat_rx(t):
    # src/Backend.h:73-77: [c, -is; -is, c], c = cos(t/2), s = sin(t/2)
    c, s = _cs(0.5 * t)
    return [complex(c), complex(0.0, -s), complex(0.0, -s), complex(c)]


def mat_ry(t):
    # src/Backend.h:79-83: [c, -s; s, c]
    c, s = _cs(0.5 * t)
    return [complex(c), complex(-s), complex(s), complex(c)]


def mat_rz(t):
    # src/Backend.h:85-88: diag(polar(1, -t/2), polar(1, t/2))
    ca, sa = _cs(-0.5 * t)
    cb, sb = _cs(0.5 * t)
    return [complex(ca, sa), 0j, 0j, complex(cb, sb)]


# python/gate_fusion.py:54-60
GATES = {
    "X": [0 + 0j, 1 + 0j, 1 + 0j, 0 + 0j],
    "H": [1 / np.sqrt(2) + 0j, 1 / np.sqrt(2) + 0j, 1 / np.sqrt(2) + 0j, -1 / np.sqrt(2) + 0j],
    "S": [1 + 0j, 0 + 0j, 0 + 0j, 0 + 1j],
    "T": [1 + 0j, 0 + 0j, 0 + 0j, np.exp(1j * np.pi / 4)],
}


def gate4(op) -> list:
    kind = op[0]
    if kind == "h":
        return mat_h()
    if kind == "x":
        return mat_x()
    if kind == "rx":
        return mat_rx(op[2])
    if kind == "ry":
        return mat_ry(op[2])
    if kind == "rz":
        return mat_rz(op[2])
    if kind == "u":
        return [complex(v) for v in op[2]]
    raise KeyError(kind)


def _g8(g4) -> np.ndarray:
    return np.array([[complex(v).real, complex(v).imag] for v in g4], dtype=np.float64).reshape(8)


# ---------------------------------------------------------------- C oracle calls
def zero_state(n: int) -> np.ndarray:
    psi = np.empty(1 << n, dtype=np.complex128)
    _check(_lib().orc_init_zero(_ptr(psi), n), "init")
    return psi


def apply_1q(psi, target, g4, standard=False):
    g = _g8(g4)
    f = _lib().orc_apply_1q_standard if standard else _lib().orc_apply_1q
    _check(f(_ptr(psi), _nq(psi), int(target), _ptr(g)), "apply_1q")


def apply_cnot(psi, control, target):
    _check(_lib().orc_apply_cnot(_ptr(psi), _nq(psi), int(control), int(target)), "cnot")


def apply_mcz(psi, qubits):
    qs = np.ascontiguousarray(qubits, dtype=np.int32)
    _check(_lib().orc_apply_mcz(_ptr(psi), _nq(psi), _ptr(qs), int(qs.size)), "mcz")


def apply_dense(psi, targets, U):
    t = np.ascontiguousarray(targets, dtype=np.int32)
    m = np.ascontiguousarray(U, dtype=np.complex128)
    assert m.shape == (1 << t.size, 1 << t.size)
    _check(_lib().orc_apply_dense(_ptr(psi), _nq(psi), int(t.size), _ptr(t), _ptr(m)), "dense")


def norm2(psi) -> float:
    return float(_lib().orc_norm2(_ptr(psi), _nq(psi)))


def probabilities(psi) -> np.ndarray:
    p = np.empty(psi.size, dtype=np.float64)
    _lib().orc_probabilities(_ptr(psi), _nq(psi), _ptr(p))
    return p


def expect_z(psi, q) -> float:
    return float(_lib().orc_expect_z(_ptr(psi), _nq(psi), int(q)))


def inner(a, b) -> complex:
    out = np.zeros(2, dtype=np.float64)
    _lib().orc_inner(_ptr(a), _ptr(b), _nq(a), _ptr(out))
    return complex(out[0], out[1])


def axpy_pauli(psi, xmask, zmask, coeff, out):
    _lib().orc_axpy_pauli(_ptr(psi), _nq(psi), int(xmask), int(zmask), float(coeff), _ptr(out))


def apply_op(psi, op):
    kind = op[0]
    if kind in ("h", "x", "rx", "ry", "rz", "u"):
        apply_1q(psi, op[1], gate4(op))
    elif kind == "cnot":
        apply_cnot(psi, op[1], op[2])
    elif kind == "cz":
        apply_mcz(psi, [op[1], op[2]])
    elif kind == "mcz":
        apply_mcz(psi, list(op[1]))
    elif kind == "dense":
        apply_dense(psi, list(op[1]), op[2])
    else:
        raise KeyError(kind)


def run(n: int, ops, psi=None) -> np.ndarray:
    """Apply ops in order to |0...0> (or a copy of psi) through the C oracle."""
    s = zero_state(n) if psi is None else np.array(psi, dtype=np.complex128, copy=True)
    for op in ops:
        apply_op(s, op)
    return s


# ---------------------------------------------------------------- brute force (tiny n)
def _embed_1q(n, q, G):
    ops = [np.eye(2, dtype=np.complex128)] * n
    ops = list(ops)
    ops[q] = np.asarray(G, dtype=np.complex128).reshape(2, 2)
    M = np.array([[1.0 + 0j]])
    for k in reversed(range(n)):  # qubit n-1 is the most significant factor
        M = np.kron(M, ops[k])
    return M


def full_operator(n: int, op) -> np.ndarray:
    """The 2^n x 2^n matrix of one op, built independently of the C oracle."""
    N = 1 << n
    kind = op[0]
    if kind in ("h", "x", "rx", "ry", "rz", "u"):
        return _embed_1q(n, op[1], gate4(op))
    idx = np.arange(N)
    if kind == "cnot":
        c, t = op[1], op[2]
        perm = np.where((idx >> c) & 1, idx ^ (1 << t), idx)
        M = np.zeros((N, N), dtype=np.complex128)
        M[perm, idx] = 1.0
        return M
    if kind in ("cz", "mcz"):
        qs = [op[1], op[2]] if kind == "cz" else list(op[1])
        mask = sum(1 << q for q in qs)
        return np.diag(np.where((idx & mask) == mask, -1.0, 1.0).astype(np.complex128))
    if kind == "dense":
        targets = list(op[1])
        U = np.asarray(op[2], dtype=np.complex128)
        k = len(targets)
        M = np.zeros((N, N), dtype=np.complex128)
        tmask = sum(1 << t for t in targets)
        for col in range(N):
            c_loc = sum(((col >> t) & 1) << j for j, t in enumerate(targets))
            rest = col & ~tmask
            for r_loc in range(1 << k):
                row = rest | sum(((r_loc >> j) & 1) << t for j, t in enumerate(targets))
                M[row, col] = U[r_loc, c_loc]
        return M
    raise KeyError(kind)


def run_bruteforce(n: int, ops) -> np.ndarray:
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    for op in ops:
        psi = full_operator(n, op) @ psi
    return psi
