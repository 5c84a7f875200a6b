"""Batched gate-level noisy trajectories on the GPU (llsv_traj_run).

After every gate, each qubit it touches goes through amplitude damping AD(gamma_q) and then
phase damping PD(lambda_q), unravelled by Kraus sampling with Philox4x32-10 streams keyed by
(seed, trajectory).  One persistent launch runs a whole batch -- many trajectories of one
circuit, or many circuits of different sizes.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native, gates as G

ZSTRIDE = 32


class TGate(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("q0", ctypes.c_int32), ("q1", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("m", ctypes.c_double * 8)]


class TItem(ctypes.Structure):
    _fields_ = [("n_qubits", ctypes.c_int32), ("first_gate", ctypes.c_int32), ("n_gates", ctypes.c_int32),
                ("first_rate", ctypes.c_int32), ("traj", ctypes.c_uint64)]


def _bind():
    L = _native.lib()
    if not getattr(L, "_traj_bound", False):
        P = ctypes.c_void_p
        L.llsv_traj_run.argtypes = [P, ctypes.c_int, P, ctypes.c_int, P, P, ctypes.c_int, ctypes.c_uint64, P, P, P, P]
        L.llsv_traj_run.restype = ctypes.c_int
        L._traj_bound = True
    return L


def pack_gates(gates, out=None, start=0):
    arr = out if out is not None else (TGate * len(gates))()
    for i, g in enumerate(gates):
        o = arr[start + i]
        if g[0] == "cnot":
            o.kind, o.q0, o.q1 = 1, int(g[1]), int(g[2])
        else:
            o.kind, o.q0 = 0, int(g[1])
            for j, v in enumerate(G.matrix(g)):
                o.m[2 * j], o.m[2 * j + 1] = v.real, v.imag
    return arr


def run_batch(circuits, seed: int):
    """circuits: list of (n, gates, gamma, lam, traj_id).  Returns (z (len, 32), jumps, hashes)."""
    total_gates = sum(len(c[1]) for c in circuits)
    gates = (TGate * total_gates)()
    items = (TItem * len(circuits))()
    gam, lam = [], []
    gpos = 0
    for i, (n, gl, gm, lm, tid) in enumerate(circuits):
        pack_gates(gl, gates, gpos)
        items[i].n_qubits, items[i].first_gate, items[i].n_gates = n, gpos, len(gl)
        items[i].first_rate, items[i].traj = len(gam), tid
        gpos += len(gl)
        gam += list(gm)
        lam += list(lm)
    return _run(items, gates, np.asarray(gam, np.float64), np.asarray(lam, np.float64), seed)


def run_trajectories(n, gates, gamma, lam, seed: int, n_traj: int, first_traj: int = 0, want_state=False):
    """n_traj trajectories (ids first_traj ..) of one circuit."""
    g = pack_gates(gates)
    items = (TItem * n_traj)()
    for i in range(n_traj):
        items[i].n_qubits, items[i].first_gate, items[i].n_gates = n, 0, len(gates)
        items[i].first_rate, items[i].traj = 0, first_traj + i
    return _run(items, g, np.asarray(gamma, np.float64), np.asarray(lam, np.float64), seed,
                state_n=n if want_state else None)


def _run(items, gates, gam, lam, seed, state_n=None):
    L = _bind()
    k = len(items)
    z = np.zeros((k, ZSTRIDE))
    j = np.zeros(k, dtype=np.uint32)
    h = np.zeros(k, dtype=np.uint64)
    st = np.zeros(1 << state_n, dtype=np.complex128) if state_n else None
    P = ctypes.c_void_p
    _native.check(L.llsv_traj_run(ctypes.cast(items, P), k, ctypes.cast(gates, P), len(gates),
                                  gam.ctypes.data_as(P), lam.ctypes.data_as(P), int(gam.size), seed,
                                  z.ctypes.data_as(P), j.ctypes.data_as(P), h.ctypes.data_as(P),
                                  st.ctypes.data_as(P) if st is not None else None))
    return (z, j, h, st) if state_n else (z, j, h)
