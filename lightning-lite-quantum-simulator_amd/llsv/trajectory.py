"""Batched gate-level noisy trajectories on the GPU (llsv_traj_run).

After every gate, each qubit it touches goes through amplitude damping AD(gamma_q) and then
phase damping PD(lambda_q), unravelled by Kraus sampling with Philox4x32-10 streams keyed by
(seed, trajectory).  One persistent launch runs a whole batch -- many trajectories of one
circuit, or many circuits of different sizes.
"""
from __future__ import annotations

import ctypes
import math
import os

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


def _pack_batch(circuits):
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
    return items, gates, np.asarray(gam, np.float64), np.asarray(lam, np.float64)


def set_default_config(threads: int, per_cu: int, fused):
    """Process-wide default launch configuration (llsv_traj_default_config); fused: 0 / False (one
    pass per gate), 1 / True (deferred operators, CNOTs held in pairs), 2 (deferred operators,
    CNOTs held in lists that run through LDS tiles)."""
    _native.check(_native.lib().llsv_traj_default_config(threads, per_cu, int(fused)))


def load_tuning(path: str):
    """Adds the tuner choices stored in a JSON file to the process cache (llsv_traj_tune_load)."""
    _native.check(_native.lib().llsv_traj_tune_load(os.fsencode(path)))


def save_tuning(path: str):
    """Writes the process's tuner choices to a JSON file (llsv_traj_tune_save)."""
    _native.check(_native.lib().llsv_traj_tune_save(os.fsencode(path)))


def tuning_entries() -> int:
    """Number of batch shapes with a tuned configuration in the process cache."""
    return int(_native.lib().llsv_traj_tune_entries())


class TrajectoryPlan:
    """A batch prepared for repeated runs (llsv_traj_plan_*): uploaded once, launch configuration
    chosen by the autotuner (tune=1 / True), the cache or the default (0), or the default only (-1);
    one run captured in a hipGraph.

    circuits: list of (n, gates, gamma, lam, traj_id), as for run_batch; or use from_circuit."""

    def __init__(self, circuits, tune=True):
        self._items, self._gates, self._gam, self._lam = _pack_batch(circuits)
        self._create(tune)

    @classmethod
    def from_circuit(cls, n, gates, gamma, lam, n_traj: int, first_traj: int = 0, tune=True):
        self = cls.__new__(cls)
        self._gates = pack_gates(gates)
        self._items = (TItem * n_traj)()
        for i in range(n_traj):
            it = self._items[i]
            it.n_qubits, it.first_gate, it.n_gates, it.first_rate, it.traj = n, 0, len(gates), 0, first_traj + i
        self._gam = np.asarray(gamma, np.float64)
        self._lam = np.asarray(lam, np.float64)
        self._create(tune)
        return self

    def _create(self, tune):
        L = _native.lib()
        P = ctypes.c_void_p
        self._lib = L
        self._h = P()
        _native.check(L.llsv_traj_plan_create(ctypes.cast(self._items, P), len(self._items), ctypes.cast(self._gates, P),
                                              len(self._gates), self._gam.ctypes.data_as(P), self._lam.ctypes.data_as(P),
                                              int(self._gam.size), int(tune), ctypes.byref(self._h)))

    def config(self):
        """(threads per trajectory, workgroups, fused mode 0 / 1 / 2, tuning ms)."""
        t, w, f, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _native.check(self._lib.llsv_traj_plan_config(self._h, ctypes.byref(t), ctypes.byref(w), ctypes.byref(f),
                                                      ctypes.byref(ms)))
        return t.value, w.value, f.value, ms.value

    def run(self, seed: int):
        """(z (items, 32), jumps, hashes) for this seed."""
        k = len(self._items)
        z = np.zeros((k, ZSTRIDE))
        j = np.zeros(k, dtype=np.uint32)
        h = np.zeros(k, dtype=np.uint64)
        P = ctypes.c_void_p
        _native.check(self._lib.llsv_traj_plan_run(self._h, seed, z.ctypes.data_as(P), j.ctypes.data_as(P),
                                                   h.ctypes.data_as(P)))
        return z, j, h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.llsv_traj_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def run_batch(circuits, seed: int):
    """circuits: list of (n, gates, gamma, lam, traj_id).  Returns (z (len, 32), jumps, hashes)."""
    items, gates, gam, lam = _pack_batch(circuits)
    return _run(items, gates, gam, lam, seed)


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
