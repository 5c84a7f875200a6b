"""Quantum trajectories with variance-reduced unravelings on the MI355X engine: the Python
surface of the reference's `quantum_sim_trajectory` module (src/bindings_trajectory.cpp:60-133,
over src/TrajectorySolver.h).

`TrajectorySolver(n_qubits)` (n <= 10) keeps the reference's interface: a Hamiltonian (matrix or
callable t -> matrix), Lindblad channels gamma (L rho L^+ - 1/2 {L^+ L, rho}), and three
unravelings -- STANDARD (quantum jumps), PROJECTOR (Pauli-like channels rewritten with
C = sqrt(g) (1 - P)) and ANALOG (homodyne diffusion).  The plan (jump operators, H_eff and the
propagator exp(-i H_eff h), or the sampled H(t) nodes) is built as in src/TrajectorySolver.h:
268-290, on the device (llsv_lplan_create_h: H_eff by one GEMM, expm by GEMMs); the trajectories
then run together on the GPU (llsv_lplan_run): every step propagates all of them with one complex
GEMM on the FP64 matrix cores.  Trajectory i draws from the reference's
stream (std::mt19937_64 seeded with splitmix64(seed ^ splitmix64(i + 1))), so results follow the
reference's draws; set `solver.rng = "philox"` for device-side draws instead.
"""
from __future__ import annotations

import ctypes
import enum
import math

import numpy as np

from llsv import _native


class Unraveling(enum.IntEnum):
    STANDARD = 0
    PROJECTOR = 1
    ANALOG = 2


class Stats:
    def __init__(self, mean=0.0, std_error=0.0, n=0):
        self.mean, self.std_error, self.n = mean, std_error, n

    def __repr__(self):
        return f"Stats(mean={self.mean:.10g}, std_error={self.std_error:.3g}, n={self.n})"


def _fail(who, msg):
    raise ValueError(f"{who}: {msg}")


def expm(a) -> np.ndarray:
    """exp(A) by the reference's scaling and squaring (src/TrajectorySolver.h:97-111), on the GPU."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    out = np.empty_like(a)
    _native.check(_native.lib().llsv_expm(a.shape[0], a.ctypes.data_as(ctypes.c_void_p),
                                          out.ctypes.data_as(ctypes.c_void_p)))
    return out


def zgemm(a, b) -> np.ndarray:
    """a @ b on the trajectory GEMM kernel (rows and columns multiples of 64, inner size of 16)."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    c = np.empty((a.shape[0], b.shape[1]), dtype=np.complex128)
    _native.check(_native.lib().llsv_zgemm(a.shape[0], b.shape[1], a.shape[1], a.ctypes.data_as(ctypes.c_void_p),
                                           b.ctypes.data_as(ctypes.c_void_p), c.ctypes.data_as(ctypes.c_void_p)))
    return c


def _matmul(a, b):
    """a @ b: on the GPU kernel from 64 x 64 up (dims are powers of two), NumPy below."""
    if a.shape[0] >= 64 and a.shape[0] % 64 == 0 and b.shape[1] % 64 == 0 and a.shape[1] % 16 == 0:
        return zgemm(a, b)
    return a @ b


def gemm_ms(M: int, N: int, K: int, reps: int = 20) -> float:
    """Average device milliseconds of one M x K by K x N complex GEMM launch (HIP events)."""
    ms = ctypes.c_float(0.0)
    _native.check(_native.lib().llsv_zgemm_time(M, N, K, reps, ctypes.byref(ms)))
    return float(ms.value)


class TrajectorySolver:
    MAX_QUBITS = 10

    def __init__(self, n_qubits: int):
        n = int(n_qubits)
        if n < 1 or n > self.MAX_QUBITS:
            _fail("TrajectorySolver", f"n_qubits must be in [1, {self.MAX_QUBITS}]")
        self._n, self._dim = n, 1 << n
        self._h_const = np.zeros((self._dim, self._dim), dtype=np.complex128)
        self._h_fn = None
        self._channels = []  # dicts: l, gamma, pauli, p, g
        self._plan_cache = None  # ((kind, steps, h), _Plan) of the last constant-H plan
        self.rng = "reference"
        self.reset()

    # -- properties
    @property
    def n_qubits(self) -> int:
        return self._n

    @property
    def dim(self) -> int:
        return self._dim

    @property
    def n_channels(self) -> int:
        return len(self._channels)

    # -- model
    def _check_h(self, h, who):
        h = np.asarray(h, dtype=np.complex128)
        if h.shape != (self._dim, self._dim):
            _fail(who, "H must be dim x dim")
        if not np.all(np.isfinite(h)):
            _fail(who, "non-finite H")
        if np.linalg.norm(h - h.conj().T) > 1e-9 * max(1.0, np.linalg.norm(h)):
            _fail(who, "H is not Hermitian")
        return np.array(h, dtype=np.complex128, order="C", copy=True)  # the model keeps its own copy

    def set_hamiltonian(self, H):
        if callable(H) and not isinstance(H, np.ndarray):
            self._check_h(H(0.0), "set_hamiltonian")
            self._h_fn = H
            self._plan_cache = None
            return
        self._h_const = self._check_h(H, "set_hamiltonian")
        self._h_fn = None
        self._plan_cache = None

    def add_lindblad(self, L, gamma: float):
        who = "add_lindblad"
        l = np.asarray(L, dtype=np.complex128)
        if l.shape != (self._dim, self._dim):
            _fail(who, "operator must be dim x dim")
        if not np.all(np.isfinite(l)):
            _fail(who, "non-finite operator")
        gamma = float(gamma)
        if not (math.isfinite(gamma) and gamma >= 0.0):
            _fail(who, "gamma must be finite and >= 0")
        if gamma == 0.0:
            return
        ch = {"l": np.array(l, order="C", copy=True), "gamma": gamma, "pauli": False, "p": None, "g": 0.0}
        ok, p, c = self._pauli_like(l)
        if ok:
            ch.update(pauli=True, p=p, g=c * gamma)
        self._channels.append(ch)
        self._plan_cache = None

    def clear_lindblad(self):
        self._channels = []
        self._plan_cache = None

    def projector_applies(self, channel: int) -> bool:
        return bool(self._channels[int(channel)]["pauli"])

    @staticmethod
    def _pauli_like(l):
        """L Hermitian with L^2 = c 1, c > 0 (src/TrajectorySolver.h:316-326)."""
        if np.linalg.norm(l - l.conj().T) > 1e-9 * max(1.0, np.linalg.norm(l)):
            return False, None, 0.0
        l2 = _matmul(l, l)
        c = float(l2[0, 0].real)
        if not c > 1e-12:
            return False, None, 0.0
        if np.linalg.norm(l2 - c * np.eye(l.shape[0])) > 1e-9 * max(1.0, np.linalg.norm(l2)):
            return False, None, 0.0
        return True, l / math.sqrt(c), c

    def reset(self):
        self._psi = np.zeros(self._dim, dtype=np.complex128)
        self._psi[0] = 1.0

    def state(self) -> np.ndarray:
        return self._psi.copy()

    # -- checks (src/TrajectorySolver.h:330-345)
    def _checked_state(self, psi0, who):
        v = np.asarray(psi0, dtype=np.complex128).reshape(-1)
        if v.size != self._dim:
            _fail(who, "psi0 must have length 2^n_qubits")
        if not np.all(np.isfinite(v)):
            _fail(who, "non-finite psi0")
        nrm = float(np.linalg.norm(v))
        if not nrm > 1e-12:
            _fail(who, "psi0 has zero norm")
        return v / nrm

    @staticmethod
    def _check_time(duration, dt, who):
        duration, dt = float(duration), float(dt)
        if not (math.isfinite(duration) and duration >= 0.0):
            _fail(who, "duration must be finite and >= 0")
        if not (math.isfinite(dt) and dt > 0.0):
            _fail(who, "dt must be finite and > 0")
        if duration / dt > 2e9:
            _fail(who, "too many time steps")
        return duration, dt

    def _eval_h(self, t):
        h = np.asarray(self._h_fn(t), dtype=np.complex128)
        if h.shape != (self._dim, self._dim):
            _fail("H(t)", "returned a matrix of the wrong size")
        if not np.all(np.isfinite(h)):
            _fail("H(t)", "returned non-finite values")
        return h

    # -- plan (src/TrajectorySolver.h:268-290)
    def _plan(self, kind, duration, dt):
        kind = Unraveling(int(kind))
        if not self._channels:
            kind = Unraveling.STANDARD  # unitary evolution: exact propagator
        steps = max(1, int(math.ceil(duration / dt - 1e-9)))
        h = duration / steps
        # a constant-H plan is kept for the next call with the same unraveling and time grid (the
        # model setters drop it); a callable H is sampled afresh every call
        key = (int(kind), steps, h)
        if self._h_fn is None and self._plan_cache is not None and self._plan_cache[0] == key:
            return self._plan_cache[1]
        dim = self._dim
        ops = []
        for ch in self._channels:
            if kind == Unraveling.PROJECTOR and ch["pauli"]:
                j = math.sqrt(ch["g"]) * (np.eye(dim, dtype=np.complex128) - ch["p"])
            else:
                j = math.sqrt(ch["gamma"]) * ch["l"]
            ops.append(np.ascontiguousarray(j))
        jump = kind != Unraveling.ANALOG
        # H_nh = -(i/2) sum C^+ C, the propagator exp(-i (H + H_nh) h) (or the RK4 / Euler nodes
        # -i (H(t_j) + H_nh)) are assembled on the device from H and the operators
        if self._h_fn is None:
            hs = self._h_const
        else:
            n_nodes = 2 * steps + 1 if jump else steps
            if n_nodes * dim * dim * 16.0 > 1.0e9:
                _fail("TrajectorySolver", "time-dependent H(t) samples exceed 1 GB; increase dt")
            hs = np.stack([self._eval_h(0.5 * h * jn if jump else h * jn) for jn in range(n_nodes)])
        hs = np.ascontiguousarray(hs, dtype=np.complex128)
        n_h = 1 if hs.ndim == 2 else hs.shape[0]
        opsa = np.ascontiguousarray(np.stack(ops)) if ops else np.zeros((1, dim, dim), dtype=np.complex128)
        lib = _native.lib()
        p = ctypes.c_void_p()
        _native.check(lib.llsv_lplan_create_h(dim, 0 if jump else 1, steps, h, n_h,
                                              hs.ctypes.data_as(ctypes.c_void_p), len(ops),
                                              opsa.ctypes.data_as(ctypes.c_void_p), ctypes.byref(p)))
        plan = _Plan(lib, p, steps, jump)
        if self._h_fn is None:
            self._plan_cache = (key, plan)
        return plan

    def _run(self, plan, psi0, n, seed, rho_only=False):
        """Final states of trajectories 0 .. n-1 (n x dim), or with rho_only the sum of their
        |psi><psi| (dim x dim, formed on the device); batched so the host random table stays small."""
        d = self._dim
        out = None if rho_only else np.empty((n, d), dtype=np.complex128)
        acc = None
        part = np.empty((d, d), dtype=np.complex128) if rho_only else None
        per = (2 * plan.steps if plan.jump else plan.steps * max(1, self.n_channels)) * 8
        batch = max(1, min(n, (256 << 20) // max(per, 1)))
        psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
        rng = 0 if self.rng == "reference" else 1
        for f in range(0, n, batch):
            b = min(batch, n - f)
            st = None if rho_only else out[f:f + b].ctypes.data_as(ctypes.c_void_p)
            rs = part.ctypes.data_as(ctypes.c_void_p) if rho_only else None
            _native.check(plan.lib.llsv_lplan_run(plan.h, psi0.ctypes.data_as(ctypes.c_void_p), b, f,
                                                  int(seed) & ((1 << 64) - 1), rng, st, rs))
            if rho_only:
                if acc is None:
                    acc, part = part, (np.empty_like(part) if f + b < n else None)
                else:
                    acc += part
        return acc if rho_only else out

    # -- evolution (src/TrajectorySolver.h:158-256)
    def run_one(self, psi0, duration, dt, seed, kind=Unraveling.STANDARD):
        who = "run_one"
        psi = self._checked_state(psi0, who)
        duration, dt = self._check_time(duration, dt, who)
        if duration == 0.0:
            return psi
        plan = self._plan(kind, duration, dt)
        return self._run(plan, psi, 1, seed)[0]

    def evolve_ensemble(self, psi0, duration, dt, n_trajectories, kind=Unraveling.STANDARD, seed=1):
        who = "evolve_ensemble"
        start = self._checked_state(psi0, who)
        duration, dt = self._check_time(duration, dt, who)
        n = int(n_trajectories)
        if n < 1:
            _fail(who, "n_trajectories must be >= 1")
        if duration == 0.0:
            return np.outer(start, start.conj())
        rho = self._run(self._plan(kind, duration, dt), start, n, seed, rho_only=True)
        rho /= float(n)
        return rho

    def estimate_observable(self, psi0, duration, dt, n_trajectories, obs, kind=Unraveling.STANDARD, seed=1):
        who = "estimate_observable"
        start = self._checked_state(psi0, who)
        duration, dt = self._check_time(duration, dt, who)
        n = int(n_trajectories)
        if n < 1:
            _fail(who, "n_trajectories must be >= 1")
        if obs is None:
            _fail(who, "empty observable")
        if not callable(obs):
            raise TypeError("obs must be callable: ndarray -> float")
        if duration > 0.0:
            states = self._run(self._plan(kind, duration, dt), start, n, seed)
        else:
            states = np.tile(start, (n, 1))
        return self._stats([float(obs(s.copy())) for s in states])

    def estimate_expectation(self, psi0, duration, dt, n_trajectories, op, kind=Unraveling.STANDARD, seed=1):
        op = np.asarray(op, dtype=np.complex128)
        if op.shape != (self._dim, self._dim):
            raise ValueError("estimate_expectation: operator must be dim x dim")
        if np.linalg.norm(op - op.conj().T) > 1e-9 * max(1.0, np.linalg.norm(op)):
            raise ValueError("estimate_expectation: operator is not Hermitian")
        who = "estimate_observable"
        start = self._checked_state(psi0, who)
        duration, dt = self._check_time(duration, dt, who)
        n = int(n_trajectories)
        if n < 1:
            _fail(who, "n_trajectories must be >= 1")
        states = self._run(self._plan(kind, duration, dt), start, n, seed) if duration > 0.0 else np.tile(start, (n, 1))
        vals = np.real(np.einsum("ti,ij,tj->t", states.conj(), op, states))
        return self._stats([float(v) for v in vals])

    @staticmethod
    def _stats(v):
        n = len(v)
        st = Stats(n=n)
        s = 0.0
        for x in v:
            s += x
        st.mean = s / n
        if n > 1:
            ss = 0.0
            for x in v:
                ss += (x - st.mean) * (x - st.mean)
            st.std_error = math.sqrt(ss / (n - 1) / n)
        return st


class _Plan:
    def __init__(self, lib, h, steps, jump):
        self.lib, self.h, self.steps, self.jump = lib, h, steps, jump

    def __del__(self):
        try:
            if self.h:
                self.lib.llsv_lplan_destroy(self.h)
                self.h = None
        except Exception:  # interpreter shutdown
            pass
