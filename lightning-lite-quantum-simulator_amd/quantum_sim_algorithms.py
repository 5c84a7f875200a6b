"""Grover, QAOA and VQE on the MI355X state-vector engine.

The Python surface of the reference's `quantum_sim_algorithms` module
(src/bindings_algorithms.cpp:31-123), so its wrapper python/algorithms.py and scripts built on it
run unchanged with this module on the path.  The drivers follow src/Algorithms.h (Grover
:57-125, QAOA :151-200, VQE :226-311) and talk only to a gate-level backend with the
ll::Backend interface (src/Backend.h:25-45); `Backend` here is that interface on one GPU.

Backend factories: `make_gpu_backend(n)`.  The reference's `make_reference_backend`,
`make_eigen_backend` and `make_statevector_backend` are kept as names so its wrapper imports,
and all return the GPU backend (up to 34 qubits instead of 26 / 24 / 25).

The GPU backend queues gates and hands them to llsv_apply_program when the state is read
(probabilities(), state()) or the queue grows long, so a Grover iteration or a QAOA layer
costs a few passes over the state instead of one pass per gate.  Results equal gate-by-gate
application up to rounding (1e-12 on amplitudes).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from llsv import _native, program as _program

_PI = 3.14159265358979323846


# ---------------------------------------------------------------- the backend
class Backend:
    """Gate-level backend on the GPU: the ll::Backend interface (src/Backend.h:25-45).

    Qubit q is bit q of the basis index.  Bad qubits, duplicate MCZ qubits and non-finite
    angles raise ValueError (std::invalid_argument in the reference)."""

    MAX_QUBITS = 34
    _FLUSH_AT = 4096  # queued gates before an early hand-over

    def __init__(self, n_qubits: int = 1):
        self._lib = _native.lib()
        self._h = ctypes.c_void_p()
        self._n = 0
        self._ops = []
        self.reset(n_qubits)

    # -- state
    def reset(self, n_qubits: int):
        n = int(n_qubits)
        if n < 1 or n > self.MAX_QUBITS:
            raise ValueError(f"GpuBackend::reset: n_qubits must be in [1, {self.MAX_QUBITS}]")
        self._ops = []
        if self._h and n == self._n:
            _native.check(self._lib.llsv_reset(self._h))
            return
        self._free()
        h = ctypes.c_void_p()
        _native.check(self._lib.llsv_create(n, ctypes.byref(h)))
        self._h, self._n = h, n

    def n_qubits(self) -> int:
        return self._n

    def probabilities(self) -> np.ndarray:
        self._flush()
        out = np.empty(1 << self._n, dtype=np.float64)
        _native.check(self._lib.llsv_probabilities(self._h, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def state(self) -> np.ndarray:
        self._flush()
        out = np.empty(1 << self._n, dtype=np.complex128)
        _native.check(self._lib.llsv_copy_to_host(self._h, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    # -- gates (src/Backend.h:33-39)
    def h(self, q: int):
        self._push(("h", self._q(q)))

    def x(self, q: int):
        self._push(("x", self._q(q)))

    def rx(self, q: int, theta: float):
        self._push(("rx", self._q(q), self._angle(theta, "rx")))

    def ry(self, q: int, theta: float):
        self._push(("ry", self._q(q), self._angle(theta, "ry")))

    def rz(self, q: int, theta: float):
        self._push(("rz", self._q(q), self._angle(theta, "rz")))

    def cnot(self, control: int, target: int):
        c, t = self._q(control), self._q(target)
        if c == t:
            raise ValueError("GpuBackend::cnot: control == target")
        self._push(("cnot", c, t))

    def mcz(self, qubits):
        qs = [self._q(q) for q in qubits]
        if not qs:
            raise ValueError("GpuBackend::mcz: empty qubit list")
        if len(set(qs)) != len(qs):
            raise ValueError("GpuBackend::mcz: duplicate qubit")
        if len(qs) > 8:
            # a program op holds at most 8 qubits: run the queue, then this MCZ on its own
            self._flush()
            arr = (ctypes.c_int * len(qs))(*qs)
            _native.check(self._lib.llsv_apply_mcz(self._h, ctypes.cast(arr, ctypes.c_void_p), len(qs)))
            return
        self._push(("mcz", qs))

    # -- internals
    def _q(self, q) -> int:
        q = int(q)
        if q < 0 or q >= self._n:
            raise ValueError(f"GpuBackend: qubit {q} outside [0, {self._n})")
        return q

    @staticmethod
    def _angle(t, who) -> float:
        t = float(t)
        if not math.isfinite(t):
            raise ValueError(f"GpuBackend::{who}: non-finite angle")
        return t

    def _push(self, op):
        self._ops.append(op)
        if len(self._ops) >= self._FLUSH_AT:
            self._flush()

    def _flush(self):
        if not self._ops:
            return
        arr, _keep = _program.to_llsv_ops(self._ops)
        self._ops = []
        _native.check(self._lib.llsv_apply_program(self._h, ctypes.cast(arr, ctypes.c_void_p), len(arr)))

    def _free(self):
        if self._h:
            self._lib.llsv_destroy(self._h)
            self._h = ctypes.c_void_p()
            self._n = 0

    def __del__(self):
        try:
            self._free()
        except Exception:  # interpreter shutdown
            pass


def make_gpu_backend(n_qubits: int) -> Backend:
    """The GPU backend (src/bindings_algorithms.cpp:46-54 shape), up to 34 qubits."""
    return Backend(int(n_qubits))


make_reference_backend = make_gpu_backend
make_eigen_backend = make_gpu_backend
make_statevector_backend = make_gpu_backend


def _fail(who: str, msg: str):
    raise ValueError(f"{who}: {msg}")


# ---------------------------------------------------------------- Grover
class GroverResult:
    def __init__(self, probabilities, marked_index, iterations):
        self.probabilities = probabilities  # size 2^n after the final iteration
        self.marked_index = marked_index    # most probable basis state
        self.iterations = iterations        # Grover iterations applied


class Grover:
    @staticmethod
    def run(backend, n_qubits: int, marked_states, iterations: int = -1) -> GroverResult:
        """iterations = -1 picks round(pi/(4 theta) - 1/2), theta = asin sqrt(M/N)
        (src/Algorithms.h:60-95).  Duplicate marked states are merged."""
        who = "Grover::run"
        n = int(n_qubits)
        if n < 1 or n > 30:
            _fail(who, "n_qubits must be in [1, 30]")
        if int(iterations) < -1:
            _fail(who, "iterations must be >= 0, or -1 for automatic")
        n_states = 1 << n
        marked = sorted(set(int(m) for m in marked_states))
        for m in marked:
            if m < 0 or m >= n_states:
                _fail(who, "marked state outside [0, 2^n)")
        if not marked:
            _fail(who, "no marked states")
        if len(marked) >= n_states:
            _fail(who, "marked states must be a proper subset of the search space")
        k = int(iterations)
        if k < 0:
            theta = math.asin(math.sqrt(len(marked) / n_states))
            k = max(0, _round_half_away(_PI / (4.0 * theta) - 0.5))
        every = list(range(n))
        backend.reset(n)
        for q in every:
            backend.h(q)
        for _ in range(k):
            for m in marked:  # |m> -> -|m>: X on the zero bits, MCZ on all, undo
                zeros = [q for q in every if not (m >> q) & 1]
                for q in zeros:
                    backend.x(q)
                backend.mcz(every)
                for q in zeros:
                    backend.x(q)
            for q in every:  # inversion about the mean (global phase -1)
                backend.h(q)
                backend.x(q)
            backend.mcz(every)
            for q in every:
                backend.x(q)
                backend.h(q)
        p = np.asarray(backend.probabilities())
        if p.size != n_states:
            raise RuntimeError("Grover::run: backend returned wrong number of probabilities")
        return GroverResult(p, int(np.argmax(p)), k)


def _round_half_away(x: float) -> int:
    """std::lround: nearest integer, halves away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


# ---------------------------------------------------------------- QAOA
class Graph:
    def __init__(self, n_nodes: int = 0, edges=()):
        self.n_nodes = int(n_nodes)
        self.edges = [(int(i), int(j), float(w)) for i, j, w in edges]  # (i, j, w_ij), i != j

    def add_edge(self, i: int, j: int, weight: float = 1.0):
        self.edges.append((int(i), int(j), float(weight)))


class QAOAResult:
    def __init__(self, probabilities, best_cut):
        self.probabilities = probabilities  # size 2^n_nodes
        self.best_cut = best_cut            # cut weight of the most probable bitstring


class QAOA:
    @staticmethod
    def run(backend, graph: Graph, gammas, betas) -> QAOAResult:
        """p = len(gammas) = len(betas) layers of U_C(gamma) then U_B(beta) on |+>^n
        (src/Algorithms.h:129-189): U_C factor per edge = CNOT(i,j) RZ_j(-gamma w) CNOT(i,j),
        U_B = RX(2 beta) on every qubit."""
        who = "QAOA::run"
        n = graph.n_nodes
        if n < 1 or n > 30:
            _fail(who, "n_nodes must be in [1, 30]")
        gammas = [float(a) for a in gammas]
        betas = [float(b) for b in betas]
        if not gammas or len(gammas) != len(betas):
            _fail(who, "gammas and betas must be non-empty and of equal length")
        for i, j, w in graph.edges:
            if i < 0 or j < 0 or i >= n or j >= n:
                _fail(who, "edge endpoint outside [0, n_nodes)")
            if i == j:
                _fail(who, "self-loop edge")
            if not math.isfinite(w):
                _fail(who, "non-finite edge weight")
        if not all(math.isfinite(a) for a in gammas):
            _fail(who, "non-finite gamma")
        if not all(math.isfinite(b) for b in betas):
            _fail(who, "non-finite beta")
        backend.reset(n)
        for q in range(n):
            backend.h(q)
        for gamma, beta in zip(gammas, betas):
            for i, j, w in graph.edges:
                backend.cnot(i, j)
                backend.rz(j, -gamma * w)
                backend.cnot(i, j)
            for q in range(n):
                backend.rx(q, 2.0 * beta)
        p = np.asarray(backend.probabilities())
        if p.size != (1 << n):
            raise RuntimeError("QAOA::run: backend returned wrong number of probabilities")
        return QAOAResult(p, QAOA.cut_value(graph, int(np.argmax(p))))

    @staticmethod
    def cut_value(graph: Graph, z: int) -> float:
        """Cut weight sum w_ij [z_i != z_j] of the bitstring with index z."""
        c = 0.0
        for i, j, w in graph.edges:
            if ((z >> i) & 1) != ((z >> j) & 1):
                c += w
        return c


# ---------------------------------------------------------------- VQE
class VQEResult:
    def __init__(self, energy, params, iterations):
        self.energy = energy          # lowest energy seen
        self.params = params          # parameters that gave it
        self.iterations = iterations  # gradient steps taken


class VQE:
    """E(theta) = <psi|H|psi>, |psi> = prod_l [CNOT ring . prod_q RY_q(theta_{l,q})] |0>^n
    (src/Algorithms.h:203-216); gradients by the parameter shift, plain gradient descent."""

    def __init__(self, n_qubits: int, n_layers: int, seed: int = 1):
        if int(n_qubits) < 1 or int(n_qubits) > 30:
            _fail("VQE", "n_qubits must be in [1, 30]")
        if int(n_layers) < 1:
            _fail("VQE", "n_layers must be >= 1")
        self._n, self._layers, self._seed = int(n_qubits), int(n_layers), int(seed)

    @property
    def n_qubits(self) -> int:
        return self._n

    @property
    def n_layers(self) -> int:
        return self._layers

    @property
    def n_params(self) -> int:
        return self._n * self._layers

    def prepare(self, backend, params):
        """Reset to |0...0> and apply the ansatz; params ordered theta[l * n + q]."""
        params = [float(t) for t in np.asarray(params, dtype=np.float64).reshape(-1)]
        if len(params) != self.n_params:
            _fail("VQE::prepare", "wrong number of parameters")
        n = self._n
        backend.reset(n)
        for layer in range(self._layers):
            for q in range(n):
                backend.ry(q, params[layer * n + q])
            if n == 2:
                backend.cnot(0, 1)
            elif n > 2:
                for q in range(n):
                    backend.cnot(q, (q + 1) % n)

    def initial_params(self) -> np.ndarray:
        """The starting point: n_params draws of std::uniform_real_distribution(-0.1, 0.1) over
        std::mt19937_64(seed), as libstdc++ computes them (src/Algorithms.h:268-271)."""
        rng = _MT19937_64(self._seed)
        return np.array([_uniform_real(rng, -0.1, 0.1) for _ in range(self.n_params)])

    def optimize(self, backend, energy_fn, max_iter: int = 200, lr: float = 0.05) -> VQEResult:
        """Gradient descent with parameter-shift gradients (src/Algorithms.h:255-307): stops when
        |grad| < 1e-6 or after max_iter steps, keeps the lowest energy seen, and leaves that state
        prepared on the backend."""
        who = "VQE::optimize"
        max_iter = int(max_iter)
        lr = float(lr)
        if max_iter < 0:
            _fail(who, "max_iter must be >= 0")
        if not (math.isfinite(lr) and lr > 0.0):
            _fail(who, "lr must be finite and > 0")
        if energy_fn is None:
            _fail(who, "energy_fn is empty")

        def ev(th):
            e = float(energy_fn(np.array(th)))
            if not math.isfinite(e):
                raise RuntimeError(f"{who}: non-finite energy")
            return e

        theta = [float(t) for t in self.initial_params()]
        best_e, best_p, steps = ev(theta), list(theta), 0
        shift = _PI / 2.0
        for _ in range(max_iter):
            grad = []
            for k in range(len(theta)):
                t0 = theta[k]
                theta[k] = t0 + shift
                ep = ev(theta)
                theta[k] = t0 - shift
                em = ev(theta)
                theta[k] = t0
                grad.append(0.5 * (ep - em))
            if math.sqrt(sum(g * g for g in grad)) < 1e-6:
                break
            theta = [t - lr * g for t, g in zip(theta, grad)]
            steps += 1
            e = ev(theta)
            if e < best_e:
                best_e, best_p = e, list(theta)
        self.prepare(backend, best_p)
        return VQEResult(best_e, np.array(best_p), steps)


class _MT19937_64:
    """std::mt19937_64 (C++ [rand.predef]: w = 64, n = 312, m = 156, r = 31, a, u, d, s, b, t, c,
    l, f as the standard lists them)."""

    _M = (1 << 64) - 1

    def __init__(self, seed: int):
        self.mt = [0] * 312
        self.mt[0] = seed & self._M
        for i in range(1, 312):
            prev = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & self._M
        self.idx = 312

    def _twist(self):
        upper, lower = 0xFFFFFFFF80000000, 0x7FFFFFFF
        mt = self.mt
        for i in range(312):
            y = (mt[i] & upper) | (mt[(i + 1) % 312] & lower)
            v = mt[(i + 156) % 312] ^ (y >> 1)
            if y & 1:
                v ^= 0xB5026F5AA96619E9
            mt[i] = v
        self.idx = 0

    def __call__(self) -> int:
        if self.idx >= 312:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & self._M


def _uniform_real(rng: _MT19937_64, a: float, b: float) -> float:
    """libstdc++ uniform_real_distribution<double>: generate_canonical<double, 53> takes one
    64-bit draw x and returns double(x) / 2^64 (1 - ulp if that rounds to 1), then
    r * (b - a) + a."""
    r = float(rng()) / 18446744073709551616.0
    if r >= 1.0:
        r = math.nextafter(1.0, 0.0)
    return r * (b - a) + a
