"""Zero-noise extrapolation on the MI355X engine: the Python surface of the reference's
`quantum_sim_mitigation` module (src/bindings_mitigation.cpp:77-176, over src/Mitigation.h and
src/Mitigation.cpp).

* `NoisyBackend(inner, p, seed=1)` wraps a gate-level backend; each gate U runs as U (U^+ U)^k at
  scale 2k + 1 with a sampled depolarizing site after every application (src/Mitigation.h:
  196-303).  Its random stream is the reference's (std::mt19937_64, same draws in the same
  order), kept in libllsv (llsv_noisy_*), so one seed gives the reference's Pauli insertions
  whether gates are applied one at a time or trajectories run in a batch.
* `ZNE(trajectories).extrapolate(noisy, circuit_fn, observable_fn, scales)` (src/Mitigation.h:
  340-390): for states of up to 20 qubits the circuit is recorded once and every scale's
  trajectories run in one GPU launch (llsv_noisy_run), each trajectory's state in LDS (up to 12
  qubits) or in a per-workgroup device-memory slot; a `Diagonal` observable is evaluated in the
  kernel, any other callable on each final state.
  Larger states run trajectory by trajectory through `noisy` exactly like the reference's loop.
  Mean, standard error (Welford, in trajectory order), Richardson weights and the weighted
  error follow src/Mitigation.h:357-390.
* `DensityNoiseBackend(p, n)` and `zne_exact` (src/Mitigation.cpp:54-246): the exact
  density-matrix counterpart, on the GPU: vec(rho) is a 2n-qubit state (rho[r, c] at index
  r + 2^n c), a gate U acts as U on the row qubits and conj(U) on the column qubits, the
  depolarizing channel as a dense 4x4 (one qubit) or 16x16 (two qubits) superoperator.

Extension: `NoisyBackend(..., rng="philox")` draws the noise with counter-based Philox on the
device instead (batched runs only), so the host draws nothing.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from llsv import _native
import quantum_sim_algorithms as _alg

MAX_ZNE_SCALES = 8
MAX_ZNE_SCALE = 99
DEFAULT_ZNE_TRAJECTORIES = 1000
MAX_DM_QUBITS = 10
MAX_BATCH_QUBITS = 20  # LLSV_ZNE_MAX_QUBITS
MAX_SITE_QUBITS = 16  # qubits of one llsv_noisy_site draw / one llsv_zgate (two code bits per qubit)
MAX_BATCH_APPS = 1 << 22  # gate applications of one folded circuit in llsv_noisy_run


class _NQ(int):
    """n_qubits as both a property value and a call (the algorithms module's backends have a
    method, the mitigation module's a property)."""

    def __call__(self) -> int:
        return int(self)


def _nq_of(b) -> int:
    v = b.n_qubits
    return int(v() if callable(v) and not isinstance(v, int) else v)


def _fail(who: str, msg: str):
    raise ValueError(f"{who}: {msg}")


# ---------------------------------------------------------------- backends
Backend = _alg.Backend


class ReferenceBackend(_alg.Backend):
    """The GPU state-vector backend (the reference binds its scalar ReferenceBackend here)."""

    def __init__(self, n_qubits: int = 1):
        super().__init__(n_qubits)

    @property
    def n_qubits(self) -> _NQ:  # property in the mitigation module (src/bindings_mitigation.cpp:89)
        return _NQ(self._n)


class Diagonal:
    """Observable sum_i w_i p_i over the basis probabilities.  Callable on any backend, like an
    observable_fn; ZNE.extrapolate evaluates it inside the trajectory kernel."""

    def __init__(self, weights):
        self.weights = np.ascontiguousarray(weights, dtype=np.float64)
        n = self.weights.size.bit_length() - 1
        if self.weights.ndim != 1 or (1 << n) != self.weights.size:
            raise ValueError("Diagonal: weights must be a 1-D array of length 2^n")
        self.n = n

    @classmethod
    def from_function(cls, n: int, f):
        """weights[i] = f(i) for every basis index i."""
        return cls(np.array([float(f(i)) for i in range(1 << n)]))

    def __call__(self, backend) -> float:
        return float(np.dot(self.weights, np.asarray(backend.probabilities(), dtype=np.float64)))


class _Handle:
    def __init__(self, p: float, seed: int, rng: int):
        self.lib = _native.lib()
        h = ctypes.c_void_p()
        _native.check(self.lib.llsv_noisy_create(float(p), int(seed) & ((1 << 64) - 1), rng, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                self.lib.llsv_noisy_destroy(self.h)
                self.h = ctypes.c_void_p()
        except Exception:  # interpreter shutdown
            pass


class NoisyBackend:
    """Folds every gate and applies sampled depolarizing noise after each application
    (src/Mitigation.h:114-303).  `inner` must outlive this object."""

    def __init__(self, inner, p: float, seed: int = 1, rng: str = "reference"):
        p = float(p)
        if not (0.0 <= p <= 1.0):
            _fail("NoisyBackend", "p must be in [0, 1]")
        if rng not in ("reference", "philox"):
            raise ValueError("NoisyBackend: rng must be 'reference' or 'philox'")
        self._inner, self._p, self._folds = inner, p, 0
        self._rng = rng
        self._z = _Handle(p, seed, 0 if rng == "reference" else 1)
        self._logical = 0
        self._physical = 0

    # -- properties (src/bindings_mitigation.cpp:112-120)
    @property
    def probability(self) -> float:
        return self._p

    @property
    def inner(self):
        return self._inner

    @property
    def rng(self) -> str:
        return self._rng

    def reseed(self, seed: int):
        _native.check(self._z.lib.llsv_noisy_reseed(self._z.h, int(seed) & ((1 << 64) - 1)))

    def set_scale(self, scale: int):
        s = int(scale) if float(scale) == int(scale) else -1
        if s < 1 or s > MAX_ZNE_SCALE or s % 2 == 0:
            _fail("NoisyBackend::set_scale", f"scale must be an odd integer in [1, {MAX_ZNE_SCALE}]")
        self._folds = (s - 1) // 2

    @property
    def scale(self) -> int:
        return 2 * self._folds + 1

    @property
    def logical_gates(self) -> int:
        return self._logical

    @property
    def physical_gates(self) -> int:
        return self._physical

    # -- Backend interface
    def reset(self, n_qubits: int):
        self._inner.reset(n_qubits)
        self._logical = 0
        self._physical = 0

    @property
    def n_qubits(self) -> _NQ:
        return _NQ(_nq_of(self._inner))

    def probabilities(self):
        return self._inner.probabilities()

    def state(self):
        return self._inner.state()

    def h(self, q):
        self._fold(lambda: self._inner.h(q), lambda: self._inner.h(q), [q])

    def x(self, q):
        self._fold(lambda: self._inner.x(q), lambda: self._inner.x(q), [q])

    def rx(self, q, t):
        self._fold(lambda: self._inner.rx(q, t), lambda: self._inner.rx(q, -t), [q])

    def ry(self, q, t):
        self._fold(lambda: self._inner.ry(q, t), lambda: self._inner.ry(q, -t), [q])

    def rz(self, q, t):
        self._fold(lambda: self._inner.rz(q, t), lambda: self._inner.rz(q, -t), [q])

    def cnot(self, control, target):
        self._fold(lambda: self._inner.cnot(control, target), lambda: self._inner.cnot(control, target),
                   [control, target])

    def mcz(self, qubits):
        qs = list(qubits)
        self._fold(lambda: self._inner.mcz(qs), lambda: self._inner.mcz(qs), qs)

    # -- internals
    def _fold(self, fwd, inv, qs):
        fwd()  # an invalid gate raises here, before any draw (src/Mitigation.h:223-225)
        self._noise(qs)
        for _ in range(self._folds):
            inv()
            self._noise(qs)
            fwd()
            self._noise(qs)
        self._logical += 1
        self._physical += 1 + 2 * self._folds

    def _noise(self, qs):
        if self._p == 0.0:
            return
        code = ctypes.c_uint32(0)
        if len(qs) <= MAX_SITE_QUBITS:
            _native.check(self._z.lib.llsv_noisy_site(self._z.h, len(qs), ctypes.byref(code)))
            c = code.value
        else:
            # a gate on >= 3 qubits gets an independent one-qubit channel per qubit
            # (src/Mitigation.h:196-204), so drawing them one qubit at a time consumes the stream
            # exactly as one wide draw would
            c = 0
            for i in range(len(qs)):
                _native.check(self._z.lib.llsv_noisy_site(self._z.h, 1, ctypes.byref(code)))
                c |= code.value << (2 * i)
        for i, q in enumerate(qs):
            w = (c >> (2 * i)) & 3
            if w in (2, 3):
                self._inner.mcz([q])
            if w in (1, 2):
                self._inner.x(q)


# ---------------------------------------------------------------- circuit recording
_ZG = {"h": 0, "x": 1, "rx": 2, "ry": 3, "rz": 4, "cnot": 5, "mcz": 6}


class ZGate(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("nq", ctypes.c_int32), ("q", ctypes.c_int8 * 16),
                ("theta", ctypes.c_double)]


class _Recorder:
    """Collects circuit_fn's gates, with the checks a backend makes."""

    def __init__(self, n: int):
        self.n = n
        self.gates = []

    @property
    def n_qubits(self) -> _NQ:
        return _NQ(self.n)

    def _q(self, q):
        q = int(q)
        if q < 0 or q >= self.n:
            raise ValueError(f"GpuBackend: qubit {q} outside [0, {self.n})")
        return q

    @staticmethod
    def _t(t, who):
        t = float(t)
        if not math.isfinite(t):
            raise ValueError(f"GpuBackend::{who}: non-finite angle")
        return t

    def h(self, q):
        self.gates.append(("h", self._q(q)))

    def x(self, q):
        self.gates.append(("x", self._q(q)))

    def rx(self, q, t):
        self.gates.append(("rx", self._q(q), self._t(t, "rx")))

    def ry(self, q, t):
        self.gates.append(("ry", self._q(q), self._t(t, "ry")))

    def rz(self, q, t):
        self.gates.append(("rz", self._q(q), self._t(t, "rz")))

    def cnot(self, c, t):
        c, t = self._q(c), self._q(t)
        if c == t:
            raise ValueError("GpuBackend::cnot: control == target")
        self.gates.append(("cnot", c, t))

    def mcz(self, qubits):
        qs = [self._q(q) for q in qubits]
        if not qs:
            raise ValueError("GpuBackend::mcz: empty qubit list")
        if len(set(qs)) != len(qs):
            raise ValueError("GpuBackend::mcz: duplicate qubit")
        self.gates.append(("mcz", qs))

    def reset(self, n_qubits):
        pass

    def probabilities(self):
        raise RuntimeError("ZNE: circuit_fn must not read the state")

    state = probabilities


def pack_gates(gates):
    """ctypes array of llsv_zgate for gate tuples ('h', q), ('rx', q, t), ('cnot', c, t), ('mcz', [qs])."""
    arr = (ZGate * max(1, len(gates)))()
    for i, g in enumerate(gates):
        o = arr[i]
        o.kind = _ZG[g[0]]
        qs = list(g[1]) if g[0] == "mcz" else ([g[1], g[2]] if g[0] == "cnot" else [g[1]])
        if len(qs) > MAX_SITE_QUBITS:
            raise ValueError(f"mcz: at most {MAX_SITE_QUBITS} qubits in one batched gate")
        o.nq = len(qs)
        for j, q in enumerate(qs):
            o.q[j] = int(q)
        o.theta = float(g[2]) if g[0] in ("rx", "ry", "rz") else 0.0
    return arr


class _FinalState:
    """A read-only backend view of one trajectory's final state."""

    def __init__(self, n: int, amps: np.ndarray):
        self._n, self._a = n, amps

    @property
    def n_qubits(self) -> _NQ:
        return _NQ(self._n)

    def state(self) -> np.ndarray:
        return self._a.copy()

    def probabilities(self) -> np.ndarray:
        return self._a.real ** 2 + self._a.imag ** 2


def batchable(gates, scale: int) -> bool:
    """Whether llsv_noisy_run takes this folded circuit: every gate's noise fits one code word
    (<= 16 qubits) and the folded circuit has <= 2^22 gate applications.  ZNE.extrapolate runs
    other scales through the per-gate loop, which continues the same random stream."""
    if len(gates) * int(scale) > MAX_BATCH_APPS:
        return False
    return all(len(g[1]) <= MAX_SITE_QUBITS for g in gates if g[0] == "mcz")


def run_trajectories(noisy: NoisyBackend, n: int, gates, scale: int, n_traj: int, observable):
    """Per-trajectory observable values of n_traj trajectories at one scale, batched on the GPU
    (n <= 20), continuing `noisy`'s stream."""
    arr = pack_gates(gates)
    L = noisy._z.lib
    vals = np.empty(int(n_traj), dtype=np.float64)
    if isinstance(observable, Diagonal):
        if observable.n != n:
            raise ValueError("Diagonal observable has the wrong number of qubits")
        _native.check(L.llsv_noisy_run(noisy._z.h, n, ctypes.cast(arr, ctypes.c_void_p), len(gates), int(scale),
                                       int(n_traj), observable.weights.ctypes.data_as(ctypes.c_void_p),
                                       vals.ctypes.data_as(ctypes.c_void_p), None))
        return vals
    chunk = max(1, (64 << 20) // (16 << n))
    for t0 in range(0, int(n_traj), chunk):
        nt = min(chunk, int(n_traj) - t0)
        amps = np.empty((nt, 1 << n), dtype=np.complex128)
        _native.check(L.llsv_noisy_run(noisy._z.h, n, ctypes.cast(arr, ctypes.c_void_p), len(gates), int(scale),
                                       nt, None, None, amps.ctypes.data_as(ctypes.c_void_p)))
        for t in range(nt):
            vals[t0 + t] = float(observable(_FinalState(n, amps[t])))
    return vals


# ---------------------------------------------------------------- ZNE
class ZNEResult:
    def __init__(self):
        self.value = 0.0
        self.std_error = 0.0
        self.scales = []
        self.values = []
        self.std_errors = []
        self.weights = []
        self.trajectories = 0

    def __repr__(self):
        return (f"ZNEResult(value={self.value:.10g}, std_error={self.std_error:.3g}, "
                f"n_scales={len(self.scales)}, trajectories={self.trajectories})")


def _welford(values):
    mean, m2 = 0.0, 0.0
    for t, v in enumerate(values):
        if not math.isfinite(v):
            _fail("ZNE::extrapolate", "observable returned a non-finite value")
        d = v - mean
        mean += d / (t + 1)
        m2 += d * (v - mean)
    return mean, m2


class ZNE:
    """Zero-noise extrapolation with Richardson weights over folded noise scales."""

    def __init__(self, trajectories: int = DEFAULT_ZNE_TRAJECTORIES):
        if int(trajectories) < 1:
            _fail("ZNE", "trajectories must be >= 1")
        self._trajectories = int(trajectories)

    @property
    def trajectories(self) -> int:
        return self._trajectories

    @staticmethod
    def default_scales():
        return [1.0, 3.0, 5.0]

    @staticmethod
    def validate_scales(scales):
        who = "ZNE::validate_scales"
        scales = [float(s) for s in scales]
        if len(scales) < 2 or len(scales) > MAX_ZNE_SCALES:
            _fail(who, f"need between 2 and {MAX_ZNE_SCALES} noise scales")
        for s in scales:
            if (not math.isfinite(s)) or s < 1.0 or s > MAX_ZNE_SCALE or s != math.floor(s) or math.fmod(s, 2.0) != 1.0:
                _fail(who, f"each scale must be an odd integer in [1, {MAX_ZNE_SCALE}]")
        if len(set(scales)) != len(scales):
            _fail(who, "scales must be distinct")

    @staticmethod
    def richardson_weights(scales):
        """c_k = prod_{j != k} lambda_j / (lambda_j - lambda_k) (src/Mitigation.h:300-314)."""
        who = "ZNE::richardson_weights"
        scales = [float(s) for s in scales]
        if len(scales) < 2:
            _fail(who, "need at least 2 scales")
        for s in scales:
            if not math.isfinite(s) or s == 0.0:
                _fail(who, "scales must be finite and nonzero")
        c = [1.0] * len(scales)
        for k in range(len(scales)):
            for j in range(len(scales)):
                if j == k:
                    continue
                if scales[j] == scales[k]:
                    _fail(who, "scales must be distinct")
                c[k] *= scales[j] / (scales[j] - scales[k])
        return c

    @staticmethod
    def richardson(scales, values):
        if len(values) != len(scales):
            _fail("ZNE::richardson", "scales and values differ in length")
        for v in values:
            if not math.isfinite(float(v)):
                _fail("ZNE::richardson", "non-finite value")
        c = ZNE.richardson_weights(scales)
        e = 0.0
        for ck, v in zip(c, values):
            e += ck * float(v)
        return e

    def extrapolate(self, noisy: NoisyBackend, circuit_fn, observable_fn, scales=None) -> ZNEResult:
        who = "ZNE::extrapolate"
        if scales is None:
            scales = self.default_scales()
        if circuit_fn is None:
            _fail(who, "circuit_fn is empty")
        if observable_fn is None:
            _fail(who, "observable_fn is empty")
        if not callable(circuit_fn):
            raise TypeError("circuit_fn must be callable")
        if not callable(observable_fn):
            raise TypeError("observable_fn must be callable")
        self.validate_scales(scales)
        saved = noisy.scale
        try:
            out = ZNEResult()
            out.scales = [float(s) for s in scales]
            out.weights = self.richardson_weights(scales)
            noisy_run = noisy.probability > 0.0
            ntr = self._trajectories if noisy_run else 1
            out.trajectories = ntr
            n = _nq_of(noisy)
            gates = None
            if n <= MAX_BATCH_QUBITS:
                rec = _Recorder(n)
                circuit_fn(rec)
                gates = rec.gates
                if not gates:
                    raise RuntimeError(f"{who}: circuit_fn applied no gates to the backend")
            for s in out.scales:
                scale = int(s)
                noisy.set_scale(scale)
                if gates is not None and batchable(gates, scale):
                    vals = run_trajectories(noisy, n, gates, scale, ntr, observable_fn)
                    noisy.reset(n)  # the counters of one trajectory, as after the reference's loop
                    noisy._logical, noisy._physical = len(gates), scale * len(gates)
                else:
                    vals = []
                    for _ in range(ntr):
                        noisy.reset(n)
                        circuit_fn(noisy)
                        if noisy.logical_gates == 0:
                            raise RuntimeError(f"{who}: circuit_fn applied no gates to the backend")
                        if noisy.physical_gates != scale * noisy.logical_gates:
                            raise RuntimeError(f"{who}: folded gate count is inconsistent")
                        vals.append(float(observable_fn(noisy)))
                mean, m2 = _welford([float(v) for v in vals])
                if ntr > 1:
                    se = math.sqrt(max(0.0, m2 / (ntr - 1)) / ntr)
                elif noisy_run:
                    se = float("nan")
                else:
                    se = 0.0
                out.values.append(mean)
                out.std_errors.append(se)
            var = 0.0
            for ck, v, se in zip(out.weights, out.values, out.std_errors):
                out.value += ck * v
                var += ck * ck * se * se
            out.std_error = math.sqrt(var)
            return out
        finally:
            noisy.set_scale(saved)


# ---------------------------------------------------------------- exact: density matrix on the GPU
_PAULI = [np.eye(2, dtype=complex), np.array([[0, 1], [1, 0]], dtype=complex),
          np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]], dtype=complex)]


def _depolarizing_superop(m: int, p: float) -> np.ndarray:
    """sum_K K (x) conj(K) for the depolarizing channel on m qubits (src/Noise.cpp:316-333), as
    an operator on (row qubits, column qubits): index bits 0..m-1 rows, m..2m-1 columns."""
    d = 1 << m
    S = np.zeros((d * d, d * d), dtype=complex)
    for code in range(4 ** m):
        P = np.array([[1.0]], dtype=complex)
        for j in range(m - 1, -1, -1):
            P = np.kron(P, _PAULI[(code >> (2 * j)) & 3])
        w = (1.0 - p) if code == 0 else p / (4 ** m - 1)
        # vec index = r + d c: operator kron(conj(P) on columns, P on rows)
        S += w * np.kron(P.conj(), P)
    return S


class DensityNoiseBackend:
    """Exact density-matrix counterpart of NoisyBackend (src/Mitigation.cpp:54-205): every
    application (folds included) is followed by the depolarizing channel on the touched qubits;
    vec(rho) is a 2n-qubit state on the GPU."""

    def __init__(self, p: float, n_qubits: int = 1):
        p = float(p)
        if not (0.0 <= p <= 1.0):
            _fail("DensityNoiseBackend", "p must be in [0, 1]")
        self._p, self._folds, self._logical, self._physical = p, 0, 0, 0
        self._sv = None
        self._S1 = _depolarizing_superop(1, p)
        self._S2 = _depolarizing_superop(2, p)
        self.reset(n_qubits)

    @property
    def probability(self) -> float:
        return self._p

    def set_scale(self, scale: int):
        s = int(scale) if float(scale) == int(scale) else -1
        if s < 1 or s > MAX_ZNE_SCALE or s % 2 == 0:
            _fail("DensityNoiseBackend::set_scale", f"scale must be an odd integer in [1, {MAX_ZNE_SCALE}]")
        self._folds = (s - 1) // 2

    @property
    def scale(self) -> int:
        return 2 * self._folds + 1

    @property
    def logical_gates(self) -> int:
        return self._logical

    @property
    def physical_gates(self) -> int:
        return self._physical

    def reset(self, n_qubits: int):
        n = int(n_qubits)
        if n < 1 or n > MAX_DM_QUBITS:
            _fail("DensityNoiseBackend::reset", f"n_qubits must be in [1, {MAX_DM_QUBITS}]")
        self._n = n
        if self._sv is None or self._sv.n_qubits() != 2 * n:
            self._sv = _alg.Backend(2 * n)
        else:
            self._sv.reset(2 * n)
        self._logical = 0
        self._physical = 0

    @property
    def n_qubits(self) -> _NQ:
        return _NQ(self._n)

    def density(self) -> np.ndarray:
        v = self._sv.state()
        d = 1 << self._n
        return v.reshape(d, d).T.copy()  # rho[r, c] at index r + d c

    def probabilities(self) -> np.ndarray:
        v = self._sv.state()
        d = 1 << self._n
        return np.ascontiguousarray(v[np.arange(d) * (d + 1)].real)

    def state(self):
        raise RuntimeError("DensityNoiseBackend::state: the state is mixed, use density()")

    # -- gates
    def _check(self, q, who="DensityNoiseBackend"):
        q = int(q)
        if q < 0 or q >= self._n:
            _fail(who, f"qubit {q} outside [0, {self._n})")
        return q

    @staticmethod
    def _angle(t, who):
        t = float(t)
        if not math.isfinite(t):
            _fail(f"DensityNoiseBackend::{who}", "non-finite angle")
        return t

    def _unitary1(self, q, m):
        n = self._n
        self._sv._push(("dense", [q], m))
        self._sv._push(("dense", [q + n], m.conj()))

    def _noise(self, qs):
        if self._p == 0.0:
            return
        n = self._n
        if len(qs) == 1:
            self._sv._push(("dense", [qs[0], qs[0] + n], self._S1))
        elif len(qs) == 2:
            self._sv._push(("dense", [qs[0], qs[1], qs[0] + n, qs[1] + n], self._S2))
        else:
            for q in qs:
                self._sv._push(("dense", [q, q + n], self._S1))

    def _fold1(self, q, fwd, inv):
        q = self._check(q)
        self._fold(lambda: self._unitary1(q, fwd), lambda: self._unitary1(q, inv), [q])

    def _fold(self, fwd, inv, qs):
        fwd()
        self._noise(qs)
        for _ in range(self._folds):
            inv()
            self._noise(qs)
            fwd()
            self._noise(qs)
        self._logical += 1
        self._physical += 1 + 2 * self._folds

    @staticmethod
    def _rot(kind, t):
        c, s = math.cos(0.5 * t), math.sin(0.5 * t)
        if kind == "rx":
            return np.array([[c, -1j * s], [-1j * s, c]])
        if kind == "ry":
            return np.array([[c, -s], [s, c]], dtype=complex)
        return np.array([[complex(c, -s), 0], [0, complex(c, s)]])

    def h(self, q):
        r = 1.0 / math.sqrt(2.0)
        u = np.array([[r, r], [r, -r]], dtype=complex)
        self._fold1(q, u, u)

    def x(self, q):
        u = np.array([[0, 1], [1, 0]], dtype=complex)
        self._fold1(q, u, u)

    def rx(self, q, t):
        t = self._angle(t, "rx")
        self._fold1(q, self._rot("rx", t), self._rot("rx", -t))

    def ry(self, q, t):
        t = self._angle(t, "ry")
        self._fold1(q, self._rot("ry", t), self._rot("ry", -t))

    def rz(self, q, t):
        t = self._angle(t, "rz")
        self._fold1(q, self._rot("rz", t), self._rot("rz", -t))

    def cnot(self, control, target):
        c, t = self._check(control), self._check(target)
        if c == t:
            _fail("DensityNoiseBackend::cnot", "control == target")
        n = self._n

        def app():
            self._sv._push(("cnot", c, t))
            self._sv._push(("cnot", c + n, t + n))
        self._fold(app, app, [c, t])

    def mcz(self, qubits):
        qs = [self._check(q) for q in qubits]
        if not qs:
            _fail("DensityNoiseBackend::mcz", "empty qubit list")
        if len(set(qs)) != len(qs):
            _fail("DensityNoiseBackend::mcz", "duplicate qubit")
        n = self._n

        def app():
            self._sv.mcz(qs)
            self._sv.mcz([q + n for q in qs])
        self._fold(app, app, qs)


def zne_exact(backend: DensityNoiseBackend, circuit_fn, observable_fn, scales=None) -> ZNEResult:
    """Exact ZNE on the density matrix (src/Mitigation.cpp:213-246): no statistical error."""
    who = "zne_exact"
    if scales is None:
        scales = ZNE.default_scales()
    if circuit_fn is None:
        _fail(who, "circuit_fn is empty")
    if observable_fn is None:
        _fail(who, "observable_fn is empty")
    ZNE.validate_scales(scales)
    saved = backend.scale
    try:
        out = ZNEResult()
        out.scales = [float(s) for s in scales]
        out.weights = ZNE.richardson_weights(scales)
        out.trajectories = 1
        for s in out.scales:
            scale = int(s)
            backend.set_scale(scale)
            backend.reset(_nq_of(backend))
            circuit_fn(backend)
            if backend.logical_gates == 0:
                raise RuntimeError(f"{who}: circuit_fn applied no gates to the backend")
            if backend.physical_gates != scale * backend.logical_gates:
                raise RuntimeError(f"{who}: folded gate count is inconsistent")
            v = float(observable_fn(backend))
            if not math.isfinite(v):
                _fail(who, "observable returned a non-finite value")
            out.values.append(v)
            out.std_errors.append(0.0)
        for ck, v in zip(out.weights, out.values):
            out.value += ck * v
        return out
    finally:
        backend.set_scale(saved)
