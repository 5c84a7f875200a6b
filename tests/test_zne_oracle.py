"""Zero-noise extrapolation: the oracle pinned, and the host side of the product (CPU only).

The oracle (oracle/zne.py) is pinned by the C++ standard's mt19937_64 known answer, by the closed
forms and hand values of the reference's own ZNE tests (src/Mitigation.h:463-700,
src/Mitigation.cpp:330-470: <Z>(lambda) = z0 r1^(3 lambda), <ZZ>(lambda) = r2^lambda, Richardson
weights (15/8, -5/4, 3/8) and (2.1875, -2.1875, 1.3125, -0.3125)), and, when oracle/_ref is built,
by the compiled reference's NoisyBackend / ZNE trajectory by trajectory.  Tolerances: 1e-12 on
deterministic values; sampled means within 5 standard errors (the reference's NSIG).
"""
import math

import numpy as np
import pytest

from oracle import zne as Z

P1 = 0.02
CIRC1 = [("ry", 0, 1.1), ("rz", 0, 0.7), ("rx", 0, 0.4)]
CIRC3 = [("h", 0), ("ry", 1, 0.7), ("cnot", 0, 2), ("rz", 2, 0.3), ("rx", 0, 0.9), ("ry", 2, -0.4),
         ("mcz", [0, 1]), ("mcz", [0, 1, 2]), ("x", 1), ("h", 2), ("rx", 1, 1.3), ("cnot", 1, 0),
         ("rz", 0, 0.8), ("mcz", [2]), ("ry", 0, 0.5)]
BELL = [("h", 0), ("cnot", 0, 1)]


def w_z(n):  # <Z_0>
    return np.array([1.0 - 2.0 * (i & 1) for i in range(2 ** n)])


def w_obs3():  # s0 s1 + 0.5 s2 (the reference's obs3)
    s = lambda i, b: 1.0 - 2.0 * ((i >> b) & 1)
    return np.array([s(i, 0) * s(i, 1) + 0.5 * s(i, 2) for i in range(8)])


def w_zz():
    return np.array([1.0, -1.0, -1.0, 1.0])


def ref_lib():
    import ctypes
    import os

    from oracle import ref

    path = os.path.join(ref.REF_DIR, "libref_mitigation.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    P, I, D, U = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint64
    lib.refm_values.argtypes = [I, I, P, P, P, P, D, U, I, I, P, P]
    lib.refm_extrapolate.argtypes = [I, I, P, P, P, P, D, U, I, I, P, P, P, P, P, P]
    lib.refm_richardson_weights.argtypes = [I, P, P]
    return lib


def encode(gates):
    kinds = {"h": 0, "x": 1, "rx": 2, "ry": 3, "rz": 4, "cnot": 5, "mcz": 6}
    G = len(gates)
    kind = np.zeros(G, dtype=np.int32)
    nq = np.zeros(G, dtype=np.int32)
    qs = np.zeros(16 * G, dtype=np.int32)
    th = np.zeros(G)
    for k, g in enumerate(gates):
        kind[k] = kinds[g[0]]
        q = Z.gate_qubits(g)
        nq[k] = len(q)
        qs[16 * k:16 * k + len(q)] = q
        if g[0] in ("rx", "ry", "rz"):
            th[k] = g[2]
    return kind, nq, qs, th


def ptr(a):
    import ctypes

    return a.ctypes.data_as(ctypes.c_void_p)


def ref_values(lib, n, gates, p, seed, scale, n_traj, w):
    kind, nq, qs, th = encode(gates)
    out = np.empty(n_traj)
    w = np.ascontiguousarray(w, dtype=float)
    rc = lib.refm_values(n, len(gates), ptr(kind), ptr(nq), ptr(qs), ptr(th), p, seed, scale, n_traj, ptr(w), ptr(out))
    assert rc == 0
    return out


def ref_extrapolate(lib, n, gates, p, seed, trajectories, scales, w):
    kind, nq, qs, th = encode(gates)
    sc = np.array(scales, dtype=float)
    w = np.ascontiguousarray(w, dtype=float)
    o3, v, se, c = np.empty(3), np.empty(len(sc)), np.empty(len(sc)), np.empty(len(sc))
    rc = lib.refm_extrapolate(n, len(gates), ptr(kind), ptr(nq), ptr(qs), ptr(th), p, seed, trajectories, len(sc),
                              ptr(sc), ptr(w), ptr(o3), ptr(v), ptr(se), ptr(c))
    assert rc == 0
    return dict(value=o3[0], std_error=o3[1], trajectories=int(o3[2]), values=list(v), std_errors=list(se),
                weights=list(c))


# ---- pins -------------------------------------------------------------------------------------

def test_mt19937_64_known_answer():
    g = Z.MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_richardson_hand_values_and_moments():
    assert np.allclose(Z.richardson_weights([1, 3, 5]), [15 / 8, -5 / 4, 3 / 8], atol=1e-13, rtol=0)
    w4 = Z.richardson_weights([1, 3, 5, 7])
    assert np.allclose(w4, [2.1875, -2.1875, 1.3125, -0.3125], atol=1e-13, rtol=0)
    assert np.allclose(Z.richardson_weights([1, 3]), [1.5, -0.5], atol=1e-15, rtol=0)
    s4 = [1.0, 3.0, 5.0, 7.0]
    assert abs(sum(w4) - 1.0) < 1e-12
    for m in (1, 2, 3):
        assert abs(sum(c * s ** m for c, s in zip(w4, s4))) < 1e-10
    a, b, q = 0.3, -0.07, 0.004
    c3 = Z.richardson_weights([1, 3, 5])
    assert abs(sum(ck * (a + b * s + q * s * s) for ck, s in zip(c3, [1, 3, 5])) - a) < 1e-12


def test_density_oracle_closed_forms():
    # one qubit, three rotations: <Z>(lambda) = z0 r1^(3 lambda)
    r1 = 1.0 - 4.0 * P1 / 3.0
    z0 = float(np.dot(w_z(1), np.real(np.diag(Z.density_run(1, CIRC1, 0.0, 1)))))
    for lam in (1, 3, 5, 7):
        e = float(np.dot(w_z(1), np.real(np.diag(Z.density_run(1, CIRC1, P1, lam)))))
        assert abs(e - z0 * r1 ** (3 * lam)) < 1e-12
    # Bell pair: <ZZ>(lambda) = r2^lambda
    p2 = 0.03
    r2 = 1.0 - 16.0 * p2 / 15.0
    for lam in (1, 3, 5):
        e = float(np.dot(w_zz(), np.real(np.diag(Z.density_run(2, BELL, p2, lam)))))
        assert abs(e - r2 ** lam) < 1e-12
    # the noiseless folded circuit is the circuit (scale 5 vs 1, all gate kinds)
    a, b = Z.density_run(3, CIRC3, 0.0, 5), Z.density_run(3, CIRC3, 0.0, 1)
    assert np.max(np.abs(a - b)) < 1e-12
    assert abs(np.trace(Z.density_run(3, CIRC3, 0.2, 3)) - 1.0) < 1e-12


def test_density_oracle_depolarizing_shrinks_paulis():
    # every non-identity Pauli expectation shrinks by r1 (1 qubit) / r2 (2 qubits) per site
    rng = np.random.default_rng(3)
    rho = np.zeros((4, 4), dtype=complex)
    v = rng.normal(size=4) + 1j * rng.normal(size=4)
    v /= np.linalg.norm(v)
    rho = np.outer(v, v.conj())
    p = 0.1
    out = Z._depolarize(rho, 2, [0, 1], p)
    r2 = 1.0 - 16.0 * p / 15.0
    for code in range(1, 16):
        P = np.kron(Z._PAULI[code >> 2], Z._PAULI[code & 3])
        assert abs(np.trace(P @ out) - r2 * np.trace(P @ rho)) < 1e-12
    out1 = Z._depolarize(rho, 2, [1], p)
    r1 = 1.0 - 4.0 * p / 3.0
    for k in (1, 2, 3):
        P = np.kron(Z._PAULI[k], Z._PAULI[0])  # qubit 1 is the high bit
        assert abs(np.trace(P @ out1) - r1 * np.trace(P @ rho)) < 1e-12


def test_noisy_oracle_matches_density_in_mean():
    p, N = 0.05, 3000
    nb = Z.NoisyOracle(3, p, 2025)
    for lam in (1, 3):
        vals = Z.noisy_values(3, CIRC3, p, 2025, lam, N, w_obs3(), oracle=nb)
        exact = float(np.dot(w_obs3(), np.real(np.diag(Z.density_run(3, CIRC3, p, lam)))))
        mean, se = Z.welford(vals)
        assert abs(mean - exact) < 5 * se


@pytest.mark.parametrize("p,seed", [(0.2, 7), (0.05, 2025)])
def test_noisy_oracle_vs_compiled_reference(p, seed):
    lib = ref_lib()
    if lib is None:
        pytest.skip("oracle/_ref not built")
    nb = Z.NoisyOracle(3, p, seed)
    for scale in (1, 3):
        ours = Z.noisy_values(3, CIRC3, p, seed, scale, 150, w_obs3(), oracle=nb)
        # the reference restarts its stream here; so does a fresh oracle
        theirs = ref_values(lib, 3, CIRC3, p, seed, scale, 150, w_obs3())
        fresh = Z.noisy_values(3, CIRC3, p, seed, scale, 150, w_obs3())
        assert np.max(np.abs(fresh - theirs)) < 1e-12
    r = ref_extrapolate(lib, 3, CIRC3, p, seed, 200, [1, 3, 5], w_obs3())
    o = Z.extrapolate(3, CIRC3, p, seed, 200, [1, 3, 5], w_obs3())
    assert abs(r["value"] - o["value"]) < 1e-12 and abs(r["std_error"] - o["std_error"]) < 1e-12
    assert np.max(np.abs(np.array(r["values"]) - o["values"])) < 1e-12
    assert np.max(np.abs(np.array(r["weights"]) - o["weights"])) < 1e-15


# ---- product host logic (no GPU) ---------------------------------------------------------------

class _Log:
    """A backend that records the gates it receives (the Paulis NoisyBackend inserts)."""

    def __init__(self, n):
        self.n, self.ops = n, []

    @property
    def n_qubits(self):
        return self.n

    def reset(self, n):
        self.ops = []

    def __getattr__(self, name):
        if name in ("h", "x", "rx", "ry", "rz", "cnot", "mcz"):
            return lambda *a: self.ops.append((name,) + tuple(tuple(x) if isinstance(x, list) else x for x in a))
        raise AttributeError(name)


def _apply(b, g):
    k = g[0]
    if k == "mcz":
        b.mcz(list(g[1]))
    elif k in ("rx", "ry", "rz", "cnot"):
        getattr(b, k)(g[1], g[2])
    else:
        getattr(b, k)(g[1])


@pytest.mark.parametrize("p,seed,scale", [(0.3, 11, 1), (0.15, 5, 3)])
def test_product_noise_stream_equals_oracle(p, seed, scale):
    """NoisyBackend's per-gate path (libllsv's std::mt19937_64 stream, no GPU) inserts exactly the
    oracle's Paulis at the oracle's sites."""
    import quantum_sim_mitigation as QM

    # gates that a Pauli (x(q) or mcz([q])) cannot be mistaken for in the log
    circ = [g for g in CIRC3 if g[0] != "x" and not (g[0] == "mcz" and len(g[1]) == 1)]
    log = _Log(3)
    nb = QM.NoisyBackend(log, p, seed)
    nb.set_scale(scale)
    orc = Z.NoisyOracle(3, p, seed)
    orc.set_scale(scale)
    for _ in range(30):
        nb.reset(3)
        orc.reset()
        for g in circ:
            _apply(nb, g)
            orc.gate(g)
        # rebuild the oracle's insertions from the log: gate applications are the non-Pauli ops
        assert nb.logical_gates == len(circ) and nb.physical_gates == scale * len(circ)
        paulis = []
        site = -1
        ops = log.ops
        expect = []
        for g in circ:
            expect += [g] * scale
        pos = 0
        for op in ops:
            if pos < len(expect) and _same(op, expect[pos]):
                site += 1
                pos += 1
            else:
                paulis.append((site, op))
        codes = {}
        for s, op in paulis:
            codes.setdefault(s, []).append(op)
        assert sorted(codes) == [s for s, _ in orc.events]
        for s, code in orc.events:
            qs = Z.gate_qubits(expect[s])
            want = []
            for j, q in enumerate(qs):
                w = (code >> (2 * j)) & 3
                if w in (2, 3):
                    want.append(("mcz", (q,)))
                if w in (1, 2):
                    want.append(("x", q))
            assert codes[s] == want


def _same(op, g):
    if op[0] != g[0]:
        return False
    if g[0] == "mcz":
        return op[1] == tuple(g[1])
    if g[0] in ("rx", "ry", "rz"):
        return op[1] == g[1] and abs(abs(op[2]) - abs(g[2])) < 1e-15
    if g[0] == "cnot":
        return op[1:] == (g[1], g[2])
    return op[1] == g[1]


def test_product_noise_stream_wide_mcz():
    """A 17-qubit mcz (one noise channel per qubit, drawn one qubit at a time by the product):
    the per-gate path inserts the oracle's Paulis at the oracle's sites, and the batched path is
    not taken for it."""
    import quantum_sim_mitigation as QM

    n, p, seed, scale = 17, 0.3, 9, 3
    wide = list(range(n))
    circ = [("h", 0), ("cnot", 0, 1), ("mcz", wide), ("rx", 3, 0.4), ("mcz", [2, 5, 16]), ("mcz", wide[::-1])]
    log = _Log(n)
    nb = QM.NoisyBackend(log, p, seed)
    nb.set_scale(scale)
    orc = Z.NoisyOracle(n, p, seed)
    orc.set_scale(scale)
    for _ in range(4):
        nb.reset(n)
        orc.reset()
        for g in circ:
            _apply(nb, g)
            orc.gate(g)
        expect = []
        for g in circ:
            expect += [g] * scale
        pos, site, paulis = 0, -1, []
        for op in log.ops:
            if pos < len(expect) and _same(op, expect[pos]):
                site += 1
                pos += 1
            else:
                paulis.append((site, op))
        assert pos == len(expect)
        codes = {}
        for s_, op in paulis:
            codes.setdefault(s_, []).append(op)
        assert sorted(codes) == [s_ for s_, _ in orc.events]
        for s_, code in orc.events:
            want = []
            for j, q in enumerate(Z.gate_qubits(expect[s_])):
                w = (code >> (2 * j)) & 3
                if w in (2, 3):
                    want.append(("mcz", (q,)))
                if w in (1, 2):
                    want.append(("x", q))
            assert codes[s_] == want
    assert any(len(Z.gate_qubits(expect[s_])) == n for s_, _ in orc.events)
    assert not QM.batchable(circ, 1) and QM.batchable(circ[:2] + circ[3:5], 99)
    assert not QM.batchable([("h", 0)] * (1 << 20), 5) and QM.batchable([("h", 0)] * (1 << 20), 3)


def test_product_host_validation_and_weights():
    import quantum_sim_mitigation as QM

    for bad in (-0.1, 1.2, float("nan")):
        with pytest.raises(ValueError):
            QM.NoisyBackend(_Log(1), bad)
    nb = QM.NoisyBackend(_Log(1), 0.02)
    for s in (2, 0, 101, 2.5):
        with pytest.raises(ValueError):
            nb.set_scale(s)
    for sc in ([1.0], [], [1.0, 3.0, 3.0], [1.0, 4.0], [1.0, 2.5], [-1.0, 1.0], [1.0, float("nan")], [1.0, 101.0],
               [1, 3, 5, 7, 9, 11, 13, 15, 17]):
        with pytest.raises(ValueError):
            QM.ZNE.validate_scales(sc)
    with pytest.raises(ValueError):
        QM.ZNE(0)
    # the folded circuit is bounded (2^22 gate applications), checked before any device work
    with pytest.raises(ValueError, match="gate applications"):
        QM.run_trajectories(nb, 1, [("h", 0)] * 42400, 99, 1, QM.Diagonal(np.ones(2)))
    with pytest.raises(ValueError):
        QM.ZNE.richardson([1.0, 3.0], [1.0])
    with pytest.raises(ValueError):
        QM.ZNE.richardson([1.0, 3.0], [1.0, float("nan")])
    with pytest.raises(ValueError):
        QM.ZNE.richardson_weights([2.0, 2.0])
    for sc in ([1, 3], [1, 3, 5], [1, 3, 5, 7], [1, 5, 9]):
        assert QM.ZNE.richardson_weights(sc) == Z.richardson_weights([float(s) for s in sc])
    # superoperator of the depolarizing channel == the oracle's channel on a random rho
    rng = np.random.default_rng(1)
    for m in (1, 2):
        d = 2 ** m
        a = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        rho = a @ a.conj().T
        rho /= np.trace(rho)
        S = QM._depolarizing_superop(m, 0.3)
        vec = rho.T.reshape(-1)  # index r + d c
        out = (S @ vec).reshape(d, d).T
        assert np.max(np.abs(out - Z._depolarize(rho, m, list(range(m)), 0.3))) < 1e-12
