"""Dense Lindblad trajectories: the oracle pinned (CPU only).

oracle/lindblad.py restates src/TrajectorySolver.h.  Pins: splitmix64's published known answer
(splitmix64(1234567) = 6457827717110365317, the first output of the reference generator for that
state) and libstdc++'s own uniform / normal draws on the solver's seeding rule
(tests/golden/lindblad_rng.json, written by tests/golden/make_lindblad_rng.cpp); scipy's expm; the
closed forms of the reference's solver tests (src/TrajectorySolver.h:589-690: T1 decay, T2 decay,
the dephasing variance ratio 1 + exp(4 gamma t)); ensemble rho vs the exact Liouvillian solution.
Sampled means within 4 standard errors (the reference's criterion).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import lindblad as LB

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lindblad_rng.json")

SM = np.array([[0, 1], [0, 0]], dtype=complex)  # sigma_minus = |0><1|
SZ = np.diag([1.0, -1.0]).astype(complex)
SX = np.array([[0, 1], [1, 0]], dtype=complex)


def test_splitmix_and_libstdcxx_draws():
    g = json.load(open(GOLD))
    for x, y in g["splitmix64"]:
        assert LB.splitmix64(int(x)) == int(y)
    assert LB.splitmix64(1234567) == 6457827717110365317
    for st in g["streams"]:
        r = LB.make_rng(st["seed"], st["index"])
        assert [r.uniform() for _ in range(40)] == st["uniform"]
        r = LB.make_rng(st["seed"], st["index"])
        assert [r.normal(0.1) for _ in range(40)] == st["normal"]


def test_expm_vs_scipy():
    from scipy.linalg import expm

    rng = np.random.default_rng(0)
    for d, scale in ((2, 0.1), (4, 3.0), (8, 20.0), (16, 0.7)):
        a = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) * scale
        ref = expm(a)
        assert np.max(np.abs(LB.expm(a) - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def mean_se(vals):
    v = np.asarray(vals)
    return float(v.mean()), float(v.std(ddof=1) / math.sqrt(len(v)))


@pytest.mark.parametrize("kind", [LB.STANDARD, LB.PROJECTOR, LB.ANALOG])
def test_t1_decay(kind):
    # <P1>(t) = exp(-t / T1) from |1>, L = sigma_-, gamma = 1 / T1 (src/TrajectorySolver.h:589-610)
    T1 = 2.0
    psi0 = np.array([0, 1], dtype=complex)
    for t in (1.0, 2.0):
        plan = LB.make_plan(np.zeros((2, 2)), [(SM, 1.0 / T1)], kind, t, 0.01 if kind == LB.ANALOG else 0.02)
        vals = [abs(LB.trajectory(psi0, plan, 7, i)[1]) ** 2 for i in range(400)]
        m, se = mean_se(vals)
        assert abs(m - math.exp(-t / T1)) < 4 * max(se, 1e-3)


def test_t2_decay_and_projector_variance():
    # pure dephasing L = Z at rate g: <X> = exp(-2 g t); Var_std / Var_proj = 1 + exp(4 g t)
    g, t = 0.5, 1.5  # g t = 0.75
    psi0 = np.array([1, 1], dtype=complex) / math.sqrt(2)
    m_exact = math.exp(-2 * g * t)
    res = {}
    for kind in (LB.STANDARD, LB.PROJECTOR):
        plan = LB.make_plan(np.zeros((2, 2)), [(SZ, g)], kind, t, 0.01)
        vals = [2 * float(np.real(np.conj(p[0]) * p[1])) for p in LB.ensemble_states(psi0, plan, 3, 600)]
        m, se = mean_se(vals)
        assert abs(m - m_exact) < 4 * se
        res[kind] = float(np.var(vals, ddof=1))
    assert abs(res[LB.STANDARD] / res[LB.PROJECTOR] - (1 + math.exp(4 * g * t))) < 0.25 * (1 + math.exp(4 * g * t))


def test_ensemble_rho_vs_liouvillian():
    # two qubits, H = X(x)X + 0.3 Z(x)1, dephasing on qubit 0 and decay on qubit 1
    H = np.kron(SX, SX) + 0.3 * np.kron(SZ, np.eye(2))
    chans = [(np.kron(np.eye(2), SZ), 0.2), (np.kron(SM, np.eye(2)), 0.4)]
    psi0 = np.array([1, 0, 0, 0], dtype=complex)
    t = 1.0
    exact = LB.lindblad_exact(H, chans, np.outer(psi0, psi0.conj()), t)
    plan = LB.make_plan(H, chans, LB.STANDARD, t, 0.01)
    S = LB.ensemble_states(psi0, plan, 11, 1500)
    rho = (S.T @ S.conj()) / len(S)
    assert np.max(np.abs(rho - exact)) < 0.05
    # the time-dependent (RK4) path with a constant H(t) agrees with the propagator path
    plan_t = LB.make_plan(lambda _t: H, chans, LB.STANDARD, t, 0.01)
    for i in range(5):
        a, b = LB.trajectory(psi0, plan, 5, i), LB.trajectory(psi0, plan_t, 5, i)
        assert np.max(np.abs(a - b)) < 1e-6


def test_pauli_like_detection():
    assert LB.pauli_like(SZ)[0] and LB.pauli_like(np.kron(SX, SX))[0]
    assert not LB.pauli_like(SM)[0] and not LB.pauli_like(np.kron(SM, np.eye(2)))[0]
    ok, p, c = LB.pauli_like(2.0 * SZ)
    assert ok and abs(c - 4.0) < 1e-15 and np.max(np.abs(p - SZ)) < 1e-15


def test_product_host_checks():
    import quantum_sim_trajectory as QT

    with pytest.raises(ValueError):
        QT.TrajectorySolver(0)
    with pytest.raises(ValueError):
        QT.TrajectorySolver(11)
    s = QT.TrajectorySolver(2)
    assert s.dim == 4 and s.n_qubits == 2 and s.MAX_QUBITS == 10
    with pytest.raises(ValueError):
        s.add_lindblad(SZ, 1.0)  # wrong size
    with pytest.raises(ValueError):
        s.add_lindblad(np.eye(4), -1.0)
    with pytest.raises(ValueError):
        s.set_hamiltonian(np.kron(SM, np.eye(2)))  # not Hermitian
    s.add_lindblad(np.kron(np.eye(2), SZ), 0.3)
    s.add_lindblad(np.kron(SM, np.eye(2)), 0.3)
    s.add_lindblad(np.eye(4), 0.0)  # gamma = 0 adds nothing
    assert s.n_channels == 2 and s.projector_applies(0) and not s.projector_applies(1)
    psi0 = np.array([1, 0, 0, 0], dtype=complex)
    with pytest.raises(ValueError):
        s.run_one(psi0, 1.0, 0.0, 1)
    with pytest.raises(ValueError):
        s.run_one(np.array([1.0, 0.0]), 1.0, 0.01, 1)
    with pytest.raises(ValueError):
        s.evolve_ensemble(psi0, 1.0, 0.01, 0)
    assert np.allclose(s.run_one(2 * psi0, 0.0, 0.01, 1), psi0)
