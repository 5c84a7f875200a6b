"""Dense Lindblad trajectories on the GPU (needs an MI355X) vs the oracle.

With the reference's random streams (the default) every trajectory takes the oracle's jump
decisions, so final states agree to rounding: 1e-9 absolute after hundreds of steps (the GPU's
expm and GEMMs sum in another order, ~1e-15 per step).  expm: 1e-12 relative to scipy.  Philox
draws: closed forms within 4 standard errors (the reference's criterion).
"""
import math
import sys

import numpy as np
import pytest

try:  # torch first, so libllsv and torch share one HIP runtime
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

from oracle import lindblad as LB

pytestmark = pytest.mark.gpu

SM = np.array([[0, 1], [0, 0]], dtype=complex)
SZ = np.diag([1.0, -1.0]).astype(complex)
SX = np.array([[0, 1], [1, 0]], dtype=complex)
SY = np.array([[0, -1j], [1j, 0]])


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("d,scale", [(2, 0.3), (4, 5.0), (64, 0.05), (128, 0.02), (1024, 0.003)])
def test_expm_vs_scipy(d, scale):
    from scipy.linalg import expm

    import quantum_sim_trajectory as QT

    rng = np.random.default_rng(d)
    a = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) * scale
    say(f"expm d={d}")
    got = QT.expm(a)
    ref = expm(a)
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(got - LB.expm(a))) < 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (128, 192, 48), (256, 64, 1024)])
def test_zgemm_vs_numpy(M, N, K):
    import quantum_sim_trajectory as QT

    rng = np.random.default_rng(M + N + K)
    a = rng.normal(size=(M, K)) + 1j * rng.normal(size=(M, K))
    b = rng.normal(size=(K, N)) + 1j * rng.normal(size=(K, N))
    ref = a @ b
    assert np.max(np.abs(QT.zgemm(a, b) - ref)) < 1e-12 * K
    with pytest.raises(ValueError):
        QT.zgemm(a[:, :8], b[:8])


def two_qubit_model():
    H = np.kron(SX, SX) + 0.3 * np.kron(SZ, np.eye(2)) + 0.2 * np.kron(np.eye(2), SY)
    chans = [(np.kron(np.eye(2), SZ), 0.2), (np.kron(SM, np.eye(2)), 0.4), (np.kron(SX, SX), 0.1)]
    return H, chans


def solver(n, H, chans):
    import quantum_sim_trajectory as QT

    s = QT.TrajectorySolver(n)
    if H is not None:
        s.set_hamiltonian(H)
    for l, g in chans:
        s.add_lindblad(l, g)
    return s


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("timedep", [False, True])
def test_trajectories_match_oracle(kind, timedep):
    import quantum_sim_trajectory as QT

    H, chans = two_qubit_model()
    Hf = (lambda t: H * math.cos(0.7 * t)) if timedep else H
    s = solver(2, Hf, chans)
    psi0 = np.array([0.6, 0.0, 0.8j, 0.0])
    t, dt, N, seed = 1.3, 0.01, 24, 2025
    say(f"lindblad kind={kind} timedep={timedep}")
    plan = LB.make_plan(Hf, chans, kind, t, dt)
    want = LB.ensemble_states(psi0, plan, seed, N)
    u = QT.Unraveling(kind)
    rho = s.evolve_ensemble(psi0, t, dt, N, u, seed)
    assert np.max(np.abs(rho - (want.T @ want.conj()) / N)) < 1e-9
    one = s.run_one(psi0, t, dt, seed, u)
    assert np.max(np.abs(one - want[0])) < 1e-9
    st = s.estimate_observable(psi0, t, dt, N, lambda p: float(abs(p[0]) ** 2), u, seed)
    vals = [abs(p[0]) ** 2 for p in want]
    assert abs(st.mean - np.mean(vals)) < 1e-9 and st.n == N


@pytest.mark.parametrize("n,steps", [(7, 12), (10, 6)])
def test_large_dims_match_oracle(n, steps):
    rng = np.random.default_rng(n)
    d = 1 << n
    a = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    H = (a + a.conj().T) / (2 * math.sqrt(d))
    l1 = np.kron(np.eye(d // 2), SM)  # decay on qubit 0
    l2 = np.kron(SZ, np.eye(d // 2))  # dephasing on the top qubit
    chans = [(l1, 0.5), (l2, 0.3)]
    s = solver(n, H, chans)
    psi0 = rng.normal(size=d) + 1j * rng.normal(size=d)
    t = 0.05 * steps
    say(f"lindblad n={n}")
    plan = LB.make_plan(H, chans, 0, t, 0.05)
    want = LB.ensemble_states(psi0, plan, 9, 4)
    rho = s.evolve_ensemble(psi0, t, 0.05, 4, 0, 9)
    assert np.max(np.abs(rho - (want.T @ want.conj()) / 4)) < 1e-9


def test_philox_closed_forms():
    import quantum_sim_trajectory as QT

    # T1 decay: <P1>(t) = exp(-t / T1)
    s = solver(1, None, [(SM, 0.5)])
    s.rng = "philox"
    for kind in (QT.Unraveling.STANDARD, QT.Unraveling.ANALOG):
        st = s.estimate_observable(np.array([0, 1.0]), 2.0, 0.01, 4000, lambda p: float(abs(p[1]) ** 2), kind, 5)
        assert abs(st.mean - math.exp(-1.0)) < 4 * max(st.std_error, 1e-3)
    # dephasing: Var_std / Var_proj = 1 + exp(4 g t), unbiased <X> = exp(-2 g t)
    g, t = 0.5, 1.5
    s = solver(1, None, [(SZ, g)])
    s.rng = "philox"
    plus = np.array([1, 1]) / math.sqrt(2)
    X = SX
    res = {}
    for kind in (QT.Unraveling.STANDARD, QT.Unraveling.PROJECTOR):
        st = s.estimate_expectation(plus, t, 0.01, 20000, X, kind, 3)
        assert abs(st.mean - math.exp(-2 * g * t)) < 4 * st.std_error
        res[kind] = st.std_error ** 2 * st.n
    ratio = res[QT.Unraveling.STANDARD] / res[QT.Unraveling.PROJECTOR]
    assert abs(ratio - (1 + math.exp(4 * g * t))) < 0.1 * (1 + math.exp(4 * g * t))


def test_reference_stream_ensemble_vs_exact():
    import quantum_sim_trajectory as QT

    H, chans = two_qubit_model()
    s = solver(2, H, chans)
    psi0 = np.array([1.0, 0, 0, 0])
    rho = s.evolve_ensemble(psi0, 1.0, 0.01, 20000, QT.Unraveling.PROJECTOR, 1)
    exact = LB.lindblad_exact(H, chans, np.outer(psi0, psi0), 1.0)
    assert np.max(np.abs(rho - exact)) < 0.03
    assert abs(np.trace(rho) - 1.0) < 1e-12


def test_norm_vanishes_raises():
    s = solver(1, None, [(SM, 1e6)])
    with pytest.raises(RuntimeError):
        s.run_one(np.array([0, 1.0]), 1.0, 1.0, 1)


def test_unitary_evolution_closed_form():
    # no channels: every unraveling is the exact propagator exp(-i H t) (no draws)
    from scipy.linalg import expm

    import quantum_sim_trajectory as QT

    H, _ = two_qubit_model()
    s = solver(2, H, [])
    psi0 = np.array([0.6, 0.0, 0.8j, 0.0])
    want = expm(-1j * H * 1.3) @ psi0
    for kind in (QT.Unraveling.STANDARD, QT.Unraveling.ANALOG):
        got = s.run_one(psi0, 1.3, 0.01, 7, kind)
        assert np.max(np.abs(got - want)) < 1e-12
    with pytest.raises(ValueError):
        s.estimate_expectation(psi0, 1.0, 0.01, 4, np.kron(SM, np.eye(2)))


def _run_plan(lib, plan, psi0, n, seed):
    import ctypes

    out = np.empty((n, psi0.size), dtype=np.complex128)
    psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
    rc = lib.llsv_lplan_run(plan, psi0.ctypes.data_as(ctypes.c_void_p), n, 0, seed, 0,
                            out.ctypes.data_as(ctypes.c_void_p), None)
    lib.llsv_lplan_destroy(plan)
    assert rc == 0
    return out


@pytest.mark.parametrize("kind,timedep,n", [(0, False, 2), (1, False, 2), (2, False, 2), (0, True, 2),
                                            (2, True, 2), (0, False, 7)])
def test_device_plan_matches_host_plan(kind, timedep, n):
    """llsv_lplan_create_h (H_eff, expm and nodes on the device) against llsv_lplan_create given
    the oracle's host-built nodes: the same trajectories to rounding."""
    import ctypes

    from llsv import _native

    lib = _native.lib()
    rng = np.random.default_rng(40 + n)
    d = 1 << n
    if n == 2:
        H, chans = two_qubit_model()
    else:
        a = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        H = (a + a.conj().T) / (2 * math.sqrt(d))
        chans = [(np.kron(np.eye(d // 2), SM), 0.5), (np.kron(SZ, np.eye(d // 2)), 0.3)]
    Hf = (lambda t: H * math.cos(0.7 * t)) if timedep else H
    t, dt = 0.4, 0.02
    plan = LB.make_plan(Hf, chans, kind, t, dt)
    jump = kind != LB.ANALOG
    nodes = [plan["prop"]] if plan["constant"] else plan["nodes"]
    nodes = np.ascontiguousarray(np.stack(nodes))
    ops = np.ascontiguousarray(np.stack(plan["op"]))
    if timedep:
        hs = np.stack([Hf(0.5 * plan["h"] * j if jump else plan["h"] * j) for j in range(len(nodes))])
    else:
        hs = H
    hs = np.ascontiguousarray(hs, dtype=np.complex128)
    say(f"device plan kind={kind} timedep={timedep} n={n}")
    ph, pd = ctypes.c_void_p(), ctypes.c_void_p()
    k = 0 if jump else 1
    _native.check(lib.llsv_lplan_create(d, k, plan["steps"], plan["h"], len(nodes),
                                        nodes.ctypes.data_as(ctypes.c_void_p), len(ops),
                                        ops.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ph)))
    _native.check(lib.llsv_lplan_create_h(d, k, plan["steps"], plan["h"], len(nodes),
                                          hs.ctypes.data_as(ctypes.c_void_p), len(ops),
                                          ops.ctypes.data_as(ctypes.c_void_p), ctypes.byref(pd)))
    psi0 = rng.normal(size=d) + 1j * rng.normal(size=d)
    psi0 /= np.linalg.norm(psi0)
    a = _run_plan(lib, ph, psi0, 16, 77)
    b = _run_plan(lib, pd, psi0, 16, 77)
    assert np.max(np.abs(a - b)) < 1e-12
    # argument checks
    bad = ctypes.c_void_p()
    assert lib.llsv_lplan_create_h(d, k, plan["steps"], plan["h"], 3, hs.ctypes.data_as(ctypes.c_void_p),
                                   len(ops), ops.ctypes.data_as(ctypes.c_void_p), ctypes.byref(bad)) != 0
    assert not bad.value


def test_plan_reuse_and_invalidation():
    """A constant-H plan is reused by the next call on the same grid (same rho bit for bit) and
    dropped by every model setter; the model keeps copies of H and the operators."""
    import quantum_sim_trajectory as QT

    H, chans = two_qubit_model()
    Hu = H.copy()
    s = solver(2, Hu, chans)
    psi0 = np.array([0.6, 0.0, 0.8j, 0.0])
    say("plan reuse")
    a = s.evolve_ensemble(psi0, 0.5, 0.01, 32, QT.Unraveling.STANDARD, 3)
    assert s._plan_cache is not None
    b = s.evolve_ensemble(psi0, 0.5, 0.01, 32, QT.Unraveling.STANDARD, 3)
    assert np.array_equal(a, b)
    Hu *= 3.0  # the caller's array: not the model's copy
    assert np.array_equal(a, s.evolve_ensemble(psi0, 0.5, 0.01, 32, QT.Unraveling.STANDARD, 3))
    plan = LB.make_plan(H, chans, LB.PROJECTOR, 0.5, 0.01)
    want = LB.ensemble_states(psi0, plan, 3, 32)
    got = s.evolve_ensemble(psi0, 0.5, 0.01, 32, QT.Unraveling.PROJECTOR, 3)  # another unraveling
    assert np.max(np.abs(got - (want.T @ want.conj()) / 32)) < 1e-9
    s.set_hamiltonian(2.0 * H)
    assert s._plan_cache is None
    plan = LB.make_plan(2.0 * H, chans, LB.STANDARD, 0.5, 0.01)
    want = LB.ensemble_states(psi0, plan, 3, 32)
    got = s.evolve_ensemble(psi0, 0.5, 0.01, 32, QT.Unraveling.STANDARD, 3)
    assert np.max(np.abs(got - (want.T @ want.conj()) / 32)) < 1e-9
    s.add_lindblad(np.kron(SX, np.eye(2)), 0.2)
    assert s._plan_cache is None
    plan = LB.make_plan(2.0 * H, chans + [(np.kron(SX, np.eye(2)), 0.2)], LB.STANDARD, 0.5, 0.01)
    want = LB.ensemble_states(psi0, plan, 3, 32)
    got = s.evolve_ensemble(psi0, 0.5, 0.01, 32, QT.Unraveling.STANDARD, 3)
    assert np.max(np.abs(got - (want.T @ want.conj()) / 32)) < 1e-9
    s.clear_lindblad()
    assert s._plan_cache is None


def test_non_finite_inputs_rejected():
    import ctypes

    import quantum_sim_trajectory as QT
    from llsv import _native

    lib = _native.lib()
    a = np.eye(4, dtype=complex)
    a[1, 2] = np.nan
    with pytest.raises(Exception):
        QT.expm(a)
    d = 64
    H = np.eye(d, dtype=complex)
    H[3, 5] = np.nan
    ops = np.ascontiguousarray(np.eye(d, dtype=complex)[None])
    p = ctypes.c_void_p()
    rc = lib.llsv_lplan_create_h(d, 0, 4, 0.01, 1, H.ctypes.data_as(ctypes.c_void_p), 1,
                                 ops.ctypes.data_as(ctypes.c_void_p), ctypes.byref(p))
    assert rc != 0 and not p.value
