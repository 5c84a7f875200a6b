"""Zero-noise extrapolation on the GPU (needs an MI355X) vs the oracle and the compiled reference.

Tolerances: per-trajectory observable values of the reference stream (same seed, same Pauli
insertions) 1e-12 absolute; ZNE values and standard errors 1e-12; the exact density-matrix path
1e-12 on rho; Philox-sampled means within 5 standard errors of the exact values (the reference's
NSIG).
"""
import math
import sys

import numpy as np
import pytest

try:  # torch first, so libllsv and torch share one HIP runtime
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

from oracle import zne as Z
from test_zne_oracle import BELL, CIRC1, CIRC3, P1, ref_extrapolate, ref_lib, ref_values, w_obs3, w_z, w_zz

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_circuit(n, G, seed, wide=True):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(G):
        k = rng.integers(0, 7 if n >= 2 else 5)
        q = int(rng.integers(0, n))
        if k <= 1:
            out.append((("h", "x")[k], q))
        elif k <= 4:
            out.append((("rx", "ry", "rz")[k - 2], q, float(rng.uniform(-3, 3))))
        elif k == 5:
            c, t = rng.choice(n, 2, replace=False)
            out.append(("cnot", int(c), int(t)))
        else:
            m = int(rng.integers(1, min(n, 4 if wide else 2) + 1))
            out.append(("mcz", [int(x) for x in rng.choice(n, m, replace=False)]))
    return out


def gpu_values(n, gates, p, seed, scale, n_traj, w, rng="reference", noisy=None):
    import quantum_sim_mitigation as QM

    nb = noisy or QM.NoisyBackend(QM.ReferenceBackend(n), p, seed, rng=rng)
    return QM.run_trajectories(nb, n, gates, scale, n_traj, QM.Diagonal(w))


@pytest.mark.parametrize("n,G,p,seed", [(1, 6, 0.3, 1), (2, 10, 0.2, 2), (3, 15, 0.2, 3), (5, 30, 0.1, 4),
                                        (8, 40, 0.05, 5), (10, 30, 0.05, 6), (12, 24, 0.04, 7),
                                        (13, 20, 0.04, 8), (15, 14, 0.03, 9), (17, 12, 0.2, 27),
                                        (20, 8, 0.2, 30)])
def test_reference_stream_matches_oracle(n, G, p, seed):
    # every trajectory: the oracle's Pauli insertions (same mt19937_64 stream), the same state
    gates = random_circuit(n, G, seed)
    w = np.random.default_rng(seed).uniform(-1, 1, 2 ** n)
    ntr = 6 if n >= 17 else 40 if n >= 10 else 120  # the oracle takes about 5 s at 20 qubits
    for scale in (1, 3):
        say(f"zne n={n} G={G} p={p} seed={seed} scale={scale}")
        ours = gpu_values(n, gates, p, seed, scale, ntr, w)
        theirs = Z.noisy_values(n, gates, p, seed, scale, ntr, w)
        assert np.max(np.abs(ours - theirs)) < 1e-12


def test_wide_mcz_between_tiled_runs():
    # 15 qubits (device-memory states): mcz gates on 14 qubits are wider than a 12-qubit tile, so they
    # split the tiled runs and take single passes, with a noise channel on each of their qubits
    n, p, seed = 15, 0.3, 21
    gates = ([("h", q) for q in range(n)] + [("mcz", list(range(14)))] + [("ry", 14, 0.3), ("cnot", 14, 2)]
             + [("mcz", list(range(1, 15)))] + [("rx", q, 0.1 * q) for q in range(6)])
    w = np.random.default_rng(seed).uniform(-1, 1, 2 ** n)
    for scale in (1, 3):
        say(f"zne wide mcz n={n} scale={scale}")
        ours = gpu_values(n, gates, p, seed, scale, 12, w)
        theirs = Z.noisy_values(n, gates, p, seed, scale, 12, w)
        assert np.max(np.abs(ours - theirs)) < 1e-12


def test_mcz_wider_than_a_code_word():
    # a 17-qubit mcz: one noise channel per qubit, drawn one qubit at a time; ZNE runs such a
    # circuit through the per-gate loop and still follows the reference stream
    import quantum_sim_mitigation as QM

    n, p, seed = 17, 0.2, 33
    gates = ([("h", q) for q in range(4)] + [("mcz", list(range(n)))] + [("cnot", 0, 16), ("rx", 16, 0.7)]
             + [("mcz", list(range(n))[::-1])] + [("ry", 2, 0.4)])
    assert not QM.batchable(gates, 1)
    w = np.random.default_rng(seed).uniform(-1, 1, 2 ** n)
    say("zne mcz on 17 qubits")
    nb = QM.NoisyBackend(QM.ReferenceBackend(n), p, seed)
    r = QM.ZNE(8).extrapolate(nb, lambda b: [_apply(b, g) for g in gates], QM.Diagonal(w), [1, 3])
    o = Z.extrapolate(n, gates, p, seed, 8, [1, 3], w)
    assert abs(r.value - o["value"]) < 1e-12 and abs(r.std_error - o["std_error"]) < 1e-12
    assert np.max(np.abs(np.array(r.values) - o["values"])) < 1e-12
    assert nb.logical_gates == len(gates) and nb.physical_gates == 3 * len(gates)


def test_reference_stream_matches_compiled_reference():
    lib = ref_lib()
    if lib is None:
        pytest.skip("oracle/_ref not built")
    for p, seed in ((0.2, 7), (0.02, 2025)):
        for scale in (1, 5):
            ours = gpu_values(3, CIRC3, p, seed, scale, 300, w_obs3())
            theirs = ref_values(lib, 3, CIRC3, p, seed, scale, 300, w_obs3())
            assert np.max(np.abs(ours - theirs)) < 1e-12


def test_extrapolate_matches_reference_driver():
    import quantum_sim_mitigation as QM

    lib = ref_lib()
    for p, seed, N in ((0.02, 2025, 2000), (0.1, 3, 500)):
        nb = QM.NoisyBackend(QM.ReferenceBackend(3), p, seed)
        r = QM.ZNE(N).extrapolate(nb, lambda b: [_apply(b, g) for g in CIRC3], QM.Diagonal(w_obs3()), [1, 3, 5])
        o = Z.extrapolate(3, CIRC3, p, seed, N, [1, 3, 5], w_obs3())
        assert abs(r.value - o["value"]) < 1e-12 and abs(r.std_error - o["std_error"]) < 1e-12
        assert np.max(np.abs(np.array(r.values) - o["values"])) < 1e-12
        assert r.trajectories == N and nb.scale == 1
        assert nb.logical_gates == len(CIRC3) and nb.physical_gates == 5 * len(CIRC3)
        if lib is not None:
            f = ref_extrapolate(lib, 3, CIRC3, p, seed, N, [1, 3, 5], w_obs3())
            assert abs(r.value - f["value"]) < 1e-12 and abs(r.std_error - f["std_error"]) < 1e-12
            assert np.max(np.abs(np.array(r.std_errors) - f["std_errors"])) < 1e-12


def _apply(b, g):
    k = g[0]
    if k == "mcz":
        b.mcz(list(g[1]))
    elif k in ("rx", "ry", "rz", "cnot"):
        getattr(b, k)(g[1], g[2])
    else:
        getattr(b, k)(g[1])


def test_callable_observable_and_per_gate_path_agree():
    # a Python observable on every final state == the in-kernel Diagonal; the per-gate decorator
    # (gates one by one through the GPU backend) == the batched run, trajectory by trajectory
    import quantum_sim_mitigation as QM

    n, p, seed = 4, 0.15, 9
    gates = random_circuit(n, 20, 9)
    w = np.random.default_rng(1).uniform(-1, 1, 2 ** n)
    circ = lambda b: [_apply(b, g) for g in gates]
    a = QM.ZNE(100).extrapolate(QM.NoisyBackend(QM.ReferenceBackend(n), p, seed), circ, QM.Diagonal(w), [1, 3])
    b = QM.ZNE(100).extrapolate(QM.NoisyBackend(QM.ReferenceBackend(n), p, seed), circ,
                                lambda be: float(np.dot(w, be.probabilities())), [1, 3])
    assert abs(a.value - b.value) < 1e-12 and abs(a.std_error - b.std_error) < 1e-12
    nb = QM.NoisyBackend(QM.ReferenceBackend(n), p, seed)
    nb.set_scale(3)
    seq = []
    for _ in range(25):
        nb.reset(n)
        circ(nb)
        seq.append(float(np.dot(w, nb.probabilities())))
    bat = gpu_values(n, gates, p, seed, 3, 25, w)
    assert np.max(np.abs(np.array(seq) - bat)) < 1e-12


def test_philox_closed_forms_and_exact():
    import quantum_sim_mitigation as QM

    # one qubit, three rotations: E(lambda) = z0 r1^(3 lambda) (src/Mitigation.h, test 2)
    r1 = 1.0 - 4.0 * P1 / 3.0
    z0 = float(np.dot(w_z(1), np.real(np.diag(Z.density_run(1, CIRC1, 0.0, 1)))))
    nb = QM.NoisyBackend(QM.ReferenceBackend(1), P1, 2024, rng="philox")
    r = QM.ZNE(60000).extrapolate(nb, lambda b: [_apply(b, g) for g in CIRC1], QM.Diagonal(w_z(1)), [1, 3, 5])
    for lam, v, se in zip(r.scales, r.values, r.std_errors):
        assert abs(v - z0 * r1 ** (3 * lam)) < 5 * se
    ana = sum(c * z0 * r1 ** (3 * lam) for c, lam in zip(r.weights, r.scales))
    assert abs(r.value - ana) < 5 * r.std_error
    assert abs(r.value - z0) < abs(r.values[0] - z0)
    # the same closed forms on wide registers (13 and 14 qubits, states in device-memory slots):
    # <Z_0> after CIRC1 on qubit 0, <Z_0 Z_(n-1)> of a Bell pair across the register
    r2w = 1.0 - 16.0 * 0.03 / 15.0
    for n in (13, 14):
        i = np.arange(2 ** n)
        z_0 = 1.0 - 2.0 * (i & 1)
        zz = z_0 * (1.0 - 2.0 * ((i >> (n - 1)) & 1))
        nbw = QM.NoisyBackend(QM.ReferenceBackend(n), P1, 31 + n, rng="philox")
        r = QM.ZNE(20000).extrapolate(nbw, lambda b: [_apply(b, g) for g in CIRC1], QM.Diagonal(z_0), [1, 3])
        for lam, v, se in zip(r.scales, r.values, r.std_errors):
            assert abs(v - z0 * r1 ** (3 * lam)) < 5 * se
        nbw = QM.NoisyBackend(QM.ReferenceBackend(n), 0.03, 47 + n, rng="philox")
        bell = [("h", 0), ("cnot", 0, n - 1)]
        r = QM.ZNE(20000).extrapolate(nbw, lambda b: [_apply(b, g) for g in bell], QM.Diagonal(zz), [1, 3])
        for lam, v, se in zip(r.scales, r.values, r.std_errors):
            assert abs(v - r2w ** lam) < 5 * se
    # Bell pair, two-qubit channel: <ZZ>(lambda) = r2^lambda
    p2 = 0.03
    r2 = 1.0 - 16.0 * p2 / 15.0
    nb2 = QM.NoisyBackend(QM.ReferenceBackend(2), p2, 99, rng="philox")
    r = QM.ZNE(40000).extrapolate(nb2, lambda b: [_apply(b, g) for g in BELL], QM.Diagonal(w_zz()), [1, 3])
    for lam, v, se in zip(r.scales, r.values, r.std_errors):
        assert abs(v - r2 ** lam) < 5 * se
    # 3 qubits, every gate kind: sampled vs exact density matrix
    p = 0.02
    nb3 = QM.NoisyBackend(QM.ReferenceBackend(3), p, 2025, rng="philox")
    mc = QM.ZNE(20000).extrapolate(nb3, lambda b: [_apply(b, g) for g in CIRC3], QM.Diagonal(w_obs3()), [1, 3, 5])
    ex = Z.zne_exact(3, CIRC3, p, [1, 3, 5], w_obs3())
    for v, se, e in zip(mc.values, mc.std_errors, ex["values"]):
        assert abs(v - e) < 5 * se
    assert abs(mc.value - ex["value"]) < 5 * mc.std_error


def test_density_backend_matches_oracle():
    import quantum_sim_mitigation as QM

    for n, gates, p in ((3, CIRC3, 0.05), (4, random_circuit(4, 25, 3), 0.1)):
        db = QM.DensityNoiseBackend(p, n)
        for scale in (1, 3):
            db.set_scale(scale)
            db.reset(n)
            for g in gates:
                _apply(db, g)
            rho = db.density()
            want = Z.density_run(n, gates, p, scale)
            assert np.max(np.abs(rho - want)) < 1e-12
            assert np.max(np.abs(db.probabilities() - np.real(np.diag(want)))) < 1e-12
    # zne_exact: closed forms (src/Mitigation.cpp tests 2 and 4)
    r1 = 1.0 - 4.0 * P1 / 3.0
    db = QM.DensityNoiseBackend(P1, 1)
    r = QM.zne_exact(db, lambda b: [_apply(b, g) for g in CIRC1], QM.Diagonal(w_z(1)), [1, 3, 5, 7])
    z0 = float(np.dot(w_z(1), np.real(np.diag(Z.density_run(1, CIRC1, 0.0, 1)))))
    for lam, v in zip(r.scales, r.values):
        assert abs(v - z0 * r1 ** (3 * lam)) < 1e-12
    with pytest.raises(RuntimeError):
        db.state()


def test_errors_and_size_limits():
    import quantum_sim_mitigation as QM

    nb = QM.NoisyBackend(QM.ReferenceBackend(2), 0.02, 1)
    with pytest.raises(RuntimeError):
        QM.ZNE(10).extrapolate(nb, lambda b: None, QM.Diagonal(w_zz()))
    with pytest.raises(ValueError):
        QM.ZNE(10).extrapolate(nb, lambda b: b.h(5), QM.Diagonal(w_zz()))
    with pytest.raises(ValueError):
        QM.ZNE(10).extrapolate(nb, lambda b: b.rx(0, float("nan")), QM.Diagonal(w_zz()))
    with pytest.raises(ValueError):
        QM.ZNE(10).extrapolate(nb, lambda b: b.h(0), lambda b: float("nan"))
    assert nb.scale == 1
    nb.set_scale(3)
    QM.ZNE(5).extrapolate(nb, lambda b: b.h(0), QM.Diagonal(w_zz()))
    assert nb.scale == 3
    with pytest.raises(ValueError):
        gpu_values(21, [("h", 0)], 0.1, 1, 1, 2, np.ones(2 ** 21))
    # 21 qubits: extrapolate runs trajectory by trajectory through the backend
    n = 21
    nb21 = QM.NoisyBackend(QM.ReferenceBackend(n), 0.2, 4)
    g21 = [("h", 0), ("cnot", 0, 20)]
    r = QM.ZNE(2).extrapolate(nb21, lambda b: [_apply(b, g) for g in g21],
                              lambda b: float(b.probabilities()[0]), [1, 3])
    o = Z.extrapolate(n, g21, 0.2, 4, 2, [1, 3], np.eye(1, 2 ** n, 0).ravel())
    assert abs(r.value - o["value"]) < 1e-12
    # 13 and 14 qubits: the batched path, one 1024-thread workgroup per trajectory and its state in a
    # device-memory slot, final states copied out for a callable observable
    for n in (13, 14):
        nbn = QM.NoisyBackend(QM.ReferenceBackend(n), 0.1, 4)
        gates = [("h", 0), ("cnot", 0, n - 1), ("ry", 5, 0.3)]
        r = QM.ZNE(6).extrapolate(nbn, lambda b: [_apply(b, g) for g in gates],
                                  lambda b: float(b.probabilities()[0]), [1, 3])
        o = Z.extrapolate(n, gates, 0.1, 4, 6, [1, 3], np.eye(1, 2 ** n, 0).ravel())
        assert abs(r.value - o["value"]) < 1e-12


def test_every_site_errs_and_one_qubit():
    # p = 1: a Pauli after every application (the reference stream draws a choice at every site)
    for n, gates, scale in ((1, [("x", 0), ("mcz", [0]), ("ry", 0, 0.4), ("h", 0)], 3),
                            (3, CIRC3, 1)):
        w = np.random.default_rng(n).uniform(-1, 1, 2 ** n)
        ours = gpu_values(n, gates, 1.0, 5, scale, 64, w)
        theirs = Z.noisy_values(n, gates, 1.0, 5, scale, 64, w)
        assert np.max(np.abs(ours - theirs)) < 1e-12
