"""The trajectory oracle pinned on the CPU: Philox4x32-10 known answers, closed forms, and the
ensemble against a brute-force density matrix (src/Noise.cpp:174-258 restated with np.kron)."""
import numpy as np
import pytest

from llsv import circuits as C
from oracle import oracle as O
from oracle import trajectory as T


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    assert T.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert T.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert T.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_rule():
    # u = (x >> 11) 2^-53, x = out0 | out1 << 32, counter (traj lo, traj hi, gate, site), key = seed
    seed, traj, gate, site = 0x1234_5678_9ABC_DEF0, (7 << 32) | 3, 11, 2
    o = T.philox4x32_10([3, 7, gate, site], [seed & 0xFFFFFFFF, seed >> 32])
    x = o[0] | (o[1] << 32)
    assert T.uniform(seed, traj, gate, site) == (x >> 11) * 2.0 ** -53
    u = np.array([T.uniform(5, t, 0, 0) for t in range(20000)])
    assert 0 <= u.min() and u.max() < 1 and abs(u.mean() - 0.5) < 0.01


def u(g):
    return ("u", g[1], O.gate4(g))


def test_amplitude_damping_closed_form():
    # X|0> = |1>, then AD(g): P(jump to |0>) = g, so E[Z] = 2 g - 1; PD leaves populations alone.
    z = np.array([T.run(1, [("u", 0, O.mat_x())], [0.3], [0.1], 1, i)[0][0] for i in range(20000)])
    assert abs(z.mean() - (2 * 0.3 - 1)) < 4 * z.std() / np.sqrt(z.size)
    assert set(np.round(z, 12)) <= {1.0, -1.0}


def noisy(n, layers, seed, lo, hi):
    rng = np.random.default_rng(seed)
    gates = []
    for layer in range(layers):
        gates += [u(("h", q)) for q in range(n)] + [u(("rz", q, float(rng.uniform(0, 6.28)))) for q in range(n)]
        gates += [("cnot", q, q + 1) for q in range(layer % 2, n - 1, 2)] + [("cnot", n - 1, 0)]
    gamma = rng.uniform(lo, hi, n)
    return gates, gamma, gamma / 2


def ensemble(n, gates, gamma, lam, count, seed=7):
    packed = T.pack(gates)
    return np.array([T.run(n, gates, gamma, lam, seed, i, packed)[0] for i in range(count)])


def test_ensemble_matches_density_matrix():
    n = 5
    gates, gamma, lam = noisy(n, 3, 1, 0.05, 0.2)
    zdm = T.density_matrix_z(n, gates, gamma, lam)
    z = ensemble(n, gates, gamma, lam, 6000)
    se = z.std(0) / np.sqrt(len(z)) + 1e-12
    assert np.all(np.abs(z.mean(0) - zdm) < 4.5 * se)


def test_planted_swapped_rates_are_caught():
    # the same statistics reject an ensemble run with gamma and lambda exchanged
    n = 4
    gates, gamma, lam = noisy(n, 3, 2, 0.3, 0.5)
    zdm = T.density_matrix_z(n, gates, gamma, lam)
    z = ensemble(n, gates, lam, gamma, 6000)
    se = z.std(0) / np.sqrt(len(z)) + 1e-12
    assert np.any(np.abs(z.mean(0) - zdm) > 6 * se)


def test_noisy_brick_builder():
    gates, gamma, lam = C.noisy_brick(3, 14, 8)
    assert len(gates) == 8 * (14 + 14) + 4 * 7 + 4 * 6
    assert np.all((1e-3 <= gamma) & (gamma <= 1e-2)) and np.allclose(lam, gamma / 2)


# ------------------------------------------ the three density-matrix restatements pinned
def brick_both_ways(n, seed, lo, hi):
    """noisy() plus CNOTs with the control below and above the target, far and near."""
    gates, gamma, lam = noisy(n, 2, seed, lo, hi)
    gates += [("cnot", 0, n - 1), ("cnot", n - 1, 1), ("cnot", 2, 0), u(("ry", 1, 0.7)), ("cnot", 1, 3)]
    return gates, gamma, lam


@pytest.mark.parametrize("n", [4, 5, 6])
def test_density_matrix_restatements_agree(n):
    gates, gamma, lam = brick_both_ways(n, 10 + n, 0.05, 0.3)
    rho_k = T.density_matrix_kron(n, gates, gamma, lam)
    rho_b = T.density_matrix_blocks(n, gates, gamma, lam)
    assert np.max(np.abs(rho_b - rho_k)) < 1e-14
    # one sweep per channel sequence is the same arithmetic as one sweep per channel
    assert np.array_equal(T.density_matrix_blocks(n, gates, gamma, lam, merged=False), rho_b)
    zk = T.density_matrix_z(n, gates, gamma, lam)
    assert np.max(np.abs(T.density_matrix_z_fast(n, gates, gamma, lam) - zk)) < 1e-14
    zb, tr = T.dm_expect_z(n, rho_b)
    assert np.max(np.abs(zb - zk)) < 1e-14 and abs(tr - 1) < 1e-13
    # a valid density matrix: Hermitian, unit trace, positive semidefinite
    assert np.max(np.abs(rho_b - rho_b.conj().T)) < 1e-15
    assert np.linalg.eigvalsh(rho_b).min() > -1e-13


def test_density_matrix_closed_forms():
    g, l = 0.3, 0.2
    X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2)
    ad, pd = T._kraus_sets([g], [l], 0)
    # AD maps |1><1| to g|0><0| + (1 - g)|1><1| (src/Noise.h:81-84): E[Z] = 2g - 1
    rho = np.zeros((2, 2), dtype=np.complex128)
    rho[0, 0] = 1
    T.dm_apply_channel(rho, 1, [0], [X])
    T.dm_apply_channel(rho, 1, [0], ad)
    assert np.allclose(rho, np.diag([g, 1 - g]), atol=1e-15, rtol=0)
    assert abs(T.dm_expect_z(1, rho)[0][0] - (2 * g - 1)) < 1e-15
    # on |+><+|: AD scales the coherence by sqrt(1-g) and moves g/2 of population to |0>;
    # PD leaves populations alone and scales the coherence by sqrt(1-l)
    rho = np.full((2, 2), 0.5, dtype=np.complex128)
    T.dm_apply_channel(rho, 1, [0], ad)
    assert np.allclose(rho, [[0.5 + g / 2, 0.5 * np.sqrt(1 - g)], [0.5 * np.sqrt(1 - g), 0.5 * (1 - g)]],
                       atol=1e-15, rtol=0)
    rho = np.full((2, 2), 0.5, dtype=np.complex128)
    T.dm_apply_channel(rho, 1, [0], pd)
    assert np.allclose(rho, [[0.5, 0.5 * np.sqrt(1 - l)], [0.5 * np.sqrt(1 - l), 0.5]], atol=1e-15, rtol=0)
    # H then CNOT, noiseless, either orientation: the Bell state |00> + |11>
    for c, t in ((0, 1), (1, 0)):
        rho = T.density_matrix_blocks(2, [("u", c, O.gate4(("h", c))), ("cnot", c, t)], [0, 0], [0, 0])
        bell = np.zeros(4, dtype=np.complex128)
        bell[[0, 3]] = 1 / np.sqrt(2)
        assert np.allclose(rho, np.outer(bell, bell.conj()), atol=1e-15, rtol=0)
    # the same AD closed form through the tensor restatement and the circuit drivers
    z = T.density_matrix_z_fast(1, [("u", 0, O.mat_x())], [g], [l])
    assert abs(z[0] - (2 * g - 1)) < 1e-15
    assert abs(T.density_matrix_z_blocks(1, [("u", 0, O.mat_x())], [g], [l])[0] - (2 * g - 1)) < 1e-15


def test_planted_density_matrix_mistakes_are_caught(monkeypatch):
    n = 5
    gates, gamma, lam = brick_both_ways(n, 3, 0.05, 0.3)
    zk = T.density_matrix_z(n, gates, gamma, lam)
    # CNOT flipped on the ket twice and never on the bra
    with monkeypatch.context() as m:
        m.setattr(T, "_qubit_axis", lambda n_, q, side: n_ - 1 - q)
        assert np.max(np.abs(T.density_matrix_z_fast(n, gates, gamma, lam) - zk)) > 1e-3
    # the target axis not shifted after the control axis is indexed away
    # (an out-of-range axis raises, a wrong in-range one must change the result)
    with monkeypatch.context() as m:
        m.setattr(T, "_axis_after_index", lambda axis, indexed: axis)
        try:
            bad = T.density_matrix_z_fast(n, gates, gamma, lam)
        except (ValueError, IndexError, np.exceptions.AxisError):
            bad = None
        assert bad is None or np.max(np.abs(bad - zk)) > 1e-3
    # control and target exchanged in the block update
    rho = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    rho[0, 0] = 1
    for g in gates:
        if g[0] == "cnot":
            T.dm_apply_channel(rho, n, [g[2], g[1]], [T._CNOT4])
            qs = [g[1], g[2]]
        else:
            T.dm_apply_channel(rho, n, [g[1]], [np.asarray(g[2], dtype=np.complex128).reshape(2, 2)])
            qs = [g[1]]
        for q in qs:
            for ks in T._kraus_sets(gamma, lam, q):
                T.dm_apply_channel(rho, n, [q], ks)
    assert np.max(np.abs(T.dm_expect_z(n, rho)[0] - zk)) > 1e-3
    # the Kraus operator's adjoint dropped (K B K instead of K B K^dag) on AD
    rho = np.zeros((2, 2), dtype=np.complex128)
    rho[1, 1] = 1
    K0, K1 = T._kraus_sets([0.3], [0.1], 0)[0]
    rho_bad = sum(K @ rho @ K for K in (K0, K1))
    assert abs(np.real(rho_bad[0, 0]) - 0.3) > 0.1



def test_cpu_trajectories_match_density_matrix_fixture():
    # two independent restatements agree on the committed fixture's 10-qubit case: 4096 CPU
    # trajectories (jump unraveling) within 3 standard errors of the block-update density matrix,
    # and that density matrix recomputed here equals the fixture
    import json
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dm_brick14.json")
    with open(path) as f:
        fx = json.load(f)["cases"][0]
    n = fx["n_qubits"]
    gates, gamma, lam = C.noisy_brick(fx["circuit_seed"], n, fx["layers"])
    gu = [g if g[0] == "cnot" else u(g) for g in gates]
    assert np.array_equal(gamma, np.array(fx["gamma"]))
    assert np.max(np.abs(T.density_matrix_z_blocks(n, gu, gamma, lam) - np.array(fx["z"]))) < 1e-13
    z = ensemble(n, gu, gamma, lam, 4096, seed=1)
    se = z.std(0, ddof=1) / np.sqrt(len(z))
    assert np.all(se > 0)
    assert np.all(np.abs(z.mean(0) - np.array(fx["z"])) < 3 * se)
