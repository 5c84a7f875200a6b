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
