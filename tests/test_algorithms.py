"""Grover, QAOA and VQE drivers (quantum_sim_algorithms) against values the reference produced
(tests/golden/algorithms_reference.json, from its docs/*.json) and closed forms, on the CPU:
the drivers run here on a backend built from the oracle (test infrastructure), so this pins
the driver logic; tests/test_gpu_algorithms.py runs the same checks on the GPU backend.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as O

FIX = json.load(open(os.path.join(GOLDEN, "algorithms_reference.json")))


class OracleBackend:
    """The ll::Backend interface over the CPU oracle (checker only)."""

    def __init__(self, n=1):
        self.reset(n)

    def reset(self, n):
        self.n = n
        self.psi = O.zero_state(n)

    def n_qubits(self):
        return self.n

    def h(self, q):
        O.apply_1q(self.psi, q, O.mat_h())

    def x(self, q):
        O.apply_1q(self.psi, q, O.mat_x())

    def rx(self, q, t):
        O.apply_1q(self.psi, q, O.mat_rx(t))

    def ry(self, q, t):
        O.apply_1q(self.psi, q, O.mat_ry(t))

    def rz(self, q, t):
        O.apply_1q(self.psi, q, O.mat_rz(t))

    def cnot(self, c, t):
        O.apply_cnot(self.psi, c, t)

    def mcz(self, qs):
        O.apply_mcz(self.psi, list(qs))

    def probabilities(self):
        return O.probabilities(self.psi)

    def state(self):
        return self.psi.copy()


def alg():
    import quantum_sim_algorithms

    return quantum_sim_algorithms


def h2_energy_fn(vqe, backend):
    g0, g1, g2, g3, g4 = FIX["vqe_h2"]["coefficients"]
    idx = np.arange(4)
    s0, s1 = 1.0 - 2.0 * (idx & 1), 1.0 - 2.0 * ((idx >> 1) & 1)
    diag = g0 + g1 * s0 + g2 * s1 + g3 * s0 * s1

    def energy(params):
        vqe.prepare(backend, params)
        psi = np.asarray(backend.state())
        p = psi.real ** 2 + psi.imag ** 2
        xx = 2.0 * float(np.real(np.conj(psi[0]) * psi[3] + np.conj(psi[1]) * psi[2]))
        return float(np.dot(p, diag)) + g4 * xx

    return energy, diag


def test_mt19937_64_known_answer():
    # C++ [rand.predef]: the 10000th output of a default-constructed (seed 5489) mt19937_64
    rng = alg()._MT19937_64(5489)
    for _ in range(9999):
        rng()
    assert rng() == 9981545732273789042


def test_h2_hamiltonian_reproduces_published_fci():
    _, diag = h2_energy_fn(None, None)
    g4 = FIX["vqe_h2"]["coefficients"][4]
    m = np.diag(diag)
    for z in range(4):
        m[z, z ^ 3] += g4
    assert abs(np.linalg.eigvalsh(m)[0] - FIX["vqe_h2"]["e_fci_electronic_published"]) < 1e-9


@pytest.mark.parametrize("row", FIX["grover_auto"], ids=lambda r: f"n{r['n']}")
def test_grover_automatic_iterations(row):
    A = alg()
    r = A.Grover.run(OracleBackend(), row["n"], [row["marked"]])
    assert r.iterations == row["iterations"] and r.marked_index == row["found"]
    theta = math.asin(math.sqrt(1.0 / (1 << row["n"])))
    closed = math.sin((2 * r.iterations + 1) * theta) ** 2
    assert abs(r.probabilities[row["marked"]] - row["peak"]) < 1e-10
    assert abs(r.probabilities[row["marked"]] - closed) < 1e-10


@pytest.mark.parametrize("row", [r for r in FIX["grover_10_iterations"]["rows"] if r["n"] <= 14],
                         ids=lambda r: f"n{r['n']}")
def test_grover_ten_iterations(row):
    A = alg()
    marked = sum(1 << q for q in range(0, row["n"], 2))
    r = A.Grover.run(OracleBackend(), row["n"], [marked], FIX["grover_10_iterations"]["iterations"])
    assert r.marked_index == row["argmax"] == marked
    assert abs(r.probabilities[marked] - row["p_success"]) < 1e-12


def qaoa_graph(A):
    q = FIX["qaoa"]
    return A.Graph(q["n_nodes"], [tuple(e) for e in q["edges"]])


@pytest.mark.parametrize("layer", FIX["qaoa"]["layers"], ids=lambda l: f"p{l['p']}")
def test_qaoa_expected_cut(layer):
    A = alg()
    g = qaoa_graph(A)
    r = A.QAOA.run(OracleBackend(), g, layer["gammas"], layer["betas"])
    cuts = np.array([A.QAOA.cut_value(g, z) for z in range(1 << g.n_nodes)])
    assert abs(float(np.dot(r.probabilities, cuts)) - layer["expected_cut"]) < 1e-10
    assert abs(r.best_cut - layer["best_cut"]) < 1e-12
    opt = FIX["qaoa"]["optimum"]
    assert abs(float(r.probabilities[np.abs(cuts - opt) < 1e-9].sum()) - layer["p_optimum"]) < 1e-10
    if layer["p"] == FIX["qaoa"]["best_distribution_p"]:
        assert np.max(np.abs(r.probabilities - np.asarray(FIX["qaoa"]["best_distribution"]))) < 1e-12


@pytest.mark.parametrize("seed", ["1", "2", "3"])
def test_vqe_h2_energy_history(seed):
    # best energy after k gradient steps, k = 0..50, lr 0.5 (the reference's recorded curves)
    A = alg()
    hist = FIX["vqe_h2"]["energy_history"][seed]
    for k in (0, 1, 2, 5, 10, 25, 50):
        vqe = A.VQE(2, FIX["vqe_h2"]["layers"], int(seed))
        be = OracleBackend(2)
        energy, _ = h2_energy_fn(vqe, be)
        res = vqe.optimize(be, energy, k, FIX["vqe_h2"]["lr"])
        assert abs(res.energy - hist[k]) < 1e-10, f"seed {seed}, k {k}"
        assert res.iterations == k


def test_vqe_prepare_and_argument_errors():
    A = alg()
    vqe = A.VQE(3, 2, 5)
    assert (vqe.n_qubits, vqe.n_layers, vqe.n_params) == (3, 2, 6)
    assert np.all(np.abs(vqe.initial_params()) <= 0.1)
    be = OracleBackend(3)
    vqe.prepare(be, np.zeros(6))  # RY(0) layers and CNOT rings leave |000>
    assert abs(be.probabilities()[0] - 1.0) < 1e-15
    with pytest.raises(ValueError):
        vqe.prepare(be, np.zeros(5))
    with pytest.raises(ValueError):
        A.VQE(0, 1)
    with pytest.raises(ValueError):
        vqe.optimize(be, lambda p: 0.0, -1)
    with pytest.raises(ValueError):
        vqe.optimize(be, lambda p: 0.0, 1, 0.0)
    with pytest.raises(ValueError):
        A.Grover.run(be, 3, [8])
    with pytest.raises(ValueError):
        A.Grover.run(be, 2, [0, 1, 2, 3])
    with pytest.raises(ValueError):
        A.QAOA.run(be, A.Graph(3, [(0, 0, 1.0)]), [0.1], [0.2])
    with pytest.raises(ValueError):
        A.QAOA.run(be, A.Graph(3, [(0, 1, 1.0)]), [0.1], [])
