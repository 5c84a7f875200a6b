"""The gradient oracle pinned on the CPU: closed forms, parameter shift (the reference's rule,
src/Algorithms.h:277-292), finite differences, Pauli algebra against np.kron."""
import math

import numpy as np
import pytest

from helpers import random_unitary
from llsv import circuits as C
from oracle import adjoint as A
from oracle import oracle as O

PAULI = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1.0, -1.0])}


def kron_pauli(n, word):
    M = np.array([[1.0 + 0j]])
    for q in reversed(range(n)):
        M = np.kron(M, PAULI[word[q]])
    return M


def masks(word):
    x = sum(1 << q for q, c in enumerate(word) if c in "XY")
    z = sum(1 << q for q, c in enumerate(word) if c in "YZ")
    return x, z


@pytest.mark.parametrize("word", ["X", "Y", "Z", "XY", "YY", "ZYX", "YIXZ", "YYYY"])
def test_pauli_apply_matches_kron(word):
    n = len(word)
    rng = np.random.default_rng(len(word))
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    x, z = masks(word)
    assert np.max(np.abs(A.pauli_apply(psi, x, z) - kron_pauli(n, word) @ psi)) < 1e-14


def test_closed_form_one_qubit():
    # E(t) = <0|RY(t)^+ Z RY(t)|0> = cos t; dE/dt = -sin t.  RX likewise; RZ|0> has dE/dt = 0.
    for t in (0.3, 1.7, -2.2):
        e, g = A.adjoint(1, [("ry", 0, 0)], [t], [(1.0, 0, 1)])
        assert abs(e - math.cos(t)) < 1e-15 and abs(g[0] + math.sin(t)) < 1e-15
        e, g = A.adjoint(1, [("rx", 0, 0)], [t], [(1.0, 0, 1)])
        assert abs(e - math.cos(t)) < 1e-15 and abs(g[0] + math.sin(t)) < 1e-15
        e, g = A.adjoint(1, [("h", 0) if False else ("u", 0, O.mat_h()), ("rz", 0, 0)], [t], [(1.0, 1, 0)])
        assert abs(e - math.cos(t)) < 1e-14 and abs(g[0] + math.sin(t)) < 1e-14  # <+|RZ^+ X RZ|+> = cos t


def random_pc(rng, n, depth):
    gates, p = [], 0
    for _ in range(depth):
        k = rng.integers(0, 6)
        q = int(rng.integers(0, n))
        if k < 3:
            gates.append((("rx", "ry", "rz")[k], q, p))
            p += 1
        elif k == 3:
            gates.append(("u", q, random_unitary(rng, 2).reshape(4).tolist()))
        elif n > 1:
            t = int(rng.integers(0, n - 1))
            gates.append((("cnot", "cz")[k - 4], q, t + (t >= q)))
    return gates, p


@pytest.mark.parametrize("n", [1, 2, 4, 6])
def test_adjoint_equals_param_shift(n):
    rng = np.random.default_rng(60 + n)
    gates, P = random_pc(rng, n, 40)
    words = ["".join(rng.choice(list("IXYZ"), n)) for _ in range(5)]
    terms = [(float(rng.standard_normal()),) + masks(w) for w in words]
    th = rng.uniform(0, 2 * np.pi, P)
    e, g = A.adjoint(n, gates, th, terms)
    H = sum(c * kron_pauli(n, w) for (c, _, _), w in zip(terms, words))
    psi = A.state(n, gates, th)
    assert abs(e - np.vdot(psi, H @ psi).real) < 1e-12
    assert np.max(np.abs(g - A.param_shift(n, gates, th, terms))) < 1e-12
    assert np.max(np.abs(g - A.finite_diff(n, gates, th, terms))) < 1e-7


def test_hea_hamiltonian_gradient():
    from llsv import adjoint as LA

    n, layers = 6, 2
    gates = LA.hea_gates(n, layers)
    terms = C.zz_x_hamiltonian(9, n)
    th = np.random.default_rng(4).uniform(0, 2 * np.pi, 2 * n * layers)
    # the gate list and the op-tuple ansatz builder agree
    psi1 = O.run(n, C.hea_ansatz(n, layers, th))
    psi2 = A.state(n, gates, th)
    assert np.max(np.abs(psi1 - psi2)) == 0
    e, g = A.adjoint(n, gates, th, terms)
    assert np.max(np.abs(g - A.param_shift(n, gates, th, terms))) < 1e-12


def test_planted_wrong_generator_sign_is_caught():
    # dE/dt computed with +Re instead of Im must disagree with parameter shift
    from llsv import adjoint as LA

    n = 3
    gates = LA.hea_gates(n, 1)
    terms = C.zz_x_hamiltonian(1, n)
    th = np.array([0.4, 1.3, -0.8, 0.2, 2.1, -1.4])
    ps = A.param_shift(n, gates, th, terms)
    e, g = A.adjoint(n, gates, th, terms)
    assert np.max(np.abs(ps)) > 1e-2
    assert np.max(np.abs(g - ps)) < 1e-13
    assert np.max(np.abs(-g - ps)) > 1e-3
