"""Adjoint gradients on the GPU vs the CPU oracle (needs an MI355X).

Tolerances: energies and adjoint gradients vs the oracle 1e-10 (absolute); central finite
differences with h = 1e-6 carry ~eps*|E|/h ~ 1e-9 of roundoff, so they are checked to 1e-7.
"""
import math
import sys

import numpy as np
import pytest

try:  # torch first, so libllsv and torch share one HIP runtime
    import torch
except ImportError:  # pragma: no cover
    torch = None

from llsv import circuits as C
from oracle import adjoint as A

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cuda(x):
    return torch.tensor(np.asarray(x), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("n,layers,batch", [(1, 1, 2), (4, 2, 3), (8, 3, 2), (12, 2, 2)])
def test_energy_and_adjoint_vs_oracle(n, layers, batch):
    from llsv import adjoint as LA

    gates = LA.hea_gates(n, layers)
    terms = C.zz_x_hamiltonian(n, n) if n > 1 else [(0.8, 1, 0), (-0.3, 0, 1)]
    circ = LA.Circuit(n, gates, terms)
    th = np.random.default_rng(n).uniform(0, 2 * np.pi, (batch, circ.n_params))
    say(f"adjoint n={n} layers={layers} batch={batch}")
    e, g = LA.energies_and_grads(circ, cuda(th))
    e2 = LA.energies(circ, cuda(th))
    e, g, e2 = e.cpu().numpy(), g.cpu().numpy(), e2.cpu().numpy()
    for b in range(batch):
        ce, cg = A.adjoint(n, gates, th[b], terms)
        assert abs(e[b] - ce) < 1e-10 and abs(e2[b] - ce) < 1e-10
        assert np.max(np.abs(g[b] - cg)) < 1e-10
    ps = A.param_shift(n, gates, th[0], terms)
    assert np.max(np.abs(g[0] - ps)) < 1e-10


def test_all_gate_kinds_and_y_terms():
    from helpers import random_unitary
    from llsv import adjoint as LA

    n = 6
    rng = np.random.default_rng(5)
    gates = [("rx", 0, 0), ("ry", 1, 1), ("rz", 2, 2), ("u", 3, random_unitary(rng, 2).reshape(4).tolist()),
             ("cnot", 0, 5), ("cz", 1, 4), ("rx", 5, 3), ("ry", 5, 0), ("cnot", 4, 2), ("rz", 0, 4)]
    terms = [(0.5, 0b000011, 0b000001), (-1.2, 0b100000, 0b100100), (0.9, 0, 0b010010), (0.3, 0b001000, 0)]
    circ = LA.Circuit(n, gates, terms)
    th = rng.uniform(0, 2 * np.pi, (2, circ.n_params))
    e, g = LA.energies_and_grads(circ, cuda(th))
    for b in range(2):
        ce, cg = A.adjoint(n, gates, th[b], terms)
        assert abs(e[b].item() - ce) < 1e-10
        assert np.max(np.abs(g[b].cpu().numpy() - cg)) < 1e-10
        assert np.max(np.abs(g[b].cpu().numpy() - A.param_shift(n, gates, th[b], terms))) < 1e-10


def test_autograd_function_and_param_shift():
    from llsv import adjoint as LA

    n, layers = 10, 2
    gates = LA.hea_gates(n, layers)
    circ = LA.Circuit(n, gates, C.zz_x_hamiltonian(2, n))
    th = cuda(np.random.default_rng(1).uniform(0, 2 * np.pi, (3, circ.n_params))).requires_grad_(True)
    w = cuda([1.0, -2.0, 0.5])
    loss = (LA.energy(th, circ) * w).sum()
    loss.backward()
    _, g = LA.energies_and_grads(circ, th.detach())
    assert torch.max(torch.abs(th.grad - w[:, None] * g)).item() < 1e-12
    ps = LA.param_shift_grads(circ, th.detach()[:1])
    assert torch.max(torch.abs(ps - g[:1])).item() < 1e-10
    # central finite differences, h = 1e-6 (looser: roundoff ~ 1e-9)
    fd = A.finite_diff(n, gates, th.detach()[0].cpu().numpy(), C.zz_x_hamiltonian(2, n), h=1e-6, which=range(6))
    assert np.max(np.abs(fd - g[0, :6].cpu().numpy())) < 1e-7


def test_20q_hea_batch_vs_cpu_adjoint():
    # the gradient workload's shape (20 qubits, HEA, ZZ + X Hamiltonian) at 2 layers, batch 4
    from llsv import adjoint as LA

    n, layers = 20, 2
    gates = LA.hea_gates(n, layers)
    terms = C.zz_x_hamiltonian(20, n)
    circ = LA.Circuit(n, gates, terms)
    th = np.random.default_rng(20).uniform(0, 2 * np.pi, (4, circ.n_params))
    say("adjoint 20q batch 4: compile the tiled sweeps")
    assert min(circ.precompile()) >= 1  # 20 qubits: tiled in the automatic mode
    say("adjoint 20q batch 4")
    e, g = LA.energies_and_grads(circ, cuda(th))
    ce, cg = A.adjoint(n, gates, th[3], terms)
    assert abs(e[3].item() - ce) < 1e-10
    assert np.max(np.abs(g[3].cpu().numpy() - cg)) < 1e-10
    # two entries by parameter shift on the CPU
    ps = A.param_shift(n, gates, th[3], terms, which=[0, 57])
    assert np.max(np.abs(ps - g[3, [0, 57]].cpu().numpy())) < 1e-10


def _kinds_circuit(n, rng, layers=3):
    from helpers import random_unitary

    gates, p = [], 0
    for layer in range(layers):
        for q in range(n):
            k = ("rx", "ry", "rz", "u")[(q + layer) % 4]
            if k == "u":
                gates.append(("u", q, random_unitary(rng, 2).reshape(4).tolist()))
            else:
                gates.append((k, q, p))
                p += 1
        for q in range(layer % 2, n - 1, 2):
            gates.append(("cnot", q, q + 1) if layer % 2 == 0 else ("cz", q, q + 1))
        gates.append(("cnot", n - 1, 0))
        gates.append(("cz", 0, n - 1))
    gates.append(("ry", n - 1, 0))  # a parameter used twice
    return gates


@pytest.mark.parametrize("n", [12, 13, 15])
def test_tiled_sweeps_match_gate_by_gate(n):
    # compiled 12-qubit register-tile sweeps (pass compiler forced on) vs one launch per gate
    from llsv import _native
    from llsv import adjoint as LA

    rng = np.random.default_rng(100 + n)
    gates = _kinds_circuit(n, rng)
    terms = C.zz_x_hamiltonian(n, n) + [(0.4, 0b11, 0b01), (-0.7, 1 << (n - 1), 1 << (n - 1))]
    circ = LA.Circuit(n, gates, terms)
    th = rng.uniform(0, 2 * np.pi, (3, circ.n_params))
    lib = _native.lib()
    prev = lib.llsv_set_program_jit(0)
    try:
        say(f"adjoint gate by gate n={n}")
        e0, g0 = LA.energies_and_grads(circ, cuda(th))
        lib.llsv_set_program_jit(2)
        npass = circ.precompile()
        assert min(npass) >= 1
        say(f"adjoint tiled n={n} passes={npass}")
        e1, g1 = LA.energies_and_grads(circ, cuda(th))
        e2 = LA.energies(circ, cuda(th))
    finally:
        lib.llsv_set_program_jit(prev)
    e0, g0, e1, g1, e2 = (t.cpu().numpy() for t in (e0, g0, e1, g1, e2))
    assert np.max(np.abs(e1 - e0)) < 1e-10 and np.max(np.abs(e2 - e0)) < 1e-10
    assert np.max(np.abs(g1 - g0)) < 1e-10
    ce, cg = A.adjoint(n, gates, th[2], terms)
    assert abs(e1[2] - ce) < 1e-10
    assert np.max(np.abs(g1[2] - cg)) < 1e-10


def test_bad_arguments():
    from llsv import adjoint as LA

    circ = LA.Circuit(3, [("ry", 3, 0)], [(1.0, 0, 1)])
    with pytest.raises(ValueError):
        LA.energies(circ, cuda(np.zeros((1, 1))))
    circ = LA.Circuit(3, [("ry", 0, 0)], [(1.0, 0, 0b1000)])
    with pytest.raises(ValueError):
        LA.energies(circ, cuda(np.zeros((1, 1))))
