"""Each workload checked at its own size (needs an MI355X).

* 28-qubit layered circuit (20 layers of RX, RY, RZ on every qubit + CZ brick, 1950 gates, 4 GiB):
  run gate by gate, as one program (register-tile passes) and fused into dense <= 5-qubit blocks
  (FP64 MFMA); the three states agree to 1e-12 everywhere, and <Z_q>, the norm and 64 amplitudes
  equal the C oracle's (tests/golden/layered28_oracle.json, written on the CPU) to 1e-12.
* 32 qubits (64 GiB per state): the 12-layer layered circuit fused into dense blocks against the
  program path, and the Haar + CNOT-ladder sweep gate by gate against the program path; the oracle
  cannot hold the state, so the two independent GPU paths are compared through all 2^32
  probabilities (1e-12) and 48 Pauli-string expectations (1e-12), which see the phases too.
* Gradients at the benchmark's size: 20 qubits, 400 parameters, batch 64; energies and adjoint
  gradients of all 64 vectors against the CPU adjoint, and 16 parameters of vectors 0 and 1
  against the CPU parameter shift (tests/golden/adjoint20_cpu.npz), to 1e-10.
"""
import json
import os
import sys

import numpy as np
import pytest

try:  # torch first, so libllsv and torch share one HIP runtime
    import torch
except ImportError:  # pragma: no cover
    torch = None

from conftest import GOLDEN
from llsv import circuits as C

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def qsv():
    import quantum_sim_v3

    return quantum_sim_v3


def run_unfused(sv, ops):
    from llsv import gates as G

    for o in ops:
        if o[0] == "cz":
            sv.cz(o[1], o[2])
        elif o[0] == "cnot":
            sv.cnot(o[1], o[2])
        else:
            sv.apply_gate_cache_optimized(o[1], G.matrix(o))


def run_fused(sv, ops, max_k=5):
    import gate_fusion as GF

    for o in GF.fuse_dense(ops, max_k):
        if o[0] == "dense":
            sv.apply_fused(list(o[1]), np.ascontiguousarray(o[2]))
        else:
            sv.apply_gate_cache_optimized(o[1], list(o[2]))


def maxdiff(a, b):
    return float(np.max(np.abs(a - b)))


def test_layered28_three_ways_and_oracle():
    with open(os.path.join(GOLDEN, "layered28_oracle.json")) as f:
        fx = json.load(f)
    n = fx["n_qubits"]
    ops = C.layered_rotations(1, n, 20)
    assert len(ops) == fx["n_ops"] == 1950
    m = qsv()
    say("layered28: gate by gate")
    a = m.StateVector(n)
    run_unfused(a, ops)
    psi = a.state()
    del a
    say("layered28: program")
    m.program_jit_precompile(n, ops)
    b = m.StateVector(n)
    b.apply_program(ops)
    d_prog = maxdiff(b.state(), psi)
    del b
    say("layered28: fused")
    c = m.StateVector(n)
    run_fused(c, ops)
    zc = [c.expect_z(q) for q in range(n)]
    d_fused = maxdiff(c.state(), psi)
    say(f"layered28: program {d_prog:.2e}, fused {d_fused:.2e}")
    assert d_prog <= TOL and d_fused <= TOL
    assert maxdiff(np.array(zc), np.array(fx["z"])) <= TOL
    assert abs(np.vdot(psi, psi).real - fx["norm"]) <= TOL
    amps = np.array(fx["amp_re"]) + 1j * np.array(fx["amp_im"])
    assert maxdiff(psi[np.array(fx["index"])], amps) <= TOL


def _pauli_terms(n, count, seed):
    rng = np.random.default_rng(seed)
    terms = []
    for _ in range(count):
        x = int(rng.integers(0, 1 << n, dtype=np.uint64))
        z = int(rng.integers(0, 1 << n, dtype=np.uint64))
        terms.append((1.0, x, z))
    for q in (0, n // 2, n - 1):  # single-qubit X and Y: phases of one bit
        terms += [(1.0, 1 << q, 0), (1.0, 1 << q, 1 << q)]
    return terms


def _compare_32q(a, b, n, what):
    terms = _pauli_terms(n, 42, 32)
    ea = np.array([a.expect_pauli([t]) for t in terms])
    eb = np.array([b.expect_pauli([t]) for t in terms])
    say(f"q32 {what}: Pauli {maxdiff(ea, eb):.2e}")
    assert maxdiff(ea, eb) <= TOL
    pa = a.probabilities()
    pb = b.probabilities()
    step = 1 << 26
    d = max(float(np.max(np.abs(pa[i:i + step] - pb[i:i + step]))) for i in range(0, pa.size, step))
    del pa, pb
    say(f"q32 {what}: probabilities {d:.2e}")
    assert d <= TOL
    assert abs(a.norm() - 1) <= TOL and abs(b.norm() - 1) <= TOL


def test_32q_fused_vs_program_and_sweep():
    n = 32
    m = qsv()
    ops = C.layered_rotations(1, n, 12)
    say("q32: layered fused")
    a = m.StateVector(n)
    run_fused(a, ops)
    say("q32: layered program")
    m.program_jit_precompile(n, ops)
    b = m.StateVector(n)
    b.apply_program(ops)
    _compare_32q(a, b, n, "layered fused vs program")
    sweep = C.haar_cnot_sweep(0, n, 1)
    a.reset()
    b.reset()
    say("q32: sweep gate by gate")
    for o in sweep:
        if o[0] == "cnot":
            a.cnot(o[1], o[2])
        else:
            a.apply_gate_cache_optimized(o[1], list(o[2]))
    say("q32: sweep program")
    b.apply_program(sweep)
    _compare_32q(a, b, n, "sweep gate by gate vs program")


def test_adjoint20_full_batch_vs_cpu():
    from llsv import adjoint as LA

    fx = np.load(os.path.join(GOLDEN, "adjoint20_cpu.npz"), allow_pickle=False)
    n, layers, batch = 20, 10, 64
    gates = LA.hea_gates(n, layers)
    circ = LA.Circuit(n, gates, C.zz_x_hamiltonian(0, n))
    th = np.random.Generator(np.random.PCG64(0)).uniform(0, 2 * np.pi, (batch, circ.n_params))
    say("adjoint20: compile the tiled sweeps")
    assert min(circ.precompile()) >= 1
    say("adjoint20: batch 64")
    e, g = LA.energies_and_grads(circ, torch.tensor(th, dtype=torch.float64, device="cuda"))
    e, g = e.cpu().numpy(), g.cpu().numpy()
    de = maxdiff(e, fx["energies"])
    dg = maxdiff(g, fx["grads"])
    ds = maxdiff(g[:2, fx["shift_params"]], fx["shift_grads"])
    say(f"adjoint20: energies {de:.2e}, gradients {dg:.2e}, vs parameter shift {ds:.2e}")
    assert de <= 1e-10 and dg <= 1e-10 and ds <= 1e-10
