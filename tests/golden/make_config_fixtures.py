"""Writes the CPU oracle's results for two workloads at their full size, so GPU tests can check them
without running the oracle on the GPU box (each takes minutes on the CPU):

* adjoint20_cpu.npz -- the gradient workload: 20 qubits, 10 layers of RY + RZ on every qubit and a
  CNOT ladder (400 parameters), H = sum a_i Z_i Z_{i+1} + sum b_i X_i (llsv.circuits.zz_x_hamiltonian
  seed 0), 64 parameter vectors uniform in [0, 2 pi) from PCG64(0).  Energies and adjoint gradients
  of all 64 vectors (oracle/adjoint.py: the C oracle's gates, NumPy Pauli sums), and the parameter
  shift of src/Algorithms.h:277-292 for 16 random parameters of vectors 0 and 1.
* layered28_oracle.json -- the 28-qubit layered circuit: 20 layers of RX, RY, RZ on every qubit and a
  CZ brick (llsv.circuits.layered_rotations(1, 28, 20), 1950 gates) through the C oracle (4 GiB
  state): <Z_q> for every qubit, the norm, and 64 amplitudes at seeded indices.

    python tests/golden/make_config_fixtures.py [adjoint20] [layered28]
"""
import json
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "lightning-lite-quantum-simulator_amd"))

from llsv import circuits as C  # noqa: E402
from oracle import adjoint as A  # noqa: E402
from oracle import oracle as O  # noqa: E402

N_ADJ, LAYERS_ADJ, BATCH = 20, 10, 64


def adjoint_case():
    """(gates, terms, theta) of the gradient workload, built as the benchmark builds it."""
    gates = []
    for layer in range(LAYERS_ADJ):
        for q in range(N_ADJ):
            gates.append(("ry", q, (layer * N_ADJ + q) * 2))
            gates.append(("rz", q, (layer * N_ADJ + q) * 2 + 1))
        for q in range(N_ADJ - 1):
            gates.append(("cnot", q, q + 1))
    terms = C.zz_x_hamiltonian(0, N_ADJ)
    theta = np.random.Generator(np.random.PCG64(0)).uniform(0, 2 * np.pi, (BATCH, 2 * N_ADJ * LAYERS_ADJ))
    return gates, terms, theta


def _adj_one(b):
    gates, terms, theta = adjoint_case()
    return A.adjoint(N_ADJ, gates, theta[b], terms)


def _shift_one(args):
    b, k = args
    gates, terms, theta = adjoint_case()
    return A.param_shift(N_ADJ, gates, theta[b], terms, which=[k])[0]


def make_adjoint20():
    t0 = time.perf_counter()
    _, _, theta = adjoint_case()
    with Pool(8) as pool:
        res = pool.map(_adj_one, range(BATCH))
        which = np.random.Generator(np.random.PCG64(1)).choice(theta.shape[1], 16, replace=False)
        shift = pool.map(_shift_one, [(b, int(k)) for b in (0, 1) for k in which])
    energies = np.array([r[0] for r in res])
    grads = np.stack([r[1] for r in res])
    np.savez_compressed(os.path.join(HERE, "adjoint20_cpu.npz"), energies=energies, grads=grads,
                        shift_params=which.astype(np.int64), shift_grads=np.array(shift).reshape(2, 16))
    print(f"adjoint20: {time.perf_counter() - t0:.0f} s", flush=True)


def make_layered28():
    t0 = time.perf_counter()
    n = 28
    ops = C.layered_rotations(1, n, 20)
    psi = O.run(n, ops)
    z = [O.expect_z(psi, q) for q in range(n)]
    idx = np.random.Generator(np.random.PCG64(28)).integers(0, 1 << n, 64)
    amps = psi[idx]
    out = {"what": "C oracle on llsv.circuits.layered_rotations(1, 28, 20)", "n_qubits": n,
           "n_ops": len(ops), "norm": O.norm2(psi), "z": z, "index": [int(i) for i in idx],
           "amp_re": [float(a.real) for a in amps], "amp_im": [float(a.imag) for a in amps],
           "seconds": round(time.perf_counter() - t0, 1)}
    with open(os.path.join(HERE, "layered28_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"layered28: {out['seconds']} s", flush=True)


if __name__ == "__main__":
    which = sys.argv[1:] or ["adjoint20", "layered28"]
    if "adjoint20" in which:
        make_adjoint20()
    if "layered28" in which:
        make_layered28()
