"""Writes tests/golden/dm_brick14.json: the exact <Z_q> of the 14-qubit noisy trajectory workload
(per layer H and RZ on every qubit and a CNOT brick, 8 layers; amplitude damping gamma_q then
phase damping gamma_q / 2 after every gate on each touched qubit -- llsv.circuits.noisy_brick(0, 14,
8), the circuit the trajectory benchmark runs) from its density matrix.

The density matrix is evolved by the oracle's C restatement of the reference's block update
KrausChannel::apply (src/Noise.cpp:174-258; oracle/dm_oracle.c), every gate as the one-operator
channel {U}.  rho is 4^14 x 16 B = 4 GiB; the run takes about ten minutes on 8 cores.  A smaller
case (10 qubits, same builder) is written beside it so the CPU tests can recompute one entry.

    python tests/golden/make_dm14_fixture.py [--qubits 14]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "lightning-lite-quantum-simulator_amd"))

from llsv import circuits as C  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle import trajectory as T  # noqa: E402


def exact_z(n, seed, layers):
    gates, gamma, lam = C.noisy_brick(seed, n, layers)
    gu = [g if g[0] == "cnot" else ("u", g[1], O.gate4(g)) for g in gates]
    t0 = time.perf_counter()
    rho = T.density_matrix_blocks(n, gu, gamma, lam)
    z, tr = T.dm_expect_z(n, rho)
    dt = time.perf_counter() - t0
    del rho
    return {"n_qubits": n, "circuit_seed": seed, "layers": layers, "n_gates": len(gates),
            "gamma": [float(x) for x in gamma], "z": [float(x) for x in z], "trace": tr,
            "seconds": round(dt, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=14)
    a = ap.parse_args()
    out = {"what": "exact <Z_q> of llsv.circuits.noisy_brick(seed, n, layers) under AD(gamma_q) then "
                   "PD(gamma_q / 2) after every gate on each touched qubit, from the density matrix "
                   "(block update of src/Noise.cpp:174-258, oracle/dm_oracle.c)",
           "cases": [exact_z(10, 0, 8), exact_z(a.qubits, 0, 8)]}
    path = os.path.join(HERE, f"dm_brick{a.qubits}.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, [c["seconds"] for c in out["cases"]])


if __name__ == "__main__":
    main()
