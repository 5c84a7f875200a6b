"""CPU baseline for bench.py -- TEST/MEASUREMENT INFRASTRUCTURE ONLY.

Times the reference's own one-qubit gate loop on the host cores: quantum_sim_v3's
``apply_gate_cache_optimized`` (src/StateVectorOptimized.cpp:56-78) compiled from the
reference sources with the README build flags into oracle/_ref/ (kind "reference").  If
oracle/_ref/ is absent it falls back to the oracle restatement liborc.so (kind "port").

Run as its own process (bench.py does): the reference module shares its name with the
product module.  Prints one JSON object.

Sample: an n-qubit complex128 state held in a NumPy array (first-touched in parallel by
the oracle's OpenMP fill), borrowed by the reference StateVector; one Haar-random 2x2 gate
on each listed target; best of `repeat` runs per target (perf_counter, as
benchmarks/compare_libraries.py:300-304).  Value = sum of algorithmic bytes
(32 * 2^n per gate) / sum of best times.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402
from oracle import ref as R  # noqa: E402


def haar(rng):
    z = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    return (q * (np.diagonal(r) / np.abs(np.diagonal(r)))).reshape(4).tolist()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--targets", default="0,5,10,15,20,25,29")
    ap.add_argument("--repeat", type=int, default=2)
    a = ap.parse_args()
    n = a.qubits
    targets = [int(t) for t in a.targets.split(",") if int(t) < n]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.Generator(np.random.PCG64(1))
    st = np.empty(1 << n, dtype=np.complex128)
    O.zero_state  # noqa: B018  (ensures liborc is loadable before the big allocation is touched)
    O._lib().orc_init_zero(O._ptr(st), n)  # parallel first touch
    kind = "reference" if R.available() else "port"
    if kind == "reference":
        sv = R.v3().StateVector(st)

        def apply(t, g):
            sv.apply_gate_cache_optimized(t, g)
    else:
        def apply(t, g):
            O.apply_1q(st, t, g)
    best = {}
    for t in targets:
        g = haar(rng)
        times = []
        for _ in range(a.repeat):
            t0 = time.perf_counter()
            apply(t, g)
            times.append(time.perf_counter() - t0)
        best[t] = min(times)
        print(f"cpu_baseline: target {t}: {best[t] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    total_bytes = 32.0 * (1 << n) * len(targets)
    total_s = sum(best.values())
    out = {
        "value": total_bytes / total_s / 1e9,
        "unit": "GB/s",
        "cores": threads,
        "kind": kind,
        "sample": (f"{n}-qubit complex128 state; one Haar 2x2 gate per target {targets} through "
                   + ("the reference quantum_sim_v3.apply_gate_cache_optimized (README build flags)"
                      if kind == "reference" else "the oracle restatement of it")
                   + f"; best of {a.repeat}; {threads} OpenMP threads; CNOT not sampled (v3 binds none)"),
        "per_target_ms": {str(t): round(v * 1e3, 2) for t, v in best.items()},
        "gates_per_s": len(targets) / total_s,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
