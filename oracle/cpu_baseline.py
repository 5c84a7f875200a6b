"""CPU baseline for bench.py -- TEST/MEASUREMENT INFRASTRUCTURE ONLY.

Times the reference's own CPU code paths (compiled from the reference sources into oracle/_ref/)
on the host cores, with the protocol of benchmarks/compare_libraries.py:300-304 (best of 3
perf_counter runs), and reports the host's physical core count (lscpu) and a STREAM-copy
bandwidth measured in the same run.  Run as its own process (bench.py does): the reference module
shares its name with the product module.  Prints one JSON object.

Legs (each a bounded sample of the benchmark's workloads):
* sweep (the headline value): the 30-qubit Haar + CNOT-ladder sweep's two gate kinds on one
  2^n-amplitude NumPy state (first-touched in parallel): one Haar 2x2 on each of 7 targets
  through quantum_sim_v3.apply_gate_cache_optimized (src/StateVectorOptimized.cpp:56-78, kind
  "reference"), and CNOT(c, c + 1) on 7 controls through the oracle's restatement of the
  reference's OpenMP CNOT loop (src/StateVectorBackend.h:109-135; the reference binds no CNOT on
  v3 and its StateVectorBackend stops at 25 qubits).  value = algorithmic bytes (32 * 2^n per
  one-qubit gate, 16 * 2^n per CNOT) / the summed best times.
* cnot25_reference: the reference's own StateVectorBackend::cnot at its 25-qubit cap.
* param_shift20: one parameter-shift gradient entry of the 20-qubit, 400-parameter workload
  (2 energies, src/Algorithms.h:277-292) through the reference's StateVectorBackend (RY / RZ /
  CNOT gates, Z Z terms from probabilities(), X terms after H on every qubit); a whole gradient is
  800 energies.
* trajectories14: trajectories of the 14-qubit noisy brick circuit through the oracle's C
  trajectory code (one core; the build's CPU restatement, the reference has no such path).

Without oracle/_ref/ the reference legs fall back to the oracle restatements (kind "port").
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "lightning-lite-quantum-simulator_amd"))

from oracle import oracle as O  # noqa: E402
from oracle import ref as R  # noqa: E402

REPEAT = 3  # benchmarks/compare_libraries.py:82


def log(msg):
    print(f"cpu_baseline: {msg}", file=sys.stderr, flush=True)


def best_of(fn, repeat=REPEAT):
    times = []
    for _ in range(repeat):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return min(times)


def physical_cores():
    """Physical cores from lscpu (cores per socket x sockets), else os.cpu_count()."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        f = {}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            f[k.strip()] = v.strip()
        return int(f["Core(s) per socket"]) * int(f["Socket(s)"]), f.get("Model name", "")
    except Exception:  # noqa: BLE001
        return os.cpu_count() or 1, ""


def stream_copy(n_doubles=1 << 27):
    """STREAM Copy over two arrays of n doubles (1 GiB each), first-touched in parallel; GB/s."""
    import ctypes

    L = O._lib()
    L.orc_stream_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
    L.orc_fill_zero.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    a = np.empty(n_doubles)
    b = np.empty(n_doubles)
    L.orc_fill_zero(O._ptr(a), n_doubles)
    L.orc_fill_zero(O._ptr(b), n_doubles)
    L.orc_stream_copy(O._ptr(a), O._ptr(b), n_doubles)  # warm
    t = best_of(lambda: L.orc_stream_copy(O._ptr(a), O._ptr(b), n_doubles), 5)
    return 16.0 * n_doubles / t / 1e9


def haar(rng):
    z = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    return (q * (np.diagonal(r) / np.abs(np.diagonal(r)))).reshape(4).tolist()


def sweep_leg(n, targets, controls):
    rng = np.random.Generator(np.random.PCG64(1))
    st = np.empty(1 << n, dtype=np.complex128)
    O._lib().orc_init_zero(O._ptr(st), n)  # parallel first touch
    kind = "reference" if R.available() else "port"
    if kind == "reference":
        sv = R.v3().StateVector(st)

        def apply(t, g):
            sv.apply_gate_cache_optimized(t, g)
    else:
        def apply(t, g):
            O.apply_1q(st, t, g)
    one_q = {}
    for t in targets:
        g = haar(rng)
        one_q[t] = best_of(lambda: apply(t, g))
        log(f"one-qubit target {t}: {one_q[t] * 1e3:.1f} ms")
    cx = {}
    for c in controls:
        cx[c] = best_of(lambda: O.apply_cnot(st, c, c + 1))
        log(f"CNOT ({c}, {c + 1}): {cx[c] * 1e3:.1f} ms")
    size = float(1 << n)
    total_bytes = 32.0 * size * len(one_q) + 16.0 * size * len(cx)
    total_s = sum(one_q.values()) + sum(cx.values())
    return kind, total_bytes / total_s / 1e9, one_q, cx


def cnot25_leg(n=25):
    if not R.available():
        return None
    b = R.RefBackend(n, kind=1)  # ll::StateVectorBackend
    b.h(0)
    t = {c: best_of(lambda: b.cnot(c, c + 1)) for c in (0, 8, 16, 23)}
    return {"kind": "reference", "n_qubits": n, "GBps": round(16.0 * (1 << n) * len(t) / sum(t.values()) / 1e9, 2),
            "per_control_ms": {str(c): round(v * 1e3, 3) for c, v in t.items()},
            "what": "ll::StateVectorBackend::cnot (src/StateVectorBackend.h:109-135) at its 25-qubit cap"}


def energy_on_backend(kind, n, gates, theta, terms):
    """<H> of the circuit through a fresh reference backend (kind 0 ReferenceBackend, 1
    StateVectorBackend): Z Z terms from probabilities(), X terms from probabilities() after H on
    every qubit (the measurement-basis change of a CPU user)."""
    b = R.RefBackend(n, kind=kind)
    for g in gates:
        if g[0] == "ry":
            b.ry(g[1], float(theta[g[2]]))
        elif g[0] == "rz":
            b.rz(g[1], float(theta[g[2]]))
        else:
            b.cnot(g[1], g[2])
    idx = np.arange(1 << n, dtype=np.uint64)
    p = b.probabilities()
    e = 0.0
    for c, x, z in terms:
        if x == 0:
            e += c * float(np.sum(p * (1.0 - 2.0 * (np.bitwise_count(idx & np.uint64(z)) & 1))))
    for q in range(n):
        b.h(q)
    p = b.probabilities()
    for c, x, z in terms:
        if x != 0:
            assert z == 0
            e += c * float(np.sum(p * (1.0 - 2.0 * (np.bitwise_count(idx & np.uint64(x)) & 1))))
    return e


def param_shift_leg(n=20, layers=10):
    from llsv import circuits as C

    if not R.available():
        return None
    gates = []
    for layer in range(layers):
        for q in range(n):
            gates.append(("ry", q, (layer * n + q) * 2))
            gates.append(("rz", q, (layer * n + q) * 2 + 1))
        for q in range(n - 1):
            gates.append(("cnot", q, q + 1))
    terms = C.zz_x_hamiltonian(0, n)
    theta = np.random.Generator(np.random.PCG64(0)).uniform(0, 2 * np.pi, 2 * n * layers)
    tp, tm = theta.copy(), theta.copy()
    tp[0] += np.pi / 2
    tm[0] -= np.pi / 2
    t0 = time.perf_counter()
    d0 = 0.5 * (energy_on_backend(1, n, gates, tp, terms) - energy_on_backend(1, n, gates, tm, terms))
    t_entry = time.perf_counter() - t0
    n_params = 2 * n * layers
    log(f"parameter shift: one entry {t_entry:.2f} s")
    return {"kind": "reference", "s_per_gradient_entry": round(t_entry, 3),
            "s_per_gradient": round(t_entry * n_params, 1), "gradients_per_s": 1.0 / (t_entry * n_params),
            "dE_dtheta0": d0,
            "what": (f"{n} qubits, {n_params} parameters: 2 energies per entry through ll::StateVectorBackend "
                     "(RY, RZ, CNOT; Z Z from probabilities, X after H on every qubit), extrapolated x "
                     f"{n_params} entries")}


def trajectories_leg(n=14, count=48):
    from llsv import circuits as C
    from oracle import trajectory as T

    gates, gamma, lam = C.noisy_brick(0, n, 8)
    gu = [g if g[0] == "cnot" else ("u", g[1], O.gate4(g)) for g in gates]
    packed = T.pack(gu)
    t0 = time.perf_counter()
    for i in range(count):
        T.run(n, gu, gamma, lam, 1, i, packed)
    t = time.perf_counter() - t0
    log(f"trajectories: {count} x {n} q in {t:.2f} s")
    return {"kind": "port", "cores": 1, "trajectories_per_s": round(count / t, 2),
            "trajectory_gates_per_s": round(count * len(gates) / t, 1),
            "what": f"{count} trajectories of the {n}-qubit noisy brick circuit, the oracle's C trajectory code"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--targets", default="0,5,10,15,20,25,29")
    ap.add_argument("--controls", default="0,5,10,15,20,25,28")
    ap.add_argument("--no-legs", action="store_true", help="only the sweep and the STREAM figure")
    a = ap.parse_args()
    n = a.qubits
    targets = [int(t) for t in a.targets.split(",") if int(t) < n]
    controls = [int(c) for c in a.controls.split(",") if int(c) < n - 1]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cores, model = physical_cores()
    stream = stream_copy()
    log(f"STREAM copy {stream:.1f} GB/s, {cores} physical cores, {threads} threads")
    kind, value, one_q, cx = sweep_leg(n, targets, controls)
    out = {
        "value": value,
        "unit": "GB/s",
        "cores": threads,
        "kind": kind,
        "sample": (f"{n}-qubit complex128 state; one Haar 2x2 gate per target {targets} through the reference "
                   f"quantum_sim_v3.apply_gate_cache_optimized (README build flags) and CNOT(c, c+1) for c in "
                   f"{controls} through the oracle's restatement of the reference's OpenMP CNOT loop; best of "
                   f"{REPEAT} each; {threads} OpenMP threads"),
        "host": {"physical_cores": cores, "model": model, "omp_threads": threads,
                 "stream_copy_GBps": round(stream, 1)},
        "one_qubit_ms": {str(t): round(v * 1e3, 2) for t, v in one_q.items()},
        "cnot_ms": {str(c): round(v * 1e3, 2) for c, v in cx.items()},
        "gates_per_s": (len(one_q) + len(cx)) / (sum(one_q.values()) + sum(cx.values())),
    }
    if not a.no_legs:
        out["cnot25_reference"] = cnot25_leg()
        out["param_shift20"] = param_shift_leg()
        out["trajectories14"] = trajectories_leg()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
