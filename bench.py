#!/usr/bin/env python3
"""Benchmark: HBM bandwidth of the streamed complex128 gate sweep on one MI355X.

Workload (one step): an n = 30 qubit complex128 state (16 GiB in HBM); a Haar-random 2x2
gate on every qubit 0..29, then a CNOT ladder (q, q+1), q = 0..28 -- 59 gates, each one
streamed pass over the state, launched one call at a time through the quantum_sim_v3
StateVector surface (apply_gate_cache_optimized / cnot), exactly as a reference user
drives it.  Inputs stay resident in HBM for the whole run.

value     = whole-job algorithmic HBM bytes / wall time (GB/s): 32 * 2^n per one-qubit
            gate, 16 * 2^n per CNOT (only the control = 1 half moves).
roofline  = the dominant kernel (the one-qubit gate): 32 * 2^n bytes / its average launch
            duration, timed with HIP events recorded on the state's own stream around every
            one-qubit launch inside the timed region; peak 8000 GB/s.
cpu_baseline = oracle/cpu_baseline.py in a subprocess (reference v3 loop on host cores).

The same line carries the other paths' measurements (skipped with --no-extra; --leg NAME runs one
alone and prints only it): latency12, layered28, adjoint20, trajectories14, zne12, zne16,
lindblad10, q32 (program28 only through --leg).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--qubits 30] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "lightning-lite-quantum-simulator_amd")
sys.path.insert(0, PKG)

METRIC = "HBM GB/s (% of peak) gate sweeps @30q c128; adjoint grads/s vs param-shift"
PEAK_HBM_GBPS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(n):
    cmd = [sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), "--qubits", str(n)]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        sys.stderr.write(r.stderr)
        if r.returncode != 0:
            return {"error": f"cpu_baseline exited {r.returncode}"}
        return json.loads(r.stdout.strip().splitlines()[-1])
    except Exception as e:  # reported, never fatal for the GPU measurement
        return {"error": str(e)}


def load_traffic(n):
    path = os.path.join(ROOT, "profiles", f"traffic_apply_1q_{n}q.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def layered_leg(V3, n=28, layers=20):
    """28-qubit layered circuit: per layer RX, RY, RZ on every qubit, then a CZ brick.
    Run unfused (one streamed pass per gate), as one program (LDS passes) and fused into
    dense <= 5-qubit blocks (FP64 MFMA for 4-5 qubit blocks)."""
    import gate_fusion as GF
    from llsv import circuits as C, gates as G

    size = 1 << n
    ops = C.layered_rotations(1, n, layers)
    n1 = sum(1 for o in ops if o[0] != "cz")
    ncz = len(ops) - n1
    alg_bytes = 32.0 * size * n1 + 8.0 * size * ncz
    calls = [(o[1], G.matrix(o)) if o[0] != "cz" else (o[1], o[2]) for o in ops]
    kinds = [o[0] != "cz" for o in ops]
    sv = V3.StateVector(n)

    def unfused():
        for one, c in zip(kinds, calls):
            if one:
                sv.apply_gate_cache_optimized(c[0], c[1])
            else:
                sv.cz(c[0], c[1])

    unfused()
    sv.sync()
    t0 = time.perf_counter()
    unfused()
    sv.sync()
    t_unf = time.perf_counter() - t0
    log(f"layered {n}q unfused: {t_unf:.3f} s")

    from llsv import program as P

    n_passes = len(P.plan(n, ops, sv.program_tile_bits())[1])

    def timed_program():
        sv.apply_program(ops)
        sv.sync()
        t0 = time.perf_counter()
        sv.apply_program(ops)
        sv.sync()
        return time.perf_counter() - t0

    prev = V3.set_program_jit(0)
    t_interp = timed_program()
    V3.set_program_jit(prev)
    tj = time.perf_counter()
    n_jit = V3.program_jit_precompile(n, ops)
    t_jit = time.perf_counter() - tj
    t_prog = timed_program()
    log(f"layered {n}q program: {n_passes} passes, {t_prog:.3f} s ({n_jit} compiled pass kernels, compile "
        f"{t_jit:.1f} s; interpreter {t_interp:.3f} s)")

    th = time.perf_counter()
    fused = GF.fuse_dense(ops, 5)
    t_host = time.perf_counter() - th
    blocks = [(list(o[1]), np.ascontiguousarray(o[2])) if o[0] == "dense" else (o[1], list(o[2])) for o in fused]
    dense = [o[0] == "dense" for o in fused]
    flops = sum(8.0 * (1 << len(o[1])) * size for o in fused if o[0] == "dense")

    def run_fused():
        for d, b in zip(dense, blocks):
            if d:
                sv.apply_fused(b[0], b[1])
            else:
                sv.apply_gate_cache_optimized(b[0], b[1])

    run_fused()
    sv.sync()
    t0 = time.perf_counter()
    sv.timer_start()
    run_fused()
    dev_ms = sv.timer_stop()
    sv.sync()
    t_fused = time.perf_counter() - t0
    k5 = sum(1 for o in fused if o[0] == "dense" and len(o[1]) == 5)
    log(f"layered {n}q fused: {len(fused)} blocks ({k5} of 5 qubits), {t_fused:.3f} s")

    # the fusion cost model (gate_fusion.plan_fusion): dense blocks only where predicted to pay
    def timed(fn):
        fn()
        sv.sync()
        t0 = time.perf_counter()
        fn()
        sv.sync()
        return time.perf_counter() - t0

    segs, t_pred = GF.plan_fusion(n, ops)
    t_auto = timed(lambda: GF.run_fusion_plan(sv, segs))
    # a deep circuit on five of the 28 qubits (300 rotations and CNOTs): one register-tile pass is
    # FP64-bound there, one dense 5-qubit block is one HBM pass on the matrix cores
    rng = np.random.Generator(np.random.PCG64(5))
    qs = [3, 7, 11, 19, 25]
    deep = []
    for _ in range(300):
        if rng.random() < 0.7:
            deep.append((("rx", "ry", "rz")[int(rng.integers(3))], int(rng.choice(qs)), float(rng.uniform(0, 2 * np.pi))))
        else:
            c, t = rng.choice(qs, 2, replace=False)
            deep.append(("cnot", int(c), int(t)))
    dsegs, d_pred = GF.plan_fusion(n, deep)
    t_deep_prog = timed(lambda: sv.apply_program(deep))
    t_deep_auto = timed(lambda: GF.run_fusion_plan(sv, dsegs))
    log(f"layered {n}q fusion plan: {[k for k, _ in segs]} {t_auto:.3f} s; deep 5-qubit circuit: program "
        f"{t_deep_prog * 1e3:.2f} ms, plan {[k for k, _ in dsegs]} {t_deep_auto * 1e3:.2f} ms")
    del sv
    return {
        "workload": f"{n}-qubit complex128 layered circuit, {layers} layers of RX/RY/RZ on every qubit + CZ brick "
                    f"({n1} rotations, {ncz} CZ)",
        "unfused": {"s": round(t_unf, 4), "gates_per_s": round(len(ops) / t_unf, 1),
                    "hbm_GBps": round(alg_bytes / t_unf / 1e9, 1)},
        "program": {"s": round(t_prog, 4), "passes": n_passes, "gates_per_s": round(len(ops) / t_prog, 1),
                    "hbm_GBps": round(32.0 * size * n_passes / t_prog / 1e9, 1),
                    "effective_GBps": round(alg_bytes / t_prog / 1e9, 1),
                    "compiled_pass_kernels": n_jit, "compile_s": round(t_jit, 2),
                    "interpreter_s": round(t_interp, 4)},
        "fused_mfma": {"s": round(t_fused, 4), "blocks": len(fused), "blocks_5q": k5,
                       "fusion_host_s": round(t_host, 3), "gates_per_s": round(len(ops) / t_fused, 1),
                       "hbm_GBps": round(32.0 * size * len(fused) / (dev_ms / 1e3) / 1e9, 1),
                       "TFLOPs": round(flops / (dev_ms / 1e3) / 1e12, 2),
                       "mfma_fraction": round(flops / (dev_ms / 1e3) / 78.6e12, 4)},
        "fused_auto": {"s": round(t_auto, 4), "segments": [k for k, _ in segs], "predicted_s": round(t_pred, 4),
                       "gates_per_s": round(len(ops) / t_auto, 1)},
        "deep5": {"what": "300 rotations and CNOTs on qubits 3, 7, 11, 19, 25 of the same 28-qubit state",
                  "program_s": round(t_deep_prog, 5), "fused_auto_s": round(t_deep_auto, 5),
                  "segments": [k for k, _ in dsegs], "predicted_s": round(d_pred, 5)},
    }


def adjoint_leg(n=20, layers=10, batch=64):
    """Hardware-efficient ansatz, 10 layers of RY+RZ on every qubit + CNOT ladder (400 params),
    H = sum a_i Z_i Z_{i+1} + sum b_i X_i; a batch of 64 parameter vectors.  Adjoint gradients
    vs GPU parameter shift (2 * 400 shifted energies per vector, batched 256 at a time)."""
    import torch

    from llsv import adjoint as LA, circuits as C

    gates = LA.hea_gates(n, layers)
    circ = LA.Circuit(n, gates, C.zz_x_hamiltonian(0, n))
    rng = np.random.Generator(np.random.PCG64(0))
    th = torch.tensor(rng.uniform(0, 2 * np.pi, (batch, circ.n_params)), dtype=torch.float64, device="cuda")
    # compile the tiled sweeps (12-qubit register-tile passes) before the timed batch
    tc = time.perf_counter()
    fwd, bwd = circ.precompile()
    t_compile = time.perf_counter() - tc
    log(f"adjoint: {fwd} forward + {bwd} backward tiled passes compiled in {t_compile:.1f} s")
    def best_of_3(fn):
        fn()  # warm-up
        torch.cuda.synchronize()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts), ts, out

    t_adj, adj_runs, (e, g) = best_of_3(lambda: LA.energies_and_grads(circ, th))
    log(f"adjoint: {batch} gradients of {circ.n_params} params at {n} q in {t_adj:.3f} s (best of 3)")
    nps = 1
    t_ps, ps_runs, ps = best_of_3(lambda: LA.param_shift_grads(circ, th[:nps]))
    err = torch.max(torch.abs(ps - g[:nps])).item()
    log(f"parameter shift: {nps} gradient in {t_ps:.3f} s (best of 3), max |adjoint - shift| {err:.2e}")
    ws = circ.workspace_bytes(batch)
    return {
        "workload": f"{n}-qubit HEA, {layers} layers RY+RZ + CNOT ladder ({circ.n_params} params), "
                    f"H = sum a ZZ + sum b X ({circ.n_terms} terms), batch {batch}, complex128",
        "adjoint_grads_per_s": round(batch / t_adj, 2),
        "adjoint_s_per_batch": round(t_adj, 4),
        "param_shift_grads_per_s": round(nps / t_ps, 4),
        "speedup_vs_param_shift": round((batch / t_adj) / (nps / t_ps), 1),
        "timing": "warm-up then best of 3 for both; runs (s): adjoint "
                  + ", ".join(f"{t:.4f}" for t in adj_runs) + "; parameter shift " + ", ".join(f"{t:.3f}" for t in ps_runs),
        "max_abs_grad_diff_vs_param_shift": err,
        "workspace_GiB": round(ws / 2**30, 2),
        "tiled_passes": [fwd, bwd],
        "compile_s": round(t_compile, 2),
        # HBM traffic per amplitude and batch element: a forward pass reads and writes psi (32 B),
        # a backward pass psi and lambda (64 B), H|psi> reads psi and writes lambda (32 B); one
        # pass per gate when the sweeps run gate by gate
        "adjoint_hbm_GBps": round(((fwd or len(gates)) * 32.0 + (bwd or len(gates)) * 64.0 + 32.0) * (1 << n) * batch
                                  / t_adj / 1e9, 1),
    }


def trajectory_leg(n=14, count=4096, mixed=2048):
    """4096 noisy trajectories of a 14-qubit brick circuit (8 layers of H, RZ, CNOT brick; AD then
    PD after every gate on each touched qubit) in one persistent launch -- one-shot with the
    default launch, then as an autotuned plan replayed as a hipGraph (the reported rate); a mixed
    batch of 2048 circuits of 10..16 qubits likewise; and the 4096-trajectory ensemble's <Z_q>
    against the exact density matrix of the same circuit (a committed fixture)."""
    from llsv import circuits as C, trajectory as LT

    gates, gamma, lam = C.noisy_brick(0, n, 8)
    LT.run_trajectories(n, gates, gamma, lam, seed=1, n_traj=64)
    t0 = time.perf_counter()
    z, j, _ = LT.run_trajectories(n, gates, gamma, lam, seed=1, n_traj=count)
    t_traj = time.perf_counter() - t0
    sites = sum(2 if g[0] == "cnot" else 1 for g in gates) * 2
    log(f"trajectories (one-shot, default launch): {count} x {n} q in {t_traj:.3f} s ({int(j.sum())} jumps)")
    # the same batch as a plan: autotuned launch configuration, one run = one hipGraph replay
    tp = time.perf_counter()
    plan = LT.TrajectoryPlan.from_circuit(n, gates, gamma, lam, count, tune=True)
    t_plan = time.perf_counter() - tp
    threads, workgroups, fused, tune_ms = plan.config()
    plan.run(1)
    reps = 5
    t0 = time.perf_counter()
    for r in range(reps):
        zp, jp, _ = plan.run(100 + r)
    t_graph = (time.perf_counter() - t0) / reps
    plan.close()
    log(f"trajectories (tuned plan, graph replay): {t_graph:.4f} s per batch ({threads} threads x {workgroups} "
        f"workgroups, fused={fused}; plan {t_plan:.2f} s incl. tuning {tune_ms / 1e3:.2f} s)")
    batch = C.mixed_batch(1, count=mixed)
    t0 = time.perf_counter()
    zb, jb, _ = LT.run_batch(batch, seed=2)
    t_mixed = time.perf_counter() - t0
    mixed_gates = sum(len(b[1]) for b in batch)
    mplan = LT.TrajectoryPlan(batch, tune=True)
    mcfg = mplan.config()
    mplan.run(2)
    t0 = time.perf_counter()
    mplan.run(3)
    t_mixed_graph = time.perf_counter() - t0
    mplan.close()
    log(f"mixed batch: {mixed} circuits of 10..16 q in {t_mixed:.3f} s one-shot, {t_mixed_graph:.3f} s tuned plan")
    # ensemble check: the same 4096 x 14-qubit batch against the exact density matrix of this
    # circuit (committed fixture, tests/golden/dm_brick14.json)
    dev = None
    fx_path = os.path.join(ROOT, "tests", "golden", "dm_brick14.json")
    if os.path.exists(fx_path):
        with open(fx_path) as f:
            fx = next(c for c in json.load(f)["cases"] if c["n_qubits"] == n)
        zt = z[:, :n]
        se = zt.std(0, ddof=1) / np.sqrt(count)
        dev = round(float(np.max(np.abs(zt.mean(0) - np.array(fx["z"])) / se)), 2)
    return {
        "workload": f"{count} trajectories, {n}-qubit brick circuit, 8 layers ({len(gates)} gates, {sites} noise "
                    f"channels per trajectory), gamma ~ U[1e-3, 1e-2], lambda = gamma / 2, Philox4x32-10",
        "s": round(t_graph, 4),
        "trajectory_gates_per_s": round(count * len(gates) / t_graph, 1),
        "trajectories_per_s": round(count / t_graph, 1),
        # units of the plan for this config: a gate or noise site on one trajectory moves 32 * 2^n
        # bytes if streamed (the states stay on chip, so this exceeds the HBM peak); a one-qubit
        # gate is 14 * 2^n flops (the noise sites' arithmetic is left out)
        "effective_GBps_if_streamed": round(count * (len(gates) + sites) * 32.0 * (1 << n) / t_graph / 1e9, 1),
        "fp64_fraction_one_qubit_gates": round(
            count * sum(1 for g in gates if g[0] != "cnot") * 14.0 * (1 << n) / t_graph / 78.6e12, 4),
        "jumps": int(j.sum()),
        "tuned_plan": {"threads": threads, "workgroups": workgroups, "fused": fused,
                       "tuning_s": round(tune_ms / 1e3, 3), "plan_create_s": round(t_plan, 3),
                       "s_per_batch_graph_replay": round(t_graph, 4)},
        "one_shot_default_launch_s": round(t_traj, 4),
        "mixed_batch": {"circuits": mixed, "qubits": "10..16", "s": round(t_mixed_graph, 4),
                        "gates_per_s": round(mixed_gates / t_mixed_graph, 1), "one_shot_s": round(t_mixed, 4),
                        "tuned": {"threads": mcfg[0], "workgroups": mcfg[1], "fused": mcfg[2],
                                  "tuning_s": round(mcfg[3] / 1e3, 3)}},
        "ensemble_vs_density_matrix_max_sigma": dev,
    }


def latency_leg(V3, n=12, n_ops=1000, seeds=10):
    """12 qubits, 1000 ops each (H on a random qubit or CNOT on a random ordered pair, p = 0.5;
    seeds 0..9): latency-bound (the 64 KiB state fits one workgroup's LDS).  Per-gate calls
    through the v3 surface (the reference's usage: one Python call per gate) and the whole
    circuit as one program (one pass, one launch)."""
    from llsv import circuits as C, gates as G

    circuits = [C.random_h_cnot(s, n, n_ops) for s in range(seeds)]
    H = G.matrix(("h", 0))

    def per_gate(ops, sv):
        for op in ops:
            if op[0] == "h":
                sv.apply_gate_cache_optimized(op[1], H)
            else:
                sv.cnot(op[1], op[2])
        sv.sync()

    sv = V3.StateVector(n)
    per_gate(circuits[0], sv)
    t0 = time.perf_counter()
    for ops in circuits:
        sv.reset()
        per_gate(ops, sv)
    t_gate = (time.perf_counter() - t0) / seeds
    sv.apply_program(circuits[0])
    sv.sync()
    t0 = time.perf_counter()
    for ops in circuits:
        sv.reset()
        sv.apply_program(ops)
        sv.sync()
    t_prog = (time.perf_counter() - t0) / seeds
    # where a program's time goes: the host call (op conversion, plan, upload, launch) and the
    # device time between events recorded on the state's stream around it
    host, dev = [], []
    for ops in circuits:
        sv.reset()
        sv.sync()
        sv.mark(0)
        th = time.perf_counter()
        sv.apply_program(ops)
        host.append(time.perf_counter() - th)
        sv.mark(1)
        dev.append(sv.mark_elapsed(0, 1) / 1e3)
    t_host, t_dev = sum(host) / seeds, sum(dev) / seeds
    log(f"latency {n}q x {n_ops} ops: per-gate {t_gate * 1e3:.2f} ms, program {t_prog * 1e3:.3f} ms per circuit "
        f"(host call {t_host * 1e3:.3f} ms, device {t_dev * 1e3:.3f} ms)")
    del sv
    return {"workload": f"{n}-qubit complex128, {n_ops} random H / CNOT ops (seeds 0..{seeds - 1})",
            "per_gate_ms_per_circuit": round(t_gate * 1e3, 3), "per_gate_ops_per_s": round(n_ops / t_gate, 1),
            "program_ms_per_circuit": round(t_prog * 1e3, 4), "program_ops_per_s": round(n_ops / t_prog, 1),
            "program_host_call_ms": round(t_host * 1e3, 4), "program_device_ms": round(t_dev * 1e3, 4)}


def zne_leg(n=12, layers=4, p=0.005, trajectories=8192):
    """Zero-noise extrapolation of <sum_i Z_i> / n on an n-qubit circuit (layers of RY + RZ on every
    qubit and a CNOT ladder), depolarizing p = 0.005 per site, scales 1, 3, 5: with the reference's
    noise stream (mt19937_64 draws on the host, the same Pauli insertions as the reference's
    NoisyBackend for the same seed) and with device Philox draws.  Leg zne12: 12 qubits, 4 layers
    (140 gates), 8192 trajectories per scale, every state in LDS; leg zne16: 16 qubits, 2 layers
    (94 gates), 1024 trajectories per scale, every state in a device-memory slot."""
    import quantum_sim_mitigation as QM

    rng = np.random.Generator(np.random.PCG64(3))
    gates = []
    for _ in range(layers):
        for q in range(n):
            gates.append(("ry", q, float(rng.uniform(0, 2 * np.pi))))
            gates.append(("rz", q, float(rng.uniform(0, 2 * np.pi))))
        for q in range(n - 1):
            gates.append(("cnot", q, q + 1))
    w = QM.Diagonal.from_function(n, lambda i: sum(1.0 - 2.0 * ((i >> q) & 1) for q in range(n)) / n)

    def circ(b):
        for g in gates:
            if g[0] == "cnot":
                b.cnot(g[1], g[2])
            else:
                getattr(b, g[0])(g[1], g[2])

    out = {"workload": f"{n}-qubit circuit, {len(gates)} gates ({layers} layers RY+RZ + CNOT ladder), "
                       f"depolarizing p = {p}, scales 1/3/5, {trajectories} trajectories per scale, complex128"}
    apps = len(gates) * (1 + 3 + 5) * trajectories
    for mode in ("reference", "philox"):
        nb = QM.NoisyBackend(QM.ReferenceBackend(n), p, 7, rng=mode)
        QM.ZNE(256).extrapolate(nb, circ, w, [1, 3, 5])  # warm-up (buffers, module load)
        t0 = time.perf_counter()
        r = QM.ZNE(trajectories).extrapolate(nb, circ, w, [1, 3, 5])
        dt = time.perf_counter() - t0
        log(f"zne ({mode}): {3 * trajectories} trajectories in {dt:.3f} s, E0 = {r.value:.6f} +- {r.std_error:.1e} "
            f"(unmitigated {r.values[0]:.6f})")
        out[mode] = {"s_per_extrapolation": round(dt, 4), "trajectories_per_s": round(3 * trajectories / dt, 1),
                     "gate_applications_per_s": round(apps / dt, 1), "value": r.value, "std_error": r.std_error,
                     "unmitigated": r.values[0]}
    # the noiseless value (p = 0: one trajectory)
    ideal = QM.ZNE(1).extrapolate(QM.NoisyBackend(QM.ReferenceBackend(n), 0.0, 1), circ, w, [1, 3]).value
    out["ideal"] = ideal
    return out


def zne16_leg():
    return zne_leg(n=16, layers=2, trajectories=1024)


def lindblad_leg(n=10, n_traj=4096, steps=20, dt=0.01):
    """Dense Lindblad trajectories at 10 qubits (dim 1024): random Hermitian H, amplitude damping on
    qubit 0 and dephasing on qubit 9, 4096 STANDARD-unraveling trajectories x 20 steps from a
    random state, the reference's random streams.  Each step propagates every trajectory with one
    4096 x 1024 by 1024 x 1024 complex GEMM (8 dim^2 T flop) on the FP64 matrix cores; timed per
    step with HIP events around the GEMM alone, and for the whole ensemble call."""
    import quantum_sim_trajectory as QT

    d = 1 << n
    rng = np.random.Generator(np.random.PCG64(10))
    a = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    H = (a + a.conj().T) / (2 * np.sqrt(d))
    sm = np.array([[0, 1], [0, 0]], dtype=complex)
    sz = np.diag([1.0, -1.0]).astype(complex)
    s = QT.TrajectorySolver(n)
    s.set_hamiltonian(H)
    s.add_lindblad(np.kron(np.eye(d // 2), sm), 0.5)
    s.add_lindblad(np.kron(sz, np.eye(d // 2)), 0.3)
    psi0 = rng.normal(size=d) + 1j * rng.normal(size=d)
    duration = steps * dt
    s.evolve_ensemble(psi0, duration, dt, 64, QT.Unraveling.STANDARD, 1)  # warm-up
    # a whole call with the plan built afresh, then a repeated call (the solver keeps the plan of a
    # constant H for the next call on the same time grid)
    s._plan_cache = None
    t0 = time.perf_counter()
    rho = s.evolve_ensemble(psi0, duration, dt, n_traj, QT.Unraveling.STANDARD, 1)
    t_all = time.perf_counter() - t0
    t0 = time.perf_counter()
    rho2 = s.evolve_ensemble(psi0, duration, dt, n_traj, QT.Unraveling.STANDARD, 1)
    t_rep = time.perf_counter() - t0
    same = bool(np.array_equal(rho, rho2))
    # the two parts of the call: the plan (H_eff, exp(-i H_eff h) on the GPU, operator upload) and
    # the trajectories themselves
    s._plan_cache = None
    t0 = time.perf_counter()
    plan = s._plan(QT.Unraveling.STANDARD, duration, dt)
    t_plan = time.perf_counter() - t0
    t0 = time.perf_counter()
    s._run(plan, psi0 / np.linalg.norm(psi0), n_traj, 1, rho_only=True)
    t_run = time.perf_counter() - t0
    # the GEMM alone, as launched by a step: n_traj trajectories x 1024 x 1024
    ms = QT.gemm_ms(n_traj, d, d, reps=20)
    flop = 8.0 * d * d * n_traj
    log(f"lindblad {n}q: {n_traj} trajectories x {steps} steps in {t_all:.3f} s (plan {t_plan:.3f} s, run "
        f"{t_run:.3f} s; repeated call {t_rep:.3f} s, same rho {same}); step GEMM {ms:.3f} ms "
        f"= {flop / ms / 1e9:.1f} TFLOP/s")
    return {"workload": f"{n}-qubit dense Lindblad trajectories (dim {d}), {n_traj} trajectories x {steps} steps, "
                        "2 channels, STANDARD unraveling, reference random streams, complex128",
            "s_per_ensemble": round(t_all, 4), "plan_s": round(t_plan, 4), "run_s": round(t_run, 4),
            "s_per_repeated_ensemble": round(t_rep, 4), "repeat_bitwise_equal": same,
            "trajectory_steps_per_s": round(n_traj * steps / t_run, 1),
            "gemm_ms_per_step": round(ms, 4), "gemm_TFLOPs": round(flop / ms / 1e9, 2),
            "gemm_frac_of_fp64_peak": round(flop / ms / 1e9 / 78.6, 3),
            "trace_rho": float(np.real(np.trace(rho)))}


def program_leg(V3, n=28, layers=20):
    """Only the program path on the 28-qubit layered circuit (for counter runs)."""
    from llsv import circuits as C, program as P

    ops = C.layered_rotations(1, n, layers)
    sv = V3.StateVector(n)
    n_jit = V3.program_jit_precompile(n, ops)
    sv.apply_program(ops)
    sv.sync()
    t0 = time.perf_counter()
    sv.apply_program(ops)
    sv.sync()
    t = time.perf_counter() - t0
    passes = len(P.plan(n, ops, sv.program_tile_bits())[1])
    del sv
    return {"passes": passes, "compiled_pass_kernels": n_jit, "s": round(t, 4), "ms_per_pass": round(t / passes * 1e3, 3)}


def q32_leg(V3, n=32, layers=12):
    """32 qubits (64 GiB): the Haar + CNOT-ladder sweep streamed gate by gate, and the layered
    circuit (12 layers) fused into <= 5-qubit dense blocks on FP64 MFMA."""
    import gate_fusion as GF
    from llsv import circuits as C, gates as G

    size = 1 << n
    sv = V3.StateVector(n)
    ops = C.haar_cnot_sweep(5, n, 1)
    alg = sum(32.0 * size if o[0] == "u" else 16.0 * size for o in ops)

    def sweep():
        for o in ops:
            if o[0] == "u":
                sv.apply_gate_cache_optimized(o[1], list(o[2]))
            else:
                sv.cnot(o[1], o[2])

    sweep()
    sv.sync()
    t0 = time.perf_counter()
    sweep()
    sv.sync()
    t_sweep = time.perf_counter() - t0
    log(f"32q sweep: {len(ops)} gates in {t_sweep:.3f} s")
    lay = C.layered_rotations(2, n, layers)
    fused = GF.fuse_dense(lay, 5)
    flops = sum(8.0 * (1 << len(o[1])) * size for o in fused if o[0] == "dense")
    blocks = [(o[0] == "dense", (list(o[1]), np.ascontiguousarray(o[2])) if o[0] == "dense" else (o[1], list(o[2])))
              for o in fused]
    sv.timer_start()
    t0 = time.perf_counter()
    for d, b in blocks:
        if d:
            sv.apply_fused(b[0], b[1])
        else:
            sv.apply_gate_cache_optimized(b[0], b[1])
    dev_ms = sv.timer_stop()
    t_f = time.perf_counter() - t0
    norm = sv.norm()
    log(f"32q fused: {len(fused)} blocks in {t_f:.3f} s, norm-1 {norm - 1:.1e}")
    # the same circuit through the fusion cost model (dense blocks only where they are predicted to pay)
    segs, t_pred = GF.plan_fusion(n, lay)
    sv.reset()
    V3.program_jit_precompile(n, [o for k, b in segs if k == "program" for o in b])
    GF.run_fusion_plan(sv, segs)
    sv.sync()
    t0 = time.perf_counter()
    GF.run_fusion_plan(sv, segs)
    sv.sync()
    t_auto = time.perf_counter() - t0
    log(f"32q fusion plan: {[k for k, _ in segs]} in {t_auto:.3f} s (predicted {t_pred:.3f} s)")
    del sv
    return {
        "workload": f"{n}-qubit complex128 (64 GiB): Haar + CNOT-ladder sweep ({len(ops)} gates) and a {layers}-layer "
                    f"RX/RY/RZ + CZ-brick circuit ({len(lay)} gates) fused into {len(fused)} dense <= 5-qubit blocks",
        "sweep": {"s": round(t_sweep, 4), "hbm_GBps": round(alg / t_sweep / 1e9, 1),
                  "pct_of_peak": round(100 * alg / t_sweep / 8e12, 1)},
        "fused_mfma": {"s": round(t_f, 4), "blocks": len(fused), "TFLOPs": round(flops / (dev_ms / 1e3) / 1e12, 2),
                       "mfma_fraction": round(flops / (dev_ms / 1e3) / 78.6e12, 4),
                       "hbm_GBps": round(32.0 * size * len(fused) / (dev_ms / 1e3) / 1e9, 1),
                       "gates_per_s": round(len(lay) / t_f, 1)},
        "fused_auto": {"s": round(t_auto, 4), "segments": [k for k, _ in segs], "predicted_s": round(t_pred, 4),
                       "gates_per_s": round(len(lay) / t_auto, 1)},
        "norm_minus_1": norm - 1,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the latency, layered, adjoint, trajectory, ZNE, Lindblad and 32-qubit legs")
    ap.add_argument("--leg", choices=["latency12", "layered28", "program28", "adjoint20", "trajectories14", "zne12",
                                      "zne16", "lindblad10", "q32"],
                    help="run only this leg (for profiling) and print its JSON")
    ap.add_argument("--probe", action="store_true",
                    help="launch each kernel a few times and exit (for rocprofv3 counter runs)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {a.gpus} is not supported (the state-vector "
              "path is single-device here; sharding is a design note, not a run)", file=sys.stderr)
        sys.exit(2)

    if not a.no_extra:
        import torch  # noqa: F401  (first, so torch and libllsv share one HIP runtime)
    import quantum_sim_v3 as V3
    from llsv import circuits as C

    if a.leg:
        import torch  # noqa: F401

        fn = {"latency12": lambda: latency_leg(V3), "layered28": lambda: layered_leg(V3),
              "program28": lambda: program_leg(V3),
              "adjoint20": adjoint_leg, "trajectories14": trajectory_leg, "zne12": zne_leg,
              "zne16": zne16_leg, "lindblad10": lindblad_leg,
              "q32": lambda: q32_leg(V3)}[a.leg]
        print(json.dumps({a.leg: fn()}), flush=True)
        return

    n = a.qubits
    size = 1 << n
    ops = C.haar_cnot_sweep(0, n, 1)
    one_q = [(op[1], list(op[2])) for op in ops if op[0] == "u"]
    cnots = [(op[1], op[2]) for op in ops if op[0] == "cnot"]
    bytes_1q = 32.0 * size
    bytes_cx = 16.0 * size
    bytes_step = bytes_1q * len(one_q) + bytes_cx * len(cnots)

    log(f"allocating {n}-qubit state ({16 * size / 2**30:.1f} GiB)")
    sv = V3.StateVector(n)

    if a.probe:
        # 2 copies (calibration: 32 B per amplitude), 4 one-qubit gates (targets 0, 7, n/2,
        # n-1: every kernel variant), 3 CNOTs and one program step, each a separate dispatch.
        other = V3.StateVector(n)
        for _ in range(2):
            other.copy_from(sv)
        other.sync()
        del other
        for q in (0, 7, n // 2, n - 1):
            sv.apply_gate_cache_optimized(q, one_q[q][1])
        sv.cnot(0, 1)
        sv.cnot(3, 4)
        sv.cnot(n - 2, n - 1)
        V3.program_jit_precompile(n, ops)
        sv.apply_program(ops)
        sv.sync()
        print(json.dumps({"probe": n, "norm": sv.norm()}), flush=True)
        return

    def step(marks=None):
        slot = 0
        for q, g in one_q:
            if marks is not None:
                sv.mark(slot)
            sv.apply_gate_cache_optimized(q, g)
            if marks is not None:
                sv.mark(slot + 1)
                slot += 2
        cx0 = slot
        for c, t in cnots:
            if marks is not None:
                sv.mark(slot)
            sv.cnot(c, t)
            if marks is not None:
                sv.mark(slot + 1)
                slot += 2
        if marks is not None:
            marks["1q"] += [sv.mark_elapsed(2 * i, 2 * i + 1) for i in range(len(one_q))]
            marks["cx"] += [sv.mark_elapsed(cx0 + 2 * i, cx0 + 2 * i + 1) for i in range(len(cnots))]

    for w in range(a.warmup):
        step()
        sv.sync()
        log(f"warmup {w + 1}/{a.warmup}")

    marks = {"1q": [], "cx": []}
    sv.sync()
    t0 = time.perf_counter()
    sv.timer_start()
    for k in range(a.steps):
        step(marks)
        if (k + 1) % max(1, a.steps // 5) == 0:
            log(f"step {k + 1}/{a.steps}")
    dev_ms = sv.timer_stop()
    sv.sync()
    wall = time.perf_counter() - t0
    norm = sv.norm()
    log(f"timed {a.steps} steps: wall {wall:.3f} s, device {dev_ms / 1e3:.3f} s, norm-1 {norm - 1:.2e}")

    # Same sweep through the program path: LDS-staged passes, several gates per HBM pass.
    from llsv import program as P

    tile_bits = sv.program_tile_bits()
    n_passes = len(P.plan(n, ops, tile_bits)[1])
    tj = time.perf_counter()
    n_jit = V3.program_jit_precompile(n, ops)
    t_jit = time.perf_counter() - tj
    sv.apply_program(ops)
    sv.sync()
    t1 = time.perf_counter()
    sv.timer_start()
    for _ in range(a.steps):
        sv.apply_program(ops)
    prog_dev_ms = sv.timer_stop()
    sv.sync()
    prog_wall = time.perf_counter() - t1
    log(f"program path: {n_passes} passes/step, wall {prog_wall:.3f} s")

    avg_1q = sum(marks["1q"]) / len(marks["1q"]) / 1e3
    avg_cx = sum(marks["cx"]) / len(marks["cx"]) / 1e3
    achieved = bytes_1q / avg_1q / 1e9
    value = bytes_step * a.steps / wall / 1e9
    gates = (len(one_q) + len(cnots)) * a.steps

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{n}-qubit complex128 sweep: Haar 2x2 on every qubit then CNOT ladder (q, q+1); "
                         f"{len(one_q) + len(cnots)} streamed gates per step, one pass each"),
            "n_qubits": n,
            "state_GiB": 16 * size / 2**30,
            "gates_per_step": len(one_q) + len(cnots),
            "bytes_per_step": bytes_step,
            "parallelism": "single GPU",
        },
        "pct_of_peak": round(100 * value / PEAK_HBM_GBPS, 1),
        "gates_per_s": round(gates / wall, 1),
        "device_GBps": round(bytes_step * a.steps / (dev_ms / 1e3) / 1e9, 1),
        "kernels": {
            "apply_1q": {"avg_ms": round(avg_1q * 1e3, 4), "GBps": round(achieved, 1),
                         "launches": len(marks["1q"])},
            "cnot": {"avg_ms": round(avg_cx * 1e3, 4), "GBps": round(bytes_cx / avg_cx / 1e9, 1),
                     "launches": len(marks["cx"])},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBPS, 4),
            "traffic": load_traffic(n),
            "kernel": "k_apply_1q (32 * 2^n bytes per launch)",
        },
        "norm_minus_1": norm - 1,
        "program_path": {
            "what": "same sweep via apply_program: gates grouped into register-tile HBM passes",
            "tile_bits": tile_bits,
            "passes_per_step": n_passes,
            "compiled_pass_kernels": n_jit,
            "compile_s": round(t_jit, 2),
            "ms_per_step": round(prog_wall / a.steps * 1e3, 3),
            "hbm_GBps": round(32.0 * size * n_passes * a.steps / prog_wall / 1e9, 1),
            "effective_GBps": round(bytes_step * a.steps / prog_wall / 1e9, 1),
            "gates_per_s": round(gates / prog_wall, 1),
            "device_ms_per_step": round(prog_dev_ms / a.steps, 3),
        },
    }
    del sv
    if not a.no_extra:
        def run_leg(name, fn):
            # a leg that raises is recorded in the line; the sweep measurement above stands
            try:
                return fn()
            except Exception as e:  # noqa: BLE001
                log(f"{name} failed: {e!r}")
                return {"error": repr(e)}

        line["latency12"] = run_leg("latency12", lambda: latency_leg(V3))
        line["layered28"] = run_leg("layered28", lambda: layered_leg(V3))
        line["adjoint20"] = run_leg("adjoint20", adjoint_leg)
        line["trajectories14"] = run_leg("trajectories14", trajectory_leg)
        line["zne12"] = run_leg("zne12", zne_leg)
        line["zne16"] = run_leg("zne16", zne16_leg)
        line["lindblad10"] = run_leg("lindblad10", lindblad_leg)
        line["q32"] = run_leg("q32", lambda: q32_leg(V3))
    if not a.no_cpu_baseline:
        log("cpu baseline (subprocess)")
        line["cpu_baseline"] = cpu_baseline(n)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
