"""Batched trajectories on the GPU vs the CPU trajectory oracle (needs an MI355X).

Per trajectory the jump decisions must be identical (same Philox streams: equal jump counts and
decision hashes) and the final <Z_q> within 1e-12; the ensemble must match the density matrix
within 4.5 standard errors."""
import json
import os
import sys

import numpy as np
import pytest

from llsv import circuits as C
from oracle import oracle as O
from oracle import trajectory as T

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def as_u(gates):
    return [g if g[0] == "cnot" else ("u", g[1], O.gate4(g)) for g in gates]


@pytest.mark.parametrize("n,lo,hi,count", [(8, 0.05, 0.3, 64), (14, 1e-3, 1e-2, 48)])
def test_trajectories_match_cpu_bit_for_bit(n, lo, hi, count):
    from llsv import trajectory as LT

    gates, _, _ = C.noisy_brick(n, n, 8)
    rng = np.random.default_rng(n)
    gamma = rng.uniform(lo, hi, n)
    lam = gamma / 2
    say(f"trajectories n={n} count={count}")
    z, j, h, st = LT.run_trajectories(n, gates, gamma, lam, seed=99, n_traj=count, want_state=True)
    packed = T.pack(as_u(gates))
    assert j.sum() > 0
    for i in range(count):
        zc, jc, hc, sc = T.run(n, as_u(gates), gamma, lam, 99, i, packed, want_state=(i == 0))
        assert j[i] == jc and h[i] == hc, f"trajectory {i}: decisions differ"
        assert np.max(np.abs(z[i, :n] - zc)) < 1e-12
        if i == 0:
            assert np.max(np.abs(st - sc)) < 1e-12


def test_ensemble_vs_density_matrix():
    from llsv import trajectory as LT

    n = 6
    gates, _, _ = C.noisy_brick(5, n, 3)
    gamma = np.random.default_rng(1).uniform(0.05, 0.2, n)
    lam = gamma / 2
    z, _, _ = LT.run_trajectories(n, gates, gamma, lam, seed=3, n_traj=8192)
    zdm = T.density_matrix_z(n, as_u(gates), gamma, lam)
    se = z[:, :n].std(0) / np.sqrt(len(z)) + 1e-12
    assert np.all(np.abs(z[:, :n].mean(0) - zdm) < 4.5 * se)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("case", [0, 1])
def test_ensemble_matches_density_matrix_fixture(case):
    # the trajectory config at full size: 4096 trajectories of the 14-qubit brick circuit (and its
    # 10-qubit sibling), ensemble <Z_q> within 3 standard errors of the exact density matrix
    # (tests/golden/dm_brick14.json, written by the block update of src/Noise.cpp:174-258)
    from llsv import trajectory as LT

    with open(os.path.join(GOLDEN, "dm_brick14.json")) as f:
        fx = json.load(f)["cases"][case]
    n, count = fx["n_qubits"], 4096
    gates, gamma, lam = C.noisy_brick(fx["circuit_seed"], n, fx["layers"])
    assert len(gates) == fx["n_gates"] and np.array_equal(gamma, np.array(fx["gamma"]))
    say(f"ensemble {count} x {n} q vs density matrix")
    z, j, _ = LT.run_trajectories(n, gates, gamma, lam, seed=1, n_traj=count)
    zt = z[:, :n]
    se = zt.std(0, ddof=1) / np.sqrt(count)
    dev = np.abs(zt.mean(0) - np.array(fx["z"])) / se
    say(f"max deviation {dev.max():.2f} sigma, jumps {int(j.sum())}")
    assert j.sum() > 0 and np.all(se > 0)
    assert np.all(dev < 3.0), dev


def test_mixed_batch_one_launch():
    from llsv import trajectory as LT

    batch = C.mixed_batch(4, count=96, n_min=10, n_max=16)
    # stronger noise so jumps happen
    batch = [(n, g, gm * 20, lm * 20, t) for (n, g, gm, lm, t) in batch]
    say("mixed batch 96")
    z, j, h = LT.run_batch(batch, seed=17)
    assert j.sum() > 0
    for i in range(0, 96, 7):
        n, g, gm, lm, t = batch[i]
        zc, jc, hc, _ = T.run(n, as_u(g), gm, lm, 17, t)
        assert j[i] == jc and h[i] == hc
        assert np.max(np.abs(z[i, :n] - zc)) < 1e-12


def test_bad_arguments():
    from llsv import trajectory as LT

    with pytest.raises(ValueError):
        LT.run_trajectories(3, [("h", 3)], [0.1] * 3, [0.1] * 3, seed=1, n_traj=2)
    with pytest.raises(ValueError):
        LT.run_trajectories(3, [("h", 0)], [1.5] * 3, [0.1] * 3, seed=1, n_traj=2)


# ------------------------------------------------------------- autotuned plans replayed as hipGraphs
CONFIGS = [(t, p, f) for t in (256, 512, 1024) for p in (1, 2, 4) for f in (0, 1, 2)
           if t * p <= 2048 and (f < 2 or p <= 2)]


@pytest.mark.parametrize("threads,per_cu,fused", CONFIGS)
def test_every_launch_configuration_matches_cpu(threads, per_cu, fused):
    # each configuration the autotuner may pick: same decisions as the CPU oracle, <Z> within 1e-12
    from llsv import trajectory as LT

    n, count = 9, 24
    gates, _, _ = C.noisy_brick(21, n, 5)
    rng = np.random.default_rng(77)
    gamma = rng.uniform(0.02, 0.2, n)
    lam = gamma / 2
    say(f"plan config {threads}x{per_cu} fused={fused}")
    LT.set_default_config(threads, per_cu, fused)
    try:
        plan = LT.TrajectoryPlan.from_circuit(n, gates, gamma, lam, count, tune=-1)
        t, w, f, _ = plan.config()
        assert (t, f) == (threads, int(fused)) and 1 <= w <= count
        z, j, h = plan.run(5)
        plan.close()
    finally:
        LT.set_default_config(256, 4, True)
    packed = T.pack(as_u(gates))
    assert j.sum() > 0
    for i in range(count):
        zc, jc, hc, _ = T.run(n, as_u(gates), gamma, lam, 5, i, packed)
        assert j[i] == jc and h[i] == hc, f"trajectory {i}: decisions differ"
        assert np.max(np.abs(z[i, :n] - zc)) < 1e-12


def test_tuned_plan_replays_graph():
    # autotuned plan: replaying the graph gives the one-shot results for each seed, and the same
    # seed twice gives identical results
    from llsv import trajectory as LT

    n, count = 12, 512
    gates, _, _ = C.noisy_brick(4, n, 6)
    gamma = np.random.default_rng(12).uniform(1e-3, 2e-2, n)
    lam = gamma / 2
    say("tuned plan n=12 x512")
    plan = LT.TrajectoryPlan.from_circuit(n, gates, gamma, lam, count, tune=True)
    threads, workgroups, fused, tune_ms = plan.config()
    assert threads in (256, 512, 1024) and 1 <= workgroups <= count and tune_ms > 0
    for seed in (1, 2):
        z, j, h = plan.run(seed)
        z2, j2, h2 = plan.run(seed)
        assert np.array_equal(z, z2) and np.array_equal(j, j2) and np.array_equal(h, h2)
        zr, jr, hr = LT.run_trajectories(n, gates, gamma, lam, seed=seed, n_traj=count)
        assert np.array_equal(j, jr) and np.array_equal(h, hr)
        assert np.max(np.abs(z - zr)) < 1e-12
    # a second plan of the same shape reuses the tuned choice without timing again
    again = LT.TrajectoryPlan.from_circuit(n, gates, gamma, lam, count, tune=True)
    assert again.config()[3] == 0.0 and again.config()[:3] == (threads, workgroups, fused)
    plan.close()
    again.close()


def test_tuned_plan_mixed_batch():
    from llsv import trajectory as LT

    batch = C.mixed_batch(6, count=96, n_min=6, n_max=11, layers=4)
    say("tuned plan mixed batch")
    plan = LT.TrajectoryPlan(batch, tune=True)
    z, j, h = plan.run(17)
    zr, jr, hr = LT.run_batch(batch, seed=17)
    assert np.array_equal(j, jr) and np.array_equal(h, hr)
    assert np.max(np.abs(z - zr)) < 1e-12
    plan.close()


_PLAN_CHILD = """
import sys
sys.path[:0] = [{root!r}, {pkg!r}]
import numpy as np
from llsv import circuits as C, trajectory as LT
n, count = 10, 384
gates, _, _ = C.noisy_brick(9, n, 5)
gamma = np.random.default_rng(9).uniform(1e-3, 2e-2, n)
plan = LT.TrajectoryPlan.from_circuit(n, gates, gamma, gamma / 2, count, tune=True)
t, w, f, ms = plan.config()
z, j, h = plan.run(5)
zr, jr, hr = LT.run_trajectories(n, gates, gamma, gamma / 2, seed=5, n_traj=count)
assert np.array_equal(j, jr) and np.array_equal(h, hr) and np.max(np.abs(z - zr)) < 1e-12
print(t, w, int(f), ms)
"""


def _child(env_extra):
    import subprocess

    from conftest import PKG, ROOT

    env = dict(os.environ, **env_extra)
    r = subprocess.run([sys.executable, "-c", _PLAN_CHILD.format(root=ROOT, pkg=PKG)], env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    t, w, f, ms = r.stdout.split()[-4:]
    return int(t), int(w), int(f), float(ms)


def test_tuning_cache_file_and_forced_config(tmp_path):
    # a process that tunes writes LLSV_TUNE_CACHE; the next process reads it and does not time
    # again; LLSV_TRAJ_CONFIG forces a configuration instead of tuning (results stay the same)
    cache = tmp_path / "tune.json"
    say("tuning cache: first process tunes")
    t1, w1, f1, ms1 = _child({"LLSV_TUNE_CACHE": str(cache)})
    assert ms1 > 0 and cache.exists()
    assert json.load(open(cache))["format"] == "llsv-traj-tune-1"
    say("tuning cache: second process reads the file")
    t2, w2, f2, ms2 = _child({"LLSV_TUNE_CACHE": str(cache)})
    assert ms2 == 0.0 and (t2, w2, f2) == (t1, w1, f1)
    say("forced configuration")
    t3, _, f3, ms3 = _child({"LLSV_TRAJ_CONFIG": "1024,1,0"})
    assert ms3 == 0.0 and (t3, f3) == (1024, 0)
