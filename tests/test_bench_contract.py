"""bench.py's contract pieces that need no GPU: the one-GPU refusal and the committed HBM-traffic
figure the roofline object carries."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_traffic_figure_is_the_copy_calibrated_one_qubit_launch():
    sys.path.insert(0, ROOT)
    import bench

    assert bench.PEAK_HBM_GBPS == 8000.0
    t = bench.load_traffic(30)
    # 32 * 2^30 algorithmic bytes per one-qubit launch; the counters read it within a few per mille
    assert t is not None and abs(t / (32.0 * 2**30) - 1.0) < 5e-3
    assert bench.load_traffic(7) is None


def test_cpu_baseline_pieces():
    # the CPU baseline's energy through the compiled reference backends equals the oracle's; its
    # host figures are sane; the whole script runs end to end on a small state
    import json

    import numpy as np
    import pytest

    sys.path.insert(0, ROOT)
    from llsv import circuits as C
    from oracle import adjoint as A, cpu_baseline as B, ref as R

    cores, _ = B.physical_cores()
    assert cores >= 1
    assert B.stream_copy(1 << 20) > 0.1
    if not R.available():
        pytest.skip("compiled reference not built")
    n, layers = 6, 2
    gates = []
    for layer in range(layers):
        for q in range(n):
            gates += [("ry", q, (layer * n + q) * 2), ("rz", q, (layer * n + q) * 2 + 1)]
        gates += [("cnot", q, q + 1) for q in range(n - 1)]
    terms = C.zz_x_hamiltonian(0, n)
    th = np.random.default_rng(3).uniform(0, 2 * np.pi, 2 * n * layers)
    for kind in (0, 1):
        assert abs(B.energy_on_backend(kind, n, gates, th, terms) - A.energy(n, gates, th, terms)) < 1e-13
    r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), "--qubits", "16",
                        "--no-legs"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["kind"] == "reference" and out["value"] > 0 and out["host"]["stream_copy_GBps"] > 0
    assert set(out["cnot_ms"]) == {"0", "5", "10"} and set(out["one_qubit_ms"]) == {"0", "5", "10", "15"}
