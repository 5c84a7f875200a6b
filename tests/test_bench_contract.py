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
