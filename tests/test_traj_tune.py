"""The trajectory tuner's cache file (llsv_traj_tune_load / _save, LLSV_TUNE_CACHE) on the host:
JSON written by the library reads back to the same entries, malformed files are refused without
changing the cache, and the environment variable loads a file on first use.  (That a plan then
uses a stored choice instead of tuning is checked on the GPU in test_gpu_trajectory.py.)"""
import json
import os
import subprocess
import sys

import pytest

from conftest import PKG, ROOT


def entry(maxn, minn, lg_items, lg_gates, threads, per_cu, fused):
    return {"maxn": maxn, "minn": minn, "lg_items": lg_items, "lg_gates": lg_gates, "threads": threads,
            "per_cu": per_cu, "fused": fused}


def write(path, entries, fmt="llsv-traj-tune-1"):
    with open(path, "w") as f:
        json.dump({"format": fmt, "entries": entries}, f)


def test_load_save_roundtrip(tmp_path):
    from llsv import trajectory as LT

    a = [entry(14, 14, 12, 8, 512, 2, 1), entry(16, 10, 11, 7, 1024, 1, 0), entry(13, 13, 12, 8, 256, 2, 2)]
    src = tmp_path / "a.json"
    write(src, a)
    before = LT.tuning_entries()
    LT.load_tuning(str(src))
    assert LT.tuning_entries() >= max(before, 3)
    out = tmp_path / "b.json"
    LT.save_tuning(str(out))
    saved = json.load(open(out))
    assert saved["format"] == "llsv-traj-tune-1"
    for e in a:
        assert e in saved["entries"]
    # a later file's entry for the same shape replaces the earlier one
    write(src, [entry(14, 14, 12, 8, 256, 4, 0)])
    LT.load_tuning(str(src))
    LT.save_tuning(str(out))
    got = [e for e in json.load(open(out))["entries"] if (e["maxn"], e["minn"], e["lg_items"], e["lg_gates"]) == (14, 14, 12, 8)]
    assert got == [entry(14, 14, 12, 8, 256, 4, 0)]


@pytest.mark.parametrize("bad", [
    {"format": "something-else", "entries": []},
    {"format": "llsv-traj-tune-1", "entries": [entry(14, 14, 12, 8, 300, 2, 1)]},   # threads not offered
    {"format": "llsv-traj-tune-1", "entries": [entry(14, 14, 12, 8, 1024, 4, 1)]},  # 4096 threads per CU
    {"format": "llsv-traj-tune-1", "entries": [entry(22, 14, 12, 8, 256, 1, 1)]},   # above 20 qubits
    {"format": "llsv-traj-tune-1", "entries": [entry(14, 14, 12, 8, 512, 2, 3)]},   # no such fused mode
    {"format": "llsv-traj-tune-1", "entries": [{"maxn": 14}]},                      # fields missing
])
def test_malformed_files_are_refused(tmp_path, bad):
    from llsv import trajectory as LT

    p = tmp_path / "bad.json"
    with open(p, "w") as f:
        json.dump(bad, f)
    before = LT.tuning_entries()
    with pytest.raises(ValueError):
        LT.load_tuning(str(p))
    assert LT.tuning_entries() == before
    with pytest.raises(ValueError):
        LT.load_tuning(str(tmp_path / "missing.json"))


def test_environment_file_is_read_on_first_use(tmp_path):
    p = tmp_path / "env.json"
    write(p, [entry(12, 12, 9, 7, 512, 2, 1), entry(10, 6, 6, 6, 256, 4, 1), entry(20, 20, 3, 9, 1024, 1, 0)])
    code = ("import sys; sys.path[:0] = [%r, %r]; from llsv import trajectory as LT; print(LT.tuning_entries())"
            % (ROOT, PKG))
    env = dict(os.environ, LLSV_TUNE_CACHE=str(p))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert int(r.stdout.strip().splitlines()[-1]) == 3
