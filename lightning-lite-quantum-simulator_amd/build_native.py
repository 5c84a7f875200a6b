"""Build the native parts of the package in-tree (no JIT cache, no install step).

  libllsv.so                       hipcc, gfx950: the HIP kernels + the C ABI of include/llsv.h
  quantum_sim_v3<EXT>, quantum_sim_v2<EXT>
                                   g++: the pybind11 StateVector surface, linked to libllsv.so

Run ``python build_native.py`` from anywhere; __graft_entry__.build() calls build().
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(os.path.dirname(PKG), "include")
OBJ = os.path.join(PKG, "_build")
ARCH = os.environ.get("LLSV_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["sv_gates.hip", "sv_reduce.hip", "sv_api.hip", "sv_program.hip", "sv_dense.hip", "sv_adjoint.hip", "sv_traj.hip", "sv_jit.hip", "sv_readout.hip", "sv_zne.hip", "sv_lindblad.hip"]
HIPCC_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-Wall", "-Wno-unused-result",
    # packed-FP32 instructions stay off: every kernel here is FP64 anyway.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INC,
]


def _run_parallel(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l]
        if p.returncode != 0:
            failed.append((c, "\n".join(lines)))
        elif lines:
            sys.stderr.write("\n".join(lines) + "\n")
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + o for c, o in failed)
        raise RuntimeError("native build failed:\n" + msg)


def build(verbose: bool = False) -> None:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append(["hipcc", *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    _run_parallel(cmds)
    lib = os.path.join(PKG, "libllsv.so")
    tmp = lib + ".tmp"
    subprocess.run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-lhiprtc"], check=True)
    os.replace(tmp, lib)

    ext = sysconfig.get_config_var("EXT_SUFFIX")
    import pybind11

    py_inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]
    cmds = []
    outs = []
    for mod, extra in (("quantum_sim_v3", []), ("quantum_sim_v2", ["-DLLSV_V2"])):
        out = os.path.join(PKG, mod + ext)
        outs.append(out)
        cmds.append(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden", *py_inc, "-I" + INC,
                     f"-DLLSV_PYMOD={mod}", *extra, os.path.join(CSRC, "py_statevector.cpp"),
                     "-L" + PKG, "-lllsv", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    _run_parallel(cmds)
    for out in outs:
        os.replace(out + ".tmp", out)
    if verbose:
        print("built", lib, *outs)


if __name__ == "__main__":
    build(verbose=True)
