"""Loader for the compiled reference in oracle/_ref/ -- TEST INFRASTRUCTURE ONLY.

``make -C oracle ref`` compiles the reference's own sources (where they lie under
the reference checkout) into oracle/_ref/:
  * quantum_sim_v3*.so  -- src/StateVectorOptimized.cpp as-is (the v3 pybind11 module)
  * quantum_sim_v2*.so  -- src/bindings_v2.cpp as-is
  * libref_backend.so   -- oracle/ref_backend_shim.cpp over src/Backend.h,
                           src/StateVectorBackend.h and src/Algorithms.h
Used to pin the oracle restatement and, in bench.py, as the CPU baseline.
Everything here returns None when oracle/_ref/ was not built.
"""
from __future__ import annotations

import ctypes
import glob
import importlib.machinery
import importlib.util
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")


def _load_ext(name: str):
    hits = glob.glob(os.path.join(REF_DIR, name + ".*.so"))
    if not hits:
        return None
    loader = importlib.machinery.ExtensionFileLoader(name, hits[0])
    spec = importlib.util.spec_from_file_location(name, hits[0], loader=loader)
    # Loading a single-phase extension registers it in sys.modules under its own name,
    # which would shadow the product module of the same name: restore the entry.
    prev = sys.modules.get(name)
    try:
        mod = importlib.util.module_from_spec(spec)
        loader.exec_module(mod)
    finally:
        if prev is None:
            sys.modules.pop(name, None)
        else:
            sys.modules[name] = prev
    return mod


_cache: dict = {}


def v3():
    """The reference quantum_sim_v3 module, or None.

    Load it only in a process that does not also import the product's quantum_sim_v3:
    CPython keeps one single-phase extension module per name, so whichever of the two
    is loaded first is returned for both (run it in a subprocess instead)."""
    if "v3" not in _cache:
        _cache["v3"] = _load_ext("quantum_sim_v3")
    return _cache["v3"]


def v2():
    if "v2" not in _cache:
        _cache["v2"] = _load_ext("quantum_sim_v2")
    return _cache["v2"]


def _backend_lib():
    if "lib" not in _cache:
        path = os.path.join(REF_DIR, "libref_backend.so")
        if not os.path.exists(path):
            _cache["lib"] = None
        else:
            lib = ctypes.CDLL(path)
            P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
            lib.refb_create.argtypes = [I, I]
            lib.refb_create.restype = P
            lib.refb_destroy.argtypes = [P]
            lib.refb_destroy.restype = None
            for nm in ("refb_h", "refb_x"):
                getattr(lib, nm).argtypes = [P, I]
            for nm in ("refb_rx", "refb_ry", "refb_rz"):
                getattr(lib, nm).argtypes = [P, I, D]
            lib.refb_cnot.argtypes = [P, I, I]
            lib.refb_mcz.argtypes = [P, P, I]
            lib.refb_state.argtypes = [P, P]
            lib.refb_probabilities.argtypes = [P, P]
            lib.refb_n_qubits.argtypes = [P]
            lib.refb_reset.argtypes = [P, I]
            lib.refb_vqe_prepare.argtypes = [P, I, I, P]
            _cache["lib"] = lib
    return _cache["lib"]


def available() -> bool:
    """True when oracle/_ref/ holds the compiled reference (nothing is loaded)."""
    return os.path.exists(os.path.join(REF_DIR, "libref_backend.so")) and bool(
        glob.glob(os.path.join(REF_DIR, "quantum_sim_v3.*.so")))


class RefBackend:
    """ll::ReferenceBackend (kind=0) or ll::StateVectorBackend (kind=1), compiled from the reference."""

    def __init__(self, n: int, kind: int = 0):
        self.lib = _backend_lib()
        if self.lib is None:
            raise RuntimeError("oracle/_ref not built")
        self.h_ = self.lib.refb_create(kind, n)
        if not self.h_:
            raise ValueError("bad qubit count")
        self.n = n

    def __del__(self):
        if getattr(self, "h_", None):
            self.lib.refb_destroy(self.h_)
            self.h_ = None

    def _rc(self, rc):
        if rc == -1:
            raise ValueError("invalid_argument")
        if rc != 0:
            raise RuntimeError("reference error")

    def h(self, q):
        self._rc(self.lib.refb_h(self.h_, q))

    def x(self, q):
        self._rc(self.lib.refb_x(self.h_, q))

    def rx(self, q, t):
        self._rc(self.lib.refb_rx(self.h_, q, t))

    def ry(self, q, t):
        self._rc(self.lib.refb_ry(self.h_, q, t))

    def rz(self, q, t):
        self._rc(self.lib.refb_rz(self.h_, q, t))

    def cnot(self, c, t):
        self._rc(self.lib.refb_cnot(self.h_, c, t))

    def mcz(self, qs):
        a = np.ascontiguousarray(qs, dtype=np.int32)
        self._rc(self.lib.refb_mcz(self.h_, a.ctypes.data_as(ctypes.c_void_p), int(a.size)))

    def vqe_prepare(self, layers, theta):
        t = np.ascontiguousarray(theta, dtype=np.float64)
        self._rc(self.lib.refb_vqe_prepare(self.h_, self.n, layers, t.ctypes.data_as(ctypes.c_void_p)))

    def state(self) -> np.ndarray:
        out = np.empty(1 << self.n, dtype=np.complex128)
        self._rc(self.lib.refb_state(self.h_, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def probabilities(self) -> np.ndarray:
        out = np.empty(1 << self.n, dtype=np.float64)
        self._rc(self.lib.refb_probabilities(self.h_, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def run(self, ops):
        for op in ops:
            k = op[0]
            if k == "h":
                self.h(op[1])
            elif k == "x":
                self.x(op[1])
            elif k == "rx":
                self.rx(op[1], op[2])
            elif k == "ry":
                self.ry(op[1], op[2])
            elif k == "rz":
                self.rz(op[1], op[2])
            elif k == "cnot":
                self.cnot(op[1], op[2])
            elif k == "cz":
                self.mcz([op[1], op[2]])
            elif k == "mcz":
                self.mcz(op[1])
            else:
                raise KeyError(k)
        return self
