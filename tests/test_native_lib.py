"""The C-ABI library and the pybind11 surface load and export what include/llsv.h declares.

CPU only: nothing here launches a kernel.
"""
import os
import re

import pytest

from conftest import PKG, ROOT


def header_functions():
    text = open(os.path.join(ROOT, "include", "llsv.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(llsv_\w+)\s*\(", text)))


def test_library_built():
    assert os.path.exists(os.path.join(PKG, "libllsv.so")), "run build_native.py"


def test_exports_every_declared_symbol():
    from llsv import _native

    lib = _native.lib()
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), f"{name} declared in include/llsv.h but not exported"
    # the ctypes table covers the whole header, so no entry point is bound untyped
    assert sorted(_native.SIGNATURES) == names


def test_version_and_error_string_without_gpu():
    from llsv import _native

    lib = _native.lib()
    assert lib.llsv_version().startswith(b"llsv")
    assert isinstance(lib.llsv_last_error(), bytes)
    # null-handle calls fail cleanly without touching a device
    assert lib.llsv_num_qubits(None) == -1
    assert lib.llsv_size(None) == 0
    assert lib.llsv_sync(None) == _native.LLSV_ERR_HANDLE
    assert b"null state handle" in lib.llsv_last_error()
    assert lib.llsv_destroy(None) == 0


def test_op_struct_layout():
    import ctypes

    from llsv import _native

    # int32 kind, int32 nq, int32 q[8], double m[8], const double* dense
    assert ctypes.sizeof(_native.LlsvOp) == 4 + 4 + 32 + 64 + 8


@pytest.mark.parametrize("mod,methods", [
    ("quantum_sim_v3", ["apply_gate_cache_optimized", "apply_gate_standard", "get_numpy_view", "get_num_qubits",
                        "get_size", "cnot", "cz", "apply_fused", "apply_program", "sync"]),
    ("quantum_sim_v2", ["apply_gate", "get_numpy_view", "get_num_qubits", "get_size", "cnot", "cz"]),
])
def test_pybind_surface(mod, methods):
    import importlib

    m = importlib.import_module(mod)
    assert os.path.dirname(m.__file__) == PKG, "the in-tree module must be the one imported"
    for name in methods:
        assert hasattr(m.StateVector, name), name


def test_no_oracle_in_product():
    # the product never imports or links the checker
    for dirpath, _, files in os.walk(PKG):
        if "_build" in dirpath:
            continue
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.sub(r"#.*|//.*", "", text), f"{f} refers to the oracle"
