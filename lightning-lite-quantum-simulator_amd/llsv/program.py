"""Op tuples -> llsv_op arrays for the C ABI, and the host-side pass planner.

Op tuples: ('h', q) ('x', q) ('rx'|'ry'|'rz', q, t) ('u', q, g4) ('cnot', c, t) ('cz', a, b)
('mcz', [q...]) ('dense', [t0...], U).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native, gates

KIND = {"1q": 0, "cnot": 1, "cz": 2, "mcz": 3, "dense": 4}


def to_llsv_ops(ops):
    """Returns (ctypes array of LlsvOp, keepalive list).  Dense matrices stay referenced by
    the keepalive list for as long as the array is used."""
    arr = (_native.LlsvOp * max(1, len(ops)))()
    keep = []
    for i, op in enumerate(ops):
        o = arr[i]
        k = op[0]
        if k in gates.ONE_QUBIT:
            o.kind, o.nq, o.q[0] = 0, 1, int(op[1])
            for j, v in enumerate(gates.matrix(op)):
                o.m[2 * j], o.m[2 * j + 1] = v.real, v.imag
        elif k in ("cnot", "cz"):
            o.kind, o.nq = KIND[k], 2
            o.q[0], o.q[1] = int(op[1]), int(op[2])
        elif k in ("mcz", "dense"):
            qs = [int(q) for q in op[1]]
            if not 1 <= len(qs) <= 8:
                raise ValueError(f"{k}: needs 1..8 qubits")
            o.kind, o.nq = KIND[k], len(qs)
            for j, q in enumerate(qs):
                o.q[j] = q
            if k == "dense":
                m = np.ascontiguousarray(op[2], dtype=np.complex128)
                if m.shape != (1 << len(qs), 1 << len(qs)):
                    raise ValueError("dense: matrix must be 2^k x 2^k")
                keep.append(m)
                o.dense = m.ctypes.data
        else:
            raise ValueError(f"unknown op kind {k!r}")
    return arr, keep


def plan(n: int, ops, tile_bits: int):
    """(order, pass_end): the scheduler's execution order of the ops and the end index of
    every pass in it (host only)."""
    arr, keep = to_llsv_ops(ops)
    L = _native.lib()
    order = (ctypes.c_uint64 * max(1, len(ops)))()
    ends = (ctypes.c_uint64 * max(1, len(ops)))()
    npass = ctypes.c_uint64(0)
    _native.check(L.llsv_plan_passes(n, ctypes.cast(arr, ctypes.c_void_p), len(ops), tile_bits,
                                     ctypes.byref(npass), ctypes.cast(order, ctypes.c_void_p),
                                     ctypes.cast(ends, ctypes.c_void_p)))
    del keep
    p = npass.value
    return [int(order[i]) for i in range(len(ops))], [int(ends[i]) for i in range(p)]
