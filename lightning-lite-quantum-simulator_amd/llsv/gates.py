"""Gate matrices in the reference's conventions (src/Backend.h:3-8, :66-88).

RX/RY/RZ(t) = exp(-i t P / 2); one-qubit gates are row-major 4-lists [g00, g01, g10, g11].
"""
from __future__ import annotations

import math

R2 = 1.0 / math.sqrt(2.0)


def h():
    return [complex(R2), complex(R2), complex(R2), complex(-R2)]


def x():
    return [0j, 1 + 0j, 1 + 0j, 0j]


def rx(t):
    c, s = math.cos(0.5 * t), math.sin(0.5 * t)
    return [complex(c), complex(0, -s), complex(0, -s), complex(c)]


def ry(t):
    c, s = math.cos(0.5 * t), math.sin(0.5 * t)
    return [complex(c), complex(-s), complex(s), complex(c)]


def rz(t):
    return [complex(math.cos(-0.5 * t), math.sin(-0.5 * t)), 0j, 0j,
            complex(math.cos(0.5 * t), math.sin(0.5 * t))]


def matrix(op):
    """2x2 (as a 4-list) of a one-qubit op tuple."""
    k = op[0]
    if k == "h":
        return h()
    if k == "x":
        return x()
    if k == "rx":
        return rx(float(op[2]))
    if k == "ry":
        return ry(float(op[2]))
    if k == "rz":
        return rz(float(op[2]))
    if k == "u":
        return [complex(v) for v in op[2]]
    raise KeyError(k)


ONE_QUBIT = ("h", "x", "rx", "ry", "rz", "u")
