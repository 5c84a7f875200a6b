"""Compiled pass kernels (hipRTC) against the interpreter kernel and the oracle (needs an MI355X).

The generated kernel runs the same arithmetic helpers in the same order as the interpreter, so
the two agree exactly (np.array_equal; a signed zero may differ where the interpreter multiplies
by an exact 1 that the compiled kernel skips); both stay within 1e-12 of the oracle.
"""
import sys

import numpy as np
import pytest

from helpers import random_ops
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


def v3():
    import quantum_sim_v3

    return quantum_sim_v3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(n, ops):
    sv = v3().StateVector(n)
    sv.apply_program(ops)
    return sv.state()


@pytest.fixture
def restore_mode():
    V3 = v3()
    prev = V3.set_program_jit(1)
    V3.set_program_jit(prev)
    yield V3
    V3.set_program_jit(prev)


KINDS = [("u",), ("h", "rz"), ("h", "cz"), ("h", "mcz"), ("h", "cnot"), ("h", "x", "cnot"),
         ("u", "cz", "cnot", "mcz", "rz", "x")]


@pytest.mark.parametrize("n", [12, 13, 15])
@pytest.mark.parametrize("kinds", KINDS)
def test_compiled_equals_interpreter(restore_mode, n, kinds):
    V3 = restore_mode
    rng = np.random.default_rng(100 * n + len(kinds))
    ops = random_ops(rng, n, 90, kinds=kinds, max_k=5)
    say(f"jit n={n} kinds={kinds}")
    V3.set_program_jit(0)
    interp = run(n, ops)
    V3.set_program_jit(2)
    c0 = V3.program_jit_stats()
    comp = run(n, ops)
    c1 = V3.program_jit_stats()
    assert c1[2] == c0[2], "compilation failed"
    assert c1[1] > c0[1], "no compiled pass kernel ran"
    assert np.array_equal(comp, interp)
    assert np.max(np.abs(comp - O.run(n, ops))) <= TOL


def test_automatic_mode_compiles_in_background(restore_mode):
    # 20 qubits: the first call runs interpreted while the kernels compile; after the wait the
    # same program runs on compiled kernels and gives the same amplitudes
    V3 = restore_mode
    V3.set_program_jit(1)
    from llsv import circuits as C

    n = 20
    ops = C.layered_rotations(11, n, 3) + [("cnot", q, q + 1) for q in range(0, n - 1, 3)] + [("mcz", [1, 9, 18])]
    say("jit automatic n=20")
    first = run(n, ops)
    V3.program_jit_wait()
    c0 = V3.program_jit_stats()
    second = run(n, ops)
    c1 = V3.program_jit_stats()
    assert c1[1] > c0[1], "no compiled pass kernel ran after the wait"
    assert np.array_equal(first, second)
    rng = np.random.default_rng(3)
    probe = rng.integers(0, 1 << n, 64)
    ref = O.run(n, ops)
    assert np.max(np.abs(second - ref)) <= TOL
    assert np.max(np.abs(second[probe] - ref[probe])) <= TOL


def test_compiled_28q_inverse_roundtrip(restore_mode):
    # a layered circuit then its inverse on 28 qubits (4 GiB): back to |0...0> on compiled kernels
    V3 = restore_mode
    V3.set_program_jit(2)
    from llsv import circuits as C, gates as G

    n = 28
    fwd = C.layered_rotations(21, n, 4)
    inv = []
    for op in reversed(fwd):
        if op[0] == "cz":
            inv.append(op)
        else:
            m = np.asarray(G.matrix(op), dtype=complex).reshape(2, 2).conj().T
            inv.append(("u", op[1], m.reshape(4).tolist()))
    say("jit 28q roundtrip")
    sv = V3.StateVector(n)
    c0 = V3.program_jit_stats()
    sv.apply_program(fwd)
    sv.apply_program(inv)
    c1 = V3.program_jit_stats()
    assert c1[1] > c0[1]
    assert abs(sv.norm() - 1.0) < 1e-12
    for q in (0, 2, 13, 26, 27):
        assert abs(sv.expect_z(q) - 1.0) < 1e-11
