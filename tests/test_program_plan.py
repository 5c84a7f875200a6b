"""The pass scheduler of llsv_apply_program, checked on the CPU through its host-only plan.

Executing the ops in the planned order (oracle) must equal executing them in circuit order:
the scheduler may only move an op past ops it shares no qubit with.
"""
import numpy as np
import pytest

from helpers import random_ops
from llsv import circuits as C
from llsv import program as P
from oracle import oracle as O


@pytest.mark.parametrize("n,tile_bits", [(4, 12), (8, 5), (10, 6), (12, 8), (14, 12)])
def test_planned_order_equals_circuit_order(n, tile_bits):
    rng = np.random.default_rng(n * 100 + tile_bits)
    ops = random_ops(rng, n, 300, max_k=min(5, tile_bits - 3 if tile_bits > 3 else 1))
    order, ends = P.plan(n, ops, tile_bits)
    assert sorted(order) == list(range(len(ops)))
    assert ends[-1] == len(ops) and all(a < b for a, b in zip([0] + ends, ends))
    ref = O.run(n, ops)
    got = O.run(n, [ops[i] for i in order])
    assert np.max(np.abs(got - ref)) < 1e-12


def test_single_tile_is_one_pass():
    ops = C.random_h_cnot(3, 12, 1000)
    order, ends = P.plan(12, ops, 12)
    assert ends == [1000] and order == list(range(1000))


def test_sweep_pass_count():
    # 30 qubits, Haar on every qubit + CNOT ladder: far fewer passes than gates
    ops = C.haar_cnot_sweep(0, 30, 1)
    order, ends = P.plan(30, ops, 12)
    assert len(ends) < len(ops) // 4
    ref_ops = C.haar_cnot_sweep(0, 12, 1)
    assert P.plan(12, ref_ops, 12)[1] == [len(ref_ops)]


def test_layered_pass_count():
    ops = C.layered_rotations(0, 28, 20)
    order, ends = P.plan(28, ops, 12)
    assert len(ends) <= 4 * 20


def test_plan_rejects_bad_ops():
    with pytest.raises(ValueError):
        P.plan(4, [("cnot", 1, 1)], 12)
    with pytest.raises(ValueError):
        P.plan(4, [("h", 4)], 12)
