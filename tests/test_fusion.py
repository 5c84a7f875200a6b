"""The fusion pre-pass: GateFuser parity with python/gate_fusion.py and fuse_dense equivalence."""
import numpy as np
import pytest

import gate_fusion as GF
from helpers import random_ops
from oracle import oracle as O


def test_multiply_order():
    # python/gate_fusion.py:11: result = g2 @ g1 ("g1 first")
    h, s = GF.GATES["H"], GF.GATES["S"]
    got = np.array(GF.GateFuser.multiply_gates(h, s)).reshape(2, 2)
    assert np.allclose(got, np.array(s).reshape(2, 2) @ np.array(h).reshape(2, 2), atol=0)
    assert isinstance(GF.GateFuser.multiply_gates(h, s), list)


def test_fuse_circuit_benchmark_shape():
    # benchmarks/benchmark_fusion.py:14-27: 10 x 20 qubits x (H, X, S, T) = 800 -> 200
    circuit = [(q, GF.GATES[g]) for _ in range(10) for q in range(20) for g in "HXST"]
    fused = GF.GateFuser.fuse_circuit(circuit)
    assert len(circuit) == 800 and len(fused) == 200
    a = O.run(20 if False else 5, [])  # keep oracle import warm
    st1 = O.run(5, [("u", q, g) for q, g in circuit if q < 5])
    st2 = O.run(5, [("u", q, g) for q, g in fused if q < 5])
    assert np.max(np.abs(st1 - st2)) < 1e-13


def test_fuse_circuit_only_adjacent():
    c = [(0, GF.GATES["H"]), (1, GF.GATES["X"]), (0, GF.GATES["H"])]
    assert len(GF.GateFuser.fuse_circuit(c)) == 3
    assert GF.GateFuser.fuse_circuit([]) == []


def test_analyze_circuit(capsys):
    out = GF.GateFuser.analyze_circuit([(0, GF.GATES["H"]), (0, GF.GATES["X"])])
    assert len(out) == 1 and "Reduction: 50.0%" in capsys.readouterr().out


@pytest.mark.parametrize("max_k", [1, 2, 3, 4, 5])
def test_fuse_dense_equivalent(max_k):
    rng = np.random.default_rng(40 + max_k)
    n = 7
    ops = random_ops(rng, n, 120, max_k=min(max_k, 3))
    fused = GF.fuse_dense(ops, max_k)
    assert len(fused) <= len(ops)
    for op in fused:
        if op[0] == "dense":
            k = len(op[1])
            assert k <= max_k
            U = np.asarray(op[2])
            assert np.max(np.abs(U.conj().T @ U - np.eye(1 << k))) < 1e-12
    assert np.max(np.abs(O.run(n, fused) - O.run(n, ops))) < 1e-12
    assert np.max(np.abs(O.run_bruteforce(n, fused) - O.run(n, ops))) < 1e-12


def test_fuse_dense_layered_block_count():
    from llsv import circuits as C

    ops = C.layered_rotations(1, 10, 2)
    fused = GF.fuse_dense(ops, 5)
    assert len(fused) < len(ops) // 4
    assert np.max(np.abs(O.run(10, fused) - O.run(10, ops))) < 1e-12


# ---------------------------------------------------------------- cost model
def deep_few_qubits(seed, qs, count):
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(count):
        if rng.random() < 0.7:
            ops.append((("rx", "ry", "rz")[int(rng.integers(3))], int(rng.choice(qs)), float(rng.uniform(0, 6.28))))
        else:
            c, t = rng.choice(qs, 2, replace=False)
            ops.append(("cnot", int(c), int(t)))
    return ops


def _segments_as_ops(segs):
    out = []
    for kind, body in segs:
        out.extend(body)
    return out


def test_plan_fusion_choices_and_equivalence():
    import gate_fusion as GF
    from llsv import circuits as C
    from oracle import oracle as O

    # a deep circuit on five qubits of a large state: one dense block beats one FP64-heavy pass
    deep = deep_few_qubits(5, [3, 7, 11, 19, 25], 300)
    segs, t = GF.plan_fusion(28, deep)
    assert [k for k, _ in segs] == ["dense"] and len(segs[0][1]) == 1
    assert t < GF.predict_pass(28, deep)
    # a layered circuit: register passes (one read and write per pass, many gates each)
    lay = C.layered_rotations(1, 28, 20)
    segs, t = GF.plan_fusion(28, lay)
    assert [k for k, _ in segs] == ["program"] and len(segs[0][1]) == len(lay)
    assert t < GF.predict_blocks(28, GF.fuse_dense(lay, 5))
    # mixed, on 8 qubits where the oracle runs it: the segments apply the same circuit
    n = 8
    mixed = deep_few_qubits(6, [0, 2, 5], 400) + C.layered_rotations(2, n, 2) + deep_few_qubits(7, [1, 4, 6, 7], 300)
    segs, t = GF.plan_fusion(n, mixed)
    assert {k for k, _ in segs} == {"dense", "program"} and t <= GF.predict_program(n, mixed)
    got = O.run(n, _segments_as_ops(segs))
    assert np.max(np.abs(got - O.run(n, mixed))) < 1e-12
    # the predicted pass time: memory plus FP64 issue, same-qubit one-qubit runs merged
    assert GF.predict_pass(20, [("rx", 0, 0.1), ("ry", 0, 0.2), ("cz", 0, 1), ("rz", 3, 0.4)]) == \
        32.0 * 2**20 / GF.PASS_BPS + (GF._OPS_GENERAL + GF._OPS_PERM + GF._OPS_DIAG) * 2**20 / GF.FP64_INSTR_PER_S


def test_plan_fusion_chooses_width():
    import gate_fusion as GF
    from llsv import circuits as C
    from oracle import oracle as O

    # max_k=None: the plan of lowest predicted time over widths 2..5
    for n, ops in [(28, deep_few_qubits(5, [3, 7, 11, 19, 25], 300)), (28, C.layered_rotations(1, 28, 4)),
                   (9, deep_few_qubits(8, [0, 1, 2, 3], 80) + deep_few_qubits(9, [4, 5, 6, 7, 8], 80))]:
        segs, t = GF.plan_fusion(n, ops, None)
        widths = {k: GF.plan_fusion(n, ops, k)[1] for k in (2, 3, 4, 5)}
        assert t == min(widths.values())
    # a five-qubit register: only width 5 makes it one block, which the choice takes
    deep = deep_few_qubits(5, [3, 7, 11, 19, 25], 300)
    segs, t = GF.plan_fusion(28, deep, None)
    assert [k for k, _ in segs] == ["dense"] and len(segs[0][1][0][1]) == 5
    # the chosen plan applies the same circuit (oracle, 9 qubits)
    ops = deep_few_qubits(8, [0, 1, 2, 3], 80) + deep_few_qubits(9, [4, 5, 6, 7, 8], 80)
    segs, _ = GF.plan_fusion(9, ops, None)
    assert np.max(np.abs(O.run(9, _segments_as_ops(segs)) - O.run(9, ops))) < 1e-12
