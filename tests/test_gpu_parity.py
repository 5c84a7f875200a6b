"""HIP path vs the oracle and the golden vectors (needs an MI355X).

Tolerance: max |delta amplitude| <= 1e-12 for every floating-point case (the HIP kernels
fuse multiply-adds, so they differ from the reference in the last bits); CNOT, CZ, MCZ and
X are permutations and sign flips and must match bit for bit.
"""
import math
import sys

import numpy as np
import pytest

from helpers import bits_equal, decode_h_cnot, load, random_ops, random_unitary
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


def qsv():
    import quantum_sim_v3

    return quantum_sim_v3


def gpu_run(n, ops, program=False):
    sv = qsv().StateVector(n)
    if program:
        sv.apply_program(ops)
    else:
        for op in ops:
            k = op[0]
            if k in ("h", "x", "rx", "ry", "rz", "u"):
                sv.apply_gate_cache_optimized(op[1], O.gate4(op))
            elif k == "cnot":
                sv.cnot(op[1], op[2])
            elif k == "cz":
                sv.cz(op[1], op[2])
            elif k == "mcz":
                sv.mcz(list(op[1]))
            elif k == "dense":
                sv.apply_fused(list(op[1]), np.asarray(op[2]))
    return sv.state()


def maxdiff(a, b):
    return float(np.max(np.abs(a - b)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------- golden vectors
@pytest.mark.parametrize("program", [False, True])
def test_config1_fixture(program):
    f = load("config1_h_cnot_12q.npz")
    for s in f["seeds"]:
        say(f"config1 seed {s} program={program}")
        st = gpu_run(12, decode_h_cnot(f["ops"][s]), program)
        assert maxdiff(st, f["states"][s]) <= TOL


def test_v3_haar_fixture_zero_copy():
    # borrowed-array mode: the user's array itself is updated after every call
    f = load("v3_haar_10q.npz")
    st = np.zeros(1024, dtype=np.complex128)
    st[0] = 1
    sv = qsv().StateVector(st)
    for t, g in zip(f["targets"], f["gates"]):
        sv.apply_gate_cache_optimized(int(t), list(g))
    assert maxdiff(st, f["state"]) <= TOL


def test_backend_cases_fixture():
    f = load("backend_cases.npz")
    i = 0
    for c in range(5):
        for t in range(5):
            if c == t:
                continue
            ops = []
            for q in range(5):
                ops += [("ry", q, 0.3 + 0.4 * q), ("rz", q, 0.1 * (q + 1))]
            ops.append(("cnot", c, t))
            assert maxdiff(gpu_run(5, ops, program=True), f["cnot_pairs"][i]) <= TOL
            i += 1
    for q in range(5):
        ops = [("ry", k, 0.5 + 0.3 * k) for k in range(5)]
        ops += [("h", q), ("x", q), ("rx", q, 0.9), ("ry", q, -1.1), ("rz", q, 2.3)]
        assert maxdiff(gpu_run(5, ops, program=True), f["every_gate"][q]) <= TOL


def test_closed_forms():
    r = 1 / math.sqrt(2)
    assert abs(gpu_run(1, [("rx", 0, math.pi)], True)[1] - (-1j)) < 1e-14
    m = gpu_run(3, [("h", 0), ("h", 1), ("h", 2), ("mcz", [0, 1, 2])], True)
    assert abs(m[7] + r ** 3) < 1e-14 and all(abs(m[i] - r ** 3) < 1e-14 for i in range(7))
    b = gpu_run(2, [("h", 0), ("cnot", 0, 1)], True)
    assert np.allclose(np.abs(b) ** 2, [0.5, 0, 0, 0.5], atol=1e-14)


# ------------------------------------------------------------- oracle, seeded random circuits
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11, 14])
def test_random_circuits_small(n):
    rng = np.random.default_rng(500 + n)
    ops = random_ops(rng, n, 200, max_k=min(n, 5))
    say(f"random n={n}")
    ref = O.run(n, ops)
    assert maxdiff(gpu_run(n, ops), ref) <= TOL
    assert maxdiff(gpu_run(n, ops, program=True), ref) <= TOL


@pytest.mark.parametrize("n", [4, 13, 17])
def test_apply_gate_standard(n):
    # the reference's mask-loop entry (src/StateVectorOptimized.cpp:81-97) runs the same kernels as
    # apply_gate_cache_optimized: identical bits, and the oracle's state to 1e-12
    rng = np.random.default_rng(700 + n)
    ops = [("u", int(q), random_unitary(rng, 2).reshape(-1).tolist()) for q in rng.integers(0, n, 60)]
    say(f"apply_gate_standard n={n}")
    a, b = qsv().StateVector(n), qsv().StateVector(n)
    for op in ops:
        a.apply_gate_standard(op[1], op[2])
        b.apply_gate_cache_optimized(op[1], op[2])
    sa = a.state()
    assert np.array_equal(sa, b.state())
    assert maxdiff(sa, O.run(n, ops)) <= TOL


@pytest.mark.parametrize("program", [False, True])
@pytest.mark.parametrize("n", [18, 21])
def test_every_target_and_pair(n, program):
    # every one-qubit gate kind on every target (dense, diagonal, phase-only paths),
    # CNOT/CZ over near and far pairs, dense blocks over low, high and mixed qubits
    rng = np.random.default_rng(n)
    ops = [("h", q) for q in range(n)]
    for q in range(n):
        ops += [("u", q, random_unitary(rng, 2).reshape(4).tolist()), ("rz", q, 0.3 * q + 0.1),
                ("u", q, [1, 0, 0, np.exp(0.7j * q)]), ("x", q)]
    for c in range(n):
        ops += [("cnot", c, (c + 1) % n), ("cnot", c, n - 1 - c) if c != n - 1 - c else ("x", c),
                ("cz", c, (c + 3) % n)]
    for targets in ([0], [n - 1], [0, 1], [n - 2, n - 1], [0, n - 1, 5], [1, 2, 3, 4], [n - 1, 0, 7, 3, 12]):
        ops.append(("dense", targets, random_unitary(rng, 1 << len(targets))))
    ops.append(("mcz", [0, 4, n - 1]))
    ops += [("cz", 0, 1), ("cz", 1, n - 1), ("mcz", [2, 9, n - 2]), ("cnot", 2, 0), ("cnot", 1, 5), ("cnot", 0, n - 1)]
    say(f"every target n={n} program={program}")
    ref = O.run(n, ops)
    assert maxdiff(gpu_run(n, ops, program=program), ref) <= TOL


@pytest.mark.parametrize("program", [False, True])
def test_permutations_bit_exact(program):
    rng = np.random.default_rng(8)
    n = 16
    start = (rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)).astype(np.complex128)
    ops = random_ops(rng, n, 300, kinds=("x", "cnot", "cz", "mcz"))
    sv = qsv().StateVector(start.copy())
    sv.defer_host_sync(True)
    if program:
        sv.apply_program(ops)
    else:
        for op in ops:
            if op[0] == "x":
                sv.apply_gate_cache_optimized(op[1], O.mat_x())
            elif op[0] == "cnot":
                sv.cnot(op[1], op[2])
            elif op[0] == "cz":
                sv.cz(op[1], op[2])
            else:
                sv.mcz(list(op[1]))
    got = sv.state()
    assert bits_equal(got, O.run(n, ops, psi=start))


# ------------------------------------------------------------- readout
def test_reductions():
    rng = np.random.default_rng(21)
    n = 20
    ops = random_ops(rng, n, 60, kinds=("u", "cnot", "ry"))
    sv = qsv().StateVector(n)
    sv.apply_program(ops)
    ref = O.run(n, ops)
    assert abs(sv.norm() - O.norm2(ref)) < 1e-12
    for q in (0, 7, 19):
        assert abs(sv.expect_z(q) - O.expect_z(ref, q)) < 1e-12
    assert np.max(np.abs(sv.probabilities() - O.probabilities(ref))) < 1e-13
    # fixed-shape reductions are bitwise reproducible
    a = [sv.norm() for _ in range(3)]
    assert a[0] == a[1] == a[2]


# ------------------------------------------------------------- surface semantics
def test_zero_copy_and_view_semantics():
    m = qsv()
    state = np.zeros(1 << 10, dtype=np.complex128)
    state[0] = 1
    sv = m.StateVector(state)
    sv.apply_gate_cache_optimized(0, [0j, 1 + 0j, 1 + 0j, 0j])
    assert state[1] == 1 and state[0] == 0  # benchmarks/test_zerocopy_proof.py:30-33
    state[1] = 0
    state[2] = 1  # host writes are seen by the next call
    sv.apply_gate_cache_optimized(1, [0j, 1 + 0j, 1 + 0j, 0j])
    assert state[0] == 1
    owned = m.StateVector(10)
    view = owned.get_numpy_view()
    assert view.shape == (1024,) and view.dtype == np.complex128 and view[0] == 1
    owned.apply_gate_cache_optimized(3, [0j, 1 + 0j, 1 + 0j, 0j])
    assert view[8] == 1 and view[0] == 0
    assert view.base.owner is owned
    del owned  # the view keeps the state alive (no double free as in the reference)
    assert view[8] == 1
    assert m.StateVector(4).get_num_qubits() == 4 and m.StateVector(4).get_size() == 16


def test_host_writes_seen_by_readouts():
    # NumPy writes to a borrowed array or a live view reach every readout, not only gate calls
    m = qsv()
    st = np.zeros(1 << 6, dtype=np.complex128)
    st[0] = 1
    sv = m.StateVector(st)
    st[0] = 0
    st[5] = 0.6
    st[9] = 0.8j
    assert abs(sv.norm() - 1.0) < 1e-15
    assert abs(sv.expect_z(0) + 1.0) < 1e-15 and abs(sv.expect_z(2) - (-0.36 + 0.64)) < 1e-15
    p = sv.probabilities()
    assert abs(p[5] - 0.36) < 1e-15 and abs(p[9] - 0.64) < 1e-15
    assert np.array_equal(sv.state(), st)
    assert abs(sv.expect_pauli([(1.0, 0, 8)]) - (0.36 - 0.64)) < 1e-15
    assert set(sv.sample(200, 3).tolist()) <= {5, 9}
    owned = m.StateVector(6)
    view = owned.get_numpy_view()
    view[0] = 0
    view[63] = 1
    assert abs(owned.expect_z(5) + 1.0) < 1e-15 and owned.sample(4, 1).tolist() == [63] * 4
    other = m.StateVector(6)
    other.copy_from(owned)  # the source's host writes travel too
    assert other.state()[63] == 1
    st[:] = 0
    st[1] = 1
    other.copy_from(sv)
    assert other.state()[1] == 1
    with pytest.raises(RuntimeError):  # a state with no probability cannot be sampled
        m.StateVector(np.zeros(8, dtype=np.complex128)).sample(2, 1)


def test_v2_surface():
    import quantum_sim_v2

    st = np.zeros(1 << 8, dtype=np.complex128)
    st[0] = 1
    sv = quantum_sim_v2.StateVector(st)
    sv.apply_gate(0, O.mat_h())
    sv.apply_gate(5, O.mat_x())
    assert maxdiff(st, O.run(8, [("h", 0), ("x", 5)])) <= TOL


def test_errors():
    m = qsv()
    sv = m.StateVector(3)
    with pytest.raises(ValueError):
        sv.apply_gate_cache_optimized(3, O.mat_h())
    with pytest.raises(ValueError):
        sv.apply_gate_cache_optimized(-1, O.mat_h())
    with pytest.raises(ValueError):
        sv.cnot(1, 1)
    with pytest.raises(ValueError):
        sv.mcz([0, 0])
    with pytest.raises(ValueError):
        sv.apply_fused([0, 1], np.eye(2))
    with pytest.raises(TypeError):
        sv.apply_gate_cache_optimized(0, (1, 0, 0, 1))  # the reference takes a list
    with pytest.raises(RuntimeError, match="1-dimensional"):
        m.StateVector(np.zeros((4, 4), dtype=np.complex128))
    with pytest.raises(RuntimeError, match="power of 2"):
        m.StateVector(np.zeros(12, dtype=np.complex128))
    with pytest.raises(ValueError):
        m.StateVector(0)


# ------------------------------------------------------------- large-n properties
def test_28q_inverse_roundtrip():
    # at the bench size: U then U^dagger returns the start state; norm stays 1
    n = 28
    rng = np.random.default_rng(28)
    ops = []
    for q in (0, 1, 5, 13, 26, 27):
        ops.append(("u", q, random_unitary(rng, 2).reshape(4).tolist()))
    ops += [("cnot", 0, 27), ("cnot", 27, 3), ("cz", 2, 26), ("dense", [0, 27, 9], random_unitary(rng, 8))]
    inv = []
    for op in reversed(ops):
        if op[0] == "u":
            inv.append(("u", op[1], np.asarray(op[2]).reshape(2, 2).conj().T.reshape(4).tolist()))
        elif op[0] == "dense":
            inv.append(("dense", op[1], np.asarray(op[2]).conj().T))
        else:
            inv.append(op)
    say("28q roundtrip")
    sv = qsv().StateVector(n)
    sv.apply_gate_cache_optimized(4, O.mat_h())
    sv.apply_program(ops)
    assert abs(sv.norm() - 1) < 1e-12
    sv.apply_program(inv)
    assert abs(sv.expect_z(4)) < 1e-12 and abs(sv.expect_z(27) - 1) < 1e-12 and abs(sv.norm() - 1) < 1e-12


# ------------------------------------------------------------- fused dense blocks (FP64 MFMA for k = 4, 5)
@pytest.mark.parametrize("n", [12, 16, 20])
@pytest.mark.parametrize("max_k", [4, 5])
def test_fused_blocks_mfma(n, max_k):
    import gate_fusion as GF
    from llsv import circuits as C

    ops = C.layered_rotations(n, n, 3)
    fused = GF.fuse_dense(ops, max_k)
    assert any(op[0] == "dense" and len(op[1]) >= 4 for op in fused)
    say(f"fused n={n} max_k={max_k}: {len(ops)} ops -> {len(fused)} blocks")
    ref = O.run(n, ops)
    assert maxdiff(gpu_run(n, fused), ref) <= TOL
    assert maxdiff(gpu_run(n, fused, program=True), ref) <= TOL


def test_fusion_plan_runs_the_circuit():
    # the cost model's mixed plan (a dense block where a deep stretch sits on few qubits, register
    # passes elsewhere), executed segment by segment, is the circuit
    import gate_fusion as GF
    from llsv import circuits as C
    from test_fusion import deep_few_qubits

    n = 16
    mixed = deep_few_qubits(6, [0, 2, 5], 400) + C.layered_rotations(2, n, 2) + deep_few_qubits(7, [1, 4, 6, 7], 300)
    segs, _ = GF.plan_fusion(n, mixed)
    assert {k for k, _ in segs} == {"dense", "program"}
    say("fusion plan n=16")
    sv = qsv().StateVector(n)
    GF.run_fusion_plan(sv, segs)
    assert maxdiff(sv.state(), O.run(n, mixed)) <= TOL


@pytest.mark.parametrize("targets", [[0, 1, 2, 3, 4], [4, 3, 2, 1, 0], [19, 2, 11, 0, 7], [15, 16, 17, 18, 19],
                                     [5, 9, 13, 17], [0, 19, 1, 18]])
def test_dense_mfma_target_orders(targets):
    # operator bit j <-> targets[j], low and high targets, sorted and unsorted
    n = 20
    rng = np.random.default_rng(sum(targets))
    start = (rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n))
    start /= np.linalg.norm(start)
    U = random_unitary(rng, 1 << len(targets))
    sv = qsv().StateVector(start.astype(np.complex128))
    sv.defer_host_sync(True)
    sv.apply_fused(targets, U)
    assert maxdiff(sv.state(), O.run(n, [("dense", targets, U)], psi=start)) <= TOL


def test_32q_roundtrip():
    # 64 GiB state: index math past 2^31 pairs, every kernel family, U then U^dagger
    n = 32
    rng = np.random.default_rng(32)
    ops = [("u", q, random_unitary(rng, 2).reshape(4).tolist()) for q in (0, 3, 17, 30, 31)]
    ops += [("cnot", 31, 0), ("cnot", 1, 31), ("cz", 0, 31), ("dense", [31, 2, 20, 0, 9], random_unitary(rng, 32))]
    inv = []
    for op in reversed(ops):
        if op[0] == "u":
            inv.append(("u", op[1], np.asarray(op[2]).reshape(2, 2).conj().T.reshape(4).tolist()))
        elif op[0] == "dense":
            inv.append(("dense", op[1], np.asarray(op[2]).conj().T))
        else:
            inv.append(op)
    say("32q roundtrip")
    sv = qsv().StateVector(n)
    sv.apply_gate_cache_optimized(6, O.mat_h())
    gpu_ops = ops + inv
    for op in gpu_ops[:5]:
        sv.apply_gate_cache_optimized(op[1], op[2])
    sv.apply_program(gpu_ops[5:])
    assert abs(sv.norm() - 1) < 1e-12
    assert abs(sv.expect_z(6)) < 1e-12 and abs(sv.expect_z(31) - 1) < 1e-12 and abs(sv.expect_z(0) - 1) < 1e-12
