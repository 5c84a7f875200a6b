"""The pass compiler of apply_program on the host: generated kernel sources compile with hipRTC
for gfx950 without a device, the cache reuses a kernel when only coefficients change, and the
mode switch validates its argument.  (The kernels' results are checked on the GPU in
test_gpu_program_jit.py.)
"""
import numpy as np
import pytest


def v3():
    import quantum_sim_v3

    return quantum_sim_v3


@pytest.fixture
def jit_always():
    V3 = v3()
    prev = V3.set_program_jit(2)
    yield V3
    V3.set_program_jit(prev)


def test_precompile_compiles_every_eligible_pass(jit_always):
    V3 = jit_always
    from llsv import circuits as C

    ops = C.layered_rotations(3, 14, 2)
    c0 = V3.program_jit_stats()
    k = V3.program_jit_precompile(14, ops)
    c1 = V3.program_jit_stats()
    assert k >= 1
    assert c1[2] == c0[2], "a generated kernel failed to compile"
    # new angles, same structure: no new compilation
    k2 = V3.program_jit_precompile(14, C.layered_rotations(4, 14, 2))
    assert k2 == k
    assert V3.program_jit_stats()[0] == c1[0]


def test_every_op_kind_compiles(jit_always):
    V3 = jit_always
    rng = np.random.default_rng(5)
    n = 13
    ops = [("h", q) for q in range(n)]
    ops += [("cnot", 0, 7), ("cnot", 9, 1), ("cnot", 12, 11), ("cz", 2, 10), ("mcz", [0, 5, 12]), ("mcz", [3]),
            ("rz", 4, 0.3), ("u", 6, [1, 0, 0, np.exp(0.4j)]), ("x", 8), ("ry", 12, 0.7)]
    ops += [("u", int(q), [complex(v) for v in rng.standard_normal(4)]) for q in rng.integers(0, n, 20)]
    c0 = V3.program_jit_stats()
    assert V3.program_jit_precompile(n, ops) >= 1
    assert V3.program_jit_stats()[2] == c0[2]


def test_mode_argument_checked():
    V3 = v3()
    with pytest.raises(ValueError):
        V3.set_program_jit(3)
    prev = V3.set_program_jit(1)
    assert V3.set_program_jit(prev) == 1


def test_small_or_dense_passes_are_not_compiled():
    V3 = v3()
    prev = V3.set_program_jit(1)  # automatic: only states of >= 20 qubits
    try:
        assert V3.program_jit_precompile(14, [("h", q) for q in range(14)]) == 0
        dense = [("dense", [0, 5, 9], np.eye(8, dtype=complex))] * 2 + [("h", 3)]
        assert V3.program_jit_precompile(24, dense) == 0
    finally:
        V3.set_program_jit(prev)


def test_adjoint_sweeps_compile_without_device():
    # the 20-qubit, 10-layer ansatz: 9 forward passes, and a backward sweep on the forward schedule
    # reversed (as many passes); its last pass holds 150 rotations, so its generated source carries
    # a 150-entry gradient-slot list -- every pass must compile
    import quantum_sim_v3 as V3
    from llsv import adjoint as LA, circuits as C

    before = V3.program_jit_stats()
    circ = LA.Circuit(20, LA.hea_gates(20, 10), C.zz_x_hamiltonian(0, 20))
    assert circ.precompile() == (9, 9)
    after = V3.program_jit_stats()
    assert after[2] == before[2], "a pass failed to compile"
