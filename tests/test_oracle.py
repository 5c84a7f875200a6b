"""The oracle pinned: golden vectors from the compiled reference, closed forms, brute force.

CPU only.  These tests are what make the oracle trustworthy as the checker of the HIP
path; each check is chosen so that a plausible slip (transposed gate, swapped
control/target, dropped term, wrong bit order) fails it -- see the planted-mistake tests.
"""
import math

import numpy as np
import pytest

from helpers import bits_equal, decode_h_cnot, load, random_ops
from llsv import circuits as C
from oracle import oracle as O
from oracle import ref as R

R2 = 1.0 / math.sqrt(2.0)


# ------------------------------------------------------------- golden vectors
def test_config1_generator_matches_fixture():
    f = load("config1_h_cnot_12q.npz")
    for s in f["seeds"]:
        ops = C.random_h_cnot(int(s), 12, 1000)
        assert decode_h_cnot(f["ops"][s]) == ops


@pytest.mark.parametrize("seed", range(10))
def test_config1_oracle_bit_exact(seed):
    # 12 qubits, 1000 H/CNOT ops: ll::ReferenceBackend output, bit for bit.
    f = load("config1_h_cnot_12q.npz")
    st = O.run(12, decode_h_cnot(f["ops"][seed]))
    assert bits_equal(st, f["states"][seed])


def test_v3_haar_oracle_bit_exact():
    # quantum_sim_v3.apply_gate_cache_optimized output (src/StateVectorOptimized.cpp:56-78).
    f = load("v3_haar_10q.npz")
    ops = [("u", int(t), list(g)) for t, g in zip(f["targets"], f["gates"])]
    assert bits_equal(O.run(10, ops), f["state"])
    # the mask-loop variant (apply_gate_standard) gives the same bits
    s = O.zero_state(10)
    for t, g in zip(f["targets"], f["gates"]):
        O.apply_1q(s, int(t), list(g), standard=True)
    assert bits_equal(s, f["state"])


def test_backend_cases_bit_exact():
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
            assert bits_equal(O.run(5, ops), f["cnot_pairs"][i])
            i += 1
    for q in range(5):
        ops = [("ry", k, 0.5 + 0.3 * k) for k in range(5)]
        ops += [("h", q), ("x", q), ("rx", q, 0.9), ("ry", q, -1.1), ("rz", q, 2.3)]
        assert bits_equal(O.run(5, ops), f["every_gate"][q])
    spec = [("h", q) for q in range(4)] + [("cnot", 0, 1), ("cnot", 2, 3), ("rz", 0, 0.3), ("rx", 1, 0.5),
                                          ("ry", 2, -0.2), ("mcz", [0, 1, 2, 3])]
    assert bits_equal(O.run(4, spec), f["spec_circuit"])


def test_vqe_ansatz_fixture():
    # ll::VQE::prepare: RY layer then CNOT ring (q, q+1 mod n), theta[l*n + q].
    f = load("backend_cases.npz")
    n, layers, th = 6, 3, f["vqe_theta"]
    ops = []
    for l in range(layers):
        ops += [("ry", q, float(th[l * n + q])) for q in range(n)]
        ops += [("cnot", q, (q + 1) % n) for q in range(n)]
    assert bits_equal(O.run(n, ops), f["vqe_state"])


# ------------------------------------------------------------- closed forms
def test_closed_forms():
    # tests/test_backends.cpp:147-203
    assert O.probabilities(O.run(4, [("x", 2)]))[4] > 1 - 1e-14
    p = O.probabilities(O.run(2, [("h", 0), ("cnot", 0, 1)]))
    assert abs(p[0] - 0.5) < 1e-14 and abs(p[3] - 0.5) < 1e-14 and p[1] < 1e-14 and p[2] < 1e-14
    assert O.probabilities(O.run(3, [("x", 2), ("cnot", 2, 0)]))[5] > 1 - 1e-14
    assert O.probabilities(O.run(3, [("x", 0), ("cnot", 0, 2)]))[5] > 1 - 1e-14
    assert O.probabilities(O.run(1, [("ry", 0, math.pi)]))[1] > 1 - 1e-14
    assert abs(O.run(1, [("rx", 0, math.pi)])[1] - (-1j)) < 1e-14
    s = O.run(1, [("h", 0), ("rz", 0, 0.7)])
    assert abs(s[0] - R2 * complex(math.cos(-0.35), math.sin(-0.35))) < 1e-14
    assert abs(s[1] - R2 * complex(math.cos(0.35), math.sin(0.35))) < 1e-14
    m = O.run(3, [("h", 0), ("h", 1), ("h", 2), ("mcz", [0, 1, 2])])
    assert abs(m[7] + R2 ** 3) < 1e-14 and all(abs(m[i] - R2 ** 3) < 1e-14 for i in range(7))
    z = O.run(2, [("h", 0), ("h", 1), ("mcz", [1])])
    assert abs(z[2] + 0.5) < 1e-14 and abs(z[0] - 0.5) < 1e-14
    # docs/BUILD_INSTRUCTIONS.md:171-187: H on qubit 0 of 2 qubits -> [.707, .707, 0, 0]
    assert np.allclose(O.run(2, [("h", 0)]), [R2, R2, 0, 0], atol=1e-15)


def test_readout_identities():
    rng = np.random.default_rng(3)
    st = O.run(7, random_ops(rng, 7, 60))
    p = O.probabilities(st)
    assert abs(p.sum() - 1) < 1e-13 and abs(O.norm2(st) - 1) < 1e-13
    for q in range(7):
        bit = (np.arange(128) >> q) & 1
        assert abs(O.expect_z(st, q) - (p[bit == 0].sum() - p[bit == 1].sum())) < 1e-13
    assert abs(O.inner(st, st) - 1) < 1e-13
    # <psi|Z0 Z1|psi> through the Pauli axpy equals the probability-weighted parity
    out = np.zeros_like(st)
    O.axpy_pauli(st, 0, 0b11, 1.0, out)
    par = 1 - 2 * (((np.arange(128) >> 0) ^ (np.arange(128) >> 1)) & 1)
    assert abs(O.inner(st, out).real - (p * par).sum()) < 1e-13
    # X_2 flips bit 2
    out = np.zeros_like(st)
    O.axpy_pauli(st, 0b100, 0, 2.0, out)
    assert np.allclose(out, 2.0 * st[np.arange(128) ^ 0b100], atol=0)


# ------------------------------------------------------------- brute force
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_bruteforce_random_circuits(n):
    rng = np.random.default_rng(100 + n)
    for _ in range(4):
        ops = random_ops(rng, n, 40)
        assert np.max(np.abs(O.run(n, ops) - O.run_bruteforce(n, ops))) < 1e-12


def test_norm_after_500_gates():
    rng = np.random.default_rng(2024)
    st = O.run(6, random_ops(rng, 6, 500))
    assert abs(O.norm2(st) - 1) < 1e-12


def test_all_pairs_and_targets_bruteforce():
    for c in range(5):
        for t in range(5):
            if c != t:
                ops = [("ry", q, 0.3 + 0.4 * q) for q in range(5)] + [("cnot", c, t), ("cz", t, c)]
                assert np.max(np.abs(O.run(5, ops) - O.run_bruteforce(5, ops))) < 1e-14


# ------------------------------------------------------------- planted mistakes
def test_fixture_catches_transposed_gate():
    f = load("v3_haar_10q.npz")
    wrong = [("u", int(t), np.asarray(g).reshape(2, 2).T.reshape(4).tolist()) for t, g in zip(f["targets"], f["gates"])]
    assert np.max(np.abs(O.run(10, wrong) - f["state"])) > 1e-3


def test_fixture_catches_swapped_cnot():
    f = load("config1_h_cnot_12q.npz")
    ops = [("cnot", o[2], o[1]) if o[0] == "cnot" else o for o in decode_h_cnot(f["ops"][0])]
    assert np.max(np.abs(O.run(12, ops) - f["states"][0])) > 1e-3


def test_bruteforce_catches_wrong_bit_order():
    from helpers import random_unitary
    U = random_unitary(np.random.default_rng(9), 4)
    ops = [("h", 0), ("ry", 1, 0.4), ("rx", 2, 1.1), ("dense", [2, 0], U)]
    flipped = ops[:3] + [("dense", [0, 2], U)]
    a = O.run_bruteforce(3, ops)
    assert np.max(np.abs(O.run(3, ops) - a)) < 1e-13
    assert np.max(np.abs(O.run(3, flipped) - a)) > 1e-3


# ------------------------------------------------------------- errors
def test_oracle_errors():
    s = O.zero_state(3)
    with pytest.raises(O.OracleError):
        O.apply_1q(s, 3, O.mat_h())
    with pytest.raises(O.OracleError):
        O.apply_cnot(s, 1, 1)
    with pytest.raises(O.OracleError):
        O.apply_mcz(s, [])
    with pytest.raises(O.OracleError):
        O.apply_mcz(s, [0, 0])
    with pytest.raises(O.OracleError):
        O.apply_dense(s, [0, 0], np.eye(4))


# ------------------------------------------------------------- live reference (this container only)
@pytest.mark.skipif(not R.available(), reason="oracle/_ref not built")
def test_oracle_vs_live_reference():
    rng = np.random.default_rng(77)
    for n in (3, 5, 8):
        ops = random_ops(rng, n, 150, kinds=("h", "x", "rx", "ry", "rz", "cnot", "cz", "mcz"))
        assert bits_equal(O.run(n, ops), R.RefBackend(n, 0).run(ops).state())
    # the reference v3 module on a borrowed array, in its own process (see oracle.ref.v3)
    import os
    import subprocess
    import sys
    import tempfile

    ops = random_ops(rng, 9, 80, kinds=("u",))
    with tempfile.TemporaryDirectory() as d:
        np.savez(os.path.join(d, "in.npz"), targets=np.array([o[1] for o in ops]),
                 gates=np.array([o[2] for o in ops], dtype=np.complex128))
        code = (
            "import sys, numpy as np; sys.path.insert(0, %r); from oracle import ref as R; "
            "f = np.load(%r); st = np.zeros(512, complex); st[0] = 1; sv = R.v3().StateVector(st)\n"
            "for t, g in zip(f['targets'], f['gates']): sv.apply_gate_cache_optimized(int(t), list(g))\n"
            "np.save(%r, st)" % (os.path.dirname(os.path.dirname(R.__file__)), os.path.join(d, "in.npz"),
                                 os.path.join(d, "out.npy")))
        subprocess.run([sys.executable, "-c", code], check=True)
        st = np.load(os.path.join(d, "out.npy"))
    assert bits_equal(O.run(9, ops), st)
