"""Readout kernels on the GPU (needs an MI355X): Pauli-sum expectations against the oracle's
H|psi> and seeded sampling against an inverse-CDF restatement on the host.

Expectations: 1e-12.  Samples: the same index for every shot (the host CDF sums in another
order, so a shot could differ only if its uniform fell within ~1e-16 of a CDF step); no
zero-probability index is ever drawn."""
import sys

import numpy as np
import pytest

from helpers import random_ops
from oracle import adjoint as A
from oracle import oracle as O
from oracle import trajectory as T

pytestmark = pytest.mark.gpu
TAG = 0x53414D50


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def v3():
    import quantum_sim_v3

    return quantum_sim_v3


def random_terms(rng, n, count):
    terms = []
    for _ in range(count):
        x = int(rng.integers(0, 1 << n)) & int(rng.integers(0, 1 << n))
        z = int(rng.integers(0, 1 << n)) & int(rng.integers(0, 1 << n))
        terms.append((float(rng.standard_normal()), x, z))
    terms.append((0.7, 0, 0))  # identity
    return terms


@pytest.mark.parametrize("n", [3, 10, 16])
def test_expect_pauli_matches_oracle(n):
    rng = np.random.default_rng(40 + n)
    ops = random_ops(rng, n, 60, kinds=("h", "rx", "ry", "rz", "u", "cnot", "cz"))
    terms = random_terms(rng, n, 37)  # 38 terms: three passes of 16
    say(f"expect_pauli n={n}")
    sv = v3().StateVector(n)
    sv.apply_program(ops)
    got = sv.expect_pauli(terms)
    psi = O.run(n, ops)
    ref = float(np.vdot(psi, A.h_apply(psi, terms)).real)
    assert abs(got - ref) < 1e-12
    # single Z terms agree with the Z readout
    for q in (0, n - 1):
        assert abs(sv.expect_pauli([(1.0, 0, 1 << q)]) - sv.expect_z(q)) < 1e-13


def host_samples(p, seed, shots):
    cdf = np.cumsum(p)
    u = np.array([T.uniform(seed, k, TAG, 0) for k in range(shots)])
    return np.searchsorted(cdf, u * cdf[-1], side="right")


@pytest.mark.parametrize("n,shots", [(1, 64), (5, 500), (12, 4000), (18, 3000)])
def test_sample_matches_inverse_cdf(n, shots):
    rng = np.random.default_rng(70 + n)
    ops = random_ops(rng, n, 40, kinds=("h", "ry", "rz", "cnot"))
    say(f"sample n={n} shots={shots}")
    sv = v3().StateVector(n)
    sv.apply_program(ops)
    p = sv.probabilities()
    got = sv.sample(shots, 1234)
    ref = host_samples(p, 1234, shots)
    assert got.dtype == np.uint64 and got.shape == (shots,)
    assert np.array_equal(got.astype(np.int64), ref)
    assert np.all(p[got.astype(np.int64)] > 0.0)
    # the state is unchanged and the draw is reproducible
    assert np.array_equal(sv.probabilities(), p)
    assert np.array_equal(sv.sample(shots, 1234), got)


def test_sample_sparse_state_and_frequencies():
    # H on 4 of 14 qubits: 16 equally likely indices out of 16384
    n = 14
    sv = v3().StateVector(n)
    for q in (1, 4, 9, 13):
        sv.apply_gate_cache_optimized(q, O.mat_h())
    shots = 64000
    s = sv.sample(shots, 7).astype(np.int64)
    support = {sum(b << q for b, q in zip(bits, (1, 4, 9, 13))) for bits in np.ndindex(2, 2, 2, 2)}
    assert set(np.unique(s)) <= support
    counts = np.array([np.sum(s == k) for k in sorted(support)])
    # binomial: mean shots/16, sd sqrt(shots p (1 - p))
    sd = np.sqrt(shots / 16 * 15 / 16)
    assert np.all(np.abs(counts - shots / 16) < 5 * sd)
    assert sv.sample(0, 1).size == 0
