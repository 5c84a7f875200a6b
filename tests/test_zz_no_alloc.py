"""llsv_sample_noalloc through the C ABI against llsv_sample (needs an MI355X).

The allocation-free variant keeps its chunk sums, prefixes and sampled indices in a fixed device
array instead of a per-call buffer; it must return exactly the indices llsv_sample returns, on
two calls in a row with the second smaller, and make no allocation (the library's allocation hook
stays armed and unused across the call)."""
import ctypes
import sys

import numpy as np
import pytest

from llsv import _native as N

pytestmark = pytest.mark.gpu
P = ctypes.c_void_p
SCRATCH = 32 << 20


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def state_from(psi):
    h = P()
    a = np.ascontiguousarray(psi, dtype=np.complex128)
    N.check(N.lib().llsv_create_from_host(a.ctypes.data_as(P), a.size, ctypes.byref(h)))
    return h


def random_state(rng, n):
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi[rng.random(1 << n) < 0.3] = 0  # zero-probability indices must never be drawn
    return psi / np.linalg.norm(psi)


def sample(fn, s, shots, seed):
    out = np.full(shots, 0xDEAD, dtype=np.uint64)
    rc = fn(s, shots, seed, out.ctypes.data_as(P))
    return rc, out


def test_noalloc_equals_sample_two_calls():
    L = N.lib()
    rng = np.random.default_rng(2024)
    cases = [(20, 200_000, 11), (13, 5_000, 12)]  # the second call smaller in state and shots
    for n, shots, seed in cases:
        say(f"noalloc sample: n={n} shots={shots}")
        psi = random_state(rng, n)
        s = state_from(psi)
        try:
            rc, ref = sample(L.llsv_sample, s, shots, seed)
            assert rc == N.LLSV_OK
            L.llsv_test_fail_allocation(0)  # any allocation inside the call would fail it
            rc, got = sample(L.llsv_sample_noalloc, s, shots, seed)
            assert L.llsv_test_fail_allocation(-1) == 0, "llsv_sample_noalloc allocated"
            assert rc == N.LLSV_OK, L.llsv_last_error()
            assert np.array_equal(got, ref)
            assert np.all(np.abs(psi[got.astype(np.int64)]) > 0)
            # same call again on the scratch the previous call left behind
            rc, again = sample(L.llsv_sample_noalloc, s, shots, seed)
            assert rc == N.LLSV_OK and np.array_equal(again, ref)
        finally:
            L.llsv_destroy(s)


def test_noalloc_limits_and_errors():
    L = N.lib()
    say("noalloc sample: limits")
    s = state_from(random_state(np.random.default_rng(3), 10))
    try:
        out = np.zeros(4, dtype=np.uint64)
        assert L.llsv_sample_noalloc(s, SCRATCH // 8 + 1, 1, out.ctypes.data_as(P)) == N.LLSV_ERR_ARG
        assert b"scratch" in L.llsv_last_error()
        assert L.llsv_sample_noalloc(s, 4, 1, None) == N.LLSV_ERR_ARG
        assert L.llsv_sample_noalloc(None, 4, 1, out.ctypes.data_as(P)) == N.LLSV_ERR_HANDLE
        assert L.llsv_sample_noalloc(s, 0, 1, None) == N.LLSV_OK
        # the largest request that fits: 10 qubits = 1 chunk -> 16 + 64 bytes besides the shots
        big = (SCRATCH - 16 - 64) // 8
        rc, got = sample(L.llsv_sample_noalloc, s, big, 5)
        assert rc == N.LLSV_OK and got.max() < 1024
        rc, ref = sample(L.llsv_sample, s, big, 5)
        assert rc == N.LLSV_OK and np.array_equal(got, ref)
        rc, _ = sample(L.llsv_sample_noalloc, s, big + 1, 5)
        assert rc == N.LLSV_ERR_ARG
    finally:
        L.llsv_destroy(s)
    z = state_from(np.zeros(1 << 6))
    try:
        for fn in (L.llsv_sample, L.llsv_sample_noalloc):
            rc, _ = sample(fn, z, 8, 1)
            assert rc == N.LLSV_ERR_NUMERIC and b"no probability" in L.llsv_last_error()
    finally:
        L.llsv_destroy(z)
