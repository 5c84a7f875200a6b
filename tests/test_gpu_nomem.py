"""Out-of-memory paths through the C ABI (needs an MI355X).

llsv_test_fail_allocation makes one of the library's device or pinned-host allocations fail, as
if memory had run out.  Each entry point that allocates must then return LLSV_ERR_NOMEM, leave
no half-built object behind, and work normally on the next call.  The hook returns its previous
setting, so every case also checks that the armed failure was used exactly when the call reported
LLSV_ERR_NOMEM (buffers the library keeps between calls may need no new allocation)."""
import ctypes
import sys

import numpy as np
import pytest

from llsv import _native as N

pytestmark = pytest.mark.gpu
P = ctypes.c_void_p


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def disarm():
    yield
    N.lib().llsv_test_fail_allocation(-1)


def armed_call(after, fn):
    """Arms the hook, runs fn() -> rc, disarms; returns (rc, failure_used)."""
    L = N.lib()
    L.llsv_test_fail_allocation(after)
    rc = fn()
    used = L.llsv_test_fail_allocation(-1) == -1
    if rc == N.LLSV_ERR_NOMEM:
        assert used, "LLSV_ERR_NOMEM without the injected failure"
        assert L.llsv_last_error()
    return rc, used


def new_state(n):
    h = P()
    N.check(N.lib().llsv_create(n, ctypes.byref(h)))
    return h


def test_create_each_allocation():
    L = N.lib()
    say("nomem: llsv_create")
    # amplitudes, partials, result word, parameter area, pinned result: each one failing
    for after in range(5):
        h = P()
        rc, used = armed_call(after, lambda: L.llsv_create(10, ctypes.byref(h)))
        assert rc == N.LLSV_ERR_NOMEM and used and not h.value, after
    h = P()
    amps = np.zeros(1 << 10, dtype=np.complex128)
    amps[3] = 1
    rc, used = armed_call(0, lambda: L.llsv_create_from_host(amps.ctypes.data_as(P), amps.size, ctypes.byref(h)))
    assert rc == N.LLSV_ERR_NOMEM and not h.value
    s = new_state(10)
    norm = ctypes.c_double()
    N.check(L.llsv_norm2(s, ctypes.byref(norm)))
    assert abs(norm.value - 1) < 1e-15
    L.llsv_destroy(s)


def test_state_buffers_allocated_on_first_use():
    L = N.lib()
    say("nomem: dense staging, probabilities, program buffers")
    s = new_state(12)
    try:
        rng = np.random.default_rng(1)
        z = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
        U = np.ascontiguousarray(np.linalg.qr(z)[0])
        tg = (ctypes.c_int * 2)(3, 7)
        rc, _ = armed_call(0, lambda: L.llsv_apply_dense(s, 2, tg, U.ctypes.data_as(P)))
        assert rc == N.LLSV_ERR_NOMEM
        N.check(L.llsv_apply_dense(s, 2, tg, U.ctypes.data_as(P)))
        probs = np.zeros(1 << 12)
        rc, _ = armed_call(0, lambda: L.llsv_probabilities(s, probs.ctypes.data_as(P), probs.size))
        assert rc == N.LLSV_ERR_NOMEM
        N.check(L.llsv_probabilities(s, probs.ctypes.data_as(P), probs.size))
        assert abs(probs.sum() - 1) < 1e-13
        ops = (N.LlsvOp * 2)()
        ops[0].kind, ops[0].nq, ops[0].q[0] = 0, 1, 5
        ops[0].m[0] = ops[0].m[6] = 1.0
        ops[1].kind, ops[1].nq, ops[1].q[0], ops[1].q[1] = 1, 2, 5, 6
        for after in (0, 1):  # device program buffer, then its pinned staging
            rc, used = armed_call(after, lambda: L.llsv_apply_program(s, ops, 2))
            assert rc == N.LLSV_ERR_NOMEM and used, after
        N.check(L.llsv_apply_program(s, ops, 2))
        N.check(L.llsv_sync(s))
        N.check(L.llsv_probabilities(s, probs.ctypes.data_as(P), probs.size))
        assert abs(probs.sum() - 1) < 1e-13
    finally:
        L.llsv_destroy(s)


def test_readout_workspaces():
    L = N.lib()
    say("nomem: expect_pauli, sample")
    s = new_state(8)
    try:
        t = np.zeros(1, dtype=[("c", "f8"), ("x", "u8"), ("z", "u8")])
        t[0] = (1.0, 0, 1)
        total = ctypes.c_double()
        rc, used = armed_call(0, lambda: L.llsv_expect_pauli(s, t.ctypes.data_as(P), 1, None, ctypes.byref(total)))
        assert rc == N.LLSV_ERR_NOMEM and used
        N.check(L.llsv_expect_pauli(s, t.ctypes.data_as(P), 1, None, ctypes.byref(total)))
        assert abs(total.value - 1) < 1e-15
        out = np.zeros(16, dtype=np.uint64)
        rc, used = armed_call(0, lambda: L.llsv_sample(s, 16, 5, out.ctypes.data_as(P)))
        assert rc == N.LLSV_ERR_NOMEM and used
        N.check(L.llsv_sample(s, 16, 5, out.ctypes.data_as(P)))
        assert np.all(out == 0)
    finally:
        L.llsv_destroy(s)


def test_trajectory_workspaces():
    from llsv import circuits as C, trajectory as LT

    say("nomem: trajectories")
    n = 6
    gates, gamma, lam = C.noisy_brick(2, n, 2)
    L = N.lib()
    L.llsv_test_fail_allocation(0)
    with pytest.raises(MemoryError):
        LT.run_trajectories(n, gates, gamma, lam, seed=1, n_traj=8)
    assert L.llsv_test_fail_allocation(-1) == -1
    z, j, h = LT.run_trajectories(n, gates, gamma, lam, seed=1, n_traj=8)
    for after in range(5):  # device buffers, then the four pinned result buffers of a plan
        L.llsv_test_fail_allocation(after)
        with pytest.raises(MemoryError):
            LT.TrajectoryPlan.from_circuit(n, gates, gamma, lam, 8, tune=-1)
        assert L.llsv_test_fail_allocation(-1) == -1
    plan = LT.TrajectoryPlan.from_circuit(n, gates, gamma, lam, 8, tune=-1)
    z2, j2, h2 = plan.run(1)
    plan.close()
    assert np.array_equal(j, j2) and np.array_equal(h, h2) and np.max(np.abs(z - z2)) < 1e-15


def test_zne_and_lindblad_workspaces():
    import quantum_sim_mitigation as M
    import quantum_sim_trajectory as QT

    L = N.lib()
    say("nomem: noisy run, Lindblad plan")
    z = P()
    N.check(L.llsv_noisy_create(0.01, 3, 1, ctypes.byref(z)))
    try:
        gates = ctypes.cast(M.pack_gates([("h", 0), ("cnot", 0, 1), ("rx", 2, 0.3)]), P)
        vals = np.zeros(4096)
        w = np.ones(8)
        for after in range(4):
            rc, used = armed_call(after, lambda: L.llsv_noisy_run(z, 3, gates, 3, 3, 4096, w.ctypes.data_as(P),
                                                                  vals.ctypes.data_as(P), None))
            assert rc in (N.LLSV_OK, N.LLSV_ERR_NOMEM)
            assert (rc == N.LLSV_ERR_NOMEM) == used
        N.check(L.llsv_noisy_run(z, 3, gates, 3, 3, 4096, w.ctypes.data_as(P), vals.ctypes.data_as(P), None))
        assert np.allclose(vals, 1.0, atol=1e-12)
    finally:
        L.llsv_noisy_destroy(z)
    s = QT.TrajectorySolver(2)
    s.set_hamiltonian(np.diag([0.0, 1.0, 1.0, 2.0]).astype(np.complex128))
    s.add_lindblad(np.kron(np.eye(2), [[0, 1], [0, 0]]).astype(np.complex128), 0.5)
    psi0 = np.zeros(4, dtype=np.complex128)
    psi0[3] = 1
    L.llsv_test_fail_allocation(0)
    with pytest.raises(MemoryError):
        s.evolve_ensemble(psi0, 0.2, 0.01, 64, QT.Unraveling.STANDARD, 1)
    assert L.llsv_test_fail_allocation(-1) == -1
    rho = s.evolve_ensemble(psi0, 0.2, 0.01, 64, QT.Unraveling.STANDARD, 1)
    assert abs(np.trace(rho) - 1) < 1e-12
