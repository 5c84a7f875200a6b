"""CPU oracle for noisy trajectories and their density-matrix limit -- TEST INFRASTRUCTURE ONLY.

* philox4x32_10 / uniform: the counter-based generator of traj_oracle.c (Salmon et al., SC'11).
* run: one trajectory through traj_oracle.c (gate, then AD and PD on each touched qubit).
* density_matrix: the same noisy circuit on a 2^n x 2^n density matrix, every gate and Kraus
  operator embedded as a full operator with np.kron (a brute-force restatement of the block
  update of src/Noise.cpp:174-258, as its own test does at src/Noise.cpp:513-535).  The ensemble
  mean of <Z_q> over trajectories converges to diag(rho) expectations.
* density_matrix_blocks / density_matrix_z_blocks: the block update of KrausChannel::apply
  (src/Noise.cpp:174-258) restated in C (dm_oracle.c), every gate as the one-operator channel
  {U}; O(4^n) per operator, so it reaches the 14-qubit trajectory workload (a 4 GiB rho).
* density_matrix_z_fast: the same evolution with rho as a (2,)*2n tensor (np.tensordot per
  operator, CNOT as an index flip on ket and bra); pinned against both of the above.

Gates: ('u', q, g4) and ('cnot', c, t); rates gamma[q], lam[q].
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import oracle as O


class TGate(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("q0", ctypes.c_int32), ("q1", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("m", ctypes.c_double * 8)]


def _lib():
    L = O._lib()
    if not hasattr(L, "_traj_bound"):
        P = ctypes.c_void_p
        L.orc_philox4x32_10.argtypes = [P, P, P]
        L.orc_philox4x32_10.restype = None
        L.orc_traj_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_traj_uniform.restype = ctypes.c_double
        L.orc_traj_run.argtypes = [ctypes.c_int, P, ctypes.c_int, P, P, ctypes.c_uint64, ctypes.c_uint64, P, P, P]
        L.orc_traj_run.restype = ctypes.c_int
        L.orc_dm_apply_channel.argtypes = [ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int, P]
        L.orc_dm_apply_channel.restype = ctypes.c_int
        L.orc_dm_apply_channel_seq.argtypes = [ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int, P, P]
        L.orc_dm_apply_channel_seq.restype = ctypes.c_int
        L.orc_dm_init_zero.argtypes = [ctypes.c_int, P]
        L.orc_dm_init_zero.restype = None
        L.orc_dm_expect_z.argtypes = [ctypes.c_int, P, P, P]
        L.orc_dm_expect_z.restype = None
        L._traj_bound = True
    return L


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    _lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def uniform(seed, traj, gate, site):
    return float(_lib().orc_traj_uniform(seed, traj, gate, site))


def pack(gates):
    arr = (TGate * max(1, len(gates)))()
    for i, g in enumerate(gates):
        if g[0] == "cnot":
            arr[i].kind, arr[i].q0, arr[i].q1 = 1, int(g[1]), int(g[2])
        else:
            arr[i].kind, arr[i].q0 = 0, int(g[1])
            for j, v in enumerate(O.gate4(g)):
                arr[i].m[2 * j], arr[i].m[2 * j + 1] = complex(v).real, complex(v).imag
    return arr


def run(n, gates, gamma, lam, seed, traj, packed=None, want_state=False):
    """(z[n], jumps, hash, state or None) for one trajectory."""
    arr = packed if packed is not None else pack(gates)
    g = np.ascontiguousarray(gamma, dtype=np.float64)
    l = np.ascontiguousarray(lam, dtype=np.float64)
    z = np.zeros(n)
    h = ctypes.c_uint64(0)
    st = np.zeros(1 << n, dtype=np.complex128) if want_state else None
    jumps = _lib().orc_traj_run(n, arr, len(gates), g.ctypes.data_as(ctypes.c_void_p),
                                l.ctypes.data_as(ctypes.c_void_p), seed, traj, z.ctypes.data_as(ctypes.c_void_p),
                                ctypes.byref(h), st.ctypes.data_as(ctypes.c_void_p) if want_state else None)
    return z, jumps, h.value, st


def _embed(n, q, K):
    M = np.array([[1.0 + 0j]])
    for k in reversed(range(n)):
        M = np.kron(M, K if k == q else np.eye(2))
    return M


def density_matrix_z(n, gates, gamma, lam):
    """<Z_q> of the exact noisy evolution (rho <- U rho U^dag, then sum_k K rho K^dag per channel)."""
    rho = density_matrix_kron(n, gates, gamma, lam)
    N = 1 << n
    p = np.real(np.diag(rho))
    idx = np.arange(N)
    return np.array([p[((idx >> q) & 1) == 0].sum() - p[((idx >> q) & 1) == 1].sum() for q in range(n)])


def density_matrix_kron(n, gates, gamma, lam):
    """rho of the exact noisy evolution, every operator embedded with np.kron (n <= 7)."""
    N = 1 << n
    rho = np.zeros((N, N), dtype=np.complex128)
    rho[0, 0] = 1.0
    for g in gates:
        U = O.full_operator(n, g)
        rho = U @ rho @ U.conj().T
        qs = [g[1], g[2]] if g[0] == "cnot" else [g[1]]
        for q in qs:
            gm, lm = gamma[q], lam[q]
            for ks in ([np.diag([1, np.sqrt(1 - gm)]), np.array([[0, np.sqrt(gm)], [0, 0]])],
                       [np.diag([1, np.sqrt(1 - lm)]), np.diag([0, np.sqrt(lm)])]):
                E = [_embed(n, q, K.astype(np.complex128)) for K in ks]
                rho = sum(e @ rho @ e.conj().T for e in E)
    return rho


def _apply_rho(rho, n, q, K, Kd):
    """rho <- K rho K^dag for a 2x2 K on qubit q, rho as a (2,)*2n tensor (ket axes first,
    most significant qubit first): the per-block update B -> K B K^dag of src/Noise.cpp:174-258."""
    ax_ket = n - 1 - q
    ax_bra = 2 * n - 1 - q
    rho = np.moveaxis(np.tensordot(K, rho, axes=([1], [ax_ket])), 0, ax_ket)
    rho = np.moveaxis(np.tensordot(rho, Kd, axes=([ax_bra], [0])), -1, ax_bra)
    return rho


def _qubit_axis(n, q, side):
    """Tensor axis of qubit q on the ket (side 0) or bra (side 1) of a (2,)*2n rho."""
    return (n - 1 - q) + side * n


def _axis_after_index(axis, indexed):
    """Position of `axis` once axis `indexed` has been removed by an integer index."""
    return axis - (1 if axis > indexed else 0)


def _cnot_rho(rho, n, c, t):
    """CNOT = |0><0|_c (x) I + |1><1|_c (x) X_t on ket and bra: flip the target axis of the
    control = 1 slice, once per side."""
    for side in (0, 1):
        ac = _qubit_axis(n, c, side)
        at = _qubit_axis(n, t, side)
        idx = [slice(None)] * (2 * n)
        idx[ac] = 1
        sub = rho[tuple(idx)]
        rho[tuple(idx)] = np.flip(sub, axis=_axis_after_index(at, ac))
    return rho


def _kraus_sets(gamma, lam, q):
    return ([np.diag([1, np.sqrt(1 - gamma[q])]), np.array([[0, np.sqrt(gamma[q])], [0, 0]])],
            [np.diag([1, np.sqrt(1 - lam[q])]), np.diag([0, np.sqrt(lam[q])])])


def density_matrix_z_fast(n, gates, gamma, lam):
    """Same result as density_matrix_z, O(4^n) per operator (n up to ~12)."""
    rho = np.zeros((2,) * (2 * n), dtype=np.complex128)
    rho[(0,) * (2 * n)] = 1.0
    for g in gates:
        if g[0] == "cnot":
            rho = _cnot_rho(rho, n, g[1], g[2])
            qs = [g[1], g[2]]
        else:
            M = np.asarray(O.gate4(g), dtype=np.complex128).reshape(2, 2)
            rho = _apply_rho(rho, n, g[1], M, M.conj().T)
            qs = [g[1]]
        for q in qs:
            for ks in _kraus_sets(gamma, lam, q):
                rho = sum(_apply_rho(rho, n, q, K.astype(np.complex128), K.conj().T.astype(np.complex128)) for K in ks)
    p = np.real(np.diagonal(rho.reshape(1 << n, 1 << n)))
    idx = np.arange(1 << n)
    return np.array([p[((idx >> q) & 1) == 0].sum() - p[((idx >> q) & 1) == 1].sum() for q in range(n)])


# ------------------------------------------------------------ block update in C (dm_oracle.c)
_CNOT4 = np.zeros((4, 4), dtype=np.complex128)
for _t in range(4):
    # operator bit 0 <-> control, bit 1 <-> target: control = 1 flips the target bit
    _CNOT4[_t ^ 2 if _t & 1 else _t, _t] = 1.0


def dm_apply_channel(rho, n, targets, ops):
    """rho (dim x dim complex128, C order) <- sum_k K_k rho K_k^dag in place; operator bit j <->
    targets[j] (src/Noise.h:58-65)."""
    ops = np.ascontiguousarray(np.asarray(ops, dtype=np.complex128).reshape(len(ops), 1 << len(targets), -1))
    tg = (ctypes.c_int * len(targets))(*[int(t) for t in targets])
    assert rho.dtype == np.complex128 and rho.flags.c_contiguous and rho.shape == (1 << n, 1 << n)
    rc = _lib().orc_dm_apply_channel(n, rho.ctypes.data_as(ctypes.c_void_p), len(targets), tg, ops.shape[0],
                                     ops.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise ValueError(f"orc_dm_apply_channel failed ({rc})")
    return rho


def dm_apply_channel_seq(rho, n, targets, channels):
    """The channels (each a list of Kraus operators), all on `targets`, applied in order in one
    sweep over rho (bit-identical to one dm_apply_channel call per channel)."""
    d = 1 << len(targets)
    flat = [np.asarray(K, dtype=np.complex128).reshape(d, d) for ch in channels for K in ch]
    ops = np.ascontiguousarray(np.stack(flat))
    counts = (ctypes.c_int * len(channels))(*[len(ch) for ch in channels])
    tg = (ctypes.c_int * len(targets))(*[int(t) for t in targets])
    assert rho.dtype == np.complex128 and rho.flags.c_contiguous and rho.shape == (1 << n, 1 << n)
    rc = _lib().orc_dm_apply_channel_seq(n, rho.ctypes.data_as(ctypes.c_void_p), len(targets), tg, len(channels),
                                         counts, ops.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise ValueError(f"orc_dm_apply_channel_seq failed ({rc})")
    return rho


def density_matrix_blocks(n, gates, gamma, lam, rho=None, merged=True):
    """The noisy circuit on a density matrix through the block update (dm_oracle.c): every gate
    U as the channel {U}, then AD(gamma_q) and PD(lam_q) on each touched qubit.  merged: a
    one-qubit gate and its two channels (same target) run in one sweep, a CNOT's channels in one
    sweep per qubit -- the same arithmetic per block as one sweep per operator set.  Returns rho
    (dim x dim; 16 * 4^n bytes)."""
    if rho is None:
        rho = np.empty((1 << n, 1 << n), dtype=np.complex128)
        _lib().orc_dm_init_zero(n, rho.ctypes.data_as(ctypes.c_void_p))
    for g in gates:
        if g[0] == "cnot":
            dm_apply_channel(rho, n, [g[1], g[2]], [_CNOT4])
            qs, first = [g[1], g[2]], []
        else:
            qs, first = [g[1]], [[np.asarray(O.gate4(g), dtype=np.complex128).reshape(2, 2)]]
        for i, q in enumerate(qs):
            chans = (first if i == 0 else []) + list(_kraus_sets(gamma, lam, q))
            if merged:
                dm_apply_channel_seq(rho, n, [q], chans)
            else:
                for ch in chans:
                    dm_apply_channel(rho, n, [q], ch)
    return rho


def dm_expect_z(n, rho):
    """(<Z_q> for q < n, trace) of a dim x dim rho."""
    z = np.zeros(n)
    tr = ctypes.c_double(0.0)
    _lib().orc_dm_expect_z(n, rho.ctypes.data_as(ctypes.c_void_p), z.ctypes.data_as(ctypes.c_void_p),
                           ctypes.byref(tr))
    return z, tr.value


def density_matrix_z_blocks(n, gates, gamma, lam):
    rho = density_matrix_blocks(n, gates, gamma, lam)
    z, _ = dm_expect_z(n, rho)
    return z
