"""Energies and gradients of parameterised circuits on the GPU (adjoint method), as a
torch.autograd.Function.

A circuit is a list of tuples: ('rx'|'ry'|'rz', q, p) rotates qubit q by params[p];
('u', q, g4) is a fixed 2x2; ('cnot', c, t); ('cz', a, b).  A Hamiltonian is a list of
(coeff, xmask, zmask) Pauli terms (X: x bit, Y: x and z bits, Z: z bit).

torch provides the device memory (parameter matrix, energies, gradients, workspace) and the
stream; all arithmetic runs in libllsv's HIP kernels (llsv_pc_energy / llsv_pc_adjoint).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native

PG = {"rx": 0, "ry": 1, "rz": 2, "u": 3, "cnot": 4, "cz": 5}


class PGate(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("q0", ctypes.c_int32), ("q1", ctypes.c_int32),
                ("param", ctypes.c_int32), ("m", ctypes.c_double * 8)]


class Pauli(ctypes.Structure):
    _fields_ = [("coeff", ctypes.c_double), ("xmask", ctypes.c_uint64), ("zmask", ctypes.c_uint64)]


def _bind():
    import torch  # noqa: F401  (torch's HIP runtime first, so libllsv shares it)

    return _native.lib()


class Circuit:
    """A parameterised circuit and Hamiltonian packed for the C ABI."""

    def __init__(self, n: int, gates, terms, n_params: int | None = None):
        self.n = int(n)
        self.gates = (PGate * max(1, len(gates)))()
        self.n_gates = len(gates)
        maxp = -1
        for i, g in enumerate(gates):
            k = g[0]
            o = self.gates[i]
            o.kind = PG[k]
            o.q0 = int(g[1])
            if k in ("rx", "ry", "rz"):
                o.param = int(g[2])
                maxp = max(maxp, o.param)
            elif k == "u":
                for j, v in enumerate(g[2]):
                    v = complex(v)
                    o.m[2 * j], o.m[2 * j + 1] = v.real, v.imag
            else:
                o.q1 = int(g[2])
        self.n_params = int(n_params) if n_params is not None else maxp + 1
        self.terms = (Pauli * len(terms))()
        for i, (c, x, z) in enumerate(terms):
            self.terms[i].coeff, self.terms[i].xmask, self.terms[i].zmask = float(c), int(x), int(z)
        self.n_terms = len(terms)

    def workspace_bytes(self, batch: int) -> int:
        return int(_native.lib().llsv_pc_workspace_bytes(self.n, batch, self.n_gates))

    def precompile(self) -> tuple[int, int]:
        """Compiles this circuit's tiled sweep kernels now, so that the first batched call does
        not wait for them (llsv_pc_precompile).  Returns (forward passes, backward passes), both
        0 when the sweeps run gate by gate (fewer than 20 qubits in the automatic compiler mode)."""
        fwd = ctypes.c_int(0)
        rc = _native.lib().llsv_pc_precompile(self.n, ctypes.cast(self.gates, ctypes.c_void_p), self.n_gates,
                                              ctypes.byref(fwd))
        if rc < 0:
            _native.check(rc)
        return int(fwd.value), int(rc) - int(fwd.value)


class _Ordered:
    """The stream to launch on, ordered with torch's current stream.  With one HIP runtime in the
    process that is torch's stream itself; otherwise (libllsv loaded before torch) torch's work is
    drained first, libllsv runs on its default stream and is drained after."""

    def __enter__(self):
        import torch

        self.shared = len(_native.hip_runtimes_loaded()) <= 1
        if self.shared:
            return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return ctypes.c_void_p(0)

    def __exit__(self, *exc):
        if not self.shared and exc[0] is None:
            _native.check(_native.lib().llsv_stream_sync(None))
        return False


def _workspace(circ: Circuit, batch: int):
    import torch

    return torch.empty(circ.workspace_bytes(batch), dtype=torch.uint8, device="cuda")


def energies(circ: Circuit, params):
    """E_b for a (batch, n_params) float64 CUDA tensor; returns a (batch,) tensor."""
    import torch

    params = params.detach().to(dtype=torch.float64).contiguous()
    B = params.shape[0]
    out = torch.empty(B, dtype=torch.float64, device=params.device)
    ws = _workspace(circ, B)
    L = _bind()
    with _Ordered() as stream:
        _native.check(L.llsv_pc_energy(circ.n, B, ctypes.cast(circ.gates, ctypes.c_void_p), circ.n_gates,
                                       circ.n_params, ctypes.c_void_p(params.data_ptr()),
                                       ctypes.cast(circ.terms, ctypes.c_void_p), circ.n_terms,
                                       ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                                       ws.numel(), stream))
    return out


def energies_and_grads(circ: Circuit, params):
    """(E (batch,), dE/dparams (batch, n_params)) by the adjoint method."""
    import torch

    params = params.detach().to(dtype=torch.float64).contiguous()
    B = params.shape[0]
    e = torch.empty(B, dtype=torch.float64, device=params.device)
    g = torch.empty(B, circ.n_params, dtype=torch.float64, device=params.device)
    ws = _workspace(circ, B)
    L = _bind()
    with _Ordered() as stream:
        _native.check(L.llsv_pc_adjoint(circ.n, B, ctypes.cast(circ.gates, ctypes.c_void_p), circ.n_gates,
                                        circ.n_params, ctypes.c_void_p(params.data_ptr()),
                                        ctypes.cast(circ.terms, ctypes.c_void_p), circ.n_terms,
                                        ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(g.data_ptr()),
                                        ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream))
    return e, g


def _autograd():
    import torch

    class AdjointEnergy(torch.autograd.Function):
        """E(params) for a batch of parameter vectors; the backward is the adjoint sweep
        (the state is un-computed with inverse gates, never stored per gate)."""

        @staticmethod
        def forward(ctx, params, circ):
            if params.requires_grad:
                e, g = energies_and_grads(circ, params)
                ctx.save_for_backward(g)
            else:
                e = energies(circ, params)
                ctx.save_for_backward(None)
            return e

        @staticmethod
        def backward(ctx, grad_e):
            (g,) = ctx.saved_tensors
            return grad_e.to(g.dtype)[:, None] * g, None

    return AdjointEnergy


_FN = None


def energy(params, circ: Circuit):
    """Differentiable E(params): params (batch, n_params) float64 on cuda -> (batch,)."""
    global _FN
    if _FN is None:
        _FN = _autograd()
    return _FN.apply(params, circ)


def param_shift_grads(circ: Circuit, params, chunk: int = 256):
    """Parameter-shift gradients on the GPU (the reference's rule, src/Algorithms.h:277-292:
    dE/dtheta_k = [E(theta + pi/2 e_k) - E(theta - pi/2 e_k)] / 2), evaluating the 2P shifted
    parameter vectors of each batch row as one batch of energies.  The baseline the adjoint
    method is measured against."""
    import torch

    params = params.detach().to(dtype=torch.float64)
    B, P = params.shape
    out = torch.empty(B, P, dtype=torch.float64, device=params.device)
    eye = torch.eye(P, dtype=torch.float64, device=params.device) * (math.pi / 2)
    for b in range(B):
        shifted = torch.cat([params[b][None, :] + eye, params[b][None, :] - eye], dim=0)
        es = torch.cat([energies(circ, shifted[i:i + chunk]) for i in range(0, 2 * P, chunk)])
        out[b] = 0.5 * (es[:P] - es[P:])
    return out


def hea_gates(n: int, layers: int):
    """Hardware-efficient ansatz: per layer RY(p) then RZ(p) on every qubit, then CNOT(q, q+1);
    parameter index (layer * n + q) * 2 + {0: ry, 1: rz} (the order of circuits.hea_ansatz)."""
    gates = []
    for layer in range(layers):
        for q in range(n):
            gates.append(("ry", q, (layer * n + q) * 2))
            gates.append(("rz", q, (layer * n + q) * 2 + 1))
        for q in range(n - 1):
            gates.append(("cnot", q, q + 1))
    return gates
