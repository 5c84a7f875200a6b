"""Gate-fusion pre-pass.

Keeps the reference's ``GateFuser`` interface (python/gate_fusion.py:3-51) so its scripts
run unchanged:
  * ``multiply_gates(g1, g2)`` -> row-major 4-list of ``g2 @ g1`` ("g1 first, then g2",
    python/gate_fusion.py:7-12);
  * ``fuse_circuit(gates)`` merges runs of *adjacent* gates on the same target
    (python/gate_fusion.py:15-37);
  * ``analyze_circuit(gates)`` prints the counts and returns the fused list
    (python/gate_fusion.py:40-51);
  * ``GATES`` holds X, H, S, T (python/gate_fusion.py:54-60).

New here: ``fuse_dense(ops, max_k=5)`` turns a circuit of op tuples (see llsv.circuits)
into dense blocks of at most ``max_k`` qubits, ``('dense', targets, U)`` with U of shape
2^k x 2^k and operator bit j <-> targets[j] (the bit convention of src/Noise.h:58-65),
or ``('u', q, g4)`` for one-qubit blocks.  Blocks are grown greedily along the circuit:
an op joins the open block while the union of qubits stays within ``max_k``; otherwise
the block is emitted and a new one opens.  The emitted circuit equals the input one up to
rounding.

``plan_fusion(n, ops, max_k=None)`` also chooses the fusion width (2..5) by the same prediction.

``plan_fusion(n, ops)`` chooses, region by region, between the two ways the engine runs a
stretch of circuit: one register-tile pass of the program path (``apply_program``: one read
and write of the state, every gate in FP64 vector arithmetic) or dense blocks of at most five
qubits (``apply_fused``: one read and write per block, the arithmetic on the FP64 matrix
cores).  The regions are the program planner's passes; each gets the option with the lower
predicted time (``predict_pass`` / ``predict_blocks``); ``run_fusion_plan`` executes the result.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["GateFuser", "GATES", "fuse_dense", "op_matrix", "plan_fusion", "run_fusion_plan", "predict_pass",
           "predict_blocks", "predict_program"]


class GateFuser:
    """Pre-processor that merges consecutive gates acting on the same qubit."""

    @staticmethod
    def multiply_gates(gate1, gate2):
        a = np.asarray(gate1, dtype=np.complex128).reshape(2, 2)
        b = np.asarray(gate2, dtype=np.complex128).reshape(2, 2)
        return (b @ a).reshape(4).tolist()

    @staticmethod
    def fuse_circuit(gates):
        out = []
        for target, gate in gates:
            if out and out[-1][0] == target:
                out[-1] = (target, GateFuser.multiply_gates(out[-1][1], gate))
            else:
                out.append((target, gate))
        return out

    @staticmethod
    def analyze_circuit(gates):
        fused = GateFuser.fuse_circuit(gates)
        before, after = len(gates), len(fused)
        print(f"Original gates: {before}")
        print(f"After fusion: {after}")
        print(f"Reduction: {(1 - after / before) * 100:.1f}%")
        return fused

    @staticmethod
    def fuse_dense(ops, max_k: int = 5):
        return fuse_dense(ops, max_k)


_R2 = 1.0 / np.sqrt(2.0)
GATES = {
    "X": [0j, 1 + 0j, 1 + 0j, 0j],
    "H": [_R2 + 0j, _R2 + 0j, _R2 + 0j, -_R2 + 0j],
    "S": [1 + 0j, 0j, 0j, 1j],
    "T": [1 + 0j, 0j, 0j, np.exp(1j * np.pi / 4)],
}


def _gate4(op):
    kind = op[0]
    if kind == "u":
        return np.asarray(op[2], dtype=np.complex128).reshape(2, 2)
    if kind == "h":
        r = 1.0 / math.sqrt(2.0)
        return np.array([[r, r], [r, -r]], dtype=np.complex128)
    if kind == "x":
        return np.array([[0, 1], [1, 0]], dtype=np.complex128)
    t = float(op[2])
    c, s = math.cos(0.5 * t), math.sin(0.5 * t)
    if kind == "rx":
        return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)
    if kind == "ry":
        return np.array([[c, -s], [s, c]], dtype=np.complex128)
    if kind == "rz":
        return np.array([[complex(math.cos(-0.5 * t), math.sin(-0.5 * t)), 0],
                         [0, complex(math.cos(0.5 * t), math.sin(0.5 * t))]], dtype=np.complex128)
    raise KeyError(kind)


def op_qubits(op):
    kind = op[0]
    if kind in ("cnot", "cz"):
        return [int(op[1]), int(op[2])]
    if kind in ("mcz", "dense"):
        return [int(q) for q in op[1]]
    return [int(op[1])]


def op_matrix(op, targets):
    """Matrix of `op` on the block qubits `targets` (op's qubits must be among them),
    2^k x 2^k with local bit j <-> targets[j]."""
    k = len(targets)
    D = 1 << k
    pos = {q: j for j, q in enumerate(targets)}
    idx = np.arange(D)
    M = np.zeros((D, D), dtype=np.complex128)
    kind = op[0]
    if kind in ("cnot",):
        c, t = pos[int(op[1])], pos[int(op[2])]
        dst = np.where((idx >> c) & 1, idx ^ (1 << t), idx)
        M[dst, idx] = 1.0
        return M
    if kind in ("cz", "mcz"):
        mask = sum(1 << pos[q] for q in op_qubits(op))
        return np.diag(np.where((idx & mask) == mask, -1.0, 1.0).astype(np.complex128))
    if kind == "dense":
        sub = [pos[int(q)] for q in op[1]]
        U = np.asarray(op[2], dtype=np.complex128)
        smask = sum(1 << j for j in sub)
        for col in range(D):
            c_loc = sum(((col >> j) & 1) << i for i, j in enumerate(sub))
            rest = col & ~smask
            for r_loc in range(1 << len(sub)):
                row = rest | sum(((r_loc >> i) & 1) << j for i, j in enumerate(sub))
                M[row, col] = U[r_loc, c_loc]
        return M
    g = _gate4(op)
    j = pos[int(op[1])]
    bit = (idx >> j) & 1
    base = idx & ~(1 << j)
    for rb in (0, 1):
        M[base | (rb << j), idx] = g[rb, bit]
    return M


def _widen(M, old, new):
    """Re-express block matrix M on qubits `old` as a matrix on `new` (old is a prefix of new)."""
    extra = len(new) - len(old)
    if extra == 0:
        return M
    # New local bits beyond len(old) are untouched: M (x) I in the new high bits.
    return np.kron(np.eye(1 << extra, dtype=np.complex128), M)


def fuse_dense(ops, max_k: int = 5):
    if max_k < 1:
        raise ValueError("max_k must be >= 1")
    out = []
    targets: list = []
    M = None

    def flush():
        nonlocal targets, M
        if M is None:
            return
        if len(targets) == 1:
            out.append(("u", targets[0], M.reshape(4).tolist()))
        else:
            out.append(("dense", list(targets), M))
        targets, M = [], None

    for op in ops:
        qs = op_qubits(op)
        if len(set(qs)) > max_k:
            flush()
            out.append(op)
            continue
        union = targets + [q for q in qs if q not in targets]
        if M is not None and len(union) > max_k:
            flush()
            union = list(dict.fromkeys(qs))
        if M is None:
            targets = union
            M = np.eye(1 << len(targets), dtype=np.complex128)
        else:
            M = _widen(M, targets, union)
            targets = union
        M = op_matrix(op, targets) @ M
    flush()
    return out


# ---------------------------------------------------------------- cost model: dense blocks or a register pass
# Rates of one MI355X, measured by the benchmark legs (bench.py): a streamed pass over the state
# (read and write, 32 B per amplitude) at the one-qubit kernels' 6.2 TB/s; FP64 vector instructions
# of a register-tile pass at 3.93e13 per second (78.6 TFLOP/s of FMAs); a dense block of k qubits
# at the bandwidth its kernel reaches (k = 4, 5: FP64 matrix cores, 5.1 / 4.8 TB/s; k <= 3:
# vector arithmetic).  A register-tile pass costs its memory time plus its FP64 issue time (the
# two overlap little at two workgroups per CU): this reproduces the 28-qubit layered circuit's 11
# passes (31.0 ms predicted, 30.2 measured, each pass within 20%) and the 30-qubit sweep's 4
# (25.8 vs 24.9 ms).
PASS_BPS = 6.2e12
FP64_INSTR_PER_S = 3.93e13
BLOCK_BPS = {1: 6.2e12, 2: 5.5e12, 3: 5.0e12, 4: 5.1e12, 5: 4.8e12}
# FP64 vector instructions per amplitude for the ops of a register-tile pass: a one-qubit matrix
# (RX RY RZ merged, Haar, H, an RX or RY alone) 4, one FMA per product in its normalised form;
# a diagonal one-qubit gate is carried into the next matrix on its qubit and costs about 1 (where a
# CNOT target or a pass end applies it); CNOT, CZ and MCZ move or negate amplitudes (about 0.5).
_OPS_GENERAL, _OPS_ROT, _OPS_DIAG, _OPS_PERM = 4.0, 4.0, 1.0, 0.5


def _fp64_per_amp(ops) -> float:
    """FP64 vector instructions per amplitude of `ops` in a register-tile pass.  One-qubit gates on
    the same qubit with no other gate on it in between are merged into one 2x2, as the planner does."""
    fp = 0.0
    run = {}  # qubit -> kinds of its current one-qubit run

    def close(q):
        nonlocal fp
        kinds = run.pop(q, None)
        if not kinds:
            return
        if len(kinds) > 1 or kinds[0] in ("h", "x", "u"):
            fp += _OPS_GENERAL
        elif kinds[0] in ("rx", "ry"):
            fp += _OPS_ROT
        else:
            fp += _OPS_DIAG

    for op in ops:
        qs = op_qubits(op)
        if op[0] in ("h", "x", "u", "rx", "ry", "rz"):
            run.setdefault(qs[0], []).append(op[0])
            continue
        for q in qs:
            close(q)
        fp += _OPS_GENERAL * (1 << len(qs)) if op[0] == "dense" else _OPS_PERM
    for q in list(run):
        close(q)
    return fp


def predict_pass(n: int, ops) -> float:
    """Seconds for one register-tile pass of `ops` on an n-qubit state: its memory time plus its
    FP64 issue time."""
    amps = float(1 << n)
    return 32.0 * amps / PASS_BPS + _fp64_per_amp(ops) * amps / FP64_INSTR_PER_S


def predict_blocks(n: int, blocks) -> float:
    """Seconds for dense blocks (fuse_dense output) on an n-qubit state: one pass per block."""
    amps = float(1 << n)
    t = 0.0
    for b in blocks:
        k = len(b[1]) if b[0] == "dense" else len(op_qubits(b))
        t += 32.0 * amps / BLOCK_BPS[min(max(k, 1), 5)]
    return t


def predict_program(n: int, ops, tile_bits: int = 12) -> float:
    """Seconds for `ops` through the program path: the planner's passes, each as predict_pass."""
    from llsv import program as P  # the planner (host only, through libllsv)

    if not ops:
        return 0.0
    order, ends = P.plan(n, ops, min(n, tile_bits))
    t, start = 0.0, 0
    for end in ends:
        t += predict_pass(n, [ops[i] for i in order[start:end]])
        start = end
    return t


def _groups(ops, max_k):
    """The runs fuse_dense would turn into one block: (targets, ops), along the circuit."""
    out, targets, cur = [], [], []
    for op in ops:
        qs = list(dict.fromkeys(op_qubits(op)))
        if len(qs) > max_k:
            if cur:
                out.append((targets, cur))
            out.append((qs, [op]))
            targets, cur = [], []
            continue
        union = targets + [q for q in qs if q not in targets]
        if cur and len(union) > max_k:
            out.append((targets, cur))
            targets, cur, union = [], [], qs
        targets, cur = union, cur + [op]
    if cur:
        out.append((targets, cur))
    return out


def plan_fusion(n: int, ops, max_k: int | None = 5, tile_bits: int = 12):
    """Segments [("program", ops) | ("dense", blocks)] that apply the circuit, and their predicted
    seconds.  The circuit is cut into fuse_dense's groups (at most max_k qubits each); a group of two
    or more qubits becomes one dense block when its FP64 work inside a register-tile pass would take
    longer than a dense pass over the state, and the other groups stay gates for the program path
    (consecutive ones in one segment, which the planner packs into passes).  If the program path
    alone is predicted faster than that plan, the plan is the whole circuit as one program.
    max_k=None chooses the fusion width as well: the plan of lowest predicted time over widths
    2..5 (the widest on a tie)."""
    if max_k is None:
        best = None
        for k in (5, 4, 3, 2):
            cand = plan_fusion(n, ops, k, tile_bits)
            if best is None or cand[1] < best[1]:
                best = cand
        return best
    segs = []
    for targets, g in _groups(ops, max_k):
        k = len(targets)
        dense = (2 <= k <= 5 and len(g) > 1 and g[0][0] != "mcz"
                 and _fp64_per_amp(g) / FP64_INSTR_PER_S > 32.0 / BLOCK_BPS[k])
        if dense:
            block = fuse_dense(g, max_k)
            if segs and segs[-1][0] == "dense":
                segs[-1][1].extend(block)
            else:
                segs.append(("dense", block))
        elif segs and segs[-1][0] == "program":
            segs[-1][1].extend(g)
        else:
            segs.append(("program", list(g)))
    t_plan = sum(predict_program(n, b, tile_bits) if k == "program" else predict_blocks(n, b) for k, b in segs)
    t_prog = predict_program(n, ops, tile_bits)
    if t_prog <= t_plan or not any(k == "dense" for k, _ in segs):
        return [("program", list(ops))], t_prog
    return segs, t_plan


def run_fusion_plan(sv, segments) -> None:
    """Applies plan_fusion's segments to a quantum_sim_v3 StateVector, in order."""
    for kind, body in segments:
        if kind == "program":
            sv.apply_program(body)
            continue
        for b in body:
            if b[0] == "dense":
                sv.apply_fused(list(b[1]), np.ascontiguousarray(b[2]))
            elif b[0] == "u":
                sv.apply_gate_cache_optimized(b[1], list(b[2]))
            else:
                sv.apply_program([b])
