"""Op tables of the LDS program kernel, built with NumPy only (no GPU, no torch).

The layout mirrors ``lsim_prog_op`` of include/lsim.h. :func:`static_table` turns an ideal static
circuit into its op table and complex128 matrix pool with array operations over all gates at
once; it is safe to call in front-end worker processes that never touch the GPU.
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence
from itertools import chain
from operator import attrgetter

import numpy as np

from lucivant_sim.circuit import Circuit, Gate, GlobalPhase

OP_DENSE, OP_DIAG, OP_MIX, OP_KRAUS, OP_MEASURE, OP_RESET, OP_COND, OP_PHASE = range(8)

MAX_FLAGS = 32
MAX_CLBITS = 64
#: widest dense (and mixture) operator of a program; diagonal ones may have 5 qubits
MAX_DENSE_QUBITS = 4


class ProgOp(ctypes.Structure):
    """Mirror of ``lsim_prog_op`` in include/lsim.h."""

    _fields_ = [
        ("kind", ctypes.c_ubyte),
        ("k", ctypes.c_ubyte),
        ("flag", ctypes.c_ubyte),
        ("negate", ctypes.c_ubyte),
        ("t", ctypes.c_ubyte * 5),
        ("pad", ctypes.c_ubyte * 3),
        ("mat", ctypes.c_int),
        ("nb", ctypes.c_int),
        ("aux", ctypes.c_int),
        ("draw", ctypes.c_int),
        ("clbit", ctypes.c_int),
        ("guard_mask", ctypes.c_uint),
        ("guard_value", ctypes.c_uint),
        ("cmask", ctypes.c_ulonglong),
        ("cpattern", ctypes.c_ulonglong),
    ]


OP_DTYPE = np.dtype(
    [
        ("kind", "u1"),
        ("k", "u1"),
        ("flag", "u1"),
        ("negate", "u1"),
        ("t", "u1", (5,)),
        ("pad", "u1", (3,)),
        ("mat", "<i4"),
        ("nb", "<i4"),
        ("aux", "<i4"),
        ("draw", "<i4"),
        ("clbit", "<i4"),
        ("guard_mask", "<u4"),
        ("guard_value", "<u4"),
        ("cmask", "<u8"),
        ("cpattern", "<u8"),
    ],
    align=True,
)
assert OP_DTYPE.itemsize == ctypes.sizeof(ProgOp)


class Unsupported(ValueError):
    """The circuit needs something the LDS program kernel doesn't do (it runs elsewhere)."""


def static_table(circuit: Circuit) -> tuple[np.ndarray, np.ndarray]:
    """(op table, complex128 matrix pool) of an ideal static circuit, one op per gate."""
    table, pool, _starts, _rows = static_tables([circuit])
    return table, pool


_QUBITS = attrgetter("qubits")
_MATRIX = attrgetter("matrix")


def static_tables(circuits: Sequence[Circuit]) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """:func:`static_table` of many circuits at once, already concatenated: (op table, matrix
    pool, first row per circuit, row count per circuit). Each circuit's ``mat`` offsets point into
    the one pool; the rows and pool entries equal those of the per-circuit tables laid end to end.
    One set of NumPy calls covers every gate of the batch and no Python object is made per gate,
    so the host cost per circuit is a few list operations."""
    gates: list[Gate] = []
    per: list[int] = []
    phases: list[complex] = []
    for c in circuits:
        ops = c.operations
        g = [op for op in ops if type(op) is Gate]
        phase = 1.0 + 0.0j
        if len(g) != len(ops):
            for op in ops:
                if type(op) is GlobalPhase:
                    phase *= np.exp(1j * op.angle)
        gates.extend(g)
        per.append(len(g))
        phases.append(phase)
    count = np.asarray(per, dtype=np.int64)
    has_phase = np.fromiter((ph != 1 for ph in phases), dtype=bool, count=len(phases))
    rows = count + has_phase
    starts = np.cumsum(rows) - rows
    table = np.zeros(int(rows.sum()), dtype=OP_DTYPE)
    table["aux"] = -1
    table["clbit"] = -1
    qubit_lists = list(map(_QUBITS, gates))
    ks = np.fromiter(map(len, qubit_lists), dtype=np.int64, count=len(gates))
    if np.any(ks > 5):
        raise Unsupported("operators on more than 5 qubits")
    sizes = np.zeros(len(gates), dtype=np.int64)
    kinds = np.zeros(len(gates), dtype=np.uint8)
    groups = []
    matrices = list(map(_MATRIX, gates))
    for k in np.unique(ks):
        idx = np.flatnonzero(ks == k)
        d = 1 << int(k)
        mats = np.concatenate([matrices[i] for i in idx]).astype(np.complex128, copy=False).reshape(len(idx), d, d)
        off = ~np.eye(d, dtype=bool)
        diag = ~np.any(mats[:, off], axis=1)
        if k > MAX_DENSE_QUBITS and not np.all(diag):
            raise Unsupported(f"dense operators on more than {MAX_DENSE_QUBITS} qubits")
        sizes[idx] = np.where(diag, d, d * d)
        kinds[idx] = np.where(diag, OP_DIAG, OP_DENSE)
        groups.append((idx, mats, diag, d))
    # pool layout per circuit: its gates' entries in order, then its phase entry (if any)
    cid = np.repeat(np.arange(len(per)), count)
    slot_extra = np.cumsum(has_phase) - has_phase  # phase slots before each circuit
    gate_off = np.cumsum(sizes) - sizes + slot_extra[cid]
    pool = np.empty(int(sizes.sum()) + int(has_phase.sum()), dtype=np.complex128)
    for idx, mats, diag, d in groups:
        if np.any(diag):
            sel = idx[diag]
            pool[gate_off[sel][:, None] + np.arange(d)[None, :]] = np.diagonal(mats[diag], axis1=1, axis2=2)
        if np.any(~diag):
            sel = idx[~diag]
            pool[gate_off[sel][:, None] + np.arange(d * d)[None, :]] = mats[~diag].reshape(-1, d * d)
    grow = np.arange(len(gates), dtype=np.int64) + slot_extra[cid]  # gate rows skip earlier phase rows
    table["kind"][grow] = kinds
    table["k"][grow] = ks
    flat_qubits = np.fromiter(chain.from_iterable(qubit_lists), dtype=np.int64, count=int(ks.sum()))
    first = np.cumsum(ks) - ks
    column = np.arange(flat_qubits.size, dtype=np.int64) - np.repeat(first, ks)
    targets = table["t"]
    targets[np.repeat(grow, ks), column] = flat_qubits
    table["t"] = targets
    table["mat"][grow] = gate_off
    if has_phase.any():
        pc = np.flatnonzero(has_phase)
        prow = starts[pc] + count[pc]
        # a circuit's phase entry follows its last gate's entries
        ends = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        first_gate = np.concatenate([[0], np.cumsum(count)]).astype(np.int64)
        pslot = ends[first_gate[pc + 1]] + slot_extra[pc]
        pool[pslot] = np.asarray(phases, dtype=np.complex128)[pc]
        table["kind"][prow] = OP_PHASE
        table["mat"][prow] = pslot
    return table, pool, starts, rows
