"""The circuit the engines run: a flat, immutable list of operations on numbered qubits and bits.

Gates carry their resolved matrix (first qubit = most significant bit), so the engines never see
OpenQASM. The types, the dynamic-circuit test, the final-measurement pairs, the canonical JSON,
the sha256 digest and the bit-string formatting behave as lucivant-sim 0.3.0
src/lucivant_sim/circuit.py:21-211; the canonical JSON is byte-identical, so digests (and the
result fingerprints built on them) agree with the reference.
"""

from __future__ import annotations

import hashlib
import json
from dataclasses import dataclass, field
from typing import Literal, Union

import numpy as np

from lucivant_sim.gates import Matrix


@dataclass(frozen=True, eq=False)
class Gate:
    """A unitary on ``qubits``; the first qubit is the matrix's most significant bit."""

    name: str
    qubits: tuple[int, ...]
    matrix: Matrix
    params: tuple[float, ...] = ()

    kind: Literal["gate"] = "gate"


@dataclass(frozen=True)
class Measure:
    qubit: int
    clbit: int | None  # None: the outcome is discarded

    kind: Literal["measure"] = "measure"


@dataclass(frozen=True)
class Reset:
    qubit: int

    kind: Literal["reset"] = "reset"


@dataclass(frozen=True)
class Barrier:
    qubits: tuple[int, ...]

    kind: Literal["barrier"] = "barrier"


@dataclass(frozen=True)
class GlobalPhase:
    angle: float

    kind: Literal["gphase"] = "gphase"


@dataclass(frozen=True)
class Condition:
    """Holds when the classical bits ``clbits`` (least significant first) read ``value``."""

    clbits: tuple[int, ...]
    value: int
    negate: bool = False

    def holds(self, bits: list[int]) -> bool:
        actual = 0
        for i, c in enumerate(self.clbits):
            actual |= bits[c] << i
        return (actual == self.value) != self.negate


@dataclass(frozen=True)
class IfElse:
    condition: Condition
    then: tuple[Operation, ...]
    otherwise: tuple[Operation, ...] = ()

    kind: Literal["if"] = "if"


Operation = Union[Gate, Measure, Reset, Barrier, GlobalPhase, IfElse]


@dataclass(frozen=True)
class Register:
    name: str
    offset: int
    size: int


_JSON_CACHE: dict[bytes, list[list[list[float]]]] = {}
_FORMAT_TABLES: dict[tuple, np.ndarray] = {}


def matrix_json(matrix: np.ndarray) -> list[list[list[float]]]:
    """[[ [re, im], ... ], ...] of the matrix rounded to 12 decimals (-0.0 folded to 0.0)."""
    key = np.ascontiguousarray(matrix, dtype=np.complex128).tobytes()
    hit = _JSON_CACHE.get(key)
    if hit is not None:
        return hit
    rounded = np.round(np.asarray(matrix, dtype=np.complex128), 12) + 0.0
    # (re, im) float64 pairs as nested Python floats, row by row
    out = np.ascontiguousarray(rounded).view(np.float64).reshape(rounded.shape[0], rounded.shape[1], 2).tolist()
    if len(_JSON_CACHE) < 8192:
        _JSON_CACHE[key] = out
    return out


def _op_json(op: Operation) -> object:
    if isinstance(op, Gate):
        return ["g", list(op.qubits), matrix_json(op.matrix)]
    if isinstance(op, Measure):
        return ["m", op.qubit, op.clbit]
    if isinstance(op, Reset):
        return ["r", op.qubit]
    if isinstance(op, Barrier):
        return ["b"]
    if isinstance(op, GlobalPhase):
        return ["p", round(op.angle, 12)]
    c = op.condition
    return ["if", [list(c.clbits), c.value, c.negate], [_op_json(o) for o in op.then], [_op_json(o) for o in op.otherwise]]


@dataclass(frozen=True)
class Circuit:
    num_qubits: int
    num_clbits: int
    operations: tuple[Operation, ...]
    qregs: tuple[Register, ...] = ()
    cregs: tuple[Register, ...] = ()
    #: OpenQASM ``input`` values the circuit was built with.
    inputs: dict[str, float] = field(default_factory=dict)

    def is_dynamic(self) -> bool:
        """True when one simulation can't serve every shot: a condition reads an outcome, or a
        qubit is used again after a mid-circuit measurement, or reset after being touched."""
        touched: set[int] = set()
        measured: set[int] = set()
        for op in self.operations:
            if isinstance(op, IfElse):
                return True
            if isinstance(op, Gate):
                if measured.intersection(op.qubits):
                    return True
                touched.update(op.qubits)
            elif isinstance(op, Reset):
                if op.qubit in touched or op.qubit in measured:
                    return True
            elif isinstance(op, Measure):
                measured.add(op.qubit)
        return False

    def final_measurements(self) -> list[tuple[int, int]]:
        """(qubit, clbit) pairs of a static circuit in first-write order; a later write to the same
        clbit replaces the qubit but keeps the position."""
        pairs: dict[int, int] = {}
        for op in self.operations:
            if isinstance(op, Measure) and op.clbit is not None:
                pairs[op.clbit] = op.qubit
        return [(q, c) for c, q in pairs.items()]

    def count_ops(self) -> dict[str, int]:
        counts: dict[str, int] = {}

        def visit(ops: tuple[Operation, ...]) -> None:
            for op in ops:
                if isinstance(op, IfElse):
                    visit(op.then)
                    visit(op.otherwise)
                    continue
                name = op.name if isinstance(op, Gate) else op.kind
                counts[name] = counts.get(name, 0) + 1

        visit(self.operations)
        return dict(sorted(counts.items()))

    def canonical(self) -> dict[str, object]:
        return {
            "qubits": self.num_qubits,
            "clbits": self.num_clbits,
            "cregs": [[r.name, r.offset, r.size] for r in self.cregs],
            "ops": [_op_json(o) for o in self.operations],
        }

    def digest(self) -> str:
        data = json.dumps(self.canonical(), separators=(",", ":"), sort_keys=True).encode()
        return "sha256:" + hashlib.sha256(data).hexdigest()

    def format_many(self, values: "np.ndarray") -> list[str]:
        """``format_bits`` of many integers; layouts of at most 16 classical bits use a cached
        table of every pattern (one lookup per value)."""
        if 0 < self.num_clbits <= 16:
            key = (self.num_clbits, tuple((r.offset, r.size) for r in self.cregs))
            table = _FORMAT_TABLES.get(key)
            if table is None:
                table = np.array([self.format_bits(v) for v in range(1 << self.num_clbits)], dtype=object)
                if len(_FORMAT_TABLES) < 64:
                    _FORMAT_TABLES[key] = table
            return table[np.asarray(values, dtype=np.int64)].tolist()
        return [self.format_bits(int(v)) for v in values]

    def format_bits(self, bits: int) -> str:
        """Registers last-declared first, space-separated, each with its bit 0 on the right."""
        if not self.cregs:
            return format(bits, f"0{self.num_clbits}b") if self.num_clbits else ""
        return " ".join(format((bits >> r.offset) & ((1 << r.size) - 1), f"0{r.size}b") for r in reversed(self.cregs))
