"""Clifford detection, used by ``automatic`` method selection.

``clifford_problem`` answers the same question as lucivant-sim 0.3.0
src/lucivant_sim/stabilizer.py:457-467 so ``choose_method`` picks the same method names (the
fingerprint includes the method): every gate must map each Pauli string to a Pauli string with a
real sign (checked, like the reference, on the matrix rounded to 12 decimals with a 1e-8
tolerance, and only for gates on at most 3 qubits), and every attached noise channel must be a
mixture of Pauli strings. The library has no tableau engine of its own.
"""

from __future__ import annotations

import itertools
from functools import lru_cache

import numpy as np

from lucivant_sim.circuit import Circuit, Gate, IfElse, Operation
from lucivant_sim.gates import Matrix, pauli_string_matrix
from lucivant_sim.noise import Channel, NoiseModel

MAX_STABILIZER_QUBITS = 2048
MAX_CLIFFORD_GATE_QUBITS = 3


@lru_cache(maxsize=4)
def _paulis(k: int) -> np.ndarray:
    return np.stack([pauli_string_matrix("".join(s)) for s in itertools.product("IXYZ", repeat=k)])


@lru_cache(maxsize=2048)
def _is_clifford_bytes(raw: bytes, k: int) -> bool:
    u = np.frombuffer(raw, dtype=np.complex128).reshape(1 << k, 1 << k)
    basis = _paulis(k)
    d = 1 << k
    # coefficient of Q in U P U^dagger: tr(Q^dagger U P U^dagger) / d
    images = np.einsum("ij,pjk,lk->pil", u, basis, u.conj())
    coeffs = np.einsum("qji,pjk->pqik", basis.conj(), images).trace(axis1=2, axis2=3) / d
    best = coeffs[np.arange(len(basis)), np.argmax(np.abs(coeffs), axis=1)]
    return bool(np.all(np.abs(np.abs(best) - 1) <= 1e-8) and np.all(np.abs(best.imag) <= 1e-8))


def is_clifford_matrix(matrix: Matrix) -> bool:
    d = matrix.shape[0]
    k = d.bit_length() - 1
    if not 0 < k <= MAX_CLIFFORD_GATE_QUBITS:
        return False
    rounded = np.round(np.ascontiguousarray(matrix, dtype=np.complex128), 12) + 0.0
    return _is_clifford_bytes(rounded.tobytes(), k)


def _is_pauli_up_to_phase(m: Matrix) -> bool:
    k = m.shape[0].bit_length() - 1
    basis = _paulis(k)
    coeffs = np.einsum("qji,ij->q", basis.conj(), m) / (1 << k)
    return bool(abs(abs(coeffs[int(np.argmax(np.abs(coeffs)))]) - 1) <= 1e-8)


def is_pauli_channel(channel: Channel) -> bool:
    return channel.mixture is not None and all(_is_pauli_up_to_phase(m) for _, m in channel.mixture)


def _gates(ops: tuple[Operation, ...]) -> list[Gate]:
    found: list[Gate] = []
    for op in ops:
        if isinstance(op, Gate):
            found.append(op)
        elif isinstance(op, IfElse):
            found.extend(_gates(op.then))
            found.extend(_gates(op.otherwise))
    return found


def clifford_problem(circuit: Circuit, noise: NoiseModel | None = None) -> str | None:
    """Why a stabilizer simulation couldn't run this circuit exactly, or None if it could."""
    if circuit.num_qubits > MAX_STABILIZER_QUBITS:
        return f"it has {circuit.num_qubits} qubits; the stabilizer engine allows up to {MAX_STABILIZER_QUBITS}"
    for gate in _gates(circuit.operations):
        if not is_clifford_matrix(np.asarray(gate.matrix)):
            return f"{gate.name} on qubits {list(gate.qubits)} isn't a Clifford gate"
        channel = noise.gate_error(gate.name, gate.qubits) if noise else None
        if channel is not None and not is_pauli_channel(channel):
            return f"the {channel.name} noise after {gate.name} isn't a Pauli channel"
    return None


def is_clifford(circuit: Circuit) -> bool:
    return clifford_problem(circuit) is None
