"""Gate unitaries as complex128 matrices, first listed qubit = most significant matrix bit.

The closed forms are the OpenQASM 2/3 standard-library definitions (lucivant-sim 0.3.0
src/lucivant_sim/gates.py:33-167 uses the same ones). Each entry is evaluated with the same
floating-point steps as the reference (``cos(t/2)``, ``exp(1j*phi) * s`` ...), because circuit
digests round matrices to 12 decimals and must come out identical for saved results to verify.
"""

from __future__ import annotations

import cmath
import math
from collections.abc import Callable
from dataclasses import dataclass

import numpy as np
from numpy.typing import NDArray

Matrix = NDArray[np.complex128]

I2 = np.eye(2, dtype=np.complex128)
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2)
PAULIS: dict[str, Matrix] = {"I": I2, "X": X, "Y": Y, "Z": Z}


def _mat(rows: list[list[complex]]) -> Matrix:
    return np.array(rows, dtype=np.complex128)


def rx(t: float) -> Matrix:
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _mat([[c, -1j * s], [-1j * s, c]])


def ry(t: float) -> Matrix:
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _mat([[c, -s], [s, c]])


def rz(t: float) -> Matrix:
    return _mat([[cmath.exp(-0.5j * t), 0], [0, cmath.exp(0.5j * t)]])


def phase(t: float) -> Matrix:
    return _mat([[1, 0], [0, cmath.exp(1j * t)]])


def u3(theta: float, phi: float, lam: float) -> Matrix:
    """U(theta, phi, lambda) of the OpenQASM specification."""
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return _mat([[c, -cmath.exp(1j * lam) * s], [cmath.exp(1j * phi) * s, cmath.exp(1j * (phi + lam)) * c]])


def controlled(u: Matrix, controls: int = 1) -> Matrix:
    """``u`` on the trailing qubits when all ``controls`` leading qubits are 1."""
    d = u.shape[0]
    size = d << controls
    out = np.eye(size, dtype=np.complex128)
    out[size - d :, size - d :] = u
    return out


def negcontrolled(u: Matrix) -> Matrix:
    """``u`` on the trailing qubits when the leading qubit is 0."""
    d = u.shape[0]
    out = np.eye(2 * d, dtype=np.complex128)
    out[:d, :d] = u
    return out


def pauli_string_matrix(label: str) -> Matrix:
    """Kronecker product of the letters, first letter = most significant bit."""
    m = np.array([[1]], dtype=np.complex128)
    for letter in label:
        m = np.kron(m, PAULIS[letter]).astype(np.complex128)
    return m


def pauli_rotation(paulis: str, t: float) -> Matrix:
    """exp(-i t/2 P) for the Pauli string P."""
    p = pauli_string_matrix(paulis)
    return math.cos(t / 2) * np.eye(p.shape[0], dtype=np.complex128) - 1j * math.sin(t / 2) * p


SX = _mat([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]) / 2
SWAP = _mat([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])
ISWAP = _mat([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]])
ECR = ((np.kron(X, I2) - np.kron(Y, X)) / math.sqrt(2)).astype(np.complex128)
DCX = _mat([[1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1], [0, 1, 0, 0]])


@dataclass(frozen=True)
class GateDef:
    """A built-in gate: number of parameters, number of qubits, matrix builder."""

    params: int
    qubits: int
    matrix: Callable[..., Matrix]


def _const(m: Matrix) -> Callable[..., Matrix]:
    return lambda: m


_CX = controlled(X)
_SWAP3 = controlled(SWAP)

GATES: dict[str, GateDef] = {
    "id": GateDef(0, 1, _const(I2)),
    "i": GateDef(0, 1, _const(I2)),
    "x": GateDef(0, 1, _const(X)),
    "y": GateDef(0, 1, _const(Y)),
    "z": GateDef(0, 1, _const(Z)),
    "h": GateDef(0, 1, _const(H)),
    "s": GateDef(0, 1, _const(phase(math.pi / 2))),
    "sdg": GateDef(0, 1, _const(phase(-math.pi / 2))),
    "t": GateDef(0, 1, _const(phase(math.pi / 4))),
    "tdg": GateDef(0, 1, _const(phase(-math.pi / 4))),
    "sx": GateDef(0, 1, _const(SX)),
    "sxdg": GateDef(0, 1, _const(SX.conj().T)),
    "rx": GateDef(1, 1, rx),
    "ry": GateDef(1, 1, ry),
    "rz": GateDef(1, 1, rz),
    "p": GateDef(1, 1, phase),
    "phase": GateDef(1, 1, phase),
    "u1": GateDef(1, 1, phase),
    "u2": GateDef(2, 1, lambda phi, lam: u3(math.pi / 2, phi, lam)),
    "u3": GateDef(3, 1, u3),
    "u": GateDef(3, 1, u3),
    "U": GateDef(3, 1, u3),
    "u0": GateDef(1, 1, lambda _gamma: I2),
    "cx": GateDef(0, 2, _const(_CX)),
    "CX": GateDef(0, 2, _const(_CX)),
    "cnot": GateDef(0, 2, _const(_CX)),
    "cy": GateDef(0, 2, _const(controlled(Y))),
    "cz": GateDef(0, 2, _const(controlled(Z))),
    "ch": GateDef(0, 2, _const(controlled(H))),
    "csx": GateDef(0, 2, _const(controlled(SX))),
    "cp": GateDef(1, 2, lambda t: controlled(phase(t))),
    "cphase": GateDef(1, 2, lambda t: controlled(phase(t))),
    "cu1": GateDef(1, 2, lambda t: controlled(phase(t))),
    "crx": GateDef(1, 2, lambda t: controlled(rx(t))),
    "cry": GateDef(1, 2, lambda t: controlled(ry(t))),
    "crz": GateDef(1, 2, lambda t: controlled(rz(t))),
    "cu3": GateDef(3, 2, lambda a, b, c: controlled(u3(a, b, c))),
    "cu": GateDef(4, 2, lambda a, b, c, g: controlled(cmath.exp(1j * g) * u3(a, b, c))),
    "swap": GateDef(0, 2, _const(SWAP)),
    "iswap": GateDef(0, 2, _const(ISWAP)),
    "ecr": GateDef(0, 2, _const(ECR)),
    "dcx": GateDef(0, 2, _const(DCX)),
    "rxx": GateDef(1, 2, lambda t: pauli_rotation("XX", t)),
    "ryy": GateDef(1, 2, lambda t: pauli_rotation("YY", t)),
    "rzz": GateDef(1, 2, lambda t: pauli_rotation("ZZ", t)),
    "rzx": GateDef(1, 2, lambda t: pauli_rotation("ZX", t)),
    "ccx": GateDef(0, 3, _const(controlled(X, 2))),
    "toffoli": GateDef(0, 3, _const(controlled(X, 2))),
    "ccz": GateDef(0, 3, _const(controlled(Z, 2))),
    "cswap": GateDef(0, 3, _const(_SWAP3)),
    "fredkin": GateDef(0, 3, _const(_SWAP3)),
    "c3x": GateDef(0, 4, _const(controlled(X, 3))),
    "c3sqrtx": GateDef(0, 4, _const(controlled(SX, 3))),
    "c4x": GateDef(0, 5, _const(controlled(X, 4))),
}

#: Other spellings of a gate; noise rules and labels use the canonical name.
ALIASES: dict[str, str] = {
    "U": "u",
    "u3": "u",
    "CX": "cx",
    "cnot": "cx",
    "phase": "p",
    "u1": "p",
    "cphase": "cp",
    "cu1": "cp",
    "toffoli": "ccx",
    "fredkin": "cswap",
    "i": "id",
}


def canonical_name(name: str) -> str:
    return ALIASES.get(name, name)


def gate_matrix(name: str, params: tuple[float, ...]) -> Matrix:
    return GATES[name].matrix(*params)


def matrix_power(u: Matrix, k: float) -> Matrix:
    """u**k: integer powers by repeated products (negative through the adjoint), fractional ones
    through the eigen-decomposition."""
    if float(k).is_integer():
        n = int(k)
        return np.linalg.matrix_power(u if n >= 0 else u.conj().T, abs(n))
    vals, vecs = np.linalg.eig(u)
    return (vecs @ np.diag(vals**k) @ np.linalg.inv(vecs)).astype(np.complex128)
