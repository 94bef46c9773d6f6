"""Noise channels after gates and readout errors on measurement (host data for the engines).

Behaviour follows lucivant-sim 0.3.0 src/lucivant_sim/noise.py:
- a :class:`Channel` is a CPTP map given by Kraus operators, checked at construction
  (noise.py:50-54); Pauli channels also carry ``mixture``, the (probability, unitary) branches in
  the order trajectories draw them: identity first when the total is below 1, then the labels in
  sorted order, zero-probability labels dropped (noise.py:66-85);
- a :class:`NoiseModel` attaches at most one channel per gate, the most specific rule winning and
  ties going to the rule added last (noise.py:163-225); frame-change gates get no default error
  (noise.py:31, 176-177);
- ``describe``/``digest`` give the same canonical JSON and sha256 as the reference
  (noise.py:236-249), so saved results verify across the two implementations.
"""

from __future__ import annotations

import hashlib
import itertools
import json
import math
from collections.abc import Iterable, Mapping, Sequence
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np

from lucivant_sim.gates import Matrix, canonical_name, pauli_string_matrix

#: Frame changes on real devices: no error attached unless a rule names them.
IDEAL_GATES = frozenset({"rz", "p", "phase", "u1", "id", "i", "u0"})


@dataclass(frozen=True, eq=False)
class Channel:
    """A completely positive, trace-preserving map given by its Kraus operators.

    ``mixture`` holds (probability, unitary) branches when the channel is a mix of unitaries, so a
    trajectory picks a branch from one draw without computing norms.
    """

    name: str
    num_qubits: int
    kraus: tuple[Matrix, ...]
    params: tuple[float, ...] = ()
    mixture: tuple[tuple[float, Matrix], ...] | None = None
    #: Set by :func:`depolarizing` only: the density matrix may use the partial-trace form.
    analytic_depolarizing: bool = False

    def __post_init__(self) -> None:
        d = 2**self.num_qubits
        total = sum(k.conj().T @ k for k in self.kraus)
        if not np.allclose(total, np.eye(d), atol=1e-9):
            raise ValueError(f"{self.name}: the Kraus operators don't preserve probability")

    def describe(self) -> list[Any]:
        return [self.name, self.num_qubits, [round(p, 15) for p in self.params]]

    def superoperator(self) -> Matrix:
        """S = sum_i K_i (x) conj(K_i) on the vectorised density matrix (row bits, then column
        bits); for the analytic depolariser, (1-p) I + (p/d) |I>><<I| in that basis. Computed
        once per channel (a noise model hands the same channel to every matching gate) and
        returned read-only."""
        cached = self.__dict__.get("_superop")
        if cached is not None:
            return cached
        d = 2**self.num_qubits
        if self.analytic_depolarizing:
            p = float(self.params[0])
            vec_id = np.eye(d, dtype=np.complex128).reshape(-1)
            s = (1 - p) * np.eye(d * d, dtype=np.complex128) + (p / d) * np.outer(vec_id, vec_id)
        else:
            s = np.zeros((d * d, d * d), dtype=np.complex128)
            for k in self.kraus:
                s += kron_conj(k)
        s.flags.writeable = False
        self.__dict__["_superop"] = s
        return s


def kron_conj(m: Matrix) -> Matrix:
    """``np.kron(m, m.conj())`` for a square matrix, as one broadcast product."""
    m = np.asarray(m, dtype=np.complex128)
    d = m.shape[0]
    return (m[:, None, :, None] * m.conj()[None, :, None, :]).reshape(d * d, d * d)


def _probability(p: float, what: str, upper: float = 1.0) -> float:
    if math.isnan(p) or not 0.0 <= p <= upper:
        raise ValueError(f"{what} must be between 0 and {upper:g}, not {p}")
    return float(p)


def pauli_channel(probabilities: Mapping[str, float]) -> Channel:
    """Apply the Pauli string P with probability ``probabilities[P]``, the identity otherwise."""
    width = len(next(iter(probabilities)))
    total = sum(probabilities.values())
    _probability(total, "the total Pauli probability")
    branches: list[tuple[float, Matrix]] = []
    if total < 1:
        branches.append((1 - total, np.eye(2**width, dtype=np.complex128)))
    for label, p in sorted(probabilities.items()):
        if len(label) != width or set(label) - set("IXYZ"):
            raise ValueError(f"{label!r} isn't a {width}-qubit Pauli string")
        if p <= 0:
            continue
        branches.append((p, pauli_string_matrix(label)))
    kraus = tuple(math.sqrt(p) * m for p, m in branches)
    params = tuple(probabilities[key] for key in sorted(probabilities))
    return Channel("pauli", width, kraus, params, tuple(branches))


def depolarizing(p: float, num_qubits: int = 1) -> Channel:
    """With probability p the qubits are replaced by the maximally mixed state: each of the
    4**k - 1 non-identity Pauli strings with probability p / 4**k."""
    d2 = 4**num_qubits
    _probability(p, "a depolarising probability", d2 / (d2 - 1))
    probs = {"".join(s): p / d2 for s in itertools.product("IXYZ", repeat=num_qubits) if set(s) != {"I"}}
    ch = pauli_channel(probs)
    return Channel("depolarizing", num_qubits, ch.kraus, (p,), ch.mixture, analytic_depolarizing=True)


def from_gate_error(error: float, num_qubits: int) -> Channel:
    """The depolarising channel with average gate infidelity ``error``: p = r d / (d - 1)."""
    d = 2**num_qubits
    return depolarizing(min(error * d / (d - 1), 4**num_qubits / (4**num_qubits - 1)), num_qubits)


def bit_flip(p: float) -> Channel:
    ch = pauli_channel({"X": _probability(p, "a bit-flip probability")})
    return Channel("bit_flip", 1, ch.kraus, (p,), ch.mixture)


def phase_flip(p: float) -> Channel:
    ch = pauli_channel({"Z": _probability(p, "a phase-flip probability")})
    return Channel("phase_flip", 1, ch.kraus, (p,), ch.mixture)


def amplitude_damping(gamma: float) -> Channel:
    """|1> decays to |0> with probability gamma."""
    g = _probability(gamma, "gamma")
    k0 = np.array([[1, 0], [0, math.sqrt(1 - g)]], dtype=np.complex128)
    k1 = np.array([[0, math.sqrt(g)], [0, 0]], dtype=np.complex128)
    return Channel("amplitude_damping", 1, (k0, k1), (g,))


def phase_damping(lam: float) -> Channel:
    """Dephasing without energy loss."""
    lam = _probability(lam, "lambda")
    k0 = np.array([[1, 0], [0, math.sqrt(1 - lam)]], dtype=np.complex128)
    k1 = np.array([[0, 0], [0, math.sqrt(lam)]], dtype=np.complex128)
    return Channel("phase_damping", 1, (k0, k1), (lam,))


def thermal_relaxation(t1: float, t2: float, time: float) -> Channel:
    """T1 relaxation composed with the extra dephasing T2 implies, over ``time``.

    gamma = 1 - exp(-t/T1), lambda = 1 - exp(-2t/T2 + t/T1); Kraus operators in the order
    [pd0 ad0, pd1 ad0, pd0 ad1, pd1 ad1] (noise.py:134-144)."""
    if t1 <= 0 or t2 <= 0 or time < 0:
        raise ValueError("T1 and T2 must be positive and the time not negative")
    if t2 > 2 * t1 + 1e-12:
        raise ValueError("T2 can't exceed 2*T1")
    gamma = 1 - math.exp(-time / t1)
    lam = max(0.0, 1 - math.exp(-2 * time / t2 + time / t1))
    ad, pd = amplitude_damping(gamma), phase_damping(min(lam, 1.0))
    kraus = tuple(b @ a for a in ad.kraus for b in pd.kraus)
    return Channel("thermal_relaxation", 1, kraus, (t1, t2, time))


@dataclass(frozen=True)
class ReadoutError:
    """p01: probability of reading 1 from a 0; p10: of reading 0 from a 1."""

    p01: float
    p10: float

    def __post_init__(self) -> None:
        _probability(self.p01, "p01")
        _probability(self.p10, "p10")

    @classmethod
    def symmetric(cls, p: float) -> ReadoutError:
        return cls(p, p)


@dataclass
class _Rule:
    channel: Channel
    gates: frozenset[str] | None
    qubits: tuple[int, ...] | None

    def score(self, name: str, qubits: tuple[int, ...]) -> float:
        """-1 when the rule doesn't apply; otherwise 2 for named qubits plus 1 for named gates."""
        if self.channel.num_qubits != len(qubits):
            return -1
        if self.gates is not None and name not in self.gates:
            return -1
        if self.qubits is not None and sorted(self.qubits) != sorted(qubits):
            return -1
        if self.gates is None and name in IDEAL_GATES:
            return -1
        return (2 if self.qubits is not None else 0) + (1 if self.gates is not None else 0)


@dataclass
class NoiseModel:
    """Which channel follows which gate, and each qubit's readout error."""

    rules: list[_Rule] = field(default_factory=list)
    readout: dict[int | None, ReadoutError] = field(default_factory=dict)
    source: str | None = None

    def add_gate_error(
        self, channel: Channel, gates: str | Iterable[str] | None = None, qubits: Sequence[int] | None = None
    ) -> NoiseModel:
        names = None
        if gates is not None:
            names = frozenset(canonical_name(g) for g in ([gates] if isinstance(gates, str) else gates))
        if qubits is not None and len(qubits) != channel.num_qubits:
            raise ValueError(f"a {channel.num_qubits}-qubit channel can't attach to {len(qubits)} qubit(s)")
        self.rules.append(_Rule(channel, names, tuple(qubits) if qubits is not None else None))
        return self

    def add_readout_error(self, error: float | ReadoutError, qubits: Iterable[int] | None = None) -> NoiseModel:
        err = error if isinstance(error, ReadoutError) else ReadoutError.symmetric(float(error))
        for q in [None] if qubits is None else qubits:
            self.readout[q] = err
        return self

    def gate_error(self, name: str, qubits: tuple[int, ...]) -> Channel | None:
        best, best_score = None, -1.0
        for rule in self.rules:
            score = rule.score(name, qubits)
            if score >= 0 and score >= best_score:
                best, best_score = rule.channel, score
        return best

    def readout_error(self, qubit: int) -> ReadoutError | None:
        return self.readout.get(qubit, self.readout.get(None))

    def has_gate_errors(self) -> bool:
        return bool(self.rules)

    def is_empty(self) -> bool:
        return not self.rules and not any(e.p01 or e.p10 for e in self.readout.values())

    def describe(self) -> dict[str, Any]:
        rules = []
        for r in self.rules:
            rules.append([sorted(r.gates) if r.gates is not None else None, list(r.qubits or []) or None, r.channel.describe()])
        readout = sorted([[-1 if q is None else q, round(e.p01, 15), round(e.p10, 15)] for q, e in self.readout.items()])
        return {"rules": rules, "readout": readout}

    def digest(self) -> str:
        data = json.dumps(self.describe(), separators=(",", ":"), sort_keys=True).encode()
        return "sha256:" + hashlib.sha256(data).hexdigest()

    @classmethod
    def uniform(cls, single: float = 0.0, two: float = 0.0, readout: float = 0.0) -> NoiseModel:
        """One average error for every 1q gate, one for every 2q gate, one readout error."""
        model = cls(source="uniform")
        if single > 0:
            model.add_gate_error(from_gate_error(single, 1))
        if two > 0:
            model.add_gate_error(from_gate_error(two, 2))
        if readout > 0:
            model.add_readout_error(readout)
        return model

    @classmethod
    def from_calibration(cls, calibration: Mapping[str, Any] | str | Path) -> NoiseModel:
        """A model from a device calibration (median errors plus optional per-qubit and per-coupler
        values), the fields read by noise.py:265-297."""
        if isinstance(calibration, (str, Path)):
            cal: dict[str, Any] = json.loads(Path(calibration).read_text(encoding="utf-8"))
        else:
            cal = dict(calibration)
        if isinstance(cal.get("calibration"), Mapping):
            cal = dict(cal["calibration"])
        model = cls(source=str(cal.get("device") or cal.get("profile_key") or "calibration"))
        single, two, readout = cal.get("single_qubit_error"), cal.get("two_qubit_error"), cal.get("readout_error")
        if single:
            model.add_gate_error(from_gate_error(float(single), 1))
        if two:
            model.add_gate_error(from_gate_error(float(two), 2))
        if readout:
            model.add_readout_error(float(readout))
        for row in cal.get("qubits") or []:
            q = int(row["qubit"])
            if row.get("single_qubit_error") is not None:
                model.add_gate_error(from_gate_error(float(row["single_qubit_error"]), 1), qubits=[q])
            if row.get("readout_error") is not None:
                model.add_readout_error(float(row["readout_error"]), qubits=[q])
        for row in cal.get("couplers") or []:
            if row.get("error") is not None:
                a, b = (int(x) for x in row["qubits"])
                model.add_gate_error(from_gate_error(float(row["error"]), 2), qubits=[a, b])
        return model
