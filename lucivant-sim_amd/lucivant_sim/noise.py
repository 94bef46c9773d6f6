"""Noise: quantum channels attached after gates, and readout errors at measurement.

This is host data for the engines. Its semantics are lucivant-sim 0.3.0's
(src/lucivant_sim/noise.py), and tests/golden/noise_cases.json records them for a corpus of
channels and models:

- a :class:`Channel` is a completely positive, trace-preserving map given by Kraus operators,
  checked when it is made; a Pauli channel also lists its (probability, unitary) branches in the
  order a trajectory draws them: the identity first when the probabilities sum below 1, then the
  Pauli strings in sorted order, zero-probability strings left out;
- a :class:`NoiseModel` attaches at most one channel to a gate. Among the rules that fit the
  gate, a rule naming qubits outranks one naming gates, and naming both outranks either; equal
  ranks go to the rule added last. Frame-change gates (``IDEAL_GATES``) fit only rules that name
  them;
- ``describe``/``digest`` give the reference's canonical JSON and its sha256, so a saved result
  verifies against either implementation.

Beyond the reference, every channel offers its superoperator on the vectorised density matrix
(:meth:`Channel.superoperator`), which the density-matrix engine applies in one pass, and a model
answers repeated lookups of one (gate, qubits) from a table it fills as it is asked.
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

#: Gates that are frame changes on real devices: no error is attached unless a rule names them.
IDEAL_GATES = frozenset({"rz", "p", "phase", "u1", "id", "i", "u0"})


def _bounded(p: float, what: str, upper: float = 1.0) -> float:
    """``p`` as a float if it lies in [0, upper]; the reference's message otherwise."""
    if math.isnan(p) or not 0.0 <= p <= upper:
        raise ValueError(f"{what} must be between 0 and {upper:g}, not {p}")
    return float(p)


@dataclass(frozen=True, eq=False)
class Channel:
    """A CPTP map given by its Kraus operators (noise.py:34-57).

    ``mixture`` holds (probability, unitary) branches when the channel is a probabilistic mix of
    unitaries, so a trajectory picks a branch from one draw without computing norms.
    ``analytic_depolarizing`` is set by :func:`depolarizing` only.
    """

    name: str
    num_qubits: int
    kraus: tuple[Matrix, ...]
    params: tuple[float, ...] = ()
    mixture: tuple[tuple[float, Matrix], ...] | None = None
    analytic_depolarizing: bool = False

    def __post_init__(self) -> None:
        d = 1 << self.num_qubits
        ops = np.asarray(self.kraus, dtype=np.complex128).reshape(-1, d, d) if self.kraus else np.zeros((0, d, d))
        gram = np.einsum("kji,kjl->il", ops.conj(), ops)  # sum of K^dagger K
        if not np.allclose(gram, np.eye(d), atol=1e-9):
            raise ValueError(f"{self.name}: the Kraus operators don't preserve probability")

    def describe(self) -> list[Any]:
        return [self.name, self.num_qubits, [round(p, 15) for p in self.params]]

    def superoperator(self) -> Matrix:
        """S = sum_i K_i (x) conj(K_i) on the vectorised density matrix (row bits, then column
        bits); for the analytic depolariser (1-p) I + (p/d) |I>><<I| in that basis, the same map
        as the partial-trace update (engines.py:122-135). Built once per channel (a model hands
        the same channel to every gate it fits) and returned read-only."""
        cached = self.__dict__.get("_superop")
        if cached is not None:
            return cached
        d = 1 << self.num_qubits
        if self.analytic_depolarizing:
            p = float(self.params[0])
            vec = np.eye(d, dtype=np.complex128).reshape(-1)
            s = (1 - p) * np.eye(d * d, dtype=np.complex128) + (p / d) * np.outer(vec, vec)
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


# --------------------------------------------------------------------------------- channels


def pauli_channel(probabilities: Mapping[str, float]) -> Channel:
    """Each Pauli string P with probability ``probabilities[P]``, the identity with the rest
    (noise.py:66-85)."""
    width = len(next(iter(probabilities)))
    total = sum(probabilities.values())
    _bounded(total, "the total Pauli probability")
    ordered = sorted(probabilities.items())
    branches: list[tuple[float, Matrix]] = [(1 - total, np.eye(1 << width, dtype=np.complex128))] if total < 1 else []
    for label, p in ordered:
        if len(label) != width or set(label) - set("IXYZ"):
            raise ValueError(f"{label!r} isn't a {width}-qubit Pauli string")
        if p > 0:
            branches.append((p, pauli_string_matrix(label)))
    return Channel(
        "pauli",
        width,
        tuple(math.sqrt(p) * u for p, u in branches),
        tuple(p for _, p in ordered),
        tuple(branches),
    )


def depolarizing(p: float, num_qubits: int = 1) -> Channel:
    """With probability p the qubits are replaced by the maximally mixed state, i.e. each of the
    4^k - 1 non-identity Pauli strings with probability p / 4^k (noise.py:88-98). A maker's
    average gate error r corresponds to p = r d / (d - 1), d = 2^k (:func:`from_gate_error`)."""
    strings = 4**num_qubits
    _bounded(p, "a depolarising probability", strings / (strings - 1))
    others = ("".join(s) for s in itertools.product("IXYZ", repeat=num_qubits))
    mix = pauli_channel({s: p / strings for s in others if s.strip("I")})
    return Channel("depolarizing", num_qubits, mix.kraus, (p,), mix.mixture, analytic_depolarizing=True)


def from_gate_error(error: float, num_qubits: int) -> Channel:
    """The depolarising channel whose average gate infidelity is ``error`` (noise.py:101-105),
    capped at the fully depolarising p = 4^k / (4^k - 1)."""
    d = 2**num_qubits
    return depolarizing(min(error * d / (d - 1), 4**num_qubits / (4**num_qubits - 1)), num_qubits)


def bit_flip(p: float) -> Channel:
    mix = pauli_channel({"X": _bounded(p, "a bit-flip probability")})
    return Channel("bit_flip", 1, mix.kraus, (p,), mix.mixture)


def phase_flip(p: float) -> Channel:
    mix = pauli_channel({"Z": _bounded(p, "a phase-flip probability")})
    return Channel("phase_flip", 1, mix.kraus, (p,), mix.mixture)


def _damping(name: str, what: str, x: float, lowering: bool) -> Channel:
    """The two Kraus operators of amplitude damping (|1> -> |0> with probability x) or phase
    damping (the |1> component loses coherence with probability x)."""
    x = _bounded(x, what)
    keep = np.diag([1.0, math.sqrt(1 - x)]).astype(np.complex128)
    jump = np.zeros((2, 2), dtype=np.complex128)
    jump[0 if lowering else 1, 1] = math.sqrt(x)
    return Channel(name, 1, (keep, jump), (x,))


def amplitude_damping(gamma: float) -> Channel:
    """Energy relaxation: |1> decays to |0> with probability gamma (noise.py:118-123)."""
    return _damping("amplitude_damping", "gamma", gamma, lowering=True)


def phase_damping(lam: float) -> Channel:
    """Loss of phase coherence without energy loss (noise.py:126-131)."""
    return _damping("phase_damping", "lambda", lam, lowering=False)


def thermal_relaxation(t1: float, t2: float, time: float) -> Channel:
    """Relaxation (T1) and dephasing (T2 <= 2 T1) over ``time``, in one unit (noise.py:134-144).
    gamma = 1 - e^(-t/T1), lambda = max(0, 1 - e^(-2t/T2 + t/T1)); the four Kraus operators are
    the phase-damping ones after the amplitude-damping ones, phase-damping index fastest."""
    if t1 <= 0 or t2 <= 0 or time < 0:
        raise ValueError("T1 and T2 must be positive and the time not negative")
    if t2 > 2 * t1 + 1e-12:
        raise ValueError("T2 can't exceed 2*T1")
    relax = amplitude_damping(1 - math.exp(-time / t1))
    dephase = phase_damping(min(max(0.0, 1 - math.exp(-2 * time / t2 + time / t1)), 1.0))
    kraus = tuple(d @ a for a, d in itertools.product(relax.kraus, dephase.kraus))
    return Channel("thermal_relaxation", 1, kraus, (t1, t2, time))


@dataclass(frozen=True)
class ReadoutError:
    """p01: the chance of reading 1 when the qubit was 0; p10: of reading 0 when it was 1."""

    p01: float
    p10: float

    def __post_init__(self) -> None:
        _bounded(self.p01, "p01")
        _bounded(self.p10, "p10")

    @classmethod
    def symmetric(cls, p: float) -> ReadoutError:
        return cls(p, p)


# ------------------------------------------------------------------------------------ models


@dataclass(frozen=True)
class _Attachment:
    """One rule: ``channel`` after the gates named in ``gates`` (None: any gate but the frame
    changes) acting on exactly the qubits ``qubits`` (None: any qubits)."""

    channel: Channel
    gates: frozenset[str] | None
    qubits: tuple[int, ...] | None

    def rank(self, name: str, width: int, qubit_set: tuple[int, ...]) -> int:
        """-1 when the rule doesn't fit the gate; else 2 for naming qubits plus 1 for naming
        gates (noise.py:169-178)."""
        if self.channel.num_qubits != width:
            return -1
        if self.gates is None:
            if name in IDEAL_GATES:
                return -1
        elif name not in self.gates:
            return -1
        if self.qubits is not None and tuple(sorted(self.qubits)) != qubit_set:
            return -1
        return (2 if self.qubits is not None else 0) + (1 if self.gates is not None else 0)


_UNSEEN = object()


@dataclass
class NoiseModel:
    """Which channel follows which gate, and which readout error each qubit has (noise.py:181).

    The most specific rule wins: a rule for a gate on given qubits, then for any gate on those
    qubits, then for a gate anywhere, then for any gate of that size. Among equally specific
    rules the last one added wins. Frame-change gates (rz, p, u1, id) are error-free unless a
    rule names them.
    """

    rules: list[_Attachment] = field(default_factory=list)
    readout: dict[int | None, ReadoutError] = field(default_factory=dict)
    source: str | None = None
    # lookups answered so far, for the rules as they were then (cleared when a rule is added)
    _answers: dict[tuple[str, tuple[int, ...]], Channel | None] = field(default_factory=dict, repr=False, compare=False)
    _answered_for: int = field(default=0, repr=False, compare=False)

    def add_gate_error(
        self, channel: Channel, gates: str | Iterable[str] | None = None, qubits: Sequence[int] | None = None
    ) -> NoiseModel:
        names = None
        if gates is not None:
            names = frozenset(canonical_name(g) for g in ((gates,) if isinstance(gates, str) else gates))
        if qubits is not None and len(qubits) != channel.num_qubits:
            raise ValueError(f"a {channel.num_qubits}-qubit channel can't attach to {len(qubits)} qubit(s)")
        self.rules.append(_Attachment(channel, names, None if qubits is None else tuple(qubits)))
        return self

    def add_readout_error(self, error: float | ReadoutError, qubits: Iterable[int] | None = None) -> NoiseModel:
        err = error if isinstance(error, ReadoutError) else ReadoutError.symmetric(float(error))
        for q in (None,) if qubits is None else qubits:
            self.readout[q] = err
        return self

    # -- lookups --------------------------------------------------------------------------

    def gate_error(self, name: str, qubits: tuple[int, ...]) -> Channel | None:
        """The channel that follows gate ``name`` (its label, e.g. ``ctrl@x``) on ``qubits``."""
        if self._answered_for != len(self.rules):
            self._answers.clear()
            self._answered_for = len(self.rules)
        key = (name, tuple(qubits))
        hit = self._answers.get(key, _UNSEEN)
        if hit is not _UNSEEN:
            return hit  # type: ignore[return-value]
        qubit_set = tuple(sorted(qubits))
        best: Channel | None = None
        best_rank = -1
        for rule in self.rules:
            r = rule.rank(name, len(qubits), qubit_set)
            if r >= 0 and r >= best_rank:
                best, best_rank = rule.channel, r
        self._answers[key] = best
        return best

    def readout_error(self, qubit: int) -> ReadoutError | None:
        return self.readout.get(qubit, self.readout.get(None))

    def has_gate_errors(self) -> bool:
        return bool(self.rules)

    def is_empty(self) -> bool:
        return not self.rules and not any(e.p01 or e.p10 for e in self.readout.values())

    def describe(self) -> dict[str, Any]:
        """The canonical description a result's noise digest hashes (noise.py:236-246)."""
        rules = []
        for r in self.rules:
            rules.append([None if r.gates is None else sorted(r.gates), list(r.qubits) if r.qubits else None, r.channel.describe()])
        readout = [[-1 if q is None else q, round(e.p01, 15), round(e.p10, 15)] for q, e in self.readout.items()]
        return {"rules": rules, "readout": sorted(readout)}

    def digest(self) -> str:
        text = json.dumps(self.describe(), separators=(",", ":"), sort_keys=True)
        return "sha256:" + hashlib.sha256(text.encode()).hexdigest()

    # -- builders -------------------------------------------------------------------------

    @classmethod
    def uniform(cls, single: float = 0.0, two: float = 0.0, readout: float = 0.0) -> NoiseModel:
        """One average gate error for every one-qubit gate, one for every two-qubit gate, and
        one symmetric readout error (noise.py:253-263); zeros add nothing."""
        model = cls(source="uniform")
        for width, error in ((1, single), (2, two)):
            if error > 0:
                model.add_gate_error(from_gate_error(error, width))
        if readout > 0:
            model.add_readout_error(readout)
        return model

    @classmethod
    def from_calibration(cls, calibration: Mapping[str, Any] | str | Path) -> NoiseModel:
        """A model from a device calibration in Lucivant's twin format (noise.py:265-297): the
        medians ``single_qubit_error``, ``two_qubit_error``, ``readout_error``, then the
        per-qubit ``qubits[]`` and per-coupler ``couplers[]`` values, which outrank them. A
        twin export wrapping the calibration in ``"calibration"`` is unwrapped; missing or null
        fields add no noise."""
        cal: dict[str, Any] = (
            json.loads(Path(calibration).read_text(encoding="utf-8")) if isinstance(calibration, (str, Path)) else dict(calibration)
        )
        inner = cal.get("calibration")
        if isinstance(inner, Mapping):
            cal = dict(inner)
        model = cls(source=str(cal.get("device") or cal.get("profile_key") or "calibration"))
        for width, key in ((1, "single_qubit_error"), (2, "two_qubit_error")):
            if cal.get(key):
                model.add_gate_error(from_gate_error(float(cal[key]), width))
        if cal.get("readout_error"):
            model.add_readout_error(float(cal["readout_error"]))
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


__all__ = [
    "IDEAL_GATES",
    "Channel",
    "NoiseModel",
    "ReadoutError",
    "amplitude_damping",
    "bit_flip",
    "depolarizing",
    "from_gate_error",
    "kron_conj",
    "pauli_channel",
    "phase_damping",
    "phase_flip",
    "thermal_relaxation",
]
