"""Running circuits: :func:`run` for counts; :func:`statevector`, :func:`density_matrix`,
:func:`probabilities`, :func:`outcome_probabilities` and :func:`expectation` for exact answers.

The public functions, their arguments, errors, method selection, seeding, sampling, result
digests and fingerprints follow lucivant-sim 0.3.0 src/lucivant_sim/simulator.py so that a
program, noise model, shot count and seed give the same counts and the same fingerprint on
either implementation (and ``verify`` accepts results saved by the other). What differs is
where the state lives: every engine here is the MI355X device engine of
:mod:`lucivant_sim.engines`, fed by the fusion planner of :mod:`lucivant_sim.plan`, and
trajectories run as one batch of states (:mod:`lucivant_sim.trajectories`).
"""

from __future__ import annotations

import hashlib
import json
import os
import secrets
import threading
import time
from collections import OrderedDict
from collections.abc import Mapping, Sequence
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Literal

import numpy as np
from numpy.typing import NDArray

from lucivant_sim import __version__
from lucivant_sim import _native as nat
from lucivant_sim.circuit import Circuit, IfElse
from lucivant_sim.clifford import clifford_problem
from lucivant_sim.engines import MAX_DENSITY_QUBITS, DensityMatrix, SimulationError, Statevector, marginal
from lucivant_sim.noise import NoiseModel
from lucivant_sim.plan import TILE_BITS, Pass, plan_density, plan_statevector, plan_tiles
from lucivant_sim.qasm import load, parse

Method = Literal[
    "automatic", "statevector", "density_matrix", "trajectories", "stabilizer", "statevector_aer", "stabilizer_stim"
]
METHODS: tuple[str, ...] = (
    "automatic",
    "statevector",
    "density_matrix",
    "trajectories",
    "stabilizer",
    "statevector_aer",
    "stabilizer_stim",
)
AUTO_DENSITY_QUBITS = 10
AUTO_STABILIZER_QUBITS = 12
MAX_SHOTS = 10_000_000

#: Identifies the engine that produced a result. It is reported beside the result, not inside
#: the fingerprint, so results stay verifiable against the reference implementation.
ENGINE = "lucivant-sim-mi355x"

#: Statevector fusion width (qubits per dense pass) when the autotuner has stored no choice for
#: this device. 5 measured best on gfx950 (26 qubits: 52-53 ms per simulation at width 5, 67 ms at
#: 4, 100 ms at 3, with merged blocks and 5-qubit operators on the matrix cores).
FUSION_WIDTH = 5

Program = str | os.PathLike[str] | Circuit | Any


#: The circuits of recently read OpenQASM text and the statevector plans of recently simulated
#: circuits (least recently used dropped first). Repeated calls on one program, such as several
#: expectation() calls, read and plan it once; every call still runs the whole simulation.
_RECENT = 8
_READ: OrderedDict[tuple[str, tuple[tuple[str, float], ...]], Circuit] = OrderedDict()
_PLANS: OrderedDict[tuple[int, int, int, int], tuple[Circuit, list[Pass]]] = OrderedDict()
_RECENT_LOCK = threading.Lock()


def to_circuit(program: Program, params: Mapping[str, float] | None = None) -> Circuit:
    """A Circuit from OpenQASM text, a .qasm path, or an already built Circuit."""
    if isinstance(program, Circuit):
        if params:
            raise ValueError("params apply to OpenQASM source, not to an already-built Circuit")
        return program
    if isinstance(program, os.PathLike):
        return load(Path(program), params)
    if isinstance(program, str):
        key = (program, tuple(sorted((str(k), float(v)) for k, v in (params or {}).items())))
        with _RECENT_LOCK:
            circuit = _READ.get(key)
            if circuit is not None:
                _READ.move_to_end(key)
                return circuit
        circuit = parse(program, params)  # a Circuit is immutable, so callers may share it
        with _RECENT_LOCK:
            _READ[key] = circuit
            while len(_READ) > _RECENT:
                _READ.popitem(last=False)
        return circuit
    module = type(program).__module__
    if module.startswith("qiskit"):
        from qiskit import qasm3  # type: ignore[import-not-found]

        return parse(qasm3.dumps(program), params)
    if module.startswith("cirq"):
        import cirq  # type: ignore[import-not-found]

        return parse(cirq.qasm(program), params)
    raise TypeError(f"can't simulate a {type(program).__name__}; give OpenQASM, a path, or a Qiskit or Cirq circuit")


def _digest(data: Any) -> str:
    return "sha256:" + hashlib.sha256(json.dumps(data, separators=(",", ":"), sort_keys=True).encode()).hexdigest()


@dataclass(frozen=True)
class Result:
    counts: dict[str, int]
    shots: int
    seed: int
    method: str
    num_qubits: int
    num_clbits: int
    circuit_digest: str
    noise_digest: str | None
    version: str
    elapsed_seconds: float
    #: Final state of an ideal static circuit before its measurements (statevector method).
    statevector: NDArray[np.complex128] | None = field(default=None, repr=False, compare=False)
    engine: str = field(default=ENGINE, compare=False)
    #: For ``precision="complex64"`` runs: the precision used and how the run compares with
    #: complex128 (first shots rerun with the same draws). Not part of the fingerprint.
    precision: dict[str, Any] | None = field(default=None, repr=False, compare=False)

    @property
    def fingerprint(self) -> str:
        """Identifies what was run: the same fingerprint gives the same counts."""
        return _digest(
            {
                "engine": "lucivant-sim",
                "version": self.version,
                "circuit": self.circuit_digest,
                "noise": self.noise_digest,
                "shots": self.shots,
                "seed": self.seed,
                "method": self.method,
            }
        )

    @property
    def counts_digest(self) -> str:
        return _digest(self.counts)

    def frequencies(self) -> dict[str, float]:
        return {k: v / self.shots for k, v in self.counts.items()} if self.shots else {}

    def most_frequent(self) -> str | None:
        return max(self.counts, key=lambda k: (self.counts[k], k)) if self.counts else None

    def to_dict(self) -> dict[str, Any]:
        return {
            "engine": "lucivant-sim",
            "version": self.version,
            "method": self.method,
            "shots": self.shots,
            "seed": self.seed,
            "num_qubits": self.num_qubits,
            "num_clbits": self.num_clbits,
            "counts": self.counts,
            "circuit_digest": self.circuit_digest,
            "noise_digest": self.noise_digest,
            "fingerprint": self.fingerprint,
            "counts_digest": self.counts_digest,
            "elapsed_seconds": self.elapsed_seconds,
        }


def choose_method(circuit: Circuit, noise: NoiseModel | None, method: str = "automatic") -> str:
    """The method that runs ``circuit``: the decision table and refusals of simulator.py:157-193."""
    if method not in METHODS:
        raise ValueError(f"method must be one of {', '.join(METHODS)}")
    gate_noise = noise is not None and noise.has_gate_errors()
    dynamic = circuit.is_dynamic()
    if method == "automatic":
        if circuit.num_qubits >= AUTO_STABILIZER_QUBITS and clifford_problem(circuit, noise) is None:
            return "stabilizer"
        if dynamic:
            return "trajectories"
        if gate_noise:
            return "density_matrix" if circuit.num_qubits <= AUTO_DENSITY_QUBITS else "trajectories"
        return "statevector"
    if method == "stabilizer":
        problem = clifford_problem(circuit, noise)
        if problem:
            raise SimulationError(f"the stabilizer method runs Clifford circuits only: {problem}")
        return method
    if method == "stabilizer_stim":
        if noise is not None and not noise.is_empty():
            raise SimulationError("stabilizer_stim runs ideal circuits; use 'stabilizer' for Pauli noise")
        if any(isinstance(op, IfElse) for op in circuit.operations):
            raise SimulationError("stabilizer_stim runs circuits without classical conditions; use 'stabilizer'")
        problem = clifford_problem(circuit, None)
        if problem:
            raise SimulationError(f"the stabilizer methods run Clifford circuits only: {problem}")
        return method
    if method in ("statevector", "statevector_aer") and (gate_noise or dynamic):
        raise SimulationError(
            "the statevector method runs ideal circuits without mid-circuit measurement; "
            "use 'density_matrix' or 'trajectories'"
        )
    if method == "density_matrix" and dynamic:
        raise SimulationError("this circuit uses measurement results mid-circuit; use 'trajectories'")
    if method == "density_matrix" and circuit.num_qubits > MAX_DENSITY_QUBITS:
        raise SimulationError(f"the density-matrix method handles up to {MAX_DENSITY_QUBITS} qubits")
    return method


def _statevector_plan(circuit: Circuit, width: int, low_bits: int, low_width: int, tile: tuple[float, int] | None = None) -> list:
    """The passes of ``circuit`` (tile passes when ``tile`` = (budget, low) is given and the
    state is wide enough, fused dense passes otherwise), planned once per recently simulated
    circuit (an entry holds the circuit itself, so an identity match cannot confuse two
    circuits)."""
    if tile is not None and circuit.num_qubits < TILE_BITS:
        tile = None
    key = (id(circuit), width, low_bits, low_width, tile)
    with _RECENT_LOCK:
        hit = _PLANS.get(key)
        if hit is not None and hit[0] is circuit:
            _PLANS.move_to_end(key)
            return hit[1]
    if tile is not None:
        passes = plan_tiles(circuit.operations, circuit.num_qubits, budget=tile[0], low=tile[1])
    else:
        passes = plan_statevector(circuit.operations, circuit.num_qubits, width, low_bits, low_width)
    with _RECENT_LOCK:
        _PLANS[key] = (circuit, passes)
        while len(_PLANS) > _RECENT:
            _PLANS.popitem(last=False)
    return passes


def _evolve(state: Statevector | DensityMatrix, circuit: Circuit, noise: NoiseModel | None, width: int | None = None) -> None:
    """Static evolution (simulator.py:229-251): gates, their noise, resets of untouched qubits;
    measurements deferred. Runs as fused device passes."""
    _use_switches(state.nbits)
    if isinstance(state, Statevector):
        if width:
            state.run(_statevector_plan(circuit, width, 0, 0))
        else:
            w, low_bits, low_width = _tuned(circuit.num_qubits)
            state.run(_statevector_plan(circuit, w, low_bits, low_width, _tile_choice(circuit.num_qubits)))
        return
    group: list[Pass] = []
    for item in plan_density(circuit.operations, circuit.num_qubits, noise):
        if isinstance(item, Pass):
            group.append(item)
            continue
        state.run(group)
        group = []
        _, (channel, qubits) = item  # type: ignore[misc]
        state.run_kraus(channel, qubits)
    state.run(group)


_TUNED: dict[int, tuple] = {}
_TUNED_LOCK = threading.Lock()


def _tuned_entry(num_qubits: int) -> tuple:
    """The autotuner's stored choice for this size (see autotune.py), read once per size:
    (width, low_bits, low_width, switches, tile) with tile = (budget, low) or None."""
    with _TUNED_LOCK:
        hit = _TUNED.get(num_qubits)
    if hit is None:
        from lucivant_sim import autotune

        try:
            hit = autotune.apply(num_qubits)
        except Exception:  # noqa: BLE001 - no device facts or an unreadable cache: the default
            hit = (0, 0, 0, autotune.DEFAULT_SWITCHES, None)
        with _TUNED_LOCK:
            _TUNED[num_qubits] = hit
    return hit


def _tuned(num_qubits: int) -> tuple[int, int, int]:
    """(width, low_bits, low_width) from the autotuner's stored choice for this size, else
    (FUSION_WIDTH, 0, 0)."""
    w, low_bits, low_width = _tuned_entry(num_qubits)[:3]
    return (w or FUSION_WIDTH, low_bits, low_width)


def _tile_choice(num_qubits: int) -> tuple[float, int] | None:
    """(budget, low) when the autotuner found tile passes faster for this size, else None."""
    tile = _tuned_entry(num_qubits)[4]
    return None if tile is None else (float(tile[0]), int(tile[1]))


def _use_switches(nbits: int) -> None:
    """Set the kernel switches (tiling, matrix cores, streaming) stored for states of ``nbits``
    bits before launching on one. They are process-wide, so a simulation never inherits the
    choice of the size simulated before it."""
    from lucivant_sim import autotune

    autotune.set_switches(_tuned_entry(nbits)[3])


def _sample(
    circuit: Circuit, probs: NDArray[np.float64], shots: int, noise: NoiseModel | None, rng: np.random.Generator
) -> dict[str, int]:
    """Counts from exact probabilities with the run's generator (simulator.py:254-284): inverse
    CDF over the measured qubits' marginal, then readout flips per (qubit, clbit) pair in
    final-measurement order, drawn only for non-zero errors."""
    pairs = circuit.final_measurements()
    if not pairs or shots == 0:
        return {}
    measured = sorted({q for q, _ in pairs})
    return _sample_marginal(circuit, marginal(probs, circuit.num_qubits, measured), shots, noise, rng)


def _sample_state(
    circuit: Circuit, state: Statevector | DensityMatrix, shots: int, noise: NoiseModel | None, rng: np.random.Generator
) -> dict[str, int]:
    """:func:`_sample` of a device state's probabilities, with the marginal over the measured
    qubits formed on the device when they are not every qubit (the same float64 values)."""
    pairs = circuit.final_measurements()
    if not pairs or shots == 0:
        return {}
    measured = sorted({q for q, _ in pairs})
    return _sample_marginal(circuit, state.marginal(measured), shots, noise, rng)


def _sample_marginal(
    circuit: Circuit, dist: NDArray[np.float64], shots: int, noise: NoiseModel | None, rng: np.random.Generator
) -> dict[str, int]:
    """The sampling of :func:`_sample` from the measured qubits' marginal ``dist``."""
    pairs = circuit.final_measurements()
    measured = sorted({q for q, _ in pairs})
    position = {q: i for i, q in enumerate(measured)}
    cumulative = _cumsum(dist)
    draws = rng.random(shots)
    wide = circuit.num_clbits > 62
    flips = noise is not None and any((e := noise.readout_error(q)) is not None and (e.p01 or e.p10) for q, _ in pairs)
    if not wide and not flips and len(dist) <= 1 << 16:
        # no readout draws: the bit pattern is a fixed function of the outcome, tabulated once,
        # and only the tally matters, so the draws are searched in sorted order (much faster)
        cumulative /= cumulative[-1]
        outcomes = np.minimum(np.searchsorted(cumulative, np.sort(draws), side="right"), len(dist) - 1)
        bits = _outcome_table(tuple(pairs), tuple(measured), len(dist))[outcomes]
        values, counts = np.unique(bits, return_counts=True)
        return dict(zip(circuit.format_many(values), counts.tolist()))
    outcomes = np.minimum(_search_normalised(cumulative, draws), len(dist) - 1).astype(np.int64)
    bits = np.zeros(shots, dtype=object if wide else np.int64)
    for q, c in pairs:
        value = (outcomes >> position[q]) & 1
        err = noise.readout_error(q) if noise else None
        if err is not None and (err.p01 or err.p10):
            flip = rng.random(shots) < np.where(value == 1, err.p10, err.p01)
            value = value ^ flip.astype(np.int64)
        if wide:
            bits = bits + np.array([int(v) << c for v in value], dtype=object)
        else:
            bits |= value << c
    values, counts = np.unique(bits, return_counts=True)
    if wide:
        return {circuit.format_bits(int(v)): int(n) for v, n in zip(values, counts)}
    return dict(zip(circuit.format_many(values), counts.tolist()))


_CUMSUM = threading.local()


def _cumsum(dist: NDArray[np.float64]) -> NDArray[np.float64]:
    """``np.cumsum(dist)`` (the same left-to-right sums, bit for bit). Large distributions
    accumulate into a per-thread buffer that is kept for the next call (a fresh 2^26-entry result
    costs more in page faults than in additions), with the library's plain loop: 2^26 entries in
    0.11 s on one host core, against 0.22 s for NumPy's accumulate into the same buffer and
    0.33 s into a fresh array."""
    if len(dist) < 1 << 20:
        return np.cumsum(dist)
    buf = getattr(_CUMSUM, "buf", None)
    if buf is None or len(buf) < len(dist):
        buf = _CUMSUM.buf = np.empty(len(dist), dtype=np.float64)
    out = buf[: len(dist)]
    src = np.ascontiguousarray(dist, dtype=np.float64)
    nat.call("lsim_host_cumsum", src.ctypes.data, out.ctypes.data, len(src))
    return out


def _search_normalised(cumulative: NDArray[np.float64], draws: NDArray[np.float64]) -> NDArray[np.int64]:
    """``searchsorted(cumulative / cumulative[-1], draws, side="right")`` without dividing the whole
    array: x / total is non-decreasing in x, so a binary search that divides only the entries it
    probes (about log2(len) per draw) lands on the same indices, division for division."""
    total = cumulative[-1]
    if len(cumulative) <= 1 << 16 or not total > 0:
        return np.searchsorted(cumulative / total, draws, side="right")
    lo = np.zeros(len(draws), dtype=np.int64)
    hi = np.full(len(draws), len(cumulative), dtype=np.int64)
    while True:
        active = lo < hi
        if not active.any():
            return lo
        mid = (lo + hi) // 2
        below = active & (cumulative[np.minimum(mid, len(cumulative) - 1)] / total <= draws)
        lo = np.where(below, mid + 1, lo)
        hi = np.where(active & ~below, mid, hi)


def _sample_many(circuits: Sequence[Any], probs: NDArray[np.float64], shots: int, noise: NoiseModel | None, seeds: Sequence[int]) -> list[dict[str, int]]:
    """``[_sample(c, probs[j], shots, noise, default_rng(seeds[j])) ...]`` for circuits of one
    width, with the array work done once for the batch where that gives the same numbers: the
    cumulative sums of all rows in one call (each row summed in order, as for one vector), and
    counts tallied by ``bincount`` instead of ``unique`` (the same sorted outcomes). Circuits whose
    clbits depend on readout draws or whose measured set isn't every qubit take ``_sample``."""
    out: list[dict[str, int] | None] = [None] * len(circuits)
    fast: list[tuple[int, tuple]] = []
    for j, c in enumerate(circuits):
        pairs = tuple(c.final_measurements())
        flips = noise is not None and any((e := noise.readout_error(q)) is not None and (e.p01 or e.p10) for q, _ in pairs)
        if shots and pairs and not flips and 0 < c.num_clbits <= 16 and len({q for q, _ in pairs}) == c.num_qubits:
            fast.append((j, pairs))
        else:
            out[j] = _sample(c, probs[j], shots, noise, np.random.default_rng(seeds[j]))
    if fast:
        size = probs.shape[1]
        cumulative = np.cumsum(probs[[j for j, _ in fast]], axis=1)
        cumulative /= cumulative[:, -1:]
        everything = tuple(range(size.bit_length() - 1))
        for row, (j, pairs) in enumerate(fast):
            c = circuits[j]
            u = np.sort(np.random.default_rng(seeds[j]).random(shots))
            outcomes = np.minimum(np.searchsorted(cumulative[row], u, side="right"), size - 1)
            bits = _outcome_table(pairs, everything, size)[outcomes]
            tally = np.bincount(bits, minlength=1 << c.num_clbits)
            values = np.flatnonzero(tally)
            out[j] = dict(zip(c.format_many(values), tally[values].tolist()))
    return out  # type: ignore[return-value]


_OUTCOME_TABLES: dict[tuple, np.ndarray] = {}


def _outcome_table(pairs: tuple[tuple[int, int], ...], measured: tuple[int, ...], size: int) -> np.ndarray:
    """Classical bits for every marginal outcome index (bit position[q] of the index -> clbit c)."""
    key = (pairs, measured, size)
    table = _OUTCOME_TABLES.get(key)
    if table is None:
        position = {q: i for i, q in enumerate(measured)}
        idx = np.arange(size, dtype=np.int64)
        table = np.zeros(size, dtype=np.int64)
        for q, c in pairs:
            table |= ((idx >> position[q]) & 1) << c
        if len(_OUTCOME_TABLES) < 256:
            _OUTCOME_TABLES[key] = table
    return table


def run(
    program: Program,
    shots: int = 1024,
    *,
    seed: int | None = None,
    noise: NoiseModel | None = None,
    method: Method | str = "automatic",
    params: Mapping[str, float] | None = None,
    precision: str = "complex128",
) -> Result:
    """Simulate ``program`` ``shots`` times and return the counts (simulator.py:320-393).

    ``precision="complex64"`` (beyond the reference) runs the trajectory method's states in
    complex64 with float32 arithmetic and float64 reductions and renormalisation; the Result's
    ``precision`` field then states how its shots compare with complex128."""
    if not 0 <= shots <= MAX_SHOTS:
        raise ValueError(f"shots must be between 0 and {MAX_SHOTS:,}")
    if precision not in ("complex128", "complex64"):
        raise ValueError("precision must be 'complex128' or 'complex64'")
    circuit = to_circuit(program, params)
    chosen = choose_method(circuit, noise, method)
    if precision == "complex64" and chosen != "trajectories":
        raise ValueError("precision='complex64' applies to the trajectory method; exact outputs stay complex128")
    seed = secrets.randbelow(2**63) if seed is None else int(seed)
    if seed < 0:
        raise ValueError("seed must not be negative")
    rng = np.random.default_rng(seed)
    started = time.perf_counter()
    vector = None
    if chosen in ("stabilizer", "stabilizer_stim", "statevector_aer"):
        raise SimulationError(
            f"the {chosen!r} method isn't part of the MI355X build; pass method='statevector' "
            "(or 'trajectories' for noisy or dynamic circuits)"
        )
    checked = None
    if chosen == "trajectories":
        from lucivant_sim.trajectories import complex64_check, run_trajectories, shot_values

        if circuit.num_qubits > 0:
            Statevector(circuit.num_qubits)  # the size check before any shot
        if precision == "complex64":
            record: dict = {}
            values, used = shot_values(circuit, noise, shots, rng, dtype=nat.C64, record=record)
            uniq, cnt = np.unique(np.asarray(values, dtype=np.int64), return_counts=True)
            tally = {int(v): int(c) for v, c in zip(uniq, cnt)}
            checked = complex64_check(circuit, noise, seed, values, used, record)
        else:
            tally = run_trajectories(circuit, noise, shots, rng)
        counts = {circuit.format_bits(b): n for b, n in sorted(tally.items())} if circuit.num_clbits else {}
    else:
        state: Statevector | DensityMatrix = (
            Statevector(circuit.num_qubits) if chosen == "statevector" else DensityMatrix(circuit.num_qubits)
        )
        _evolve(state, circuit, noise)
        if isinstance(state, Statevector):
            vector = state.vector()
        counts = _sample_state(circuit, state, shots, noise, rng)
    elapsed = round(time.perf_counter() - started, 6)
    return Result(
        counts=dict(sorted(counts.items())),
        shots=shots,
        seed=seed,
        method=chosen,
        num_qubits=circuit.num_qubits,
        num_clbits=circuit.num_clbits,
        circuit_digest=circuit.digest(),
        noise_digest=noise.digest() if noise is not None and not noise.is_empty() else None,
        version=__version__,
        elapsed_seconds=elapsed,
        statevector=vector,
        precision=checked,
    )


def _static(circuit: Circuit, what: str) -> None:
    if circuit.is_dynamic():
        raise SimulationError(f"{what} needs a circuit without mid-circuit measurement or conditions")


def statevector(program: Program, params: Mapping[str, float] | None = None) -> NDArray[np.complex128]:
    """The ideal final state (measurements ignored), index bit q = qubit q."""
    circuit = to_circuit(program, params)
    _static(circuit, "a statevector")
    state = Statevector(circuit.num_qubits)
    _evolve(state, circuit, None)
    return state.vector()


def density_matrix(
    program: Program, noise: NoiseModel | None = None, params: Mapping[str, float] | None = None
) -> NDArray[np.complex128]:
    """The final mixed state under ``noise`` (measurements ignored)."""
    circuit = to_circuit(program, params)
    _static(circuit, "a density matrix")
    state = DensityMatrix(circuit.num_qubits)
    _evolve(state, circuit, noise)
    return state.matrix()


def _exact_state(circuit: Circuit, noise: NoiseModel | None) -> Statevector | DensityMatrix:
    gate_noise = noise is not None and noise.has_gate_errors()
    state: Statevector | DensityMatrix = DensityMatrix(circuit.num_qubits) if gate_noise else Statevector(circuit.num_qubits)
    _evolve(state, circuit, noise)
    return state


def probabilities(
    program: Program, noise: NoiseModel | None = None, params: Mapping[str, float] | None = None, *, threshold: float = 1e-12
) -> dict[str, float]:
    """Probabilities of every basis state above ``threshold`` (qubit 0 rightmost), before readout
    error; the density matrix when there is gate noise, else the statevector."""
    circuit = to_circuit(program, params)
    _static(circuit, "exact probabilities")
    probs = _exact_state(circuit, noise).probabilities()
    return _basis_dict(probs, circuit.num_qubits, threshold)


def _basis_dict(probs: NDArray[np.float64], n: int, threshold: float) -> dict[str, float]:
    """``{format(i, f"0{n}b"): probs[i] for every i with probs[i] > threshold}`` in index order
    (simulator.py:440-442). The keys are joined from two tables of half-width bit strings, which
    builds a 2^26-entry dict in about half the time of formatting every index."""
    keep = np.flatnonzero(probs > threshold)
    values = probs[keep].tolist()
    if n <= 12:
        return {format(int(i), f"0{n}b"): v for i, v in zip(keep.tolist(), values)}
    low = n // 2
    lo = [format(i, f"0{low}b") for i in range(1 << low)]
    hi = [format(i, f"0{n - low}b") for i in range(1 << (n - low))]
    keys = map(str.__add__, map(hi.__getitem__, (keep >> low).tolist()), map(lo.__getitem__, (keep & ((1 << low) - 1)).tolist()))
    return dict(zip(keys, values))


def outcome_probabilities(
    program: Program, noise: NoiseModel | None = None, params: Mapping[str, float] | None = None, *, threshold: float = 1e-12
) -> dict[str, float]:
    """Probabilities of the classical outcomes, keyed as Result.counts, with readout error
    (simulator.py:445-486)."""
    circuit = to_circuit(program, params)
    _static(circuit, "exact outcome probabilities")
    pairs = circuit.final_measurements()
    if not pairs:
        return {}
    state = _exact_state(circuit, noise)
    measured = sorted({q for q, _ in pairs})
    dist = state.marginal(measured)
    recorded = _recorded_distribution(dist, measured, pairs, noise, threshold)
    return {circuit.format_bits(bits): w for bits, w in recorded}


def _recorded_distribution(
    dist: NDArray[np.float64], measured: list[int], pairs: list[tuple[int, int]], noise: NoiseModel | None, threshold: float
) -> list[tuple[int, float]]:
    """(bits, probability) of every recorded outcome above ``threshold``, in increasing bits, for
    the marginal ``dist`` over ``measured`` (index bit i = measured[i]) and the final (qubit,
    clbit) ``pairs`` (simulator.py:467-486). Each recorded bit is its qubit's bit, flipped with
    probability p01 (a 0) or p10 (a 1) of that qubit's readout error, independently of the other
    recorded bits; entries of ``dist`` at or below ``threshold`` are left out first.

    One contraction per measured qubit: its axis of the [2]*m distribution is replaced by one
    axis per clbit it is recorded in, through the 2 x 2 flip matrix of its readout error, so the
    cost is O(m 2^m) instead of a loop over every outcome and flip pattern."""
    m = len(measured)
    w = np.where(dist > threshold, dist, 0.0).reshape([2] * m)
    clbits: list[int] = []  # the clbit of each read axis, in the order they are appended
    for q in reversed(measured):  # C-order axis 0 is the highest index bit: measured[m-1] first
        err = noise.readout_error(q) if noise is not None else None
        flip = np.eye(2) if err is None else np.array([[1.0 - err.p01, err.p01], [err.p10, 1.0 - err.p10]])
        mine = [c for qq, c in pairs if qq == q]
        factor = np.ones(2)
        for j, _ in enumerate(mine):
            factor = factor[..., None] * flip.reshape((2,) + (1,) * j + (2,))
        w = np.tensordot(w, factor, axes=([0], [0]))
        clbits += mine
    # order the read axes by clbit, highest first, so the flat index packs them in increasing order
    order = sorted(range(len(clbits)), key=lambda a: -clbits[a])
    flat = np.transpose(w, order).reshape(-1) if clbits else w.reshape(-1)
    ascending = sorted(clbits)
    keep = np.flatnonzero(flat > threshold)
    if ascending and ascending[-1] < 63:
        bits = np.zeros(keep.size, dtype=np.int64)
        for i, c in enumerate(ascending):
            bits |= ((keep >> i) & 1) << c
        return list(zip(bits.tolist(), flat[keep].tolist()))
    out = []
    for x, v in zip(keep.tolist(), flat[keep].tolist()):
        out.append((sum(((x >> i) & 1) << c for i, c in enumerate(ascending)), v))
    return out


def expectation(
    program: Program,
    observable: str | Mapping[str, float],
    noise: NoiseModel | None = None,
    params: Mapping[str, float] | None = None,
) -> float:
    """<O> for a Pauli string ("ZZ", qubit 0 rightmost) or a weighted sum {"ZI": 0.5, "XX": -1}.
    Every term is evaluated by one device reduction over the final state (simulator.py:489-527)."""
    circuit = to_circuit(program, params)
    _static(circuit, "an expectation value")
    terms = {observable: 1.0} if isinstance(observable, str) else dict(observable)
    n = circuit.num_qubits
    for label in terms:
        if len(label) != n or set(label.upper()) - set("IXYZ"):
            raise ValueError(f"{label!r} must be a Pauli string of length {n} (qubit 0 rightmost)")
    state = _exact_state(circuit, noise)
    labels = list(terms)
    values = state.expectations(labels) if labels else np.zeros(0)
    total = 0.0
    for label, v in zip(labels, values):
        total += terms[label] * float(v)
    return total


def verify(saved: Mapping[str, Any] | str | Path, program: Program, noise: NoiseModel | None = None) -> bool:
    """Rerun a saved result (``Result.to_dict()``) and check the counts are identical."""
    data = json.loads(Path(saved).read_text(encoding="utf-8")) if isinstance(saved, (str, Path)) else dict(saved)
    again = run(program, int(data["shots"]), seed=int(data["seed"]), noise=noise, method=str(data["method"]))
    if again.circuit_digest != data.get("circuit_digest"):
        return False
    if (again.noise_digest or None) != (data.get("noise_digest") or None):
        return False
    return again.counts_digest == data.get("counts_digest") and again.counts == data.get("counts")
