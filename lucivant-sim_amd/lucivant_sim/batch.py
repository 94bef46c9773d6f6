"""Many circuits at once: every small circuit of a batch in one launch of the LDS program kernel.

:func:`run_many` returns exactly what ``[run(p, shots, seed=s, noise=noise) for p, s in ...]``
returns (same counts, fingerprints and digests; only ``elapsed_seconds`` differs), and
:func:`expectation_many` what ``[expectation(p, obs) ...]`` returns. Circuits that fit the
program kernel (statevector or trajectory method, at most :func:`program.max_qubits` qubits) of
the same width run together, one workgroup per circuit (per shot for trajectories); anything else
runs through the single-circuit path.
"""

from __future__ import annotations

import secrets
import time
from collections.abc import Mapping, Sequence
from typing import Any

import numpy as np

from lucivant_sim import __version__, frontend
from lucivant_sim import _native as nat
from lucivant_sim import program as pg
from lucivant_sim.circuit import Circuit
from lucivant_sim.engines import SimulationError, pauli_masks
from lucivant_sim.frontend import Compiled
from lucivant_sim.noise import NoiseModel
from lucivant_sim.optable import static_table, static_tables
from lucivant_sim.simulator import METHODS, MAX_SHOTS, Result, _sample_many, _static, choose_method, expectation, run, to_circuit


def _result(circuit: Any, counts: dict[str, int], shots: int, seed: int, method: str, noise: NoiseModel | None, t0: float, vector=None) -> Result:  # type: ignore[no-untyped-def]
    digest = circuit.digest if isinstance(circuit, Compiled) else circuit.digest()
    return Result(
        counts=dict(sorted(counts.items())),
        shots=shots,
        seed=seed,
        method=method,
        num_qubits=circuit.num_qubits,
        num_clbits=circuit.num_clbits,
        circuit_digest=digest,
        noise_digest=noise.digest() if noise is not None and not noise.is_empty() else None,
        version=__version__,
        elapsed_seconds=round(time.perf_counter() - t0, 6),
        statevector=vector,
    )


def counts_many(
    circuits: Sequence[Circuit],
    shots: int,
    seeds: Sequence[int],
    noise: NoiseModel | None = None,
    dtype: int = nat.C128,
) -> list[dict[str, int]]:
    """Counts of ideal static circuits of one width (each with its own seed) from one
    LDS-program launch: the counts ``run`` would return, without building Results."""
    if not circuits:
        return []
    table, pool, starts, rows = static_tables(circuits)
    out = pg.run_table(circuits[0].num_qubits, table, pool, starts, rows, dtype=dtype, want_probs=True)
    assert out.probs is not None
    return _sample_many(circuits, out.probs, shots, noise, seeds)


def run_many(
    programs: Sequence[Any],
    shots: int = 1024,
    *,
    seeds: Sequence[int] | None = None,
    noise: NoiseModel | None = None,
    method: str = "automatic",
    params: Mapping[str, float] | None = None,
    dtype: int = nat.C128,
    keep_statevector: bool = False,
) -> list[Result]:
    """``run`` for every program; same-width small circuits share one LDS-program launch. When
    the front-end pool runs (``frontend.start``) and every program is OpenQASM text, the programs
    are read in the pool's worker processes."""
    if not 0 <= shots <= MAX_SHOTS:
        raise ValueError(f"shots must be between 0 and {MAX_SHOTS:,}")
    t0 = time.perf_counter()
    if frontend.active() and programs and all(isinstance(p, str) for p in programs):
        return _run_many_frontend(list(programs), shots, seeds, noise, method, params, dtype, keep_statevector, t0)
    circuits = [to_circuit(p, params) for p in programs]
    seeds = [secrets.randbelow(2**63) for _ in circuits] if seeds is None else [int(s) for s in seeds]
    if len(seeds) != len(circuits) or any(s < 0 for s in seeds):
        raise ValueError("give one non-negative seed per program")
    methods = [choose_method(c, noise, method) for c in circuits]
    results: list[Result | None] = [None] * len(circuits)
    groups: dict[tuple[str, int], list[int]] = {}
    for i, (c, m) in enumerate(zip(circuits, methods)):
        fits = 1 <= c.num_qubits <= pg.max_qubits(dtype)
        if fits and m in ("statevector", "trajectories"):
            groups.setdefault((m, c.num_qubits), []).append(i)
        else:
            results[i] = run(c, shots, seed=seeds[i], noise=noise, method=m)
    for (m, n), idx in groups.items():
        if m == "statevector":
            try:
                table, pool, starts, rows = static_tables([circuits[i] for i in idx])
            except pg.Unsupported:
                # the circuits the program kernel can't take run one by one; the rest together
                keep = []
                for i in idx:
                    try:
                        static_table(circuits[i])
                        keep.append(i)
                    except pg.Unsupported:
                        results[i] = run(circuits[i], shots, seed=seeds[i], noise=noise, method=m)
                idx = keep
                if not idx:
                    continue
                table, pool, starts, rows = static_tables([circuits[i] for i in idx])
            group = [circuits[i] for i in idx]
            out = pg.run_table(n, table, pool, starts, rows, dtype=dtype, want_probs=True, want_state=keep_statevector)
            assert out.probs is not None
            counts = _sample_many(group, out.probs, shots, noise, [seeds[i] for i in idx])
            for j, i in enumerate(idx):
                vec = out.state[j].astype(np.complex128) if out.state is not None else None
                results[i] = _result(circuits[i], counts[j], shots, seeds[i], m, noise, t0, vec)
        else:
            _trajectory_group(circuits, idx, seeds, shots, noise, dtype, results, t0)
    return [r for r in results if r is not None]


def _run_many_frontend(programs, shots, seeds, noise, method, params, dtype, keep_statevector, t0) -> list[Result]:  # type: ignore[no-untyped-def]
    seeds = [secrets.randbelow(2**63) for _ in programs] if seeds is None else [int(s) for s in seeds]
    if len(seeds) != len(programs) or any(s < 0 for s in seeds):
        raise ValueError("give one non-negative seed per program")
    if method not in METHODS:
        raise ValueError(f"method must be one of {', '.join(METHODS)}")
    lds = pg.max_qubits(dtype)
    results: list[Result | None] = [None] * len(programs)
    # chunks small enough that reading (in the workers) and launching, sampling and building
    # Results (here) overlap, large enough that each launch still fills a good part of the GPU
    chunk = max(32, -(-len(programs) // (4 * max(frontend.workers(), 1))))
    for first, compiled in frontend.compile_chunks(programs, dict(params) if params else None, noise, method, lds, chunk):
        for c in compiled:
            if c.error is not None:
                raise c.error
        tabled: dict[int, list[int]] = {}
        traj: dict[int, list[int]] = {}
        for j, c in enumerate(compiled):
            i = first + j
            if c.table is not None:
                tabled.setdefault(c.num_qubits, []).append(j)
            elif c.method == "trajectories" and 1 <= c.num_qubits <= lds:
                traj.setdefault(c.num_qubits, []).append(j)
            else:
                results[i] = run(c.circuit, shots, seed=seeds[i], noise=noise, method=c.method)
        for n, idx in tabled.items():
            out = pg.run_tables(n, [compiled[j].table for j in idx], [compiled[j].pool for j in idx], dtype=dtype, want_probs=True, want_state=keep_statevector)
            counts = _sample_many([compiled[j] for j in idx], out.probs, shots, noise, [seeds[first + j] for j in idx])
            for r, j in enumerate(idx):
                vec = out.state[r].astype(np.complex128) if out.state is not None else None
                results[first + j] = _result(compiled[j], counts[r], shots, seeds[first + j], "statevector", noise, t0, vec)
        if traj:
            circuits = [c.circuit for c in compiled]
            chunk_seeds = seeds[first : first + len(compiled)]
            chunk_results: list[Result | None] = [None] * len(compiled)
            for _n, idx in traj.items():
                _trajectory_group(circuits, idx, chunk_seeds, shots, noise, dtype, chunk_results, t0)
            for j, r in enumerate(chunk_results):
                if r is not None:
                    results[first + j] = r
    return [r for r in results if r is not None]


def trajectory_counts(circuits: Sequence[Any], programs: Sequence[pg.Program], shots: int, seeds: Sequence[int], dtype: int = nat.C128) -> list[dict[str, int]]:
    """Counts of every circuit's ``shots`` trajectories from one LDS-program launch, one
    workgroup per shot. ``programs`` come from ``program.compile_trajectory``; each circuit's
    draws are its own seeded stream in the reference's order (``rng.random(shots * D)``)."""
    if not programs or shots == 0:
        return [{} for _ in programs]
    width = max(p.draws for p in programs)
    draws = np.zeros((len(programs) * shots, max(width, 1)))
    for j, (p, seed) in enumerate(zip(programs, seeds)):
        if p.draws:
            draws[j * shots : (j + 1) * shots, : p.draws] = np.random.default_rng(seed).random(shots * p.draws).reshape(shots, p.draws)
    out = pg.run_programs(list(programs), repeat=shots, dtype=dtype, draws=draws, want_clbits=True)
    assert out.clbits is not None
    counts = []
    for j, c in enumerate(circuits):
        bits = out.clbits[j * shots : (j + 1) * shots]
        uniq, cnt = np.unique(bits, return_counts=True)
        counts.append(dict(zip(c.format_many(uniq), cnt.tolist())) if c.num_clbits else {})
    return counts


def _trajectory_group(circuits, idx, seeds, shots, noise, dtype, results, t0) -> None:  # type: ignore[no-untyped-def]
    progs, ok = [], []
    for i in idx:
        try:
            progs.append(pg.compile_trajectory(circuits[i], noise))
            ok.append(i)
        except pg.Unsupported:
            results[i] = run(circuits[i], shots, seed=seeds[i], noise=noise, method="trajectories")
    counts = trajectory_counts([circuits[i] for i in ok], progs, shots, [seeds[i] for i in ok], dtype)
    for i, cnt in zip(ok, counts):
        results[i] = _result(circuits[i], cnt, shots, seeds[i], "trajectories", noise, t0)


def expectation_many(
    programs: Sequence[Any],
    observables: Sequence[str | Mapping[str, float]],
    *,
    params: Mapping[str, float] | None = None,
    dtype: int = nat.C128,
) -> list[float]:
    """``expectation`` of each ideal program's observable; small circuits share one launch."""
    circuits = [to_circuit(p, params) for p in programs]
    if len(observables) != len(circuits):
        raise ValueError("give one observable per program")
    out: list[float | None] = [None] * len(circuits)
    groups: dict[tuple[int, int], list[int]] = {}
    terms: list[dict[str, float]] = []
    for i, (c, obs) in enumerate(zip(circuits, observables)):
        _static(c, "an expectation value")
        t = {obs: 1.0} if isinstance(obs, str) else dict(obs)
        for label in t:
            if len(label) != c.num_qubits or set(label.upper()) - set("IXYZ"):
                raise ValueError(f"{label!r} must be a Pauli string of length {c.num_qubits} (qubit 0 rightmost)")
        terms.append(t)
        if 1 <= c.num_qubits <= pg.max_qubits(dtype) and 0 < len(t) <= nat.MAX_TERMS:
            groups.setdefault((c.num_qubits, len(t)), []).append(i)
        else:
            out[i] = expectation(c, t)
    for (_n, _nt), idx in groups.items():
        progs = [pg.compile_static(circuits[i]) for i in idx]
        masks = [[pauli_masks(lbl) for lbl in terms[i]] for i in idx]
        res = pg.run_programs(progs, dtype=dtype, pauli=masks)
        assert res.expect is not None
        for j, i in enumerate(idx):
            total = 0.0
            for w, v in zip(terms[i].values(), res.expect[j]):
                total += w * float(v)
            out[i] = total
    return [float(v) for v in out]  # type: ignore[arg-type]


__all__ = ["SimulationError", "expectation_many", "run_many"]
