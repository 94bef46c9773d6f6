"""Front-end worker processes: the host-side reading of circuits, in parallel.

For a batch of small circuits the device side is one LDS-program launch; what remains per
circuit is host work in Python: reading the OpenQASM text, checking the circuit, choosing its
method, hashing it for the result digest, and building its op table. :func:`start` forks a pool
of worker processes that do this work for circuits given as text (workers use NumPy only and
never touch the GPU); :func:`compile_texts` returns, per program, a :class:`Compiled` summary that
carries everything the launching process needs: the op table and matrix pool for the LDS program
kernel, the measured (qubit, clbit) pairs and register layout for sampling, and the digest.
``batch.run_many`` and ``serving.Server`` use the pool when it is running.

:func:`start` also forks trajectory workers: processes that run wide trajectory requests for the
server on the GPU, each in its own interpreter. A trajectory batch's host side (choices, fused
per-combination operators, shared-state bookkeeping) is Python; on a thread of the serving process
it competes for one interpreter lock with the scheduler, and at saturation it was the slowest
path through the whole stream.

The pools are forked, so they must be started before this process first uses the GPU (children
must not inherit a live device context): :func:`start` returns 0 and starts nothing when the GPU
is already in use, and callers then keep the in-process path.
"""

from __future__ import annotations

import multiprocessing
import os
from concurrent.futures import Future, ProcessPoolExecutor
from concurrent.futures.process import BrokenProcessPool
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from lucivant_sim.circuit import Circuit, Register

_POOL: Any = None
_WORKERS = 0
_TRAJ_POOL: Any = None
_TRAJ_WORKERS = 0


@dataclass
class Compiled:
    """One program as read by a worker. ``table``/``pool`` are set for ideal static circuits the
    LDS program kernel can run; ``passes`` for wider statevectors and ``program`` (an LDS
    trajectory program) for small noisy or dynamic circuits, when fusion widths were given;
    ``circuit`` when the launching process needs the whole circuit (dynamic circuits, noisy runs,
    wide states otherwise); ``terms`` are an expectation request's checked Pauli terms; ``error``
    holds the exception reading the program raised, re-raised by the caller."""

    num_qubits: int = 0
    num_clbits: int = 0
    cregs: tuple[Register, ...] = ()
    pairs: tuple[tuple[int, int], ...] = ()
    digest: str = ""
    method: str = ""
    table: np.ndarray | None = None
    pool: np.ndarray | None = None
    circuit: Circuit | None = None
    terms: dict[str, float] | None = None
    passes: list | None = None
    program: Any = None
    error: BaseException | None = None
    _shell: Circuit | None = field(default=None, repr=False)

    # the circuit interface that sampling and results use
    def final_measurements(self) -> list[tuple[int, int]]:
        return list(self.pairs)

    def _formatter(self) -> Circuit:
        if self._shell is None:
            self._shell = Circuit(self.num_qubits, self.num_clbits, (), (), self.cregs)
        return self._shell

    def format_many(self, values: np.ndarray) -> list[str]:
        return self._formatter().format_many(values)

    def format_bits(self, bits: int) -> str:
        return self._formatter().format_bits(bits)


def default_workers() -> int:
    """Worker count: the CPUs this process may run on, less one for the launching process, at
    most 15."""
    try:
        cpus = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover - non-Linux
        cpus = os.cpu_count() or 2
    env = os.environ.get("LSIM_FRONTEND_WORKERS")
    if env:
        return max(1, int(env))
    return max(1, min(15, cpus - 1))


def default_trajectory_workers() -> int:
    """Trajectory worker count: ``LSIM_TRAJECTORY_WORKERS``, else 2."""
    return max(0, int(os.environ.get("LSIM_TRAJECTORY_WORKERS", "2")))


def start(workers: int | None = None, trajectory_workers: int | None = None) -> int:
    """Start the worker pool and the trajectory workers; returns the pool's size (0: not started
    because the GPU is in use)."""
    global _POOL, _WORKERS, _TRAJ_POOL, _TRAJ_WORKERS
    if _POOL is not None:
        return _WORKERS
    import torch

    from lucivant_sim import _native as nat

    if torch.cuda.is_initialized() or (nat.loaded() and torch.cuda.is_available()):
        return 0
    n = workers or default_workers()
    ctx = multiprocessing.get_context("fork")
    _POOL = ctx.Pool(n)
    _WORKERS = n
    t = default_trajectory_workers() if trajectory_workers is None else trajectory_workers
    # (counting devices does not open one)
    if t > 0 and torch.cuda.device_count() > 0:
        _TRAJ_POOL = ProcessPoolExecutor(t, mp_context=ctx, initializer=_trajectory_init)
        # the executor forks all its workers at the first submission: make that happen now,
        # before this process opens the GPU
        _TRAJ_POOL.submit(int).result()
        _TRAJ_WORKERS = t
    return n


def stop() -> None:
    global _POOL, _WORKERS, _TRAJ_POOL, _TRAJ_WORKERS
    if _POOL is not None:
        _POOL.close()
        _POOL.join()
    if _TRAJ_POOL is not None:
        _TRAJ_POOL.shutdown(wait=True)
    _POOL, _WORKERS, _TRAJ_POOL, _TRAJ_WORKERS = None, 0, None, 0


def active() -> bool:
    return _POOL is not None


def workers() -> int:
    return _WORKERS


def trajectory_workers() -> int:
    return _TRAJ_WORKERS


def compile_one(
    src: Any,
    params: dict | None,
    noise: Any,
    method: str,
    lds_qubits: int,
    observable: Any = None,
    widths: dict[int, tuple[int, int, int]] | None = None,
    digest: bool = True,
) -> Compiled:
    """Read one program (runs in a worker; also callable in-process).

    ``observable`` set: an expectation request (labels checked here; the statevector method, or
    the density matrix when ``noise`` has gate errors, as ``expectation`` chooses).
    ``widths`` set: statevectors wider than ``lds_qubits`` also get their passes, planned with
    ``widths[num_qubits]`` = (fusion width, low bits, low width[, tile]): tile passes when
    ``tile`` = (budget, low) is given, fused dense passes otherwise. ``digest`` False skips the
    circuit digest (callers that build no Result)."""
    from lucivant_sim._native import MAX_TERMS
    from lucivant_sim.engines import SimulationError
    from lucivant_sim.optable import Unsupported, static_table
    from lucivant_sim.plan import TILE_BITS, plan_statevector, plan_tiles
    from lucivant_sim.simulator import choose_method, to_circuit

    try:
        c = to_circuit(src, params)
        terms = None
        if observable is not None:
            # the checks of expectation(), in its order (simulator.py:496-502)
            if c.is_dynamic():
                raise SimulationError("an expectation value needs a circuit without mid-circuit measurement or conditions")
            terms = {observable: 1.0} if isinstance(observable, str) else {str(k): float(v) for k, v in dict(observable).items()}
            for label in terms:
                if len(label) != c.num_qubits or set(label.upper()) - set("IXYZ"):
                    raise ValueError(f"{label!r} must be a Pauli string of length {c.num_qubits} (qubit 0 rightmost)")
            # as expectation() does (simulator.py:503-516): the density matrix under gate noise
            chosen = "density_matrix" if noise is not None and noise.has_gate_errors() else "statevector"
        else:
            chosen = choose_method(c, noise, method)
        out = Compiled(
            num_qubits=c.num_qubits,
            num_clbits=c.num_clbits,
            cregs=c.cregs,
            digest=c.digest() if digest else "",
            method=chosen,
            terms=terms,
        )
        if chosen != "statevector":
            out.circuit = c
            if chosen == "trajectories" and widths is not None and 1 <= c.num_qubits <= lds_qubits:
                # the server runs small trajectory requests as LDS programs
                from lucivant_sim import program as pg

                try:
                    out.program = pg.compile_trajectory(c, noise)
                except pg.Unsupported:
                    pass
            return out
        out.pairs = tuple(c.final_measurements())
        # the program kernel takes 1..MAX_TERMS Pauli strings per instance; other term counts
        # (none, or more) are evaluated over a state in HBM in chunks of MAX_TERMS
        lds_terms = terms is None or 0 < len(terms) <= MAX_TERMS
        if 1 <= c.num_qubits <= lds_qubits and lds_terms:
            try:
                out.table, out.pool = static_table(c)
                return out
            except Unsupported:
                pass
        if widths is not None:
            w, low_bits, low_width, *rest = widths.get(c.num_qubits, (4, 0, 0))
            tile = rest[0] if rest else None
            if tile is not None and c.num_qubits >= TILE_BITS:
                out.passes = plan_tiles(c.operations, c.num_qubits, budget=tile[0], low=tile[1])
            else:
                out.passes = plan_statevector(c.operations, c.num_qubits, w, low_bits, low_width)
        else:
            out.circuit = c
        return out
    except Exception as exc:  # noqa: BLE001 - handed back to the caller, raised there in order
        return Compiled(error=exc)


def _job(args: tuple) -> Compiled:
    return compile_one(*args)


def compile_texts(sources: list[str], params: dict | None, noise: Any, method: str, lds_qubits: int) -> list[Compiled]:
    """Read many programs in the workers (results in program order)."""
    if _POOL is None:
        raise RuntimeError("the front-end pool isn't running")
    jobs = [(s, params, noise, method, lds_qubits) for s in sources]
    chunk = max(1, len(jobs) // (4 * _WORKERS))
    return _POOL.map(_job, jobs, chunksize=chunk)


def _chunk_job(args: tuple) -> list[Compiled]:
    sources, rest = args
    return [compile_one(src, *rest) for src in sources]


def compile_chunks(sources: list[str], params: dict | None, noise: Any, method: str, lds_qubits: int, chunk: int):  # type: ignore[no-untyped-def]
    """Read many programs in the workers, ``chunk`` per task; yields ``(first index, [Compiled])``
    in program order as chunks complete, so the caller can launch early chunks while the workers
    read later ones."""
    if _POOL is None:
        raise RuntimeError("the front-end pool isn't running")
    rest = (params, noise, method, lds_qubits)
    tasks = [(sources[i : i + chunk], rest) for i in range(0, len(sources), chunk)]
    for j, part in enumerate(_POOL.imap(_chunk_job, tasks)):
        yield j * chunk, part


def submit(*args: Any) -> Any:
    """:func:`compile_one` of ``args`` in a worker, asynchronously: an ``AsyncResult`` whose
    ``get()`` is the :class:`Compiled`."""
    if _POOL is None:
        raise RuntimeError("the front-end pool isn't running")
    return _POOL.apply_async(_job, (args,))


# ------------------------------------------------------------------------------- trajectory workers

#: a trajectory worker's own stream (set up by its initializer)
_WORKER_STREAM: Any = None
#: cached device memory a trajectory worker keeps between requests (bytes); beyond it the
#: worker returns its cache, so memory the server admitted by footprint is free again
WORKER_CACHE_BYTES = 4 << 30


def _trajectory_init() -> None:
    """In a new trajectory worker: open the device and load the library, so the first request
    does not pay for it."""
    global _WORKER_STREAM
    import torch

    from lucivant_sim import _native as nat

    try:
        nat.load()
        _WORKER_STREAM = torch.cuda.Stream()
    except Exception:  # noqa: BLE001 - an initializer that raises breaks the whole executor;
        # the first request then raises the same error to its requester instead
        _WORKER_STREAM = None


def _trajectory_job(circuit: Circuit, noise: Any, shots: int, seed: Any, dtype: int) -> dict[str, int]:
    import torch

    from lucivant_sim.trajectories import run_trajectories

    if _WORKER_STREAM is None:
        _trajectory_init()
    try:
        with torch.cuda.stream(_WORKER_STREAM):
            tally = run_trajectories(circuit, noise, shots, np.random.default_rng(seed), dtype=dtype)
    finally:
        if torch.cuda.memory_reserved() > WORKER_CACHE_BYTES:
            torch.cuda.empty_cache()
    return {circuit.format_bits(b): k for b, k in sorted(tally.items())} if circuit.num_clbits else {}


def submit_trajectories(circuit: Circuit, noise: Any, shots: int, seed: Any, dtype: int) -> Future:
    """A wide trajectory request's counts (formatted bit strings), computed by a trajectory
    worker: the same ``run_trajectories`` call the serving process would make, so the counts are
    the same. If a worker has died, the requests it held fail with ``BrokenProcessPool``, and this
    raises it too and reports no trajectory workers from then on (callers then run trajectories
    themselves)."""
    global _TRAJ_WORKERS
    if _TRAJ_POOL is None or _TRAJ_WORKERS == 0:
        raise RuntimeError("the trajectory workers aren't running")
    try:
        return _TRAJ_POOL.submit(_trajectory_job, circuit, noise, shots, seed, dtype)
    except BrokenProcessPool:
        _TRAJ_WORKERS = 0
        raise
