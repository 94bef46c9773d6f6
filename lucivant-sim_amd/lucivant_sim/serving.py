"""A serving loop with continuous batching of circuit requests on one MI355X.

Requests (an expectation of a Pauli sum, or seeded counts, of one OpenQASM program) arrive at any
time. The scheduler runs in ticks; at every tick it

1. admits waiting requests in arrival order while their HBM footprint fits the budget
   (statevector: 16 B x 2^n; trajectories: shots x 16 B x 2^n, chunked), so requests that arrive
   while others are on the GPU join as memory frees up;
2. hands every admitted request of at most ``program.max_qubits`` qubits that hasn't run yet to
   the small-batch thread, which runs them grouped by width as one LDS-program launch each (one
   workgroup per circuit or per shot) on a stream of its own, waits for the outputs and samples
   the counts, while the scheduler goes on launching;
3. advances resident wide requests by up to ``chunk_passes`` fused passes each on their own
   HIP streams (states stay resident in HBM between ticks), so wide circuits overlap one another
   and small batches slip in between their chunks. At most ``max_launching`` wide requests are
   being launched at once, started in priority order: cheap ones first, then count requests by
   decreasing width (their host sampling of 2^n probabilities then overlaps the device work of
   the rest), then the others in arrival order;
4. completes finished requests: probabilities or expectations come back to the host, counts are
   sampled with the request's own seeded generator exactly as ``run`` samples them.

Noisy count requests wider than an LDS program run as trajectory batches, whose host side is
Python between launches: on the trajectory workers that ``frontend.start`` forks (processes of
their own, so that work does not contend for this process's interpreter lock) when they run,
else on a thread of this process, one request at a time.

A request's result doesn't depend on what else is in flight: each circuit runs in its own
workgroup or its own state buffer with the same kernels, so a result alone and in the mix is the
same (tests/test_gpu_serving.py checks it). Latency is measured from arrival to result.
"""

from __future__ import annotations

import itertools
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Any, Mapping

import numpy as np
import torch

from lucivant_sim import _native as nat
from lucivant_sim import frontend
from lucivant_sim import program as pg
from lucivant_sim.engines import MAX_DENSITY_QUBITS, MAX_STATEVECTOR_QUBITS, SimulationError, Statevector, launch_expectations, pauli_masks, plan_launches, ptr, require_device
from lucivant_sim.frontend import Compiled, compile_one
from lucivant_sim.noise import NoiseModel
from lucivant_sim.simulator import FUSION_WIDTH, _sample, _sample_many, _tile_choice, _tuned, _use_switches

_ids = itertools.count()


@dataclass(eq=False)  # identity: the scheduler's membership tests must not compare fields
class Request:
    """One unit of work. ``observable`` set: an expectation; otherwise counts of ``shots``."""

    program: Any
    observable: str | Mapping[str, float] | None = None
    shots: int = 1000
    seed: int = 0
    noise: NoiseModel | None = None
    method: str = "automatic"
    id: int = field(default_factory=lambda: next(_ids))
    arrival: float = 0.0
    # filled in by the server
    done: float = 0.0
    result: Any = None
    error: BaseException | None = None
    event: threading.Event = field(default_factory=threading.Event, repr=False)
    # the program as read (in a front-end worker when the pool runs, else in ``submit``)
    _reading: Any = field(default=None, repr=False)
    _compiled: Compiled | None = field(default=None, repr=False)
    #: when the request was read, admitted and first launched, and its outputs queued
    stamps: dict[str, float] = field(default_factory=dict, repr=False)

    def wait(self, timeout: float | None = None) -> Any:
        self.event.wait(timeout)
        if self.error is not None:
            raise self.error
        return self.result

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass(eq=False)
class _Job:
    req: Request
    c: Compiled
    method: str
    footprint: int
    terms: dict[str, float] | None = None
    # wide statevector requests: resident state and the remaining passes
    state: Statevector | None = None
    passes: list | None = None
    next_pass: int = 0
    stream: torch.cuda.Stream | None = None
    mats: Any = None
    calls: list[tuple[Any, tuple]] | None = None
    # outputs queued behind the last pass: (event, device tensor), read once the event completes
    finishing: tuple[torch.cuda.Event, torch.Tensor, Any] | None = None
    seq: int = 0  # admission order
    # wide trajectory requests: their run on a trajectory worker (process or thread); small
    # requests: the batch run on the small-batch thread that carries them
    future: Any = None


class Server:
    """Continuous-batching scheduler; ``submit`` from any thread, ``start``/``stop`` the loop
    (or drive it with ``run_until_idle``)."""

    def __init__(
        self,
        hbm_budget: int = 64 << 30,
        streams: int = 4,
        chunk_passes: int = 64,
        dtype: int = nat.C128,
        fusion_width: int | None = None,
        max_launching: int = 2,
    ) -> None:
        require_device()
        self.budget = hbm_budget
        self.max_launching = max(1, max_launching)
        self._seq = itertools.count()
        self.used = 0
        self.dtype = dtype
        self.width = fusion_width or FUSION_WIDTH
        self.chunk = chunk_passes
        self.streams = [torch.cuda.Stream() for _ in range(max(1, streams))]
        # small batches (LDS programs) run on a stream of their own, beside the wide requests
        self.small_stream = torch.cuda.Stream()
        self._rr = 0
        self.inbox: queue.Queue[Request] = queue.Queue()
        self.reading: list[Request] = []
        self.waiting: list[_Job] = []
        self.resident: list[_Job] = []
        self.completed: list[Request] = []
        self.max_lds = pg.max_qubits(dtype)
        # fusion choice per width for the passes of wide requests (planned where they are read)
        self.widths = {n: ((self.width, 0, 0, None) if fusion_width else (*_tuned(n), _tile_choice(n))) for n in range(self.max_lds + 1, 34)}
        self._stop = threading.Event()
        self._thread: threading.Thread | None = None
        # host sampling of wide count requests (a cumulative sum over 2^n probabilities) runs
        # beside the scheduler, so launches continue meanwhile
        self._sampler = ThreadPoolExecutor(max_workers=4, thread_name_prefix="lsim-sample")
        # trajectory batches wider than an LDS program (host-driven: they read reduced density
        # matrices back between launches) run on the front end's trajectory worker processes
        # when they run, else one at a time on a thread and stream of their own, so the
        # scheduler keeps launching the other requests meanwhile
        self._trajectories = ThreadPoolExecutor(max_workers=1, thread_name_prefix="lsim-traj")
        self._traj_stream = torch.cuda.Stream()
        # small batches (LDS programs: tables, draws, the launch, waiting for its outputs and
        # sampling) run on a thread of their own too, one batch at a time on small_stream
        self._small = ThreadPoolExecutor(max_workers=1, thread_name_prefix="lsim-small")
        self._small_runs: list[Any] = []
        self._lock = threading.Lock()
        self._sampling: list[Any] = []
        self.ticks = 0
        self.peak_used = 0
        #: scheduler-thread time per phase (admission, handing small batches to their thread,
        #: wide launches, completion)
        self.host_seconds = {"admit": 0.0, "small": 0.0, "wide_launch": 0.0, "wide_outputs": 0.0, "wide_finish": 0.0, "copy_out": 0.0}

    # -- submission ---------------------------------------------------------------------

    def submit(self, req: Request) -> Request:
        """Queue a request. Its program is read (parsed, checked, tabulated or planned) in a
        front-end worker process when the pool runs, else here in the caller's thread; either way
        that host work overlaps what the scheduler has on the GPU."""
        if not req.arrival:
            req.arrival = time.perf_counter()
        args = (req.program, None, req.noise, req.method, self.max_lds, req.observable, self.widths, False)
        if frontend.active() and isinstance(req.program, str):
            req._reading = frontend.submit(*args)
        else:
            req._compiled = compile_one(*args)
        self.inbox.put(req)
        return req

    def start(self) -> None:
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()

    def stop(self) -> None:
        self._stop.set()
        if self._thread is not None:
            self._thread.join()
        for job in list(self.resident):
            if job.future is not None:
                job.future.exception()  # wait; the error, if any, was reported to the requester
        for f in self._small_runs:
            f.result()
        self._small_runs = []
        for f in self._sampling:
            f.result()
        self._sampling = []

    def _loop(self) -> None:
        while not self._stop.is_set():
            if not self.tick():
                time.sleep(0.0005)

    def run_until_idle(self) -> None:
        while True:
            worked = self.tick()
            if not (worked or self.reading or self.waiting or self.resident or not self.inbox.empty()):
                break
            if not worked:
                time.sleep(0.0005)
        for f in self._small_runs:
            f.result()
        self._small_runs = []
        for f in self._sampling:
            f.result()
        self._sampling = []

    # -- scheduling ---------------------------------------------------------------------

    def _prepare(self, req: Request, c: Compiled) -> _Job:
        if c.method not in ("statevector", "trajectories", "density_matrix"):
            raise ValueError(f"the server runs statevector, density-matrix and trajectory requests, not {c.method!r}")
        # the engines' own size guards and messages (engines.py:29-31)
        if c.method == "density_matrix" and c.num_qubits > MAX_DENSITY_QUBITS:
            raise SimulationError(
                f"{c.num_qubits} qubits need more memory than the density-matrix engine allows "
                f"(up to {MAX_DENSITY_QUBITS}); use method='trajectories'"
            )
        if c.num_qubits > MAX_STATEVECTOR_QUBITS:
            raise SimulationError(
                f"{c.num_qubits} qubits need more memory than the statevector engine allows (up to {MAX_STATEVECTOR_QUBITS})"
            )
        size = 16 << c.num_qubits
        if c.method == "density_matrix":
            size = 16 << (2 * c.num_qubits)
        footprint = size * (req.shots if c.method == "trajectories" and c.num_qubits > self.max_lds else 1)
        return _Job(req, c, c.method, footprint, c.terms, seq=next(self._seq))

    def _admit(self) -> bool:
        while True:
            try:
                self.reading.append(self.inbox.get_nowait())
            except queue.Empty:
                break
        unread = []
        for req in self.reading:
            if req._compiled is None:
                if not req._reading.ready():
                    unread.append(req)
                    continue
                req._compiled = req._reading.get()
                req._reading = None
            c = req._compiled
            req.stamps.setdefault("read", time.perf_counter())
            try:
                if c.error is not None:
                    raise c.error
                self.waiting.append(self._prepare(req, c))
            except BaseException as exc:  # noqa: BLE001 - reported to the requester
                self._finish_error(req, exc)
        moved = len(unread) < len(self.reading)
        self.reading = unread
        still = []
        for job in self.waiting:
            if self.used + job.footprint <= self.budget or not self.resident:
                self.used += job.footprint
                self.peak_used = max(self.peak_used, self.used)
                self.resident.append(job)
                job.req.stamps["admitted"] = time.perf_counter()
                moved = True
            else:
                still.append(job)
        self.waiting = still
        return moved

    def tick(self) -> bool:
        """One scheduling round; True if it did any work (waiting on the device or on the
        trajectory thread is not work: the caller then sleeps briefly, which lets the host
        threads that do the waited-for work take the interpreter)."""
        t0 = time.perf_counter()
        admitted = self._admit()
        t1 = time.perf_counter()
        self.host_seconds["admit"] += t1 - t0
        if not self.resident:
            return admitted
        self.ticks += 1
        progress = admitted
        small: list[_Job] = []
        wide: list[_Job] = []
        for j in self.resident:
            (small if j.state is None and (j.c.table is not None or j.c.program is not None) else wide).append(j)
        fresh = [j for j in small if j.future is None]
        if fresh:
            run = self._small.submit(self._run_small_on_stream, fresh)
            for j in fresh:
                j.future = run
            self._small_runs.append(run)
            progress = True
        self._small_runs = [f for f in self._small_runs if not f.done()]
        t2 = time.perf_counter()
        self.host_seconds["small"] += t2 - t1
        launching = sum(1 for j in wide if j.state is not None and j.passes is not None and j.next_pass < len(j.passes))
        for job in sorted(wide, key=self._priority):
            if job.finishing is not None:
                continue
            if job.state is None and job.method == "statevector":
                if launching >= self.max_launching:
                    continue
                launching += 1
            progress |= self._advance_wide(job)
        t3 = time.perf_counter()
        self.host_seconds["wide_launch"] += t3 - t2
        for job in wide:
            if job.finishing is None and job.passes is not None and job.next_pass >= len(job.passes) and job.state is not None:
                progress = True
                try:
                    self._queue_outputs(job)
                except BaseException as exc:  # noqa: BLE001 - e.g. no memory for the outputs
                    self._finish_error(job.req, exc)
                    self._release(job)
        t4 = time.perf_counter()
        self.host_seconds["wide_outputs"] += t4 - t3
        for job in wide:
            if job.finishing is not None and job.finishing[0].query():
                progress = True
                self._complete_wide(job)
        self._sampling = [f for f in self._sampling if not f.done()]
        self.host_seconds["wide_finish"] += time.perf_counter() - t4
        return progress

    @staticmethod
    def _priority(job: _Job) -> tuple:
        """Start order of wide requests: cheap ones (under about 16 passes' worth of a 24-qubit
        state) first, since they finish at once; then count requests, widest first, whose host
        sampling is longest and should overlap the rest; then the others by admission."""
        n = job.c.num_qubits
        work = len(job.c.passes or ()) << n
        count = job.terms is None and job.method == "statevector"
        return (0 if work < (16 << 24) else 1, -(1 << n) if count else 0, job.seq)

    def _run_small_on_stream(self, jobs: list[_Job]) -> None:
        with torch.cuda.stream(self.small_stream):
            self._run_small(jobs)

    def _run_small(self, jobs: list[_Job]) -> None:
        by_key: dict[tuple, list[_Job]] = {}
        for j in jobs:
            kind = j.method if j.terms is None else "expect"
            key = (j.c.num_qubits, kind, len(j.terms or ()), j.req.shots, id(j.req.noise) if kind != "expect" else 0)
            by_key.setdefault(key, []).append(j)
        for (n, kind, _nterms, shots, _noise), group in by_key.items():
            try:
                if kind == "expect":
                    masks = [[pauli_masks(lbl) for lbl in j.terms] for j in group]  # type: ignore[union-attr]
                    out = pg.run_tables(n, [j.c.table for j in group], [j.c.pool for j in group], dtype=self.dtype, pauli=masks)
                    for idx, j in enumerate(group):
                        total = 0.0
                        for w, v in zip(j.terms.values(), out.expect[idx]):  # type: ignore[union-attr,index]
                            total += w * float(v)
                        self._finish(j, total)
                elif kind == "statevector":
                    out = pg.run_tables(n, [j.c.table for j in group], [j.c.pool for j in group], dtype=self.dtype, want_probs=True)
                    counts = _sample_many([j.c for j in group], out.probs, shots, group[0].req.noise, [j.req.seed for j in group])
                    for j, cnt in zip(group, counts):
                        self._finish(j, cnt)
                else:
                    from lucivant_sim.batch import trajectory_counts

                    counts = trajectory_counts([j.c for j in group], [j.c.program for j in group], shots, [j.req.seed for j in group], self.dtype)
                    for j, cnt in zip(group, counts):
                        self._finish(j, cnt)
            except BaseException as exc:  # noqa: BLE001
                for j in group:
                    self._finish_error(j.req, exc)
                    self._release(j)

    def _advance_wide(self, job: _Job) -> bool:
        """Advance one wide request; True if it launched or finished anything."""
        if job.method == "density_matrix":
            from lucivant_sim.simulator import expectation, run

            try:
                job.passes, job.next_pass = [], 0
                if job.terms is not None:
                    # a noisy expectation: Tr(P rho) of the density matrix, as expectation() gives
                    self._finish(job, expectation(job.c.circuit, job.terms, job.req.noise))
                    return True
                res = run(job.c.circuit, job.req.shots, seed=job.req.seed, noise=job.req.noise, method="density_matrix")
                self._finish(job, res.counts)
            except BaseException as exc:  # noqa: BLE001
                self._finish_error(job.req, exc)
                self._release(job)
            return True
        if job.method == "trajectories":
            if job.future is None:
                job.req.stamps.setdefault("launched", time.perf_counter())
                if frontend.trajectory_workers():
                    # a worker process of its own (its own interpreter lock), same counts
                    assert job.c.circuit is not None
                    try:
                        job.future = frontend.submit_trajectories(job.c.circuit, job.req.noise, job.req.shots, job.req.seed, self.dtype)
                    except Exception:  # noqa: BLE001 - a worker died: run it on this process's thread
                        job.future = None
                if job.future is None:
                    job.future = self._trajectories.submit(self._run_trajectories, job)
                return True
            if not job.future.done():
                return False
            try:
                counts = job.future.result()
                job.passes, job.next_pass = [], 0
                self._finish(job, counts)
            except BaseException as exc:  # noqa: BLE001
                self._finish_error(job.req, exc)
                self._release(job)
            return True
        lib = nat.load()
        if job.state is None:
            job.stream = self.streams[self._rr % len(self.streams)]
            self._rr += 1
            try:
                with torch.cuda.stream(job.stream):
                    job.state = Statevector(job.c.num_qubits, dtype=self.dtype)
                    job.passes = job.c.passes
                    assert job.passes is not None
                    # every launch's function and arguments, built once (matrices uploaded)
                    job.mats, job.calls = plan_launches(
                        job.state.buffer, self.dtype, job.c.num_qubits, 1, job.passes, job.stream.cuda_stream
                    )
            except BaseException as exc:  # noqa: BLE001 - reported to the requester
                self._finish_error(job.req, exc)
                self._release(job)
                return True
        assert job.calls is not None
        job.req.stamps.setdefault("launched", time.perf_counter())
        _use_switches(job.c.num_qubits)
        stop = min(len(job.calls), job.next_pass + self.chunk)
        for fn, args in job.calls[job.next_pass : stop]:
            if fn(*args):
                exc = nat.NativeError(lib.lsim_last_error().decode(errors="replace"))
                self._finish_error(job.req, exc)
                self._release(job)
                return True
        launched = stop > job.next_pass
        job.next_pass = stop
        return launched

    def _run_trajectories(self, job: _Job) -> dict[str, int]:
        """A wide trajectory request's counts, on the trajectory thread and stream."""
        from lucivant_sim.trajectories import run_trajectories

        c = job.c.circuit
        assert c is not None
        with torch.cuda.stream(self._traj_stream):
            tally = run_trajectories(c, job.req.noise, job.req.shots, np.random.default_rng(job.req.seed), dtype=self.dtype)
        return {c.format_bits(b): k for b, k in sorted(tally.items())} if c.num_clbits else {}

    def _queue_outputs(self, job: _Job) -> None:
        """Queue a wide request's outputs behind its last pass, on its stream: Pauli
        expectations, or |psi|^2 for sampling; an event marks when they are ready."""
        assert job.state is not None and job.stream is not None
        with torch.cuda.stream(job.stream):
            sv = job.state
            if job.terms is not None:
                dev = launch_expectations("lsim_sv_pauli_expect", sv.buffer, sv.dtype, sv.n, sv.n, list(job.terms))
            else:
                dev = torch.empty(1 << sv.n, dtype=torch.float64, device=sv.buffer.device)
                nat.call("lsim_sv_probabilities", ptr(sv.buffer), sv.dtype, sv.n, 1, ptr(dev), job.stream.cuda_stream)
            # straight into pinned host memory by DMA, behind the kernels on the same stream, so that
            # completing the request issues no device work (nothing that could wait on other streams)
            out = torch.empty(dev.numel(), dtype=torch.float64, pin_memory=True)
            out.copy_(dev, non_blocking=True)
            done = torch.cuda.Event()
            done.record(job.stream)
        job.finishing = (done, out, dev)  # dev stays alive until the copy has completed
        job.req.stamps["outputs_queued"] = time.perf_counter()

    def _complete_wide(self, job: _Job) -> None:
        if job.req.done:
            return
        assert job.finishing is not None
        _, out, _dev = job.finishing
        t0 = time.perf_counter()
        values = out.numpy()  # the event has completed: the values are in pinned host memory
        self.host_seconds["copy_out"] += time.perf_counter() - t0
        job.finishing = None
        if job.terms is not None:
            total = 0.0
            for w, v in zip(job.terms.values(), values):
                total += w * float(v)
            self._finish(job, total)
            return
        # the state's HBM is free for admission now; the host sampling runs beside the scheduler
        self._release(job)
        self._sampling.append(self._sampler.submit(self._sample_wide, job, values))

    def _sample_wide(self, job: _Job, probs: np.ndarray) -> None:
        try:
            self._finish(job, _sample(job.c, probs, job.req.shots, job.req.noise, np.random.default_rng(job.req.seed)))
        except BaseException as exc:  # noqa: BLE001
            self._finish_error(job.req, exc)

    # -- bookkeeping ----------------------------------------------------------------------

    def _finish(self, job: _Job, result: Any) -> None:
        job.req.result = result
        job.req.done = time.perf_counter()
        self._release(job)
        with self._lock:
            self.completed.append(job.req)
        job.req.event.set()

    def _finish_error(self, req: Request, exc: BaseException) -> None:
        req.error = exc
        req.done = time.perf_counter()
        with self._lock:
            self.completed.append(req)
        req.event.set()

    def _release(self, job: _Job) -> None:
        with self._lock:
            if job in self.resident:
                self.resident.remove(job)
                self.used -= job.footprint
        job.state = None
        job.mats = None
        job.calls = None

    def stats(self) -> dict[str, float]:
        lat = np.array([r.latency for r in self.completed if r.error is None])
        if lat.size == 0:
            return {"completed": 0}
        span = max(r.done for r in self.completed) - min(r.arrival for r in self.completed)
        phases: dict[str, list[float]] = {"read": [], "wait": [], "device": [], "finish": []}
        for r in self.completed:
            st = r.stamps
            if r.error is not None or "read" not in st:
                continue
            phases["read"].append(st["read"] - r.arrival)
            if "admitted" in st:
                phases["wait"].append(st["admitted"] - st["read"])
            if "launched" in st and "outputs_queued" in st:
                phases["device"].append(st["outputs_queued"] - st["launched"])
                phases["finish"].append(r.done - st["outputs_queued"])
        return {
            "completed": int(lat.size),
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "circuits_per_s": float(lat.size / span) if span > 0 else float("inf"),
            "ticks": self.ticks,
            "peak_hbm_bytes": self.peak_used,
            "scheduler_seconds": {k: round(v, 3) for k, v in self.host_seconds.items()},
            # median seconds per phase: arrival -> read, read -> admitted, launches of wide
            # requests, outputs queued -> result
            "phase_median_s": {k: round(float(np.median(v)), 4) for k, v in phases.items() if v},
        }


def serve_stream(server: Server, requests: list[Request], arrivals: list[float]) -> dict[str, float]:
    """Submit ``requests`` at the given offsets (seconds from now) from a feeder thread while the
    server loop runs; returns the server's latency and throughput statistics."""
    t0 = time.perf_counter()

    def feed() -> None:
        for req, at in zip(requests, arrivals):
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            req.arrival = time.perf_counter()
            server.submit(req)

    feeder = threading.Thread(target=feed, daemon=True)
    server.start()
    feeder.start()
    feeder.join()
    for r in requests:
        r.wait()
    server.stop()
    return server.stats()


def plain_stream(requests: list[Request], arrivals: list[float], *, fused: bool = False) -> dict[str, float]:
    """The same stream through the single-request API, one request at a time with no batching
    or overlap: ``lucivant_sim.run`` / ``expectation``. ``fused=False`` launches one kernel per
    gate at every size (fusion width 1, whatever the autotuner stored; runs of one-qubit gates on
    a qubit are still multiplied into one operator, as the reference's own evolution does);
    ``fused=True`` keeps the library's fused plans (the autotuner's choice per size)."""
    from lucivant_sim import simulator

    old, old_tuned, old_tile = simulator.FUSION_WIDTH, simulator._tuned, simulator._tile_choice
    if not fused:
        simulator.FUSION_WIDTH = 1
        simulator._tuned = lambda num_qubits: (1, 0, 0)
        simulator._tile_choice = lambda num_qubits: None
    t0 = time.perf_counter()
    lat = []
    try:
        for req, at in zip(requests, arrivals):
            now = time.perf_counter() - t0
            if now < at:
                time.sleep(at - now)
            arrival = t0 + at
            if req.observable is not None:
                req.result = simulator.expectation(req.program, req.observable, req.noise)
            else:
                res = simulator.run(req.program, req.shots, seed=req.seed, noise=req.noise, method=req.method)
                req.result = res.counts
            req.done = time.perf_counter()
            lat.append(req.done - arrival)
    finally:
        simulator.FUSION_WIDTH, simulator._tuned, simulator._tile_choice = old, old_tuned, old_tile
    arr = np.array(lat)
    span = time.perf_counter() - t0
    return {
        "completed": len(lat),
        "p50_ms": float(np.percentile(arr, 50) * 1e3),
        "p99_ms": float(np.percentile(arr, 99) * 1e3),
        "circuits_per_s": len(lat) / span,
    }
