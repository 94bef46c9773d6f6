"""A serving loop with continuous batching of circuit requests on one MI355X.

Requests (an expectation of a Pauli sum, or seeded counts, of one OpenQASM program) arrive at any
time. The scheduler runs in ticks; at every tick it

1. admits waiting requests in arrival order while their HBM footprint fits the budget
   (statevector: 16 B x 2^n; trajectories: shots x 16 B x 2^n, chunked), so requests that arrive
   while others are on the GPU join as memory frees up;
2. launches every admitted request of at most ``program.max_qubits`` qubits that hasn't run yet,
   grouped by width, as one LDS-program launch (one workgroup per circuit or per shot);
3. advances every resident wide request by up to ``chunk_passes`` fused passes on its own HIP
   stream (states stay resident in HBM between ticks), so wide circuits overlap one another and
   small batches slip in between their chunks;
4. completes finished requests: probabilities or expectations come back to the host, counts are
   sampled with the request's own seeded generator exactly as ``run`` samples them.

A request's result doesn't depend on what else is in flight: each circuit runs in its own
workgroup or its own state buffer with the same kernels, so a result alone and in the mix is the
same (tests/test_gpu_serving.py checks it). Latency is measured from arrival to result.
"""

from __future__ import annotations

import itertools
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import Any, Mapping

import numpy as np
import torch

from lucivant_sim import _native as nat
from lucivant_sim import program as pg
from lucivant_sim.circuit import Circuit
from lucivant_sim.engines import Statevector, pauli_masks, ptr, require_device, upload
from lucivant_sim.noise import NoiseModel
from lucivant_sim.plan import plan_statevector
from lucivant_sim.simulator import FUSION_WIDTH, _sample, choose_method, to_circuit

_ids = itertools.count()


@dataclass
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

    def wait(self, timeout: float | None = None) -> Any:
        self.event.wait(timeout)
        if self.error is not None:
            raise self.error
        return self.result

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Job:
    req: Request
    circuit: Circuit
    method: str
    footprint: int
    terms: dict[str, float] | None = None
    # wide statevector requests: resident state and the remaining passes
    state: Statevector | None = None
    passes: list | None = None
    next_pass: int = 0
    stream: torch.cuda.Stream | None = None
    mats: Any = None
    addrs: list[int] | None = None


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
    ) -> None:
        require_device()
        self.budget = hbm_budget
        self.used = 0
        self.dtype = dtype
        self.width = fusion_width or FUSION_WIDTH
        self.chunk = chunk_passes
        self.streams = [torch.cuda.Stream() for _ in range(max(1, streams))]
        self._rr = 0
        self.inbox: queue.Queue[Request] = queue.Queue()
        self.waiting: list[_Job] = []
        self.resident: list[_Job] = []
        self.completed: list[Request] = []
        self.max_lds = pg.max_qubits(dtype)
        self._stop = threading.Event()
        self._thread: threading.Thread | None = None
        self.ticks = 0
        self.peak_used = 0

    # -- submission ---------------------------------------------------------------------

    def submit(self, req: Request) -> Request:
        """Queue a request. Its program is parsed here, in the caller's thread, so the host work
        of arriving requests overlaps what the scheduler has on the GPU."""
        if not req.arrival:
            req.arrival = time.perf_counter()
        try:
            req.program = to_circuit(req.program)
        except BaseException as exc:  # noqa: BLE001 - reported to the requester
            self._finish_error(req, exc)
            return req
        self.inbox.put(req)
        return req

    def start(self) -> None:
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()

    def stop(self) -> None:
        self._stop.set()
        if self._thread is not None:
            self._thread.join()

    def _loop(self) -> None:
        while not self._stop.is_set():
            if not self.tick():
                time.sleep(0.0005)

    def run_until_idle(self) -> None:
        while self.tick() or self.waiting or self.resident or not self.inbox.empty():
            pass

    # -- scheduling ---------------------------------------------------------------------

    def _prepare(self, req: Request) -> _Job:
        circuit = to_circuit(req.program)
        if req.observable is not None:
            terms = {req.observable: 1.0} if isinstance(req.observable, str) else dict(req.observable)
            for label in terms:
                if len(label) != circuit.num_qubits or set(label.upper()) - set("IXYZ"):
                    raise ValueError(f"{label!r} must be a Pauli string of length {circuit.num_qubits}")
            if circuit.is_dynamic():
                raise ValueError("an expectation value needs a circuit without mid-circuit measurement")
            method = "statevector"
        else:
            terms = None
            method = choose_method(circuit, req.noise, req.method)
            if method not in ("statevector", "trajectories", "density_matrix"):
                raise ValueError(f"the server runs statevector, density-matrix and trajectory requests, not {method!r}")
        size = 16 << circuit.num_qubits
        if method == "density_matrix":
            size = 16 << (2 * circuit.num_qubits)
        footprint = size * (req.shots if method == "trajectories" and circuit.num_qubits > self.max_lds else 1)
        return _Job(req, circuit, method, footprint, terms)

    def _admit(self) -> None:
        while True:
            try:
                req = self.inbox.get_nowait()
            except queue.Empty:
                break
            try:
                self.waiting.append(self._prepare(req))
            except BaseException as exc:  # noqa: BLE001 - reported to the requester
                self._finish_error(req, exc)
        still = []
        for job in self.waiting:
            if self.used + job.footprint <= self.budget or not self.resident:
                self.used += job.footprint
                self.peak_used = max(self.peak_used, self.used)
                self.resident.append(job)
            else:
                still.append(job)
        self.waiting = still

    def tick(self) -> bool:
        """One scheduling round; True if it did any work."""
        self._admit()
        if not self.resident:
            return False
        self.ticks += 1
        small = [
            j
            for j in self.resident
            if j.state is None and j.method != "density_matrix" and 1 <= j.circuit.num_qubits <= self.max_lds
        ]
        self._run_small(small)
        wide = [j for j in self.resident if j not in small]
        for job in wide:
            self._advance_wide(job)
        for job in wide:
            if job.passes is not None and job.next_pass >= len(job.passes):
                self._complete_wide(job)
        return True

    def _run_small(self, jobs: list[_Job]) -> None:
        by_key: dict[tuple[int, str, int], list[_Job]] = {}
        for j in jobs:
            key = (j.circuit.num_qubits, j.method if j.terms is None else "expect", len(j.terms or ()))
            by_key.setdefault(key, []).append(j)
        for (n, kind, nterms), group in by_key.items():
            try:
                if kind == "expect":
                    progs = [pg.compile_static(j.circuit) for j in group]
                    masks = [[pauli_masks(lbl) for lbl in j.terms] for j in group]  # type: ignore[union-attr]
                    out = pg.run_programs(progs, dtype=self.dtype, pauli=masks)
                    for idx, j in enumerate(group):
                        total = 0.0
                        for w, v in zip(j.terms.values(), out.expect[idx]):  # type: ignore[union-attr,index]
                            total += w * float(v)
                        self._finish(j, total)
                elif kind == "statevector":
                    progs = [pg.compile_static(j.circuit) for j in group]
                    out = pg.run_programs(progs, dtype=self.dtype, want_probs=True)
                    for idx, j in enumerate(group):
                        rng = np.random.default_rng(j.req.seed)
                        self._finish(j, _sample(j.circuit, out.probs[idx], j.req.shots, j.req.noise, rng))  # type: ignore[index]
                else:
                    from lucivant_sim.batch import run_many

                    res = run_many([j.circuit for j in group], group[0].req.shots, seeds=[j.req.seed for j in group], noise=group[0].req.noise, method="trajectories", dtype=self.dtype) if _same_shots_noise(group) else [
                        run_many([j.circuit], j.req.shots, seeds=[j.req.seed], noise=j.req.noise, method="trajectories", dtype=self.dtype)[0] for j in group
                    ]
                    for j, r in zip(group, res):
                        self._finish(j, r.counts)
            except BaseException as exc:  # noqa: BLE001
                for j in group:
                    self._finish_error(j.req, exc)
                    self._release(j)

    def _advance_wide(self, job: _Job) -> None:
        if job.method == "density_matrix":
            from lucivant_sim.simulator import run

            try:
                res = run(job.circuit, job.req.shots, seed=job.req.seed, noise=job.req.noise, method="density_matrix")
                job.passes, job.next_pass = [], 0
                self._finish(job, res.counts)
            except BaseException as exc:  # noqa: BLE001
                self._finish_error(job.req, exc)
                self._release(job)
            return
        if job.method == "trajectories":
            from lucivant_sim.trajectories import run_trajectories

            try:
                tally = run_trajectories(job.circuit, job.req.noise, job.req.shots, np.random.default_rng(job.req.seed), dtype=self.dtype)
                c = job.circuit
                counts = {c.format_bits(b): k for b, k in sorted(tally.items())} if c.num_clbits else {}
                job.passes, job.next_pass = [], 0
                self._finish(job, counts)
            except BaseException as exc:  # noqa: BLE001
                self._finish_error(job.req, exc)
                self._release(job)
            return
        if job.state is None:
            job.stream = self.streams[self._rr % len(self.streams)]
            self._rr += 1
            with torch.cuda.stream(job.stream):
                job.state = Statevector(job.circuit.num_qubits, dtype=self.dtype)
                job.passes = plan_statevector(job.circuit.operations, job.circuit.num_qubits, self.width)
                job.mats, job.addrs = upload([p.matrix for p in job.passes], self.dtype)
        assert job.passes is not None and job.addrs is not None and job.stream is not None
        lib = nat.load()
        hs = job.stream.cuda_stream
        state = ptr(job.state.buffer)
        stop = min(len(job.passes), job.next_pass + self.chunk)
        for p, addr in zip(job.passes[job.next_pass : stop], job.addrs[job.next_pass : stop]):
            if p.kind == "scale":
                z = complex(p.matrix[0])
                lib.lsim_sv_scale(state, self.dtype, job.circuit.num_qubits, 1, z.real, z.imag, hs)
                continue
            fn = lib.lsim_sv_apply_dense if p.kind == "dense" else lib.lsim_sv_apply_diag
            fn(state, self.dtype, job.circuit.num_qubits, 1, p.k, nat.int_array(p.bits), addr, None, hs)
        job.next_pass = stop

    def _complete_wide(self, job: _Job) -> None:
        if job.req.done:
            return
        assert job.state is not None and job.stream is not None
        with torch.cuda.stream(job.stream):
            if job.terms is not None:
                vals = job.state.expectations(list(job.terms))
                total = 0.0
                for w, v in zip(job.terms.values(), vals):
                    total += w * float(v)
                self._finish(job, total)
            else:
                probs = job.state.probabilities()
                self._finish(job, _sample(job.circuit, probs, job.req.shots, job.req.noise, np.random.default_rng(job.req.seed)))

    # -- bookkeeping ----------------------------------------------------------------------

    def _finish(self, job: _Job, result: Any) -> None:
        job.req.result = result
        job.req.done = time.perf_counter()
        self.completed.append(job.req)
        job.req.event.set()
        self._release(job)

    def _finish_error(self, req: Request, exc: BaseException) -> None:
        req.error = exc
        req.done = time.perf_counter()
        self.completed.append(req)
        req.event.set()

    def _release(self, job: _Job) -> None:
        if job in self.resident:
            self.resident.remove(job)
            self.used -= job.footprint
        job.state = None
        job.mats = None

    def stats(self) -> dict[str, float]:
        lat = np.array([r.latency for r in self.completed if r.error is None])
        if lat.size == 0:
            return {"completed": 0}
        span = max(r.done for r in self.completed) - min(r.arrival for r in self.completed)
        return {
            "completed": int(lat.size),
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "circuits_per_s": float(lat.size / span) if span > 0 else float("inf"),
            "ticks": self.ticks,
            "peak_hbm_bytes": self.peak_used,
        }


def _same_shots_noise(group: list[_Job]) -> bool:
    return len({j.req.shots for j in group}) == 1 and len({id(j.req.noise) for j in group}) == 1


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


def plain_stream(requests: list[Request], arrivals: list[float]) -> dict[str, float]:
    """The same stream through the plain path: one request at a time, one launch per gate
    (``lucivant_sim.run`` / ``expectation`` with fusion width 1), no batching or overlap."""
    from lucivant_sim import simulator

    old = simulator.FUSION_WIDTH
    simulator.FUSION_WIDTH = 1
    t0 = time.perf_counter()
    lat = []
    try:
        for req, at in zip(requests, arrivals):
            now = time.perf_counter() - t0
            if now < at:
                time.sleep(at - now)
            arrival = t0 + at
            if req.observable is not None:
                req.result = simulator.expectation(req.program, req.observable)
            else:
                res = simulator.run(req.program, req.shots, seed=req.seed, noise=req.noise, method=req.method)
                req.result = res.counts
            req.done = time.perf_counter()
            lat.append(req.done - arrival)
    finally:
        simulator.FUSION_WIDTH = old
    arr = np.array(lat)
    span = time.perf_counter() - t0
    return {
        "completed": len(lat),
        "p50_ms": float(np.percentile(arr, 50) * 1e3),
        "p99_ms": float(np.percentile(arr, 99) * 1e3),
        "circuits_per_s": len(lat) / span,
    }
