"""The serving loop: every request's result is the same alone and in a mix of concurrent
requests, and equals the single-circuit API's (expectations within 1e-10, counts exactly)."""

from __future__ import annotations

import numpy as np
import pytest

import lucivant_sim as ls
from lucivant_sim import workloads as wl
from lucivant_sim.serving import Request, Server, serve_stream

pytestmark = pytest.mark.gpu


def _requests() -> list[Request]:
    rng = np.random.default_rng(5)
    reqs = []
    for i, n in enumerate([10, 16, 12, 14, 11, 15, 13, 10]):
        if i % 2 == 0:
            src = wl.hea_ansatz(n, 3 + i, 40 + i)
            obs = {lbl: float(w) for lbl, w in zip(wl.pauli_strings(n, 16, 70 + i), rng.uniform(-1, 1, 16))}
            reqs.append(Request(src, observable=obs))
        else:
            src = wl.random_layers(n, 4 + i, 90 + i)
            reqs.append(Request(src, shots=1000, seed=300 + i))
    reqs.append(Request(wl.teleport_chain(9, 3), shots=500, seed=11, noise=wl.noise_c4()))
    return reqs


def _fresh(r: Request) -> Request:
    return Request(r.program, observable=r.observable, shots=r.shots, seed=r.seed, noise=r.noise)


def test_results_alone_and_in_the_mix_are_the_same() -> None:
    reqs = _requests()
    server = Server(streams=3, chunk_passes=8)
    for r in reqs:
        server.submit(r)
    server.run_until_idle()
    for r in reqs:
        alone = Server(streams=1)
        a = alone.submit(_fresh(r))
        alone.run_until_idle()
        assert r.error is None and a.error is None
        if r.observable is not None:
            assert abs(a.result - r.result) < 1e-12
            assert abs(r.result - ls.expectation(r.program, r.observable)) < 1e-10
        else:
            assert a.result == r.result
            assert r.result == ls.run(r.program, r.shots, seed=r.seed, noise=r.noise).counts


def test_streamed_arrivals_are_all_served() -> None:
    reqs = _requests()
    arrivals = list(np.cumsum(np.random.default_rng(1).exponential(0.01, size=len(reqs))))
    stats = serve_stream(Server(streams=2, chunk_passes=4), reqs, arrivals)
    assert stats["completed"] == len(reqs)
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
    assert all(r.error is None for r in reqs)


def test_admission_respects_the_hbm_budget() -> None:
    # budget for one 16-qubit state: the second wide request waits until the first completes
    reqs = [Request(wl.random_layers(16, 3, s), shots=100, seed=s) for s in range(3)]
    server = Server(hbm_budget=16 << 16, streams=2, chunk_passes=2)
    for r in reqs:
        server.submit(r)
    server.run_until_idle()
    assert server.peak_used <= 16 << 16
    for r in reqs:
        assert r.result == ls.run(r.program, 100, seed=r.seed).counts


def test_bad_requests_fail_alone_and_the_rest_are_served() -> None:
    """Requests that cannot run -- a syntax error, a state too wide for the device engines, a
    method the server does not run -- come back with their error; the good requests in the same
    mix come back with the single-circuit API's results."""
    good = _requests()[:3]
    bad = [
        Request("OPENQASM 3;\nqubit[2] q;\nfoo q[0];\n", shots=10, seed=1),
        Request("OPENQASM 3;\ninclude \"stdgates.inc\";\nqubit[30] q;\nh q;\nt q;\n", shots=10, seed=2),
        Request(wl.random_layers(4, 2, 3), shots=10, seed=3, method="stabilizer"),
    ]
    server = Server(streams=2, chunk_passes=8)
    for r in [good[0], bad[0], good[1], bad[1], bad[2], good[2]]:
        server.submit(r)
    server.run_until_idle()
    assert isinstance(bad[0].error, ls.QasmError) and bad[0].error.line == 3
    assert isinstance(bad[1].error, (ls.SimulationError, ValueError))
    assert isinstance(bad[2].error, (ls.SimulationError, ValueError))
    for r in good:
        assert r.error is None
        if r.observable is not None:
            assert abs(r.result - ls.expectation(r.program, r.observable)) < 1e-10
        else:
            assert r.result == ls.run(r.program, r.shots, seed=r.seed, noise=r.noise).counts
