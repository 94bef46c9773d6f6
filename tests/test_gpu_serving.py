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


def _alone(r: Request) -> Request:
    server = Server(streams=1)
    a = server.submit(_fresh(r))
    server.run_until_idle()
    return a


def test_wide_requests_alone_and_in_the_mix_match_the_oracle() -> None:
    """Requests of 17-26 qubits beside small ones. The 26-qubit configuration (40 layers of u3 +
    cx/cz) as a weighted expectation of its 8 Pauli strings is checked against the oracle's own
    full run (tests/golden/full_c2_oracle.json) within 1e-10; a 20-qubit count request against
    the oracle's counts for the same seed; an 18-qubit ansatz expectation against the oracle's;
    17- and 22-qubit requests against the single-circuit API. Each result is also the same alone
    and in the mix."""
    import json
    from pathlib import Path

    from oracle import engine as orc

    rec = json.loads((Path(__file__).parent / "golden" / "full_c2_oracle.json").read_text())
    src26, labels = wl.config_c2()
    assert labels == rec["labels"]
    weights = np.random.default_rng(8).uniform(-1, 1, len(labels))
    c2 = Request(src26, observable={lbl: float(w) for lbl, w in zip(labels, weights)})
    c2_want = float(np.dot(weights, rec["expectations"]))

    counts20 = Request(wl.random_layers(20, 6, 123), shots=1000, seed=9)
    counts20_want = orc.run_counts(ls.parse(counts20.program), 1000, 9)

    rng = np.random.default_rng(21)
    hea18 = Request(wl.hea_ansatz(18, 12, 5), observable={lbl: float(w) for lbl, w in zip(wl.pauli_strings(18, 16, 3), rng.uniform(-1, 1, 16))})
    psi18 = orc.evolve_sv(ls.parse(hea18.program))
    hea18_want = sum(w * orc.expectation_sv(psi18, lbl, 18) for lbl, w in hea18.observable.items())  # type: ignore[union-attr]

    counts17 = Request(wl.random_layers(17, 30, 44), shots=1000, seed=12)
    hea22 = Request(wl.hea_ansatz(22, 40, 6), observable={lbl: float(w) for lbl, w in zip(wl.pauli_strings(22, 16, 4), rng.uniform(-1, 1, 16))})

    mix = [c2, *_requests()[:4], counts20, hea18, counts17, hea22]
    server = Server(streams=4, chunk_passes=16)
    for r in mix:
        server.submit(r)
    server.run_until_idle()
    for r in mix:
        assert r.error is None, r.error
    assert abs(c2.result - c2_want) < 1e-10
    assert counts20.result == counts20_want
    assert abs(hea18.result - hea18_want) < 1e-10
    assert counts17.result == ls.run(counts17.program, 1000, seed=12).counts
    assert abs(hea22.result - ls.expectation(hea22.program, hea22.observable)) < 1e-10  # type: ignore[arg-type]
    for r in (c2, counts20, hea18, counts17, hea22):
        a = _alone(r)
        assert a.error is None
        if r.observable is not None:
            assert abs(a.result - r.result) < 1e-12
        else:
            assert a.result == r.result


def test_noisy_expectation_requests_use_the_density_matrix() -> None:
    """An expectation request with gate noise is Tr(P rho) of the density matrix, as
    expectation() computes it (simulator.py:503-516), not the ideal statevector's value; wider
    than the density-matrix engine's 13 qubits it fails with the engine's error."""
    from oracle import engine as orc

    noise = wl.noise_c3()
    src = wl.hea_ansatz(6, 4, 11)
    labels = wl.pauli_strings(6, 5, 12)
    obs = {lbl: float(w) for lbl, w in zip(labels, np.linspace(-1, 1, 5))}
    rho = orc.evolve_dm(ls.parse(src), noise)
    want = sum(w * orc.expectation_dm(rho, lbl, 6) for lbl, w in obs.items())
    ideal = ls.expectation(src, obs)
    assert abs(want - ideal) > 1e-6  # the noise changes the value
    noisy = Request(src, observable=obs, noise=noise)
    wide = Request(wl.hea_ansatz(14, 2, 3), observable="Z" * 14, noise=noise)
    server = Server(streams=2)
    for r in (noisy, wide, *_requests()[:2]):
        server.submit(r)
    server.run_until_idle()
    assert noisy.error is None and abs(noisy.result - want) < 1e-10
    assert abs(noisy.result - ls.expectation(src, obs, noise)) < 1e-12
    assert isinstance(wide.error, ls.SimulationError)


def test_expectation_requests_with_many_or_no_terms() -> None:
    """65 Pauli strings (more than one program launch takes) and an empty observable, on a
    10-qubit circuit: the same values as expectation() (0.0 for no terms), and the other
    requests of the mix are unaffected."""
    src = wl.hea_ansatz(10, 5, 31)
    labels = wl.pauli_strings(10, 65, 32)
    obs = {lbl: float(w) for lbl, w in zip(labels, np.random.default_rng(33).uniform(-1, 1, 65))}
    many = Request(src, observable=obs)
    none = Request(src, observable={})
    rest = _requests()[:3]
    server = Server(streams=2)
    for r in (many, none, *rest):
        server.submit(r)
    server.run_until_idle()
    assert many.error is None and abs(many.result - ls.expectation(src, obs)) < 1e-10
    assert none.error is None and none.result == 0.0
    for r in rest:
        assert r.error is None
        if r.observable is not None:
            assert abs(r.result - ls.expectation(r.program, r.observable)) < 1e-10
        else:
            assert r.result == ls.run(r.program, r.shots, seed=r.seed).counts


def test_tile_plans_chosen_by_the_autotuner_match_the_oracle(monkeypatch: pytest.MonkeyPatch) -> None:
    """When the stored choice for a size is tile passes, ``run``, ``expectation`` and the server
    all plan that size's circuits into tile passes: 18- and 20-qubit results still equal the
    oracle's (counts exactly, expectations within 1e-10)."""
    from lucivant_sim import autotune, simulator
    from oracle import engine as orc

    for n in (18, 20):
        monkeypatch.setitem(simulator._TUNED, n, (4, 0, 0, autotune.DEFAULT_SWITCHES, (24.0, 4)))
    assert simulator._tile_choice(20) == (24.0, 4)
    counts20 = Request(wl.random_layers(20, 6, 123), shots=1000, seed=9)
    counts20_want = orc.run_counts(ls.parse(counts20.program), 1000, 9)
    rng = np.random.default_rng(22)
    hea18 = Request(wl.hea_ansatz(18, 12, 7), observable={lbl: float(w) for lbl, w in zip(wl.pauli_strings(18, 16, 5), rng.uniform(-1, 1, 16))})
    psi18 = orc.evolve_sv(ls.parse(hea18.program))
    hea18_want = sum(w * orc.expectation_sv(psi18, lbl, 18) for lbl, w in hea18.observable.items())  # type: ignore[union-attr]

    assert ls.run(counts20.program, 1000, seed=9).counts == counts20_want
    assert abs(ls.expectation(hea18.program, hea18.observable) - hea18_want) < 1e-10  # type: ignore[arg-type]
    server = Server(streams=2, chunk_passes=8)
    assert server.widths[20][3] == (24.0, 4)
    mix = [counts20, hea18, *_requests()[:2]]
    for r in mix:
        server.submit(r)
    server.run_until_idle()
    for r in mix:
        assert r.error is None, r.error
    assert counts20.result == counts20_want
    assert abs(hea18.result - hea18_want) < 1e-10


def test_noisy_wide_count_requests_run_beside_the_rest() -> None:
    """Noisy count requests too wide for an LDS program (14 and 16 qubits, trajectories through
    the HBM batch) run on the server's trajectory thread while the other requests proceed: each
    result equals the single-circuit API's alone and in the mix, and the 14-qubit one equals the
    oracle's per-shot loop for the same seed."""
    from oracle import engine as orc

    noise = ls.NoiseModel.uniform(single=1e-3, two=1e-2)
    w14 = Request(wl.random_layers(14, 3, 501), shots=300, seed=21, noise=noise, method="trajectories")
    w16 = Request(wl.teleport_chain(16, 2), shots=400, seed=22, noise=wl.noise_c4(), method="trajectories")
    mix = [w14, *_requests()[:5], w16]
    server = Server(streams=3, chunk_passes=8)
    for r in mix:
        server.submit(r)
    server.run_until_idle()
    for r in mix:
        assert r.error is None, r.error
    want14 = orc.trajectory_counts(ls.parse(w14.program), noise, 300, np.random.default_rng(21))
    assert w14.result == want14
    for r in (w14, w16):
        assert r.result == ls.run(r.program, r.shots, seed=r.seed, noise=r.noise, method="trajectories").counts
        alone = Request(r.program, shots=r.shots, seed=r.seed, noise=r.noise, method="trajectories")
        server = Server(streams=1)
        server.submit(alone)
        server.run_until_idle()
        assert alone.error is None and alone.result == r.result
