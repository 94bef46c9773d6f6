"""Child process of tests/test_gpu_frontend.py: starts the front-end pool (and its trajectory
workers) before touching the GPU, runs a mixed batch through ``run_many`` and a request stream
(noisy count requests of 14 and 15 qubits among them, one of them a dynamic circuit) through the
server with the pool, then the same work in this process without it, and prints one JSON line
comparing them."""

from __future__ import annotations

import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path[:0] = [str(ROOT / "lucivant-sim_amd"), str(ROOT)]

from lucivant_sim import frontend  # noqa: E402

workers = frontend.start(3)

import numpy as np  # noqa: E402

import lucivant_sim as ls  # noqa: E402
from lucivant_sim import workloads as wl  # noqa: E402
from lucivant_sim.batch import run_many  # noqa: E402
from lucivant_sim.serving import Request, Server  # noqa: E402


def main() -> None:
    texts = [wl.random_layers(8, 5, 700 + i) for i in range(12)] + [
        wl.random_layers(11, 4, 50),
        wl.random_layers(15, 3, 51),  # wider than the LDS program: the single-circuit path
        wl.hea_ansatz(9, 3, 52),
    ]
    seeds = list(range(100, 100 + len(texts)))
    pooled = run_many(texts, 700, seeds=seeds)
    alone = [ls.run(t, 700, seed=s) for t, s in zip(texts, seeds)]
    batch_same = all(a.counts == b.counts and a.fingerprint == b.fingerprint for a, b in zip(pooled, alone))

    noise = wl.noise_c4()
    noisy = [wl.teleport_chain(9, 3), wl.random_layers(6, 3, 5)]
    pooled_noisy = run_many(noisy, 300, seeds=[1, 2], noise=noise, method="trajectories")
    alone_noisy = [ls.run(t, 300, seed=s, noise=noise, method="trajectories") for t, s in zip(noisy, [1, 2])]
    noisy_same = all(a.counts == b.counts for a, b in zip(pooled_noisy, alone_noisy))

    error_kept = False
    try:
        run_many([texts[0], "OPENQASM 2.0;\nqreg q[2];\nh q[5];\n"], 10, seeds=[1, 2])
    except ls.QasmError as exc:
        error_kept = exc.line == 3
    # an error in a later chunk of a long batch (chunks of 32): the first error in program
    # order is raised, and the pool serves the next batch normally
    many = [wl.random_layers(6, 2, 900 + i) for i in range(80)]
    many[67] = "OPENQASM 3;\nqubit[2] q;\n\nfoo q[0];\n"
    many[75] = "OPENQASM 2.0;\nqreg q[2];\nh q[5];\n"
    try:
        run_many(many, 10, seeds=list(range(80)))
        error_kept = False
    except ls.QasmError as exc:
        error_kept = error_kept and exc.line == 4
    again = run_many(texts[:3], 700, seeds=seeds[:3])
    error_kept = error_kept and all(a.counts == b.counts for a, b in zip(again, alone[:3]))

    rng = np.random.default_rng(3)
    reqs = []
    for i, n in enumerate([10, 16, 12, 14]):
        if i % 2 == 0:
            obs = {lbl: float(w) for lbl, w in zip(wl.pauli_strings(n, 8, 80 + i), rng.uniform(-1, 1, 8))}
            reqs.append(Request(wl.hea_ansatz(n, 3, 60 + i), observable=obs))
        else:
            reqs.append(Request(wl.random_layers(n, 4, 90 + i), shots=500, seed=i))
    # noisy count requests wider than an LDS program: run by the trajectory worker processes
    # (one of them with mid-circuit measurements and conditioned corrections)
    for i, n in enumerate([14, 15]):
        reqs.append(Request(wl.random_layers(n, 6, 95 + i), shots=400, seed=20 + i, noise=noise, method="trajectories"))
    reqs.append(Request(wl.teleport_chain(14, 3), shots=300, seed=31, noise=noise, method="trajectories"))
    server = Server(streams=2, chunk_passes=8)
    for r in reqs:
        server.submit(r)
    server.run_until_idle()
    serve_same = True
    for r in reqs:
        if r.error is not None:
            serve_same = False
        elif r.observable is not None:
            serve_same &= abs(r.result - ls.expectation(r.program, r.observable)) < 1e-10
        elif r.noise is not None:
            serve_same &= r.result == ls.run(r.program, r.shots, seed=r.seed, noise=r.noise, method=r.method).counts
        else:
            serve_same &= r.result == ls.run(r.program, r.shots, seed=r.seed).counts
    trajectory_workers = frontend.trajectory_workers()
    frontend.stop()
    print(json.dumps({"workers": workers, "trajectory_workers": trajectory_workers, "batch_same": batch_same, "noisy_same": noisy_same, "error_kept": error_kept, "serve_same": bool(serve_same)}), flush=True)


if __name__ == "__main__":
    main()
