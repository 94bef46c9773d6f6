"""Benchmark: 26-qubit complex128 gate application on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--width 4] [--no-cpu]

Workload (the 26-qubit statevector configuration): a random circuit of 40 layers, each u3 on all
26 qubits followed by cx or cz on a random disjoint pairing (seed 1002), 1,560 gates, planned
into fused passes of at most ``--width`` qubits. One step = one whole simulation from |0...0>
(state initialisation plus every fused pass) with the state resident in HBM (1 GiB).

Printed as one JSON line:
- ``value``: whole-job gate-apply throughput in GB/s = passes x 2 x 16 B x 2^26 / step time;
- ``roofline``: the dense fused-gate kernel, bytes per launch 2 x 16 x 2^26 over its average
  launch duration measured with HIP events on the launch stream, against the 8 TB/s HBM peak;
  ``traffic`` comes from profiles/ when a rocprofv3 counter summary is committed there;
- ``cpu_baseline``: the reference's own CPU route restated in the oracle (one-qubit runs fused,
  then its size-dependent slice dispatch, engines.py:45-119) applying the circuit's first gates
  at 26 qubits on one host thread, in the same unit;
- ``reference_runs``: the reference's own recorded seconds on the 12-, 13- and 20-qubit
  configurations beside the GPU seconds of the same runs here.
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "lucivant-sim_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "26q gate-apply HBM GB/s; batched circuits/s + p99 latency vs plain, 1 GPU"
_T0 = time.perf_counter()


def progress(what: str) -> None:
    """One line on stderr per finished piece of work (the JSON line stays alone on stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {what}", file=sys.stderr, flush=True)
HBM_PEAK_GBS = 8000.0
N = 26
BYTES_PER_PASS = 2 * 16 * (1 << N)


def cpu_baseline(circuit, budget_s: float = 15.0) -> dict:  # type: ignore[no-untyped-def]
    """The reference's own CPU route for this circuit, restated in the oracle, on one host thread:
    runs of one-qubit gates fused (simulator.py:196-236, as its ideal statevector evolution does),
    then each gate through its size-dependent dispatch at 26 qubits (engines.py:45-119): a batched
    2x2 product for one-qubit gates with at least 8 entries below their axis, slice moves for
    cx/cz, slice combinations for the lowest three qubits. Gates are applied in circuit order
    until the budget is spent; the per-gate mean gives the GB/s figure and the estimated seconds
    per whole simulation."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from lucivant_sim.circuit import Gate
    from oracle import engine as orc

    gates = [op for op in orc.fuse_one_qubit_runs(circuit.operations) if isinstance(op, Gate)]
    rng = np.random.default_rng(0)
    t = (rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)).astype(np.complex128)
    t /= np.linalg.norm(t)
    t = t.reshape([2] * N)
    done, elapsed = 0, 0.0
    routes = {"batched_2x2": 0, "slices": 0}
    with threadpool_limits(limits=1):
        while done < len(gates) and elapsed < budget_s:
            g = gates[done]
            axes = [N - 1 - q for q in g.qubits]
            low = len(axes) == 1 and (1 << (N - 1 - axes[0])) < orc.BATCHED_FROM_INNER
            routes["slices" if len(axes) == 2 or low else "batched_2x2"] += 1
            t0 = time.perf_counter()
            t = orc.apply_slices(t, g.matrix, axes)
            elapsed += time.perf_counter() - t0
            done += 1
    per_gate = elapsed / done
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {
        "value": round(done * BYTES_PER_PASS / elapsed / 1e9, 3),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": f"first {done} of the {len(gates)} gates left after one-qubit-run fusion of the 26-qubit circuit, "
        f"each through the reference's slice dispatch ({routes['batched_2x2']} batched 2x2 products, "
        f"{routes['slices']} slice updates), {elapsed:.1f} s on one host thread; GB/s counts 2 x 16 B x 2^26 per gate",
        "seconds_per_gate": round(per_gate, 4),
        "est_seconds_per_simulation": round(per_gate * len(gates), 1),
        "host_cpus": os.cpu_count(),
        "host_cpus_in_affinity": affinity,
        "threads": "threadpoolctl limit 1 (BLAS and OpenMP pools)",
    }


def reference_runs(gpu: dict) -> dict:
    """The reference's own seconds on the configurations it could run, recorded when the golden
    fixtures were written (tests/golden/runs.json, full_c3_reference.json: lucivant-sim 0.3.0 on an
    8-CPU container, thread pools unset), beside the GPU seconds of the same runs here."""
    gold = ROOT / "tests" / "golden"
    runs = json.loads((gold / "runs.json").read_text())
    c3 = json.loads((gold / "full_c3_reference.json").read_text())
    env = runs["env"]
    out = {
        "reference_host": f"{env['cpus']} CPUs, numpy {env['numpy']}, thread pools unset",
        "c1_12q_1000_shots": {"reference_s": round(runs["runs"]["c1_12q_seed7"]["wall_seconds"], 4), "gpu_s": gpu.get("c1")},
        "c3_13q_density_4000_shots": {"reference_s": round(c3["wall_seconds"], 1), "gpu_s": gpu.get("c3")},
        "c4_20q_256_shots": {"reference_s": round(runs["runs"]["c4_20q_256shots_seed7"]["wall_seconds"], 1), "gpu_s": gpu.get("c4")},
    }
    for row in out.values():
        if isinstance(row, dict) and row.get("gpu_s"):
            row["ratio"] = round(row["reference_s"] / row["gpu_s"], 1)
    return out


def plain_counts(circuit, shots: int, seed: int):  # type: ignore[no-untyped-def]
    """The plain GPU path for one circuit: a fresh HBM statevector, one kernel launch per gate
    (matrices uploaded once), probabilities back to the host, the reference's sampling rule."""
    import numpy as np

    from lucivant_sim.circuit import Gate
    from lucivant_sim.engines import Statevector, run_passes
    from lucivant_sim.plan import Pass, is_diagonal
    from lucivant_sim.simulator import _sample

    passes = []
    for op in circuit.operations:
        if isinstance(op, Gate):
            m = np.asarray(op.matrix)
            passes.append(Pass("diag", op.qubits, np.diag(m)) if is_diagonal(m) else Pass("dense", op.qubits, m))
    sv = Statevector(circuit.num_qubits)
    run_passes(sv.buffer, sv.dtype, sv.nbits, 1, passes)
    return _sample(circuit, sv.probabilities(), shots, None, np.random.default_rng(seed))


def plain_run(text: str, shots: int, seed: int):  # type: ignore[no-untyped-def]
    """``run`` on the plain GPU path: read the text, one launch per gate, sample, build the
    Result (with its circuit digest), one circuit at a time."""
    import time as _time

    import lucivant_sim as ls
    from lucivant_sim.batch import _result

    t0 = _time.perf_counter()
    c = ls.parse(text)
    return _result(c, plain_counts(c, shots, seed), shots, seed, "statevector", None, t0)


def bench_batched(count: int = 1024, plain_count: int = 48) -> dict:
    """1024 distinct 12-qubit circuits of the 12-qubit configuration's shape (20 layers of u3 on
    every qubit + cx on a random pairing, 1000 shots each, distinct seeds), from OpenQASM text to
    Results. plain = ``plain_run`` per circuit; batched = ``run_many`` over all texts: read in the
    front-end worker processes in chunks, one LDS-program launch and one sampling pass per chunk in
    this process, overlapped with the reading of later chunks. ``in_process`` = the same launch and sampling from already-read circuits, with no
    worker processes (host work in one thread)."""
    import numpy as np

    import lucivant_sim as ls
    from lucivant_sim import frontend, workloads
    from lucivant_sim.batch import counts_many, run_many

    texts = [workloads.random_layers(12, 20, 5000 + i) for i in range(count)]
    seeds = list(range(count))
    run_many(texts[:8], 1000, seeds=seeds[:8])  # warm-up (kernel load, worker imports)
    plain_run(texts[0], 1000, 0)
    t0 = time.perf_counter()
    lat = []
    plain_results = []
    for i in range(plain_count):
        a = time.perf_counter()
        plain_results.append(plain_run(texts[i], 1000, seeds[i]))
        lat.append(time.perf_counter() - a)
    plain = plain_count / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    res = run_many(texts, 1000, seeds=seeds)
    dt = time.perf_counter() - t0
    batched = count / dt
    same = all(res[i].counts == plain_results[i].counts and res[i].fingerprint == plain_results[i].fingerprint for i in range(plain_count))
    circuits = [ls.parse(t) for t in texts]
    counts_many(circuits[:8], 1000, seeds[:8])
    t0 = time.perf_counter()
    inproc = counts_many(circuits, 1000, seeds)
    dt_in = time.perf_counter() - t0
    same = same and all(inproc[i] == res[i].counts for i in range(count))
    kernel = program_roofline(circuits)
    return {
        "workload": f"{count} distinct 12-qubit circuits, 20 layers u3 + cx, 1000 shots each, complex128, OpenQASM text to Results",
        "plain_circuits_per_s": round(plain, 1),
        "plain_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "plain_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "batched_circuits_per_s": round(batched, 1),
        "batched_latency_ms": round(dt * 1e3, 2),
        "speedup": round(batched / plain, 2),
        "frontend_workers": frontend.workers(),
        "in_process_circuits_per_s": round(count / dt_in, 1),
        "in_process_note": "from already-read circuits, counts only (no reading, digests or Results): one thread of host work",
        "results_equal_plain": same,
        "roofline": kernel,
    }


#: MI355X dense FP64 peak (vector and matrix cores alike)
FP64_PEAK_TFLOPS = 78.6


def program_roofline(circuits, reps: int = 5) -> dict:  # type: ignore[no-untyped-def]
    """Duration of the one ``k_program`` launch that simulates every circuit of the batch (each
    workgroup holds one 12-qubit state in LDS and applies its op table), from HIP events recorded
    on the launch stream right around the launch; the flop count is that of the op table the
    kernel runs: 8 x 2^k flop per amplitude for a dense k-qubit operator, 6 for a diagonal one
    (one complex multiply), 3 per amplitude for the probabilities at the end."""
    import numpy as np

    from lucivant_sim import program as pg
    from lucivant_sim.optable import static_tables

    n = circuits[0].num_qubits
    table, pool, starts, rows = static_tables(circuits)
    kind, k = table["kind"].astype(np.int64), table["k"].astype(np.int64)
    per_amp = np.where(kind == 0, 8 * (1 << k), np.where(kind == 1, 6, 0))
    flops = float(per_amp.sum() + 3 * len(circuits)) * (1 << n)
    times = []
    for _ in range(reps):
        timer: list = []
        pg.run_table(n, table, pool, starts, rows, want_probs=True, timer=timer)
        times.append(timer[0].elapsed_time(timer[1]) * 1e-3)
    t = float(np.median(times))
    achieved = flops / t / 1e12
    dense_rows = int(np.sum(kind == 0))
    return {
        "bound": "fp64",
        "kernel": "lsim::k_program<double>: one workgroup per circuit, state in LDS",
        "achieved": round(achieved, 2),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
        "launch_ms": round(t * 1e3, 3),
        "ops": int(len(table)),
        "dense_ops": dense_rows,
        "mean_dense_width": round(float(k[kind == 0].mean()), 2) if dense_rows else 0,
        "flop_per_launch": flops,
    }


def bench_batched_shots(shots: int = 20000, plain_shots: int = 40) -> dict:
    """Trajectory shots of the 20-qubit configuration's recipe at 9 qubits (teleport chain with
    noise): plain = one shot at a time, one launch per operation (the reference's per-shot loop
    on the device); batched = every shot in one LDS-program launch (one workgroup per shot), in
    complex128 and complex64 (FP32 arithmetic, FP64 reductions and renormalisation)."""
    import numpy as np

    import lucivant_sim as ls
    from lucivant_sim import _native as nat
    from lucivant_sim import workloads
    from lucivant_sim.trajectories import run_lds, run_sequential

    c = ls.parse(workloads.teleport_chain(9, 3))
    noise = workloads.noise_c4()
    run_sequential(c, noise, 2, np.random.default_rng(0))
    t0 = time.perf_counter()
    plain_bits = run_sequential(c, noise, plain_shots, np.random.default_rng(7))
    plain = plain_shots / (time.perf_counter() - t0)
    out: dict = {"workload": f"9-qubit teleport chain with noise, {shots} shots", "plain_shots_per_s": round(plain, 1)}
    for name, dt in (("complex128", nat.C128), ("complex64", nat.C64)):
        run_lds(c, noise, 64, np.random.default_rng(0), dtype=dt)
        t0 = time.perf_counter()
        bits = run_lds(c, noise, shots, np.random.default_rng(7), dtype=dt)
        el = time.perf_counter() - t0
        assert bits is not None
        out[name] = {"shots_per_s": round(shots / el, 1), "seconds": round(el, 4), "speedup_vs_plain": round(shots / el / plain, 1)}
        if name == "complex128":
            ref_bits = bits
            out[name]["first_shots_equal_plain"] = [int(v) for v in bits[:plain_shots]] == plain_bits
        else:
            out[name]["shots_differing_from_complex128"] = int(np.sum(bits != ref_bits))
    return out


#: arrival rates of the serving stream (requests per second; 1000 = all at once, saturation)
SERVING_QPS = (10.0, 20.0, 30.0, 40.0, 1000.0)
#: requests per rate
SERVING_REQUESTS = 500
#: requests of the saturation subset that the single-request baselines also serve
COMPARED_REQUESTS = 64
#: widest count request that carries the noise model (run as trajectories)
NOISY_MAX_QUBITS = 16


def serving_requests(count: int, seed: int):  # type: ignore[no-untyped-def]
    """The serving mix: widths uniform in 10..26, depths in 10..100; even requests are
    hardware-efficient ansatz expectations (16-term Pauli sums, weights uniform in [-1, 1]), odd
    ones random u3 + cx circuits asking 1000 seeded shots, with the uniform noise model
    (depolarising 1e-3 / 1e-2, as trajectories) up to NOISY_MAX_QUBITS qubits."""
    import numpy as np

    from lucivant_sim import workloads
    from lucivant_sim.noise import NoiseModel
    from lucivant_sim.serving import Request

    rng = np.random.default_rng(seed)
    noise = NoiseModel.uniform(single=1e-3, two=1e-2)
    reqs = []
    for i in range(count):
        n = int(rng.integers(10, 27))
        depth = int(rng.integers(10, 101))
        if i % 2 == 0:
            src = workloads.hea_ansatz(n, depth, 7000 + i)
            obs = {lbl: float(w) for lbl, w in zip(workloads.pauli_strings(n, 16, 8000 + i), rng.uniform(-1, 1, 16))}
            reqs.append(Request(src, observable=obs))
        else:
            src = workloads.random_layers(n, depth, 9000 + i)
            noisy = n <= NOISY_MAX_QUBITS
            reqs.append(Request(src, shots=1000, seed=i, noise=noise if noisy else None, method="trajectories" if noisy else "automatic"))
    return reqs


def bench_serving(seed: int = 1005) -> dict:
    """The serving stream at each rate of SERVING_QPS: SERVING_REQUESTS requests of the mix
    (serving_requests) with Poisson arrivals, through the continuous-batching server. Then the
    first COMPARED_REQUESTS requests, all arriving at once, through the server and through the
    single-request API one at a time, both unfused (one launch per gate) and fused (the
    library's plans): circuits/s and latency of each, and the speedups of serving over both. The
    three must return the same results (counts equal, expectations within 1e-9)."""
    import numpy as np

    from lucivant_sim.serving import Request, Server, plain_stream, serve_stream

    points = []
    for qps in SERVING_QPS:
        reqs = serving_requests(SERVING_REQUESTS, seed)
        arrivals = list(np.cumsum(np.random.default_rng(seed + int(qps)).exponential(1.0 / qps, size=len(reqs))))
        stats = serve_stream(Server(), reqs, arrivals)
        errors = sum(1 for r in reqs if r.error is not None)
        progress(f"serving at {qps:g} requests/s: {stats.get('circuits_per_s', 0):.1f} circuits/s, p99 {stats.get('p99_ms', 0):.0f} ms")
        points.append({"qps": qps, "requests": len(reqs), "errors": errors, **{k: round(v, 3) if isinstance(v, float) else v for k, v in stats.items()}})

    def fresh(reqs):  # type: ignore[no-untyped-def]
        return [Request(r.program, observable=r.observable, shots=r.shots, seed=r.seed, noise=r.noise, method=r.method) for r in reqs]

    base = serving_requests(COMPARED_REQUESTS, seed)
    at_once = [0.0] * len(base)
    served = fresh(base)
    served_stats = serve_stream(Server(), served, at_once)
    progress("serving subset: served")
    unfused = fresh(base)
    unfused_stats = plain_stream(unfused, at_once, fused=False)
    progress("serving subset: single requests, unfused")
    fused = fresh(base)
    fused_stats = plain_stream(fused, at_once, fused=True)
    progress("serving subset: single requests, fused")

    def same(a, b):  # type: ignore[no-untyped-def]
        return all((abs(x.result - y.result) < 1e-9) if x.observable is not None else x.result == y.result for x, y in zip(a, b))

    def rnd(d):  # type: ignore[no-untyped-def]
        return {k: round(v, 3) if isinstance(v, float) else v for k, v in d.items()}

    return {
        "workload": f"{SERVING_REQUESTS} requests per rate: widths 10-26, depths 10-100; half 16-term Pauli-sum "
        f"expectations of HEA circuits, half 1000-shot counts of random u3 + cx circuits (noisy trajectories up to "
        f"{NOISY_MAX_QUBITS} qubits); latency from arrival",
        "points": points,
        "saturation_subset": {
            "requests": len(base),
            "served": rnd(served_stats),
            "single_request_unfused": rnd(unfused_stats),
            "single_request_fused": rnd(fused_stats),
            "speedup_vs_unfused": round(served_stats["circuits_per_s"] / unfused_stats["circuits_per_s"], 2),
            "speedup_vs_fused": round(served_stats["circuits_per_s"] / fused_stats["circuits_per_s"], 2),
            "results_equal": same(served, unfused) and same(served, fused),
        },
    }


def bench_c2_outputs() -> dict:
    """The 26-qubit configuration's outputs through the public API: ``probabilities(src)`` (one
    simulation, |psi|^2 on the device, 512 MiB to the host, then the reference's dict of every
    outcome above 1e-12) and ``expectation(src, P_j)`` of its 8 Pauli strings, one call each (each
    call simulates the circuit again, as the reference's does), plus all 8 as one weighted
    observable. Checks: the probabilities sum to 1, and the weighted call equals the weighted sum
    of the single calls."""
    import numpy as np

    import lucivant_sim as ls
    from lucivant_sim import workloads
    from lucivant_sim.simulator import _exact_state

    src, labels = workloads.config_c2()
    circuit = ls.parse(src)
    t0 = time.perf_counter()
    probs = _exact_state(circuit, None).probabilities()
    device_s = time.perf_counter() - t0
    total = float(np.sum(probs))
    del probs
    t0 = time.perf_counter()
    table = ls.probabilities(src)
    api_s = time.perf_counter() - t0
    entries = len(table)
    del table
    t0 = time.perf_counter()
    values = [ls.expectation(src, label) for label in labels]
    each_s = time.perf_counter() - t0
    weights = np.random.default_rng(1003).uniform(-1, 1, len(labels))
    t0 = time.perf_counter()
    combined = ls.expectation(src, dict(zip(labels, weights.tolist())))
    one_s = time.perf_counter() - t0
    linear = abs(combined - float(np.dot(weights, values))) < 1e-10
    return {
        "workload": "26-qubit configuration (40 layers u3 + cx/cz, seed 1002): probabilities() and expectation() of 8 Pauli strings (seed 1003)",
        "probabilities_api_s": round(api_s, 3),
        "probabilities_device_s": round(device_s, 3),
        "probabilities_note": "device_s = simulation, |psi|^2 and the copy to host memory; the rest of api_s is the reference's Python dict of every outcome above 1e-12",
        "probabilities_entries": entries,
        "probability_sum_minus_1": float(f"{total - 1:.3g}"),
        "expectation_8_calls_s": round(each_s, 3),
        "expectation_note": "every call simulates the whole circuit again, as the reference's does; the text is read and planned once and reused by later calls (recent-program cache)",
        "expectation_one_weighted_call_s": round(one_s, 4),
        "expectation_values": [round(v, 12) for v in values],
        "weighted_equals_sum_of_calls": linear,
    }


def bench_graph(reps: int = 20) -> dict:
    """Whole simulations replayed as one captured HIP graph against the same launches issued one
    by one from Python: random u3 + cx/cz circuits of 14, 18 and 22 qubits, depth 40, fused at
    width 5, where passes last a few microseconds. Also a variational sweep step: new angles
    copied into the captured matrices, then a replay. Both the per-simulation wall time and the
    agreement of replay and eager states are reported."""
    import numpy as np
    import torch

    import lucivant_sim as ls
    from lucivant_sim import workloads
    from lucivant_sim.graph import CompiledCircuit

    out = []
    for n in (14, 18, 22):
        c = ls.parse(workloads.random_layers(n, 40, 6000 + n, two_qubit=("cx", "cz"), measure=False))
        cc = CompiledCircuit(c, width=5)
        eager_state = cc.run_eager().buffer.clone()
        torch.cuda.synchronize()
        timings = {}
        for name, fn in (("eager", cc.run_eager), ("graph", cc.replay)):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            timings[name] = (time.perf_counter() - t0) / reps
        same = bool(torch.equal(cc.state.buffer, eager_state))
        out.append(
            {
                "qubits": n,
                "passes": len(cc.passes),
                "eager_ms": round(timings["eager"] * 1e3, 3),
                "graph_ms": round(timings["graph"] * 1e3, 3),
                "speedup": round(timings["eager"] / timings["graph"], 2),
                "replay_equals_eager": same,
            }
        )
        del cc, eager_state
    # a variational sweep on the hardware-efficient ansatz: update the angles, replay
    a = ls.parse(workloads.hea_ansatz(18, 20, 1))
    cc = CompiledCircuit(a, width=5)
    steps = [ls.parse(workloads.hea_ansatz(18, 20, 2 + i)) for i in range(reps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for b in steps:
        cc.update(b)
        cc.replay()
    torch.cuda.synchronize()
    sweep = (time.perf_counter() - t0) / reps
    return {
        "workload": "random u3 + cx/cz circuits, depth 40, fused at width 5: one captured HIP graph per circuit vs eager launches",
        "circuits": out,
        "sweep_18q_hea_20_layers_ms_per_step": round(sweep * 1e3, 3),
        "sweep_note": "per step: the fused matrices of new angles planned on the host and copied into the captured buffer, then one replay",
    }


def bench_trajectories(shots: int = 4096) -> dict:
    """The 20-qubit trajectory configuration (teleport chain, 6 mid-circuit measurements,
    uniform depolarising 1e-3/1e-2 and amplitude damping 0.01 on h) in complex128 and complex64."""
    import numpy as np

    import lucivant_sim as ls
    from lucivant_sim import _native as nat
    from lucivant_sim import workloads
    from lucivant_sim.trajectories import run_batched

    c = ls.parse(workloads.teleport_chain(20, 3))
    noise = workloads.noise_c4()
    out = {}
    bits = {}
    from lucivant_sim import trajectories

    for name, dt in (("complex128", nat.C128), ("complex64", nat.C64)):
        t0 = time.perf_counter()
        bits[name] = run_batched(c, noise, shots, np.random.default_rng(7), dtype=dt)
        el = time.perf_counter() - t0
        st = trajectories.STATS
        out[name] = {
            "seconds": round(el, 3),
            "shots_per_s": round(shots / el, 1),
            "host_round_trips": int(st["syncs"]),
            "launches_over_the_batch": int(st["passes"]),
            "seconds_waiting_for_the_device": round(st["wait_seconds"], 4),
            "seconds_of_host_work": round(el - st["wait_seconds"], 4),
        }
    a, b = bits["complex128"], bits["complex64"]
    ua, ca = np.unique(a, return_counts=True)
    ub, cb = np.unique(b, return_counts=True)
    pa = dict(zip(ua.tolist(), (ca / shots).tolist()))
    pb = dict(zip(ub.tolist(), (cb / shots).tolist()))
    keys = set(pa) | set(pb)
    out["workload"] = f"20-qubit teleport chain, {shots} shots in one batch (room for {shots} x 16 MiB complex128 states; shots with equal histories share one)"
    out["complex64_vs_complex128"] = {
        "shots_with_different_bits": int(np.sum(a != b)),
        "tvd": round(0.5 * sum(abs(pa.get(k, 0) - pb.get(k, 0)) for k in keys), 6),
    }
    return out


def bench_reference_configs() -> dict:
    """GPU seconds of the runs whose reference seconds are recorded in tests/golden: the 12-qubit
    configuration (1000 shots, seed 7) and the 20-qubit trajectory configuration (256 shots,
    seed 7), each through ls.run from OpenQASM text, after one warm-up; counts must equal the
    reference's."""
    import lucivant_sim as ls
    from lucivant_sim import workloads

    gold = ROOT / "tests" / "golden"
    progs = json.loads((gold / "programs.json").read_text())
    runs = json.loads((gold / "runs.json").read_text())["runs"]
    out = {}
    for key, name, noise in (("c1", "c1_12q_seed7", None), ("c4", "c4_20q_256shots_seed7", workloads.noise_c4())):
        rec = runs[name]
        ls.run(progs[name], rec["shots"], seed=rec["seed"], noise=noise, method=rec["method"])
        t0 = time.perf_counter()
        res = ls.run(progs[name], rec["shots"], seed=rec["seed"], noise=noise, method=rec["method"])
        out[key] = round(time.perf_counter() - t0, 4)
        assert res.counts == rec["counts"], name
    return out


def bench_density_matrix() -> dict:
    """The 13-qubit density-matrix configuration: depth 30 (seed 1004) under the uniform noise
    model with thermal relaxation on u3 and readout error; exact probabilities, then 4000 seeded
    shots with the density-matrix method. The 4^13-entry matrix is 1 GiB complex128."""
    import lucivant_sim as ls
    from lucivant_sim import workloads
    from lucivant_sim.plan import Pass, plan_density

    src = workloads.config_c3()
    noise = workloads.noise_c3()
    circuit = ls.parse(src)
    ls.probabilities(workloads.config_c3(4, 2), noise)  # warm-up
    t0 = time.perf_counter()
    probs = ls.probabilities(circuit, noise)
    t_probs = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = ls.run(circuit, 4000, seed=7, noise=noise, method="density_matrix")
    t_run = time.perf_counter() - t0
    passes = [p for p in plan_density(circuit.operations, 13, noise) if isinstance(p, Pass)]
    kernel = passes_roofline(passes, 26)
    return {
        "workload": "13-qubit density matrix, depth 30 u3 + cx/cz, depolarising 1e-3/1e-2, thermal relaxation on u3, readout 0.01/0.03",
        "superoperator_passes": len(passes),
        "roofline": kernel,
        "probabilities_seconds": round(t_probs, 3),
        "run_4000_shots_seconds": round(t_run, 3),
        "distinct_outcomes": len(res.counts),
        "probability_sum": round(sum(probs.values()), 12),
    }


def passes_roofline(passes, nbits: int) -> dict:  # type: ignore[no-untyped-def]
    """Per-launch durations (HIP events on the launch stream) of ``passes`` run once on a
    scratch vector of 2^nbits complex128 entries from |0>: the dense and diagonal launches,
    each a read and a write of the vector, 2 x 16 B x 2^nbits."""
    import numpy as np
    import torch

    from lucivant_sim import _native as nat
    from lucivant_sim.engines import Statevector, plan_launches

    sv = Statevector(nbits)
    stream = torch.cuda.current_stream()
    keep, calls = plan_launches(sv.buffer, nat.C128, nbits, 1, passes, stream.cuda_stream)
    events = []
    for fn, a in calls:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        assert fn(*a) == 0
        e1.record(stream)
        events.append((e0, e1))
    torch.cuda.synchronize()
    durations = [a.elapsed_time(b) * 1e-3 for a, b in events]
    del keep, sv
    torch.cuda.empty_cache()
    nbytes = 2 * 16 * (1 << nbits)
    avg = float(np.mean(durations))
    return {
        "bound": "hbm",
        "kernel": "lsim::k_dense / k_dense_mfma / k_diag over the 2^26-entry vector (superoperators on row and column bits)",
        "achieved": round(nbytes / avg / 1e9, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(nbytes / avg / 1e9 / HBM_PEAK_GBS, 4),
        "avg_launch_ms": round(avg * 1e3, 4),
        "launches": len(durations),
        "evolution_device_ms": round(sum(durations) * 1e3, 2),
        "bytes_per_launch": nbytes,
    }


def diagonal_strings(n: int, count: int, seed: int) -> list[str]:
    """``count`` random strings of only I and Z on ``n`` qubits (each letter Z with probability 1/2)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    return ["".join("Z" if z else "I" for z in rng.integers(0, 2, n)) for _ in range(count)]


def bench_pauli(buf, labels, reps: int = 20) -> dict:  # type: ignore[no-untyped-def]
    """Pauli expectations on the resident 26-qubit state (2^26 complex128 amplitudes, 1 GiB):
    the 26-qubit configuration's 8 strings in one ``lsim_sv_pauli_expect`` call (strings whose X
    masks span few enough bits share one read of the state), and the read-only ceiling: a single
    Z string, whose reduction reads every amplitude once and writes nothing. HIP events on the
    launch stream bracket each call; GB/s counts 16 B x 2^26 per read of the state."""
    import numpy as np
    import torch

    from lucivant_sim import _native as nat
    from lucivant_sim.engines import launch_expectations

    stream = torch.cuda.current_stream()
    state_bytes = 16 * (1 << N)

    def timed(lbls) -> float:  # type: ignore[no-untyped-def]
        launch_expectations("lsim_sv_pauli_expect", buf, nat.C128, N, N, lbls)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            launch_expectations("lsim_sv_pauli_expect", buf, nat.C128, N, N, lbls)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e-3 / reps

    t8 = timed(labels)
    t1 = timed(["Z" + "I" * (N - 1)])
    t16 = timed(diagonal_strings(N, 16, 1004))
    return {
        "workload": "26-qubit state resident in HBM: 8 Pauli strings (seed 1003) in one call, one Z string, and 16 strings of I and Z (seed 1004) in one call",
        "strings_8_ms": round(t8 * 1e3, 4),
        "string_1_ms": round(t1 * 1e3, 4),
        "diagonal_strings_16_ms": round(t16 * 1e3, 4),
        "read_only_GBps": round(state_bytes / t1 / 1e9, 1),
        "strings_8_over_1": round(t8 / t1, 2),
        "roofline": {
            "bound": "hbm",
            "kernel": "lsim::k_pauli_diag (one Z string): one read of the state per group of strings",
            "achieved": round(state_bytes / t1 / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(state_bytes / t1 / 1e9 / HBM_PEAK_GBS, 4),
        },
    }


def bench_tile_plan(circuit, steps: int, dense_ms: float, budget: float = 64.0, low: int = 4) -> dict:  # type: ignore[no-untyped-def]
    """The same 26-qubit simulation planned into tile passes instead: each pass loads tiles of
    4096 amplitudes (12 state bits, the 4 lowest among them) into registers, applies a run of
    up to 64 one- and two-qubit operators to them (regrouping registers through LDS when an
    operator's bits are not held per thread) and stores them back, so one sweep of HBM covers
    more gates than a 5-qubit dense pass. Per-launch durations of ``k_tile`` are measured with
    HIP events on the launch stream; GB/s counts 2 x 16 B x 2^26 per pass."""
    import numpy as np
    import torch

    from lucivant_sim import _native as nat
    from lucivant_sim.engines import Statevector, launch_all, plan_launches
    from lucivant_sim.plan import TilePass, plan_tiles

    plan = plan_tiles(circuit.operations, N, budget=budget, low=low)
    tiles = [p for p in plan if isinstance(p, TilePass)]
    sv = Statevector(N)
    stream = torch.cuda.current_stream()
    hs = stream.cuda_stream
    keep, calls = plan_launches(sv.buffer, nat.C128, N, 1, plan, hs)
    lib = nat.load()
    state = sv.buffer.data_ptr()

    def step() -> None:
        assert lib.lsim_sv_set_basis(state, nat.C128, N, 1, 0, hs) == 0
        launch_all(calls)

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    events = []
    lib.lsim_sv_set_basis(state, nat.C128, N, 1, 0, hs)
    for fn, a in calls:
        if fn is lib.lsim_sv_apply_tile:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn(*a)
            e1.record(stream)
            events.append((e0, e1))
        else:
            fn(*a)
    torch.cuda.synchronize()
    avg = float(np.mean([a.elapsed_time(b) * 1e-3 for a, b in events])) if events else float("nan")
    norm = float(torch.linalg.vector_norm(sv.buffer).item())
    del keep, sv
    torch.cuda.empty_cache()
    ops = [len(t.ops) for t in tiles]
    return {
        "workload": "the same 26-qubit simulation, planned into tile passes (planner budget "
        f"{budget:g}, {low} lowest bits always in the tile)",
        "passes": len(plan),
        "tile_passes": len(tiles),
        "ops_per_tile_pass": round(float(np.mean(ops)), 1) if ops else 0,
        "ms_per_step": round(ms, 3),
        "speedup_vs_dense_plan": round(dense_ms / ms, 3),
        "k_tile_avg_launch_ms": round(avg * 1e3, 4),
        "k_tile_GBps": round(BYTES_PER_PASS / avg / 1e9, 1),
        "k_tile_frac_of_hbm_peak": round(BYTES_PER_PASS / avg / 1e9 / HBM_PEAK_GBS, 4),
        "norm_minus_1": float(f"{norm - 1:.3g}"),
        "note": "fewer sweeps of HBM at a lower rate per sweep: the headline line stays on the dense plan, whose "
        "kernel runs at the HBM roofline; the library's autotuner picks whichever plan is faster per size",
    }


PROFILE_LEGS = ("dense", "tile", "pauli", "density", "trajectories", "program", "marginal")


def profile_leg(leg: str, circuit, labels, steps: int) -> None:  # type: ignore[no-untyped-def]
    """The kernels of one bench leg, ``steps`` times, for a profiler run (rocprofv3 names the
    kernels; each leg that sweeps a 1 GiB vector adds one k_scale pass per step, a streaming
    read and write of exactly 2^30 B, to calibrate the HBM counters):
    tile = the 26-qubit simulation as tile passes (budget 64, 4 low bits);
    pauli = the 8 Pauli strings and one Z string on the resident 26-qubit state;
    density = the 13-qubit density-matrix probabilities (superoperator passes over 4^13 entries);
    trajectories = 1024 shots of the 20-qubit teleport chain with noise in one batch;
    program = 1024 distinct 12-qubit circuits in one LDS-program launch;
    marginal = device marginals of the 26-qubit configuration's probabilities (2^26 float64)
    over its top 10 qubits (pieces of 8192 entries) and its lowest 10 (chains of 2^16)."""
    import numpy as np
    import torch

    import lucivant_sim as ls
    from lucivant_sim import _native as nat
    from lucivant_sim import workloads
    from lucivant_sim.engines import Statevector, launch_all, launch_expectations, plan_launches, run_plan
    from lucivant_sim.plan import plan_statevector, plan_tiles

    lib = nat.load()
    hs = torch.cuda.current_stream().cuda_stream
    if leg in ("tile", "pauli"):
        sv = Statevector(N)
        state = sv.buffer.data_ptr()
        if leg == "tile":
            plan = plan_tiles(circuit.operations, N, budget=64.0, low=4)
            keep, calls = plan_launches(sv.buffer, nat.C128, N, 1, plan, hs)
            for _ in range(steps):
                assert lib.lsim_sv_set_basis(state, nat.C128, N, 1, 0, hs) == 0
                launch_all(calls)
                assert lib.lsim_sv_scale(state, nat.C128, N, 1, 1.0, 0.0, hs) == 0
            del keep
            what = f"{steps} steps of {len(plan)} tile-plan passes"
        else:
            run_plan(sv.buffer, nat.C128, N, 1, plan_statevector(circuit.operations, N, 5))
            diagonal = diagonal_strings(N, 16, 1004)
            for _ in range(steps):
                launch_expectations("lsim_sv_pauli_expect", sv.buffer, nat.C128, N, N, labels)
                launch_expectations("lsim_sv_pauli_expect", sv.buffer, nat.C128, N, N, ["Z" + "I" * (N - 1)])
                launch_expectations("lsim_sv_pauli_expect", sv.buffer, nat.C128, N, N, diagonal)
                assert lib.lsim_sv_scale(state, nat.C128, N, 1, 1.0, 0.0, hs) == 0
            what = f"{steps} x (8 strings, 1 string, 16 diagonal strings, 1 k_scale)"
    elif leg == "density":
        src, noise = workloads.config_c3(), workloads.noise_c3()
        c = ls.parse(src)
        for _ in range(steps):
            ls.probabilities(c, noise)
        dm = Statevector(26)
        assert lib.lsim_sv_scale(dm.buffer.data_ptr(), nat.C128, 26, 1, 1.0, 0.0, hs) == 0
        what = f"{steps} density-matrix probability runs + 1 k_scale"
    elif leg == "marginal":
        sv = Statevector(N)
        state = sv.buffer.data_ptr()
        run_plan(sv.buffer, nat.C128, N, 1, plan_statevector(circuit.operations, N, 5))
        probs = sv.device_probabilities()
        out = torch.empty(1 << 10, dtype=torch.float64, device=probs.device)
        ws = torch.empty(1 << 13, dtype=torch.float64, device=probs.device)
        for _ in range(steps):
            for keep in (list(range(16, 26)), list(range(10))):
                karr = nat.int_array(keep)
                assert lib.lsim_marginal_workspace_bytes(N, 1, len(keep), karr) <= ws.numel() * 8
                assert lib.lsim_marginal(probs.data_ptr(), N, 1, len(keep), karr, out.data_ptr(), ws.data_ptr(), hs) == 0
            assert lib.lsim_sv_scale(state, nat.C128, N, 1, 1.0, 0.0, hs) == 0
        what = f"{steps} x (marginal of the top 10 qubits, marginal of the lowest 10, 1 k_scale)"
    elif leg == "trajectories":
        from lucivant_sim.trajectories import run_batched

        c = ls.parse(workloads.teleport_chain(20, 3))
        for _ in range(steps):
            run_batched(c, workloads.noise_c4(), 1024, np.random.default_rng(7))
        what = f"{steps} x 1024 trajectory shots"
    else:
        from lucivant_sim.batch import counts_many

        circuits = [ls.parse(workloads.random_layers(12, 20, 5000 + i)) for i in range(1024)]
        for _ in range(steps):
            counts_many(circuits, 1000, list(range(1024)))
        what = f"{steps} x 1024 programs"
    torch.cuda.synchronize()
    print(f"profiled {leg}: {what}", flush=True)


def read_traffic() -> float | None:
    """HBM bytes per dense launch from the committed rocprofv3 counter summary, if present
    (its ``dense_bytes_per_launch:`` line: FETCH_SIZE and WRITE_SIZE scaled by a plain copy)."""
    path = ROOT / "profiles" / "dense_gate_26q_traffic.txt"
    if not path.exists():
        return None
    m = re.search(r"^dense_bytes_per_launch:\s*([0-9.]+)", path.read_text(), re.MULTILINE)
    return float(m.group(1)) if m else None


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=0, help="fusion width (0: run the autotuner and use its choice)")
    ap.add_argument("--tiling", type=int, choices=(0, 1), default=None, help="force the LDS-tiled dense kernel off/on")
    ap.add_argument("--mfma", type=int, default=None, help="mask of operator widths on the matrix cores (bit k: k qubits)")
    ap.add_argument("--streaming", type=int, choices=(0, 1, 2), default=None, help="non-temporal accesses: 0 never, 1 by state size, 2 always")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched, serving and trajectory legs")
    ap.add_argument(
        "--profile",
        nargs="?",
        const="dense",
        default=None,
        choices=PROFILE_LEGS,
        help="only run one leg's kernels for rocprofv3 (dense: the headline passes, each step "
        "followed by one k_scale pass, a read+write of the 1 GiB state used to calibrate HBM "
        "counters); prints no JSON",
    )
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(
            "bench.py measures one MI355X: the statevector path runs one state per GPU and the library has no "
            "multi-GPU sharding (run with --gpus 1)",
            file=sys.stderr,
        )
        return 2

    if not args.no_extras and not args.profile:
        # the front-end workers are forked before this process first uses the GPU
        from lucivant_sim import frontend

        frontend.start()

    import numpy as np
    import torch

    import lucivant_sim as ls
    from lucivant_sim import _native as nat
    from lucivant_sim import simulator, workloads
    from lucivant_sim.engines import Statevector, upload
    from lucivant_sim.plan import plan_statevector

    src, _labels = workloads.config_c2()
    circuit = ls.parse(src)
    if args.profile not in (None, "dense"):
        profile_leg(args.profile, circuit, _labels, max(args.steps, 1))
        return 0
    tuned = None
    if args.tiling is not None:
        nat.load().lsim_set_dense_tiling(args.tiling)
    if args.mfma is not None:
        nat.load().lsim_set_dense_mfma(args.mfma)
    if args.streaming is not None:
        nat.load().lsim_set_dense_streaming(args.streaming)
    if args.width:
        width = args.width
    elif args.profile:
        width = simulator.FUSION_WIDTH
    else:
        from lucivant_sim import autotune

        tuned = autotune.tune(circuit, store=False)
        progress("autotuner done")
        width = tuned["width"]
        nat.load().lsim_set_dense_tiling(tuned["tiled"])
        nat.load().lsim_set_dense_mfma(tuned["mfma"])
        nat.load().lsim_set_dense_streaming(tuned["streaming"])
    gates = sum(1 for op in circuit.operations if isinstance(op, ls.circuit.Gate))
    low_bits = tuned["low_bits"] if tuned else 0
    low_width = tuned["low_width"] if tuned else 0
    passes = plan_statevector(circuit.operations, N, width, low_bits, low_width)
    dense = sum(1 for p in passes if p.kind == "dense")

    sv = Statevector(N)
    stream = torch.cuda.current_stream()
    hs = stream.cuda_stream
    mats, addrs = upload([p.matrix for p in passes])
    calls = []
    for p, addr in zip(passes, addrs):
        fn = "lsim_sv_apply_dense" if p.kind == "dense" else "lsim_sv_apply_diag"
        calls.append((getattr(nat.load(), fn), p.kind, p.k, nat.int_array(p.bits), addr))
    state = sv.buffer.data_ptr()
    lib = nat.load()

    def step() -> None:
        rc = lib.lsim_sv_set_basis(state, nat.C128, N, 1, 0, hs)
        assert rc == 0
        for fn, _kind, k, bits, addr in calls:
            rc = fn(state, nat.C128, N, 1, k, bits, addr, None, hs)
            assert rc == 0, lib.lsim_last_error()

    if args.profile:
        for _ in range(max(args.steps, 1)):
            step()
            lib.lsim_sv_scale(state, nat.C128, N, 1, 1.0, 0.0, hs)
        torch.cuda.synchronize()
        print(f"profiled {args.steps} steps of {len(passes)} passes ({dense} dense) plus one scale pass each", flush=True)
        return 0

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = len(passes) * BYTES_PER_PASS / (elapsed / args.steps) / 1e9

    # dominant kernel: every dense launch bracketed by HIP events on the launch stream
    durations = {1: [], 2: [], 3: [], 4: [], 5: []}
    all_dense = []
    for _ in range(2):
        lib.lsim_sv_set_basis(state, nat.C128, N, 1, 0, hs)
        for fn, kind, k, bits, addr in calls:
            if kind != "dense":
                fn(state, nat.C128, N, 1, k, bits, addr, None, hs)
                continue
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn(state, nat.C128, N, 1, k, bits, addr, None, hs)
            b.record(stream)
            all_dense.append((k, a, b))
    # the attainable stream rate on this device: lsim_sv_scale reads and writes the same 1 GiB
    # once (k_scale), timed the same way right after the dense launches
    copies = []
    for _ in range(16):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        lib.lsim_sv_scale(state, nat.C128, N, 1, 1.0, 0.0, hs)
        b.record(stream)
        copies.append((a, b))
    # an out-of-place copy of the state into a second 1 GiB buffer (torch's vectorised copy
    # kernel): the same read + write bytes with distinct source and destination
    other = torch.empty_like(sv.buffer)
    outs = []
    for _ in range(8):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        other.copy_(sv.buffer)
        b.record(stream)
        outs.append((a, b))
    torch.cuda.synchronize()
    del other
    for k, a, b in all_dense:
        durations[k].append(a.elapsed_time(b) * 1e-3)
    copy_avg = float(np.mean([a.elapsed_time(b) * 1e-3 for a, b in copies]))
    out_of_place_avg = float(np.mean([a.elapsed_time(b) * 1e-3 for a, b in outs[1:]]))
    dense_times = [t for ts in durations.values() for t in ts]
    avg = float(np.mean(dense_times))
    achieved = BYTES_PER_PASS / avg / 1e9
    traffic = read_traffic()

    # sanity: the state stays normalised
    norm = float(torch.linalg.vector_norm(sv.buffer).item())
    assert abs(norm - 1) < 1e-9, norm
    del mats
    pauli = None if args.no_extras else bench_pauli(sv.buffer, _labels)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "26-qubit complex128 statevector, 40 layers of u3 on every qubit + cx/cz on a random pairing "
            "(seed 1002), one full simulation from |0> per step",
            "qubits": N,
            "gates": gates,
            "fusion_width": width,
            "low_qubit_width": low_width or None,
            "matrix_core_widths": [k for k in (3, 4, 5) if (nat.load().lsim_set_dense_mfma(-1) >> k) & 1],
            "streaming_mode": int(nat.load().lsim_set_dense_streaming(-1)),
            "passes": len(passes),
            "dense_passes": dense,
            "bytes_per_pass": BYTES_PER_PASS,
            "effective_unfused_GBps": round(gates * BYTES_PER_PASS / (elapsed / args.steps) / 1e9, 1),
            "autotune_ms": {k: round(v * 1e3, 2) for k, v in tuned["timings"].items()} if tuned else None,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "lsim::k_dense<K,double,NT> and lsim::k_dense_mfma<K,NT>: every dense pass of the step",
            "avg_launch_ms": round(avg * 1e3, 4),
            "per_k_avg_ms": {str(k): round(float(np.mean(v)) * 1e3, 4) for k, v in durations.items() if v},
            "copy_kernel_GBps": round(BYTES_PER_PASS / copy_avg / 1e9, 1),
            "frac_of_copy": round(copy_avg / avg, 4),
            "out_of_place_copy_GBps": round(BYTES_PER_PASS / out_of_place_avg / 1e9, 1),
            "frac_of_out_of_place_copy": round(out_of_place_avg / avg, 4),
        },
    }
    del sv
    torch.cuda.empty_cache()
    if pauli is not None:
        # one Pauli-Z string is a read-only reduction of the state: the read ceiling
        out["roofline"]["read_only_GBps"] = pauli["read_only_GBps"]
    progress(f"headline: {ms_per_step:.2f} ms per step, {value:.0f} GB/s")
    if not args.no_extras:
        legs = [
            ("pauli_26q", lambda: pauli),
            ("tile_plan", lambda: bench_tile_plan(circuit, min(args.steps, 10), ms_per_step)),
            ("batched", bench_batched),
            ("batched_shots", bench_batched_shots),
            ("serving", bench_serving),
            ("outputs_26q", bench_c2_outputs),
            ("graph", bench_graph),
            ("trajectories_20q", bench_trajectories),
            ("density_matrix_13q", bench_density_matrix),
        ]
        for name, leg in legs:
            out[name] = leg()
            progress(f"{name} done")
    if not args.no_extras:
        gpu_s = bench_reference_configs()
        gpu_s["c3"] = out["density_matrix_13q"]["run_4000_shots_seconds"]
        out["reference_runs"] = reference_runs(gpu_s)
        progress("reference_runs done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(circuit)
        progress("cpu_baseline done")
    print(json.dumps(out), flush=True)
    if not args.no_extras:
        from lucivant_sim import frontend

        frontend.stop()
    return 0


if __name__ == "__main__":
    sys.exit(main())
