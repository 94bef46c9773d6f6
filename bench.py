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
- ``cpu_baseline``: the oracle's NumPy tensordot (the reference's own contraction,
  engines.py:38-42) applying the circuit's first gates at 26 qubits on one host thread, in the
  same unit.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "lucivant-sim_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "26q gate-apply HBM GB/s; batched circuits/s + p99 latency vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
N = 26
BYTES_PER_PASS = 2 * 16 * (1 << N)


def cpu_baseline(circuit, budget_s: float = 15.0) -> dict:  # type: ignore[no-untyped-def]
    """The oracle applying gates one by one at 26 qubits (NumPy tensordot, one thread)."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from lucivant_sim.circuit import Gate
    from oracle import engine as orc

    gates = [op for op in circuit.operations if isinstance(op, Gate)]
    rng = np.random.default_rng(0)
    psi = (rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)).astype(np.complex128)
    psi /= np.linalg.norm(psi)
    done, elapsed = 0, 0.0
    with threadpool_limits(limits=1):
        while done < len(gates) and elapsed < budget_s:
            g = gates[done]
            t0 = time.perf_counter()
            psi = orc.apply_sv(psi, g.matrix, g.qubits, N)
            elapsed += time.perf_counter() - t0
            done += 1
    return {
        "value": round(done * BYTES_PER_PASS / elapsed / 1e9, 3),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": f"first {done} of {len(gates)} gates of the 26-qubit circuit, one NumPy tensordot per gate (unfused), "
        f"{elapsed:.1f} s on one host thread; GB/s counts 2 x 16 B x 2^26 per gate",
    }


def read_traffic() -> float | None:
    """HBM bytes per dense launch from the committed rocprofv3 counter summary, if present."""
    path = ROOT / "profiles" / "dense_gate_26q_traffic.json"
    if not path.exists():
        return None
    try:
        return float(json.loads(path.read_text())["bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=0, help="fusion width (0: the tuned default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(
            "bench.py measures one MI355X: the statevector path runs one state per GPU and this build has no "
            "multi-GPU sharding (run with --gpus 1)",
            file=sys.stderr,
        )
        return 2

    import numpy as np
    import torch

    import lucivant_sim as ls
    from lucivant_sim import _native as nat
    from lucivant_sim import simulator, workloads
    from lucivant_sim.engines import Statevector, upload
    from lucivant_sim.plan import plan_statevector

    width = args.width or simulator.FUSION_WIDTH
    src, _labels = workloads.config_c2()
    circuit = ls.parse(src)
    gates = sum(1 for op in circuit.operations if isinstance(op, ls.circuit.Gate))
    passes = plan_statevector(circuit.operations, N, width)
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
    torch.cuda.synchronize()
    for k, a, b in all_dense:
        durations[k].append(a.elapsed_time(b) * 1e-3)
    dense_times = [t for ts in durations.values() for t in ts]
    avg = float(np.mean(dense_times))
    achieved = BYTES_PER_PASS / avg / 1e9
    traffic = read_traffic()

    # sanity: the state stays normalised
    norm = float(torch.linalg.vector_norm(sv.buffer).item())
    assert abs(norm - 1) < 1e-9, norm
    del mats

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
            "passes": len(passes),
            "dense_passes": dense,
            "bytes_per_pass": BYTES_PER_PASS,
            "effective_unfused_GBps": round(gates * BYTES_PER_PASS / (elapsed / args.steps) / 1e9, 1),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "lsim::k_dense<K,double>",
            "avg_launch_ms": round(avg * 1e3, 4),
            "per_k_avg_ms": {str(k): round(float(np.mean(v)) * 1e3, 4) for k, v in durations.items() if v},
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(circuit)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
