"""Full-size fixtures: the 26-qubit statevector and 13-qubit density-matrix configurations.

    python tests/golden/make_full_size.py oracle-c2      # ~15 min, NumPy oracle
    python tests/golden/make_full_size.py reference-c3  # ~50 min, the reference itself

``oracle-c2`` runs the oracle (oracle/engine.py, pinned by tests/test_oracle.py) on the 26-qubit
configuration and records what a device run must reproduce within 1e-10: 4096 amplitudes at
seeded indices, probability checksums, and the 8 Pauli expectations.

``reference-c3`` imports lucivant-sim 0.3.0 from the reference checkout (this container only)
and records the 13-qubit density-matrix configuration's counts for 4000 shots, seed 7 (counts
are pinned by the reference alone), with its fingerprint and its wall time on this CPU.
"""

from __future__ import annotations

import importlib.util
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]


def load_workloads():  # type: ignore[no-untyped-def]
    spec = importlib.util.spec_from_file_location("lsim_workloads", REPO / "lucivant-sim_amd" / "lucivant_sim" / "workloads.py")
    assert spec is not None and spec.loader is not None
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def oracle_c2(out: Path) -> None:
    sys.path.insert(0, str(REPO / "lucivant-sim_amd"))
    sys.path.insert(0, str(REPO))
    import lucivant_sim as ls
    from oracle import engine as orc

    wl = load_workloads()
    src, labels = wl.config_c2()
    t0 = time.perf_counter()
    psi = orc.evolve_sv(ls.parse(src))
    evolve_s = time.perf_counter() - t0
    n = 26
    idx = np.sort(np.random.default_rng(2026).choice(1 << n, size=4096, replace=False))
    probs = np.abs(psi) ** 2
    expect = [orc.expectation_sv(psi, lbl, n) for lbl in labels]
    rec = {
        "indices": idx.tolist(),
        "amplitudes": [[float(z.real), float(z.imag)] for z in psi[idx]],
        "prob_sum": float(probs.sum()),
        "prob_index_mean": float(np.dot(probs, np.arange(1 << n, dtype=np.float64)) / (1 << n)),
        "prob_max": float(probs.max()),
        "prob_argmax": int(probs.argmax()),
        "labels": labels,
        "expectations": expect,
        "oracle_evolve_seconds": evolve_s,
    }
    (out / "full_c2_oracle.json").write_text(json.dumps(rec))
    print("wrote full_c2_oracle.json", evolve_s)


def reference_c3(out: Path, reference: str) -> None:
    sys.path.insert(0, reference)
    import lucivant_sim as ref

    wl = load_workloads()
    src = wl.config_c3()
    t0 = time.perf_counter()
    res = ref.run(src, 4000, seed=7, noise=wl.noise_c3(), method="density_matrix")
    dt = time.perf_counter() - t0
    rec = res.to_dict()
    rec["wall_seconds"] = dt
    (out / "full_c3_reference.json").write_text(json.dumps(rec))
    print("wrote full_c3_reference.json", dt)


if __name__ == "__main__":
    what = sys.argv[1]
    dest = Path(sys.argv[2]) if len(sys.argv) > 2 else HERE
    if what == "oracle-c2":
        oracle_c2(dest)
    elif what == "reference-c3":
        reference_c3(dest, "/root/reference/src")
    else:
        raise SystemExit(f"unknown fixture {what!r}")
