"""Write the golden fixtures: what only the reference implementation itself can pin.

    python tests/golden/make_golden.py [--reference /root/reference/src]

Runs lucivant-sim 0.3.0 (imported from the reference checkout, never shipped) on the project's
seeded workload programs and records, as JSON under tests/golden/:
- exact counts of ``run`` for fixed seeds (statevector sampling, density-matrix sampling with
  readout flips, batched-order and sequential trajectories), with the circuit digest, noise
  digest, method, fingerprint and counts digest of each result;
- circuit digests of programs covering the OpenQASM reader's features;
- ``choose_method``'s answer for each configuration;
- wall time of each call on this container's CPU (the reference CPU baseline).

Amplitudes, probabilities and expectations are not recorded: the oracle pins those from the
mathematics (see oracle/__init__.py). The workload generator is loaded from its file so the
programs are byte-identical to the ones the device build runs.
"""

from __future__ import annotations

import argparse
import importlib.util
import json
import os
import platform
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]


def load_workloads():  # type: ignore[no-untyped-def]
    spec = importlib.util.spec_from_file_location("lsim_workloads", REPO / "lucivant-sim_amd" / "lucivant_sim" / "workloads.py")
    assert spec is not None and spec.loader is not None
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


BELL = """
OPENQASM 3;
include "stdgates.inc";
qubit[2] q;
bit[2] c;
h q[0];
cx q[0], q[1];
c = measure q;
"""

TELEPORT = """OPENQASM 3.0;
include "stdgates.inc";
qubit[3] q;
bit[2] m;
bit[1] out;
ry(1.234) q[0];
h q[1];
cx q[1], q[2];
cx q[0], q[1];
h q[0];
m[0] = measure q[0];
m[1] = measure q[1];
if (m[1]) x q[2];
if (m[0]) z q[2];
out[0] = measure q[2];
"""

READER_PROGRAMS = {
    "bell3": BELL,
    "bell2": 'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[2];\ncreg c[2];\nh q[0];\nCX q[0],q[1];\nmeasure q -> c;',
    "broadcast": "OPENQASM 3; qubit[4] q; qubit[4] r; h q; cx q, r; x q[1:2]; z q[{0, 3}]; y q[-1];",
    "modifiers": """
    OPENQASM 3;
    include "stdgates.inc";
    input float theta;
    const float half = theta / 2;
    gate rot(a) t { ry(a) t; }
    gate pair(a) x, y { rot(a) x; ctrl @ rot(2 * a) x, y; }
    qubit[3] q;
    pair(half) q[0], q[1];
    inv @ s q[2];
    pow(2) @ t q[2];
    negctrl @ x q[0], q[2];
    """,
    "teleport": TELEPORT,
    "if_else": """
    OPENQASM 3; include "stdgates.inc";
    qubit[2] q; bit[2] c;
    x q[0];
    c[0] = measure q[0];
    if (c[0] == 1) { c[0] = measure q[1]; } else { x q[1]; }
    c[1] = measure q[1];
    """,
    "registers": "OPENQASM 2.0; qreg q[2]; creg a[1]; creg b[1]; x q[1]; measure q[0] -> a[0]; measure q[1] -> b[0];",
    "physical": "OPENQASM 3; x $2; h $0;",
    "gphase": "OPENQASM 3;\npragma something\nqubit q;\ngphase(pi);\nx q;",
    "qelib_wide": 'OPENQASM 2.0; include "qelib1.inc"; qreg q[5]; h q; ccx q[0],q[1],q[2]; c3x q[0],q[1],q[2],q[3]; '
    "c4x q[0],q[1],q[2],q[3],q[4]; rccx q[1],q[2],q[3]; rc3x q[0],q[1],q[2],q[4]; cswap q[0],q[3],q[4]; "
    "cu(0.1,0.2,0.3,0.4) q[1],q[0]; rzz(0.7) q[2],q[3]; ecr q[0],q[1]; iswap q[1],q[2]; sx q[3]; u2(0.5,0.6) q[4];",
}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference/src")
    ap.add_argument("--out", default=str(HERE))
    ap.add_argument("--skip-slow", action="store_true", help="skip the 20-qubit trajectory fixture")
    args = ap.parse_args()
    sys.path.insert(0, args.reference)
    import lucivant_sim as ref  # the reference implementation
    from lucivant_sim.simulator import choose_method

    wl = load_workloads()
    out = Path(args.out)
    env = {
        "numpy": __import__("numpy").__version__,
        "python": platform.python_version(),
        "cpus": os.cpu_count(),
        "threads_env": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")},
        "reference_version": ref.__version__,
    }

    def result_record(res, seconds: float) -> dict:  # type: ignore[no-untyped-def]
        d = res.to_dict()
        d["wall_seconds"] = seconds
        return d

    def timed_run(src: str, **kw):  # type: ignore[no-untyped-def]
        t0 = time.perf_counter()
        res = ref.run(src, **kw)
        return res, time.perf_counter() - t0

    runs: dict[str, dict] = {}
    programs: dict[str, str] = {}

    # statevector sampling: the README Bell run and the 12-qubit configuration (median of 5 times)
    res, dt = timed_run(BELL, shots=1000, seed=7)
    runs["bell_seed7"] = result_record(res, dt)
    programs["bell_seed7"] = BELL
    c1 = wl.config_c1()
    times = []
    for _ in range(5):
        res, dt = timed_run(c1, shots=1000, seed=7)
        times.append(dt)
    rec = result_record(res, sorted(times)[2])
    rec["wall_seconds_all"] = times
    runs["c1_12q_seed7"] = rec
    programs["c1_12q_seed7"] = c1

    # density-matrix sampling with readout flips, the 13-qubit configuration's recipe at 6 qubits
    dm_src = wl.config_c3(6, 8)
    res, dt = timed_run(dm_src, shots=4000, seed=7, noise=wl.noise_c3(), method="density_matrix")
    runs["c3_recipe_6q_seed7"] = result_record(res, dt)
    programs["c3_recipe_6q_seed7"] = dm_src

    # trajectories, draw-balanced (batched on the device): the 20-qubit chain's recipe at 9 qubits
    tc = wl.teleport_chain(9, 3)
    res, dt = timed_run(tc, shots=3000, seed=7, noise=wl.noise_c4(), method="trajectories")
    runs["c4_recipe_9q_seed7"] = result_record(res, dt)
    programs["c4_recipe_9q_seed7"] = tc

    # trajectories with unbalanced arms (sequential on the device): the teleport example, noisy
    noise_t = ref.NoiseModel.uniform(single=0.01, two=0.02, readout=0.01)
    res, dt = timed_run(TELEPORT, shots=400, seed=3, noise=noise_t)
    runs["teleport_noisy_seed3"] = result_record(res, dt)
    programs["teleport_noisy_seed3"] = TELEPORT

    # an ideal dynamic circuit (automatic -> trajectories, no noise draws)
    res, dt = timed_run(TELEPORT, shots=2000, seed=4)
    runs["teleport_ideal_seed4"] = result_record(res, dt)
    programs["teleport_ideal_seed4"] = TELEPORT

    # noisy Bell, automatic -> density matrix (reference test_sim.py:255-266's model)
    noise_b = ref.NoiseModel.uniform(single=0.01, two=0.03, readout=0.01)
    res, dt = timed_run(BELL, shots=3000, seed=42, noise=noise_b)
    runs["bell_noisy_seed42"] = result_record(res, dt)
    programs["bell_noisy_seed42"] = BELL

    if not args.skip_slow:
        # the 20-qubit trajectory configuration, 256 shots (CPU time scales linearly in shots)
        c4 = wl.teleport_chain(20, 3)
        res, dt = timed_run(c4, shots=256, seed=7, noise=wl.noise_c4(), method="trajectories")
        runs["c4_20q_256shots_seed7"] = result_record(res, dt)
        programs["c4_20q_256shots_seed7"] = c4

    noises = {
        "c3": wl.noise_c3(),
        "c4": wl.noise_c4(),
        "uniform_bell": noise_b,
        "uniform_teleport": noise_t,
        "calibration": ref.NoiseModel.from_calibration(
            {
                "device": "example_device",
                "single_qubit_error": 0.001,
                "two_qubit_error": 0.01,
                "readout_error": 0.02,
                "qubits": [{"qubit": 1, "single_qubit_error": 0.3, "readout_error": None}],
                "couplers": [{"qubits": [0, 1], "error": 0.05}],
            }
        ),
    }
    noise_digests = {k: v.digest() for k, v in noises.items()}

    reader = {}
    for name, src in READER_PROGRAMS.items():
        params = {"theta": 1.1} if name == "modifiers" else None
        circ = ref.parse(src, params)
        reader[name] = {
            "source": src,
            "params": params,
            "digest": circ.digest(),
            "num_qubits": circ.num_qubits,
            "num_clbits": circ.num_clbits,
            "labels": [[getattr(op, "name", op.kind), list(getattr(op, "qubits", ()))] for op in circ.operations],
        }

    c2_src, c2_labels = wl.config_c2()
    methods = {
        "c1": choose_method(ref.parse(c1), None),
        "c2": choose_method(ref.parse(c2_src), None),
        "c3_automatic": choose_method(ref.parse(wl.config_c3()), wl.noise_c3()),
        "c4": choose_method(ref.parse(wl.teleport_chain(20, 3)), wl.noise_c4()),
        "clifford_12q": choose_method(ref.parse("OPENQASM 3; qubit[12] q; h q; cx q[0], q[1]; s q[3];"), None),
        "clifford_12q_noisy": choose_method(
            ref.parse("OPENQASM 3; qubit[12] q; h q; cx q[0], q[1];"), ref.NoiseModel.uniform(single=0.01)
        ),
        "clifford_12q_ampdamp": choose_method(
            ref.parse("OPENQASM 3; qubit[12] q; h q;"), ref.NoiseModel().add_gate_error(ref.amplitude_damping(0.1))
        ),
        "t_12q": choose_method(ref.parse("OPENQASM 3; qubit[12] q; h q; t q[0];"), None),
        "noisy_11q": choose_method(ref.parse("OPENQASM 3; qubit[11] q; h q;"), ref.NoiseModel.uniform(single=0.01)),
        "noisy_10q": choose_method(ref.parse("OPENQASM 3; qubit[10] q; h q;"), ref.NoiseModel.uniform(single=0.01)),
    }
    digests = {"c2": ref.parse(c2_src).digest(), "c3": ref.parse(wl.config_c3()).digest(), "c4": ref.parse(wl.teleport_chain(20, 3)).digest()}

    (out / "runs.json").write_text(json.dumps({"env": env, "runs": runs}, indent=1, sort_keys=True))
    (out / "programs.json").write_text(json.dumps(programs, indent=1, sort_keys=True))
    (out / "host.json").write_text(
        json.dumps(
            {"noise_digests": noise_digests, "reader": reader, "methods": methods, "config_digests": digests, "c2_labels": c2_labels},
            indent=1,
            sort_keys=True,
        )
    )
    print("wrote", ", ".join(sorted(runs)))


if __name__ == "__main__":
    main()
