"""Write tests/golden/noise_cases.json: what lucivant-sim 0.3.0's noise module gives for a
corpus of channels and noise models.

    python tests/golden/make_noise_cases.py [--reference /root/reference/src]

The reference is imported from its checkout in this container only, to record data: each
channel's description, Kraus operators and mixture branches (probabilities and unitaries, in the
order trajectories draw them); each model's digest, the channel its rules attach to every gate
of a grid of names and qubits, its readout errors; and the message of each invalid input. This
build's noise module is checked against these records (tests/test_noise.py).
"""

from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent

GATE_GRID = [
    (name, qubits)
    for name in ("x", "sx", "h", "u", "rz", "p", "id", "u0", "ry", "cx", "cz", "swap", "ccx", "ctrl@x", "g")
    for qubits in ((0,), (1,), (2,), (0, 1), (1, 0), (1, 2), (0, 1, 2))
]


def _mat(m) -> list:  # type: ignore[no-untyped-def]
    return [[[float(v.real), float(v.imag)] for v in row] for row in m]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference/src")
    args = ap.parse_args()
    sys.path.insert(0, args.reference)
    import lucivant_sim as ref  # the reference implementation

    channels = {
        "depolarizing_1": ref.depolarizing(0.03),
        "depolarizing_2": ref.depolarizing(0.07, 2),
        "depolarizing_max_1": ref.depolarizing(4 / 3),
        "from_gate_error_1": ref.from_gate_error(0.01, 1),
        "from_gate_error_2": ref.from_gate_error(0.02, 2),
        "from_gate_error_cap": ref.from_gate_error(0.9, 1),
        "bit_flip": ref.bit_flip(0.1),
        "phase_flip": ref.phase_flip(0.2),
        "amplitude_damping": ref.amplitude_damping(0.3),
        "phase_damping": ref.phase_damping(0.4),
        "thermal": ref.thermal_relaxation(120, 90, 0.05),
        "thermal_t2_2t1": ref.thermal_relaxation(10, 20, 3),
        "thermal_zero_time": ref.thermal_relaxation(10, 5, 0),
        "pauli_partial": ref.pauli_channel({"ZX": 0.1, "XI": 0.05, "YY": 0.0, "IZ": 0.2}),
        "pauli_full": ref.pauli_channel({"X": 0.5, "Z": 0.5}),
        "custom": ref.Channel("custom", 1, ref.amplitude_damping(0.2).kraus, (0.2,)),
    }
    chan_out = {}
    for name, ch in channels.items():
        chan_out[name] = {
            "describe": ch.describe(),
            "num_qubits": ch.num_qubits,
            "analytic_depolarizing": bool(ch.analytic_depolarizing),
            "kraus": [_mat(k) for k in ch.kraus],
            "mixture": None if ch.mixture is None else [[float(p), _mat(u)] for p, u in ch.mixture],
        }

    models = {
        "uniform": ref.NoiseModel.uniform(single=1e-3, two=1e-2, readout=2e-2),
        "uniform_zero": ref.NoiseModel.uniform(),
        "specific": ref.NoiseModel()
        .add_gate_error(ref.depolarizing(0.01), gates=["x", "u3"])
        .add_gate_error(ref.depolarizing(0.02), qubits=[1])
        .add_gate_error(ref.depolarizing(0.03), gates="sx", qubits=[1])
        .add_gate_error(ref.depolarizing(0.04))
        .add_gate_error(ref.depolarizing(0.05, 2), qubits=[1, 0])
        .add_gate_error(ref.depolarizing(0.06, 2), gates="CX")
        .add_gate_error(ref.bit_flip(0.07), gates=["rz"])
        .add_gate_error(ref.depolarizing(0.08, 3))
        .add_gate_error(ref.depolarizing(0.09, 2), gates="ctrl@x", qubits=[2, 1])
        .add_readout_error(0.01)
        .add_readout_error(ref.ReadoutError(0.02, 0.05), qubits=[1, 2])
        .add_readout_error(0.0, qubits=[0]),
        "ties": ref.NoiseModel().add_gate_error(ref.bit_flip(0.1)).add_gate_error(ref.phase_flip(0.2)).add_gate_error(ref.bit_flip(0.3), gates="h").add_gate_error(ref.phase_flip(0.4), gates=["h"]),
        "readout_zero": ref.NoiseModel().add_readout_error(0.0),
        "calibration": ref.NoiseModel.from_calibration(
            {
                "calibration": {
                    "profile_key": "twin_profile",
                    "single_qubit_error": 0.002,
                    "two_qubit_error": None,
                    "readout_error": 0.03,
                    "qubits": [{"qubit": 0, "single_qubit_error": 0.01, "readout_error": 0.04}, {"qubit": 2, "single_qubit_error": None}],
                    "couplers": [{"qubits": [2, 1], "error": 0.06}, {"qubits": [0, 1], "error": None}],
                }
            }
        ),
        "calibration_empty": ref.NoiseModel.from_calibration({}),
    }
    model_out = {}
    for name, m in models.items():
        lookups = []
        for g, q in GATE_GRID:
            ch = m.gate_error(g, q)
            lookups.append([g, list(q), None if ch is None else ch.describe()])
        model_out[name] = {
            "digest": m.digest(),
            "describe": m.describe(),
            "has_gate_errors": m.has_gate_errors(),
            "is_empty": m.is_empty(),
            "source": m.source,
            "readout": [[q, None if (e := m.readout_error(q)) is None else [e.p01, e.p10]] for q in range(4)],
            "lookups": lookups,
        }

    bad = {
        "bit_flip_high": lambda: ref.bit_flip(1.5),
        "bit_flip_nan": lambda: ref.bit_flip(float("nan")),
        "phase_flip_negative": lambda: ref.phase_flip(-0.1),
        "depolarizing_high": lambda: ref.depolarizing(1.5),
        "depolarizing_2_high": lambda: ref.depolarizing(1.1, 2),
        "gamma": lambda: ref.amplitude_damping(2.0),
        "lambda": lambda: ref.phase_damping(-1.0),
        "t2": lambda: ref.thermal_relaxation(1.0, 3.0, 0.1),
        "t1_zero": lambda: ref.thermal_relaxation(0.0, 1.0, 0.1),
        "time_negative": lambda: ref.thermal_relaxation(1.0, 1.0, -0.1),
        "pauli_total": lambda: ref.pauli_channel({"X": 0.7, "Z": 0.6}),
        "pauli_label": lambda: ref.pauli_channel({"X": 0.1, "XY": 0.1}),
        "pauli_letter": lambda: ref.pauli_channel({"A": 0.1}),
        "not_cptp": lambda: ref.Channel("bad", 1, (ref.amplitude_damping(0.2).kraus[0],)),
        "readout_high": lambda: ref.ReadoutError(1.2, 0.0),
        "readout_p10": lambda: ref.ReadoutError(0.0, -0.2),
        "attach_width": lambda: ref.NoiseModel().add_gate_error(ref.depolarizing(0.1, 2), qubits=[0]),
    }
    errors = {}
    for name, fn in bad.items():
        try:
            fn()
        except ValueError as exc:
            errors[name] = str(exc)
        else:
            raise SystemExit(f"{name} raised nothing")
    out = {"reference_version": ref.__version__, "channels": chan_out, "models": model_out, "errors": errors}
    (HERE / "noise_cases.json").write_text(json.dumps(out, indent=0, sort_keys=True))
    print(f"wrote {len(chan_out)} channels, {len(model_out)} models, {len(errors)} errors")


if __name__ == "__main__":
    main()
