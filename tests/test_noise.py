"""The noise module against lucivant-sim 0.3.0's own: channel descriptions, Kraus operators and
mixture branches (their order decides which draw picks which branch), the channel each model
attaches to every gate of a grid of names and qubits, readout errors, digests, and the messages
of invalid inputs (tests/golden/noise_cases.json, written by tests/golden/make_noise_cases.py)."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import lucivant_sim as ls

CASES = json.loads((Path(__file__).parent / "golden" / "noise_cases.json").read_text())


def _channels() -> dict[str, ls.Channel]:
    return {
        "depolarizing_1": ls.depolarizing(0.03),
        "depolarizing_2": ls.depolarizing(0.07, 2),
        "depolarizing_max_1": ls.depolarizing(4 / 3),
        "from_gate_error_1": ls.from_gate_error(0.01, 1),
        "from_gate_error_2": ls.from_gate_error(0.02, 2),
        "from_gate_error_cap": ls.from_gate_error(0.9, 1),
        "bit_flip": ls.bit_flip(0.1),
        "phase_flip": ls.phase_flip(0.2),
        "amplitude_damping": ls.amplitude_damping(0.3),
        "phase_damping": ls.phase_damping(0.4),
        "thermal": ls.thermal_relaxation(120, 90, 0.05),
        "thermal_t2_2t1": ls.thermal_relaxation(10, 20, 3),
        "thermal_zero_time": ls.thermal_relaxation(10, 5, 0),
        "pauli_partial": ls.pauli_channel({"ZX": 0.1, "XI": 0.05, "YY": 0.0, "IZ": 0.2}),
        "pauli_full": ls.pauli_channel({"X": 0.5, "Z": 0.5}),
        "custom": ls.Channel("custom", 1, ls.amplitude_damping(0.2).kraus, (0.2,)),
    }


def _models() -> dict[str, ls.NoiseModel]:
    return {
        "uniform": ls.NoiseModel.uniform(single=1e-3, two=1e-2, readout=2e-2),
        "uniform_zero": ls.NoiseModel.uniform(),
        "specific": ls.NoiseModel()
        .add_gate_error(ls.depolarizing(0.01), gates=["x", "u3"])
        .add_gate_error(ls.depolarizing(0.02), qubits=[1])
        .add_gate_error(ls.depolarizing(0.03), gates="sx", qubits=[1])
        .add_gate_error(ls.depolarizing(0.04))
        .add_gate_error(ls.depolarizing(0.05, 2), qubits=[1, 0])
        .add_gate_error(ls.depolarizing(0.06, 2), gates="CX")
        .add_gate_error(ls.bit_flip(0.07), gates=["rz"])
        .add_gate_error(ls.depolarizing(0.08, 3))
        .add_gate_error(ls.depolarizing(0.09, 2), gates="ctrl@x", qubits=[2, 1])
        .add_readout_error(0.01)
        .add_readout_error(ls.ReadoutError(0.02, 0.05), qubits=[1, 2])
        .add_readout_error(0.0, qubits=[0]),
        "ties": ls.NoiseModel().add_gate_error(ls.bit_flip(0.1)).add_gate_error(ls.phase_flip(0.2)).add_gate_error(ls.bit_flip(0.3), gates="h").add_gate_error(ls.phase_flip(0.4), gates=["h"]),
        "readout_zero": ls.NoiseModel().add_readout_error(0.0),
        "calibration": ls.NoiseModel.from_calibration(
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
        "calibration_empty": ls.NoiseModel.from_calibration({}),
    }


def _mat(rec: list) -> np.ndarray:
    return np.array([[complex(re, im) for re, im in row] for row in rec])


@pytest.mark.parametrize("name", sorted(CASES["channels"]))
def test_channels_equal_the_reference(name: str) -> None:
    rec = CASES["channels"][name]
    ch = _channels()[name]
    assert json.dumps(ch.describe()) == json.dumps(rec["describe"])
    assert ch.num_qubits == rec["num_qubits"] and ch.analytic_depolarizing == rec["analytic_depolarizing"]
    assert len(ch.kraus) == len(rec["kraus"])
    for k, want in zip(ch.kraus, rec["kraus"]):
        assert np.array_equal(k, _mat(want))
    if rec["mixture"] is None:
        assert ch.mixture is None
    else:
        assert ch.mixture is not None and len(ch.mixture) == len(rec["mixture"])
        for (p, u), (wp, wu) in zip(ch.mixture, rec["mixture"]):
            assert p == wp and np.array_equal(u, _mat(wu))


@pytest.mark.parametrize("name", sorted(CASES["models"]))
def test_models_attach_what_the_reference_attaches(name: str) -> None:
    rec = CASES["models"][name]
    m = _models()[name]
    assert m.digest() == rec["digest"]
    assert json.dumps(m.describe(), sort_keys=True) == json.dumps(rec["describe"], sort_keys=True)
    assert (m.has_gate_errors(), m.is_empty(), m.source) == (rec["has_gate_errors"], rec["is_empty"], rec["source"])
    for q, want in rec["readout"]:
        e = m.readout_error(q)
        assert (None if e is None else [e.p01, e.p10]) == want
    for _ in range(2):  # the second pass answers from the lookup table
        for g, q, want in rec["lookups"]:
            ch = m.gate_error(g, tuple(q))
            assert (None if ch is None else json.loads(json.dumps(ch.describe()))) == want, (g, q)


def test_lookups_follow_rules_added_later() -> None:
    m = ls.NoiseModel.uniform(single=1e-3)
    assert m.gate_error("x", (0,)).params[0] == pytest.approx(2e-3)  # type: ignore[union-attr]
    m.add_gate_error(ls.bit_flip(0.25), gates="x")
    assert m.gate_error("x", (0,)).name == "bit_flip"  # type: ignore[union-attr]


_BAD = {
    "bit_flip_high": lambda: ls.bit_flip(1.5),
    "bit_flip_nan": lambda: ls.bit_flip(float("nan")),
    "phase_flip_negative": lambda: ls.phase_flip(-0.1),
    "depolarizing_high": lambda: ls.depolarizing(1.5),
    "depolarizing_2_high": lambda: ls.depolarizing(1.1, 2),
    "gamma": lambda: ls.amplitude_damping(2.0),
    "lambda": lambda: ls.phase_damping(-1.0),
    "t2": lambda: ls.thermal_relaxation(1.0, 3.0, 0.1),
    "t1_zero": lambda: ls.thermal_relaxation(0.0, 1.0, 0.1),
    "time_negative": lambda: ls.thermal_relaxation(1.0, 1.0, -0.1),
    "pauli_total": lambda: ls.pauli_channel({"X": 0.7, "Z": 0.6}),
    "pauli_label": lambda: ls.pauli_channel({"X": 0.1, "XY": 0.1}),
    "pauli_letter": lambda: ls.pauli_channel({"A": 0.1}),
    "not_cptp": lambda: ls.Channel("bad", 1, (ls.amplitude_damping(0.2).kraus[0],)),
    "readout_high": lambda: ls.ReadoutError(1.2, 0.0),
    "readout_p10": lambda: ls.ReadoutError(0.0, -0.2),
    "attach_width": lambda: ls.NoiseModel().add_gate_error(ls.depolarizing(0.1, 2), qubits=[0]),
}


@pytest.mark.parametrize("name", sorted(CASES["errors"]))
def test_invalid_inputs_give_the_reference_message(name: str) -> None:
    with pytest.raises(ValueError) as err:
        _BAD[name]()
    assert str(err.value) == CASES["errors"][name]
