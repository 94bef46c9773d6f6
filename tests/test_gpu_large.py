"""Full-size configurations on the MI355X, checked through size-independent properties (the
oracle would take tens of minutes at these sizes): unitarity round trips, norms, traces,
probability sums and expectations of diagonal observables against the probabilities."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest
import torch

import lucivant_sim as ls
from lucivant_sim import _native as nat
from lucivant_sim import workloads as wl
from lucivant_sim.engines import DensityMatrix, Statevector
from lucivant_sim.plan import plan_density, plan_statevector

pytestmark = pytest.mark.gpu


def _inverse(circuit: ls.Circuit) -> ls.Circuit:
    from lucivant_sim.circuit import Gate

    ops = tuple(Gate(g.name + "_dg", g.qubits, np.asarray(g.matrix).conj().T) for g in reversed(circuit.operations) if isinstance(g, Gate))
    return ls.Circuit(circuit.num_qubits, 0, ops)


def test_26_qubit_circuit_round_trip_and_observables() -> None:
    src, labels = wl.config_c2()
    c = ls.parse(src)
    sv = Statevector(26)
    sv.run(plan_statevector(c.operations, 26, 4))
    probs = sv.probabilities()
    assert abs(probs.sum() - 1) < 1e-10
    # <Z...Z> on a diagonal string equals the signed sum of probabilities
    zlabel = "I" * 13 + "Z" * 13  # Z on qubits 0..12 (the rightmost letter is qubit 0)
    idx = np.arange(2**26, dtype=np.int64)
    sign = 1 - 2 * (np.bitwise_count((idx & ((1 << 13) - 1)).astype(np.uint64)) & 1).astype(np.float64)
    e = sv.expectations([zlabel] + labels)
    assert abs(e[0] - float(np.dot(sign, probs))) < 1e-10
    assert np.all(np.abs(e[1:]) <= 1 + 1e-12)
    sv.run(plan_statevector(_inverse(c).operations, 26, 5))
    amp0 = complex(sv.buffer[0].item())
    rest = float(torch.linalg.vector_norm(sv.buffer[1:]).item())
    assert abs(amp0 - 1) < 1e-10 and rest < 1e-10


@pytest.mark.parametrize("width", [4, 5, "tile"])
def test_26_qubit_circuit_matches_the_oracle_fixture(width) -> None:  # type: ignore[no-untyped-def]
    """The 26-qubit configuration against the oracle's own full run (tests/golden/
    make_full_size.py oracle-c2): 4096 amplitudes at seeded indices, probability checksums and the
    8 Pauli expectations, within 1e-10, for three plans (fused widths 4 and 5, tile passes)."""
    from lucivant_sim.plan import plan_tiles

    rec = json.loads((Path(__file__).parent / "golden" / "full_c2_oracle.json").read_text())
    src, labels = wl.config_c2()
    assert labels == rec["labels"]
    c = ls.parse(src)
    sv = Statevector(26)
    sv.run(plan_tiles(c.operations, 26) if width == "tile" else plan_statevector(c.operations, 26, width))
    idx = torch.tensor(rec["indices"], dtype=torch.int64, device=sv.buffer.device)
    got = sv.buffer[idx].cpu().numpy()
    want = np.array([complex(re, im) for re, im in rec["amplitudes"]])
    assert np.abs(got - want).max() < 1e-10
    probs = sv.probabilities()
    assert abs(probs.sum() - rec["prob_sum"]) < 1e-10
    index_mean = float(np.dot(probs, np.arange(2**26, dtype=np.float64)) / 2**26)
    assert abs(index_mean - rec["prob_index_mean"]) < 1e-10
    assert int(probs.argmax()) == rec["prob_argmax"] and abs(float(probs.max()) - rec["prob_max"]) < 1e-12
    e = sv.expectations(labels)
    assert np.abs(e - np.array(rec["expectations"])).max() < 1e-10
    if width == 5:
        # the public API re-simulates per call (simulator.py:517-519)
        assert abs(ls.expectation(src, labels[0]) - rec["expectations"][0]) < 1e-10


def test_13_qubit_density_matrix_counts_equal_the_reference() -> None:
    """The 13-qubit density-matrix configuration (noise on every gate, readout errors), 4000 shots
    at seed 7, against lucivant-sim 0.3.0's own run of it (tests/golden/make_full_size.py
    reference-c3): identical counts, digests and fingerprint."""
    rec = json.loads((Path(__file__).parent / "golden" / "full_c3_reference.json").read_text())
    res = ls.run(wl.config_c3(), 4000, seed=7, noise=wl.noise_c3(), method="density_matrix")
    assert res.circuit_digest == rec["circuit_digest"]
    assert res.noise_digest == rec["noise_digest"]
    assert res.fingerprint == rec["fingerprint"]
    assert res.counts == rec["counts"]
    assert res.counts_digest == rec["counts_digest"]


def test_13_qubit_density_matrix_trace_and_round_trip() -> None:
    src = wl.config_c3(13, 6)
    c = ls.parse(src)
    dm = DensityMatrix(13)
    items = plan_density(c.operations, 13, wl.noise_c3())
    dm.run([p for p in items if not isinstance(p, tuple)])
    p = dm.probabilities()
    assert abs(p.sum() - 1) < 1e-10 and p.min() >= 0
    ideal = DensityMatrix(13)
    ideal.run(plan_density(c.operations, 13, None))
    ideal.run(plan_density(_inverse(c).operations, 13, None))
    d = ideal.probabilities()
    assert abs(d[0] - 1) < 1e-10


def test_20_qubit_trajectory_batch_is_normalised_and_reproducible() -> None:
    from lucivant_sim.trajectories import run_batched

    c = ls.parse(wl.teleport_chain(20, 3))
    a = run_batched(c, wl.noise_c4(), 64, np.random.default_rng(5))
    b = run_batched(c, wl.noise_c4(), 64, np.random.default_rng(5), batch_bytes=16 << 20 << 4)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
def test_shared_trajectory_states_give_the_bits_of_one_state_per_shot(dtype: int) -> None:
    """Shots with equal histories share one device state, copied where their choices part; every
    shot's bits equal those of the batch that keeps one state per shot from the start (noisy
    teleport chain with mid-circuit measurements, and random layers with strong noise)."""
    from lucivant_sim.trajectories import run_batched

    cases = [
        (ls.parse(wl.teleport_chain(16, 3)), wl.noise_c4(), 512),
        (ls.parse(wl.random_layers(12, 8, 77, measure=True)), ls.NoiseModel.uniform(single=0.02, two=0.05, readout=0.02), 300),
    ]
    for c, noise, shots in cases:
        shared = run_batched(c, noise, shots, np.random.default_rng(17), dtype=dtype)
        alone = run_batched(c, noise, shots, np.random.default_rng(17), dtype=dtype, share=False)
        assert np.array_equal(shared, alone)


def test_trajectory_rescale_of_small_norms_keeps_the_bits(monkeypatch) -> None:  # type: ignore[no-untyped-def]
    """Raising the rescale threshold so that the device batch rescales shared states at nearly
    every decision leaves every shot's bits as they were (dynamic circuit with noise)."""
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).parent))
    from dynamic_circuits import DYNAMIC, dynamic_noise
    from lucivant_sim import trajectories as tr

    c = ls.parse(DYNAMIC)
    want = tr.run_batched(c, dynamic_noise(), 400, np.random.default_rng(19))
    monkeypatch.setattr(tr, "_RENORMALISE_BELOW", {nat.C128: 0.999, nat.C64: 0.999})
    got = tr.run_batched(c, dynamic_noise(), 400, np.random.default_rng(19))
    assert np.array_equal(got, want)
