"""The public API on the MI355X against the oracle and the reference's own counts.

Amplitudes, probabilities and expectations: within 1e-10 of the oracle (complex128). Counts:
identical to the reference's golden counts for the same program, noise and seed (the draw
stream and sampling rule are the reference's; a mismatch could only come from a draw landing
within rounding of a CDF or branch boundary, which none of these fixtures has). Statistical
behaviour follows the reference tests (test_sim.py) with their tolerances.
"""

from __future__ import annotations

import json
import math

import numpy as np
import pytest

import lucivant_sim as ls
from lucivant_sim import _native as nat
from lucivant_sim import workloads as wl
from lucivant_sim.trajectories import run_batched, run_sequential
from dynamic_circuits import DYNAMIC, dynamic_noise
from oracle import engine as orc

pytestmark = pytest.mark.gpu

BELL = """
OPENQASM 3;
include "stdgates.inc";
qubit[2] q;
bit[2] c;
h q[0];
cx q[0], q[1];
c = measure q;
"""

NOISES = {
    "bell_seed7": None,
    "c1_12q_seed7": None,
    "c3_recipe_6q_seed7": "c3",
    "c4_recipe_9q_seed7": "c4",
    "teleport_noisy_seed3": "teleport",
    "teleport_ideal_seed4": None,
    "bell_noisy_seed42": "bell",
    "c4_20q_256shots_seed7": "c4",
}


def noise_for(tag: str | None) -> ls.NoiseModel | None:
    if tag is None:
        return None
    if tag == "c3":
        return wl.noise_c3()
    if tag == "c4":
        return wl.noise_c4()
    if tag == "teleport":
        return ls.NoiseModel.uniform(single=0.01, two=0.02, readout=0.01)
    return ls.NoiseModel.uniform(single=0.01, two=0.03, readout=0.01)


def tvd(counts: dict[str, int], probs: dict[str, float]) -> float:
    return orc.tvd(counts, probs)


@pytest.mark.parametrize("name", sorted(NOISES))
def test_counts_and_fingerprints_equal_the_reference(name: str, golden_runs: dict, golden_programs: dict) -> None:
    if name not in golden_runs:
        pytest.skip("fixture not generated")
    rec = golden_runs[name]
    res = ls.run(golden_programs[name], rec["shots"], seed=rec["seed"], noise=noise_for(NOISES[name]), method=rec["method"])
    assert res.method == rec["method"]
    assert res.counts == rec["counts"]
    assert res.fingerprint == rec["fingerprint"]
    assert res.counts_digest == rec["counts_digest"]
    assert ls.verify(rec, golden_programs[name], noise_for(NOISES[name]))


def test_c1_statevector_within_1e10_of_the_oracle() -> None:
    src = wl.config_c1()
    want = orc.evolve_sv(ls.parse(src))
    from lucivant_sim import simulator

    for width in (1, 2, 3, 4, 5):
        old = simulator.FUSION_WIDTH
        simulator.FUSION_WIDTH = width
        try:
            got = ls.statevector(src)
        finally:
            simulator.FUSION_WIDTH = old
        assert np.abs(got - want).max() < 1e-10, width
    res = ls.run(src, shots=10, seed=1)
    assert res.statevector is not None and np.abs(res.statevector - want).max() < 1e-10


def test_density_matrix_within_1e10_of_the_oracle() -> None:
    src = wl.config_c3(6, 8)
    noise = wl.noise_c3()
    want = orc.evolve_dm(ls.parse(src), noise)
    got = ls.density_matrix(src, noise)
    assert np.abs(got - want).max() < 1e-10
    probs = ls.probabilities(src, noise)
    p = orc.probabilities_dm(want)
    assert set(probs) == {format(i, "06b") for i in range(64) if p[i] > 1e-12}
    assert max(abs(probs[k] - p[int(k, 2)]) for k in probs) < 1e-10


def test_wide_gates_and_channels_on_the_density_matrix() -> None:
    # a reset is static only on a qubit nothing has touched yet (circuit.py:121-138)
    src = 'OPENQASM 3; include "stdgates.inc"; qubit[5] q; reset q[4]; h q[0:3]; ccx q[0], q[1], q[2]; cswap q[3], q[0], q[2]; rx(0.3) q[1]; cx q[1], q[4];'
    noise = ls.NoiseModel().add_gate_error(ls.depolarizing(0.1, 3), gates="ccx").add_gate_error(ls.amplitude_damping(0.2), gates="rx")
    want = orc.evolve_dm(ls.parse(src), noise)
    assert np.abs(ls.density_matrix(src, noise) - want).max() < 1e-10


def test_expectations_within_1e10() -> None:
    src = wl.random_layers(10, 6, 21, two_qubit=("cx", "cz"), measure=False)
    psi = orc.evolve_sv(ls.parse(src))
    labels = wl.pauli_strings(10, 16, 3)
    rng = np.random.default_rng(4)
    obs = {lbl: float(w) for lbl, w in zip(labels, rng.uniform(-1, 1, size=16))}
    want = sum(w * orc.expectation_sv(psi, lbl, 10) for lbl, w in obs.items())
    assert abs(ls.expectation(src, obs) - want) < 1e-10
    noise = wl.noise_c3()
    rho = orc.evolve_dm(ls.parse(src), noise)
    want_dm = sum(w * orc.expectation_dm(rho, lbl, 10) for lbl, w in obs.items())
    assert abs(ls.expectation(src, obs, noise) - want_dm) < 1e-10


def test_expectation_with_noise_is_exact() -> None:
    source = 'OPENQASM 3; include "stdgates.inc"; qubit[2] q; h q[0]; cx q[0], q[1];'
    noise = ls.NoiseModel().add_gate_error(ls.depolarizing(0.2, 2), gates="cx")
    assert ls.expectation(source, "XX", noise) == pytest.approx(0.8)
    assert ls.expectation(source, {"ZZ": 1.0, "ZI": 3.0}) == pytest.approx(1.0)
    with pytest.raises(ValueError, match="Pauli string of length 2"):
        ls.expectation(source, "ZZZ")


def test_thermal_relaxation_decays_one_to_zero() -> None:
    noise = ls.NoiseModel().add_gate_error(ls.thermal_relaxation(t1=10.0, t2=8.0, time=3.0), gates="x")
    probs = ls.probabilities("OPENQASM 3; qubit q; x q;", noise)
    assert probs["1"] == pytest.approx(math.exp(-0.3))


def test_bell_counts_are_balanced_and_correlated() -> None:
    result = ls.run(BELL, shots=20_000, seed=11)
    assert set(result.counts) == {"00", "11"}
    assert tvd(result.counts, {"00": 0.5, "11": 0.5}) < 0.01
    assert result.method == "statevector" and result.statevector is not None


def test_noisy_counts_match_exact_noisy_probabilities() -> None:
    noise = ls.NoiseModel.uniform(single=0.02, two=0.06)
    exact = ls.probabilities(BELL, noise)
    sampled = ls.run(BELL, shots=40_000, seed=5, noise=noise)
    assert sampled.method == "density_matrix"
    assert tvd(sampled.counts, exact) < 0.01


def test_trajectories_agree_with_the_density_matrix() -> None:
    noise = ls.NoiseModel().add_gate_error(ls.amplitude_damping(0.15), gates="h").add_gate_error(ls.depolarizing(0.1, 2), gates="cx")
    exact = ls.probabilities(BELL, noise)
    traj = ls.run(BELL, shots=20_000, seed=9, noise=noise, method="trajectories")
    assert tvd(traj.counts, exact) < 0.015


def test_readout_error_flips_the_record() -> None:
    noise = ls.NoiseModel().add_readout_error(ls.ReadoutError(p01=0.1, p10=0.3))
    result = ls.run("OPENQASM 3; qubit[2] q; bit[2] c; x q[1]; c = measure q;", shots=40_000, seed=2, noise=noise)
    expected = {"10": 0.9 * 0.7, "00": 0.9 * 0.3, "11": 0.1 * 0.7, "01": 0.1 * 0.3}
    assert result.method == "statevector"
    assert tvd(result.counts, expected) < 0.01


def test_teleportation_with_mid_circuit_measurement_and_conditions() -> None:
    theta = 1.234
    source = f"""
    OPENQASM 3;
    include "stdgates.inc";
    qubit[3] q;
    bit[2] m;
    bit[1] out;
    ry({theta}) q[0];
    h q[1]; cx q[1], q[2];
    cx q[0], q[1]; h q[0];
    m[0] = measure q[0];
    m[1] = measure q[1];
    if (m[1]) x q[2];
    if (m[0] == 1) {{ z q[2]; }}
    out[0] = measure q[2];
    """
    result = ls.run(source, shots=20_000, seed=4)
    assert result.method == "trajectories"
    ones = sum(n for k, n in result.counts.items() if k.split(" ")[0] == "1")
    assert ones / 20_000 == pytest.approx(math.sin(theta / 2) ** 2, abs=0.012)


def test_reset_and_if_else_exact_counts() -> None:
    assert ls.run("OPENQASM 3; qubit q; bit c; x q; reset q; c = measure q;", shots=100, seed=1).counts == {"0": 100}
    source = """
    OPENQASM 3; include "stdgates.inc";
    qubit[2] q; bit[2] c;
    x q[0];
    c[0] = measure q[0];
    if (c[0] == 1) { c[0] = measure q[1]; } else { x q[1]; }
    c[1] = measure q[1];
    """
    assert ls.run(source, shots=50, seed=1).counts == {"00": 50}


def test_outcome_probabilities_include_readout_error() -> None:
    noise = ls.NoiseModel.uniform(two=0.05, readout=0.1)
    exact = ls.outcome_probabilities(BELL, noise)
    assert sum(exact.values()) == pytest.approx(1)
    assert tvd(ls.run(BELL, shots=40_000, seed=8, noise=noise).counts, exact) < 0.01
    several = "OPENQASM 2.0; qreg q[2]; creg a[1]; creg b[1]; x q[1]; measure q[0] -> a[0]; measure q[1] -> b[0];"
    assert ls.outcome_probabilities(several) == {"1 0": 1.0}


def test_reproducibility_and_verification(tmp_path) -> None:  # type: ignore[no-untyped-def]
    noise = ls.NoiseModel.uniform(single=0.01, two=0.03, readout=0.01)
    a = ls.run(BELL, shots=3000, seed=42, noise=noise)
    b = ls.run(BELL, shots=3000, seed=42, noise=noise)
    assert a.counts == b.counts and a.fingerprint == b.fingerprint
    assert ls.run(BELL, shots=3000, seed=43, noise=noise).counts != a.counts
    saved = tmp_path / "result.json"
    saved.write_text(json.dumps(a.to_dict()))
    assert ls.verify(saved, BELL, noise)
    assert not ls.verify(saved, BELL, None)
    assert not ls.verify(a.to_dict() | {"counts": {"00": 1}}, BELL, noise)


def test_an_unseeded_run_records_its_seed() -> None:
    """test_sim.py:269-271: without a seed, run draws one and records it; rerunning with the
    recorded seed gives the same counts and fingerprint."""
    noise = ls.NoiseModel.uniform(single=0.01, two=0.03)
    a = ls.run(BELL, shots=500, noise=noise)
    assert isinstance(a.seed, int) and 0 <= a.seed < 2**63
    b = ls.run(BELL, shots=500, seed=a.seed, noise=noise)
    assert a.counts == b.counts and a.fingerprint == b.fingerprint


def test_limits_and_physical_qubits() -> None:
    with pytest.raises(ls.SimulationError, match="up to 26"):
        ls.run("OPENQASM 3; qubit[30] q; h q; t q;", shots=1)
    sv = ls.statevector("OPENQASM 3; x $2;")
    assert sv.shape == (8,) and abs(sv[4]) == pytest.approx(1)
    assert np.allclose(ls.statevector("OPENQASM 3;\npragma something\nqubit q;\ngphase(pi);\nx q;"), [0, -1])


def test_batched_sequential_and_oracle_trajectories_agree_exactly() -> None:
    src = wl.teleport_chain(9, 3)
    circuit = ls.parse(src)
    noise = wl.noise_c4()
    shots = 300
    batched = run_batched(circuit, noise, shots, np.random.default_rng(17))
    sequential = run_sequential(circuit, noise, shots, np.random.default_rng(17))
    assert [int(v) for v in batched] == sequential
    want = orc.trajectory_counts(circuit, noise, shots, np.random.default_rng(17))
    uniq, cnt = np.unique(batched, return_counts=True)
    assert {circuit.format_bits(int(v)): int(c) for v, c in zip(uniq, cnt)} == want


def test_batched_trajectories_with_fused_queues_equal_the_sequential_order() -> None:
    """On the shared dynamic circuit, the batch -- which fuses chosen operators into per-shot
    blocks and carries reduced density matrices on the host -- gives every shot the bits the
    reference's sequential order gives it."""
    circuit = ls.parse(DYNAMIC)
    noise = dynamic_noise()
    shots = 120
    batched = run_batched(circuit, noise, shots, np.random.default_rng(23))
    sequential = run_sequential(circuit, noise, shots, np.random.default_rng(23))
    assert [int(v) for v in batched] == sequential
    assert len(set(sequential)) > 4


def test_complex64_trajectories_stay_close_to_complex128() -> None:
    src = wl.teleport_chain(9, 3)
    circuit = ls.parse(src)
    noise = wl.noise_c4()
    a = run_batched(circuit, noise, 2000, np.random.default_rng(3), dtype=nat.C128)
    b = run_batched(circuit, noise, 2000, np.random.default_rng(3), dtype=nat.C64)
    # same draws: shots can only differ where a draw lies within float32 rounding of a boundary
    assert np.mean(a != b) < 0.01


def test_complex64_trajectories_state_their_error() -> None:
    """precision='complex64' runs trajectories in complex64 and the Result states how its shots
    compare with complex128 (the first shots rerun with the same draws). The draws are identical,
    so outcomes can differ only where a probability sits within float32 rounding of a draw."""
    src, noise = wl.teleport_chain(9, 3), wl.noise_c4()
    r64 = ls.run(src, 2000, seed=3, noise=noise, method="trajectories", precision="complex64")
    r128 = ls.run(src, 2000, seed=3, noise=noise, method="trajectories")
    assert r64.precision is not None and r64.precision["dtype"] == "complex64"
    assert r64.precision["checked_shots"] == 2000
    assert r64.precision["shots_differing_from_complex128"] <= 2
    # the LDS program kernel records every decision: measurement p1 and Kraus running weights
    assert r64.precision["decisions_compared"] > 2000 * 6
    assert 0 < r64.precision["max_abs_dp"] < 1e-5
    assert r64.precision["counts_tvd"] <= r64.precision["shots_differing_from_complex128"] / 2000
    assert r128.precision is None
    tvd = 0.5 * sum(abs(r64.counts.get(k, 0) - r128.counts.get(k, 0)) for k in set(r64.counts) | set(r128.counts)) / 2000
    assert tvd <= 2 / 2000
    assert r64.fingerprint == r128.fingerprint  # precision is not part of what was run
    with pytest.raises(ValueError):
        ls.run(wl.random_layers(5, 2, 1), 10, seed=3, method="statevector", precision="complex64")


def test_c4_20_qubit_complex64_against_the_reference_counts(golden_runs: dict, golden_programs: dict) -> None:
    """The 20-qubit trajectory configuration (teleport chain, 6 mid-circuit measurements with
    if/else corrections, depolarising noise and amplitude damping), 256 shots at seed 7, in both
    precisions against lucivant-sim 0.3.0's own counts for it. complex128 equals them exactly.
    complex64 (float32 states, float64 decisions) may differ only in shots where a decision took
    another branch because its draw lay within float32 error of the branch boundary: each such
    shot is listed with the distance of its draw from the complex128 boundary, which must be
    below 1e-5; decision probabilities that did not flip agree within 1e-5."""
    from lucivant_sim.trajectories import compare_decisions

    name = "c4_20q_256shots_seed7"
    rec = golden_runs[name]
    circuit = ls.parse(golden_programs[name])
    noise = wl.noise_c4()
    r128: dict = {}
    r64: dict = {}
    b128 = run_batched(circuit, noise, 256, np.random.default_rng(7), record=r128)
    b64 = run_batched(circuit, noise, 256, np.random.default_rng(7), dtype=nat.C64, record=r64)

    def counts(bits: np.ndarray) -> dict[str, int]:
        uniq, cnt = np.unique(bits, return_counts=True)
        return {circuit.format_bits(int(v)): int(c) for v, c in zip(uniq, cnt)}

    assert counts(b128) == rec["counts"]
    assert np.array_equal(r64["draws"], r128["draws"])
    stats = compare_decisions(r64["decisions"], r128["decisions"], r128["draws"], margins=256)
    differing = np.flatnonzero(b64 != b128)
    for s in differing:
        print(f"shot {s}: complex64 {int(b64[s]):b} vs complex128 {int(b128[s]):b}")
    print(f"flipped decisions: {stats['flipped_decisions']}, margins {stats['flip_margins']}, max |dp| {stats['max_abs_dp']:.3e}")
    assert len(differing) <= stats["flipped_decisions"]
    assert all(m < 1e-5 for m in stats["flip_margins"])
    assert stats["decisions_compared"] >= 256 * 6
    assert stats["max_abs_dp"] < 1e-5
    got = counts(b64)
    tvd = 0.5 * sum(abs(got.get(k, 0) - rec["counts"].get(k, 0)) for k in set(got) | set(rec["counts"])) / 256
    assert tvd <= len(differing) / 256
    res = ls.run(golden_programs[name], 256, seed=7, noise=noise, method="trajectories", precision="complex64")
    assert res.precision is not None and res.precision["checked_shots"] == 256
    assert res.precision["shots_differing_from_complex128"] == len(differing)
    assert res.precision["max_abs_dp"] == stats["max_abs_dp"]


def test_partial_measurement_counts_use_the_device_marginal_and_equal_the_oracle() -> None:
    """A 20-qubit circuit measuring 6 of its qubits: the marginal is summed on the device, and
    counts and outcome probabilities equal the oracle's (host NumPy marginal of the same seeded
    run); a noisy density-matrix run measuring 3 of 8 qubits likewise."""
    from lucivant_sim.engines import device_marginal_pays

    n = 20
    measured = [2, 5, 11, 12, 16, 19]
    body = wl.random_layers(n, 4, 77, measure=False).replace(f"qubit[{n}] q;", f"qubit[{n}] q;\nbit[6] c;")
    src = body + "".join(f"c[{i}] = measure q[{q}];\n" for i, q in enumerate(measured))
    c = ls.parse(src)
    assert device_marginal_pays(n, measured)
    res = ls.run(src, 3000, seed=11)
    assert res.counts == orc.run_counts(c, 3000, 11)
    want = orc.marginal(orc.probabilities_sv(orc.evolve_sv(c)), n, measured)
    got = ls.outcome_probabilities(src)
    assert abs(sum(got.values()) - 1) < 1e-12
    for i, p in enumerate(want):
        if p > 1e-12:
            assert abs(got[c.format_bits(i)] - p) < 1e-12
    noise = wl.noise_c3()
    part = wl.config_c3(8, 3).replace("c = measure q;", "c[0] = measure q[3];\nc[1] = measure q[6];\nc[2] = measure q[7];")
    assert device_marginal_pays(8, [3, 6, 7])
    r = ls.run(part, 2000, seed=5, noise=noise, method="density_matrix")
    assert r.counts == orc.run_counts(ls.parse(part), 2000, 5, noise=noise, method="density_matrix")


def test_permutation_layers_run_as_moves_and_equal_the_oracle() -> None:
    """A circuit whose fused blocks include pure permutations (layers of x, cx, swap and ccx,
    with s and t phases between them): statevector and density-matrix results equal the oracle's
    within 1e-10, and the plan does contain permutation passes."""
    from lucivant_sim.plan import plan_statevector

    src = 'OPENQASM 3.0;\ninclude "stdgates.inc";\nqubit[9] q;\n'
    for layer in range(4):
        src += f"h q[{layer}];\nx q[{(layer + 2) % 9}];\ncx q[{layer}], q[{(layer + 3) % 9}];\nswap q[{(layer + 4) % 9}], q[{(layer + 5) % 9}];\n"
        src += f"ccx q[{(layer + 6) % 9}], q[{(layer + 7) % 9}], q[{(layer + 8) % 9}];\ns q[{(layer + 1) % 9}];\nt q[{(layer + 4) % 9}];\n"
    c = ls.parse(src)
    passes = plan_statevector(c.operations, 9, 3)
    assert any(p.moves() is not None for p in passes)
    assert np.abs(ls.statevector(src) - orc.evolve_sv(c)).max() < 1e-10
    noise = ls.NoiseModel.uniform(single=0.01, two=0.02)
    assert np.abs(ls.density_matrix(src, noise) - orc.evolve_dm(c, noise)).max() < 1e-10
