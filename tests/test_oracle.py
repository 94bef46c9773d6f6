"""Pins the oracle (oracle/engine.py) before anything is compared against it.

Each check is chosen so that a plausible slip in the restatement (wrong axis for a qubit, matrix
bits in the wrong order, a transposed or unconjugated operand, a dropped Kraus term, a different
draw order) fails: brute-force operators, closed forms, invariants, and the counts the reference
itself produced for fixed seeds (tests/golden/runs.json, written by tests/golden/make_golden.py).
"""

from __future__ import annotations

import math
from itertools import permutations

import numpy as np
import pytest

import lucivant_sim as ls
from lucivant_sim import workloads as wl
from oracle import engine as orc

pytestmark = pytest.mark.filterwarnings("ignore::DeprecationWarning")


def _state(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    return v / np.linalg.norm(v)


def _unitary(k: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(2**k, 2**k)) + 1j * rng.normal(size=(2**k, 2**k)))
    return q


MATRICES = {
    1: {
        "general": _unitary(1, 1),
        "projector": np.array([[1, 0], [0, 0]], dtype=np.complex128),
        "y": orc.PAULI["Y"],
    },
    2: {
        "general": _unitary(2, 2),
        "cx": np.eye(4, dtype=np.complex128)[[0, 1, 3, 2]],
        "lower": np.array([[1, 0, 0, 0], [0.5, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=np.complex128),
    },
    3: {"general": _unitary(3, 3)},
}


@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("n", [3, 4])
def test_contraction_matches_the_brute_force_operator_on_every_placement(n: int, k: int) -> None:
    for name, m in MATRICES[k].items():
        for qubits in permutations(range(n), k):
            psi = _state(n, 31 * n + sum(qubits))
            got = orc.apply_sv(psi, m, qubits, n)
            want = orc.full_operator(m, qubits, n) @ psi
            assert np.allclose(got, want, atol=1e-13), (name, qubits)


_S = np.diag([1, 1j]).astype(np.complex128)
_SWAP = np.eye(4, dtype=np.complex128)[[0, 2, 1, 3]]
_CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)
_ISWAPISH = np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, -1, 0, 0], [0, 0, 0, 1j]], dtype=np.complex128)


@pytest.mark.parametrize(
    "n, matrix, qubits",
    [
        (13, _unitary(1, 5), (4,)),  # below 2^14 entries: the contraction
        (15, _unitary(1, 6), (3,)),  # 8 entries below the axis: the batched 2x2 product
        (15, _unitary(1, 6), (14,)),
        (15, _unitary(1, 7), (1,)),  # general, too few entries below and below 2^18: contraction
        (18, _unitary(1, 7), (1,)),  # general slices
        (18, _unitary(1, 8), (0,)),
        (18, MATRICES[1]["projector"], (2,)),  # an all-zero row clears its slice
        (14, orc.PAULI["X"], (0,)),  # moves
        (14, orc.PAULI["Y"], (13,)),
        (14, _S, (6,)),
        (14, MATRICES[2]["cx"], (0, 9)),  # asymmetric: control and target must not swap
        (14, MATRICES[2]["cx"], (9, 0)),
        (14, _SWAP, (2, 11)),
        (14, _CZ, (5, 1)),
        (14, _ISWAPISH, (12, 3)),
        (15, _unitary(2, 9), (3, 10)),  # general two-qubit below 2^18: contraction
        (18, _unitary(2, 9), (3, 10)),  # general two-qubit slices
        (18, MATRICES[2]["lower"], (17, 0)),
    ],
)
def test_reference_dispatch_equals_the_contraction(n: int, matrix: np.ndarray, qubits: tuple[int, ...]) -> None:
    """Every route of the reference's size-dependent dispatch gives the contraction's result,
    which the brute-force operators pin; the contraction is applied to an untouched copy because
    the slice routes work in place."""
    psi = _state(n, 3 * n + qubits[0]).reshape([2] * n)
    axes = [n - 1 - q for q in qubits]
    want = orc.contract(psi.copy(), matrix, axes)
    got = orc.apply_slices(psi.copy(), matrix, axes)
    assert got.shape == psi.shape
    assert np.allclose(got, want, rtol=0, atol=1e-14)


def test_one_qubit_runs_fuse_into_their_product() -> None:
    """Runs of one-qubit gates on a qubit become one gate; a two-qubit gate on the qubit ends the
    run, an if/else or barrier ends all of them, and the fused circuit's state equals the
    unfused one's."""
    from lucivant_sim.circuit import Barrier, Gate

    src = "OPENQASM 3.0;\ninclude \"stdgates.inc\";\nqubit[3] q;\nh q[0];\nt q[0];\nrx(0.3) q[1];\ncx q[0], q[1];\ns q[0];\nsx q[0];\nry(0.7) q[2];\nrz(0.2) q[2];\n"
    circuit = ls.parse(src)
    fused = orc.fuse_one_qubit_runs(circuit.operations)
    assert [len(op.qubits) for op in fused if isinstance(op, Gate)] == [1, 1, 2, 1, 1]
    assert [op.name for op in fused if isinstance(op, Gate)] == ["fused", "rx", "cx", "fused", "fused"]
    n = 3
    psi = np.zeros(8, dtype=np.complex128)
    psi[0] = 1
    for op in fused:
        psi = orc.apply_sv(psi, op.matrix, op.qubits, n)
    assert np.allclose(psi, orc.evolve_sv(circuit), atol=1e-14)
    h = circuit.operations[0]
    split = orc.fuse_one_qubit_runs((h, Barrier(qubits=(1,)), h))
    assert len([op for op in split if isinstance(op, Gate)]) == 2


def test_full_operator_is_little_endian_with_the_first_qubit_most_significant() -> None:
    # cx with control qubit 0 and target qubit 1 maps |01> (index 1: qubit 0 set) to |11> (index 3)
    cx = MATRICES[2]["cx"]
    op = orc.full_operator(cx, (0, 1), 2)
    assert op[3, 1] == 1 and op[1, 3] == 1 and op[0, 0] == 1 and op[2, 2] == 1


def _mixed(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(2**n, 2**n)) + 1j * rng.normal(size=(2**n, 2**n))
    rho = a @ a.conj().T
    return rho / np.trace(rho)


def test_density_gate_is_u_rho_u_dagger() -> None:
    n = 3
    rho = _mixed(n, 4)
    for qubits in [(0,), (2,), (1, 2), (2, 0)]:
        u = _unitary(len(qubits), 9)
        full = orc.full_operator(u, qubits, n)
        assert np.allclose(orc.apply_dm(rho, u, qubits, n), full @ rho @ full.conj().T, atol=1e-13)


@pytest.mark.parametrize("qubits", [(0,), (2,), (0, 2), (2, 1)])
@pytest.mark.parametrize("p", [0.0, 0.05, 0.5, 1.0])
def test_analytic_depolarizing_equals_the_kraus_sum(qubits: tuple[int, ...], p: float) -> None:
    n = 3
    rho = _mixed(n, 7 + len(qubits))
    ch = ls.depolarizing(p, len(qubits))
    explicit = sum(orc.apply_dm(rho, k, qubits, n) for k in ch.kraus)
    assert np.allclose(orc.depolarizing_trace(rho, p, qubits, n), explicit, atol=1e-12)


@pytest.mark.parametrize("k", [1, 2])
def test_depolarizing_invariants_and_composition(k: int) -> None:
    n = 3
    rho = _mixed(n, 11)
    q = tuple(range(k))
    assert np.array_equal(orc.apply_channel_dm(rho, ls.depolarizing(0.0, k), q, n), rho)
    for p in (0.01, 0.3, 1.0, 4**k / (4**k - 1)):
        out = orc.apply_channel_dm(rho, ls.depolarizing(p, k), q, n)
        assert np.trace(out) == pytest.approx(1, abs=1e-12)
        assert np.allclose(out, out.conj().T, atol=1e-13)
        assert np.linalg.eigvalsh(out).min() > -1e-12
    small = _mixed(k, 3)
    assert np.allclose(orc.apply_channel_dm(small, ls.depolarizing(1.0, k), q, k), np.eye(2**k) / 2**k, atol=1e-12)
    twice = orc.apply_channel_dm(orc.apply_channel_dm(rho, ls.depolarizing(0.1, k), q, n), ls.depolarizing(0.2, k), q, n)
    assert np.allclose(twice, orc.apply_channel_dm(rho, ls.depolarizing(1 - 0.9 * 0.8, k), q, n), atol=1e-12)


def test_custom_channel_named_depolarizing_uses_its_own_kraus() -> None:
    ch = ls.Channel("depolarizing", 1, (orc.PAULI["X"],))
    rho = np.diag([1.0, 0.0]).astype(np.complex128)
    assert np.allclose(orc.apply_channel_dm(rho, ch, (0,), 1), np.diag([0, 1]))


def test_analytic_values_of_the_reference_tests() -> None:
    bell = ls.parse('OPENQASM 3; include "stdgates.inc"; qubit[2] q; h q[0]; cx q[0], q[1];')
    noise = ls.NoiseModel().add_gate_error(ls.depolarizing(0.2, 2), gates="cx")
    rho = orc.evolve_dm(bell, noise)
    assert orc.expectation_dm(rho, "XX", 2) == pytest.approx(0.8)
    psi = orc.evolve_sv(bell)
    assert orc.expectation_sv(psi, "ZZ", 2) + 3 * orc.expectation_sv(psi, "ZI", 2) == pytest.approx(1.0)
    one = ls.parse("OPENQASM 3; qubit q; x q;")
    relax = ls.NoiseModel().add_gate_error(ls.thermal_relaxation(t1=10.0, t2=8.0, time=3.0), gates="x")
    assert orc.probabilities_dm(orc.evolve_dm(one, relax))[1] == pytest.approx(math.exp(-0.3))


def test_pauli_labels_put_qubit_0_rightmost() -> None:
    """Closed forms that fix the label order (simulator.py:511, 524: qubit 0 is the rightmost
    letter): after x q[0] the state is |01>, so <IZ> = -1 and <ZI> = +1; after rx(pi/2) q[0],
    <IY> = -sin(pi/2) = -1 and <YI> = 0; after h q[1], <XI> = 1 and <IX> = 0."""
    one = orc.evolve_sv(ls.parse("OPENQASM 3; qubit[2] q; x q[0];"))
    assert orc.expectation_sv(one, "IZ", 2) == pytest.approx(-1.0)
    assert orc.expectation_sv(one, "ZI", 2) == pytest.approx(1.0)
    rot = orc.evolve_sv(ls.parse('OPENQASM 3; include "stdgates.inc"; qubit[2] q; rx(pi/2) q[0];'))
    assert orc.expectation_sv(rot, "IY", 2) == pytest.approx(-1.0)
    assert orc.expectation_sv(rot, "YI", 2) == pytest.approx(0.0, abs=1e-15)
    had = orc.evolve_sv(ls.parse('OPENQASM 3; include "stdgates.inc"; qubit[2] q; h q[1];'))
    assert orc.expectation_sv(had, "XI", 2) == pytest.approx(1.0)
    assert orc.expectation_sv(had, "IX", 2) == pytest.approx(0.0, abs=1e-15)
    rho = orc.evolve_dm(ls.parse("OPENQASM 3; qubit[2] q; x q[0];"), None)
    assert orc.expectation_dm(rho, "IZ", 2) == pytest.approx(-1.0)
    assert orc.expectation_dm(rho, "ZI", 2) == pytest.approx(1.0)


def test_expectations_match_dense_pauli_matrices() -> None:
    n = 3
    psi = _state(n, 5)
    rho = _mixed(n, 6)
    for label in orc.all_labels(n):
        p = orc.pauli_matrix(label)
        assert orc.expectation_sv(psi, label, n) == pytest.approx(float(np.vdot(psi, p @ psi).real), abs=1e-12)
        assert orc.expectation_dm(rho, label, n) == pytest.approx(float(np.trace(p @ rho).real), abs=1e-12)


def test_ideal_density_matrix_is_the_pure_state() -> None:
    c = ls.parse(wl.random_layers(4, 3, 77, two_qubit=("cx", "cz"), measure=False))
    psi = orc.evolve_sv(c)
    assert np.allclose(orc.evolve_dm(c, None), np.outer(psi, psi.conj()), atol=1e-12)


GOLDEN_METHODS = {
    "bell_seed7": ("statevector", None),
    "c1_12q_seed7": ("statevector", None),
    "c3_recipe_6q_seed7": ("density_matrix", "c3"),
    "c4_recipe_9q_seed7": ("trajectories", "c4"),
    "teleport_noisy_seed3": ("trajectories", "teleport"),
    "teleport_ideal_seed4": ("trajectories", None),
    "bell_noisy_seed42": ("density_matrix", "bell"),
}


def golden_noise(tag: str | None) -> ls.NoiseModel | None:
    """The noise model each golden run was made with (tests/golden/make_golden.py)."""
    if tag is None:
        return None
    if tag == "c3":
        return wl.noise_c3()
    if tag == "c4":
        return wl.noise_c4()
    if tag == "teleport":
        return ls.NoiseModel.uniform(single=0.01, two=0.02, readout=0.01)
    return ls.NoiseModel.uniform(single=0.01, two=0.03, readout=0.01)


@pytest.mark.parametrize("name", sorted(GOLDEN_METHODS))
def test_oracle_counts_equal_the_reference_counts(name: str, golden_runs: dict, golden_programs: dict) -> None:
    rec = golden_runs[name]
    method, tag = GOLDEN_METHODS[name]
    assert rec["method"] == method
    circuit = ls.parse(golden_programs[name])
    counts = orc.run_counts(circuit, rec["shots"], rec["seed"], golden_noise(tag), method)
    assert counts == rec["counts"]


def test_readme_bell_counts(golden_runs: dict) -> None:
    assert golden_runs["bell_seed7"]["counts"] == {"00": 502, "11": 498}


def test_oracle_trajectories_agree_with_the_oracle_density_matrix() -> None:
    bell = ls.parse('OPENQASM 3; include "stdgates.inc"; qubit[2] q; bit[2] c; h q[0]; cx q[0], q[1]; c = measure q;')
    noise = ls.NoiseModel().add_gate_error(ls.amplitude_damping(0.15), gates="h").add_gate_error(ls.depolarizing(0.1, 2), gates="cx")
    p = orc.probabilities_dm(orc.evolve_dm(bell, noise))
    exact = {format(i, "02b"): float(v) for i, v in enumerate(p) if v > 1e-12}
    traj = orc.trajectory_counts(bell, noise, 4000, np.random.default_rng(9))
    assert orc.tvd(traj, exact) < 0.03


def test_recorded_distribution_known_answers() -> None:
    """The oracle's readout expansion: closed forms for one qubit (p01 applies to a true 0, p10 to
    a true 1), a fan-out into two clbits (both reads flip independently), a pure relabelling of
    bits without error, and total probability kept."""
    one = orc.recorded_distribution(np.array([0.4, 0.6]), [2], [(2, 0)], {2: (0.1, 0.3)})
    assert one == pytest.approx({0: 0.4 * 0.9 + 0.6 * 0.3, 1: 0.4 * 0.1 + 0.6 * 0.7}, rel=1e-15)
    fan = orc.recorded_distribution(np.array([1.0, 0.0]), [0], [(0, 0), (0, 3)], {0: (0.1, 0.0)})
    assert fan == pytest.approx({0: 0.81, 1: 0.09, 8: 0.09, 9: 0.01}, rel=1e-14)
    moved = orc.recorded_distribution(np.array([0.1, 0.2, 0.3, 0.4]), [1, 4], [(4, 0), (1, 5)], {})
    assert moved == pytest.approx({0: 0.1, 32: 0.2, 1: 0.3, 33: 0.4}, rel=1e-15)
    rng = np.random.default_rng(3)
    d = rng.random(8)
    d /= d.sum()
    total = orc.recorded_distribution(d, [0, 1, 2], [(0, 1), (1, 0), (2, 2), (2, 3)], {0: (0.2, 0.05), 2: (0.01, 0.3)}, threshold=-1.0)
    assert sum(total.values()) == pytest.approx(1.0, rel=1e-14)
