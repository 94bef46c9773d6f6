"""The LDS program kernel (one workgroup per circuit instance or shot) against the oracle, the
HBM path and the reference's golden counts."""

from __future__ import annotations

import numpy as np
import pytest

import lucivant_sim as ls
from lucivant_sim import _native as nat
from lucivant_sim import program as pg
from lucivant_sim import workloads as wl
from lucivant_sim.batch import expectation_many, run_many
from lucivant_sim.engines import pauli_masks
from lucivant_sim.trajectories import run_batched, run_lds
from oracle import engine as orc

pytestmark = pytest.mark.gpu


def test_program_kernel_fits_thirteen_qubits() -> None:
    assert pg.max_qubits(nat.C128) >= 13
    assert pg.max_qubits(nat.C64) >= 14


@pytest.mark.parametrize("fuse", [0, 2, 4])
def test_static_programs_equal_the_oracle(fuse: int) -> None:
    srcs = [wl.random_layers(10, 5, s, two_qubit=("cx", "cz", "swap")) for s in range(5)]
    circuits = [ls.parse(s) for s in srcs]
    progs = [pg.compile_static(c, fuse) for c in circuits]
    out = pg.run_programs(progs, want_state=True, want_probs=True)
    for j, c in enumerate(circuits):
        want = orc.evolve_sv(c)
        assert np.abs(out.state[j] - want).max() < 1e-10  # type: ignore[index]
        assert np.abs(out.probs[j] - np.abs(want) ** 2).max() < 1e-10  # type: ignore[index]


def test_thirteen_qubit_program_and_wide_gates() -> None:
    src = wl.random_layers(13, 3, 9) + "ccx q[0], q[5], q[12];\nc3x q[1], q[2], q[3], q[11];\ngphase(0.4);\n"
    c = ls.parse(src)
    out = pg.run_programs([pg.compile_static(c)], want_state=True)
    assert np.abs(out.state[0] - orc.evolve_sv(c)).max() < 1e-10  # type: ignore[index]


def test_five_qubit_dense_gates_leave_the_program_path() -> None:
    """Programs take dense operators of at most 4 qubits; a circuit with a 5-qubit gate is refused
    by the program compiler and the batch API runs it on the single-circuit path instead."""
    from lucivant_sim.batch import run_many

    wide = wl.random_layers(8, 2, 1) + "c4x q[1], q[2], q[3], q[4], q[7];\n"
    with pytest.raises(pg.Unsupported):
        pg.compile_static(ls.parse(wide))
    narrow = wl.random_layers(8, 2, 2)
    res = run_many([wide, narrow], 500, seeds=[4, 5])
    assert [r.counts for r in res] == [ls.run(wide, 500, seed=4).counts, ls.run(narrow, 500, seed=5).counts]


def test_complex64_program_within_1e5() -> None:
    c = ls.parse(wl.random_layers(12, 10, 3))
    out = pg.run_programs([pg.compile_static(c)], dtype=nat.C64, want_state=True)
    assert np.abs(out.state[0] - orc.evolve_sv(c)).max() < 1e-5  # type: ignore[index]


def test_in_kernel_expectations() -> None:
    srcs = [wl.hea_ansatz(9, 4, s) for s in range(3)]
    labels = [wl.pauli_strings(9, 16, 10 + s) for s in range(3)]
    progs = [pg.compile_static(ls.parse(s)) for s in srcs]
    out = pg.run_programs(progs, pauli=[[pauli_masks(lbl) for lbl in ls_] for ls_ in labels])
    for j, s in enumerate(srcs):
        psi = orc.evolve_sv(ls.parse(s))
        for t, lbl in enumerate(labels[j]):
            assert abs(out.expect[j, t] - orc.expectation_sv(psi, lbl, 9)) < 1e-10  # type: ignore[index]


@pytest.mark.parametrize("name", ["c4_recipe_9q_seed7", "teleport_ideal_seed4"])
def test_lds_trajectories_reproduce_the_golden_counts(name: str, golden_runs: dict, golden_programs: dict) -> None:
    rec = golden_runs[name]
    c = ls.parse(golden_programs[name])
    noise = wl.noise_c4() if "c4" in name else None
    bits = run_lds(c, noise, rec["shots"], np.random.default_rng(rec["seed"]))
    assert bits is not None
    uniq, cnt = np.unique(bits, return_counts=True)
    assert {c.format_bits(int(v)): int(k) for v, k in zip(uniq, cnt)} == rec["counts"]


def test_lds_and_hbm_trajectories_agree_shot_by_shot() -> None:
    c = ls.parse(wl.teleport_chain(10, 3))
    noise = wl.noise_c4().add_readout_error(ls.ReadoutError(0.02, 0.05)).add_gate_error(ls.thermal_relaxation(50, 40, 2.0), gates="cx")
    a = run_lds(c, noise, 500, np.random.default_rng(8))
    b = run_batched(c, noise, 500, np.random.default_rng(8))
    assert a is not None and np.array_equal(a, b)


def test_run_many_equals_run() -> None:
    srcs = [wl.random_layers(8, 4, s) for s in range(6)]
    seeds = list(range(100, 106))
    many = run_many(srcs, 500, seeds=seeds)
    for src, seed, r in zip(srcs, seeds, many):
        one = ls.run(src, 500, seed=seed)
        assert r.counts == one.counts and r.fingerprint == one.fingerprint
    tele = [wl.teleport_chain(9, 3, theta=0.3 * s) for s in range(1, 4)]
    noise = wl.noise_c4()
    many = run_many(tele, 200, seeds=[1, 2, 3], noise=noise)
    for src, seed, r in zip(tele, [1, 2, 3], many):
        assert r.counts == ls.run(src, 200, seed=seed, noise=noise).counts


def test_expectation_many_equals_expectation() -> None:
    srcs = [wl.hea_ansatz(10, 3, s) for s in range(4)]
    rng = np.random.default_rng(2)
    obs = [{lbl: float(w) for lbl, w in zip(wl.pauli_strings(10, 16, s), rng.uniform(-1, 1, 16))} for s in range(4)]
    got = expectation_many(srcs, obs)
    for src, o, g in zip(srcs, obs, got):
        assert abs(g - ls.expectation(src, o)) < 1e-10
