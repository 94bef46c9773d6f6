"""Host logic on the CPU: the OpenQASM reader, noise models, method selection, digests and
fingerprints against the reference's own outputs, and the fusion planner against the oracle."""

from __future__ import annotations

import numpy as np
import pytest

import lucivant_sim as ls
from lucivant_sim import workloads as wl
from lucivant_sim.plan import Pass, plan_density, plan_statevector
from lucivant_sim.simulator import Result, choose_method
from lucivant_sim.trajectories import _kraus_choice, _mixture_choice, compile_program, draw_count
from oracle import engine as orc

# ------------------------------------------------------------------------------------ reader


def test_reader_digests_and_labels_equal_the_reference(golden_host: dict) -> None:
    for name, rec in golden_host["reader"].items():
        c = ls.parse(rec["source"], rec["params"])
        assert c.digest() == rec["digest"], name
        assert (c.num_qubits, c.num_clbits) == (rec["num_qubits"], rec["num_clbits"]), name
        labels = [[getattr(op, "name", op.kind), list(getattr(op, "qubits", ()))] for op in c.operations]
        assert labels == rec["labels"], name


def test_config_programs_have_the_reference_digests(golden_host: dict) -> None:
    src, labels = wl.config_c2()
    assert ls.parse(src).digest() == golden_host["config_digests"]["c2"]
    assert labels == golden_host["c2_labels"]
    assert ls.parse(wl.config_c3()).digest() == golden_host["config_digests"]["c3"]
    assert ls.parse(wl.teleport_chain(20, 3)).digest() == golden_host["config_digests"]["c4"]


def test_broadcast_slices_and_sets() -> None:
    c = ls.parse("OPENQASM 3; qubit[4] q; qubit[4] r; h q; cx q, r; x q[1:2]; z q[{0, 3}]; y q[-1];")
    names = [(op.name, op.qubits) for op in c.operations]
    assert names[:4] == [("h", (0,)), ("h", (1,)), ("h", (2,)), ("h", (3,))]
    assert names[4:8] == [("cx", (0, 4)), ("cx", (1, 5)), ("cx", (2, 6)), ("cx", (3, 7))]
    assert names[8:] == [("x", (1,)), ("x", (2,)), ("z", (0,)), ("z", (3,)), ("y", (3,))]


def test_missing_and_unknown_inputs_are_reported() -> None:
    with pytest.raises(ls.QasmError, match="needs a value for its input 'theta'"):
        ls.parse("OPENQASM 3; input float theta; qubit q; rx(theta) q;")
    with pytest.raises(ls.QasmError, match="no input named phi"):
        ls.parse("OPENQASM 3; qubit q; h q;", params={"phi": 1.0})


@pytest.mark.parametrize(
    ("source", "message"),
    [
        ("OPENQASM 3; qubit q; foo q;", "line 1, column 22: unknown gate 'foo'"),
        ("OPENQASM 3;\nqubit[2] q;\ncx q[0], q[0];", "line 3, column 1: cx is given the same qubit twice"),
        ("OPENQASM 3; qubit[2] q; h q[2];", "index 2 is outside 'q'"),
        ("OPENQASM 3; qubit q; for uint i in [0:3] { h q; }", "'for' isn't supported"),
        ("OPENQASM 3; qubit q; rx q;", "rx takes 1 parameter"),
        ("OPENQASM 3; qubit[2] q; qubit[3] r; cx q, r;", "different sizes"),
        ('OPENQASM 2.0; include "mine.inc"; qreg q[1];', "only qelib1.inc and stdgates.inc"),
        ("OPENQASM 4; qubit q;", "only OpenQASM 2 and 3"),
        ("OPENQASM 3; qubit q; h q", "expected ';'"),
    ],
)
def test_errors_point_at_the_problem(source: str, message: str) -> None:
    with pytest.raises(ls.QasmError, match=message.replace("(", r"\(").replace("[", r"\[")):
        ls.parse(source)


def test_includes_resolve_only_inside_the_folder(tmp_path) -> None:  # type: ignore[no-untyped-def]
    (tmp_path / "mine.inc").write_text("gate flip a { x a; }")
    (tmp_path / "main.qasm").write_text('OPENQASM 3; include "mine.inc"; qubit q; flip q;')
    c = ls.load(tmp_path / "main.qasm")
    assert np.allclose(orc.evolve_sv(c), [0, 1])
    (tmp_path / "bad.qasm").write_text('OPENQASM 3; include "../../etc/passwd"; qubit q;')
    with pytest.raises(ls.QasmError, match="outside"):
        ls.load(tmp_path / "bad.qasm")


# ------------------------------------------------------------------------------------- noise


def test_noise_digests_equal_the_reference(golden_host: dict) -> None:
    models = {
        "c3": wl.noise_c3(),
        "c4": wl.noise_c4(),
        "uniform_bell": ls.NoiseModel.uniform(single=0.01, two=0.03, readout=0.01),
        "uniform_teleport": ls.NoiseModel.uniform(single=0.01, two=0.02, readout=0.01),
        "calibration": ls.NoiseModel.from_calibration(
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
    for name, model in models.items():
        assert model.digest() == golden_host["noise_digests"][name], name


def test_calibration_noise_prefers_per_qubit_values() -> None:
    cal = {
        "device": "example_device",
        "single_qubit_error": 0.001,
        "two_qubit_error": 0.01,
        "readout_error": 0.02,
        "qubits": [{"qubit": 1, "single_qubit_error": 0.3, "readout_error": None}],
        "couplers": [{"qubits": [0, 1], "error": 0.05}],
    }
    model = ls.NoiseModel.from_calibration(cal)
    assert model.gate_error("sx", (1,)).params[0] == pytest.approx(0.3 * 2)  # type: ignore[union-attr]
    assert model.gate_error("sx", (0,)).params[0] == pytest.approx(0.001 * 2)  # type: ignore[union-attr]
    assert model.gate_error("cx", (1, 0)).params[0] == pytest.approx(0.05 * 4 / 3)  # type: ignore[union-attr]
    assert model.gate_error("rz", (0,)) is None
    assert model.readout_error(1) == ls.ReadoutError(0.02, 0.02)


def test_invalid_channels() -> None:
    with pytest.raises(ValueError, match="T2 can't exceed"):
        ls.thermal_relaxation(1.0, 3.0, 0.1)
    with pytest.raises(ValueError, match="between 0 and"):
        ls.bit_flip(1.5)


@pytest.mark.parametrize("k", [1, 2])
def test_superoperators_equal_the_oracle_channel(k: int) -> None:
    n = 3
    rng = np.random.default_rng(3)
    a = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    rho = a @ a.conj().T
    rho /= np.trace(rho)
    qubits = (2,) if k == 1 else (2, 0)
    channels = [ls.depolarizing(0.3, k), ls.Channel("generic", k, ls.depolarizing(0.3, k).kraus)]
    if k == 1:
        channels += [ls.amplitude_damping(0.2), ls.thermal_relaxation(120, 90, 0.05), ls.phase_damping(0.4)]
    for ch in channels:
        want = orc.apply_channel_dm(rho, ch, qubits, n)
        bits = tuple(n + q for q in qubits) + qubits
        got = orc.apply_sv(rho.reshape(-1), ch.superoperator(), bits, 2 * n).reshape(8, 8)
        assert np.allclose(got, want, atol=1e-13), ch.name


# ---------------------------------------------------------------------------- method choice


def test_method_choice_equals_the_reference(golden_host: dict) -> None:
    g = golden_host["methods"]
    c2_src, _ = wl.config_c2()
    cases = {
        "c1": (wl.config_c1(), None),
        "c2": (c2_src, None),
        "c3_automatic": (wl.config_c3(), wl.noise_c3()),
        "c4": (wl.teleport_chain(20, 3), wl.noise_c4()),
        "clifford_12q": ("OPENQASM 3; qubit[12] q; h q; cx q[0], q[1]; s q[3];", None),
        "clifford_12q_noisy": ("OPENQASM 3; qubit[12] q; h q; cx q[0], q[1];", ls.NoiseModel.uniform(single=0.01)),
        "clifford_12q_ampdamp": ("OPENQASM 3; qubit[12] q; h q;", ls.NoiseModel().add_gate_error(ls.amplitude_damping(0.1))),
        "t_12q": ("OPENQASM 3; qubit[12] q; h q; t q[0];", None),
        "noisy_11q": ("OPENQASM 3; qubit[11] q; h q;", ls.NoiseModel.uniform(single=0.01)),
        "noisy_10q": ("OPENQASM 3; qubit[10] q; h q;", ls.NoiseModel.uniform(single=0.01)),
    }
    for name, (src, noise) in cases.items():
        assert choose_method(ls.parse(src), noise) == g[name], name


def test_explicit_method_refusals() -> None:
    bell = ls.parse('OPENQASM 3; include "stdgates.inc"; qubit[2] q; bit[2] c; h q[0]; cx q[0], q[1]; c = measure q;')
    with pytest.raises(ls.SimulationError, match="statevector method"):
        choose_method(bell, ls.NoiseModel.uniform(single=0.01), "statevector")
    with pytest.raises(ls.SimulationError, match="up to 13"):
        choose_method(ls.parse("OPENQASM 3; qubit[14] q; h q;"), None, "density_matrix")
    with pytest.raises(ValueError, match="method must be one of"):
        choose_method(bell, None, "magic")


def test_fingerprints_equal_the_reference(golden_runs: dict) -> None:
    for name, rec in golden_runs.items():
        r = Result(
            counts=rec["counts"],
            shots=rec["shots"],
            seed=rec["seed"],
            method=rec["method"],
            num_qubits=rec["num_qubits"],
            num_clbits=rec["num_clbits"],
            circuit_digest=rec["circuit_digest"],
            noise_digest=rec["noise_digest"],
            version=ls.__version__,
            elapsed_seconds=0.0,
        )
        assert r.fingerprint == rec["fingerprint"], name
        assert r.counts_digest == rec["counts_digest"], name


def test_program_digests_equal_the_golden_runs(golden_runs: dict, golden_programs: dict) -> None:
    for name, rec in golden_runs.items():
        assert ls.parse(golden_programs[name]).digest() == rec["circuit_digest"], name


# ------------------------------------------------------------------------------------ planner


def _run_passes_on_cpu(psi: np.ndarray, passes: list[Pass], nbits: int) -> np.ndarray:
    for p in passes:
        if p.kind == "scale":
            psi = psi * p.matrix[0]
        elif p.kind == "diag":
            psi = orc.apply_sv(psi, np.diag(p.matrix), p.bits, nbits)
        else:
            psi = orc.apply_sv(psi, p.matrix, p.bits, nbits)
    return psi


@pytest.mark.parametrize("low_width", [0, 2, 3])
@pytest.mark.parametrize("width", [1, 2, 3, 4, 5])
def test_fused_statevector_passes_equal_sequential_gates(width: int, low_width: int) -> None:
    src = wl.random_layers(7, 6, 11, two_qubit=("cx", "cz", "swap", "iswap"), measure=True)
    src = src.replace("c = measure q;", "gphase(0.3); ccx q[0], q[3], q[6]; c = measure q;")
    c = ls.parse(src)
    passes = plan_statevector(c.operations, c.num_qubits, width, 3, low_width)
    assert all(p.k <= max(width, 3) for p in passes)
    if low_width:
        assert all(p.k <= max(low_width, 3) for p in passes if min(p.bits, default=9) < 3)
    psi0 = np.zeros(2**7, dtype=np.complex128)
    psi0[0] = 1
    assert np.allclose(_run_passes_on_cpu(psi0, passes, 7), orc.evolve_sv(c), atol=1e-12)
    assert len(passes) < sum(1 for op in c.operations if isinstance(op, ls.circuit.Gate))
    unmerged = plan_statevector(c.operations, c.num_qubits, width, 3, low_width, merge=False)
    assert len(passes) <= len(unmerged)
    assert np.allclose(_run_passes_on_cpu(psi0, unmerged, 7), orc.evolve_sv(c), atol=1e-12)


@pytest.mark.parametrize("width", [1, 2])
def test_fused_density_passes_equal_the_oracle(width: int) -> None:
    n = 4
    src = wl.random_layers(n, 4, 5, two_qubit=("cx", "cz")) + "reset q[2];\nccx q[0], q[1], q[3];\n"
    c = ls.parse(src)
    noise = wl.noise_c3().add_gate_error(ls.depolarizing(0.05, 3), gates="ccx")
    items = plan_density(c.operations, n, noise, width)
    rho = np.zeros(4**n, dtype=np.complex128)
    rho[0] = 1
    for item in items:
        if isinstance(item, Pass):
            rho = _run_passes_on_cpu(rho, [item], 2 * n)
        else:
            _, (ch, qubits) = item
            rho = orc.apply_channel_dm(rho.reshape(2**n, 2**n), ch, qubits, n).reshape(-1)
    want = orc.evolve_dm(c, noise)
    assert np.allclose(rho.reshape(2**n, 2**n), want, atol=1e-12)


# ------------------------------------------------------------------------------ trajectories


def test_draw_ledger() -> None:
    c4 = ls.parse(wl.teleport_chain(20, 3))
    noise = wl.noise_c4()
    steps, d, nflags = compile_program(c4, noise)
    gates = sum(1 for s in steps if s.kind == "gate" and not s.guard)
    # every unconditioned gate is noisy (depolarising or amplitude damping), plus 6 measurements,
    # 3 final measurements, and per if/else one noisy one-qubit gate
    assert d == gates + 6 + 3 + 6
    assert nflags == 6
    assert draw_count(c4.operations, noise) == d
    teleport = ls.parse(
        'OPENQASM 3; include "stdgates.inc"; qubit[3] q; bit[2] m; h q[0]; m[0] = measure q[0]; if (m[0]) x q[2];'
    )
    assert draw_count(teleport.operations, None) == 1
    assert draw_count(teleport.operations, ls.NoiseModel.uniform(single=0.01)) is None


def test_batched_decisions_equal_the_scalar_rule() -> None:
    rng = np.random.default_rng(1)
    ch = ls.depolarizing(0.4, 2)
    r = rng.random(2000)
    sel, table = _mixture_choice(ch, r)
    for i in range(len(r)):
        acc, want = 0.0, -1
        for p, u in ch.mixture:  # type: ignore[union-attr]
            acc += p
            if r[i] < acc:
                want = -1 if np.allclose(u, np.eye(4)) else 1
                if want == 1:
                    got = table[sel[i]]
                    assert np.array_equal(got, u)
                break
        assert (sel[i] >= 0) == (want == 1)
    amp = ls.thermal_relaxation(120, 90, 30.0)
    a = rng.normal(size=(500, 2, 2)) + 1j * rng.normal(size=(500, 2, 2))
    rho = a @ np.conj(np.transpose(a, (0, 2, 1)))
    rho /= np.trace(rho, axis1=1, axis2=2)[:, None, None]
    r = rng.random(500)
    mats, sel = _kraus_choice(amp, rho, r)
    for i in range(500):
        acc, pick, w = 0.0, None, 0.0
        for j, k in enumerate(amp.kraus):
            w = float(np.trace(k @ rho[i] @ k.conj().T).real)
            acc += w
            if r[i] < acc and w > 0:
                pick = (j, w)
                break
        if pick is None:
            pick = (len(amp.kraus) - 1, w)
        assert sel[i] == i
        assert np.allclose(mats[i], amp.kraus[pick[0]] / np.sqrt(pick[1]))


# ----------------------------------------------------------------------- program validation


def test_program_tables_are_bounds_checked_before_launch() -> None:
    from lucivant_sim import program as pg

    c = ls.parse(wl.teleport_chain(9, 3))
    prog = pg.compile_trajectory(c, wl.noise_c4())
    table = np.array(prog.ops, dtype=pg.OP_DTYPE)
    reals = np.asarray(prog.reals)
    pg._validate(table, 9, prog.mat_size, reals, prog.draws)
    planted = [
        ("t", lambda t: t["t"].__setitem__((0, 0), 9), "beyond the state"),
        ("mat", lambda t: t["mat"].__setitem__(0, prog.mat_size), "matrix pool"),
        ("draw", lambda t: t["draw"].__setitem__(np.flatnonzero(t["kind"] == pg.OP_MEASURE)[0], prog.draws), "draw"),
        ("k", lambda t: t["k"].__setitem__(0, 0), "no target"),
        ("kind", lambda t: t["kind"].__setitem__(0, 9), "unknown op"),
    ]
    for name, plant, message in planted:
        bad = table.copy()
        plant(bad)
        with pytest.raises(ValueError, match=message):
            pg._validate(bad, 9, prog.mat_size, reals, prog.draws)
    with pytest.raises(ValueError, match="draw"):
        pg._validate(table, 9, prog.mat_size, reals, 0)
    # a mixture's branch offsets (kept among the reals, shared by the ops of one channel): each
    # must be -1 or a whole matrix inside the pool
    mixes = np.flatnonzero(table["kind"] == pg.OP_MIX)
    assert mixes.size
    for row in (mixes[0], mixes[-1]):
        aux, nb = int(table["aux"][row]), int(table["nb"][row])
        for off in (prog.mat_size, -2):
            bad_reals = reals.copy()
            bad_reals[aux + nb + nb - 1] = off
            with pytest.raises(ValueError, match="mixture branch"):
                pg._validate(table, 9, prog.mat_size, bad_reals, prog.draws)


def test_large_distributions_sample_as_the_oracle_with_a_reused_buffer() -> None:
    """Wide distributions accumulate into a kept per-thread buffer; counts still equal the
    oracle's sampler, also when a smaller distribution follows a larger one in the same buffer
    (a stale tail must not leak into the search)."""
    from lucivant_sim.simulator import _cumsum, _sample

    rng = np.random.default_rng(5)
    for n in (21, 20):
        p = rng.random(1 << n) ** 3
        p /= p.sum()
        assert np.array_equal(_cumsum(p), np.cumsum(p))
        src = f"OPENQASM 3.0;\ninclude \"stdgates.inc\";\nqubit[{n}] q;\nbit[{n}] c;\nc = measure q;\n"
        c = ls.parse(src)
        got = _sample(c, p, 3000, None, np.random.default_rng(11))
        want = orc.sample(c, p, 3000, None, np.random.default_rng(11))
        assert got == want, n


def test_basis_dict_equals_formatting_every_index() -> None:
    """probabilities() keys: the half-width tables give the same dict, in the same order, as
    formatting each kept index (simulator.py:440-442), with zero plateaus and exact-threshold
    entries dropped."""
    from lucivant_sim.simulator import _basis_dict

    rng = np.random.default_rng(2)
    for n in (1, 5, 12, 13, 17):
        p = rng.random(1 << n)
        p[rng.integers(0, 1 << n, size=max(1, (1 << n) // 7))] = 0.0
        p[0] = 1e-12
        p /= p.sum()
        want = {format(int(i), f"0{n}b"): float(p[i]) for i in np.flatnonzero(p > 1e-12)}
        got = _basis_dict(p, n, 1e-12)
        assert got == want and list(got) == list(want), n


def test_trajectory_queue_fuses_blocks_equal_to_program_order() -> None:
    """The trajectory batch fuses queued operators into blocks of at most 4 qubits with one
    operator per combination of per-shot choices. Applying the blocks in order, with each shot's
    combination, equals applying every item in program order (the oracle's tensordot), for fixed
    gates, mixture-like tables with identity choices (-1) and guarded gates."""
    from lucivant_sim.trajectories import _fuse, _group, _Queued

    rng = np.random.default_rng(9)
    n, shots = 6, 40

    def unitary(k: int) -> np.ndarray:
        q, _ = np.linalg.qr(rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k)))
        return q

    items = []
    for _ in range(30):
        k = int(rng.integers(1, 3))
        qubits = tuple(int(x) for x in rng.choice(n, size=k, replace=False))
        kind = rng.integers(0, 3)
        if kind == 0:
            items.append(_Queued(qubits, [unitary(k)]))
        else:
            table = [unitary(k) for _ in range(int(rng.integers(1, 4)))]
            idx = rng.integers(-1, len(table), size=shots)
            items.append(_Queued(qubits, table, idx))
    blocks = _group(items, 4)
    assert all(len(q) <= 4 for q, _ in blocks)
    assert sum(len(b) for _, b in blocks) == len(items)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    for b in range(0, shots, 7):
        want = psi0.copy()
        for it in items:
            j = 0 if it.idx is None else int(it.idx[b])
            if j >= 0:
                want = orc.apply_sv(want, it.table[j], it.qubits, n)
        got = psi0.copy()
        for qubits, members in blocks:
            per_shot = [it for it in members if it.idx is not None]
            combo = np.array([[int(it.idx[b]) for it in per_shot]], dtype=np.int64).reshape(1, len(per_shot))
            op = _fuse(qubits, members, combo)[0]
            got = orc.apply_sv(got, op, tuple(qubits), n)
        assert np.abs(got - want).max() < 1e-12, b


def test_blocks_split_when_host_fusion_would_outcost_a_pass() -> None:
    """Many distinct per-shot choices make host products dear: the block is cut into consecutive
    parts (order kept, nothing lost) for a small batch, and kept whole for a large one."""
    from lucivant_sim.trajectories import _Queued, _split_for_host

    rng = np.random.default_rng(4)
    items = [_Queued((i % 4,), [np.eye(2)] * 4, rng.integers(-1, 4, size=4096)) for i in range(8)]
    items.insert(3, _Queued((0, 1), [np.eye(4)]))
    small = _split_for_host(items, 4, pass_seconds=1e-5)
    assert len(small) > 1 and [it for part in small for it in part] == items
    large = _split_for_host(items, 4, pass_seconds=10.0)
    assert large == [items]


def test_server_starts_cheap_then_widest_count_requests_first() -> None:
    """The serving loop's start order for wide requests: cheap ones first, then count requests
    widest first (their host sampling overlaps the rest), then the others by admission."""
    from types import SimpleNamespace

    from lucivant_sim.serving import Server

    def job(seq: int, n: int, passes: int, terms=None):  # type: ignore[no-untyped-def]
        return SimpleNamespace(seq=seq, c=SimpleNamespace(num_qubits=n, passes=[0] * passes), terms=terms, method="statevector")

    jobs = [
        job(0, 26, 300, {"Z" * 26: 1.0}),
        job(1, 24, 200),
        job(2, 14, 80),
        job(3, 26, 310),
        job(4, 25, 90, {"Z" * 25: 1.0}),
        job(5, 16, 40, {"Z" * 16: 1.0}),
    ]
    order = [j.seq for j in sorted(jobs, key=Server._priority)]
    assert order == [2, 5, 3, 1, 0, 4]


class _HostBatch:
    """The trajectory batch's five device operations restated on the host (the oracle's
    tensordot for operators), so its queueing, fusion, kept reduced density matrices and shared
    states (slots) can be checked against the reference's sequential rule on the CPU."""

    @staticmethod
    def make(n: int, nshots: int, nflags: int, num_clbits: int, share: bool = True):  # type: ignore[no-untyped-def]
        from lucivant_sim import trajectories as tr

        class Batch(tr._Batch):
            passes = 0
            reductions = 0

            def _allocate(self) -> None:
                self.states = np.full((self.b, 1 << self.n), np.nan, dtype=np.complex128)
                self.states[: self.s] = 0.0
                self.states[: self.s, 0] = 1.0

            def run_pass(self, p) -> None:  # type: ignore[no-untyped-def]
                Batch.passes += 1
                m = np.diag(p.matrix) if p.kind == "diag" else p.matrix
                for b in range(self.s):
                    self.states[b] = orc.apply_sv(self.states[b], m, p.bits, self.n)

            def apply_selected(self, kind, qubits, table, sel) -> None:  # type: ignore[no-untyped-def]
                if not np.any(sel >= 0):
                    return
                Batch.passes += 1
                for b in np.flatnonzero(sel >= 0):
                    m = table[sel[b]]
                    m = np.diag(m) if kind == "diag" else m
                    self.states[b] = orc.apply_sv(self.states[b], m, tuple(qubits), self.n)

            def reduced(self, qubits):  # type: ignore[no-untyped-def]
                Batch.reductions += 1
                k = len(qubits)
                out = np.zeros((self.s, 1 << k, 1 << k), dtype=np.complex128)
                axes = [self.n - 1 - q for q in qubits]
                for b in range(self.s):
                    t = np.moveaxis(self.states[b].reshape([2] * self.n), axes, list(range(k))).reshape(1 << k, -1)
                    out[b] = t @ t.conj().T
                return out

            def copy_slots(self, src, dst) -> None:  # type: ignore[no-untyped-def]
                assert np.all(dst < self.s) and not set(src.tolist()) & set(dst.tolist())
                self.states[dst] = self.states[src]

            def finish(self) -> None:
                pass

        return Batch(n, nshots, 0, nflags, num_clbits, share)


def _host_trajectory_counts(circuit, noise, shots: int, seed: int, share: bool = True):  # type: ignore[no-untyped-def]
    from lucivant_sim import trajectories as tr

    steps, per_shot, nflags = tr.compile_program(circuit, noise)
    draws = np.random.default_rng(seed).random(shots * per_shot).reshape(shots, per_shot)
    batch = _HostBatch.make(circuit.num_qubits, shots, nflags, circuit.num_clbits, share)
    bits = tr._run_chunk(steps, circuit, draws, 0, nflags, batch)
    uniq, cnt = np.unique(bits, return_counts=True)
    return {circuit.format_bits(int(v)): int(c) for v, c in zip(uniq, cnt)}, batch


@pytest.mark.parametrize("which", ["dynamic", "teleport"])
def test_trajectory_batch_logic_equals_the_reference_rule(which: str) -> None:
    """The batch's host logic -- queued and fused operators, projections and Kraus operators
    applied unnormalised, reduced density matrices computed ahead over two qubits and carried on
    the host -- run over host states, gives the counts of the oracle's sequential per-shot loop
    (the reference's rule, pinned by the golden counts) for the same draws, and saves passes."""
    from dynamic_circuits import DYNAMIC, dynamic_noise
    from lucivant_sim.trajectories import compile_program

    if which == "dynamic":
        circuit = ls.parse(DYNAMIC)
        noise = dynamic_noise()
    else:
        circuit = ls.parse(wl.teleport_chain(9, 3))
        noise = wl.noise_c4()
    shots = 150
    got, batch = _host_trajectory_counts(circuit, noise, shots, 31)
    want = orc.trajectory_counts(circuit, noise, shots, np.random.default_rng(31))
    assert got == want
    ops = sum(1 for s in compile_program(circuit, noise)[0] if s.kind in ("gate", "measure", "reset"))
    assert batch.passes < ops
    # shots with equal histories share states: slots in use stay below the shot count, and the
    # counts equal those of one state per shot
    assert batch.s < shots, batch.s
    alone, _ = _host_trajectory_counts(circuit, noise, shots, 31, share=False)
    assert alone == got


def test_search_over_normalised_cumulative_sums_without_dividing_all() -> None:
    """The sampler's binary search divides only the entries it probes, and lands where
    searchsorted over the divided array does: with zero-probability plateaus, exact ties and draws
    equal to normalised entries."""
    from lucivant_sim.simulator import _search_normalised

    rng = np.random.default_rng(0)
    for n in (10, 17, 20):
        p = rng.random(1 << n) ** 4
        p[5:50] = 0.0
        p /= p.sum()
        cum = np.cumsum(p)
        normalised = cum / cum[-1]
        u = rng.random(3000)
        u[:100] = normalised[rng.integers(0, len(cum), 100)]
        u[100] = 0.0
        assert np.array_equal(_search_normalised(cum, u), np.searchsorted(normalised, u, side="right")), n


def test_recorded_outcome_distribution_equals_the_definition() -> None:
    """outcome_probabilities' readout expansion (one contraction per measured qubit) gives the
    oracle's brute-force definition: a qubit recorded in two clbits, clbits never written, qubits
    with and without readout error, p01 != p10, and marginal entries at the threshold dropped."""
    from lucivant_sim.simulator import _recorded_distribution

    rng = np.random.default_rng(12)
    measured = [0, 2, 3, 5]
    pairs = [(3, 0), (0, 2), (5, 1), (3, 4), (2, 6)]
    noise = ls.NoiseModel().add_readout_error(ls.ReadoutError(p01=0.05, p10=0.2), qubits=[3]).add_readout_error(ls.ReadoutError(p01=0.3, p10=0.01), qubits=[5])
    readout = {3: (0.05, 0.2), 5: (0.3, 0.01)}
    dist = rng.random(16) ** 3
    dist[[1, 6]] = 0.0
    dist[9] = 1e-13
    dist /= dist.sum()
    got = _recorded_distribution(dist, measured, pairs, noise, 1e-12)
    want = orc.recorded_distribution(dist, measured, pairs, readout, 1e-12)
    assert [b for b, _ in got] == list(want)
    assert np.allclose([w for _, w in got], list(want.values()), rtol=1e-13, atol=0)
    # no readout error: the recorded bits are the measured bits, moved to their clbits
    plain = _recorded_distribution(dist, measured, pairs, None, 1e-12)
    assert dict(plain) == pytest.approx(orc.recorded_distribution(dist, measured, pairs, {}, 1e-12), rel=1e-14)
    # a uniform symmetric error on every qubit, and a single measured qubit
    sym = ls.NoiseModel().add_readout_error(0.1)
    got1 = dict(_recorded_distribution(np.array([0.25, 0.75]), [4], [(4, 1)], sym, 1e-12))
    assert got1 == pytest.approx({0: 0.25 * 0.9 + 0.75 * 0.1, 2: 0.25 * 0.1 + 0.75 * 0.9}, rel=1e-15)


def test_repeated_programs_are_read_and_planned_once() -> None:
    """The same OpenQASM text with the same parameters reads to the same (immutable) Circuit, and
    a circuit is planned once per fusion setting; other text or parameters read anew, and only
    the most recent entries are kept."""
    from lucivant_sim import simulator as sim

    src = 'OPENQASM 3; include "stdgates.inc"; input float t; qubit[3] q; rx(t) q[0]; cx q[0], q[2];'
    a = sim.to_circuit(src, {"t": 0.5})
    assert sim.to_circuit(src, {"t": 0.5}) is a
    assert sim.to_circuit(src, {"t": 0.25}) is not a
    p = sim._statevector_plan(a, 5, 0, 0)
    assert sim._statevector_plan(a, 5, 0, 0) is p
    assert sim._statevector_plan(a, 4, 0, 0) is not p
    want = plan_statevector(a.operations, 3, 5)
    assert len(p) == len(want) and all(np.array_equal(x.matrix, y.matrix) and x.bits == y.bits for x, y in zip(p, want))
    for i in range(sim._RECENT + 2):
        sim.to_circuit(f"OPENQASM 3; qubit[{i + 1}] q; x q[0];")
    assert len(sim._READ) == sim._RECENT
    assert sim.to_circuit(src, {"t": 0.5}) is not a  # dropped as least recently used


def test_trajectory_batch_rescales_small_norms_per_slot(monkeypatch) -> None:  # type: ignore[no-untyped-def]
    """With the rescale threshold raised so that every decision rescales the states whose squared
    norm fell below it (one diagonal pass over the slots that need it), the host-state batch still
    gives the oracle's counts for the same draws."""
    from dynamic_circuits import DYNAMIC, dynamic_noise
    from lucivant_sim import trajectories as tr

    monkeypatch.setattr(tr, "_RENORMALISE_BELOW", {0: 0.999, 1: 0.999})
    circuit = ls.parse(DYNAMIC)
    noise = dynamic_noise()
    got, _ = _host_trajectory_counts(circuit, noise, 120, 8)
    assert got == orc.trajectory_counts(circuit, noise, 120, np.random.default_rng(8))


@pytest.mark.parametrize("budget", [6.0, 24.0, 1000.0])
def test_tile_passes_equal_sequential_gates(budget: float) -> None:
    """Applying every tile pass's operators, on the state bits its tile bits name, gives the
    circuit's state (oracle application, within 1e-12): random layers with cx/cz, a wide gate
    (run as its own dense pass) and a global phase. Every pass's tile bits include the lowest
    ones, and its operator cost respects the budget."""
    from lucivant_sim.plan import TILE_LOW, TilePass, _tile_cost, plan_tiles

    n = 14
    src = wl.random_layers(n, 6, 91, two_qubit=("cx", "cz"), measure=False)
    src += "ccx q[0], q[7], q[13];\nh q[13];\ngphase(0.3);\ncx q[13], q[2];\n"
    c = ls.parse(src)
    want = orc.evolve_sv(c)
    plan = plan_tiles(c.operations, n, budget=budget)
    psi = np.zeros(2**n, dtype=np.complex128)
    psi[0] = 1
    gates = 0
    for p in plan:
        if isinstance(p, TilePass):
            assert len(p.bits) == 12 and p.bits[:TILE_LOW] == tuple(range(TILE_LOW)) and list(p.bits) == sorted(p.bits)
            assert sum(map(_tile_cost, p.ops)) <= budget or len(p.ops) == 1
            for op in p.ops:
                qubits = tuple(p.bits[b] for b in op.bits)
                m = np.diag(op.matrix) if op.kind == "diag" else op.matrix
                psi = orc.apply_sv(psi, m, qubits, n)
            gates += p.gates
        else:
            m = np.diag(p.matrix) if p.kind == "diag" else p.matrix
            psi = orc.apply_sv(psi, m, p.bits, n) if p.kind != "scale" else psi * p.matrix[0]
            gates += p.gates
    assert np.abs(psi - want).max() < 1e-12
    assert gates == sum(1 for op in c.operations if isinstance(op, ls.circuit.Gate))


def test_permutation_passes_are_recognised() -> None:
    """Dense passes whose matrix has one non-zero per row and column run as permutations with
    phases (lsim_sv_apply_perm); anything else stays dense."""
    import numpy as np

    from lucivant_sim.plan import Pass

    cx = np.eye(4, dtype=np.complex128)[[0, 1, 3, 2]]
    mv = Pass("dense", (1, 0), cx).moves()
    assert mv is not None and mv[0] == (0, 1, 3, 2) and mv[2] == 0b1111
    y = np.array([[0, -1j], [1j, 0]])
    src, entries, unit = Pass("dense", (3,), y).moves()
    assert src == (1, 0) and unit == 0 and np.allclose(entries, [-1j, 1j])
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    assert Pass("dense", (0,), h).moves() is None
    assert Pass("dense", (0,), np.array([[1, 0], [0, 0]], dtype=np.complex128)).moves() is None
    assert Pass("diag", (0,), np.array([1, 1j])).moves() is None
    hp = Pass("dense", (0,), h)
    assert hp.payload() is hp.matrix
    assert np.array_equal(Pass("dense", (1, 0), cx).payload(), np.ones(4))


def _same_passes(got: list, want: list) -> None:
    from lucivant_sim.plan import TilePass

    assert len(got) == len(want)
    for x, y in zip(got, want):
        assert type(x) is type(y)
        if isinstance(x, TilePass):
            assert x.bits == y.bits and x.gates == y.gates
            _same_passes(x.ops, y.ops)
        else:
            assert (x.kind, x.bits, x.gates) == (y.kind, y.bits, y.gates)
            assert np.abs(np.asarray(x.matrix) - np.asarray(y.matrix)).max() < 1e-13


@pytest.mark.parametrize("seed", range(4))
def test_native_fusion_equals_the_python_steps(seed: int) -> None:
    """lsim_host_plan_fused forms the passes that the plan module's Python steps (one-qubit runs,
    greedy, merge, descending targets) form -- the same blocks, targets, kinds and gate counts,
    matrices within 1e-13 -- at every width, low-qubit cap and with and without the merge, on
    circuits with swap, iswap, a three-qubit gate and a global phase."""
    from plan_reference import plan_statevector_reference

    src = wl.random_layers(9, 8, 100 + seed, two_qubit=("cx", "cz", "swap", "iswap"), measure=True)
    src = src.replace("c = measure q;", "gphase(0.3); ccx q[0], q[3], q[6]; c = measure q;")
    c = ls.parse(src)
    for width in (1, 2, 3, 4, 5):
        for low_width in (0, 2, 3):
            for merge in (True, False):
                got = plan_statevector(c.operations, 9, width, 3, low_width, merge)
                _same_passes(got, plan_statevector_reference(c.operations, 9, width, 3, low_width, merge))


@pytest.mark.parametrize("budget", [6.0, 24.0, 48.0])
def test_native_tile_plan_equals_the_python_planner(budget: float) -> None:
    """lsim_host_plan_tiles chooses the tile bits, the gates and the operators that the Python
    tile planner (counting runnable gates by its rule) chooses: 14 qubits with a wide gate and a
    global phase, a 20-qubit ansatz and a 22-qubit random circuit."""
    from lucivant_sim.plan import plan_tiles
    from plan_reference import plan_tiles_reference

    cases = [
        (14, wl.random_layers(14, 10, 5, two_qubit=("cx", "cz"), measure=False) + "ccx q[0], q[7], q[13];\nh q[13];\ngphase(0.3);\ncx q[13], q[2];\n"),
        (20, wl.hea_ansatz(20, 12, 3)),
        (22, wl.random_layers(22, 12, 9)),
    ]
    for n, src in cases:
        c = ls.parse(src)
        _same_passes(plan_tiles(c.operations, n, budget=budget), plan_tiles_reference(c.operations, n, budget=budget))


@pytest.mark.parametrize("seed", range(3))
def test_native_plans_of_random_gates_equal_sequential_gates(seed: int) -> None:
    """Random gates on 1-5 qubits (random unitaries, diagonals and permutations among them) on 12
    qubits: the fused passes at every width and the tile passes at two budgets apply, in order,
    to the state the gates give one by one (oracle application, within 1e-12); every gate is
    counted once."""
    from lucivant_sim.circuit import Gate
    from lucivant_sim.plan import TilePass, plan_tiles

    rng = np.random.default_rng(600 + seed)
    n = 12
    ops = []
    for _ in range(160):
        k = int(rng.choice([1, 1, 2, 2, 2, 3, 4, 5]))
        qs = tuple(int(q) for q in rng.choice(n, size=k, replace=False))
        d = 1 << k
        kind = rng.integers(3)
        if kind == 0:
            m = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))[0]
        elif kind == 1:
            m = np.diag(np.exp(1j * rng.uniform(0, 6.3, d)))
        else:
            m = np.eye(d, dtype=np.complex128)[rng.permutation(d)] * np.exp(1j * rng.uniform(0, 6.3, d))[:, None]
        ops.append(Gate("unitary", qs, np.ascontiguousarray(m, dtype=np.complex128)))
    psi0 = np.zeros(1 << n, dtype=np.complex128)
    psi0[0] = 1
    want = psi0
    for op in ops:
        want = orc.apply_sv(want, op.matrix, op.qubits, n)
    for width in (1, 2, 3, 4, 5):
        passes = plan_statevector(tuple(ops), n, width)
        assert sum(p.gates for p in passes) == len(ops)
        assert np.abs(_run_passes_on_cpu(psi0, passes, n) - want).max() < 1e-12
    for budget in (6.0, 48.0):
        psi = psi0
        gates = 0
        for p in plan_tiles(tuple(ops), n, budget=budget):
            gates += p.gates
            if isinstance(p, TilePass):
                for op in p.ops:
                    m = np.diag(op.matrix) if op.kind == "diag" else op.matrix
                    psi = orc.apply_sv(psi, m, tuple(p.bits[b] for b in op.bits), n)
            else:
                psi = _run_passes_on_cpu(psi, [p], n)
        assert gates == len(ops)
        assert np.abs(psi - want).max() < 1e-12


def test_native_planners_refuse_bad_input() -> None:
    """Out-of-range widths, qubits and tile shapes come back as errors with the library's
    message, not as plans."""
    from lucivant_sim import _native as nat
    from lucivant_sim.plan import plan_tiles

    c = ls.parse(wl.random_layers(14, 2, 1))
    with pytest.raises(ValueError):
        plan_statevector(c.operations, 14, 6)
    with pytest.raises(nat.NativeError, match="low bits"):
        plan_tiles(c.operations, 14, low=11)
    with pytest.raises(ValueError):
        plan_tiles(c.operations, 14, tile_bits=15)
