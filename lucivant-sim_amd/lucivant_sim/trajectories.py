"""Trajectories: every shot of a noisy or dynamic circuit, run as one batch of statevectors.

The reference runs one fresh statevector per shot and walks the circuit in Python, drawing from
one seeded generator in program order (simulator.py:287-317, 364-371): one draw per gate that has
a noise channel, one per measurement plus one more when the qubit has a readout error, one per
reset; an if/else contributes the draws of the arm it takes (engines.py:159-205).

When every if/else has arms with the same number of draws, each shot consumes a constant number D
of draws and shot k reads stream positions [kD, (k+1)D). This module then draws
``rng.random(shots * D)`` once (the same stream as the scalar draws, in the same order) and runs
all shots together in a batch of states ``[B][2^n]`` in HBM: ideal gates as fused passes shared by
every shot; noise branches, collapses and conditioned gates as per-shot operators selected on the
host from the draws and from per-shot reduced density matrices, then applied by one batched
launch. Each decision uses the reference's float64 rule (cumulative sums in the same order), so
counts equal the reference's except where a draw falls within rounding of a branch boundary.

Circuits whose arms draw differently are run shot after shot with the same device engine
(:func:`run_sequential`), which reproduces the reference's draw order exactly.
"""

from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np
import torch

from lucivant_sim import _native as nat
from lucivant_sim.circuit import Barrier, Circuit, Condition, Gate, GlobalPhase, IfElse, Measure, Operation, Reset
from lucivant_sim.engines import WORKSPACE, Statevector, _NP_DTYPE, _TORCH_DTYPE, ptr, require_device, run_passes, stream_handle, upload
from lucivant_sim.noise import Channel, NoiseModel, ReadoutError
from lucivant_sim.plan import plan_statevector

#: HBM the batch of trajectory states may take (bytes); shots beyond it run in further chunks.
BATCH_BYTES = int(os.environ.get("LSIM_TRAJECTORY_BYTES", str(64 << 30)))


# ------------------------------------------------------------------------------- draw ledger


def draw_count(ops: tuple[Operation, ...], noise: NoiseModel | None) -> int | None:
    """Draws one pass over ``ops`` consumes, or None when an if/else's arms differ."""
    total = 0
    for op in ops:
        if isinstance(op, Gate):
            if noise is not None and noise.gate_error(op.name, op.qubits) is not None:
                total += 1
        elif isinstance(op, Measure):
            total += 1
            if noise is not None and noise.readout_error(op.qubit) is not None:
                total += 1
        elif isinstance(op, Reset):
            total += 1
        elif isinstance(op, IfElse):
            a, b = draw_count(op.then, noise), draw_count(op.otherwise, noise)
            if a is None or b is None or a != b:
                return None
            total += a
    return total


# ------------------------------------------------------------------------------- the program


@dataclass
class _Step:
    kind: str  # gate, channel, measure, reset, phase, cond
    guard: tuple[tuple[int, bool], ...]
    qubits: tuple[int, ...] = ()
    matrix: np.ndarray | None = None
    channel: Channel | None = None
    column: int = -1  # draw column
    clbit: int | None = None
    readout: ReadoutError | None = None
    condition: Condition | None = None
    flag: int = -1
    angle: float = 0.0


def compile_program(circuit: Circuit, noise: NoiseModel | None) -> tuple[list[_Step], int, int]:
    """Linear steps (if/else flattened into guarded steps), draws per shot, number of flags."""
    steps: list[_Step] = []
    nflags = 0

    def walk(ops: tuple[Operation, ...], guard: tuple[tuple[int, bool], ...], column: int) -> int:
        nonlocal nflags
        for op in ops:
            if isinstance(op, Gate):
                steps.append(_Step("gate", guard, tuple(op.qubits), np.asarray(op.matrix, dtype=np.complex128)))
                ch = noise.gate_error(op.name, op.qubits) if noise is not None else None
                if ch is not None:
                    steps.append(_Step("channel", guard, tuple(op.qubits), channel=ch, column=column))
                    column += 1
            elif isinstance(op, Measure):
                err = noise.readout_error(op.qubit) if noise is not None else None
                steps.append(_Step("measure", guard, (op.qubit,), column=column, clbit=op.clbit, readout=err))
                column += 2 if err is not None else 1
            elif isinstance(op, Reset):
                steps.append(_Step("reset", guard, (op.qubit,), column=column))
                column += 1
            elif isinstance(op, GlobalPhase):
                steps.append(_Step("phase", guard, angle=op.angle))
            elif isinstance(op, IfElse):
                flag = nflags
                nflags += 1
                steps.append(_Step("cond", guard, condition=op.condition, flag=flag))
                end_then = walk(op.then, guard + ((flag, True),), column)
                end_else = walk(op.otherwise, guard + ((flag, False),), column)
                if end_then != end_else:
                    raise ValueError("if/else arms consume different numbers of draws")
                column = end_then
            elif isinstance(op, Barrier):
                continue
        return column

    d = walk(circuit.operations, (), 0)
    return steps, d, nflags


# ------------------------------------------------------------------------------- execution


class _Batch:
    """B states of n qubits in HBM plus their classical bits and if/else flags."""

    def __init__(self, n: int, batch: int, dtype: int, nflags: int, num_clbits: int) -> None:
        self.n, self.b, self.dtype = n, batch, dtype
        self.buf = torch.empty(batch << n, dtype=_TORCH_DTYPE[dtype], device=require_device())
        nat.call("lsim_sv_set_basis", ptr(self.buf), dtype, n, batch, 0, stream_handle())
        self.wide = num_clbits > 62
        self.clbits = np.zeros(batch, dtype=object if self.wide else np.int64)
        self.flags = np.zeros((max(nflags, 1), batch), dtype=bool)
        self.pending: list[Gate] = []

    def active(self, guard: tuple[tuple[int, bool], ...]) -> np.ndarray | None:
        if not guard:
            return None
        mask = np.ones(self.b, dtype=bool)
        for flag, polarity in guard:
            mask &= self.flags[flag] == polarity
        return mask

    def flush(self) -> None:
        if self.pending:
            passes = plan_statevector(tuple(self.pending), self.n, 4)
            run_passes(self.buf, self.dtype, self.n, self.b, passes)
            self.pending = []

    def apply_selected(self, kind: str, qubits: tuple[int, ...], table: np.ndarray, sel: np.ndarray) -> None:
        """Per-state operator: state b gets table[sel[b]] (dense matrices or diagonals), or is
        left alone when sel[b] < 0."""
        if not np.any(sel >= 0):
            return
        dev, (addr,) = upload([table], self.dtype)
        sel_dev = torch.from_numpy(np.ascontiguousarray(sel, dtype=np.int32)).to(self.buf.device)
        fn = "lsim_sv_apply_dense" if kind == "dense" else "lsim_sv_apply_diag"
        nat.call(fn, ptr(self.buf), self.dtype, self.n, self.b, len(qubits), nat.int_array(qubits), addr, ptr(sel_dev), stream_handle())
        dev.record_stream(torch.cuda.current_stream())
        sel_dev.record_stream(torch.cuda.current_stream())

    def reduced(self, qubits: tuple[int, ...]) -> np.ndarray:
        """Per-state reduced density matrices [B][2^k][2^k] on ``qubits`` (float64 sums)."""
        k = len(qubits)
        out = torch.empty(self.b << (2 * k), dtype=torch.complex128, device=self.buf.device)
        ws = WORKSPACE.get(nat.load().lsim_reduce_workspace_bytes(self.n, self.b, k))
        nat.call(
            "lsim_sv_reduced_dm", ptr(self.buf), self.dtype, self.n, self.b, k, nat.int_array(qubits), ptr(out), ptr(ws), stream_handle()
        )
        return out.cpu().numpy().reshape(self.b, 1 << k, 1 << k)

    def set_bits(self, clbit: int, outcome: np.ndarray, mask: np.ndarray | None) -> None:
        if self.wide:
            for i in range(self.b):
                if mask is None or mask[i]:
                    v = int(self.clbits[i])
                    self.clbits[i] = (v & ~(1 << clbit)) | (int(outcome[i]) << clbit)
            return
        cleared = self.clbits & ~np.int64(1 << clbit)
        updated = cleared | (outcome.astype(np.int64) << clbit)
        self.clbits = updated if mask is None else np.where(mask, updated, self.clbits)

    def clbit_values(self, clbits: tuple[int, ...]) -> np.ndarray:
        actual = np.zeros(self.b, dtype=object if self.wide else np.int64)
        for i, c in enumerate(clbits):
            if self.wide:
                actual = actual + np.array([((int(v) >> c) & 1) << i for v in self.clbits], dtype=object)
            else:
                actual |= ((self.clbits >> c) & 1) << i
        return actual


def _mixture_choice(channel: Channel, r: np.ndarray) -> tuple[np.ndarray, list[np.ndarray]]:
    """Branch index per shot into the non-identity unitaries (-1: identity or no branch), with
    the cumulative sums formed in the reference's order (engines.py:162-170)."""
    assert channel.mixture is not None
    cumulative, table, index = [], [], []
    acc = 0.0
    for p, unitary in channel.mixture:
        acc += p
        cumulative.append(acc)
        if np.allclose(unitary, np.eye(unitary.shape[0])):
            index.append(-1)
        else:
            index.append(len(table))
            table.append(unitary)
    branch = np.searchsorted(np.array(cumulative), r, side="right")
    lookup = np.array(index + [-1], dtype=np.int64)
    return lookup[branch], table


def _kraus_choice(channel: Channel, rho: np.ndarray, r: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Per-shot normalised Kraus operator (engines.py:171-183): the first i with r < sum of
    weights so far and w_i > 0, else the last one if its weight is positive, else none."""
    nshots = r.shape[0]
    d = channel.kraus[0].shape[0]
    mats = np.zeros((nshots, d, d), dtype=np.complex128)
    sel = np.full(nshots, -1, dtype=np.int64)
    chosen = np.zeros(nshots, dtype=bool)
    acc = np.zeros(nshots)
    weight = np.zeros(nshots)
    for kraus in channel.kraus:
        weight = np.einsum("ra,sac,rc->s", kraus, rho, kraus.conj()).real
        acc = acc + weight
        pick = (~chosen) & (r < acc) & (weight > 0)
        if np.any(pick):
            mats[pick] = kraus[None] / np.sqrt(weight[pick])[:, None, None]
            sel[pick] = np.flatnonzero(pick)
            chosen |= pick
    rest = (~chosen) & (weight > 0)
    if np.any(rest):
        last = channel.kraus[-1]
        mats[rest] = last[None] / np.sqrt(weight[rest])[:, None, None]
        sel[rest] = np.flatnonzero(rest)
    return mats, sel


def _collapse_factors(p1: np.ndarray, outcome: np.ndarray) -> np.ndarray:
    """1/sqrt(p) of the observed branch (1 where p = 0, as engines.py:195-201 skips the divide)."""
    p = np.where(outcome == 1, p1, 1 - p1)
    safe = np.where(p > 0, p, 1.0)
    return np.where(p > 0, 1.0 / np.sqrt(safe), 1.0)


def _run_chunk(steps: list[_Step], circuit: Circuit, draws: np.ndarray, dtype: int, nflags: int) -> np.ndarray:
    nshots = draws.shape[0]
    st = _Batch(circuit.num_qubits, nshots, dtype, nflags, circuit.num_clbits)
    everyone = np.arange(nshots)
    for step in steps:
        mask = st.active(step.guard)
        if step.kind == "gate" and mask is None:
            st.pending.append(Gate("fused", step.qubits, step.matrix))  # type: ignore[arg-type]
            continue
        st.flush()
        if step.kind == "gate":
            st.apply_selected("dense", step.qubits, step.matrix[None], np.where(mask, 0, -1))  # type: ignore[index]
        elif step.kind == "phase":
            z = np.exp(1j * step.angle)
            d = np.array([[z, z]], dtype=np.complex128)
            sel = np.zeros(nshots, dtype=np.int64) if mask is None else np.where(mask, 0, -1)
            if circuit.num_qubits:
                st.apply_selected("diag", (0,), d, sel)
        elif step.kind == "channel":
            ch = step.channel
            assert ch is not None
            r = draws[:, step.column]
            if ch.mixture is not None:
                sel, table = _mixture_choice(ch, r)
                if mask is not None:
                    sel = np.where(mask, sel, -1)
                if table:
                    st.apply_selected("dense", step.qubits, np.stack(table), sel)
            else:
                mats, sel = _kraus_choice(ch, st.reduced(step.qubits), r)
                if mask is not None:
                    sel = np.where(mask, sel, -1)
                st.apply_selected("dense", step.qubits, mats, sel)
        elif step.kind in ("measure", "reset"):
            q = step.qubits[0]
            rho = st.reduced((q,))
            p1 = np.minimum(np.maximum(rho[:, 1, 1].real, 0.0), 1.0)
            outcome = (draws[:, step.column] < p1).astype(np.int64)
            scale = _collapse_factors(p1, outcome)
            sel = everyone if mask is None else np.where(mask, everyone, -1)
            if step.kind == "measure":
                diags = np.zeros((nshots, 2), dtype=np.complex128)
                diags[everyone, outcome] = scale
                st.apply_selected("diag", (q,), diags, sel)
                if step.readout is not None:
                    err = step.readout
                    flip = draws[:, step.column + 1] < np.where(outcome == 1, err.p10, err.p01)
                    outcome = outcome ^ flip.astype(np.int64)
                if step.clbit is not None:
                    st.set_bits(step.clbit, outcome, mask)
            else:
                mats = np.zeros((nshots, 2, 2), dtype=np.complex128)
                mats[:, 0, 0] = np.where(outcome == 0, scale, 0)
                mats[:, 0, 1] = np.where(outcome == 1, scale, 0)
                st.apply_selected("dense", (q,), mats, sel)
        elif step.kind == "cond":
            c = step.condition
            assert c is not None
            holds = (st.clbit_values(c.clbits) == c.value) != c.negate
            st.flags[step.flag] = holds.astype(bool)
    st.flush()
    torch.cuda.current_stream().synchronize()
    return st.clbits


def run_batched(
    circuit: Circuit,
    noise: NoiseModel | None,
    shots: int,
    rng: np.random.Generator,
    *,
    dtype: int = nat.C128,
    batch_bytes: int | None = None,
) -> np.ndarray:
    """Classical bits of every shot (int64, or Python ints beyond 62 clbits), all shots batched.
    Requires draw-balanced if/else arms (see :func:`draw_count`)."""
    steps, per_shot, nflags = compile_program(circuit, noise)
    item = np.dtype(_NP_DTYPE[dtype]).itemsize
    limit = batch_bytes if batch_bytes is not None else BATCH_BYTES
    chunk = max(1, min(shots, limit // (item << circuit.num_qubits)))
    out = []
    for start in range(0, shots, chunk):
        b = min(chunk, shots - start)
        draws = rng.random(b * per_shot).reshape(b, per_shot) if per_shot else np.zeros((b, 0))
        out.append(_run_chunk(steps, circuit, draws, dtype, nflags))
    return np.concatenate(out) if out else np.zeros(0, dtype=np.int64)


#: Shots per program launch (bounds the draw buffer and output sizes of one launch).
LDS_CHUNK = 1 << 16


def run_lds(
    circuit: Circuit, noise: NoiseModel | None, shots: int, rng: np.random.Generator, *, dtype: int = nat.C128
) -> np.ndarray | None:
    """All shots as LDS programs, one workgroup per shot, when the state fits a workgroup's LDS
    and the circuit fits the program kernel; None otherwise (the caller uses :func:`run_batched`).
    Draws are taken from ``rng`` exactly as :func:`run_batched` takes them."""
    from lucivant_sim import program as pg

    if circuit.num_qubits < 1 or circuit.num_qubits > pg.max_qubits(dtype) or circuit.num_clbits > 62:
        return None
    try:
        prog = pg.compile_trajectory(circuit, noise)
    except pg.Unsupported:
        return None
    out = []
    for start in range(0, shots, LDS_CHUNK):
        b = min(LDS_CHUNK, shots - start)
        draws = rng.random(b * prog.draws).reshape(b, prog.draws) if prog.draws else None
        res = pg.run_programs([prog], repeat=b, dtype=dtype, draws=draws, want_clbits=True)
        assert res.clbits is not None
        out.append(res.clbits.astype(np.int64))
    return np.concatenate(out)


def run_sequential(circuit: Circuit, noise: NoiseModel | None, shots: int, rng: np.random.Generator) -> list[int]:
    """One device statevector per shot, in the reference's exact draw order (any circuit)."""
    results = []
    for _ in range(shots):
        state = Statevector(circuit.num_qubits)
        clbits = [0] * circuit.num_clbits

        def run_ops(ops: tuple[Operation, ...]) -> None:
            for op in ops:
                if isinstance(op, Gate):
                    state.apply(op.matrix, op.qubits)
                    channel = noise.gate_error(op.name, op.qubits) if noise else None
                    if channel is not None:
                        state.apply_channel(channel, op.qubits, rng)
                elif isinstance(op, Measure):
                    outcome = state.measure(op.qubit, rng)
                    err = noise.readout_error(op.qubit) if noise else None
                    if err is not None and rng.random() < (err.p10 if outcome else err.p01):
                        outcome ^= 1
                    if op.clbit is not None:
                        clbits[op.clbit] = outcome
                elif isinstance(op, Reset):
                    state.reset(op.qubit, rng)
                elif isinstance(op, GlobalPhase):
                    state.scale(np.exp(1j * op.angle))
                elif isinstance(op, IfElse):
                    run_ops(op.then if op.condition.holds(clbits) else op.otherwise)

        run_ops(circuit.operations)
        results.append(sum(b << i for i, b in enumerate(clbits)))
    return results


def shot_values(
    circuit: Circuit, noise: NoiseModel | None, shots: int, rng: np.random.Generator, *, dtype: int = nat.C128
) -> tuple[list[int] | np.ndarray, int]:
    """The classical bit pattern of every shot, in shot order, and the precision it was computed
    in (circuits whose if/else arms draw unequally run shot after shot in complex128)."""
    if shots == 0:
        return [], dtype
    if draw_count(circuit.operations, noise) is None:
        return run_sequential(circuit, noise, shots, rng), nat.C128
    values = run_lds(circuit, noise, shots, rng, dtype=dtype)
    if values is None:
        values = run_batched(circuit, noise, shots, rng, dtype=dtype)
    return values, dtype


def run_trajectories(
    circuit: Circuit, noise: NoiseModel | None, shots: int, rng: np.random.Generator, *, dtype: int = nat.C128
) -> dict[int, int]:
    """Tally of classical bit patterns over ``shots`` trajectories."""
    values, _ = shot_values(circuit, noise, shots, rng, dtype=dtype)
    if isinstance(values, list):
        tally: dict[int, int] = {}
        for v in values:
            tally[v] = tally.get(v, 0) + 1
        return tally
    uniq, counts = np.unique(values, return_counts=True)
    return {int(v): int(c) for v, c in zip(uniq, counts)}


#: shots of a complex64 run that are rerun in complex128 to state its error
PRECISION_CHECK_SHOTS = 256


def complex64_check(circuit: Circuit, noise: NoiseModel | None, seed: int, values: list[int] | np.ndarray, used: int) -> dict[str, object]:
    """How a complex64 trajectory run compares with complex128: its first shots are rerun in
    complex128 with the same draws (the same seeded stream) and compared shot by shot. The shots
    are deterministic given their draws, so any difference is the precision's doing."""
    if used != nat.C64:
        return {"dtype": "complex128", "note": "this circuit runs shot after shot in complex128"}
    m = min(len(values), PRECISION_CHECK_SHOTS)
    ref, _ = shot_values(circuit, noise, m, np.random.default_rng(seed), dtype=nat.C128)
    differing = int(np.sum(np.asarray(values[:m], dtype=np.int64) != np.asarray(ref, dtype=np.int64)))
    return {"dtype": "complex64", "checked_shots": m, "shots_differing_from_complex128": differing}


__all__ = ["complex64_check", "compile_program", "draw_count", "run_batched", "run_sequential", "run_trajectories", "shot_values"]
