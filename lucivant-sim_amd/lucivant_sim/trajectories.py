"""Trajectories: every shot of a noisy or dynamic circuit, run as one batch of statevectors.

The reference runs one fresh statevector per shot and walks the circuit in Python, drawing from
one seeded generator in program order (simulator.py:287-317, 364-371): one draw per gate that has
a noise channel, one per measurement plus one more when the qubit has a readout error, one per
reset; an if/else contributes the draws of the arm it takes (engines.py:159-205).

When every if/else has arms with the same number of draws, each shot consumes a constant number D
of draws and shot k reads stream positions [kD, (k+1)D). This module then draws
``rng.random(shots * D)`` once (the same stream as the scalar draws, in the same order) and runs
all shots together in a batch of states ``[B][2^n]`` in HBM (:class:`_Batch`), where shots with
equal histories share one state: gates, noise branches, projections and conditioned gates are
queued as per-shot choices and fused into shared passes, and decisions read per-shot reduced
density matrices, computed on the device or carried on the host. Each decision uses the reference's float64 rule (cumulative sums in the same order),
so counts equal the reference's except where a draw falls within rounding of a branch boundary.

Circuits whose arms draw differently are run shot after shot with the same device engine
(:func:`run_sequential`), which reproduces the reference's draw order exactly.
"""

from __future__ import annotations

import os
import time
from dataclasses import dataclass

import numpy as np
import torch

from lucivant_sim import _native as nat
from lucivant_sim.circuit import Barrier, Circuit, Condition, Gate, GlobalPhase, IfElse, Measure, Operation, Reset
from lucivant_sim.engines import WORKSPACE, Statevector, _NP_DTYPE, _TORCH_DTYPE, ptr, require_device, run_passes, stream_handle, to_device, upload
from lucivant_sim.noise import Channel, NoiseModel, ReadoutError
from lucivant_sim.plan import Pass, _as_pass, embed

#: Counters of the last :func:`run_batched`: host round trips (reduced density matrices read
#: back), seconds the host waited in them for the device, and launches over the batch.
STATS: dict[str, float] = {"syncs": 0, "wait_seconds": 0.0, "passes": 0}

#: HBM the batch of trajectory states may take (bytes); shots beyond it run in further chunks.
BATCH_BYTES = int(os.environ.get("LSIM_TRAJECTORY_BYTES", str(64 << 30)))

#: Widest block of queued operators fused into one pass over the batch (the vector kernel's
#: scalar-operand limit; a per-shot operator table then holds 2^4 x 2^4 matrices).
FUSE_WIDTH = 4


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


@dataclass
class _Queued:
    """An operator waiting to be fused: ``table[idx[b]]`` for shot b (nothing where idx[b] < 0),
    or ``table[0]`` for every shot when ``idx`` is None. Projections and Kraus operators are
    queued unnormalised (``unitary`` False)."""

    qubits: tuple[int, ...]
    table: list[np.ndarray]
    idx: np.ndarray | None = None
    unitary: bool = True


def _group(items: list[_Queued], width: int) -> list[tuple[list[int], list[_Queued]]]:
    """Consecutive queued operators grouped into blocks of at most ``width`` qubits: each joins
    the latest block that touches one of its qubits when the union stays within ``width`` (blocks
    that touch none of its qubits commute with it), else it opens a new block."""
    blocks: list[tuple[list[int], list[_Queued]]] = []
    last: dict[int, int] = {}
    for it in items:
        j = max((last.get(q, -1) for q in it.qubits), default=-1)
        if j >= 0 and len(set(blocks[j][0]) | set(it.qubits)) <= width:
            blocks[j][0].extend(q for q in it.qubits if q not in blocks[j][0])
            blocks[j][1].append(it)
        else:
            blocks.append((list(it.qubits), [it]))
            j = len(blocks) - 1
        for q in it.qubits:
            last[q] = j
    return blocks


def _merge_groups(blocks: list[tuple[list[int], list[_Queued]]], width: int) -> list[tuple[list[int], list[_Queued]]]:
    """Fewer blocks, as :func:`plan._merge` does for fused passes: two blocks join when their
    union fits and no block between them shares a qubit with the one that moves (forward: an
    earlier block moves to a later one; backward: a later one moves to an earlier one)."""
    while True:
        before = len(blocks)
        for backward in (False, True):
            alive = [True] * len(blocks)
            order = range(len(blocks) - 1, -1, -1) if backward else range(len(blocks))
            for j in order:
                if not alive[j]:
                    continue
                mine = set(blocks[j][0])
                others = range(j - 1, -1, -1) if backward else range(j + 1, len(blocks))
                for j2 in others:
                    if not alive[j2]:
                        continue
                    other = set(blocks[j2][0])
                    if len(mine | other) <= width:
                        first, second = (blocks[j2], blocks[j]) if backward else (blocks[j], blocks[j2])
                        blocks[j2] = (first[0] + [q for q in second[0] if q not in first[0]], first[1] + second[1])
                        alive[j] = False
                        break
                    if mine & other:
                        break
            blocks = [b for b, a in zip(blocks, alive) if a]
        if len(blocks) == before:
            return blocks


def _fuse(qubits: list[int], items: list[_Queued], combos: np.ndarray) -> np.ndarray:
    """The block's operator for each combination of per-shot choices (a row of ``combos`` holds
    one table index per per-shot item, in order; -1 is the identity): the product of the items in
    program order, embedded on ``qubits`` (first = most significant matrix bit)."""
    k = len(qubits)
    d = 1 << k
    out = np.broadcast_to(np.eye(d, dtype=np.complex128), (len(combos), d, d)).copy()
    column = 0
    for it in items:
        pos = tuple(qubits.index(q) for q in it.qubits)
        if it.idx is None:
            out = embed(it.table[0], list(pos), k) @ out
            continue
        out = _choices(it.table, pos, k)[combos[:, column] + 1] @ out
        column += 1
    return out


_CHOICES: dict[tuple, tuple[list[np.ndarray], np.ndarray]] = {}


def _choices(table: list[np.ndarray], pos: tuple[int, ...], k: int) -> np.ndarray:
    """The identity followed by each table entry embedded at ``pos`` of a k-position operator.
    Tables of channel branches recur from block to block, so the stack is kept per (table, pos,
    k) while the table object lives."""
    key = (id(table), pos, k)
    hit = _CHOICES.get(key)
    if hit is not None and hit[0] is table:
        return hit[1]
    d = 1 << k
    stack = np.stack([np.eye(d, dtype=np.complex128)] + [embed(m, list(pos), k) for m in table])
    if len(_CHOICES) > 512:
        _CHOICES.clear()
    _CHOICES[key] = (table, stack)
    return stack


def _combinations(columns: list[np.ndarray]) -> tuple[np.ndarray, np.ndarray]:
    """Distinct rows of the per-shot choice columns (each in -1..T-1) and each shot's row index:
    the rows are packed into one integer per shot (mixed radix) when that fits, which is far
    faster than sorting rows."""
    keys = np.stack(columns, axis=1)
    radix = [int(c.max()) + 2 for c in columns]
    if float(np.prod(np.array(radix, dtype=np.float64))) < 2.0**62:
        code = np.zeros(keys.shape[0], dtype=np.int64)
        for c, r in zip(columns, radix):
            code = code * r + (c.astype(np.int64) + 1)
        _, first, inverse = np.unique(code, return_index=True, return_inverse=True)
        return keys[first], np.asarray(inverse, dtype=np.int64).reshape(-1)
    combos, inverse = np.unique(keys, axis=0, return_inverse=True)
    return combos, np.asarray(inverse, dtype=np.int64).reshape(-1)


#: Host seconds per (combination x operator) of a block fused on 2^k x 2^k matrices, per d^3.
_FUSE_SECONDS_PER_D3 = 0.7e-9


def _split_for_host(items: list[_Queued], k: int, pass_seconds: float) -> list[list[_Queued]]:
    """A block's operators in consecutive parts, each one pass, so that fusing a part on the host
    costs no more than a pass on the device: with many per-shot choices (strong noise) every shot
    may need its own fused operator, and one more pass is then cheaper than the host products."""
    d3 = float(1 << (3 * k))
    budget = max(pass_seconds, 2e-4)
    parts: list[list[_Queued]] = []
    part: list[_Queued] = []
    # each shot's combination of the part's choices so far, as an index into the distinct ones
    # (extended one operator at a time: the pairs (combination, new choice) are the new ones)
    inverse: np.ndarray | None = None
    for it in items:
        grown, count = _extend(inverse, it.idx)
        if part and count * (len(part) + 1) * d3 * _FUSE_SECONDS_PER_D3 > budget:
            parts.append(part)
            part = []
            grown, count = _extend(None, it.idx)
        part.append(it)
        inverse = grown
    if part:
        parts.append(part)
    return parts


def _extend(inverse: np.ndarray | None, idx: np.ndarray | None) -> tuple[np.ndarray | None, int]:
    """Distinct combinations after one more per-shot choice column (None: a shared operator):
    each shot's new combination index and the number of combinations."""
    if idx is None:
        return inverse, 1 if inverse is None else int(inverse.max()) + 1
    if inverse is None:
        uniq, inv = np.unique(idx, return_inverse=True)
        return inv.reshape(-1), len(uniq)
    code = inverse.astype(np.int64) * (int(idx.max()) + 2) + (idx.astype(np.int64) + 1)
    uniq, inv = np.unique(code, return_inverse=True)
    return inv.reshape(-1), len(uniq)


def _descending_many(mats: np.ndarray, bits: tuple[int, ...]) -> tuple[np.ndarray, tuple[int, ...]]:
    """:func:`plan.descending` for a stack of operators on the same targets."""
    k = len(bits)
    order = sorted(range(k), key=lambda i: -bits[i])
    if order == list(range(k)):
        return np.ascontiguousarray(mats), bits
    idx = np.arange(1 << k)
    old = np.zeros(1 << k, dtype=np.int64)
    for j, i in enumerate(order):
        old |= ((idx >> (k - 1 - j)) & 1) << (k - 1 - i)
    return np.ascontiguousarray(mats[:, old][:, :, old]), tuple(bits[i] for i in order)


def _partial(rho: np.ndarray, qubits: tuple[int, ...], keep: tuple[int, ...]) -> np.ndarray:
    """Per-shot reduced density matrices on ``keep`` (a subset of ``qubits``, in its own order)
    from those on ``qubits``; first listed = most significant index bit in both."""
    if keep == qubits:
        return rho
    k = len(qubits)
    letters = "abcdefgh"
    rows, cols = list(letters[:k]), list(letters[k : 2 * k])
    for i, q in enumerate(qubits):
        if q not in keep:
            cols[i] = rows[i]  # traced out
    out_rows = "".join(rows[qubits.index(q)] for q in keep)
    out_cols = "".join(cols[qubits.index(q)] for q in keep)
    t = rho.reshape((rho.shape[0],) + (2,) * (2 * k))
    r = np.einsum(f"z{''.join(rows)}{''.join(cols)}->z{out_rows}{out_cols}", t)
    d = 1 << len(keep)
    return r.reshape(rho.shape[0], d, d)


#: A shot's unnormalised state is rescaled once its squared norm falls below this (dtype: floor).
_RENORMALISE_BELOW = {nat.C128: 1e-150, nat.C64: 1e-8}


class _Batch:
    """B states of n qubits in HBM plus their classical bits and if/else flags.

    Every operator is queued and fused, up to :data:`FUSE_WIDTH` qubits at a time, into one pass
    over the batch: each shot gets the product of its own choices, from a table holding one fused
    operator per distinct combination of choices in the batch. A choice is a mixture branch (from
    the draws), an if/else arm (from the flags), a measurement's projection or a Kraus operator
    (from the draws and the reduced density matrix). Projections and Kraus operators are applied
    without renormalising: each decision reads the reduced density matrix divided by its trace,
    which is the squared norm of the shot's state.

    A decision needs the reduced density matrix of its qubits. It is computed on the device over
    up to two qubits -- those of the decision and of the next ones, chosen ahead so that the
    operators in between stay inside the set or miss it -- and then kept on the host, where each
    queued operator inside the set transforms it, so the following decisions read it without
    another pass. An operator that straddles the set, or a projection or Kraus operator outside it,
    drops it.

    Shots with the same history share one device state, a slot (``slot[b]`` for shot b; slots
    0..s-1 of the buffer are in use, and every pass covers those only). When the shots of one slot
    choose differently, the slot's state is first copied to a new slot for each further choice.
    The device work then follows the number of distinct histories, not the number of shots: at
    1e-3..1e-2 noise most shots take the same branches for most of the circuit. A slot's state
    is the state each of its shots would hold on its own: the same operators in the same order."""

    def __init__(self, n: int, batch: int, dtype: int, nflags: int, num_clbits: int, share: bool = True) -> None:
        self.n, self.b, self.dtype = n, batch, dtype
        # the slot holding each shot's state, and the slots in use (share=False: one per shot)
        self.slot = np.zeros(batch, dtype=np.int64) if share else np.arange(batch, dtype=np.int64)
        self.s = 1 if share else batch
        self._allocate()
        self.wide = num_clbits > 62
        self.clbits = np.zeros(batch, dtype=object if self.wide else np.int64)
        self.flags = np.zeros((max(nflags, 1), batch), dtype=bool)
        self.pending: list[_Queued] = []
        self.known: tuple[tuple[int, ...], np.ndarray] | None = None  # (qubits, unnormalised rho)

    # -- device operations -------------------------------------------------------------------

    def _allocate(self) -> None:
        """Room for one state per shot; the slots in use start as |0...0>."""
        self.buf = torch.empty(self.b << self.n, dtype=_TORCH_DTYPE[self.dtype], device=require_device())
        nat.call("lsim_sv_set_basis", ptr(self.buf), self.dtype, self.n, self.s, 0, stream_handle())

    def run_pass(self, p: Pass) -> None:
        STATS["passes"] += 1
        run_passes(self.buf, self.dtype, self.n, self.s, [p])

    def finish(self) -> None:
        torch.cuda.current_stream().synchronize()

    def apply_selected(self, kind: str, qubits: tuple[int, ...], table: np.ndarray, sel: np.ndarray) -> None:
        """Per-slot operator: slot j gets table[sel[j]] (dense matrices or diagonals), or is left
        alone when sel[j] < 0."""
        if not np.any(sel >= 0):
            return
        STATS["passes"] += 1
        # the operators and the selectors go over in one copy: the table's bytes, then the int32
        # selectors (the table's byte size is a multiple of 8, so they stay aligned)
        mats = np.ascontiguousarray(table, dtype=_NP_DTYPE[self.dtype]).reshape(-1).view(np.uint8)
        sels = np.ascontiguousarray(sel, dtype=np.int32).view(np.uint8)
        dev = to_device(np.concatenate([mats, sels]))
        base = ptr(dev)
        fn = "lsim_sv_apply_dense" if kind == "dense" else "lsim_sv_apply_diag"
        nat.call(fn, ptr(self.buf), self.dtype, self.n, self.s, len(qubits), nat.int_array(qubits), base, base + mats.size, stream_handle())
        dev.record_stream(torch.cuda.current_stream())

    def reduced(self, qubits: tuple[int, ...]) -> np.ndarray:
        """Per-slot reduced density matrices [s][2^k][2^k] on ``qubits`` (float64 sums)."""
        k = len(qubits)
        out = torch.empty(self.s << (2 * k), dtype=torch.complex128, device=self.buf.device)
        ws = WORKSPACE.get(nat.load().lsim_reduce_workspace_bytes(self.n, self.s, k))
        nat.call(
            "lsim_sv_reduced_dm", ptr(self.buf), self.dtype, self.n, self.s, k, nat.int_array(qubits), ptr(out), ptr(ws), stream_handle()
        )
        t0 = time.perf_counter()
        host = out.cpu().numpy().reshape(self.s, 1 << k, 1 << k)
        STATS["syncs"] += 1
        STATS["wait_seconds"] += time.perf_counter() - t0
        return host

    def copy_slots(self, src: np.ndarray, dst: np.ndarray) -> None:
        """State dst[j] <- state src[j] (slots in use; no slot is both a source and a target)."""
        both = to_device(np.ascontiguousarray(np.stack([src, dst]), dtype=np.int32))
        nat.call("lsim_sv_copy_states", ptr(self.buf), self.dtype, self.n, self.s, len(src), ptr(both[0]), ptr(both[1]), stream_handle())
        both.record_stream(torch.cuda.current_stream())

    def _to_slots(self, sel: np.ndarray, choices: int) -> np.ndarray:
        """The per-slot selection for a per-shot one (sel[b] in -1..choices-1). Within a slot, the
        shots of the lowest choice stay; those of each further choice move to a new slot that
        starts as a copy of theirs."""
        key = self.slot * (choices + 1) + (sel.astype(np.int64) + 1)
        uniq, inverse = np.unique(key, return_inverse=True)
        owner, pick = uniq // (choices + 1), uniq % (choices + 1) - 1
        moved = np.flatnonzero(np.r_[False, owner[1:] == owner[:-1]])
        target = owner.copy()
        if moved.size:
            target[moved] = self.s + np.arange(moved.size)
            self.s += int(moved.size)
            self.copy_slots(owner[moved], target[moved])
        self.slot = target[np.asarray(inverse).reshape(-1)]
        out = np.full(self.s, -1, dtype=np.int64)
        out[target] = pick
        return out

    # -- the queue ---------------------------------------------------------------------------

    def active(self, guard: tuple[tuple[int, bool], ...]) -> np.ndarray | None:
        if not guard:
            return None
        mask = np.ones(self.b, dtype=bool)
        for flag, polarity in guard:
            mask &= self.flags[flag] == polarity
        return mask

    def queue(self, qubits: tuple[int, ...], table: list[np.ndarray], idx: np.ndarray | None = None, unitary: bool = True) -> None:
        if idx is not None and not np.any(idx >= 0):
            return
        it = _Queued(qubits, table, idx, unitary)
        self.pending.append(it)
        self._follow(it)

    def _follow(self, it: _Queued) -> None:
        """Carry the kept reduced density matrix through a queued operator, or drop it."""
        if self.known is None:
            return
        kept, rho = self.known
        span = set(it.qubits)
        if not span <= set(kept):
            if not (span.isdisjoint(kept) and it.unitary):
                self.known = None
            return
        pos = [kept.index(q) for q in it.qubits]
        k = len(kept)
        ops = [embed(m, pos, k) for m in it.table]
        if it.idx is None:
            rho = ops[0] @ rho @ ops[0].conj().T
        else:
            rho = rho.copy()
            for j, e in enumerate(ops):
                pick = it.idx == j
                if np.any(pick):
                    rho[pick] = e @ rho[pick] @ e.conj().T
        self.known = (kept, rho)

    def _pass_seconds(self) -> float:
        """About what one pass over the batch costs on the device (read and write at 5 TB/s)."""
        return 2.0 * (self.s << self.n) * np.dtype(_NP_DTYPE[self.dtype]).itemsize / 5e12

    def flush(self) -> None:
        pending, self.pending = self.pending, []
        for qubits, items in _merge_groups(_group(pending, FUSE_WIDTH), FUSE_WIDTH):
            for part in _split_for_host(items, len(qubits), self._pass_seconds()):
                self._apply_block(qubits, part)

    def _apply_block(self, qubits: list[int], items: list[_Queued]) -> None:
        per_shot = [it for it in items if it.idx is not None]
        if not per_shot:
            op = _fuse(qubits, items, np.zeros((1, 0), dtype=np.int64))[0]
            self.run_pass(_as_pass(op, tuple(qubits), len(items)))
            return
        combos, sel = _combinations([it.idx for it in per_shot])
        if len(per_shot) == len(items):
            # shots whose choices are all the identity are left alone
            sel = np.where(np.all(combos < 0, axis=1)[sel], -1, sel)
        mats, bits = _descending_many(_fuse(qubits, items, combos), tuple(qubits))
        self.apply_selected("dense", bits, mats, self._to_slots(sel, len(mats)))

    def density(self, qubits: tuple[int, ...], ahead: tuple[int, ...]) -> np.ndarray:
        """Unnormalised per-shot reduced density matrices on ``qubits`` of the state with every
        queued operator applied; ``ahead`` (containing ``qubits``) is the set to compute on the
        device when the kept matrices do not cover ``qubits``."""
        if self.known is not None and set(qubits) <= set(self.known[0]):
            return _partial(self.known[1], self.known[0], qubits)
        # operators at the end of the queue that act inside the set wait for the next pass:
        # the host carries the reduced density matrix through them instead
        tail: list[_Queued] = []
        while self.pending and set(self.pending[-1].qubits) <= set(ahead):
            tail.insert(0, self.pending.pop())
        self.flush()
        rho = self.reduced(ahead)
        trace = np.trace(rho, axis1=1, axis2=2).real
        low = trace < _RENORMALISE_BELOW[self.dtype]
        if np.any(low & (trace > 0)) and self.n:
            # keep unnormalised states away from underflow: rescale those slots to norm 1
            fix = low & (trace > 0)
            s = np.where(fix, 1.0 / np.sqrt(np.where(fix, trace, 1.0)), 1.0)
            self.apply_selected("diag", (0,), np.stack([s, s], axis=1).astype(np.complex128), np.where(fix, np.arange(self.s), -1))
            rho = rho * np.where(fix, 1.0 / np.where(fix, trace, 1.0), 1.0)[:, None, None]
        self.known = (ahead, rho[self.slot])
        for it in tail:
            self.pending.append(it)
            self._follow(it)
        assert self.known is not None
        return _partial(self.known[1], self.known[0], qubits)

    # -- classical bits ----------------------------------------------------------------------

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


_MIXTURES: dict[int, tuple[Channel, np.ndarray, np.ndarray, list[np.ndarray]]] = {}


def _mixture_tables(channel: Channel) -> tuple[np.ndarray, np.ndarray, list[np.ndarray]]:
    """The channel's cumulative branch probabilities in the reference's order, each branch's
    index into the non-identity unitaries (-1: the identity, and one more -1 past the last
    branch), and those unitaries; built once per channel object."""
    hit = _MIXTURES.get(id(channel))
    if hit is not None and hit[0] is channel:
        return hit[1], hit[2], hit[3]
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
    out = (np.array(cumulative), np.array(index + [-1], dtype=np.int64), table)
    if len(_MIXTURES) > 256:
        _MIXTURES.clear()
    _MIXTURES[id(channel)] = (channel, *out)
    return out


def _mixture_choice(channel: Channel, r: np.ndarray) -> tuple[np.ndarray, list[np.ndarray]]:
    """Branch index per shot into the non-identity unitaries (-1: identity or no branch), with
    the cumulative sums formed in the reference's order (engines.py:162-170)."""
    cumulative, lookup, table = _mixture_tables(channel)
    branch = np.searchsorted(cumulative, r, side="right")
    return lookup[branch], table


def _kraus_index(channel: Channel, rho: np.ndarray, r: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Per-shot Kraus operator index (engines.py:171-183) given normalised reduced density
    matrices: the first i with r < sum of weights so far and w_i > 0, else the last one if its
    weight is positive, else -1 (the state is left alone). Also returns the chosen weights and
    the running sum of weights at the chosen branch (what the draw was compared with)."""
    nshots = r.shape[0]
    idx = np.full(nshots, -1, dtype=np.int64)
    chosen_w = np.zeros(nshots)
    chosen_acc = np.zeros(nshots)
    acc = np.zeros(nshots)
    weight = np.zeros(nshots)
    for i, kraus in enumerate(channel.kraus):
        weight = np.einsum("ra,sac,rc->s", kraus, rho, kraus.conj()).real
        acc = acc + weight
        pick = (idx < 0) & (r < acc) & (weight > 0)
        idx[pick] = i
        chosen_w[pick] = weight[pick]
        chosen_acc[pick] = acc[pick]
    rest = (idx < 0) & (weight > 0)
    idx[rest] = len(channel.kraus) - 1
    chosen_w[rest] = weight[rest]
    chosen_acc[rest] = acc[rest]
    return idx, chosen_w, chosen_acc


def _kraus_choice(channel: Channel, rho: np.ndarray, r: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Per-shot normalised Kraus operator K_i / sqrt(w_i) for :func:`_kraus_index`'s choice
    (``sel[b] = b``), or ``sel[b] = -1`` when no operator applies."""
    idx, w, _ = _kraus_index(channel, rho, r)
    d = channel.kraus[0].shape[0]
    mats = np.zeros((r.shape[0], d, d), dtype=np.complex128)
    for i, kraus in enumerate(channel.kraus):
        pick = idx == i
        mats[pick] = kraus[None] / np.sqrt(w[pick])[:, None, None]
    return mats, np.where(idx >= 0, np.arange(r.shape[0]), -1)


def _reads(step: _Step) -> bool:
    """True for the steps whose decision reads the state."""
    return step.kind in ("measure", "reset") or (step.kind == "channel" and step.channel is not None and step.channel.mixture is None)


def _ahead(steps: list[_Step], t: int) -> tuple[int, ...]:
    """Qubits to reduce over at reading step t: its own, then those of the next reading steps
    while the set holds at most two qubits and every operator in between either stays inside the
    set or misses it (global phases and if/else conditions touch no qubit)."""
    chosen = list(steps[t].qubits)
    between: list[set[int]] = []
    for s in steps[t + 1 :]:
        if not _reads(s):
            if s.kind in ("gate", "channel"):
                between.append(set(s.qubits))
            continue
        union = set(chosen) | set(s.qubits)
        if len(union) > 2 or any(not (b <= union or b.isdisjoint(union)) for b in between):
            break
        chosen += [q for q in s.qubits if q not in chosen]
    return tuple(chosen)


_PROJECT = [np.array([[1, 0], [0, 0]], dtype=np.complex128), np.array([[0, 0], [0, 1]], dtype=np.complex128)]
_RESET = [np.array([[1, 0], [0, 0]], dtype=np.complex128), np.array([[0, 1], [0, 0]], dtype=np.complex128)]


def _run_chunk(
    steps: list[_Step],
    circuit: Circuit,
    draws: np.ndarray,
    dtype: int,
    nflags: int,
    batch: _Batch | None = None,
    share: bool = True,
    decisions: np.ndarray | None = None,
) -> np.ndarray:
    """Every shot of one chunk (``draws``: [shots, D]); returns the classical bits. When given,
    ``decisions`` ([shots, D, 2], NaN-filled) receives per draw column the probability a
    state-dependent decision compared its draw with and the branch it took, as the program
    kernel records them."""
    nshots = draws.shape[0]
    st = batch if batch is not None else _Batch(circuit.num_qubits, nshots, dtype, nflags, circuit.num_clbits, share)
    for t, step in enumerate(steps):
        mask = st.active(step.guard)
        if step.kind == "gate":
            assert step.matrix is not None
            st.queue(step.qubits, [step.matrix], None if mask is None else np.where(mask, 0, -1))
        elif step.kind == "phase":
            continue  # a global phase changes no probability, and only bits leave the batch
        elif step.kind == "cond":
            c = step.condition
            assert c is not None
            holds = (st.clbit_values(c.clbits) == c.value) != c.negate
            st.flags[step.flag] = holds.astype(bool)
        elif step.kind == "channel" and step.channel is not None and step.channel.mixture is not None:
            sel, table = _mixture_choice(step.channel, draws[:, step.column])
            if table:
                st.queue(step.qubits, table, sel if mask is None else np.where(mask, sel, -1))
        else:
            rho = st.density(step.qubits, _ahead(steps, t))
            trace = np.trace(rho, axis1=1, axis2=2).real
            norm = np.where(trace > 0, trace, 1.0)
            r = draws[:, step.column]
            if step.kind == "channel":
                ch = step.channel
                assert ch is not None
                idx, _, acc = _kraus_index(ch, rho / norm[:, None, None], r)
                st.queue(step.qubits, list(ch.kraus), idx if mask is None else np.where(mask, idx, -1), unitary=False)
                if decisions is not None:
                    _record(decisions, step.column, acc, idx, mask)
                continue
            p1 = np.minimum(np.maximum(rho[:, 1, 1].real / norm, 0.0), 1.0)
            outcome = (r < p1).astype(np.int64)
            if decisions is not None:
                _record(decisions, step.column, p1, outcome, mask)
            table = _PROJECT if step.kind == "measure" else _RESET
            st.queue(step.qubits, table, outcome if mask is None else np.where(mask, outcome, -1), unitary=False)
            if step.kind == "measure":
                if step.readout is not None:
                    err = step.readout
                    flip = draws[:, step.column + 1] < np.where(outcome == 1, err.p10, err.p01)
                    outcome = outcome ^ flip.astype(np.int64)
                if step.clbit is not None:
                    st.set_bits(step.clbit, outcome, mask)
    # operators still queued change no recorded bit: they are dropped
    st.finish()
    return st.clbits


def _record(decisions: np.ndarray, column: int, prob: np.ndarray, choice: np.ndarray, mask: np.ndarray | None) -> None:
    rows = slice(None) if mask is None else mask
    decisions[rows, column, 0] = prob[rows]
    decisions[rows, column, 1] = choice[rows]


def run_batched(
    circuit: Circuit,
    noise: NoiseModel | None,
    shots: int,
    rng: np.random.Generator,
    *,
    dtype: int = nat.C128,
    batch_bytes: int | None = None,
    share: bool = True,
    record: dict | None = None,
) -> np.ndarray:
    """Classical bits of every shot (int64, or Python ints beyond 62 clbits), all shots batched.
    Requires draw-balanced if/else arms (see :func:`draw_count`). ``share=False`` gives every shot
    its own state from the start (the same results, with more device work). ``record`` (a dict)
    receives ``draws`` [shots, D] and ``decisions`` [shots, D, 2] (see :func:`_run_chunk`)."""
    STATS.update(syncs=0, wait_seconds=0.0, passes=0)
    steps, per_shot, nflags = compile_program(circuit, noise)
    item = np.dtype(_NP_DTYPE[dtype]).itemsize
    limit = batch_bytes if batch_bytes is not None else BATCH_BYTES
    chunk = max(1, min(shots, limit // (item << circuit.num_qubits)))
    out, all_draws, all_dec = [], [], []
    for start in range(0, shots, chunk):
        b = min(chunk, shots - start)
        draws = rng.random(b * per_shot).reshape(b, per_shot) if per_shot else np.zeros((b, 0))
        dec = np.full((b, per_shot, 2), np.nan) if record is not None else None
        out.append(_run_chunk(steps, circuit, draws, dtype, nflags, share=share, decisions=dec))
        if record is not None:
            all_draws.append(draws)
            all_dec.append(dec)
    if record is not None:
        record["draws"] = np.concatenate(all_draws) if all_draws else np.zeros((0, per_shot))
        record["decisions"] = np.concatenate(all_dec) if all_dec else np.zeros((0, per_shot, 2))
    return np.concatenate(out) if out else np.zeros(0, dtype=np.int64)


#: Shots per program launch (bounds the draw buffer and output sizes of one launch).
LDS_CHUNK = 1 << 16


def run_lds(
    circuit: Circuit,
    noise: NoiseModel | None,
    shots: int,
    rng: np.random.Generator,
    *,
    dtype: int = nat.C128,
    record: dict | None = None,
) -> np.ndarray | None:
    """All shots as LDS programs, one workgroup per shot, when the state fits a workgroup's LDS
    and the circuit fits the program kernel; None otherwise (the caller uses :func:`run_batched`).
    Draws are taken from ``rng`` exactly as :func:`run_batched` takes them; ``record`` as there."""
    from lucivant_sim import program as pg

    if circuit.num_qubits < 1 or circuit.num_qubits > pg.max_qubits(dtype) or circuit.num_clbits > 62:
        return None
    try:
        prog = pg.compile_trajectory(circuit, noise)
    except pg.Unsupported:
        return None
    out, all_draws, all_dec = [], [], []
    # the kernel's draw offsets are 32-bit: a launch holds fewer than 2^31 draws
    chunk = max(1, min(LDS_CHUNK, ((1 << 31) - 1) // max(prog.draws, 1)))
    for start in range(0, shots, chunk):
        b = min(chunk, shots - start)
        draws = rng.random(b * prog.draws).reshape(b, prog.draws) if prog.draws else None
        res = pg.run_programs([prog], repeat=b, dtype=dtype, draws=draws, want_clbits=True, want_decisions=record is not None)
        assert res.clbits is not None
        out.append(res.clbits.astype(np.int64))
        if record is not None:
            all_draws.append(draws if draws is not None else np.zeros((b, 0)))
            all_dec.append(res.decisions if res.decisions is not None else np.zeros((b, 0, 2)))
    if record is not None:
        record["draws"] = np.concatenate(all_draws)
        record["decisions"] = np.concatenate(all_dec)
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
    circuit: Circuit,
    noise: NoiseModel | None,
    shots: int,
    rng: np.random.Generator,
    *,
    dtype: int = nat.C128,
    record: dict | None = None,
) -> tuple[list[int] | np.ndarray, int]:
    """The classical bit pattern of every shot, in shot order, and the precision it was computed
    in (circuits whose if/else arms draw unequally run shot after shot in complex128, and record
    nothing). ``record`` receives the draws and decisions of a batched run (:func:`run_batched`)."""
    if shots == 0:
        return [], dtype
    if draw_count(circuit.operations, noise) is None:
        return run_sequential(circuit, noise, shots, rng), nat.C128
    values = run_lds(circuit, noise, shots, rng, dtype=dtype, record=record)
    if values is None:
        values = run_batched(circuit, noise, shots, rng, dtype=dtype, record=record)
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
PRECISION_CHECK_SHOTS = 4096


def complex64_check(
    circuit: Circuit, noise: NoiseModel | None, seed: int, values: list[int] | np.ndarray, used: int, record: dict | None = None
) -> dict[str, object]:
    """The error of a complex64 trajectory run against complex128 (float64 arithmetic, the
    reference's): its first shots are rerun in complex128 with the same draws (the same seeded
    stream) and compared decision by decision. Shots are deterministic given their draws, so any
    difference is the precision's doing. Reported:

    - ``shots_differing_from_complex128`` and ``counts_tvd``, the total variation distance
      between the two tallies of the checked shots;
    - ``max_abs_dp``: the largest |p64 - p128| over every measurement probability p1 and Kraus
      running weight a draw was compared with, up to each shot's first flipped decision (later
      decisions of that shot follow another history);
    - ``flipped_decisions``: shots in which some decision took another branch, and for the first
      few of them ``flip_margins``, the distance of the draw from the complex128 boundary (a flip
      is a tolerance-level case when it is within about the float32 error of the state)."""
    if used != nat.C64:
        return {"dtype": "complex128", "note": "this circuit runs shot after shot in complex128"}
    m = min(len(values), PRECISION_CHECK_SHOTS)
    ref_record: dict = {}
    ref, _ = shot_values(circuit, noise, m, np.random.default_rng(seed), dtype=nat.C128, record=ref_record)
    mine = np.asarray(values[:m], dtype=np.int64)
    theirs = np.asarray(ref, dtype=np.int64)
    u64, c64 = np.unique(mine, return_counts=True)
    u128, c128 = np.unique(theirs, return_counts=True)
    tally: dict[int, int] = {}
    for u, c in zip(u64.tolist(), c64.tolist()):
        tally[u] = tally.get(u, 0) + c
    for u, c in zip(u128.tolist(), c128.tolist()):
        tally[u] = tally.get(u, 0) - c
    out: dict[str, object] = {
        "dtype": "complex64",
        "against": "complex128 rerun of the same shots with the same draws",
        "checked_shots": m,
        "shots_differing_from_complex128": int(np.sum(mine != theirs)),
        "counts_tvd": 0.5 * sum(abs(v) for v in tally.values()) / m if m else 0.0,
    }
    if record is None or "decisions" not in record or "decisions" not in ref_record:
        return out
    out.update(compare_decisions(record["decisions"][:m], ref_record["decisions"], ref_record["draws"]))
    return out


def compare_decisions(d64: np.ndarray, d128: np.ndarray, draws: np.ndarray, margins: int = 16) -> dict[str, object]:
    """Decision-by-decision comparison of two runs of the same shots and draws (records laid out
    as :func:`_run_chunk` fills them)."""
    shots, cols = d128.shape[:2]
    if shots == 0 or cols == 0:
        return {"decisions_compared": 0, "max_abs_dp": 0.0, "flipped_decisions": 0, "flip_margins": []}
    made = ~np.isnan(d128[:, :, 1]) & ~np.isnan(d64[:, :, 1])
    flip = made & (d64[:, :, 1] != d128[:, :, 1])
    flipped = flip.any(axis=1)
    first = np.where(flipped, flip.argmax(axis=1), cols)
    same = made & (np.arange(cols)[None, :] <= first[:, None]) & ~flip
    dp = np.abs(d64[:, :, 0] - d128[:, :, 0])
    where = np.flatnonzero(flipped)[:margins]
    return {
        "decisions_compared": int(same.sum()),
        "max_abs_dp": float(dp[same].max()) if same.any() else 0.0,
        "flipped_decisions": int(flipped.sum()),
        "flip_margins": [float(abs(draws[s, first[s]] - d128[s, first[s], 0])) for s in where],
    }


__all__ = ["compare_decisions", "complex64_check", "compile_program", "draw_count", "run_batched", "run_sequential", "run_trajectories", "shot_values"]
