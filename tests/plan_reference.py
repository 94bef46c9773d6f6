"""The pass planners restated in Python on NumPy, as the library planned before its native
planner (lsim_host_plan_fused / lsim_host_plan_tiles) took over: the specification those are
checked against in test_host.py. The fusion steps (one-qubit runs, greedy, merge, descending
targets) are the plan module's own Python functions, which the density-matrix planner still uses;
the tile planner's choice of bits counts runnable gates with the rule written out below."""

from __future__ import annotations

import numpy as np

from lucivant_sim.circuit import Barrier, Gate, GlobalPhase, Measure, Operation, Reset
from lucivant_sim.gates import Matrix
from lucivant_sim.plan import (
    KERNEL_MAX_BITS,
    TILE_BITS,
    TILE_BUDGET,
    TILE_LOW,
    TILE_MAX_OPS,
    Pass,
    TilePass,
    _as_pass,
    _greedy,
    _merge,
    _one_qubit_runs,
    descending,
    is_diagonal,
)


def plan_statevector_reference(
    ops: tuple[Operation, ...], num_qubits: int, width: int = 4, low_bits: int = 0, low_width: int = 0, merge: bool = True
) -> list[Pass]:
    """Passes for an ideal static circuit on a statevector (measurements deferred, resets of
    untouched qubits and barriers are no-ops, as in simulator.py:229-251). ``low_bits`` and
    ``low_width`` optionally narrow the blocks that touch the lowest qubits; ``merge`` joins
    blocks that fit together after the greedy pass (:func:`_merge`)."""
    if not 1 <= width <= KERNEL_MAX_BITS:
        raise ValueError(f"fusion width must be between 1 and {KERNEL_MAX_BITS}")
    items: list[tuple[Matrix, tuple[int, ...]]] = []
    phase = 1.0 + 0.0j
    for op in ops:
        if isinstance(op, Gate):
            if len(op.qubits) > KERNEL_MAX_BITS:
                raise ValueError(f"{op.name} acts on {len(op.qubits)} qubits; gates of up to {KERNEL_MAX_BITS} are supported")
            items.append((np.asarray(op.matrix, dtype=np.complex128), tuple(op.qubits)))
        elif isinstance(op, GlobalPhase):
            phase *= np.exp(1j * op.angle)
        elif isinstance(op, (Measure, Reset, Barrier)):
            continue
        else:
            raise ValueError(f"a static statevector can't run {type(op).__name__}")
    passes: list[Pass] = []
    blocks = _greedy(_one_qubit_runs(items), width, low_bits, low_width)
    if merge:
        blocks = _merge(blocks, width, low_bits, low_width)
    for blk in blocks:
        passes.append(_as_pass(blk.matrix, tuple(blk.qubits), blk.gates))
    if phase != 1:
        if passes:
            p = passes[0]
            p.matrix = p.matrix * phase
        elif num_qubits > 0:
            passes.append(Pass("scale", (), np.array([phase], dtype=np.complex128), 0))
    return passes



def tile_structure(qubits: list[tuple[int, ...]], diagonal: list[bool]) -> tuple[float, int]:
    """The cost (in :func:`_tile_cost` units) and the number of the operators
    fuse_in_tile forms from gates on ``qubits``, from the qubits alone: the one-qubit runs,
    greedy and merge at width 2 decide by qubit sets only. An operator counts as diagonal when
    all its gates are (a product of diagonals is diagonal; an accidentally diagonal product of
    other gates counts as dense, so the cost is never underestimated)."""
    runs: list[list] = []  # [qubits, diagonal]
    open_run: dict[int, int] = {}
    for qs, dg in zip(qubits, diagonal):
        if len(qs) == 1:
            at = open_run.get(qs[0])
            if at is None:
                open_run[qs[0]] = len(runs)
                runs.append([qs, dg])
            else:
                runs[at][1] = runs[at][1] and dg
            continue
        for q in qs:
            open_run.pop(q, None)
        runs.append([qs, dg])
    blocks: list[list] = []  # [qubit list, diagonal]
    last: dict[int, int] = {}
    for qs, dg in runs:
        j = max((last.get(q, -1) for q in qs), default=-1)
        if j >= 0 and len(set(blocks[j][0]) | set(qs)) <= 2:
            blocks[j][0].extend(q for q in qs if q not in blocks[j][0])
            blocks[j][1] = blocks[j][1] and dg
        else:
            blocks.append([list(qs), dg])
            j = len(blocks) - 1
        for q in qs:
            last[q] = j
    while True:  # the merge of _merge at width 2, on qubit sets
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
                    if len(mine | other) <= 2:
                        first, second = (blocks[j2], blocks[j]) if backward else (blocks[j], blocks[j2])
                        blocks[j2] = [first[0] + [q for q in second[0] if q not in first[0]], first[1] and second[1]]
                        alive[j] = False
                        break
                    if mine & other:
                        break
            blocks = [b for b, a in zip(blocks, alive) if a]
        if len(blocks) == before:
            break
    cost = sum((0.5 * len(b[0])) if b[1] else (1.0 if len(b[0]) == 1 else 2.5) for b in blocks)
    return cost, len(blocks)


def fuse_in_tile(items: list[tuple[Matrix, tuple[int, ...]]], bits: tuple[int, ...]) -> list[Pass]:
    """The gates of one tile pass as operators on at most 2 tile bits, in an order that keeps
    every dependency (the statevector greedy and merge at width 2, then tile-local bits)."""
    local = {q: i for i, q in enumerate(bits)}
    blocks = _merge(_greedy(_one_qubit_runs(items), 2), 2)
    out = []
    for blk in blocks:
        m, qs = descending(np.asarray(blk.matrix, dtype=np.complex128), tuple(blk.qubits))
        tb = tuple(local[q] for q in qs)  # the higher tile bit first
        if is_diagonal(m):
            out.append(Pass("diag", tb, np.ascontiguousarray(np.diag(m)), blk.gates))
        else:
            out.append(Pass("dense", tb, np.ascontiguousarray(m), blk.gates))
    return out


def plan_tiles_reference(
    ops: tuple[Operation, ...], num_qubits: int, budget: float = TILE_BUDGET, tile_bits: int = TILE_BITS, low: int = TILE_LOW
) -> list[TilePass | Pass]:
    """Passes for an ideal static circuit as tile passes (statevectors of at least ``tile_bits``
    qubits; the same gates and order as :func:`plan_statevector`, so results agree to rounding).

    Each pass takes the pending gates that can run next on one set of tile bits: the ``low``
    lowest bits, then bits added one at a time (two when a two-qubit gate needs both), each time
    the bit that lets the most pending gates run, then improved by single swaps. A gate can run
    when its qubits are tile bits and every earlier pending gate on its qubits can run. The pass
    is cut where its fused operators would cost more than ``budget``. Gates on more than two
    qubits run as their own dense pass."""
    n = num_qubits
    if n < tile_bits:
        raise ValueError(f"tile passes need at least {tile_bits} qubits")
    items: list[tuple[Matrix, tuple[int, ...]]] = []
    phase = 1.0 + 0.0j
    for op in ops:
        if isinstance(op, Gate):
            if len(op.qubits) > KERNEL_MAX_BITS:
                raise ValueError(f"{op.name} acts on {len(op.qubits)} qubits; gates of up to {KERNEL_MAX_BITS} are supported")
            items.append((np.asarray(op.matrix, dtype=np.complex128), tuple(op.qubits)))
        elif isinstance(op, GlobalPhase):
            phase *= np.exp(1j * op.angle)
        elif isinstance(op, (Measure, Reset, Barrier)):
            continue
        else:
            raise ValueError(f"a static statevector can't run {type(op).__name__}")
    masks = [sum(1 << q for q in qs) for _, qs in items]
    diagonal = [is_diagonal(m) for m, _ in items]
    full = (1 << n) - 1
    low_mask = (1 << low) - 1
    window = 64 * n  # pending gates looked at when choosing bits (those beyond are blocked)

    def runnable(sets: list[int], window_masks: np.ndarray) -> list[int]:
        """For each candidate set: how many pending gates can run on it (a gate runs when its
        qubits lie in the set and no earlier gate that cannot run shares a qubit with it; the
        count stops once every qubit is blocked)."""
        out = []
        for s in sets:
            blocked, runs = 0, 0
            for m in window_masks.tolist():
                if m & blocked or m & ~s:
                    blocked |= m
                    if blocked == full:
                        break
                else:
                    runs += 1
            out.append(runs)
        return out

    def first_best(sets: list[int], window_masks: np.ndarray) -> tuple[int, int]:
        vals = runnable(sets, window_masks)
        best = max(range(len(sets)), key=vals.__getitem__)  # the first of the largest
        return sets[best], vals[best]

    def choose(pending: list[int]) -> int:
        wm = np.array([masks[g] for g in pending[:window]], dtype=np.uint64)
        s = low_mask | masks[pending[0]]
        while bin(s).count("1") < tile_bits:
            outside = [q for q in range(n) if not (s >> q) & 1]
            base = runnable([s], wm)[0]
            best, value = first_best([s | (1 << q) for q in outside], wm)
            if value == base and bin(s).count("1") <= tile_bits - 2:
                pairs = [s | (1 << a) | (1 << b) for i, a in enumerate(outside) for b in outside[i + 1 :]]
                s = first_best(pairs, wm)[0]
            else:
                s = best
        cur = runnable([s], wm)[0]
        for _ in range(3):
            better = False
            for a in [q for q in range(low, n) if (s >> q) & 1 and not (masks[pending[0]] >> q) & 1]:
                swaps = [(s & ~(1 << a)) | (1 << b) for b in range(n) if not (s >> b) & 1]
                vals = runnable(swaps, wm)
                for t, v in zip(swaps, vals):
                    if v > cur:
                        s, cur, better = t, v, True
                        break
            if not better:
                break
        return s

    out: list[TilePass | Pass] = []
    pending = list(range(len(items)))
    while pending:
        first = pending[0]
        if len(items[first][1]) > 2:
            m, qs = items[first]
            out.append(_as_pass(m, qs, 1))
            pending = pending[1:]
            continue
        s = choose(pending)
        bits = tuple(q for q in range(n) if (s >> q) & 1)
        blocked = 0
        taken: list[int] = []
        for gi in pending:
            m = masks[gi]
            if m & blocked or m & ~s or len(items[gi][1]) > 2:
                blocked |= m
            else:
                taken.append(gi)
        # the longest prefix of the runnable gates whose operators fit the budget and the table,
        # searched on the operators' structure (which gates fuse into which operator depends on
        # qubits only); the matrices are multiplied once, for the prefix chosen
        lo_n, hi_n = 1, len(taken)
        while lo_n < hi_n:
            mid = (lo_n + hi_n + 1) // 2
            cost, count = tile_structure([items[g][1] for g in taken[:mid]], [diagonal[g] for g in taken[:mid]])
            if cost <= budget and count <= TILE_MAX_OPS:
                lo_n = mid
            else:
                hi_n = mid - 1
        best_ops = fuse_in_tile([items[g] for g in taken[:lo_n]], bits)
        used = set(taken[:lo_n])
        out.append(TilePass(bits, best_ops, lo_n))
        pending = [g for g in pending if g not in used]
    if phase != 1:
        for p in out:
            if p.ops if isinstance(p, TilePass) else True:
                target = p.ops[0] if isinstance(p, TilePass) else p
                target.matrix = target.matrix * phase
                break
        else:
            out.append(Pass("scale", (), np.array([phase], dtype=np.complex128), 0))
    return out
