"""Fusion planning: turning a circuit's operations into the device passes that run it.

A pass is one read and one write of the whole state in HBM, so the planner's job is to apply as
many gates as possible per pass. Statevector: consecutive gates are merged greedily into dense
operators on at most ``width`` qubits (the kernel handles up to 5); a gate joins the latest block
that touches any of its qubits when the union still fits, which is always legal because no later
block touches those qubits. Diagonal results run on the diagonal kernel. Global phases fold into
the first pass. Density matrix: every gate and its noise channel become superoperators on the
2k vector bits of their qubits (U (x) conj(U) for a gate, sum K (x) conj(K) for a channel,
the analytic depolariser as (1-p) I + (p/d)|I>><<I|) and are merged into blocks of at most
``width`` qubits (2 by default, a 16 x 16 superoperator on 4 bits).

The reference runs one NumPy sweep per gate (engines.py:38-119 via simulator.py:229-251) and
merges only runs of one-qubit gates (simulator.py:196-226). Fusing more gates changes results by
rounding only.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from functools import lru_cache

import numpy as np

from lucivant_sim import _native as nat
from lucivant_sim.circuit import Barrier, Gate, GlobalPhase, Measure, Operation, Reset
from lucivant_sim.gates import Matrix
from lucivant_sim.noise import Channel, NoiseModel, kron_conj

#: Widest operator the dense kernel applies (bits).
KERNEL_MAX_BITS = 5


@lru_cache(maxsize=4096)
def _embedding(positions: tuple[int, ...], k: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Index arrays of :func:`embed`: the sub-operator row/column of every (row, column) of the
    k-position operator, and where the untouched positions agree."""
    m = len(positions)
    d = 1 << k
    idx = np.arange(d)
    sub_idx = np.zeros(d, dtype=np.int64)
    mask = 0
    for i, p in enumerate(positions):
        sub_idx |= ((idx >> (k - 1 - p)) & 1) << (m - 1 - i)
        mask |= 1 << (k - 1 - p)
    rest = idx & ~mask
    same = rest[:, None] == rest[None, :]
    rows = np.broadcast_to(sub_idx[:, None], (d, d))
    cols = np.broadcast_to(sub_idx[None, :], (d, d))
    for a in (same, rows, cols):
        a.flags.writeable = False
    return same, rows, cols


def embed(sub: Matrix, positions: list[int], k: int) -> Matrix:
    """``sub`` acting on local positions ``positions`` (first = its most significant bit) of a
    k-position operator whose position j is index bit k-1-j, as a 2^k x 2^k matrix."""
    if len(positions) == k and list(positions) == list(range(k)):
        return np.asarray(sub, dtype=np.complex128)
    same, rows, cols = _embedding(tuple(positions), k)
    return np.where(same, np.asarray(sub, dtype=np.complex128)[rows, cols], 0)


def is_diagonal(m: Matrix) -> bool:
    return not np.any(m - np.diag(np.diag(m)))


@dataclass
class Pass:
    """One device pass: ``matrix`` (2^k x 2^k) on vector ``bits`` (first = most significant
    matrix bit). ``kind`` is "dense", "diag" (matrix holds the diagonal only) or "scale"."""

    kind: str
    bits: tuple[int, ...]
    matrix: np.ndarray
    gates: int = 1  # reference operations this pass replaces

    @property
    def k(self) -> int:
        return len(self.bits)

    def moves(self) -> tuple[tuple[int, ...], np.ndarray, int] | None:
        """For a dense pass whose matrix has one non-zero entry per row and column (a
        permutation with phases, the reference's slice moves): the source column of each row,
        each row's entry, and the mask of rows whose entry is exactly 1; else None. Such a pass
        runs as lsim_sv_apply_perm. Computed once per pass."""
        if self.kind != "dense":
            return None
        cached = self.__dict__.get("_moves", _UNSET)
        if cached is _UNSET:
            m = np.asarray(self.matrix)
            nz = m != 0
            cached = None
            if np.all(nz.sum(axis=1) == 1) and np.all(nz.sum(axis=0) == 1):
                src = nz.argmax(axis=1)
                entries = np.ascontiguousarray(m[np.arange(len(src)), src], dtype=np.complex128)
                unit = sum(1 << r for r, e in enumerate(entries) if e == 1)
                cached = (tuple(int(c) for c in src), entries, unit)
            self.__dict__["_moves"] = cached
        return cached

    def payload(self) -> np.ndarray:
        """What the pass's launch reads from device memory: the matrix, the diagonal, the scale
        factor, or a permutation's row entries."""
        mv = self.moves()
        return self.matrix if mv is None else mv[1]


_UNSET = object()


@dataclass
class _Block:
    qubits: list[int]
    matrix: Matrix
    gates: int = 0

    def extend(self, qubits: tuple[int, ...]) -> None:
        new = [q for q in qubits if q not in self.qubits]
        if new:
            # kron(matrix, I): the new qubits become the least significant positions
            d, e = self.matrix.shape[0], 1 << len(new)
            out = np.zeros((d * e, d * e), dtype=np.complex128)
            diag = np.arange(e)
            out.reshape(d, e, d, e)[:, diag, :, diag] = self.matrix
            self.matrix = out
            self.qubits.extend(new)

    def absorb(self, matrix: Matrix, qubits: tuple[int, ...], count: int = 1) -> None:
        self.extend(qubits)
        k = len(self.qubits)
        if len(qubits) == 1:
            # a one-qubit gate on position p (index bit k-1-p) mixes row pairs p apart
            p = self.qubits.index(qubits[0])
            d = 1 << k
            rows = self.matrix.reshape(1 << p, 2, (d >> (p + 1)) * d)
            self.matrix = np.matmul(np.asarray(matrix, dtype=np.complex128), rows).reshape(d, d)
        else:
            pos = [self.qubits.index(q) for q in qubits]
            self.matrix = embed(matrix, pos, k) @ self.matrix
        self.gates += count


def _one_qubit_runs(items: list[tuple[Matrix, tuple[int, ...]]]) -> list[tuple[Matrix, tuple[int, ...], int]]:
    """The same operators with each run of one-qubit gates on a qubit (no other gate on that
    qubit in between) multiplied into one 2x2 matrix, placed where the run's first gate was: the
    later gates of the run commute with everything in between, which acts on other qubits. The
    greedy then forms the same blocks in the same order, from fewer operators. Returns (matrix,
    qubits, gates it stands for)."""
    out: list[tuple[Matrix, tuple[int, ...], int]] = []
    run: dict[int, int] = {}  # qubit -> index in ``out`` of its open run
    for matrix, qubits in items:
        if len(qubits) == 1:
            q = qubits[0]
            at = run.get(q)
            if at is None:
                run[q] = len(out)
                out.append((matrix, qubits, 1))
            else:
                m, _, count = out[at]
                out[at] = (matrix @ m, qubits, count + 1)
            continue
        for q in qubits:
            run.pop(q, None)
        out.append((matrix, qubits, 1))
    return out


def _greedy(
    items: list[tuple[Matrix, tuple[int, ...]]] | list[tuple[Matrix, tuple[int, ...], int]],
    width: int,
    low_bits: int = 0,
    low_width: int = 0,
) -> list[_Block]:
    """Merge each item into the latest block touching its qubits while the union has at most
    ``width`` qubits -- at most ``low_width`` when the union includes a qubit below ``low_bits``
    (operators on the lowest index bits cost more per pass; the autotuner sets both)."""
    blocks: list[_Block] = []
    last: dict[int, int] = {}  # qubit -> index of the latest block touching it
    for matrix, qubits, *rest in items:
        count = rest[0] if rest else 1
        j = max((last.get(q, -1) for q in qubits), default=-1)
        union = set(blocks[j].qubits) | set(qubits) if j >= 0 else set()
        limit = low_width if low_width and low_bits and min(union, default=low_bits) < low_bits else width
        if j >= 0 and len(union) <= limit:
            blocks[j].absorb(matrix, qubits, count)
        else:
            blk = _Block(list(qubits), np.eye(1 << len(qubits), dtype=np.complex128))
            blk.absorb(matrix, qubits, count)
            blocks.append(blk)
            j = len(blocks) - 1
        for q in qubits:
            last[q] = j
    return blocks


def _static_gates(ops: tuple[Operation, ...]) -> tuple[list[tuple[Matrix, tuple[int, ...]]], complex]:
    """The gates of an ideal static circuit in program order, and the product of its global
    phases (measurements deferred, resets of untouched qubits and barriers are no-ops, as in
    simulator.py:229-251)."""
    items: list[tuple[Matrix, tuple[int, ...]]] = []
    phase = 1.0 + 0.0j
    for op in ops:
        if isinstance(op, Gate):
            if len(op.qubits) > KERNEL_MAX_BITS:
                raise ValueError(f"{op.name} acts on {len(op.qubits)} qubits; gates of up to {KERNEL_MAX_BITS} are supported")
            items.append((op.matrix, tuple(op.qubits)))
        elif isinstance(op, GlobalPhase):
            phase *= np.exp(1j * op.angle)
        elif isinstance(op, (Measure, Reset, Barrier)):
            continue
        else:
            raise ValueError(f"a static statevector can't run {type(op).__name__}")
    return items, phase


def _gate_arrays(items: list[tuple[Matrix, tuple[int, ...]]]) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The native planner's input: qubit counts, 5 qubit slots per gate, matrices back to back."""
    pad = (0,) * KERNEL_MAX_BITS
    nq = np.fromiter((len(qs) for _, qs in items), dtype=np.int32, count=len(items))
    qubits = np.array([(qs + pad)[:KERNEL_MAX_BITS] for _, qs in items], dtype=np.int32).reshape(len(items), KERNEL_MAX_BITS)
    mats = np.concatenate([np.ravel(m) for m, _ in items]).astype(np.complex128, copy=False) if items else np.zeros(1, dtype=np.complex128)
    return nq, qubits, mats


def _native_plan(fn: str, items: list[tuple[Matrix, tuple[int, ...]]], *args: object) -> list[Pass | TilePass]:
    """Runs a native planner over ``items`` and turns its plan into passes: one :class:`Pass` per
    operator pass, one :class:`TilePass` (its operators on tile bits) per tile pass."""
    lib = nat.load()
    nq, qubits, mats = _gate_arrays(items)
    handle = ctypes.c_void_p()
    rc = getattr(lib, fn)(len(items), nq.ctypes.data, qubits.ctypes.data, mats.ctypes.data, *args, ctypes.byref(handle))
    if rc:
        raise nat.NativeError(f"{fn}: {lib.lsim_last_error().decode(errors='replace')}")
    try:
        npass, nops, nvals = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        lib.lsim_host_plan_sizes(handle, ctypes.byref(npass), ctypes.byref(nops), ctypes.byref(nvals))
        tiles = np.zeros(max(npass.value, 1), dtype=np.uint64)
        pass_ops = np.zeros(2 * max(npass.value, 1), dtype=np.int32)
        info = np.zeros(8 * max(nops.value, 1), dtype=np.int32)
        at = np.zeros(max(nops.value, 1), dtype=np.int64)
        values = np.zeros(max(nvals.value, 1), dtype=np.complex128)
        rc = lib.lsim_host_plan_read(handle, tiles.ctypes.data, pass_ops.ctypes.data, info.ctypes.data, at.ctypes.data, values.ctypes.data)
        if rc:
            raise nat.NativeError(f"lsim_host_plan_read: {lib.lsim_last_error().decode(errors='replace')}")
    finally:
        lib.lsim_host_plan_free(handle)
    ops: list[Pass] = []
    flat = info.tolist()
    offsets = at.tolist()
    for i in range(nops.value):
        kind, k = flat[8 * i], flat[8 * i + 1]
        bits = tuple(flat[8 * i + 2 : 8 * i + 2 + k])
        d = 1 << k
        if kind == nat.PLAN_DIAG:
            ops.append(Pass("diag", bits, values[offsets[i] : offsets[i] + d], flat[8 * i + 7]))
        else:
            ops.append(Pass("dense", bits, values[offsets[i] : offsets[i] + d * d].reshape(d, d), flat[8 * i + 7]))
    out: list[Pass | TilePass] = []
    po = pass_ops.tolist()
    for j, tile in enumerate(tiles.tolist()[: npass.value]):
        first, count = po[2 * j], po[2 * j + 1]
        if tile == 0:
            out.append(ops[first])
        else:
            mine = ops[first : first + count]
            out.append(TilePass(tuple(q for q in range(64) if (tile >> q) & 1), mine, sum(op.gates for op in mine)))
    return out


def plan_statevector(
    ops: tuple[Operation, ...], num_qubits: int, width: int = 4, low_bits: int = 0, low_width: int = 0, merge: bool = True
) -> list[Pass]:
    """Passes for an ideal static circuit on a statevector (measurements deferred, resets of
    untouched qubits and barriers are no-ops, as in simulator.py:229-251), planned by the
    library's native planner (``lsim_host_plan_fused``): runs of one-qubit gates multiplied, then
    greedy blocks of at most ``width`` qubits, then (``merge``) blocks joined across blocks they
    commute with -- the steps of :func:`_one_qubit_runs`, :func:`_greedy` and :func:`_merge`.
    ``low_bits`` and ``low_width`` optionally narrow the blocks that touch the lowest qubits."""
    if not 1 <= width <= KERNEL_MAX_BITS:
        raise ValueError(f"fusion width must be between 1 and {KERNEL_MAX_BITS}")
    items, phase = _static_gates(ops)
    passes: list[Pass] = []
    if items:
        nbits = max(num_qubits, 1 + max(max(qs) for _, qs in items))
        passes = _native_plan("lsim_host_plan_fused", items, nbits, width, low_bits, low_width, 1 if merge else 0)  # type: ignore[assignment]
    if phase != 1:
        if passes:
            p = passes[0]
            p.matrix = p.matrix * phase
        elif num_qubits > 0:
            passes.append(Pass("scale", (), np.array([phase], dtype=np.complex128), 0))
    return passes


def _join(first: _Block, second: _Block) -> _Block:
    """One block applying ``first`` then ``second``."""
    joined = _Block(list(first.qubits), np.eye(1 << len(first.qubits), dtype=np.complex128))
    joined.absorb(first.matrix, tuple(first.qubits))
    joined.absorb(second.matrix, tuple(second.qubits))
    joined.gates = first.gates + second.gates
    return joined


def _merge(blocks: list[_Block], width: int, low_bits: int = 0, low_width: int = 0) -> list[_Block]:
    """Fewer blocks: two blocks join when their union fits the width (with the same low-qubit cap
    as :func:`_greedy`) and no block between them shares a qubit with the one that moves. Forward,
    block j moves to a later block j2 (commuting past the blocks between); backward, block j2
    moves to an earlier block j. Repeated until nothing merges. The greedy leaves small blocks
    behind (the first gates on fresh qubits, the rest of a layer), and each merge saves one pass
    over the state."""

    def fits(a: set[int], b: set[int]) -> bool:
        union = a | b
        limit = low_width if low_width and low_bits and min(union) < low_bits else width
        return len(union) <= limit

    while True:
        before = len(blocks)
        for backward in (False, True):
            alive = [True] * len(blocks)
            order = range(len(blocks) - 1, -1, -1) if backward else range(len(blocks))
            for j in order:
                if not alive[j]:
                    continue
                mine = set(blocks[j].qubits)
                others = range(j - 1, -1, -1) if backward else range(j + 1, len(blocks))
                for j2 in others:
                    if not alive[j2]:
                        continue
                    other = set(blocks[j2].qubits)
                    if fits(mine, other):
                        first, second = (j2, j) if backward else (j, j2)
                        blocks[j2] = _join(blocks[first], blocks[second])
                        alive[j] = False
                        break
                    if mine & other:
                        break
            blocks = [b for b, a in zip(blocks, alive) if a]
        if len(blocks) == before:
            return blocks


def descending(matrix: Matrix, bits: tuple[int, ...]) -> tuple[Matrix, tuple[int, ...]]:
    """The same operator with its targets listed from the highest bit down (the lowest bit is
    then the matrix's least significant one). The dense kernel loads the amplitudes of a group in
    matrix-index order, so amplitudes that share a 128-byte line -- those differing only in the
    lowest target bits -- are then fetched by consecutive instructions."""
    k = len(bits)
    order = sorted(range(k), key=lambda i: -bits[i])
    if order == list(range(k)):
        return matrix, bits
    idx = np.arange(1 << k)
    old = np.zeros(1 << k, dtype=np.int64)
    for j, i in enumerate(order):
        old |= ((idx >> (k - 1 - j)) & 1) << (k - 1 - i)
    return np.ascontiguousarray(matrix[np.ix_(old, old)]), tuple(bits[i] for i in order)


def _as_pass(matrix: Matrix, bits: tuple[int, ...], gates: int) -> Pass:
    matrix, bits = descending(matrix, bits)
    if is_diagonal(matrix):
        return Pass("diag", bits, np.ascontiguousarray(np.diag(matrix)), gates)
    return Pass("dense", bits, np.ascontiguousarray(matrix), gates)


def superop_of_gate(u: Matrix) -> Matrix:
    """U (x) conj(U): rho -> U rho U^dagger on the row-then-column vectorisation."""
    return kron_conj(u)


def plan_density(
    ops: tuple[Operation, ...], num_qubits: int, noise: NoiseModel | None, width: int = 2
) -> list[Pass | tuple[str, object]]:
    """Passes for a static circuit on an n-qubit density matrix (vector bits: n+q row, q column).

    Each gate and channel on at most ``width`` qubits becomes a superoperator on the vector bits
    (n+q..., q...) of its qubits; these are fused by the same greedy rule as statevector gates,
    with the width counted in vector bits (2 per qubit). A wider gate becomes two passes, U on its
    row bits then conj(U) on its column bits; a wider channel is returned as
    ("kraus", (channel, qubits)) for the engine's out-of-place Kraus sum."""
    n = num_qubits
    if not 1 <= width <= 2:
        raise ValueError("density-matrix fusion width is 1 or 2 qubits")
    out: list[Pass | tuple[str, object]] = []
    items: list[tuple[Matrix, tuple[int, ...]]] = []

    def vbits(q: tuple[int, ...]) -> tuple[int, ...]:
        return tuple(n + x for x in q) + tuple(q)

    def flush() -> None:
        for blk in _merge(_greedy(items, 2 * width), 2 * width):
            out.append(_as_pass(blk.matrix, tuple(blk.qubits), blk.gates))
        items.clear()

    def channel(ch: Channel, q: tuple[int, ...]) -> None:
        if len(q) <= width:
            items.append((ch.superoperator(), vbits(q)))
        else:
            flush()
            out.append(("kraus", (ch, q)))

    for op in ops:
        if isinstance(op, Gate):
            u = np.asarray(op.matrix, dtype=np.complex128)
            q = tuple(op.qubits)
            if len(q) <= width:
                items.append((superop_of_gate(u), vbits(q)))
            else:
                flush()
                if len(q) > KERNEL_MAX_BITS:
                    raise ValueError(f"{op.name} acts on {len(q)} qubits; gates of up to {KERNEL_MAX_BITS} are supported")
                out.append(_as_pass(u, tuple(n + x for x in q), 1))
                out.append(_as_pass(np.ascontiguousarray(u.conj()), q, 0))
            ch = noise.gate_error(op.name, q) if noise is not None else None
            if ch is not None:
                channel(ch, q)
        elif isinstance(op, Reset):
            channel(RESET_CHANNEL, (op.qubit,))
        elif isinstance(op, (Measure, Barrier, GlobalPhase)):
            continue
        else:
            raise ValueError(f"a density matrix can't run {type(op).__name__} (it needs trajectories)")
    flush()
    return out


RESET_CHANNEL = Channel(
    "reset",
    1,
    (np.array([[1, 0], [0, 0]], dtype=np.complex128), np.array([[0, 1], [0, 0]], dtype=np.complex128)),
)


# ------------------------------------------------------------------------------ tile passes

#: State bits per tile of a tile pass (the kernel's LSIM_TILE_BITS).
TILE_BITS = 12
#: Lowest state bits every tile includes: a tile is read in runs of 2^TILE_LOW contiguous
#: amplitudes (256 bytes in complex128).
TILE_LOW = 4
#: Operators one tile pass may hold (the kernel's LSIM_MAX_TILE_OPS).
TILE_MAX_OPS = 64
#: Default operator budget of a tile pass, in the units of :func:`_tile_cost` (about one
#: 1-qubit dense operator each). Each unit costs one read and write of the tile in LDS; the
#: budget keeps a pass's LDS and FP64 work near the time its sweep of HBM takes.
TILE_BUDGET = 24.0


@dataclass
class TilePass:
    """One sweep of the state in tiles: ``bits`` (ascending) are the tile's state bits (tile bit
    i = state bit bits[i]); ``ops`` are applied to every tile in order, each a :class:`Pass`
    whose ``bits`` are tile bits (first = most significant matrix bit) and whose kind is "dense"
    or "diag"."""

    bits: tuple[int, ...]
    ops: list[Pass]
    gates: int = 0

    kind = "tile"


def _tile_cost(p: Pass) -> float:
    if p.kind == "diag":
        return 0.5 * p.k
    return 1.0 if p.k == 1 else 2.5


def plan_tiles(
    ops: tuple[Operation, ...], num_qubits: int, budget: float = TILE_BUDGET, tile_bits: int = TILE_BITS, low: int = TILE_LOW
) -> list[TilePass | Pass]:
    """Passes for an ideal static circuit as tile passes (statevectors of at least ``tile_bits``
    qubits; the same gates and order as :func:`plan_statevector`, so results agree to rounding),
    planned by the library's native planner (``lsim_host_plan_tiles``).

    Each pass takes the pending gates that can run next on one set of tile bits: the ``low``
    lowest bits, then bits added one at a time (two when a two-qubit gate needs both), each time
    the bit that lets the most pending gates run, then improved by single swaps. A gate can run
    when its qubits are tile bits and every earlier pending gate on its qubits can run. The pass
    is cut where its fused operators would cost more than ``budget`` (:func:`_tile_cost` units)
    or number more than TILE_MAX_OPS. Gates on more than two qubits run as their own dense pass."""
    n = num_qubits
    if n < tile_bits:
        raise ValueError(f"tile passes need at least {tile_bits} qubits")
    items, phase = _static_gates(ops)
    out: list[TilePass | Pass] = []
    if items:
        out = _native_plan("lsim_host_plan_tiles", items, n, float(budget), tile_bits, low, TILE_MAX_OPS)
    if phase != 1:
        for p in out:
            if p.ops if isinstance(p, TilePass) else True:
                target = p.ops[0] if isinstance(p, TilePass) else p
                target.matrix = target.matrix * phase
                break
        else:
            out.append(Pass("scale", (), np.array([phase], dtype=np.complex128), 0))
    return out
