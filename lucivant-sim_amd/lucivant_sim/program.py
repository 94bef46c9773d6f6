"""Compiling circuits into LDS programs: whole circuits run by one workgroup each.

For states that fit a workgroup's LDS (``max_qubits``: 13 qubits complex128 on gfx950), a circuit
instance -- one circuit of a batch, or one trajectory shot -- runs from |0...0> to its outputs in
a single launch of ``lsim_program_run`` with no HBM traffic in between and no host round trips.
This module turns a :class:`~lucivant_sim.circuit.Circuit` (plus noise) into the op table the
kernel interprets (see include/lsim.h), and runs batches of them.

Channel decisions are those of the reference (engines.py:159-205): the mixture's cumulative
probabilities are summed here in the reference's order and passed as float64; Kraus weights,
measurement probabilities and collapses are computed in the kernel in float64.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from lucivant_sim import _native as nat
from lucivant_sim.circuit import Circuit
from lucivant_sim.engines import ptr, require_device, stream_handle, to_device
from lucivant_sim.noise import NoiseModel
from lucivant_sim.plan import is_diagonal

from lucivant_sim.optable import (  # noqa: F401  (re-exported)
    MAX_CLBITS,
    MAX_DENSE_QUBITS,
    MAX_FLAGS,
    OP_COND,
    OP_DENSE,
    OP_DIAG,
    OP_DTYPE,
    OP_KRAUS,
    OP_MEASURE,
    OP_MIX,
    OP_PHASE,
    OP_RESET,
    ProgOp,
    Unsupported,
    static_table,
)


_MAX_QUBITS: dict[int, int] = {}


def max_qubits(dtype: int = nat.C128) -> int:
    """Widest state the program kernel holds in one workgroup's LDS."""
    if dtype not in _MAX_QUBITS:
        require_device()
        _MAX_QUBITS[dtype] = int(nat.load().lsim_program_max_qubits(dtype))
    return _MAX_QUBITS[dtype]


@dataclass
class Program:
    """One circuit's op table with its matrix pool (complex128) and float64 side data.
    ``draws`` is the number of draws one run consumes."""

    n: int
    ops: list[tuple] = field(default_factory=list)
    mats: list[np.ndarray] = field(default_factory=list)
    reals: list[float] = field(default_factory=list)
    mat_size: int = 0
    draws: int = 0
    #: prebuilt op table and matrix pool (set by compile_static's vectorised path)
    table: np.ndarray | None = None
    pool: np.ndarray | None = None

    def add_matrix(self, m: np.ndarray) -> int:
        off = self.mat_size
        flat = np.ascontiguousarray(m, dtype=np.complex128).reshape(-1)
        self.mats.append(flat)
        self.mat_size += flat.size
        return off

    def op(self, kind: int, qubits=(), *, mat=0, nb=0, aux=-1, draw=0, clbit=-1, guard=(0, 0), flag=0, negate=0, cmask=0, cpattern=0) -> None:  # type: ignore[no-untyped-def]
        if len(qubits) > 5 or (kind in (OP_DENSE, OP_MIX) and len(qubits) > MAX_DENSE_QUBITS):
            raise Unsupported(f"dense operators on more than {MAX_DENSE_QUBITS} qubits")
        t = tuple(qubits) + (0,) * (5 - len(qubits))
        self.ops.append((kind, len(qubits), flag, negate, t, (0, 0, 0), mat, nb, aux, draw, clbit, guard[0], guard[1], cmask, cpattern))


def _gate(prog: Program, matrix: np.ndarray, qubits: tuple[int, ...], guard: tuple[int, int]) -> None:
    m = np.asarray(matrix, dtype=np.complex128)
    if is_diagonal(m):
        prog.op(OP_DIAG, qubits, mat=prog.add_matrix(np.diag(m)), guard=guard)
    else:
        prog.op(OP_DENSE, qubits, mat=prog.add_matrix(m), guard=guard)


def compile_static(circuit: Circuit, fuse_width: int = 0) -> Program:
    """An ideal static circuit (measurements deferred): gates only, optionally fused.

    Unfused programs are built with array operations over all gates at once (a few dozen NumPy
    calls per circuit whatever its length), which keeps host time per batched circuit small."""
    if not fuse_width:
        return _compile_static_vectorised(circuit)
    from lucivant_sim.plan import plan_statevector

    prog = Program(circuit.num_qubits)
    for p in plan_statevector(circuit.operations, circuit.num_qubits, fuse_width):
        if p.kind == "scale":
            prog.op(OP_PHASE, (), mat=prog.add_matrix(p.matrix))
        elif p.kind == "diag":
            prog.op(OP_DIAG, p.bits, mat=prog.add_matrix(p.matrix))
        else:
            prog.op(OP_DENSE, p.bits, mat=prog.add_matrix(p.matrix))
    return prog


def _compile_static_vectorised(circuit: Circuit) -> Program:
    table, pool = static_table(circuit)
    prog = Program(circuit.num_qubits)
    prog.table, prog.pool, prog.mat_size = table, pool, pool.size
    return prog


def compile_trajectory(circuit: Circuit, noise: NoiseModel | None) -> Program:
    """A dynamic or noisy circuit as one shot's program (draw-balanced if/else arms only)."""
    from lucivant_sim.trajectories import compile_program

    if circuit.num_clbits > MAX_CLBITS:
        raise Unsupported(f"more than {MAX_CLBITS} classical bits")
    steps, draws, nflags = compile_program(circuit, noise)
    if nflags > MAX_FLAGS:
        raise Unsupported(f"more than {MAX_FLAGS} if/else blocks")
    prog = Program(circuit.num_qubits, draws=draws)
    # a noise model attaches the same channel object to every gate it matches: its tables (the
    # cumulative weights and unitaries of a mixture, the Kraus stack) go into the program once and
    # every op of that channel points at them
    shared: dict[int, tuple[object, int]] = {}
    for s in steps:
        mask = value = 0
        for flag, polarity in s.guard:
            mask |= 1 << flag
            if polarity:
                value |= 1 << flag
        guard = (mask, value)
        if s.kind == "gate":
            _gate(prog, s.matrix, s.qubits, guard)  # type: ignore[arg-type]
        elif s.kind == "phase":
            prog.op(OP_PHASE, (), mat=prog.add_matrix(np.array([np.exp(1j * s.angle)])), guard=guard)
        elif s.kind == "channel":
            ch = s.channel
            assert ch is not None
            known = shared.get(id(ch))
            if ch.mixture is not None:
                if known is None:
                    aux = len(prog.reals)
                    acc = 0.0
                    offsets = []
                    for p, unitary in ch.mixture:
                        acc += p
                        prog.reals.append(acc)
                        identity = np.allclose(unitary, np.eye(unitary.shape[0]))
                        offsets.append(-1.0 if identity else float(prog.add_matrix(unitary)))
                    prog.reals.extend(offsets)
                    known = shared[id(ch)] = (ch, aux)
                prog.op(OP_MIX, s.qubits, nb=len(ch.mixture), aux=known[1], draw=s.column, guard=guard)
            else:
                if len(s.qubits) > 2:
                    raise Unsupported("Kraus channels on more than 2 qubits")
                if known is None:
                    known = shared[id(ch)] = (ch, prog.add_matrix(np.stack(ch.kraus)))
                prog.op(OP_KRAUS, s.qubits, mat=known[1], nb=len(ch.kraus), draw=s.column, guard=guard)
        elif s.kind == "measure":
            aux = -1
            if s.readout is not None:
                aux = len(prog.reals)
                prog.reals.extend([s.readout.p01, s.readout.p10])
            prog.op(OP_MEASURE, s.qubits, aux=aux, draw=s.column, clbit=-1 if s.clbit is None else s.clbit, guard=guard)
        elif s.kind == "reset":
            prog.op(OP_RESET, s.qubits, draw=s.column, guard=guard)
        elif s.kind == "cond":
            c = s.condition
            assert c is not None
            if len(set(c.clbits)) != len(c.clbits) or c.value < 0 or c.value >= (1 << len(c.clbits)):
                # a value the bits can't hold never matches; repeated bits need the general rule
                if len(set(c.clbits)) != len(c.clbits):
                    raise Unsupported("a condition repeats a classical bit")
                cmask, cpattern = 0, 1  # (x & 0) == 1 never holds
            else:
                cmask = cpattern = 0
                for i, cb in enumerate(c.clbits):
                    cmask |= 1 << cb
                    if (c.value >> i) & 1:
                        cpattern |= 1 << cb
            prog.op(OP_COND, (), flag=s.flag, negate=int(c.negate), cmask=cmask, cpattern=cpattern, guard=guard)
    return prog


@dataclass
class BatchOutputs:
    clbits: np.ndarray | None = None
    probs: np.ndarray | None = None
    state: np.ndarray | None = None
    expect: np.ndarray | None = None
    #: [instances, D, 2]: per draw column, the probability a state-dependent decision compared its
    #: draw with and the branch it took (NaN where the column decides nothing state-dependent)
    decisions: np.ndarray | None = None


def run_programs(
    programs: list[Program],
    *,
    repeat: int = 1,
    dtype: int = nat.C128,
    draws: np.ndarray | None = None,
    want_clbits: bool = False,
    want_probs: bool = False,
    want_state: bool = False,
    pauli: list[list[tuple[int, int]]] | None = None,
    want_decisions: bool = False,
) -> BatchOutputs:
    """Run every program (each ``repeat`` times, e.g. once per shot) as one launch.

    ``draws``: float64 [instances, D] with instance i = program i // repeat's run i % repeat.
    ``pauli``: per program, the (x, z) masks of the Pauli strings whose expectations to return
    (every program the same number of strings)."""
    if not programs:
        return BatchOutputs()
    n = programs[0].n
    if any(p.n != n for p in programs):
        raise ValueError("programs of one launch must have the same width")
    table = np.zeros(sum(len(p.table) if p.table is not None else len(p.ops) for p in programs), dtype=OP_DTYPE)
    starts, counts = [], []
    mats, reals = [], []
    mat_base = real_base = row = 0
    for p in programs:
        if p.table is not None:
            arr = p.table.copy()
            arr["mat"] += mat_base
            table[row : row + len(arr)] = arr
            starts.append(row)
            counts.append(len(arr))
            row += len(arr)
            mats.append(p.pool)
            reals.append(np.zeros(0))
            mat_base += p.mat_size
            continue
        if p.ops:
            arr = np.array(p.ops, dtype=OP_DTYPE)
            # rebase offsets into the concatenated pools
            arr["mat"] += mat_base
            uses_reals = np.isin(arr["kind"], [OP_MIX, OP_MEASURE]) & (arr["aux"] >= 0)
            arr["aux"] = np.where(uses_reals, arr["aux"] + real_base, arr["aux"])
            table[row : row + len(arr)] = arr
        starts.append(row)
        counts.append(len(p.ops))
        row += len(p.ops)
        mats.extend(p.mats)
        rp = np.asarray(p.reals, dtype=np.float64)
        if p.reals:
            # MIX branch offsets stored as reals point into this program's matrix pool
            rp = _rebase_mix_offsets(p, rp, mat_base)
        reals.append(rp)
        mat_base += p.mat_size
        real_base += len(p.reals)
    mat_host = np.concatenate(mats) if mats else np.zeros(0, dtype=np.complex128)
    real_host = np.concatenate(reals) if reals else np.zeros(0)
    return run_table(
        n,
        table,
        mat_host,
        np.asarray(starts, dtype=np.int64),
        np.asarray(counts, dtype=np.int64),
        reals=real_host,
        repeat=repeat,
        dtype=dtype,
        draws=draws,
        want_clbits=want_clbits,
        want_probs=want_probs,
        want_state=want_state,
        pauli=pauli,
        want_decisions=want_decisions,
    )


def run_table(
    n: int,
    table: np.ndarray,
    pool: np.ndarray,
    starts: np.ndarray,
    counts: np.ndarray,
    *,
    reals: np.ndarray | None = None,
    repeat: int = 1,
    dtype: int = nat.C128,
    draws: np.ndarray | None = None,
    want_clbits: bool = False,
    want_probs: bool = False,
    want_state: bool = False,
    pauli: list[list[tuple[int, int]]] | None = None,
    want_decisions: bool = False,
    timer: list | None = None,
) -> BatchOutputs:
    """:func:`run_programs` over programs already laid end to end: program j is rows
    ``starts[j] : starts[j] + counts[j]`` of ``table``, whose matrix offsets index ``pool``
    (``optable.static_tables`` builds this form directly). ``timer`` (a list) receives a pair of
    timing events recorded on the launch stream right around the kernel launch."""
    programs = len(starts)
    if programs == 0:
        return BatchOutputs()
    if n > max_qubits(dtype):
        raise Unsupported(f"{n} qubits don't fit one workgroup's LDS")
    device = require_device()
    mat_base = int(pool.size)
    instances = programs * repeat
    if draws is not None and draws.size and draws.shape[0] != instances:
        raise ValueError(f"draws has {draws.shape[0]} rows for {instances} instances")
    if draws is not None and draws.size >= 1 << 31:
        # the kernel indexes draws with 32-bit offsets (draw_start[b] + column)
        raise ValueError(f"{draws.size} draws in one launch exceed the kernel's 32-bit draw offsets; launch fewer instances")
    if pauli is not None:
        if len(pauli) != programs or len({len(t) for t in pauli}) > 1:
            raise ValueError("give every program the same number of Pauli strings")
        if any(x >> n or z >> n for terms in pauli for x, z in terms):
            raise ValueError("a Pauli mask names a qubit beyond the state")
    if table.dtype != OP_DTYPE or pool.dtype != np.complex128:
        raise ValueError("the op table must have OP_DTYPE rows and the pool complex128 entries")
    if np.any(starts < 0) or np.any(counts < 0) or np.any(starts + counts > len(table)):
        raise ValueError("a program's rows lie outside the op table")
    op_start = np.repeat(starts.astype(np.int32), repeat)
    op_count = np.repeat(counts.astype(np.int32), repeat)
    mat_host = pool if pool.size else np.zeros(1, dtype=np.complex128)
    real_host = reals if reals is not None and reals.size else np.zeros(0)
    ndraws = draws.shape[1] if draws is not None and draws.size else 0
    _validate(table, n, mat_base, real_host, ndraws)
    if real_host.size == 0:
        real_host = np.zeros(1)

    def up(a: np.ndarray) -> torch.Tensor:
        return to_device(a)

    d_ops = up(table.view(np.uint8)) if table.size else up(np.zeros(ctypes.sizeof(ProgOp), dtype=np.uint8))
    d_start, d_count = up(op_start), up(op_count)
    d_mats, d_reals = up(mat_host), up(real_host)
    d_draws = d_dstart = None
    if draws is not None and draws.size:
        d_draws = up(draws.astype(np.float64).reshape(-1))
        d_dstart = up((np.arange(instances, dtype=np.int64) * draws.shape[1]).astype(np.int32))
    out = BatchOutputs()
    size = 1 << n
    clb = torch.empty(instances, dtype=torch.int64, device=device) if want_clbits else None
    prb = torch.empty(instances * size, dtype=torch.float64, device=device) if want_probs else None
    sdt = torch.complex128 if dtype == nat.C128 else torch.complex64
    stt = torch.empty(instances * size, dtype=sdt, device=device) if want_state else None
    nterms = len(pauli[0]) if pauli else 0
    px = pz = ex = None
    if nterms:
        masks = np.array(pauli, dtype=np.uint64).reshape(programs, nterms, 2)
        masks = np.repeat(masks, repeat, axis=0)
        px = up(np.ascontiguousarray(masks[:, :, 0]))
        pz = up(np.ascontiguousarray(masks[:, :, 1]))
        ex = torch.empty(instances * nterms, dtype=torch.float64, device=device)
    dec = None
    if want_decisions and d_draws is not None:
        dec = torch.full((instances * ndraws * 2,), float("nan"), dtype=torch.float64, device=device)
    if timer is not None:
        timer.append(torch.cuda.Event(enable_timing=True))
        timer[-1].record(torch.cuda.current_stream())
    nat.call(
        "lsim_program_run",
        dtype,
        n,
        instances,
        ptr(d_ops),
        ptr(d_start),
        ptr(d_count),
        ptr(d_mats),
        ptr(d_reals),
        ptr(d_draws) if d_draws is not None else None,
        ptr(d_dstart) if d_dstart is not None else None,
        ptr(clb) if clb is not None else None,
        ptr(prb) if prb is not None else None,
        ptr(stt) if stt is not None else None,
        nterms,
        ptr(px) if px is not None else None,
        ptr(pz) if pz is not None else None,
        ptr(ex) if ex is not None else None,
        ptr(dec) if dec is not None else None,
        stream_handle(),
    )
    if timer is not None:
        timer.append(torch.cuda.Event(enable_timing=True))
        timer[-1].record(torch.cuda.current_stream())
    if clb is not None:
        out.clbits = clb.cpu().numpy()
    if prb is not None:
        out.probs = prb.cpu().numpy().reshape(instances, size)
    if stt is not None:
        out.state = stt.cpu().numpy().reshape(instances, size)
    if ex is not None:
        out.expect = ex.cpu().numpy().reshape(instances, nterms)
    if dec is not None:
        out.decisions = dec.cpu().numpy().reshape(instances, ndraws, 2)
    return out


def run_tables(n: int, tables: list[np.ndarray], pools: list[np.ndarray], **kwargs) -> BatchOutputs:  # type: ignore[no-untyped-def]
    """:func:`run_table` over per-program op tables, each with its own matrix pool (as
    ``optable.static_table`` returns them)."""
    rows = np.fromiter(map(len, tables), dtype=np.int64, count=len(tables))
    sizes = np.fromiter((p.size for p in pools), dtype=np.int64, count=len(pools))
    table = np.concatenate(tables) if tables else np.zeros(0, dtype=OP_DTYPE)
    table["mat"] += np.repeat(np.cumsum(sizes) - sizes, rows).astype(np.int32)
    pool = np.concatenate(pools) if pools else np.zeros(0, dtype=np.complex128)
    return run_table(n, table, pool, np.cumsum(rows) - rows, rows, **kwargs)


def _validate(table: np.ndarray, n: int, mat_size: int, reals: np.ndarray, ndraws: int) -> None:
    """Every index the kernel will follow stays inside its buffer (checked before launch)."""
    if not table.size:
        return
    real_size = reals.size
    kind, k = table["kind"], table["k"].astype(np.int64)
    if np.any(kind > OP_PHASE):
        raise ValueError("unknown op kind")
    if np.any(k > np.minimum(5, n)):
        raise ValueError("an operator is wider than 5 qubits or than the state")
    if np.any(np.isin(kind, [OP_DENSE, OP_MIX]) & (k > MAX_DENSE_QUBITS)):
        raise ValueError(f"a dense operator is wider than {MAX_DENSE_QUBITS} qubits")
    targeted = np.isin(kind, [OP_DENSE, OP_DIAG, OP_MIX, OP_KRAUS, OP_MEASURE, OP_RESET])
    if np.any(targeted & (k < 1)) or np.any(np.isin(kind, [OP_MEASURE, OP_RESET]) & (k != 1)):
        raise ValueError("an operator has no target qubit (or a measurement more than one)")
    cols = np.arange(5)[None, :]
    used = cols < k[:, None]
    if np.any(np.where(used, table["t"], 0) >= n):
        raise ValueError("an operator targets a qubit beyond the state")
    d = 1 << k
    span = np.select([kind == OP_DENSE, kind == OP_DIAG, kind == OP_KRAUS, kind == OP_PHASE], [d * d, d, table["nb"] * d * d, 1], 0)
    if np.any((span > 0) & ((table["mat"] < 0) | (table["mat"] + span > mat_size))):
        raise ValueError("an op's matrix lies outside the matrix pool")
    if np.any((kind == OP_KRAUS) & (k > 2)):
        raise ValueError("Kraus channels are limited to 2 qubits")
    mix = kind == OP_MIX
    if np.any(mix & ((table["aux"] < 0) | (table["aux"] + 2 * table["nb"] > real_size))):
        raise ValueError("a mixture's probabilities lie outside the reals")
    rows = np.flatnonzero(mix)
    if rows.size:
        # the ops of one channel share its reals: check each distinct (aux, nb, 4^k) once
        trip = np.stack([table["aux"][rows].astype(np.int64), table["nb"][rows].astype(np.int64), d[rows] * d[rows]], axis=1)
        for aux, nb, dd in np.unique(trip, axis=0).tolist():
            offs = reals[aux + nb : aux + 2 * nb]
            if np.any((offs >= 0) & (offs + dd > mat_size)) or np.any((offs < 0) & (offs != -1)):
                raise ValueError("a mixture branch lies outside the matrix pool")
    meas = kind == OP_MEASURE
    if np.any(meas & (table["aux"] >= 0) & (table["aux"] + 2 > real_size)):
        raise ValueError("a readout error lies outside the reals")
    drawing = mix | (kind == OP_KRAUS) | meas | (kind == OP_RESET)
    extra = (meas & (table["aux"] >= 0)).astype(np.int64)
    if np.any(drawing & ((table["draw"] < 0) | (table["draw"] + extra >= ndraws))):
        raise ValueError("an op reads a draw beyond the instance's draws")
    if np.any(meas & (table["clbit"] >= 64)) or np.any((kind == OP_COND) & (table["flag"] >= 32)):
        raise ValueError("classical bit or flag out of range")


def _rebase_mix_offsets(prog: Program, reals: np.ndarray, mat_base: int) -> np.ndarray:
    """Shift the matrix offsets a MIX op stores among its reals (-1 stays -1)."""
    out = reals.copy()
    done = set()  # ops of one channel share its reals: each segment moves once
    for o in prog.ops:
        if o[0] == OP_MIX and o[8] not in done:
            aux, nb = o[8], o[7]
            done.add(aux)
            seg = out[aux + nb : aux + 2 * nb]
            out[aux + nb : aux + 2 * nb] = np.where(seg >= 0, seg + mat_base, seg)
    return out
