"""NumPy restatement of the reference engines and drivers (test infrastructure only).

Follows lucivant-sim 0.3.0:
- state layout, gate contraction: src/lucivant_sim/engines.py:8-9, 38-42, 138-157;
- trajectory channel choice, measurement, collapse, reset: engines.py:159-205;
- density matrix, channels, analytic depolarising: engines.py:122-135, 215-264;
- marginal: engines.py:267-275;
- static evolution, sampling, trajectories, expectation: simulator.py:229-317, 489-533;
- recorded outcome probabilities with readout error: simulator.py:467-486.

The circuit comes in as lucivant_sim's Circuit IR (plain host data: gates already resolved to
matrices). Every state update here is written independently of the device engine: tensordot on
the [2]*n tensor, exactly as the reference contracts it, with no fusion.
"""

from __future__ import annotations

import itertools

import numpy as np

PAULI = {
    "I": np.eye(2, dtype=np.complex128),
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


# --------------------------------------------------------------------------- gate application


def contract(tensor: np.ndarray, matrix: np.ndarray, axes: list[int]) -> np.ndarray:
    """U on tensor axes ``axes`` (first = matrix MSB): engines.py:38-42."""
    k = len(axes)
    out = np.tensordot(np.asarray(matrix).reshape([2] * (2 * k)), tensor, axes=(list(range(k, 2 * k)), axes))
    return np.moveaxis(out, list(range(k)), axes)


def apply_sv(psi: np.ndarray, matrix: np.ndarray, qubits: tuple[int, ...], n: int) -> np.ndarray:
    """psi <- U psi on ``qubits`` (qubit q = axis n-1-q = index bit q)."""
    if n == 0:
        return psi * np.asarray(matrix).reshape(-1)[0]
    t = psi.reshape([2] * n)
    return contract(t, matrix, [n - 1 - q for q in qubits]).reshape(-1)


def full_operator(matrix: np.ndarray, qubits: tuple[int, ...], n: int) -> np.ndarray:
    """The 2^n x 2^n operator of a k-qubit matrix on ``qubits``, built entry by entry:
    <i|O|j> = U[r(i), r(j)] when i and j agree outside ``qubits``, where r(x) reads x's bits at
    ``qubits`` with the first qubit most significant. The brute-force check of ``apply_sv``."""
    d = 1 << n
    k = len(qubits)
    idx = np.arange(d)
    r = np.zeros(d, dtype=np.int64)
    mask = 0
    for i, q in enumerate(qubits):
        r |= ((idx >> q) & 1) << (k - 1 - i)
        mask |= 1 << q
    same = (idx[:, None] & ~mask) == (idx[None, :] & ~mask)
    return np.where(same, np.asarray(matrix)[r[:, None], r[None, :]], 0)


# ------------------------------------------------- the reference's dispatch on large states
#
# The reference applies gates through a size-dependent dispatch (engines.py:45-119) rather than
# always through the tensordot above. Its results equal the contraction's up to rounding; it is
# restated here because it sets the reference's CPU cost, which the CPU baseline times.

#: state sizes (entries) from which slices replace the contraction: permutation-with-phase
#: gates from 2^14, general matrices from 2^18 (engines.py:20-24)
SLICE_FROM_MOVES = 1 << 14
SLICE_FROM_GENERAL = 1 << 18
#: a general one-qubit gate is one batched 2x2 product when at least this many entries lie below
#: its axis (engines.py:25-27)
BATCHED_FROM_INNER = 8


def _is_permutation_with_phases(m: np.ndarray) -> bool:
    nz = m != 0
    return bool(np.all(nz.sum(axis=0) == 1) and np.all(nz.sum(axis=1) == 1))


def apply_slices(tensor: np.ndarray, matrix: np.ndarray, axes: list[int]) -> np.ndarray:
    """U on ``axes`` of the [2]*n tensor, choosing the reference's route (engines.py:45-119):

    - no axes, more than two, or fewer than 2^14 entries: the contraction;
    - a one-qubit gate that is not a permutation with phases, with at least 8 entries below its
      axis: one batched product of U with the (outer, 2, inner) view;
    - below 2^14 entries (permutations) or 2^18 (general): the contraction;
    - otherwise the 2^k slices of the tensor cut by the gate's axes (first axis = matrix MSB)
      are rewritten in place: a permutation with phases moves (and scales) slices, a general
      matrix forms each new slice as the U-weighted sum of copies of the old ones, skipping
      zero entries (an all-zero row clears its slice).

    Returns the updated tensor (the input itself on the in-place route)."""
    k = len(axes)
    if k == 0 or k > 2 or tensor.size < SLICE_FROM_MOVES:
        return contract(tensor, matrix, axes)
    u = np.asarray(matrix, dtype=np.complex128)
    moves = _is_permutation_with_phases(u)
    if k == 1 and not moves:
        below = 1 << (tensor.ndim - 1 - axes[0])
        if below >= BATCHED_FROM_INNER:
            blocks = np.matmul(u, tensor.reshape(-1, 2, below))
            return blocks.reshape(tensor.shape)
    if tensor.size < (SLICE_FROM_MOVES if moves else SLICE_FROM_GENERAL):
        return contract(tensor, matrix, axes)
    t = tensor if tensor.flags.c_contiguous else np.ascontiguousarray(tensor)
    # a view with the gate's axes leading, first axis first: view[b0, b1] is a writable slice
    view = np.moveaxis(t, axes, list(range(k)))
    d = 1 << k

    def part(r: int) -> np.ndarray:
        return view[tuple((r >> (k - 1 - i)) & 1 for i in range(k))]

    if moves:
        source = [int(np.flatnonzero(u[r])[0]) for r in range(d)]
        kept = {c: part(c).copy() for r, c in enumerate(source) if c != r}
        for r, c in enumerate(source):
            w = u[r, c]
            if c == r:
                if w != 1:
                    part(r)[...] *= w
            elif w == 1:
                part(r)[...] = kept[c]
            else:
                np.multiply(kept[c], w, out=part(r))
        return t
    old = [part(c).copy() for c in range(d)]
    for r in range(d):
        dst = part(r)
        cols = [c for c in range(d) if u[r, c] != 0]
        if not cols:
            dst[...] = 0
            continue
        np.multiply(old[cols[0]], u[r, cols[0]], out=dst)
        for c in cols[1:]:
            dst += u[r, c] * old[c]
    return t


def fuse_one_qubit_runs(ops):  # type: ignore[no-untyped-def]
    """Consecutive one-qubit gates on a qubit become one gate, the product of their matrices
    (later @ earlier); a wider gate, a measurement or a reset on the qubit ends its run, and an
    if/else or barrier ends every run (simulator.py:196-226, applied to ideal statevectors)."""
    from lucivant_sim.circuit import Barrier, Gate, IfElse, Measure, Reset

    out: list = []
    open_run: dict[int, int] = {}
    for op in ops:
        if isinstance(op, Gate) and len(op.qubits) == 1:
            q = op.qubits[0]
            if q in open_run:
                at = open_run[q]
                fused = np.asarray(op.matrix, dtype=np.complex128) @ np.asarray(out[at].matrix, dtype=np.complex128)
                out[at] = Gate(name="fused", qubits=(q,), matrix=fused)
            else:
                open_run[q] = len(out)
                out.append(op)
            continue
        if isinstance(op, (IfElse, Barrier)):
            open_run.clear()
        elif isinstance(op, Gate):
            for q in op.qubits:
                open_run.pop(q, None)
        elif isinstance(op, (Measure, Reset)):
            open_run.pop(op.qubit, None)
        out.append(op)
    return out


# ------------------------------------------------------------------------------- statevector


def evolve_sv(circuit) -> np.ndarray:  # type: ignore[no-untyped-def]
    """Ideal static evolution, one contraction per gate (simulator.py:229-251)."""
    from lucivant_sim.circuit import Gate, GlobalPhase

    n = circuit.num_qubits
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    for op in circuit.operations:
        if isinstance(op, Gate):
            psi = apply_sv(psi, op.matrix, op.qubits, n)
        elif isinstance(op, GlobalPhase):
            psi = psi * np.exp(1j * op.angle)
    return psi


def probabilities_sv(psi: np.ndarray) -> np.ndarray:
    return np.abs(psi) ** 2


def pauli_matrix(label: str) -> np.ndarray:
    """Dense matrix of a label whose rightmost letter is qubit 0 (so the leftmost is the MSB)."""
    m = np.array([[1]], dtype=np.complex128)
    for letter in label.upper():
        m = np.kron(m, PAULI[letter])
    return m


def expectation_sv(psi: np.ndarray, label: str, n: int) -> float:
    """<psi|P|psi> by applying the letters one by one (simulator.py:517-527)."""
    probe = psi.copy()
    for q, letter in enumerate(reversed(label.upper())):
        if letter != "I":
            probe = apply_sv(probe, PAULI[letter], (q,), n)
    return float(np.vdot(psi, probe).real)


# --------------------------------------------------------------------------- density matrix


def apply_dm(rho: np.ndarray, matrix: np.ndarray, qubits: tuple[int, ...], n: int) -> np.ndarray:
    """U on the row axes, conj(U) on the column axes (engines.py:233-235)."""
    t = rho.reshape([2] * (2 * n))
    t = contract(t, matrix, [n - 1 - q for q in qubits])
    t = contract(t, np.asarray(matrix).conj(), [2 * n - 1 - q for q in qubits])
    return t.reshape(1 << n, 1 << n)


def depolarizing_trace(rho: np.ndarray, p: float, qubits: tuple[int, ...], n: int) -> np.ndarray:
    """(1 - p) rho + p I/d (x) Tr_targets(rho) (engines.py:122-135)."""
    k = len(qubits)
    d = 1 << k
    rows = [n - 1 - q for q in qubits]
    cols = [2 * n - 1 - q for q in qubits]
    t = rho.reshape([2] * (2 * n))
    moved = np.moveaxis(t, rows + cols, range(2 * k))
    m = moved.reshape((d, d, *moved.shape[2 * k :]))
    partial = np.trace(m, axis1=0, axis2=1)
    out = (1 - p) * m
    for i in range(d):
        out[i, i] += (p / d) * partial
    return np.moveaxis(out.reshape(moved.shape), range(2 * k), rows + cols).reshape(1 << n, 1 << n)


def apply_channel_dm(rho: np.ndarray, channel, qubits: tuple[int, ...], n: int) -> np.ndarray:  # type: ignore[no-untyped-def]
    """engines.py:240-250: the analytic depolariser for k in (1, 2), else the Kraus sum."""
    if channel.analytic_depolarizing and len(qubits) in (1, 2):
        return depolarizing_trace(rho, channel.params[0], qubits, n)
    total = np.zeros_like(rho)
    for k in channel.kraus:
        total = total + apply_dm(rho, k, qubits, n)
    return total


RESET_KRAUS = (np.array([[1, 0], [0, 0]], dtype=np.complex128), np.array([[0, 1], [0, 0]], dtype=np.complex128))


def evolve_dm(circuit, noise) -> np.ndarray:  # type: ignore[no-untyped-def]
    """Static evolution of rho with gate noise and resets (simulator.py:229-251)."""
    from lucivant_sim.circuit import Gate, Reset

    n = circuit.num_qubits
    rho = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    rho[0, 0] = 1
    for op in circuit.operations:
        if isinstance(op, Gate):
            rho = apply_dm(rho, op.matrix, op.qubits, n)
            ch = noise.gate_error(op.name, op.qubits) if noise is not None else None
            if ch is not None:
                rho = apply_channel_dm(rho, ch, op.qubits, n)
        elif isinstance(op, Reset):
            rho = sum(apply_dm(rho, k, (op.qubit,), n) for k in RESET_KRAUS)
    return rho


def probabilities_dm(rho: np.ndarray) -> np.ndarray:
    return np.clip(np.diagonal(rho).real, 0.0, None)


def expectation_dm(rho: np.ndarray, label: str, n: int) -> float:
    """Tr(P rho) by left-multiplying the letters (simulator.py:505-516, 530-533)."""
    t = rho.reshape([2] * (2 * n))
    for q, letter in enumerate(reversed(label.upper())):
        if letter != "I":
            t = contract(t, PAULI[letter], [n - 1 - q])
    return float(np.trace(t.reshape(1 << n, 1 << n)).real)


# ---------------------------------------------------------------------------------- sampling


def marginal(probs: np.ndarray, n: int, qubits: list[int]) -> np.ndarray:
    """engines.py:267-275."""
    keep = sorted(qubits)
    if n == 0:
        return probs
    t = probs.reshape([2] * n)
    drop = tuple(n - 1 - q for q in range(n) if q not in keep)
    return np.asarray(t.sum(axis=drop) if drop else t).reshape(-1)


def sample(circuit, probs: np.ndarray, shots: int, noise, rng: np.random.Generator) -> dict[str, int]:  # type: ignore[no-untyped-def]
    """simulator.py:254-284 (counts from exact probabilities, readout flips)."""
    pairs = circuit.final_measurements()
    if not pairs or shots == 0:
        return {}
    measured = sorted({q for q, _ in pairs})
    pos = {q: i for i, q in enumerate(measured)}
    dist = marginal(probs, circuit.num_qubits, measured)
    cum = np.cumsum(dist)
    cum /= cum[-1]
    idx = np.minimum(np.searchsorted(cum, rng.random(shots), side="right").astype(np.int64), len(dist) - 1)
    bits = np.zeros(shots, dtype=np.int64)
    for q, c in pairs:
        v = (idx >> pos[q]) & 1
        err = noise.readout_error(q) if noise else None
        if err is not None and (err.p01 or err.p10):
            v = v ^ (rng.random(shots) < np.where(v == 1, err.p10, err.p01)).astype(np.int64)
        bits |= v << c
    values, counts = np.unique(bits, return_counts=True)
    return {circuit.format_bits(int(v)): int(k) for v, k in zip(values, counts)}


# ------------------------------------------------------------------------------ trajectories


class _Shot:
    """One trajectory's statevector with the reference's per-shot operations."""

    def __init__(self, n: int) -> None:
        self.n = n
        self.psi = np.zeros(1 << n, dtype=np.complex128)
        self.psi[0] = 1

    def apply(self, m: np.ndarray, qubits: tuple[int, ...]) -> None:
        self.psi = apply_sv(self.psi, m, qubits, self.n)

    def channel(self, ch, qubits: tuple[int, ...], rng: np.random.Generator) -> None:  # type: ignore[no-untyped-def]
        """engines.py:159-183."""
        r = rng.random()
        if ch.mixture is not None:
            acc = 0.0
            for p, u in ch.mixture:
                acc += p
                if r < acc:
                    if not np.allclose(u, np.eye(u.shape[0])):
                        self.apply(u, qubits)
                    return
            return
        acc = 0.0
        cand = self.psi
        for k in ch.kraus:
            cand = apply_sv(self.psi, k, qubits, self.n)
            w = float(np.vdot(cand, cand).real)
            acc += w
            if r < acc and w > 0:
                self.psi = cand / np.sqrt(w)
                return
        norm = float(np.vdot(cand, cand).real)
        if norm > 0:
            self.psi = cand / np.sqrt(norm)

    def p1(self, q: int) -> float:
        one = self.psi[(np.arange(1 << self.n) >> q) & 1 == 1]
        return float(np.vdot(one, one).real)

    def measure(self, q: int, rng: np.random.Generator) -> int:
        """engines.py:189-201."""
        p1 = min(max(self.p1(q), 0.0), 1.0)
        o = int(rng.random() < p1)
        p = p1 if o else 1 - p1
        self.psi[((np.arange(1 << self.n) >> q) & 1) != o] = 0
        if p > 0:
            self.psi = self.psi / np.sqrt(p)
        return o


def recorded_distribution(
    dist: np.ndarray, measured: list[int], pairs: list[tuple[int, int]], readout: dict[int, tuple[float, float]], threshold: float = 1e-12
) -> dict[int, float]:
    """Probabilities of the recorded clbits with readout flips (simulator.py:467-486), from the
    definition rather than the reference's branch expansion: for every true index t of ``dist``
    (bit i = measured[i]) above ``threshold`` and every read pattern r over ``pairs``, the weight
    is dist[t] times, per pair (q, c), 1 - f or f as r_c equals q's true bit or not, where f is
    p01 for a true 0 and p10 for a true 1 of q's readout error (``readout``: qubit -> (p01, p10);
    no entry: no flips). Outcomes at or below ``threshold`` are dropped. Brute force over
    2^m x 2^len(pairs) patterns: small cases only."""
    out: dict[int, float] = {}
    for t, p in enumerate(dist):
        if p <= threshold:
            continue
        truth = {q: (t >> i) & 1 for i, q in enumerate(measured)}
        for r in itertools.product((0, 1), repeat=len(pairs)):
            w = float(p)
            bits = 0
            for (q, c), read in zip(pairs, r):
                p01, p10 = readout.get(q, (0.0, 0.0))
                f = p10 if truth[q] else p01
                w *= f if read != truth[q] else 1.0 - f
                bits |= read << c
            out[bits] = out.get(bits, 0.0) + w
    return {b: w for b, w in sorted(out.items()) if w > threshold}


def trajectory_counts(circuit, noise, shots: int, rng: np.random.Generator) -> dict[str, int]:  # type: ignore[no-untyped-def]
    """The per-shot loop of simulator.py:287-317, 364-371."""
    from lucivant_sim.circuit import Gate, GlobalPhase, IfElse, Measure, Reset

    tally: dict[int, int] = {}
    for _ in range(shots):
        s = _Shot(circuit.num_qubits)
        bits = [0] * circuit.num_clbits

        def run(ops) -> None:  # type: ignore[no-untyped-def]
            for op in ops:
                if isinstance(op, Gate):
                    s.apply(op.matrix, op.qubits)
                    ch = noise.gate_error(op.name, op.qubits) if noise else None
                    if ch is not None:
                        s.channel(ch, op.qubits, rng)
                elif isinstance(op, Measure):
                    o = s.measure(op.qubit, rng)
                    err = noise.readout_error(op.qubit) if noise else None
                    if err is not None and rng.random() < (err.p10 if o else err.p01):
                        o ^= 1
                    if op.clbit is not None:
                        bits[op.clbit] = o
                elif isinstance(op, Reset):
                    if s.measure(op.qubit, rng):
                        s.apply(PAULI["X"], (op.qubit,))
                elif isinstance(op, GlobalPhase):
                    s.psi = s.psi * np.exp(1j * op.angle)
                elif isinstance(op, IfElse):
                    run(op.then if op.condition.holds(bits) else op.otherwise)

        run(circuit.operations)
        v = sum(b << i for i, b in enumerate(bits))
        tally[v] = tally.get(v, 0) + 1
    if not circuit.num_clbits:
        return {}
    return {circuit.format_bits(b): c for b, c in sorted(tally.items())}


def run_counts(circuit, shots: int, seed: int, noise=None, method: str = "statevector") -> dict[str, int]:  # type: ignore[no-untyped-def]
    """Counts of ``run`` for an explicit method (simulator.py:320-393), seeded the same way."""
    rng = np.random.default_rng(seed)
    if method == "trajectories":
        return trajectory_counts(circuit, noise, shots, rng)
    if method == "statevector":
        probs = probabilities_sv(evolve_sv(circuit))
    else:
        probs = probabilities_dm(evolve_dm(circuit, noise))
    return dict(sorted(sample(circuit, probs, shots, noise, rng).items()))


def tvd(counts: dict[str, int], probs: dict[str, float]) -> float:
    total = sum(counts.values())
    keys = set(counts) | set(probs)
    return 0.5 * sum(abs(counts.get(k, 0) / total - probs.get(k, 0.0)) for k in keys)


def all_labels(n: int) -> list[str]:
    return ["".join(s) for s in itertools.product("IXYZ", repeat=n)]
