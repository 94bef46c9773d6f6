"""The statevector and density-matrix engines, resident in MI355X HBM.

Both classes keep the duck-typed engine protocol of lucivant-sim 0.3.0
src/lucivant_sim/engines.py:138-264 (``apply``, ``scale``, ``apply_channel``, ``measure``,
``collapse``, ``reset``, ``probabilities``, ``vector``/``matrix``) so the drivers in
:mod:`lucivant_sim.simulator` read as the reference's do. The state lives in a torch device buffer
(complex128 by default, qubit q = amplitude index bit q); every operation is a call into
liblsim.so, the hand-written gfx950 kernels, on torch's current stream. Only scalars and the
answers the API returns cross back to the host.

:meth:`Statevector.run` and :meth:`DensityMatrix.run` execute a whole planned program (see
:mod:`lucivant_sim.plan`): every pass's matrix is uploaded in one copy, then the passes are
launched back to back with no host synchronisation.
"""

from __future__ import annotations

import collections
import ctypes
import math
import threading
from typing import Sequence

import numpy as np
import torch
from numpy.typing import NDArray

from lucivant_sim import _native as nat
from lucivant_sim.gates import Matrix
from lucivant_sim.noise import Channel
from lucivant_sim.plan import RESET_CHANNEL, Pass, TilePass, descending, is_diagonal, superop_of_gate

#: Memory guards of the reference (engines.py:29-31): 2^26 amplitudes and 4^13 entries, 1 GiB each.
MAX_STATEVECTOR_QUBITS = 26
MAX_DENSITY_QUBITS = 13


class SimulationError(RuntimeError):
    """The circuit can't be simulated as asked (too large, or the method doesn't fit it)."""


def require_device() -> torch.device:
    """The GPU the engines run on; raises when none is visible (there is no CPU fallback)."""
    if not torch.cuda.is_available():
        raise nat.NativeError("no GPU is visible: the lucivant_sim engines run on the MI355X (gfx950) only")
    nat.load()
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


_TORCH_DTYPE = {nat.C128: torch.complex128, nat.C64: torch.complex64}
_NP_DTYPE = {nat.C128: np.complex128, nat.C64: np.complex64}


class Workspace:
    """Device scratch for the reductions, grown on demand and reused; one buffer per stream, so
    reductions queued on different streams never share scratch (a buffer is allocated on, and
    only ever used by, its stream)."""

    def __init__(self) -> None:
        self.bufs: dict[int, torch.Tensor] = {}

    def get(self, nbytes: int) -> torch.Tensor:
        key = torch.cuda.current_stream().cuda_stream
        buf = self.bufs.get(key)
        if buf is None or buf.numel() * 8 < nbytes:
            buf = torch.empty(max(nbytes // 8 + 1, 1024), dtype=torch.float64, device=require_device())
            self.bufs[key] = buf
        return buf


WORKSPACE = Workspace()


class _Staging:
    """Pinned host memory reused by every upload: each upload copies its bytes into the next free
    span of a ring and queues one asynchronous copy from there on the current stream, so it
    neither waits for the work already queued on the stream (a copy from pageable memory would)
    nor allocates pinned memory (which costs about a millisecond when no cached block fits). A
    span is reused only after the event recorded behind its copy has completed."""

    ALIGN = 256

    def __init__(self, nbytes: int) -> None:
        self.size = nbytes
        self.buf: torch.Tensor | None = None
        self.host: np.ndarray | None = None
        self.head = 0
        self.live: collections.deque = collections.deque()  # (start, end, event), oldest first
        self.lock = threading.Lock()

    def _reserve(self, n: int) -> int:
        while True:
            while self.live and self.live[0][2].query():
                self.live.popleft()
            start = -(-self.head // self.ALIGN) * self.ALIGN
            if start + n > self.size:
                start = 0
            if not any(a < start + n and start < b for a, b, _ in self.live):
                self.head = start + n
                return start
            _, _, oldest = self.live.popleft()
            oldest.synchronize()

    def upload(self, host: np.ndarray, device: torch.device) -> torch.Tensor:
        host = np.ascontiguousarray(host)
        n = host.nbytes
        if n == 0 or n > self.size // 4:
            staged = torch.from_numpy(host).pin_memory()
            return staged.to(device, non_blocking=True)
        with self.lock:
            if self.buf is None:
                self.buf = torch.empty(self.size, dtype=torch.uint8, pin_memory=True)
                self.host = self.buf.numpy()
            start = self._reserve(n)
            self.host[start : start + n] = host.reshape(-1).view(np.uint8)  # type: ignore[index]
            dev = torch.empty(n, dtype=torch.uint8, device=device)
            dev.copy_(self.buf[start : start + n], non_blocking=True)
            done = torch.cuda.Event()
            done.record()
            self.live.append((start, start + n, done))
        return dev.view(torch.from_numpy(host[:0].reshape(-1)).dtype).reshape(host.shape)


_STAGING = _Staging(64 << 20)


def to_device(host: np.ndarray) -> torch.Tensor:
    """A device copy of a host array, queued on the current stream behind the work already
    there, without waiting for it (through the pinned staging ring)."""
    return _STAGING.upload(host, require_device())


def upload(arrays: Sequence[np.ndarray], dtype: int = nat.C128) -> tuple[torch.Tensor, list[int]]:
    """Copy several small complex arrays to the device in one transfer; returns the buffer and
    each array's device address."""
    npdt = _NP_DTYPE[dtype]
    flat = [np.ascontiguousarray(a, dtype=npdt).reshape(-1) for a in arrays]
    sizes = [f.size for f in flat]
    host = np.concatenate(flat) if flat else np.zeros(1, dtype=npdt)
    dev = to_device(host)
    base = ptr(dev)
    item = np.dtype(npdt).itemsize
    addrs, off = [], 0
    for n in sizes:
        addrs.append(base + off * item)
        off += n
    return dev, addrs


def run_passes(buf: torch.Tensor, dtype: int, nbits: int, batch: int, passes: Sequence[Pass]) -> None:
    """Launch ``passes`` in order on ``batch`` states of ``nbits`` bits held in ``buf``."""
    if not passes:
        return
    dev, addrs = upload([p.payload() for p in passes], dtype)
    stream = stream_handle()
    state = ptr(buf)
    lib = nat.load()
    for p, addr in zip(passes, addrs):
        fn, args = pass_launch(lib, p, state, dtype, nbits, batch, addr, stream)
        if fn(*args):
            raise nat.NativeError(f"{fn.__name__}: {lib.lsim_last_error().decode(errors='replace')}")
    # keep the matrices alive until the launches that read them have run
    dev.record_stream(torch.cuda.current_stream())


def plan_launches(buf: torch.Tensor, dtype: int, nbits: int, batch: int, plan: Sequence[Pass | TilePass], stream: int) -> tuple[list, list[tuple]]:
    """Every launch of ``plan`` (fused dense and diagonal passes, global phases, tile passes) on
    ``batch`` states of ``nbits`` bits in ``buf``, as (function, arguments) pairs that only
    enqueue work on ``stream``; and the objects the arguments point into, which must stay alive
    until the launches have run. All matrices are uploaded in one copy per kind (dense passes in
    the state's dtype, tile operators in complex128)."""
    lib = nat.load()
    state = ptr(buf)
    keep: list = []
    calls: list[tuple] = []
    flat = [p for p in plan if not isinstance(p, TilePass)]
    tiles = [p for p in plan if isinstance(p, TilePass)]
    addrs: list[int] = []
    if flat:
        dev, addrs = upload([p.payload() for p in flat], dtype)
        keep.append(dev)
    tile_tables = []
    if tiles:
        host = np.concatenate([op.matrix.reshape(-1) for tp in tiles for op in tp.ops] or [np.zeros(1, dtype=np.complex128)])
        tdev = to_device(np.ascontiguousarray(host, dtype=np.complex128))
        keep.append(tdev)
        at = 0
        for tp in tiles:
            table = (nat.TileOp * max(len(tp.ops), 1))()
            for j, op in enumerate(tp.ops):
                table[j].kind = nat.TILE_DIAG if op.kind == "diag" else nat.TILE_DENSE
                table[j].k = op.k
                table[j].t[0] = op.bits[0]
                table[j].t[1] = op.bits[1] if op.k == 2 else 0
                table[j].mat = at
                at += op.matrix.size
            bits = nat.int_array(tp.bits)
            keep += [table, bits]
            tile_tables.append((bits, len(tp.ops), ctypes.addressof(table), ptr(tdev)))
    fi = ti = 0
    for p in plan:
        if isinstance(p, TilePass):
            bits, nops, table, mats = tile_tables[ti]
            ti += 1
            calls.append((lib.lsim_sv_apply_tile, (state, dtype, nbits, batch, bits, nops, table, mats, stream)))
            continue
        addr = addrs[fi]
        fi += 1
        fn, args = pass_launch(lib, p, state, dtype, nbits, batch, addr, stream)
        keep.append(args)
        calls.append((fn, args))
    return keep, calls


def pass_launch(lib, p: Pass, state: int, dtype: int, nbits: int, batch: int, addr: int, stream: int) -> tuple:  # type: ignore[no-untyped-def]
    """The entry point and arguments of one pass over ``batch`` states (``addr``: the device
    copy of ``p.payload()``): a scale, a diagonal, a permutation with phases, or a dense
    operator. The host arrays among the arguments must stay alive until the call."""
    if p.kind == "scale":
        z = complex(p.matrix[0])
        return lib.lsim_sv_scale, (state, dtype, nbits, batch, z.real, z.imag, stream)
    targets = nat.int_array(p.bits)
    if p.kind == "diag":
        return lib.lsim_sv_apply_diag, (state, dtype, nbits, batch, p.k, targets, addr, None, stream)
    mv = p.moves()
    if mv is not None:
        src, _entries, unit = mv
        return lib.lsim_sv_apply_perm, (state, dtype, nbits, batch, p.k, targets, nat.int_array(src), unit, addr, None, stream)
    return lib.lsim_sv_apply_dense, (state, dtype, nbits, batch, p.k, targets, addr, None, stream)


def launch_all(calls: list[tuple]) -> None:
    """Issue launches made by :func:`plan_launches`; raises NativeError on the first failure."""
    lib = nat.load()
    for fn, args in calls:
        if fn(*args):
            raise nat.NativeError(f"{fn.__name__}: {lib.lsim_last_error().decode(errors='replace')}")


def run_plan(buf: torch.Tensor, dtype: int, nbits: int, batch: int, plan: Sequence[Pass | TilePass]) -> None:
    """Launch a plan that mixes fused dense passes and tile passes (:func:`plan.plan_tiles`), in
    order, on ``batch`` states of ``nbits`` bits in ``buf``."""
    if not any(isinstance(p, TilePass) for p in plan):
        run_passes(buf, dtype, nbits, batch, plan)  # type: ignore[arg-type]
        return
    keep, calls = plan_launches(buf, dtype, nbits, batch, plan, stream_handle())
    launch_all(calls)
    for t in keep:
        if isinstance(t, torch.Tensor):
            t.record_stream(torch.cuda.current_stream())


def _single(matrix: Matrix, bits: tuple[int, ...]) -> Pass:
    m, bits = descending(np.asarray(matrix, dtype=np.complex128), tuple(bits))
    if is_diagonal(m):
        return Pass("diag", bits, np.ascontiguousarray(np.diag(m)))
    return Pass("dense", bits, np.ascontiguousarray(m))


def pauli_masks(label: str) -> tuple[int, int]:
    """(x, z) bit masks of a Pauli label whose rightmost letter is qubit 0."""
    x = z = 0
    for q, letter in enumerate(reversed(label.upper())):
        if letter in "XY":
            x |= 1 << q
        if letter in "ZY":
            z |= 1 << q
    return x, z


class _DeviceState:
    nbits: int
    dtype: int
    buffer: torch.Tensor

    def _stream(self) -> int:
        return stream_handle()

    def device_probabilities(self) -> torch.Tensor:
        raise NotImplementedError

    def probabilities(self) -> NDArray[np.float64]:
        """The measurement distribution of every basis state, as float64 on the host."""
        return self.device_probabilities().cpu().numpy()

    def marginal(self, qubits: Sequence[int]) -> NDArray[np.float64]:
        """Probabilities of ``qubits`` alone (index bit i = the i-th smallest of them): the
        reference's marginal (engines.py:267-275). When they are not every qubit, the others are
        summed out on the device in NumPy's own order (bit-identical to the host marginal), and
        only the 2^m results are copied to the host."""
        keep = sorted(set(int(q) for q in qubits))
        probs = self.device_probabilities()
        if len(keep) == self.n or not device_marginal_pays(self.n, keep):
            return marginal(probs.cpu().numpy(), self.n, keep)
        return device_marginal(probs, self.n, keep, self._stream())

    def run(self, passes: Sequence[Pass | TilePass]) -> None:
        run_plan(self.buffer, self.dtype, self.nbits, 1, passes)

    def _reduced(self, bits: tuple[int, ...]) -> NDArray[np.complex128]:
        k = len(bits)
        out = torch.empty(1 << (2 * k), dtype=torch.complex128, device=self.buffer.device)
        ws = WORKSPACE.get(nat.load().lsim_reduce_workspace_bytes(self.nbits, 1, k))
        nat.call(
            "lsim_sv_reduced_dm",
            ptr(self.buffer),
            self.dtype,
            self.nbits,
            1,
            k,
            nat.int_array(bits),
            ptr(out),
            ptr(ws),
            self._stream(),
        )
        return out.cpu().numpy().reshape(1 << k, 1 << k)


class Statevector(_DeviceState):
    """2^n amplitudes in HBM, starting in |0...0>."""

    def __init__(self, num_qubits: int, *, dtype: int = nat.C128) -> None:
        if num_qubits > MAX_STATEVECTOR_QUBITS:
            raise SimulationError(
                f"{num_qubits} qubits need more memory than the statevector engine allows "
                f"(up to {MAX_STATEVECTOR_QUBITS})"
            )
        device = require_device()
        self.n = self.nbits = num_qubits
        self.dtype = dtype
        self.buffer = torch.empty(1 << num_qubits, dtype=_TORCH_DTYPE[dtype], device=device)
        nat.call("lsim_sv_set_basis", ptr(self.buffer), dtype, num_qubits, 1, 0, self._stream())

    # -- the engine protocol (engines.py:150-212) ---------------------------------------------

    def apply(self, matrix: Matrix, qubits: tuple[int, ...]) -> None:
        if not qubits:
            self.scale(complex(np.asarray(matrix).reshape(-1)[0]))
            return
        if len(qubits) > nat.MAX_K:
            raise SimulationError(f"gates on more than {nat.MAX_K} qubits aren't supported by the device engine")
        self.run([_single(matrix, tuple(qubits))])

    def scale(self, factor: complex) -> None:
        f = complex(factor)
        nat.call("lsim_sv_scale", ptr(self.buffer), self.dtype, self.nbits, 1, f.real, f.imag, self._stream())

    def apply_channel(self, channel: Channel, qubits: tuple[int, ...], rng: np.random.Generator) -> None:
        """One branch of the channel with its Born probability, consuming exactly one draw
        (engines.py:159-183): a mixture picks by cumulative probability (the identity branch
        leaves the state alone); Kraus operators take the first branch whose cumulative weight
        passes the draw, with every weight from one reduced density matrix."""
        r = rng.random()
        if channel.mixture is not None:
            acc = 0.0
            for p, unitary in channel.mixture:
                acc += p
                if r < acc:
                    if not np.allclose(unitary, np.eye(unitary.shape[0])):
                        self.apply(unitary, qubits)
                    return
            return
        rho = self._reduced(tuple(qubits))
        acc = 0.0
        weight = 0.0
        for kraus in channel.kraus:
            weight = float(np.real(np.trace(kraus @ rho @ kraus.conj().T)))
            acc += weight
            if r < acc and weight > 0:
                self.apply(kraus / math.sqrt(weight), qubits)
                return
        if weight > 0:
            self.apply(channel.kraus[-1] / math.sqrt(weight), qubits)

    def probability_of_one(self, qubit: int) -> float:
        return float(self._reduced((qubit,))[1, 1].real)

    def measure(self, qubit: int, rng: np.random.Generator) -> int:
        p1 = min(max(self.probability_of_one(qubit), 0.0), 1.0)
        outcome = int(rng.random() < p1)
        self.collapse(qubit, outcome, p1 if outcome else 1 - p1)
        return outcome

    def collapse(self, qubit: int, outcome: int, probability: float) -> None:
        scale = 1.0 / math.sqrt(probability) if probability > 0 else 1.0
        d = np.zeros(2, dtype=np.complex128)
        d[outcome] = scale
        self.run([Pass("diag", (qubit,), d)])

    def reset(self, qubit: int, rng: np.random.Generator) -> None:
        if self.measure(qubit, rng):
            self.apply(np.array([[0, 1], [1, 0]], dtype=np.complex128), (qubit,))

    def device_probabilities(self) -> torch.Tensor:
        out = torch.empty(1 << self.n, dtype=torch.float64, device=self.buffer.device)
        nat.call("lsim_sv_probabilities", ptr(self.buffer), self.dtype, self.n, 1, ptr(out), self._stream())
        return out

    def vector(self) -> NDArray[np.complex128]:
        return self.buffer.cpu().numpy().astype(np.complex128, copy=True)

    @property
    def tensor(self) -> NDArray[np.complex128]:
        """A host copy shaped [2]*n (axis n-1-q is qubit q); assigning one uploads it."""
        v = self.vector()
        return v.reshape([2] * self.n) if self.n else v

    @tensor.setter
    def tensor(self, value: NDArray[np.complex128]) -> None:
        host = np.ascontiguousarray(np.asarray(value).reshape(-1), dtype=_NP_DTYPE[self.dtype])
        self.buffer.copy_(torch.from_numpy(host))

    # -- beyond the protocol -------------------------------------------------------------------

    def expectations(self, labels: Sequence[str]) -> NDArray[np.float64]:
        """<psi|P|psi> for each Pauli label (qubit 0 rightmost), from one launch per 64 labels."""
        return _expect("lsim_sv_pauli_expect", self.buffer, self.dtype, self.n, self.n, labels)


class DensityMatrix(_DeviceState):
    """A 2^n x 2^n density matrix in HBM as a vector of 2n bits (row bits above column bits)."""

    def __init__(self, num_qubits: int, *, dtype: int = nat.C128) -> None:
        if num_qubits > MAX_DENSITY_QUBITS:
            raise SimulationError(
                f"{num_qubits} qubits need more memory than the density-matrix engine allows "
                f"(up to {MAX_DENSITY_QUBITS}); use method='trajectories'"
            )
        device = require_device()
        self.n = num_qubits
        self.nbits = 2 * num_qubits
        self.dtype = dtype
        self.buffer = torch.empty(1 << self.nbits, dtype=_TORCH_DTYPE[dtype], device=device)
        nat.call("lsim_sv_set_basis", ptr(self.buffer), dtype, self.nbits, 1, 0, self._stream())

    def _bits(self, qubits: tuple[int, ...]) -> tuple[int, ...]:
        return tuple(self.n + q for q in qubits) + tuple(qubits)

    def apply(self, matrix: Matrix, qubits: tuple[int, ...]) -> None:
        """rho <- U rho U^dagger: one superoperator pass for up to two qubits, else U on the row
        bits and conj(U) on the column bits (engines.py:233-235)."""
        u = np.asarray(matrix, dtype=np.complex128)
        q = tuple(qubits)
        if not q:
            return
        if len(q) <= 2:
            self.run([_single(superop_of_gate(u), self._bits(q))])
        else:
            if len(q) > nat.MAX_K:
                raise SimulationError(f"gates on more than {nat.MAX_K} qubits aren't supported by the device engine")
            self.run([_single(u, tuple(self.n + x for x in q)), _single(u.conj(), q)])

    def scale(self, factor: complex) -> None:
        """A global phase leaves a density matrix unchanged."""

    def apply_channel(self, channel: Channel, qubits: tuple[int, ...]) -> None:
        q = tuple(qubits)
        if len(q) <= 2:
            self.run([_single(channel.superoperator(), self._bits(q))])
            return
        self._kraus_sum(channel, q)

    def _kraus_sum(self, channel: Channel, q: tuple[int, ...]) -> None:
        """sum_i K_i rho K_i^dagger out of place, for channels on more than two qubits."""
        if len(q) > nat.MAX_K:
            raise SimulationError(f"channels on more than {nat.MAX_K} qubits aren't supported by the device engine")
        total = torch.zeros_like(self.buffer)
        stream = self._stream()
        for k in channel.kraus:
            term = self.buffer.clone()
            run_passes(term, self.dtype, self.nbits, 1, [_single(k, tuple(self.n + x for x in q)), _single(k.conj(), q)])
            nat.call("lsim_sv_axpy", ptr(total), ptr(term), self.dtype, self.nbits, 1, 1.0, 0.0, stream)
        self.buffer = total

    def reset(self, qubit: int) -> None:
        self.apply_channel(RESET_CHANNEL, (qubit,))

    def run_kraus(self, channel: Channel, qubits: tuple[int, ...]) -> None:
        self._kraus_sum(channel, tuple(qubits))

    def device_probabilities(self) -> torch.Tensor:
        out = torch.empty(1 << self.n, dtype=torch.float64, device=self.buffer.device)
        nat.call("lsim_dm_probabilities", ptr(self.buffer), self.dtype, self.n, 1, ptr(out), self._stream())
        return out

    def matrix(self) -> NDArray[np.complex128]:
        d = 1 << self.n
        return self.buffer.cpu().numpy().astype(np.complex128, copy=True).reshape(d, d)

    @property
    def tensor(self) -> NDArray[np.complex128]:
        m = self.matrix()
        return m.reshape([2] * (2 * self.n)) if self.n else m

    @tensor.setter
    def tensor(self, value: NDArray[np.complex128]) -> None:
        host = np.ascontiguousarray(np.asarray(value).reshape(-1), dtype=_NP_DTYPE[self.dtype])
        self.buffer.copy_(torch.from_numpy(host))

    def expectations(self, labels: Sequence[str]) -> NDArray[np.float64]:
        """Re Tr(P rho) for each Pauli label (qubit 0 rightmost)."""
        return _expect("lsim_dm_pauli_expect", self.buffer, self.dtype, self.n, 2 * self.n, labels)


def _expect(fn: str, buf: torch.Tensor, dtype: int, n: int, nbits: int, labels: Sequence[str]) -> NDArray[np.float64]:
    labels = list(labels)
    if n == 0:
        # a 0-qubit state is the scalar 1 (or rho = [[1]]): every (empty) string has value 1
        return np.ones(len(labels))
    return launch_expectations(fn, buf, dtype, n, nbits, labels).cpu().numpy()


def launch_expectations(fn: str, buf: torch.Tensor, dtype: int, n: int, nbits: int, labels: Sequence[str]) -> torch.Tensor:
    """Queue the Pauli-expectation reductions of ``labels`` on the current stream; returns the
    device tensor that will hold the values (read it after the stream reaches this point)."""
    labels = list(labels)
    res = torch.empty(max(len(labels), 1), dtype=torch.float64, device=buf.device)
    stream = stream_handle()
    for start in range(0, len(labels), nat.MAX_TERMS):
        chunk = labels[start : start + nat.MAX_TERMS]
        masks = [pauli_masks(lbl) for lbl in chunk]
        xs = nat.ull_array([m[0] for m in masks])
        zs = nat.ull_array([m[1] for m in masks])
        ws = WORKSPACE.get(nat.load().lsim_reduce_workspace_bytes(nbits, 1, len(chunk)))
        nat.call(fn, ptr(buf), dtype, n, 1, len(chunk), xs, zs, ptr(res) + 8 * start, ptr(ws), stream)
    return res[: len(labels)]


#: NumPy's iterator buffer (np.getbufsize()): the longest piece of a contiguous run of dropped
#: entries that NumPy's reduction sums pairwise before adding it to the output
NUMPY_CHUNK_BITS = 13
#: log2 of the longest chain of ordered additions per marginal entry left to the device (each
#: entry adds its pieces one after another, as NumPy does; longer chains stay on the host)
DEVICE_MARGINAL_CHAIN_BITS = 16


def device_marginal_pays(n: int, keep: Sequence[int]) -> bool:
    """Whether :func:`device_marginal` suits this keep set: fewer than every qubit, and at most
    2^DEVICE_MARGINAL_CHAIN_BITS ordered additions per entry (each entry adds the pieces of
    2^min(low, 13) entries that share its kept bits, ``low`` being the number of dropped qubits
    below the lowest kept one)."""
    low = 0
    while low < n and low not in keep:
        low += 1
    chain_bits = n - min(low, NUMPY_CHUNK_BITS) - len(keep)
    return 0 < len(keep) < n and chain_bits <= DEVICE_MARGINAL_CHAIN_BITS


def device_marginal(probs: torch.Tensor, n: int, keep: Sequence[int], stream: int) -> NDArray[np.float64]:
    """The marginal of float64 device probabilities over the sorted qubits ``keep``, summed on
    the device by lsim_marginal in NumPy's order, so it equals :func:`marginal` of the same
    values bit for bit; returns the 2^m values on the host."""
    keep_arr = nat.int_array(list(keep))
    lib = nat.load()
    nbytes = lib.lsim_marginal_workspace_bytes(n, 1, len(keep), keep_arr)
    out = torch.empty(1 << len(keep), dtype=torch.float64, device=probs.device)
    ws = torch.empty(max(nbytes // 8, 1), dtype=torch.float64, device=probs.device)
    nat.call("lsim_marginal", ptr(probs), n, 1, len(keep), keep_arr, ptr(out), ptr(ws), stream)
    return out.cpu().numpy()


def marginal(probabilities: NDArray[np.float64], num_qubits: int, qubits: list[int]) -> NDArray[np.float64]:
    """Probabilities of ``qubits`` alone (index bit i = i-th smallest of them), by summing the
    others out exactly as engines.py:267-275 does, so sampling sees the same numbers."""
    keep = sorted(qubits)
    if num_qubits == 0:
        return probabilities
    tensor = probabilities.reshape([2] * num_qubits)
    drop = tuple(num_qubits - 1 - q for q in range(num_qubits) if q not in keep)
    reduced = tensor.sum(axis=drop) if drop else tensor
    return np.asarray(reduced).reshape(-1)


__all__ = [
    "MAX_DENSITY_QUBITS",
    "MAX_STATEVECTOR_QUBITS",
    "DensityMatrix",
    "SimulationError",
    "Statevector",
    "marginal",
    "pauli_masks",
]
