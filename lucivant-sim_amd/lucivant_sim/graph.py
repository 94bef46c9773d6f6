"""Compiled circuits: plan once, upload once, replay as one HIP graph.

A :class:`CompiledCircuit` holds a statevector in HBM, the fused passes of a circuit, their
matrices in one device buffer, and a HIP graph of the whole sequence -- state initialisation
then every pass -- captured and owned by the library (``lsim_graph_capture_begin`` /
``lsim_graph_capture_end`` / ``lsim_graph_launch`` in include/lsim.h). A replay runs the
simulation with one launch from the host, so mid-size states whose passes last a few
microseconds are no longer bound by per-launch host overhead.

:meth:`CompiledCircuit.update` swaps in a circuit of the same structure (the same gates on the
same qubits, other angles: a variational sweep) by recomputing the fused matrices on the host and
copying them into the same device buffer; the captured graph reads them from there, so no
recapture is needed. The graph contains only this library's kernels (no memset nodes: the state
is initialised by a kernel).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from lucivant_sim import _native as nat
from lucivant_sim.circuit import Circuit, Gate
from lucivant_sim.engines import _NP_DTYPE, Statevector, pass_launch, ptr
from lucivant_sim.plan import plan_statevector


def _structure(circuit: Circuit) -> tuple:
    return tuple((op.qubits if isinstance(op, Gate) else type(op).__name__) for op in circuit.operations)


class CompiledCircuit:
    def __init__(self, circuit: Circuit, width: int = 4, dtype: int = nat.C128) -> None:
        if circuit.is_dynamic():
            raise ValueError("a compiled circuit is static (no mid-circuit measurement or conditions)")
        self.circuit = circuit
        self.width = width
        self.dtype = dtype
        self.state = Statevector(circuit.num_qubits, dtype=dtype)
        self.passes = plan_statevector(circuit.operations, circuit.num_qubits, width)
        self._layout = [(p.kind, p.bits, p.payload().size, p.moves() and p.moves()[0]) for p in self.passes]
        host = self._pack(self.passes)
        self.mats = torch.from_numpy(host).to(self.state.buffer.device)
        self.graph: int | None = None  # an lsim_graph handle
        self._capture()

    def _pack(self, passes: list) -> np.ndarray:
        npdt = _NP_DTYPE[self.dtype]
        flat = [np.ascontiguousarray(p.payload(), dtype=npdt).reshape(-1) for p in passes]
        return np.concatenate(flat) if flat else np.zeros(1, dtype=npdt)

    def _launch_all(self, stream: int) -> None:
        lib = nat.load()
        n = self.circuit.num_qubits
        state = ptr(self.state.buffer)
        rc = lib.lsim_sv_set_basis(state, self.dtype, n, 1, 0, stream)
        if rc:
            raise nat.NativeError(lib.lsim_last_error().decode())
        base = ptr(self.mats)
        item = np.dtype(_NP_DTYPE[self.dtype]).itemsize
        off = 0
        for p in self.passes:
            addr = base + off * item
            off += p.payload().size
            fn, args = pass_launch(lib, p, state, self.dtype, n, 1, addr, stream)
            if fn(*args):
                raise nat.NativeError(lib.lsim_last_error().decode())

    def _capture(self) -> None:
        # the launches run once eagerly first (module loading stays out of the graph), then are
        # recorded on a stream of their own
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self._launch_all(side.cuda_stream)
        torch.cuda.synchronize()
        handle = ctypes.c_void_p()
        nat.call("lsim_graph_capture_begin", side.cuda_stream)
        try:
            self._launch_all(side.cuda_stream)
        finally:
            nat.call("lsim_graph_capture_end", side.cuda_stream, ctypes.byref(handle))
        self.graph = handle.value
        self.nodes = int(nat.load().lsim_graph_nodes(self.graph))
        torch.cuda.synchronize()

    def replay(self) -> Statevector:
        """Run the whole simulation again (one graph launch on the current stream); returns the
        state object."""
        assert self.graph is not None
        nat.call("lsim_graph_launch", self.graph, torch.cuda.current_stream().cuda_stream)
        return self.state

    def __del__(self) -> None:
        graph, self.graph = getattr(self, "graph", None), None
        if graph is not None:
            try:
                torch.cuda.synchronize()
                nat.load().lsim_graph_destroy(graph)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass

    def run_eager(self) -> Statevector:
        """The same launches without the graph (for comparison)."""
        self._launch_all(torch.cuda.current_stream().cuda_stream)
        return self.state

    def update(self, circuit: Circuit) -> None:
        """Load the matrices of a circuit with the same structure (other gate angles)."""
        if _structure(circuit) != _structure(self.circuit):
            raise ValueError("update needs a circuit with the same gates on the same qubits")
        passes = plan_statevector(circuit.operations, circuit.num_qubits, self.width)
        layout = [(p.kind, p.bits, p.payload().size, p.moves() and p.moves()[0]) for p in passes]
        if layout != self._layout:
            # a diagonal block may turn dense for other angles: the graph would not match
            raise ValueError("the new angles change the pass structure; compile the circuit anew")
        self.mats.copy_(torch.from_numpy(self._pack(passes)))
        self.passes = passes
        self.circuit = circuit
