"""Compiled circuits: plan once, upload once, replay as one HIP graph.

A :class:`CompiledCircuit` holds a statevector in HBM, the fused passes of a circuit, their
matrices in one device buffer, and a HIP graph (captured through ``torch.cuda.CUDAGraph``, which
is HIP's graph API on ROCm) of the whole sequence: state initialisation then every pass. A
replay runs the simulation with one launch from the host, so mid-size states whose passes last a
few microseconds are no longer bound by per-launch host overhead.

:meth:`CompiledCircuit.update` swaps in a circuit of the same structure (the same gates on the
same qubits, other angles: a variational sweep) by recomputing the fused matrices on the host and
copying them into the same device buffer; the captured graph reads them from there, so no
recapture is needed. The graph contains only this library's kernels (no memset nodes: the state
is initialised by a kernel).
"""

from __future__ import annotations

import numpy as np
import torch

from lucivant_sim import _native as nat
from lucivant_sim.circuit import Circuit, Gate
from lucivant_sim.engines import _NP_DTYPE, Statevector, ptr
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
        self._layout = [(p.kind, p.bits, p.matrix.size) for p in self.passes]
        host = self._pack(self.passes)
        self.mats = torch.from_numpy(host).to(self.state.buffer.device)
        self.graph: torch.cuda.CUDAGraph | None = None
        self._capture()

    def _pack(self, passes: list) -> np.ndarray:
        npdt = _NP_DTYPE[self.dtype]
        flat = [np.ascontiguousarray(p.matrix, dtype=npdt).reshape(-1) for p in passes]
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
            off += p.matrix.size
            if p.kind == "scale":
                z = complex(p.matrix[0])
                rc = lib.lsim_sv_scale(state, self.dtype, n, 1, z.real, z.imag, stream)
            else:
                fn = lib.lsim_sv_apply_dense if p.kind == "dense" else lib.lsim_sv_apply_diag
                rc = fn(state, self.dtype, n, 1, p.k, nat.int_array(p.bits), addr, None, stream)
            if rc:
                raise nat.NativeError(lib.lsim_last_error().decode())

    def _capture(self) -> None:
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self._launch_all(side.cuda_stream)  # warm-up outside the graph (module loads)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._launch_all(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()

    def replay(self) -> Statevector:
        """Run the whole simulation again (one graph launch); returns the state object."""
        assert self.graph is not None
        self.graph.replay()
        return self.state

    def run_eager(self) -> Statevector:
        """The same launches without the graph (for comparison)."""
        self._launch_all(torch.cuda.current_stream().cuda_stream)
        return self.state

    def update(self, circuit: Circuit) -> None:
        """Load the matrices of a circuit with the same structure (other gate angles)."""
        if _structure(circuit) != _structure(self.circuit):
            raise ValueError("update needs a circuit with the same gates on the same qubits")
        passes = plan_statevector(circuit.operations, circuit.num_qubits, self.width)
        layout = [(p.kind, p.bits, p.matrix.size) for p in passes]
        if layout != self._layout:
            # a diagonal block may turn dense for other angles: the graph would not match
            raise ValueError("the new angles change the pass structure; compile the circuit anew")
        self.mats.copy_(torch.from_numpy(self._pack(passes)))
        self.passes = passes
        self.circuit = circuit
