"""The allocation-free Pauli expectation (lsim_sv_pauli_expect_noalloc) against
lsim_sv_pauli_expect with a caller workspace: the same kernels in the same order, so the values
are bit-identical. Two calls in a row on one GPU share the library's scratch; the second has a
smaller input (fewer states, fewer strings, fewer bits) and must not see the first's partials."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from lucivant_sim import _native as nat
from lucivant_sim import workloads as wl
from lucivant_sim.engines import Statevector, pauli_masks, ptr, require_device

pytestmark = pytest.mark.gpu


def _states(nbits: int, batch: int, seed: int, dtype: torch.dtype) -> torch.Tensor:
    g = torch.Generator().manual_seed(seed)
    v = torch.randn(batch, 1 << nbits, dtype=torch.float64, generator=g) + 1j * torch.randn(batch, 1 << nbits, dtype=torch.float64, generator=g)
    v = v / torch.linalg.vector_norm(v, dim=1, keepdim=True)
    return v.to(dtype).to(require_device()).contiguous()


def _both(states: torch.Tensor, dtype: int, nbits: int, batch: int, labels: list[str]) -> tuple[np.ndarray, np.ndarray]:
    lib = nat.load()
    masks = [pauli_masks(lbl) for lbl in labels]
    xs = nat.ull_array([m[0] for m in masks])
    zs = nat.ull_array([m[1] for m in masks])
    stream = torch.cuda.current_stream().cuda_stream
    want = torch.empty(batch * len(labels), dtype=torch.float64, device=states.device)
    ws = torch.empty(lib.lsim_reduce_workspace_bytes(nbits, batch, len(labels)) // 8 + 1, dtype=torch.float64, device=states.device)
    nat.call("lsim_sv_pauli_expect", ptr(states), dtype, nbits, batch, len(labels), xs, zs, ptr(want), ptr(ws), stream)
    got = torch.full((batch * len(labels),), float("nan"), dtype=torch.float64, device=states.device)
    nat.call("lsim_sv_pauli_expect_noalloc", ptr(states), dtype, nbits, batch, len(labels), xs, zs, ptr(got), stream)
    torch.cuda.synchronize()
    return got.cpu().numpy(), want.cpu().numpy()


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
def test_noalloc_expectation_equals_the_workspace_entry_point(dtype: int) -> None:
    tdt = torch.complex128 if dtype == nat.C128 else torch.complex64
    print(f"no-alloc expectation, dtype {dtype}: 22 bits x 2 states x 24 strings, then 12 bits x 1 state x 5 strings", flush=True)
    big = _states(22, 2, 1, tdt)
    got, want = _both(big, dtype, 22, 2, wl.pauli_strings(22, 24, 2) + ["Z" * 22])
    assert np.array_equal(got, want) and np.all(np.isfinite(got))
    small = _states(12, 1, 3, tdt)
    got, want = _both(small, dtype, 12, 1, wl.pauli_strings(12, 5, 4))
    assert np.array_equal(got, want) and np.all(np.isfinite(got))


def test_noalloc_expectation_at_26_qubits_and_in_a_loop() -> None:
    """One 26-qubit state with 64 strings (2 MiB of partials), then repeated small calls."""
    print("no-alloc expectation: 26 bits x 64 strings, then 10 x (14 bits x 3 strings)", flush=True)
    sv = Statevector(26)
    src, _ = wl.config_c2()
    import lucivant_sim as ls

    from lucivant_sim.simulator import _evolve

    _evolve(sv, ls.parse(src), None)
    labels = wl.pauli_strings(26, 63, 9) + ["I" * 26]
    got, want = _both(sv.buffer, nat.C128, 26, 1, labels)
    assert np.array_equal(got, want)
    assert got[-1] == pytest.approx(1.0, abs=1e-12)
    small = _states(14, 1, 5, torch.complex128)
    for i in range(10):
        got, want = _both(small, nat.C128, 14, 1, wl.pauli_strings(14, 3, 20 + i))
        assert np.array_equal(got, want)
