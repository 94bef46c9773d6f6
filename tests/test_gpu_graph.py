"""Captured HIP graphs of whole simulations: every replay, with eager work and synchronisations
in between, equals the eager result and the oracle; parameter updates reuse the graph."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import lucivant_sim as ls
from lucivant_sim import workloads as wl
from lucivant_sim.graph import CompiledCircuit
from oracle import engine as orc

pytestmark = pytest.mark.gpu


def test_replays_equal_eager_and_oracle() -> None:
    c = ls.parse(wl.random_layers(16, 6, 12, two_qubit=("cx", "cz")))
    want = orc.evolve_sv(c)
    cc = CompiledCircuit(c, width=4)
    eager = cc.run_eager().vector()
    assert np.abs(eager - want).max() < 1e-10
    for i in range(4):
        got = cc.replay().vector()
        assert np.array_equal(got, eager), i
        # unrelated eager work between replays
        x = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        del x


def test_update_reuses_the_graph_for_new_angles() -> None:
    a = ls.parse(wl.hea_ansatz(14, 4, 1))
    b = ls.parse(wl.hea_ansatz(14, 4, 2))  # same structure, other angles
    cc = CompiledCircuit(a, width=3)
    assert np.abs(cc.replay().vector() - orc.evolve_sv(a)).max() < 1e-10
    cc.update(b)
    assert np.abs(cc.replay().vector() - orc.evolve_sv(b)).max() < 1e-10
    assert np.abs(cc.replay().vector() - orc.evolve_sv(b)).max() < 1e-10


def test_library_graph_holds_one_node_per_launch_and_refuses_the_null_stream() -> None:
    """The graph is the library's own (lsim_graph_*): one kernel node per recorded launch (the
    initialisation and every pass), and capture on the null stream is refused before anything is
    recorded."""
    from lucivant_sim import _native as nat

    c = ls.parse(wl.random_layers(12, 3, 5))
    cc = CompiledCircuit(c, width=3)
    assert cc.nodes == 1 + len(cc.passes)
    lib = nat.load()
    assert lib.lsim_graph_capture_begin(None) == 1
    assert "null stream" in lib.lsim_last_error().decode()
    assert lib.lsim_graph_nodes(None) == -1
    assert lib.lsim_graph_destroy(None) == 0


def test_graph_recorded_through_the_c_abi_replays_with_new_data() -> None:
    """Record two passes on a batch of 3 states through the C ABI directly, then replay after
    overwriting the states and the matrices in place: each replay reads the current data, with
    eager launches and synchronisations between replays."""
    import ctypes

    from lucivant_sim import _native as nat
    from lucivant_sim.engines import ptr

    n, batch = 10, 3
    rng = np.random.default_rng(4)

    def states() -> np.ndarray:
        v = rng.normal(size=(batch, 1 << n)) + 1j * rng.normal(size=(batch, 1 << n))
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    def unitary(k: int) -> np.ndarray:
        q, _ = np.linalg.qr(rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k)))
        return q

    buf = torch.from_numpy(states().reshape(-1)).cuda()
    u2, u3 = torch.from_numpy(unitary(2)).cuda(), torch.from_numpy(unitary(3)).cuda()
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    h = ctypes.c_void_p()
    nat.call("lsim_graph_capture_begin", side.cuda_stream)
    nat.call("lsim_sv_apply_dense", ptr(buf), nat.C128, n, batch, 2, nat.int_array([7, 1]), ptr(u2), None, side.cuda_stream)
    nat.call("lsim_sv_apply_dense", ptr(buf), nat.C128, n, batch, 3, nat.int_array([0, 9, 4]), ptr(u3), None, side.cuda_stream)
    nat.call("lsim_graph_capture_end", side.cuda_stream, ctypes.byref(h))
    try:
        assert nat.load().lsim_graph_nodes(h) == 2
        for rep in range(3):
            psi = states()
            m2, m3 = unitary(2), unitary(3)
            buf.copy_(torch.from_numpy(psi.reshape(-1)))
            u2.copy_(torch.from_numpy(m2))
            u3.copy_(torch.from_numpy(m3))
            torch.cuda.synchronize()
            nat.call("lsim_graph_launch", h, torch.cuda.current_stream().cuda_stream)
            got = buf.cpu().numpy().reshape(batch, -1)
            for b in range(batch):
                want = orc.apply_sv(orc.apply_sv(psi[b], m2, (7, 1), n), m3, (0, 9, 4), n)
                assert np.abs(got[b] - want).max() < 1e-12, (rep, b)
            x = torch.randn(1 << 18, device="cuda") * 2  # eager work between replays
            torch.cuda.synchronize()
            del x
    finally:
        nat.load().lsim_graph_destroy(h)
