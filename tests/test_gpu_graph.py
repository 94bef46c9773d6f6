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
