"""The front-end worker pool: programs read in worker processes give the same op tables,
matrix pools, digests, measured pairs and methods as reading them in this process, and errors
come back with their type, message, line and column."""

from __future__ import annotations

import numpy as np
import pytest

import lucivant_sim as ls
from lucivant_sim import frontend, qasm
from lucivant_sim import workloads as wl
from lucivant_sim.noise import NoiseModel
from lucivant_sim.optable import static_table

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def pool():  # type: ignore[no-untyped-def]
    if torch.cuda.is_initialized():
        pytest.skip("the pool forks before the GPU is used")
    n = frontend.start(2)
    assert n == 2 and frontend.active()
    yield n
    frontend.stop()
    assert not frontend.active()


def test_workers_read_what_this_process_reads(pool) -> None:  # type: ignore[no-untyped-def]
    texts = [wl.random_layers(6 + i % 3, 3, 40 + i) for i in range(9)] + [wl.hea_ansatz(5, 2, 1)]
    got = frontend.compile_texts(texts, None, None, "automatic", 13)
    assert len(got) == len(texts)
    for src, c in zip(texts, got):
        ref = ls.parse(src)
        assert c.error is None and c.method == "statevector"
        assert c.digest == ref.digest()
        assert (c.num_qubits, c.num_clbits, c.cregs) == (ref.num_qubits, ref.num_clbits, ref.cregs)
        assert c.final_measurements() == ref.final_measurements()
        table, pool_ = static_table(ref)
        assert c.table is not None and c.pool is not None
        assert c.table.tobytes() == table.tobytes() and np.array_equal(c.pool, pool_)
        assert c.format_many(np.array([0, 1, 5])) == ref.format_many(np.array([0, 1, 5]))


def test_wide_noisy_and_dynamic_programs_keep_the_circuit(pool) -> None:  # type: ignore[no-untyped-def]
    noise = NoiseModel.uniform(single=1e-3, two=1e-2)
    texts = [wl.random_layers(14, 2, 3), wl.teleport_chain(9, 3), wl.random_layers(6, 2, 4)]
    got = frontend.compile_texts(texts, None, noise, "automatic", 13)
    assert got[0].table is None and got[0].circuit is not None  # wider than the LDS program holds
    assert got[1].method == "trajectories" and got[1].circuit is not None
    assert got[2].method == "density_matrix" and got[2].circuit is not None
    assert got[0].circuit.digest() == ls.parse(texts[0]).digest()


def test_errors_come_back_whole(pool) -> None:  # type: ignore[no-untyped-def]
    bad = 'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[2];\nh q[0];\n  cx q[0], q[7];\n'
    got = frontend.compile_texts([wl.random_layers(4, 1, 1), bad], None, None, "automatic", 13)
    assert got[0].error is None
    err = got[1].error
    assert isinstance(err, qasm.QasmError)
    with pytest.raises(qasm.QasmError) as ref:
        ls.parse(bad)
    assert (str(err), err.line, err.column) == (str(ref.value), ref.value.line, ref.value.column)
    got = frontend.compile_texts([wl.random_layers(4, 1, 1)], None, None, "bogus", 13)
    assert isinstance(got[0].error, ValueError)


def test_async_submit(pool) -> None:  # type: ignore[no-untyped-def]
    src = wl.random_layers(5, 2, 9)
    c = frontend.submit(src, None, None, "automatic", 13).get(timeout=60)
    assert c.digest == ls.parse(src).digest()


def test_trajectory_workers_that_died_are_reported_and_dropped(monkeypatch) -> None:  # type: ignore[no-untyped-def]
    """A trajectory worker that dies breaks its executor: the request it held fails with
    BrokenProcessPool, the next submission raises it too, and from then on no trajectory workers
    are reported (the server then runs trajectory requests on its own thread)."""
    import multiprocessing
    import os
    from concurrent.futures import ProcessPoolExecutor
    from concurrent.futures.process import BrokenProcessPool

    ex = ProcessPoolExecutor(1, mp_context=multiprocessing.get_context("fork"))
    try:
        held = ex.submit(os._exit, 3)
        with pytest.raises(BrokenProcessPool):
            held.result(timeout=60)
        monkeypatch.setattr(frontend, "_TRAJ_POOL", ex)
        monkeypatch.setattr(frontend, "_TRAJ_WORKERS", 1)
        assert frontend.trajectory_workers() == 1
        with pytest.raises(BrokenProcessPool):
            frontend.submit_trajectories(ls.parse(wl.random_layers(3, 1, 1)), None, 10, 1, 0)
        assert frontend.trajectory_workers() == 0
    finally:
        ex.shutdown()
