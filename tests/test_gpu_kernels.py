"""Parity of each gfx950 kernel, called through the C ABI, against the oracle.

complex128 results must agree with the oracle to 1e-12 absolute (the 1e-10 budget of the
reference's complex128 contract, with margin); complex64 to 1e-5. Integer/index work (which
amplitudes move, selectors, skipped states) is exact by construction of these checks.
"""

from __future__ import annotations

from itertools import permutations

import numpy as np
import pytest
import torch

from lucivant_sim import _native as nat
from lucivant_sim.engines import WORKSPACE, pauli_masks, ptr, stream_handle
from oracle import engine as orc

pytestmark = pytest.mark.gpu

TOL = {nat.C128: 1e-12, nat.C64: 1e-5}
NPDT = {nat.C128: np.complex128, nat.C64: np.complex64}


def dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def states(n: int, batch: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(batch, 2**n)) + 1j * rng.normal(size=(batch, 2**n))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def unitary(k: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(2**k, 2**k)) + 1j * rng.normal(size=(2**k, 2**k)))
    return q


def apply_dense(psi: np.ndarray, n: int, mats: np.ndarray, targets: tuple[int, ...], dtype: int, sel=None) -> np.ndarray:  # type: ignore[no-untyped-def]
    s = dev(psi.astype(NPDT[dtype]).reshape(-1))
    m = dev(mats.astype(NPDT[dtype]))
    sd = dev(np.asarray(sel, dtype=np.int32)) if sel is not None else None
    batch = psi.shape[0]
    nat.call("lsim_sv_apply_dense", ptr(s), dtype, n, batch, len(targets), nat.int_array(targets), ptr(m), ptr(sd) if sd is not None else None, stream_handle())
    return s.cpu().numpy().reshape(batch, -1)


@pytest.mark.parametrize("k,targets", [(3, (12, 7, 1)), (4, (11, 6, 2, 0)), (5, (12, 9, 5, 3, 1)), (5, (12, 11, 10, 9, 8))])
def test_matrix_cores_on_a_large_batch_of_states(k: int, targets: tuple[int, ...]) -> None:
    """2048 states of 13 qubits, each with its own operator or none: the matrix-core kernel then
    runs few wavefronts per state, each striding over several steps of its state; it agrees with
    the vector kernel to rounding and leaves unselected states untouched."""
    lib = nat.load()
    n, batch = 13, 2048
    rng = np.random.default_rng(40 + k)
    psi = rng.normal(size=(batch, 2**n)) + 1j * rng.normal(size=(batch, 2**n))
    mats = np.stack([unitary(k, 300 + j) for j in range(3)])
    sel = rng.integers(-1, 3, size=batch).astype(np.int32)
    old = lib.lsim_set_dense_mfma(-1)
    try:
        outs = []
        for mask in (1 << k, 0):
            lib.lsim_set_dense_mfma(mask)
            outs.append(apply_dense(psi, n, mats, targets, nat.C128, sel))
    finally:
        lib.lsim_set_dense_mfma(old)
    assert np.abs(outs[0] - outs[1]).max() < 1e-12
    untouched = sel < 0
    assert np.array_equal(outs[0][untouched], psi[untouched])
    for b in np.flatnonzero(~untouched)[:3]:
        assert np.abs(outs[0][b] - orc.apply_sv(psi[b], mats[sel[b]], targets, n)).max() < 1e-10


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_dense_every_placement(k: int, dtype: int) -> None:
    n = 6
    for i, targets in enumerate(permutations(range(n), k)):
        psi = states(n, 1, 100 + i)
        u = unitary(k, 7 + i)
        got = apply_dense(psi, n, u[None], targets, dtype)[0]
        want = orc.apply_sv(psi[0], u, targets, n)
        assert np.abs(got - want).max() < TOL[dtype], targets


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
@pytest.mark.parametrize("k", [4, 5])
def test_dense_wide_operators(k: int, dtype: int) -> None:
    n = 12
    rng = np.random.default_rng(k)
    placements = [tuple(range(k)), tuple(range(n - k, n)), tuple(reversed(range(k)))]
    placements += [tuple(int(x) for x in rng.choice(n, size=k, replace=False)) for _ in range(6)]
    for i, targets in enumerate(placements):
        psi = states(n, 1, 200 + i)
        u = unitary(k, 30 + i)
        got = apply_dense(psi, n, u[None], targets, dtype)[0]
        want = orc.apply_sv(psi[0], u, targets, n)
        assert np.abs(got - want).max() < TOL[dtype] * 4, targets


def test_dense_non_unitary_and_large_state() -> None:
    n = 18
    psi = states(n, 1, 5)
    lower = np.array([[1, 0, 0, 0], [0.5, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=np.complex128)
    for targets in [(0, 17), (17, 3), (9, 10)]:
        got = apply_dense(psi, n, lower[None], targets, nat.C128)[0]
        assert np.abs(got - orc.apply_sv(psi[0], lower, targets, n)).max() < 1e-12


def test_batched_selection() -> None:
    n, batch = 8, 7
    psi = states(n, batch, 9)
    mats = np.stack([unitary(2, s) for s in range(3)])
    sel = [0, 2, -1, 1, 1, -1, 0]
    got = apply_dense(psi, n, mats, (5, 1), nat.C128, sel)
    for b in range(batch):
        want = psi[b] if sel[b] < 0 else orc.apply_sv(psi[b], mats[sel[b]], (5, 1), n)
        assert np.abs(got[b] - want).max() < 1e-12


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
def test_diagonal(dtype: int) -> None:
    n, batch = 9, 3
    psi = states(n, batch, 3)
    rng = np.random.default_rng(0)
    for targets in [(0,), (8, 2), (4, 0, 7), (1, 2, 3, 4, 5)]:
        d = np.exp(1j * rng.uniform(0, 6.28, size=(2, 2 ** len(targets))))
        s = dev(psi.astype(NPDT[dtype]).reshape(-1))
        dd = dev(d.astype(NPDT[dtype]))
        sel = dev(np.array([1, -1, 0], dtype=np.int32))
        nat.call("lsim_sv_apply_diag", ptr(s), dtype, n, batch, len(targets), nat.int_array(targets), ptr(dd), ptr(sel), stream_handle())
        got = s.cpu().numpy().reshape(batch, -1)
        for b, m in enumerate([1, -1, 0]):
            want = psi[b] if m < 0 else orc.apply_sv(psi[b], np.diag(d[m]), targets, n)
            assert np.abs(got[b] - want).max() < TOL[dtype]


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
def test_streaming_modes_of_diagonal_and_scale(dtype: int) -> None:
    """Every streaming mode (0 never, 1 by size, 2 always) gives bit-identical diagonal and scale
    results, equal to the oracle's."""
    lib = nat.load()
    n, batch = 9, 3
    psi = states(n, batch, 4)
    d = np.exp(1j * np.random.default_rng(1).uniform(0, 6.28, size=(2, 8)))
    targets = (6, 1, 4)
    old = lib.lsim_set_dense_streaming(-1)
    outs = []
    try:
        for mode in (0, 1, 2):
            assert lib.lsim_set_dense_streaming(mode) in (0, 1, 2) and lib.lsim_set_dense_streaming(-1) == mode
            s = dev(psi.astype(NPDT[dtype]).reshape(-1))
            dd = dev(d.astype(NPDT[dtype]))
            sel = dev(np.array([1, -1, 0], dtype=np.int32))
            nat.call("lsim_sv_apply_diag", ptr(s), dtype, n, batch, 3, nat.int_array(targets), ptr(dd), ptr(sel), stream_handle())
            nat.call("lsim_sv_scale", ptr(s), dtype, n, batch, 0.6, 0.8, stream_handle())
            outs.append(s.cpu().numpy().reshape(batch, -1))
    finally:
        lib.lsim_set_dense_streaming(old)
    for b, m in enumerate([1, -1, 0]):
        want = (psi[b] if m < 0 else orc.apply_sv(psi[b], np.diag(d[m]), targets, n)) * (0.6 + 0.8j)
        assert np.abs(outs[0][b] - want).max() < TOL[dtype] * 2
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_scale_of_a_state_beyond_the_last_level_cache() -> None:
    """A 25-qubit complex128 state (512 MiB) takes the streaming scale kernel under the default
    mode; it multiplies every amplitude exactly as the cached kernel does."""
    lib = nat.load()
    n = 25
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(1 << n, dtype=torch.complex128, device="cuda", generator=g)
    old = lib.lsim_set_dense_streaming(-1)
    try:
        got = []
        for mode in (1, 0):
            lib.lsim_set_dense_streaming(mode)
            y = x.clone()
            nat.call("lsim_sv_scale", ptr(y), nat.C128, n, 1, 0.6, -0.8, stream_handle())
            got.append(y)
    finally:
        lib.lsim_set_dense_streaming(old)
    torch.cuda.synchronize()
    assert torch.equal(got[0], got[1])
    assert float((got[0] - x * (0.6 - 0.8j)).abs().max()) < 1e-14


def test_scale_basis_probabilities_axpy() -> None:
    n, batch = 10, 2
    s = torch.empty(batch << n, dtype=torch.complex128, device="cuda")
    nat.call("lsim_sv_set_basis", ptr(s), nat.C128, n, batch, 37, stream_handle())
    host = s.cpu().numpy().reshape(batch, -1)
    want = np.zeros((batch, 2**n))
    want[:, 37] = 1
    assert np.array_equal(host, want)
    psi = states(n, batch, 1)
    s = dev(psi.reshape(-1))
    nat.call("lsim_sv_scale", ptr(s), nat.C128, n, batch, 0.3, -0.4, stream_handle())
    assert np.abs(s.cpu().numpy().reshape(batch, -1) - psi * (0.3 - 0.4j)).max() < 1e-15
    p = torch.empty(batch << n, dtype=torch.float64, device="cuda")
    nat.call("lsim_sv_probabilities", ptr(s), nat.C128, n, batch, ptr(p), stream_handle())
    assert np.abs(p.cpu().numpy() - np.abs(psi.reshape(-1) * (0.3 - 0.4j)) ** 2).max() < 1e-15
    y = dev(psi.reshape(-1))
    x = dev(psi[::-1].reshape(-1).copy())
    nat.call("lsim_sv_axpy", ptr(y), ptr(x), nat.C128, n, batch, 0.5, 2.0, stream_handle())
    assert np.abs(y.cpu().numpy() - (psi.reshape(-1) + (0.5 + 2j) * psi[::-1].reshape(-1))).max() < 1e-14


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
@pytest.mark.parametrize("targets", [(0,), (9,), (3, 7), (7, 3), (0, 9)])
def test_reduced_density_matrix(targets: tuple[int, ...], dtype: int) -> None:
    n, batch = 10, 3
    psi = states(n, batch, 4)
    s = dev(psi.astype(NPDT[dtype]).reshape(-1))
    k = len(targets)
    out = torch.empty(batch << (2 * k), dtype=torch.complex128, device="cuda")
    ws = WORKSPACE.get(nat.load().lsim_reduce_workspace_bytes(n, batch, k))
    nat.call("lsim_sv_reduced_dm", ptr(s), dtype, n, batch, k, nat.int_array(targets), ptr(out), ptr(ws), stream_handle())
    got = out.cpu().numpy().reshape(batch, 2**k, 2**k)
    for b in range(batch):
        v = psi[b].astype(NPDT[dtype]).astype(np.complex128)
        full = np.outer(v, v.conj())
        # reduce with the oracle's index convention: row index bits = targets (first = MSB)
        idx = np.arange(2**n)
        r = np.zeros(2**n, dtype=np.int64)
        for i, q in enumerate(targets):
            r |= ((idx >> q) & 1) << (k - 1 - i)
        rest = idx & ~sum(1 << q for q in targets)
        want = np.zeros((2**k, 2**k), dtype=np.complex128)
        same = rest[:, None] == rest[None, :]
        np.add.at(want, (r[:, None].repeat(2**n, 1)[same], r[None, :].repeat(2**n, 0)[same]), full[same])
        assert np.abs(got[b] - want).max() < 1e-12


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
def test_pauli_expectations_statevector(dtype: int) -> None:
    n = 3
    psi = states(n, 2, 8)
    labels = orc.all_labels(n)
    s = dev(psi.astype(NPDT[dtype]).reshape(-1))
    masks = [pauli_masks(lbl) for lbl in labels]
    out = torch.empty(2 * len(labels), dtype=torch.float64, device="cuda")
    ws = WORKSPACE.get(nat.load().lsim_reduce_workspace_bytes(n, 2, len(labels)))
    nat.call(
        "lsim_sv_pauli_expect",
        ptr(s), dtype, n, 2, len(labels),
        nat.ull_array([m[0] for m in masks]), nat.ull_array([m[1] for m in masks]),
        ptr(out), ptr(ws), stream_handle(),
    )  # fmt: skip
    got = out.cpu().numpy().reshape(2, -1)
    for b in range(2):
        v = psi[b].astype(NPDT[dtype]).astype(np.complex128)
        for t, lbl in enumerate(labels):
            assert abs(got[b, t] - orc.expectation_sv(v, lbl, n)) < TOL[dtype], lbl


def test_pauli_expectations_wide() -> None:
    from lucivant_sim import workloads as wl
    from lucivant_sim.engines import Statevector

    n = 14
    psi = states(n, 1, 2)[0]
    labels = wl.pauli_strings(n, 16, 5) + ["I" * n, "Z" * n]
    sv = Statevector(n)
    sv.tensor = psi
    got = sv.expectations(labels)
    for t, lbl in enumerate(labels):
        assert abs(got[t] - orc.expectation_sv(psi, lbl, n)) < 1e-12, lbl


def test_density_matrix_diagonal_and_pauli() -> None:
    from lucivant_sim.engines import DensityMatrix

    n = 3
    rng = np.random.default_rng(6)
    a = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    rho = a @ a.conj().T
    rho /= np.trace(rho)
    dm = DensityMatrix(n)
    dm.tensor = rho
    assert np.abs(dm.probabilities() - orc.probabilities_dm(rho)).max() < 1e-15
    labels = orc.all_labels(n)
    got = dm.expectations(labels)
    for t, lbl in enumerate(labels):
        assert abs(got[t] - orc.expectation_dm(rho, lbl, n)) < 1e-12, lbl


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_dense_tiled_placements_equal_the_oracle_and_the_direct_kernel(k: int, dtype: int) -> None:
    """Placements with low target bits take the LDS-tiled kernel; it must agree with the oracle
    and with the direct kernel (tiling switched off) on every kind of placement."""
    n = 12
    rng = np.random.default_rng(40 + k)
    placements = [tuple(range(k)), tuple(reversed(range(k))), (0,) + tuple(range(n - k + 1, n)), (1, 0) + tuple(range(7, 7 + k - 2))]
    placements += [tuple(int(x) for x in rng.choice(n, size=k, replace=False)) for _ in range(10)]
    lib = nat.load()
    for i, targets in enumerate(placements):
        psi = states(n, 2, 300 + i)
        u = unitary(k, 60 + i)
        want = orc.apply_sv(psi[1], u, targets, n)
        tiled = apply_dense(psi, n, np.stack([u, u]), targets, dtype, [1, 0])[1]
        old = lib.lsim_set_dense_tiling(0)
        try:
            direct = apply_dense(psi, n, np.stack([u, u]), targets, dtype, [1, 0])[1]
        finally:
            lib.lsim_set_dense_tiling(old)
        assert np.abs(tiled - want).max() < TOL[dtype] * 4, targets
        assert np.abs(tiled - direct).max() < TOL[dtype] * 4, targets


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
@pytest.mark.parametrize("k", [3, 4, 5])
@pytest.mark.parametrize("extra", [4, 5, 7])
def test_dense_streaming_and_matrix_core_variants_agree(k: int, dtype: int, extra: int) -> None:
    """Every kernel variant (vector or matrix cores, streaming or cached accesses) gives the oracle's
    state on high, low, mixed and scattered placements, leaves unselected states untouched, and
    the switches read back. complex64 always takes the vector kernel. ``extra`` = 5 is the
    smallest state the matrix-core kernel takes (two steps of 16 groups); 4 stays on the vector
    kernel."""
    n = k + extra
    lib = nat.load()
    rng = np.random.default_rng(10 * k)
    placements = [tuple(range(n - 1, n - 1 - k, -1)), tuple(range(k - 1, -1, -1)), tuple(range(k)), (n - 1, 0) + tuple(range(2, 2 + k - 2))]
    placements += [tuple(int(x) for x in rng.choice(n, size=k, replace=False)) for _ in range(3)]
    placements += [tuple(int(x) for x in rng.choice(range(3, n), size=k, replace=False)) for _ in range(3)]
    old_m, old_s = lib.lsim_set_dense_mfma(-1), lib.lsim_set_dense_streaming(-1)
    try:
        for i, targets in enumerate(placements):
            psi = states(n, 3, 700 + 10 * k + i)
            u = unitary(k, 120 + i)
            want0 = orc.apply_sv(psi[0], u, targets, n)
            want2 = orc.apply_sv(psi[2], u, targets, n)
            outs = []
            for mask in (0, 1 << k):
                for streaming in (0, 2):
                    lib.lsim_set_dense_mfma(mask)
                    lib.lsim_set_dense_streaming(streaming)
                    assert lib.lsim_set_dense_mfma(-1) == mask and lib.lsim_set_dense_streaming(-1) == streaming
                    got = apply_dense(psi, n, np.stack([u, u]), targets, dtype, [1, -1, 0])
                    assert np.abs(got[0] - want0).max() < TOL[dtype] * 4, (targets, mask, streaming)
                    assert np.abs(got[2] - want2).max() < TOL[dtype] * 4, (targets, mask, streaming)
                    assert np.array_equal(got[1], psi[1].astype(NPDT[dtype])), (targets, mask, streaming)
                    outs.append(got)
            # streaming changes nothing but the cache policy: bit-identical results
            assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[2], outs[3]), targets
    finally:
        lib.lsim_set_dense_mfma(old_m)
        lib.lsim_set_dense_streaming(old_s)


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
def test_copy_states_within_a_batch(dtype: int) -> None:
    """lsim_sv_copy_states copies whole states to other slots of the batch, bit for bit, leaves
    every other state alone and skips pairs with an index outside the batch."""
    n, batch = 11, 7
    psi = states(n, batch, 21).astype(NPDT[dtype])
    s = dev(psi.reshape(-1))
    src = np.array([0, 2, 2, 1, 9], dtype=np.int32)
    dst = np.array([4, 5, 6, -1, 3], dtype=np.int32)
    d_src, d_dst = dev(src), dev(dst)
    nat.call("lsim_sv_copy_states", ptr(s), dtype, n, batch, len(src), ptr(d_src), ptr(d_dst), stream_handle())
    got = s.cpu().numpy().reshape(batch, -1)
    want = psi.copy()
    want[4], want[5], want[6] = psi[0], psi[2], psi[2]
    assert np.array_equal(got, want)
    nat.call("lsim_sv_copy_states", ptr(s), dtype, n, batch, 0, 0, 0, stream_handle())
    with pytest.raises(nat.NativeError, match="count"):
        nat.call("lsim_sv_copy_states", ptr(s), dtype, n, batch, batch + 1, ptr(d_src), ptr(d_dst), stream_handle())


def _structured_labels(n: int, seed: int) -> list[str]:
    """Strings that exercise the grouping of the expectation kernel: X masks only in the low
    bits, only in the high bits, spans of exactly 8 dimensions above bit 4 and beyond (a group
    split), all-Z and identity strings, Y on both ends, and random strings."""
    from lucivant_sim import workloads as wl

    rng = np.random.default_rng(seed)

    def label(xmask: int, zmask: int) -> str:
        out = []
        for q in range(n):
            x, z = (xmask >> q) & 1, (zmask >> q) & 1
            out.append("Y" if x and z else "X" if x else "Z" if z else "I")
        return "".join(reversed(out))

    labels = ["I" * n, "Z" * n, "Y" + "I" * (n - 2) + "Y", label(0b1011, 0b0110), label(1 << (n - 1), 1)]
    for i in range(10):  # ten independent high masks: more than 8 dimensions above bit 4
        labels.append(label((1 << (4 + i % (n - 4))) | (1 << (n - 1 - i % 3)), int(rng.integers(1 << n))))
    labels += wl.pauli_strings(n, 40, seed)
    return labels


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
@pytest.mark.parametrize("n", [12, 14])
def test_diagonal_pauli_strings_interleaved_with_others(dtype: int, n: int) -> None:
    """Strings of only I and Z are read by their own groups (no partner amplitudes), ahead of the
    others: 20 of them (two groups) shuffled among 20 random strings, each value back in its own
    position, for a 2-state batch, against the oracle."""
    from lucivant_sim import workloads as wl

    print(f"diagonal Pauli groups: n={n} dtype={dtype}", flush=True)
    rng = np.random.default_rng(70 + n)
    diag = ["I" * n, "Z" + "I" * (n - 1), "I" * (n - 1) + "Z", "Z" * n]
    while len(diag) < 20:
        diag.append("".join("Z" if rng.random() < 0.5 else "I" for _ in range(n)))
    labels = diag + wl.pauli_strings(n, 20, 80 + n)
    labels = [labels[i] for i in rng.permutation(len(labels))]
    batch = 2
    psi = states(n, batch, 90 + n)
    s = dev(psi.astype(NPDT[dtype]).reshape(-1))
    masks = [pauli_masks(lbl) for lbl in labels]
    out = torch.empty(batch * len(labels), dtype=torch.float64, device="cuda")
    ws = WORKSPACE.get(nat.load().lsim_reduce_workspace_bytes(n, batch, len(labels)))
    nat.call(
        "lsim_sv_pauli_expect",
        ptr(s), dtype, n, batch, len(labels),
        nat.ull_array([m[0] for m in masks]), nat.ull_array([m[1] for m in masks]),
        ptr(out), ptr(ws), stream_handle(),
    )  # fmt: skip
    got = out.cpu().numpy().reshape(batch, -1)
    for b in range(batch):
        v = psi[b].astype(NPDT[dtype]).astype(np.complex128)
        want = np.array([orc.expectation_sv(v, lbl, n) for lbl in labels])
        assert np.abs(got[b] - want).max() < TOL[dtype]


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
@pytest.mark.parametrize("n", [12, 13, 16])
def test_grouped_pauli_strings_equal_the_oracle(dtype: int, n: int) -> None:
    """From 12 qubits on, the strings of a call are read in groups (one pass over the state per
    group of up to 16 strings whose X masks span at most 8 dimensions above bit 4): every value
    of a 3-state batch against the oracle."""
    print(f"grouped Pauli expectation: n={n} dtype={dtype}", flush=True)
    batch = 3
    psi = states(n, batch, 40 + n)
    labels = _structured_labels(n, n)
    assert len(labels) <= nat.MAX_TERMS
    s = dev(psi.astype(NPDT[dtype]).reshape(-1))
    masks = [pauli_masks(lbl) for lbl in labels]
    out = torch.empty(batch * len(labels), dtype=torch.float64, device="cuda")
    ws = WORKSPACE.get(nat.load().lsim_reduce_workspace_bytes(n, batch, len(labels)))
    nat.call(
        "lsim_sv_pauli_expect",
        ptr(s), dtype, n, batch, len(labels),
        nat.ull_array([m[0] for m in masks]), nat.ull_array([m[1] for m in masks]),
        ptr(out), ptr(ws), stream_handle(),
    )  # fmt: skip
    got = out.cpu().numpy().reshape(batch, -1)
    for b in range(batch):
        v = psi[b].astype(NPDT[dtype]).astype(np.complex128)
        want = np.array([orc.expectation_sv(v, lbl, n) for lbl in labels])
        assert np.abs(got[b] - want).max() < TOL[dtype]
    assert np.all(np.abs(got[:, 0] - 1) < TOL[dtype])  # the identity


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
@pytest.mark.parametrize(("n", "budget"), [(12, 24.0), (13, 6.0), (16, 24.0), (16, 1000.0), (22, 24.0)])
def test_tile_passes_equal_the_oracle(dtype: int, n: int, budget: float) -> None:
    """Tile passes (lsim_sv_apply_tile: a tile's amplitudes staged in LDS, a run of 1-2 qubit
    operators applied, written back) on a batch of two random states: every amplitude against the
    oracle applying the circuit's gates one by one (random layers with cx/cz, a wide gate that
    runs as its own dense pass, a global phase). At 22 qubits each workgroup of a launch takes
    two of the 1024 tiles, the second one prefetched while the first is computed."""
    import lucivant_sim as ls
    from lucivant_sim import workloads as wl
    from lucivant_sim.engines import run_plan
    from lucivant_sim.plan import TilePass, plan_tiles

    print(f"tile passes: n={n} budget={budget} dtype={dtype}", flush=True)
    src = wl.random_layers(n, 6 if n < 20 else 3, 17 + n, two_qubit=("cx", "cz"), measure=False)
    src += f"ccx q[0], q[{n // 2}], q[{n - 1}];\nh q[{n - 1}];\ngphase(0.3);\ncx q[{n - 1}], q[2];\n"
    c = ls.parse(src)
    plan = plan_tiles(c.operations, n, budget=budget)
    assert any(isinstance(p, TilePass) for p in plan)
    psi = states(n, 2, 70 + n)
    s = dev(psi.astype(NPDT[dtype]).reshape(-1))
    run_plan(s, dtype, n, 2, plan)
    got = s.cpu().numpy().reshape(2, -1)
    for b in range(2):
        want = psi[b].astype(NPDT[dtype]).astype(np.complex128)
        for op in c.operations:
            if isinstance(op, ls.circuit.Gate):
                want = orc.apply_sv(want, np.asarray(op.matrix), tuple(op.qubits), n)
            elif isinstance(op, ls.circuit.GlobalPhase):
                want = want * np.exp(1j * op.angle)
        assert np.abs(got[b] - want).max() < (1e-12 if dtype == nat.C128 else 2e-5)


@pytest.mark.parametrize(
    "n, keep",
    [
        (6, [1, 4]),  # pieces of 2 entries
        (10, [0, 7]),  # qubit 0 kept: entry by entry
        (12, [3, 11]),  # pieces of 8
        (14, [7, 8, 13]),  # pieces of 128
        (16, [9]),  # pieces of 512: a tree over four 128-entry blocks
        (18, [13, 17]),  # pieces of 8192
        (20, [16, 19]),  # runs of 2^16 cut into pieces of 8192
        (22, [0, 1, 2, 3, 4, 21]),
        (22, list(range(8, 22))),
    ],
)
def test_device_marginal_equals_numpy_bit_for_bit(n: int, keep: list[int]) -> None:
    """lsim_marginal over two states of float64 probabilities against the reference's marginal
    (NumPy's tensor.sum over the dropped axes, engines.py:267-275): identical float64 values,
    for chunk sizes from 1 to 8192 entries and runs longer than NumPy's buffer."""
    from lucivant_sim.engines import marginal

    print(f"marginal: n={n} keep={keep}", flush=True)
    rng = np.random.default_rng(n)
    p = rng.random((2, 1 << n)) ** 3
    p /= p.sum(axis=1, keepdims=True)
    probs = dev(p.reshape(-1))
    karr = nat.int_array(keep)
    nbytes = nat.load().lsim_marginal_workspace_bytes(n, 2, len(keep), karr)
    ws = torch.empty(max(nbytes // 8, 1), dtype=torch.float64, device="cuda")
    out = torch.empty(2 << len(keep), dtype=torch.float64, device="cuda")
    nat.call("lsim_marginal", ptr(probs), n, 2, len(keep), karr, ptr(out), ptr(ws), stream_handle())
    got = out.cpu().numpy().reshape(2, -1)
    for b in range(2):
        assert np.array_equal(got[b], marginal(p[b], n, keep)), b


@pytest.mark.parametrize("dtype", [nat.C128, nat.C64])
@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_permutations_with_phases_equal_the_oracle(k: int, dtype: int) -> None:
    """lsim_sv_apply_perm on a batch of three states: a random permutation of 2^k rows, random
    phases except for the rows flagged as plain moves, targets in a scrambled order (one of them
    the lowest bit); the third state's selector skips it. Against the oracle applying the same
    matrix as a dense operator."""
    n = 11
    rng = np.random.default_rng(40 + k)
    d = 1 << k
    perm = rng.permutation(d)
    unit = rng.random(d) < 0.5
    phases = np.where(unit, 1.0 + 0j, np.exp(1j * rng.uniform(0, 2 * np.pi, d)))
    mask = sum(1 << r for r in range(d) if unit[r])
    m = np.zeros((d, d), dtype=np.complex128)
    m[np.arange(d), perm] = phases
    targets = tuple(int(x) for x in rng.permutation([0, 3, 5, 8, 10][:k]))
    psi = states(n, 3, 7 + k)
    s = dev(psi.astype(NPDT[dtype]).reshape(-1))
    ph = dev(np.stack([phases, np.conj(phases)]).astype(NPDT[dtype]))
    sel = dev(np.array([0, 1, -1], dtype=np.int32))
    nat.call("lsim_sv_apply_perm", ptr(s), dtype, n, 3, k, nat.int_array(targets), nat.int_array([int(x) for x in perm]), mask, ptr(ph), ptr(sel), stream_handle())
    got = s.cpu().numpy().reshape(3, -1)
    m_conj = np.zeros_like(m)
    m_conj[np.arange(d), perm] = np.where(unit, 1.0 + 0j, np.conj(phases))
    for b, mat in ((0, m), (1, m_conj)):
        want = orc.apply_sv(psi[b].astype(NPDT[dtype]).astype(np.complex128), mat, targets, n)
        assert np.abs(got[b] - want).max() < TOL[dtype], b
    assert np.array_equal(got[2], psi[2].astype(NPDT[dtype]))


@pytest.mark.gpu
def test_uploads_through_the_staging_ring_keep_their_contents() -> None:
    """Uploads queued behind long device work, several times the ring's size in total and of
    mixed sizes and dtypes, each read back equal to its host array: no staging span is reused
    before the copy that reads it has run."""
    from lucivant_sim import engines

    print("staging ring: 400 uploads behind queued work", flush=True)
    ring = engines._Staging(1 << 20)  # a small ring, so it wraps many times
    device = engines.require_device()
    rng = np.random.default_rng(11)
    big = torch.randn(4096, 4096, dtype=torch.float64, device=device)
    hosts, devs = [], []
    for i in range(400):
        if i % 50 == 0:
            big = big @ big * 1e-3  # keeps the stream busy while uploads queue behind it
        n = 100_000 if i % 97 == 0 else int(rng.integers(1, 12_000))  # the large ones bypass the ring
        kind = i % 3
        host = rng.random(n) if kind == 0 else (rng.random(n) + 1j * rng.random(n)) if kind == 1 else rng.integers(-5, 5, n).astype(np.int32)
        hosts.append(host)
        devs.append(ring.upload(host, device))
    torch.cuda.synchronize()
    for host, dev in zip(hosts, devs):
        assert np.array_equal(dev.cpu().numpy(), host)
