"""The summation order the device marginal (lsim_marginal) follows, pinned against NumPy on the
CPU: for every keep set, the marginal is 0 + S(c_0) + S(c_1) + ... over the chunks of
C = 2^min(low, 13) contiguous entries whose kept bits spell the output index, in increasing
order (low = dropped qubits below the lowest kept one), each chunk summed by NumPy's pairwise
rule. The model below restates that order in plain NumPy operations; it must equal the
reference's `tensor.sum(axis=drop)` (engines.py:267-275) bit for bit, which is what lets sampling
from the device marginal give the reference's counts. The GPU test compares the kernel itself
with NumPy."""

from __future__ import annotations

import itertools

import numpy as np
import pytest

from lucivant_sim.engines import NUMPY_CHUNK_BITS, device_marginal_pays, marginal


def _pairwise(rows: np.ndarray) -> np.ndarray:
    """NumPy's pairwise sum of each row (power-of-two length), vectorised over rows."""
    m = rows.shape[1]
    if m < 8:
        acc = np.zeros(rows.shape[0])
        for i in range(m):
            acc = acc + rows[:, i]
        return acc
    if m <= 128:
        r = [rows[:, j].copy() for j in range(8)]
        for i in range(8, m, 8):
            for j in range(8):
                r[j] = r[j] + rows[:, i + j]
        return ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]))
    return _pairwise(rows[:, : m // 2]) + _pairwise(rows[:, m // 2 :])


def _model(probs: np.ndarray, n: int, keep: list[int]) -> np.ndarray:
    low = 0
    while low not in keep:
        low += 1
    c = min(low, NUMPY_CHUNK_BITS)
    sums = _pairwise(probs.reshape(-1, 1 << c)) if c else probs.copy()
    chunk = np.arange(len(sums))
    j = np.zeros(len(sums), dtype=np.int64)
    for i, q in enumerate(keep):
        j |= ((chunk >> (q - c)) & 1) << i
    order = np.argsort(j, kind="stable")  # per output, chunks stay in increasing order
    per = len(sums) >> len(keep)
    grouped = sums[order].reshape(1 << len(keep), per)
    out = np.zeros(1 << len(keep))
    for t in range(per):
        out = out + grouped[:, t]
    return out


def _probs(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = rng.random(1 << n) ** 3  # uneven magnitudes, so the order of additions shows
    return p / p.sum()


@pytest.mark.parametrize("n", [5, 8])
def test_model_equals_numpy_for_every_keep_set(n: int) -> None:
    p = _probs(n, n)
    for m in range(1, n):
        for keep in itertools.combinations(range(n), m):
            assert np.array_equal(_model(p, n, list(keep)), marginal(p, n, list(keep))), keep


@pytest.mark.parametrize(
    "n, keep",
    [
        (16, [15]),
        (16, [0]),
        (18, [17]),  # runs of 2^17 entries: pieces of 8192
        (18, [16, 17]),
        (18, [0, 17]),
        (18, [14, 15]),  # runs of 2^14: two pieces each
        (20, [13]),  # a run of exactly 8192
        (20, [3, 19]),
        (20, [2, 5, 9, 12]),
        (20, [0, 1, 2]),
        (20, [13, 14, 15, 16, 17, 18, 19]),
    ],
)
def test_model_equals_numpy_on_long_runs(n: int, keep: list[int]) -> None:
    p = _probs(n, 3 * n + keep[0])
    assert np.array_equal(_model(p, n, keep), marginal(p, n, keep))


def test_a_sequential_or_fully_pairwise_order_would_differ() -> None:
    """The pin has teeth: summing each output's entries left to right, or pairwise over all of
    them, misses NumPy's values for some keep set."""
    n = 18
    p = _probs(n, 1)
    seq_differs = pair_differs = False
    for keep in ([17], [16, 17], [14, 15], [3, 17]):
        want = marginal(p, n, keep)
        moved = np.moveaxis(p.reshape([2] * n), [n - 1 - q for q in keep], range(len(keep))).reshape(1 << len(keep), -1)
        seq = np.array([np.add.accumulate(np.ascontiguousarray(r))[-1] for r in moved])
        pair = _pairwise(np.ascontiguousarray(moved))
        seq_differs |= not np.array_equal(seq, want)
        pair_differs |= not np.array_equal(pair, want)
    assert seq_differs and pair_differs


def test_device_marginal_is_chosen_for_short_chains_only() -> None:
    assert device_marginal_pays(26, [25])  # chunks of 8192, 2^12 per entry
    assert device_marginal_pays(26, list(range(10, 26)))
    assert device_marginal_pays(26, list(range(10)))  # 2^16 entries added in order per output
    assert not device_marginal_pays(26, list(range(6)))  # 2^20: left to the host
    assert not device_marginal_pays(26, [0])  # 2^25 entries added in order
    assert not device_marginal_pays(26, list(range(26)))  # nothing to sum
