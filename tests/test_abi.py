"""The C ABI on the CPU: liblsim.so loads, exports every function include/lsim.h declares, and
refuses bad arguments with a status and a message before touching a GPU."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "lsim.h"
LIB = ROOT / "lucivant-sim_amd" / "lucivant_sim" / "liblsim.so"


def declared() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(lsim_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib() -> ctypes.CDLL:
    if not LIB.exists():
        pytest.fail(f"{LIB} is not built: run python lucivant-sim_amd/build.py")
    import torch  # noqa: F401  (the HIP runtime the library binds to)

    from lucivant_sim import _native as nat

    return nat.load()


def test_every_declared_symbol_is_exported(lib: ctypes.CDLL) -> None:
    names = declared()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name


def test_ctypes_signatures_cover_the_header() -> None:
    from lucivant_sim import _native as nat

    assert sorted(nat.SIGNATURES) == declared()


def test_abi_version(lib: ctypes.CDLL) -> None:
    assert lib.lsim_abi_version() == 3


def _err(lib: ctypes.CDLL) -> str:
    return lib.lsim_last_error().decode()


def test_argument_errors_are_reported_without_a_gpu(lib: ctypes.CDLL) -> None:
    from lucivant_sim import _native as nat

    t = nat.int_array([0, 0])
    fake = 0x1000  # never dereferenced: validation fails first
    assert lib.lsim_sv_apply_dense(None, 0, 4, 1, 1, nat.int_array([0]), fake, None, None) == 1
    assert "NULL" in _err(lib)
    assert lib.lsim_sv_apply_dense(fake, 0, 4, 1, 6, nat.int_array([0] * 6), fake, None, None) == 1
    assert "k must be between 1 and 5" in _err(lib)
    assert lib.lsim_sv_apply_dense(fake, 0, 4, 1, 2, t, fake, None, None) == 1
    assert "given twice" in _err(lib)
    assert lib.lsim_sv_apply_dense(fake, 0, 4, 1, 1, nat.int_array([4]), fake, None, None) == 1
    assert "outside" in _err(lib)
    assert lib.lsim_sv_apply_diag(fake, 7, 4, 1, 1, nat.int_array([0]), fake, None, None) == 1
    assert "dtype" in _err(lib)
    assert lib.lsim_sv_reduced_dm(fake, 0, 4, 1, 3, nat.int_array([0, 1, 2]), fake, fake, None) == 1
    assert "1 or 2 bits" in _err(lib)
    xs = nat.ull_array([1 << 5])
    assert lib.lsim_sv_pauli_expect(fake, 0, 4, 1, 1, xs, nat.ull_array([0]), fake, fake, None) == 1
    assert "beyond" in _err(lib)
    assert lib.lsim_sv_set_basis(fake, 0, 4, 1, 16, None) == 1
    assert "outside the state" in _err(lib)
    # the allocation-free expectation refuses partials beyond its 4 MiB scratch before any GPU call
    x64 = nat.ull_array([1] * 64)
    z64 = nat.ull_array([0] * 64)
    assert lib.lsim_sv_pauli_expect_noalloc(fake, 0, 20, 1024, 64, x64, z64, fake, None) == 1
    assert "scratch" in _err(lib)
    assert lib.lsim_sv_pauli_expect_noalloc(fake, 0, 4, 1, 1, xs, nat.ull_array([0]), fake, None) == 1
    assert "beyond" in _err(lib)
    assert lib.lsim_reduce_workspace_bytes(26, 1, 1) > 0


def test_variant_switches_read_set_and_clamp(lib: ctypes.CDLL) -> None:
    """The tuning switches return their previous value, read without changing on a negative
    argument, and the streaming mode clamps to 0..2."""
    old_s, old_m, old_t = lib.lsim_set_dense_streaming(-1), lib.lsim_set_dense_mfma(-1), lib.lsim_set_dense_tiling(1)
    try:
        assert old_s == 1 and old_m == (1 << 3) | (1 << 4) | (1 << 5)
        assert lib.lsim_set_dense_streaming(7) == old_s
        assert lib.lsim_set_dense_streaming(-1) == 2
        assert lib.lsim_set_dense_streaming(0) == 2 and lib.lsim_set_dense_streaming(-1) == 0
        assert lib.lsim_set_dense_mfma(1 << 5) == old_m and lib.lsim_set_dense_mfma(-1) == 1 << 5
        assert lib.lsim_set_dense_tiling(0) == 1 and lib.lsim_set_dense_tiling(-1) == 0
    finally:
        lib.lsim_set_dense_streaming(old_s)
        lib.lsim_set_dense_mfma(old_m)
        lib.lsim_set_dense_tiling(old_t)


def test_host_cumsum_is_numpys_bit_for_bit(lib: ctypes.CDLL) -> None:
    """The sampler's host running sum: the same float64 additions in the same order as
    numpy.cumsum, including a leading negative zero, single elements and empty input."""
    import numpy as np

    rng = np.random.default_rng(8)
    for x in (rng.random(1000) ** 6, rng.normal(size=4097), np.array([-0.0, 1e-300, 1.0]), np.array([2.5]), np.zeros(0)):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.full(x.size, np.nan)
        assert lib.lsim_host_cumsum(x.ctypes.data, out.ctypes.data, x.size) == 0
        want = np.cumsum(x)
        assert np.array_equal(out, want) and np.array_equal(np.signbit(out), np.signbit(want))
    assert lib.lsim_host_cumsum(None, None, 3) == 1
    assert "NULL" in _err(lib)


def test_engines_refuse_to_run_without_a_gpu() -> None:
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import lucivant_sim as ls
    from lucivant_sim._native import NativeError

    with pytest.raises(NativeError, match="no GPU"):
        ls.statevector('OPENQASM 3; include "stdgates.inc"; qubit q; h q;')
