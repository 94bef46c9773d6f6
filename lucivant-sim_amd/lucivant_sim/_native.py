"""ctypes binding of liblsim.so, the gfx950 kernels behind the engines.

The library is loaded once, after torch: torch brings its HIP runtime (libamdhip64.so.7) into the
process first and the library binds to that same runtime, so device pointers of torch tensors are
valid arguments. There is no CPU fallback: if the library or a GPU is missing, every engine call
raises :class:`NativeError`.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Any

import torch  # noqa: F401  (loads the HIP runtime the library links against)

LIB_PATH = Path(os.environ.get("LSIM_LIBRARY", Path(__file__).with_name("liblsim.so")))

C128, C64 = 0, 1
#: status codes of include/lsim.h
OK, EINVAL, EHIP = 0, 1, 2
MAX_K = 5
MAX_TERMS = 64

_c_void = ctypes.c_void_p
_c_int = ctypes.c_int
_c_ll = ctypes.c_longlong
_c_double = ctypes.c_double
_c_ull_p = ctypes.POINTER(ctypes.c_ulonglong)
_c_int_p = ctypes.POINTER(ctypes.c_int)

# name -> (argtypes, restype); the order of each list is the C prototype in include/lsim.h
SIGNATURES: dict[str, tuple[list[Any], Any]] = {
    "lsim_last_error": ([], ctypes.c_char_p),
    "lsim_abi_version": ([], _c_int),
    "lsim_device_info": ([_c_int_p, ctypes.POINTER(_c_ll), ctypes.c_char_p, _c_int], _c_int),
    "lsim_host_cumsum": ([_c_void, _c_void, ctypes.c_size_t], _c_int),
    "lsim_sv_set_basis": ([_c_void, _c_int, _c_int, _c_ll, _c_ll, _c_void], _c_int),
    "lsim_sv_apply_dense": ([_c_void, _c_int, _c_int, _c_ll, _c_int, _c_int_p, _c_void, _c_void, _c_void], _c_int),
    "lsim_set_dense_tiling": ([_c_int], _c_int),
    "lsim_set_dense_mfma": ([_c_int], _c_int),
    "lsim_set_dense_streaming": ([_c_int], _c_int),
    "lsim_sv_apply_diag": ([_c_void, _c_int, _c_int, _c_ll, _c_int, _c_int_p, _c_void, _c_void, _c_void], _c_int),
    "lsim_sv_apply_tile": ([_c_void, _c_int, _c_int, _c_ll, _c_int_p, _c_int, _c_void, _c_void, _c_void], _c_int),
    "lsim_sv_scale": ([_c_void, _c_int, _c_int, _c_ll, _c_double, _c_double, _c_void], _c_int),
    "lsim_sv_copy_states": ([_c_void, _c_int, _c_int, _c_ll, _c_ll, _c_void, _c_void, _c_void], _c_int),
    "lsim_sv_axpy": ([_c_void, _c_void, _c_int, _c_int, _c_ll, _c_double, _c_double, _c_void], _c_int),
    "lsim_sv_probabilities": ([_c_void, _c_int, _c_int, _c_ll, _c_void, _c_void], _c_int),
    "lsim_reduce_workspace_bytes": ([_c_int, _c_ll, _c_int], ctypes.c_size_t),
    "lsim_sv_reduced_dm": ([_c_void, _c_int, _c_int, _c_ll, _c_int, _c_int_p, _c_void, _c_void, _c_void], _c_int),
    "lsim_sv_pauli_expect": (
        [_c_void, _c_int, _c_int, _c_ll, _c_int, _c_ull_p, _c_ull_p, _c_void, _c_void, _c_void],
        _c_int,
    ),
    "lsim_sv_pauli_expect_noalloc": (
        [_c_void, _c_int, _c_int, _c_ll, _c_int, _c_ull_p, _c_ull_p, _c_void, _c_void],
        _c_int,
    ),
    "lsim_dm_probabilities": ([_c_void, _c_int, _c_int, _c_ll, _c_void, _c_void], _c_int),
    "lsim_sv_apply_perm": ([_c_void, _c_int, _c_int, _c_ll, _c_int, _c_int_p, _c_int_p, ctypes.c_uint, _c_void, _c_void, _c_void], _c_int),
    "lsim_host_plan_fused": ([_c_int, _c_void, _c_void, _c_void, _c_int, _c_int, _c_int, _c_int, _c_int, _c_void], _c_int),
    "lsim_host_plan_tiles": ([_c_int, _c_void, _c_void, _c_void, _c_int, ctypes.c_double, _c_int, _c_int, _c_int, _c_void], _c_int),
    "lsim_host_plan_sizes": ([_c_void, _c_void, _c_void, _c_void], _c_int),
    "lsim_host_plan_read": ([_c_void, _c_void, _c_void, _c_void, _c_void, _c_void], _c_int),
    "lsim_host_plan_free": ([_c_void], None),
    "lsim_graph_capture_begin": ([_c_void], _c_int),
    "lsim_graph_capture_end": ([_c_void, ctypes.POINTER(ctypes.c_void_p)], _c_int),
    "lsim_graph_nodes": ([_c_void], _c_ll),
    "lsim_graph_launch": ([_c_void, _c_void], _c_int),
    "lsim_graph_destroy": ([_c_void], _c_int),
    "lsim_marginal_workspace_bytes": ([_c_int, _c_ll, _c_int, _c_void], ctypes.c_size_t),
    "lsim_marginal": ([_c_void, _c_int, _c_ll, _c_int, _c_void, _c_void, _c_void, _c_void], _c_int),
    "lsim_program_max_qubits": ([_c_int], _c_int),
    "lsim_program_run": (
        [_c_int, _c_int, _c_ll, _c_void, _c_void, _c_void, _c_void, _c_void, _c_void, _c_void]
        + [_c_void, _c_void, _c_void, _c_int, _c_void, _c_void, _c_void, _c_void, _c_void],
        _c_int,
    ),
    "lsim_dm_pauli_expect": (
        [_c_void, _c_int, _c_int, _c_ll, _c_int, _c_ull_p, _c_ull_p, _c_void, _c_void, _c_void],
        _c_int,
    ),
}


class TileOp(ctypes.Structure):
    """``lsim_tile_op`` of include/lsim.h."""

    _fields_ = [("kind", _c_int), ("k", _c_int), ("t", _c_int * 2), ("mat", _c_ll)]


TILE_DENSE, TILE_DIAG = 0, 1
#: Operator kinds of a host plan (LSIM_PLAN_DENSE, LSIM_PLAN_DIAG)
PLAN_DENSE, PLAN_DIAG = 0, 1


class NativeError(RuntimeError):
    """The gfx950 library is missing or one of its calls failed."""


_lib: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
    """The library with every prototype declared (raises if it isn't built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise NativeError(f"{LIB_PATH} is not built: run `python lucivant-sim_amd/build.py` (hipcc, gfx950)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def loaded() -> bool:
    """Whether this process has loaded the library (and so may hold device state)."""
    return _lib is not None


def call(name: str, *args: Any) -> None:
    """Call an entry point and raise NativeError with the library's message on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.lsim_last_error().decode(errors="replace")
        raise NativeError(f"{name}: {msg} (status {rc})")


def int_array(values: list[int] | tuple[int, ...]) -> ctypes.Array[ctypes.c_int]:
    return (ctypes.c_int * len(values))(*values)


def ull_array(values: list[int]) -> ctypes.Array[ctypes.c_ulonglong]:
    return (ctypes.c_ulonglong * len(values))(*values)


def device_info() -> dict[str, Any]:
    lib = load()
    cu = ctypes.c_int()
    hbm = ctypes.c_longlong()
    arch = ctypes.create_string_buffer(64)
    call("lsim_device_info", ctypes.byref(cu), ctypes.byref(hbm), arch, 64)
    del lib
    return {"compute_units": cu.value, "hbm_bytes": hbm.value, "arch": arch.value.decode()}
