"""ctypes binding of ``libmatvis_hip.so`` (C ABI in ``include/matvis_hip.h``).

The library is built in-tree by ``matvis_amd/csrc/build.py``.  There is no
fallback: if it is missing or fails to load, :func:`lib` raises.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "csrc" / "libmatvis_hip.so"

#: the library's mv_abi_version() this binding is written for
ABI_VERSION = 3

MV_OK = 0
MV_ERR_ARG = -1
MV_ERR_HIP = -2
MV_ERR_CAPACITY = -3

FORM_C128 = 0
FORM_C64 = 1
FORM_SPLIT_BF16 = 2
FORM_SPLIT_F16 = 3

_P = C.c_void_p
_I = C.c_int
_L = C.c_int64
_D = C.c_double
_S = C.c_size_t

#: Every exported entry point with its argument types (restype is int unless
#: listed in _RESTYPES).  Kept in the order of include/matvis_hip.h.
SIGNATURES: dict[str, list] = {
    "mv_last_error": [C.c_char_p, _S],
    "mv_abi_version": [],
    "mv_rotate": [_P, _L, _P, _L, _P, _I, _L, _P],
    "mv_rotate_cull_workspace_bytes": [_L],
    "mv_rotate_cull": [_P, _L, _P, _P, _L, _L, _L, _P, _L, _P, _I, _P, _P,
                       _P, _S, _P],
    "mv_bcrs_update": [_P, _L, _L, _P, _D, _P, _L, _P],
    "mv_coherency_rotate": [_P, _P, _L, _P, _P, _L, _I, _P, _L, _P, _P, _L,
                            _I, _I, _P, _P],
    "mv_beam_prefilter": [_P, _I, _L, _I, _I, _P],
    "mv_beam_interp": [_P, _I, _I, _I, _I, _I, _I, _P, _I, _I, _P, _I, _L,
                       _L, _P, _P, _L, _P],
    "mv_beam_interp_azza": [_P, _I, _I, _I, _I, _I, _I, _P, _I, _I, _P, _P,
                            _L, _P, _L, _P],
    "mv_beam_interp_banded_workspace_bytes": [_I, _I, _I, _L],
    "mv_beam_interp_banded": [_P, _I, _I, _I, _I, _I, _I, _P, _I, _I, _P, _I,
                              _L, _L, _P, _P, _L, _P, _S, _P],
    "mv_beam_gaussian": [_P, _I, _I, _P, _I, _L, _L, _P, _P, _I, _L, _P],
    "mv_exptau": [_P, _I, _D, _P, _I, _L, _L, _P, _I, _L, _P],
    "mv_zmat_parts": [_P, _P, _P, _I, _P, _I, _I, _I, _L, _P, _I, _P],
    "mv_zmat_fused": [_P, _I, _D, _P, _L, _P, _P, _P, _I, _L, _I, _P, _I,
                      _I, _L, _P, _I, _I, _D, _P, _P, _P, _L, _L, _P],
    "mv_split_z": [_P, _I, _L, _I, _L, _L, _L, _I, _I, _D, _P, _P, _P, _L,
                   _L, _P],
    "mv_herk_tile": [_I],
    "mv_herk_workspace_bytes": [_I, _I, _I],
    "mv_herk_max_slices": [_I, _I, _L],
    "mv_herk_slices": [_I, _I, _I, _I, _I, _L, _I],
    "mv_herk": [_I, _I, _I, _P, _P, _P, _I, _L, _I, _L, _L, _P, _I, _D, _P,
                _L, _I, _P, _S, _P],
    "mv_herk_noalloc": [_I, _I, _I, _P, _P, _P, _I, _L, _I, _L, _L, _P, _D,
                        _P, _L, _I, _P],
    "mv_herk_compute": [_I, _I, _I, _P, _P, _P, _I, _L, _I, _L, _L, _P, _I,
                        _D, _P, _L, _I, _P, _S, _P],
    "mv_herk_reduce": [_I, _I, _I, _P, _P, _P, _I, _L, _I, _L, _L, _P, _I,
                       _D, _P, _L, _I, _P, _S, _P],
    "mv_herk_problem_bytes": [],
    "mv_herk_grouped_workspace_bytes": [_I, _I, _P, _I],
    "mv_herk_grouped": [_I, _I, _I, _I, _P, _P, _I, _P, _S, _P],
    "mv_vacc_sum": [_P, _I, _L, _P, _L, _P],
    "mv_vis_expand": [_P, _L, _I, _I, _I, _P, _L, _P, _I, _P],
}

_RESTYPES = {
    "mv_rotate_cull_workspace_bytes": _S,
    "mv_beam_interp_banded_workspace_bytes": _S,
    "mv_herk_workspace_bytes": _S,
    "mv_herk_problem_bytes": _S,
    "mv_herk_grouped_workspace_bytes": _S,
}

_lock = threading.Lock()
_lib = None


class HipError(RuntimeError):
    """A libmatvis_hip entry point returned a non-zero status."""

    def __init__(self, name: str, code: int, msg: str):
        super().__init__(f"{name} failed ({code}): {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load (once) and return the library handle; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m matvis_amd.csrc.build` (hipcc, gfx950). "
                "matvis_amd has no CPU fallback."
            )
        handle = C.CDLL(str(LIB_PATH))
        for name, argtypes in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = argtypes
            fn.restype = _RESTYPES.get(name, _I)
        got = handle.mv_abi_version()
        if got != ABI_VERSION:
            raise ImportError(
                f"{LIB_PATH} has ABI version {got}, this binding needs "
                f"{ABI_VERSION}: rebuild it with `python -m "
                "matvis_amd.csrc.build`")
        _lib = handle
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(512)
    lib().mv_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def call(name: str, *args) -> int:
    """Call an int-status entry point; raise :class:`HipError` on failure."""
    rc = getattr(lib(), name)(*args)
    if rc != MV_OK:
        err = last_error()
        if rc == MV_ERR_ARG:
            raise ValueError(f"{name}: {err}")
        raise HipError(name, rc, err)
    return rc
