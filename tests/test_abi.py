"""The C ABI of libmatvis_hip.so: every declared symbol is exported and bound,
and argument checks fail with MV_ERR_ARG before anything reaches the GPU."""

import ctypes as C
import re
from pathlib import Path

import pytest

from matvis_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "matvis_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|size_t)\s+(mv_[a-z0-9_]+)\(", text,
                                 re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 18
    for name in names:
        assert hasattr(lib, name), name


def test_ctypes_table_matches_header():
    assert sorted(_lib.SIGNATURES) == declared_symbols()
    text = HEADER.read_text()
    for name, argtypes in _lib.SIGNATURES.items():
        m = re.search(rf"\b{name}\(([^;]*)\);", text, re.S)
        params = [p for p in m.group(1).split(",") if p.strip() not in ("", "void")]
        assert len(params) == len(argtypes), name


def test_abi_version():
    assert _lib.lib().mv_abi_version() == _lib.ABI_VERSION == 3


def test_tiles_and_workspace():
    lib = _lib.lib()
    assert lib.mv_herk_tile(_lib.FORM_C128) == 64
    assert lib.mv_herk_tile(_lib.FORM_C64) == 64
    assert lib.mv_herk_tile(_lib.FORM_SPLIT_BF16) == 128
    assert lib.mv_herk_tile(_lib.FORM_SPLIT_F16) == 128
    assert lib.mv_herk_tile(7) == 0
    # m = 700 -> 11 tiles of 64 -> 66 upper tiles; 3 slices of c128 tiles
    assert lib.mv_herk_workspace_bytes(_lib.FORM_C128, 700, 3) == 3 * 66 * 64 * 64 * 16
    # split forms: FP32 partial tiles plus FP64 row norms of the diagonal
    assert lib.mv_herk_workspace_bytes(_lib.FORM_SPLIT_BF16, 700, 2) == (
        2 * 21 * 128 * 128 * 8 + 2 * 21 * 128 * 8)
    assert lib.mv_rotate_cull_workspace_bytes(0) == 256
    assert lib.mv_rotate_cull_workspace_bytes(2048 * 70) == 512
    # za bands: complex64 rows of 363 -> 128 KB holds 45 rows, bands of
    # 42; 91 rows -> 3 bands: counts, band starts, workgroup starts (256 B
    # each, aligned), then (az, za) and band, rank, order per position
    n = 60000
    per = lambda b: -(-b // 256) * 256  # noqa: E731
    assert lib.mv_beam_interp_banded_workspace_bytes(0, 91, 363, n) == (
        3 * 256 + 2 * per(8 * n) + 3 * per(4 * n))
    # complex128 rows of 4100 do not fit 4 in 128 KB: the gather path
    assert lib.mv_beam_interp_banded_workspace_bytes(1, 91, 4100, n) == 0


def _err():
    return _lib.last_error()


def test_argument_errors_do_not_launch():
    lib = _lib.lib()
    rc = lib.mv_herk(9, 1, 1, None, None, None, 10, 10, 1, 10, 10, None, 1,
                     1.0, None, 64, 0, None, 0, None)
    assert rc == _lib.MV_ERR_ARG and "unknown form" in _err()
    rc = lib.mv_rotate_cull(None, 10, None, None, 0, 10, 10, None, 10, None, 1,
                            None, None, None, 0, None)
    assert rc == _lib.MV_ERR_ARG and "null pointer" in _err()
    dummy = C.c_void_p(16)
    rc = lib.mv_herk(_lib.FORM_SPLIT_BF16, 2, 5, dummy, dummy, None, 10, 16, 1,
                     16, 16, None, 1, 1.0, dummy, 128, 0, dummy, 1 << 30, None)
    assert rc == _lib.MV_ERR_ARG and "unsupported split" in _err()
    rc = lib.mv_beam_interp(dummy, 1, 1, 1, 1, 4, 4, dummy, 2, 0, dummy, 1, 4,
                            4, None, dummy, 4, None)
    assert rc == _lib.MV_ERR_ARG and "order must be 1 or 3" in _err()
    rc = lib.mv_beam_interp_banded(dummy, 0, 1, 2, 2, 91, 363, dummy, 3, 0,
                                   dummy, 1, 100, 100, None, dummy, 100,
                                   dummy, 0, None)
    assert rc == _lib.MV_ERR_ARG and "workspace too small" in _err()
    rc = lib.mv_zmat_fused(dummy, 4, 1e8, dummy, 8, dummy, None, dummy, 1, 8,
                           1, None, 2, 2, 8, None, 0, 1, 1.0, dummy, None,
                           None, 16, 8, None)
    assert rc == _lib.MV_ERR_ARG and "seg_stride" in _err()
    rc = lib.mv_coherency_rotate(None, None, 10, None, None, 10, 1, None, 10,
                                 None, None, 1, 1, 0, None, None)
    assert rc == _lib.MV_ERR_ARG and "null pointer" in _err()
    rc = lib.mv_coherency_rotate(dummy, dummy, 10, dummy, dummy, 5, 1, dummy,
                                 10, dummy, dummy, 1, 1, 0, dummy, None)
    assert rc == _lib.MV_ERR_ARG and "bad sizes" in _err()


def test_call_raises_value_error():
    with pytest.raises(ValueError, match="mv_vis_expand"):
        _lib.call("mv_vis_expand", None, 4, 64, 2, 1, None, 4, None, 1, None)
