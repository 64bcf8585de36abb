"""The library's own HIP sources built for the CPU under AddressSanitizer
(tests/cpu_emu) and run through two GPU tests' code paths: the split HERK
on a 129-row problem (direct LDS copies, lane-half exchanges, partial
tiles, the reduce) and the BCRS corrections.  Any access outside the
buffers the product hands a kernel, or a barrier under divergent control
flow, fails here without a GPU.  Skipped where clang's sanitizer runtime is
absent."""

import sys
from pathlib import Path

import pytest

EMU = Path(__file__).parent / "cpu_emu"


def _runtime():
    sys.path.insert(0, str(EMU))
    try:
        import build_emu

        path = build_emu.asan_runtime()
    except Exception:  # no clang here
        return None
    return path if path and Path(path).exists() else None


@pytest.mark.skipif(_runtime() is None, reason="no AddressSanitizer runtime")
def test_kernels_memory_safe_on_cpu():
    import run_on_cpu

    sel = ("herk_split_accuracy and 129-77-2-2-3 or "
           "bcrs_update_matches_oracle and 0")
    r = run_on_cpu.run(["-m", "gpu", "-x", "-q", "-k", sel,
                        str(EMU.parent / "test_gpu_kernels.py"),
                        str(EMU.parent / "test_erfa.py")], timeout=900)
    assert not r.asan_reports, r.asan_reports[0][:3000]
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert " passed" in r.stdout
