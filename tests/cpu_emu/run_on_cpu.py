"""Run the GPU tests' code paths on the CPU against libmatvis_emu.so.

Usage (AddressSanitizer must be preloaded; `run()` below does that):

    python tests/cpu_emu/run_on_cpu.py [pytest args...]

The product's device layer is pointed at CPU memory: ``matvis_amd._device``
allocates torch CPU tensors (malloc'd, so AddressSanitizer guards them) and
``torch.cuda`` streams/events become no-ops.  Every kernel runs through the
emulated library built from the same sources (build_emu.py), so any access
outside the buffers the product code hands a kernel aborts the run with a
report naming the kernel and the test.  Wave ballots, shuffles and MFMA are
emulated with the instruction's operand layout and one FP32 rounding per MFMA
instruction, so the tests' value assertions hold here too; a value failure
here is a layout or indexing mistake, not emulation noise.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]


def patch_device_layer():
    import contextlib

    import torch

    import matvis_amd._device as D
    import matvis_amd._lib as L

    L.LIB_PATH = HERE / "_build" / "libmatvis_emu.so"
    cpu = torch.device("cpu")
    D.require_gpu = lambda: cpu
    D.free_memory = lambda: 1 << 36
    D.stream_ptr = lambda stream=None: None

    def to_dev(a, dtype=None):
        # a host->device copy never aliases its source; on the CPU it must
        # not either
        import numpy as np

        a = np.array(a if dtype is None else np.asarray(a, dtype=dtype),
                     copy=True, order="C")
        return torch.from_numpy(a)

    D.to_dev = to_dev

    class Event:
        def __init__(self, *a, **k):
            pass

        def record(self, *a, **k):
            pass

        def synchronize(self):
            pass

        def elapsed_time(self, other):
            return 1.0  # no device clock here; a fixed nonzero stand-in

    class Stream:
        cuda_stream = 0

        def __init__(self, *a, **k):
            pass

        def wait_event(self, e):
            pass

        def wait_stream(self, s):
            pass

        def synchronize(self):
            pass

    torch.cuda.Event = Event
    torch.cuda.Stream = Stream
    torch.cuda.current_stream = lambda *a, **k: Stream()
    torch.cuda.stream = lambda s: contextlib.nullcontext()
    torch.cuda.is_available = lambda: True
    torch.cuda.set_device = lambda d: None
    torch.cuda.synchronize = lambda *a, **k: None
    torch.cuda.empty_cache = lambda: None
    real_empty = torch.empty

    def empty(*a, pin_memory=False, **k):
        return real_empty(*a, **k)

    torch.empty = empty


def main(argv):
    sys.path.insert(0, str(ROOT))
    sys.path.insert(0, str(ROOT / "tests"))
    patch_device_layer()
    import pytest

    args = argv or ["-m", "gpu", str(ROOT / "tests" / "test_gpu_kernels.py"),
                    str(ROOT / "tests" / "test_gpu_simulate.py")]
    return pytest.main(["-p", "no:cacheprovider", "-q", "-rN",
                        "--tb=no", "-v", *args])


def run(pytest_args=(), timeout=3000) -> subprocess.CompletedProcess:
    """Build the emulated library and run main() under AddressSanitizer."""
    sys.path.insert(0, str(HERE))
    import build_emu

    build_emu.build()
    env = dict(os.environ)
    # the sanitizer runtime must load first; keep whatever else is preloaded
    env["LD_PRELOAD"] = " ".join(
        p for p in (build_emu.asan_runtime(), os.environ.get("LD_PRELOAD", ""))
        if p)
    log = HERE / "_build" / "asan"
    for old in (HERE / "_build").glob("asan.*"):
        old.unlink()
    env["ASAN_OPTIONS"] = ("detect_leaks=0:halt_on_error=1:abort_on_error=0:"
                           f"log_path={log}")
    env["MV_CPU_EMU"] = "1"
    r = subprocess.run([sys.executable, __file__, "-p", "no:faulthandler",
                        *pytest_args], env=env, capture_output=True,
                       text=True, timeout=timeout)
    reports = [p.read_text() for p in (HERE / "_build").glob("asan.*")]
    r.asan_reports = [t for t in reports if "ERROR: AddressSanitizer" in t
                      or "emu: barrier" in t]
    return r


if __name__ == "__main__":
    if os.environ.get("MV_CPU_EMU") != "1":
        r = run(sys.argv[1:])
        sys.stdout.write(r.stdout[-20000:])
        sys.stderr.write(r.stderr[-5000:])
        for t in r.asan_reports:
            sys.stderr.write(t[:6000])
        sys.exit(r.returncode)
    sys.exit(main(sys.argv[1:]))
