"""Build libmatvis_hip.so for gfx950 with hipcc, in-tree.

Usage: ``python -m matvis_amd.csrc.build`` or ``build()`` from Python.
Compiles every ``*.hip`` file in this directory into one shared library next
to them.  Flags: ``--offload-arch=gfx950`` only, ``-O3``, and packed-FP32
instruction generation disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_NAME = "libmatvis_hip.so"
SOURCES = ["api.hip", "rotate.hip", "beam.hip", "zmat.hip", "herk.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def build(verbose: bool = False, jobs: int = 4) -> Path:
    """Compile all sources and link ``libmatvis_hip.so``; return its path."""
    out = HERE / LIB_NAME
    with tempfile.TemporaryDirectory(prefix="mvbuild") as tmp:
        objs = []
        procs = []
        for src in SOURCES:
            obj = Path(tmp) / (Path(src).stem + ".o")
            cmd = [HIPCC, *FLAGS, "-c", str(HERE / src), "-o", str(obj)]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.STDOUT)))
            objs.append(str(obj))
            if len([p for _, p in procs if p.poll() is None]) >= jobs:
                procs[0][1].wait()
        failed = []
        for src, p in procs:
            text, _ = p.communicate()
            text = text.decode(errors="replace")
            # hipcc's host pass does not know the packed-fp32 feature flag
            lines = [ln for ln in text.splitlines()
                     if "packed-fp32-ops" not in ln and ln.strip()]
            if p.returncode != 0:
                failed.append(src)
                print("\n".join(lines), file=sys.stderr)
            elif verbose and lines:
                print("\n".join(lines))
        if failed:
            raise RuntimeError(f"hipcc failed for {failed}")
        tmp_out = Path(tmp) / LIB_NAME
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs,
               "-o", str(tmp_out)]
        subprocess.run(cmd, check=True)
        os.replace(tmp_out, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
