"""Build libmatvis_emu.so: libmatvis_hip's own sources compiled for the CPU.

Every ``kernel<<<grid, block, shmem, stream>>>(args)`` launch in the sources
is rewritten to ``emu_launch(&kernel, emu_cfg(grid, block, shmem, stream),
args)``, which runs each thread of each block in turn (hip/hip_runtime.h in
this directory provides the device vocabulary).  The result has the same C ABI
as the GPU library and is built with AddressSanitizer, so a kernel that reads
or writes outside the buffers it is given is reported on the CPU.

Wave intrinsics and MFMA are emulated as wave collectives over the documented
operand layouts (hip/hip_runtime.h).  Output goes to tests/cpu_emu/_build/
(git-ignored).
"""

from __future__ import annotations

import re
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
SRC = ROOT / "matvis_amd" / "csrc"
OUT = HERE / "_build"
CLANG = "/opt/rocm/llvm/bin/clang++"
SOURCES = ["api.hip", "rotate.hip", "beam.hip", "zmat.hip", "herk.hip"]


def _name_start(text: str, end: int) -> int:
    """Index where the kernel name ending just before `end` begins."""
    i = end
    while i > 0 and text[i - 1] in " \t\n\\":
        i -= 1
    depth = 0
    while i > 0:
        c = text[i - 1]
        if c == ">":
            depth += 1
        elif c == "<":
            depth -= 1
        elif depth == 0 and not (c.isalnum() or c in "_:"):
            break
        i -= 1
    return i


def rewrite_launches(text: str) -> str:
    out = []
    pos = 0
    for m in re.finditer(r"<<<", text):
        if m.start() < pos:
            continue
        start = _name_start(text, m.start())
        close = text.index(">>>", m.end())
        cfg = text[m.end():close]
        after = close + 3
        j = after
        while text[j] in " \t\n\\":
            j += 1
        if text[j] != "(":
            raise ValueError(f"launch without argument list near {text[start:j]!r}")
        name = text[start:m.start()].rstrip(" \t\n\\")
        out.append(text[pos:start])
        out.append(f"emu_launch(&{name}, emu_cfg({cfg}), ")
        pos = j + 1
    out.append(text[pos:])
    return "".join(out)


def build(verbose: bool = False) -> Path:
    OUT.mkdir(exist_ok=True)
    gen = OUT / "src"
    gen.mkdir(exist_ok=True)
    for f in SRC.glob("*.h"):
        (gen / f.name).write_text(f.read_text())
    srcs = []
    for name in SOURCES:
        text = rewrite_launches((SRC / name).read_text())
        dst = gen / (Path(name).stem + ".cpp")
        dst.write_text(text.replace('#include "../../include/matvis_hip.h"',
                                    f'#include "{ROOT}/include/matvis_hip.h"'))
        srcs.append(str(dst))
    srcs.append(str(HERE / "emu_runtime.cpp"))
    lib = OUT / "libmatvis_emu.so"
    cmd = [CLANG, "-std=c++17", "-O1", "-g", "-fPIC", "-shared",
           "-fsanitize=address", "-fno-omit-frame-pointer",
           "-Wno-unused-function", "-I", str(HERE), "-I", str(gen),
           *srcs, "-o", str(lib)]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return lib


def asan_runtime() -> str:
    out = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True, check=True).stdout
    return out.strip()


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
