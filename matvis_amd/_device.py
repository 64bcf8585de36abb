"""Device buffers and streams (PyTorch-ROCm as the allocator, nothing more).

All compute runs in ``libmatvis_hip.so``; torch provides device memory, the
current stream and pinned host buffers.  Without a GPU these raise.
"""

from __future__ import annotations

import numpy as np
import torch

_NP2TORCH = {
    np.dtype(np.float64): torch.float64,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.complex128): torch.complex128,
    np.dtype(np.complex64): torch.complex64,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.uint8): torch.uint8,
}


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError(
            "matvis_amd runs on an AMD GPU through libmatvis_hip.so; "
            "torch.cuda.is_available() is False"
        )
    return torch.device("cuda", torch.cuda.current_device())


def tdtype(dtype) -> torch.dtype:
    return _NP2TORCH[np.dtype(dtype)]


def empty(shape, dtype) -> torch.Tensor:
    return torch.empty(shape, dtype=tdtype(dtype), device=require_gpu())


def zeros(shape, dtype) -> torch.Tensor:
    return torch.zeros(shape, dtype=tdtype(dtype), device=require_gpu())


def to_dev(a, dtype=None) -> torch.Tensor:
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(a).to(require_gpu())


def as_dev(x, dtype) -> torch.Tensor:
    """A contiguous device tensor of `dtype` holding x (numpy or torch)."""
    td = tdtype(dtype)
    if isinstance(x, torch.Tensor):
        t = x.to(device=require_gpu(), dtype=td)
        return t.contiguous()
    return to_dev(np.asarray(x), dtype)


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def stream_ptr(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def to_host(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


def free_memory() -> int:
    require_gpu()
    return int(torch.cuda.mem_get_info()[0])
