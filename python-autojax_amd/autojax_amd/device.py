"""Device plumbing: torch-ROCm tensors as device buffers and the current stream.

torch only allocates memory and supplies the stream; every computation runs
in libautojax_amd.  Nothing here falls back to the CPU: without a GPU the
helpers raise.
"""

from __future__ import annotations

import numpy as np
import torch

_NP2T = {
    np.dtype(np.float64): torch.float64,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.complex128): torch.complex128,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.int32): torch.int32,
}


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("autojax_amd needs a ROCm GPU (torch.cuda.is_available() is False); there is no CPU fallback")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def ptr(t) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def to_device(a, dtype=None) -> torch.Tensor:
    """numpy array or tensor -> contiguous device tensor (complex kept as complex128)."""
    dev = require_gpu()
    if isinstance(a, torch.Tensor):
        t = a
    else:
        arr = np.ascontiguousarray(a) if dtype is None else np.ascontiguousarray(a, dtype=dtype)
        t = torch.from_numpy(arr)
    return t.to(dev, non_blocking=False).contiguous()


def empty(shape, dtype=torch.float64) -> torch.Tensor:
    return torch.empty(shape, dtype=dtype, device=require_gpu())


def workspace(nbytes: int) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=require_gpu())


def to_host(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


def synchronize():
    torch.cuda.current_stream().synchronize()
