"""Device plumbing: torch tensors as the array carrier, HIP streams, checks."""

from __future__ import annotations

import numpy as np
import torch

from . import _lib


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("pycwb_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(stream: torch.cuda.Stream | None = None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def as_dev(x, dtype: torch.dtype, device: torch.device) -> torch.Tensor:
    """Contiguous device tensor of ``dtype`` (copies host arrays once)."""
    if isinstance(x, torch.Tensor):
        t = x.to(device=device, dtype=dtype)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x))).to(device=device, dtype=dtype)
    return t.contiguous()


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else int(t.data_ptr())


def lib():
    return _lib.load()
