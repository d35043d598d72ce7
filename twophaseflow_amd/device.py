"""Array hand-off between callers and libtpf.

The HIP entry points take raw device pointers.  Callers may pass torch
tensors already on the GPU (zero copy, the fast path) or NumPy arrays (one
host->device copy in and one device->host copy out, the drop-in path for
code written against the reference's NumPy backend).  Outputs come back in
the caller's kind.  PyTorch is used only for device memory and the stream.
"""

from __future__ import annotations

import numpy as np
import torch


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError(
            "twophaseflow_amd needs a HIP device; no CPU fallback exists")


def stream_handle() -> int:
    """hipStream_t of torch's current stream, as an integer for ctypes."""
    return int(torch.cuda.current_stream().cuda_stream)


class Handoff:
    """Remembers whether a call came in as NumPy or torch."""

    def __init__(self, template):
        self.numpy = not isinstance(template, torch.Tensor)
        self.device = None if self.numpy else template.device

    def to_device(self, x, shape=None) -> torch.Tensor:
        require_gpu()
        if isinstance(x, torch.Tensor):
            t = x
            if t.device.type != "cuda":
                t = t.to("cuda")
            if t.dtype != torch.float64:
                t = t.to(torch.float64)
            t = t.contiguous()
        else:
            t = torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float64))).to("cuda")
        if shape is not None and tuple(t.shape) != tuple(shape):
            t = t.reshape(shape)
        return t

    def back(self, t: torch.Tensor):
        if self.numpy:
            return t.cpu().numpy()
        return t


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous() and t.dtype == torch.float64
    return t.data_ptr()
