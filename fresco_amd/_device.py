"""Device-memory plumbing: torch-ROCm tensors carry x / grad to the C ABI.

torch is used only for device memory and streams; all arithmetic of the path
runs in libfresco_hip's HIP kernels.
"""
from __future__ import annotations

import numpy as np
import torch


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("fresco_amd needs a ROCm GPU (torch.cuda.is_available() is False)")


def device():
    require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle():
    return ctypes_stream(torch.cuda.current_stream())


def ctypes_stream(stream):
    import ctypes

    return ctypes.c_void_p(stream.cuda_stream)


def as_device_f64(x, n_expected=None):
    """Return (tensor, was_torch): a contiguous float64 device tensor of x.

    Accepts torch tensors (any device / dtype), numpy arrays, lists."""
    if isinstance(x, torch.Tensor):
        t = x.detach()
        if t.device.type != "cuda" or t.dtype != torch.float64 or not t.is_contiguous():
            t = t.to(device=device(), dtype=torch.float64).contiguous()
        t = t.reshape(-1)
        was_torch = True
    else:
        a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())
        t = torch.from_numpy(a).to(device())
        was_torch = False
    if n_expected is not None and t.numel() != n_expected:
        raise ValueError("expected %d coordinates, got %d" % (n_expected, t.numel()))
    return t, was_torch
