"""Array plumbing between caller arrays (NumPy or torch) and device tensors.

PyTorch is used only for device memory and streams; the arithmetic runs in
the HIP kernels of libwilson_amd.
"""

from __future__ import annotations

from typing import Any, Tuple

import numpy as np
import torch

_NP2T = {np.dtype(np.complex64): torch.complex64, np.dtype(np.complex128): torch.complex128}


def to_device(x: Any, dtype=None) -> Tuple[torch.Tensor, str, Any]:
    """Return (contiguous complex CUDA tensor, kind, original-device-or-None)."""
    if isinstance(x, torch.Tensor):
        t = x
        kind = "torch"
        orig = t.device
        if not t.is_complex():
            t = t.to(torch.complex128)
        if dtype is not None and t.dtype != dtype:
            t = t.to(dtype)
        if not t.is_cuda:
            t = t.to("cuda")
        return t.contiguous(), kind, orig
    a = np.asarray(x)
    if not np.iscomplexobj(a):
        a = a.astype(np.complex128)
    if a.dtype not in _NP2T:
        a = a.astype(np.complex128)
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t, "numpy", None


def from_device(t: torch.Tensor, kind: str, orig) -> Any:
    if kind == "numpy":
        return t.cpu().numpy()
    if orig is not None and orig.type != "cuda":
        return t.to(orig)
    return t


def torch_dtype_of(x: Any):
    if isinstance(x, torch.Tensor):
        return x.dtype if x.is_complex() else torch.complex128
    a = np.asarray(x)
    return _NP2T.get(a.dtype, torch.complex128)
