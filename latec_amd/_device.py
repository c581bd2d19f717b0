"""Device buffers: torch-ROCm tensors used purely as HBM allocations."""

from __future__ import annotations

import numpy as np

_TORCH = None


def torch():
    global _TORCH
    if _TORCH is None:
        import torch as t

        _TORCH = t
    return _TORCH


def device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("latec_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return t.device("cuda", t.cuda.current_device())


def to_device(a: np.ndarray):
    """Host numpy -> device tensor (same bytes)."""
    t = torch()
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        return t.from_numpy(a.view(np.int16)).to(device())
    if a.dtype == np.uint32:
        return t.from_numpy(a.view(np.int32)).to(device())
    return t.from_numpy(a).to(device())


def to_host(x, dtype=None) -> np.ndarray:
    """Device tensor -> numpy, reinterpreting to ``dtype`` (e.g. uint16)."""
    a = x.detach().cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def empty(n, dtype):
    t = torch()
    tmap = {np.float32: t.float32, np.float64: t.float64, np.uint16: t.int16,
            np.int16: t.int16, np.uint32: t.int32, np.int32: t.int32, np.uint8: t.uint8,
            np.int64: t.int64}
    return t.empty(int(n), dtype=tmap[np.dtype(dtype).type], device=device())


def zeros(n, dtype):
    z = empty(n, dtype)
    z.zero_()
    return z
