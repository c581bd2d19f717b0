"""Device buffers: torch-ROCm tensors used purely as HBM allocations."""

from __future__ import annotations

import threading

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


_STAGING = threading.local()


def staging(n, dtype):
    """(pinned host tensor, numpy view) of ``n`` elements of numpy ``dtype``:
    host memory that copies to and from HBM at the link's pinned rate.  One
    buffer per thread and torch dtype, grown to the largest request; the
    views stay valid until this thread's next request of the same dtype."""
    t = torch()
    bufs = getattr(_STAGING, "bufs", None)
    if bufs is None:
        bufs = _STAGING.bufs = {}
    td = _tdtype(dtype)
    buf = bufs.get(td)
    if buf is None or buf.numel() < n:
        buf = bufs[td] = None
        buf = bufs[td] = t.empty(max(int(n), 1), dtype=td, pin_memory=True)
    view = buf[:int(n)]
    return view, view.numpy().view(dtype)


def to_host_staged(x, dtype=None) -> np.ndarray:
    """Device tensor -> numpy through the pinned staging buffer of x's dtype
    (a synchronous copy); the array is valid as ``staging`` says."""
    view, a = staging(x.numel(), _NPDTYPE[x.dtype])
    view.copy_(x.detach().reshape(-1))
    return a.view(dtype) if dtype is not None else a


def to_host_pinned(x, dtype=None) -> np.ndarray:
    """Device tensor -> a new numpy array in pinned host memory (one copy at
    the link's pinned rate; torch caches freed pinned blocks, so a repeated
    call of the same size does not pin pages again).  The array keeps its
    buffer alive."""
    t = torch()
    host = t.empty(x.shape, dtype=x.dtype, pin_memory=True)
    host.copy_(x.detach())
    a = host.numpy()
    return a.view(dtype) if dtype is not None else a


def _tdtype(dtype):
    t = torch()
    tmap = {np.float32: t.float32, np.float64: t.float64, np.uint16: t.int16,
            np.int16: t.int16, np.uint32: t.int32, np.int32: t.int32, np.uint8: t.uint8,
            np.int64: t.int64}
    return tmap[np.dtype(dtype).type]


class _NpDtypes(dict):
    def __missing__(self, td):
        t = torch()
        m = {t.float32: np.float32, t.float64: np.float64, t.int16: np.int16,
             t.int32: np.int32, t.uint8: np.uint8, t.int64: np.int64}
        self.update(m)
        return m[td]


_NPDTYPE = _NpDtypes()


def empty(n, dtype):
    return torch().empty(int(n), dtype=_tdtype(dtype), device=device())


def zeros(n, dtype):
    z = empty(n, dtype)
    z.zero_()
    return z


def minmax(d_x, segments) -> np.ndarray:
    """(nseg, 2) float32 [min, max] of device float32 segments (offset, length)
    of d_x, NaN-propagating as NumPy (latec_minmax)."""
    from ._lib import check, lib, ptr

    off = np.ascontiguousarray([int(a) for a, _ in segments], np.int64)
    ln = np.ascontiguousarray([int(b) for _, b in segments], np.int64)
    words = int(lib().latec_minmax_words(len(off)))
    scratch = empty(words, np.float32)
    out = empty(2 * len(off), np.float32)
    check(lib().latec_minmax(len(off), ptr(d_x), ptr(off), ptr(ln), ptr(out), ptr(scratch),
                             words, ptr(None)))
    return to_host(out, np.float32).reshape(-1, 2)
