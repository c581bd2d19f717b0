"""Linear error-bounded quantizer (the API of latec/quantizer.py) on the GPU.

code = rint((x - pred) / 2eb) + R in [1, 2R), 0 = outlier stored exactly;
the bound is verified with the decoder's own arithmetic in float32
(quantizer.py:28-120).  Both directions run as liblatec kernels.
"""

from __future__ import annotations

import numpy as np

from . import _device as D
from ._lib import InterpOpts, check, lib, ptr, stream_ptr

DEFAULT_RADIUS = 1 << 15


def opts_for(radius: int, round_output=False, cubic=1, center=0, end_mode=0) -> InterpOpts:
    o = InterpOpts()
    o.cubic, o.center, o.end_mode = int(cubic), int(center), int(end_mode)
    o.radius = int(radius)
    o.code_bytes = 2 if radius <= 32768 else 4
    if round_output is True:
        o.round_mode = 1
    elif round_output:
        span, offset = round_output
        o.round_mode, o.span, o.offset = 2, float(span), float(offset)
    return o


def _opts_ptr(o: InterpOpts):
    import ctypes

    return ctypes.cast(ctypes.pointer(o), ctypes.c_void_p)


def quantize_device(dx, dpred, n: int, eb: float, o: InterpOpts, d_hist=None):
    """Device quantize: (d_codes, d_out_rank, d_out_val, n_out) with the
    outliers sorted by position."""
    codes = D.empty(n, np.uint16 if o.code_bytes == 2 else np.uint32)
    rank = D.empty(max(n, 1), np.uint32)
    val = D.empty(max(n, 1), np.float32)
    cnt = D.zeros(1, np.uint32)
    hist = d_hist if d_hist is not None else D.zeros(lib().latec_hist_words(o.radius), np.uint32)
    check(lib().latec_quantize(ptr(dx), ptr(dpred), int(n), float(eb), _opts_ptr(o),
                               ptr(codes), ptr(rank), ptr(val), ptr(cnt), ptr(hist), stream_ptr()))
    k = int(cnt.item())
    if k > 1:  # the kernel appends outliers in any order: list them by position
        nn = np.array([n], np.int64)
        words = int(lib().latec_zero_ranks_words(1, ptr(nn)))
        scratch = D.empty(words, np.uint32)
        check(lib().latec_order_outliers(1, ptr(codes), o.code_bytes, ptr(np.zeros(1, np.int64)),
                                         ptr(nn), ptr(dx), ptr(rank), ptr(val), ptr(cnt),
                                         ptr(scratch), words, stream_ptr()))
    return codes, rank, val, k, hist


def quantize(x, pred, eb, radius=DEFAULT_RADIUS, round_output=False):
    """(codes uint32, outlier values float32 in scan order): quantizer.py:28-87."""
    if eb <= 0:
        raise ValueError(f"error bound must be > 0, got {eb}")
    x = np.ascontiguousarray(x, np.float32).ravel()
    pred = np.ascontiguousarray(pred, np.float32).ravel()
    if x.shape != pred.shape:
        raise ValueError(f"shape mismatch: x {x.shape} vs pred {pred.shape}")
    if len(x) == 0:
        return np.zeros(0, np.uint32), np.zeros(0, np.float32)
    o = opts_for(radius, round_output)
    codes, _rank, val, k, _ = quantize_device(D.to_device(x), D.to_device(pred), len(x), eb, o)
    hc = D.to_host(codes, np.uint16 if o.code_bytes == 2 else np.uint32).astype(np.uint32)
    return hc, D.to_host(val[:k]).astype(np.float32)


def dequantize(pred, codes, outlier_vals, eb, radius=DEFAULT_RADIUS):
    """quantizer.py:90-111."""
    pred = np.ascontiguousarray(pred, np.float32).ravel()
    codes = np.asarray(codes, np.uint32).ravel()
    outlier_vals = np.ascontiguousarray(outlier_vals, np.float32).ravel()
    zeros = np.flatnonzero(codes == 0).astype(np.uint32)
    if len(zeros) != len(outlier_vals):
        raise ValueError(
            f"outlier count mismatch: {len(zeros)} zeros vs {len(outlier_vals)} values")
    if len(pred) == 0:
        return np.zeros(0, np.float32)
    o = opts_for(radius)
    hc = codes.astype(np.uint16) if o.code_bytes == 2 else codes
    out = D.empty(len(pred), np.float32)
    # keep every device buffer referenced until the kernel has finished
    d_pred, d_codes = D.to_device(pred), D.to_device(hc)
    d_rank = D.to_device(zeros if len(zeros) else np.zeros(1, np.uint32))
    d_val = D.to_device(outlier_vals if len(outlier_vals) else np.zeros(1, np.float32))
    cnt = D.to_device(np.array([len(zeros)], np.uint32))
    check(lib().latec_dequantize(ptr(d_pred), ptr(d_codes), len(pred), ptr(d_rank), ptr(d_val),
                                 ptr(cnt), float(eb), _opts_ptr(o), ptr(out), stream_ptr()))
    D.torch().cuda.current_stream().synchronize()
    return D.to_host(out)
