"""SZ3-style interpolation predictor (the API of latec/predictor.py).

``InterpPredictor`` keeps the reference's plugin interface (``predict(recon,
known, pos)``, ``stage_masks``, ``stream_flag``, ``center``/``end_mode``,
``tunable``/``fast_eb_ratio``).  Its prediction runs on the GPU
(``latec_interp_predict``); inside ``latec_amd.compress``/``decompress`` the
codec does not call ``predict`` at all but runs the fused
predict+quantize stage kernels over the whole schedule.
"""

from __future__ import annotations

import atexit
import ctypes

import numpy as np

from . import _device as D
from ._lib import check, lib, ptr
from .bitstream import FLAG_CUBIC, FLAG_INTERP
from .levels import LazyStageMasks, stage_counts

END_QUAD = 1 << 0
END_EXTRAP = 1 << 1
END_MODE_MAX = END_QUAD | END_EXTRAP


def default_interp_center(ndim: int) -> int:
    """predictor.py:21-27: average the odd axes up to 3-D, first axis above."""
    return 1 if ndim >= 4 else 0


def default_interp_end_mode(ndim: int) -> int:
    """predictor.py:51-66: quadratic line ends by default."""
    return END_QUAD


class Plan:
    """Owner of one native schedule plan (shape, channels, levels, stride, block)."""

    def __init__(self, shape, channels, levels, stride, block):
        self.shape = tuple(int(n) for n in shape)
        self.channels = int(channels)
        self.levels, self.stride, self.block = int(levels), int(stride), int(block)
        arr = np.asarray(self.shape, np.int64)
        h = ctypes.c_void_p()
        check(lib().latec_plan_create(len(self.shape), ptr(arr), self.channels, self.levels,
                                      self.stride, self.block, ctypes.byref(h)))
        self.handle = h
        self.n_stages = lib().latec_plan_n_stages(h)
        self.counts, self.strides, self.weights, self.n_classes = [], [], [], []
        c, s, w, k = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        for i in range(self.n_stages):
            check(lib().latec_plan_stage_info(h, i, ctypes.byref(c), ctypes.byref(s),
                                              ctypes.byref(w), ctypes.byref(k)))
            self.counts.append(c.value)
            self.strides.append(s.value)
            self.weights.append(w.value)
            self.n_classes.append(k.value)
        self.offsets = [lib().latec_plan_stage_offset(h, i) for i in range(self.n_stages + 1)]
        self.npoints = self.offsets[-1]

    def class_info(self, stage, cls):
        out = np.zeros(4 + 3 * len(self.shape), np.int64)
        check(lib().latec_plan_class_info(self.handle, stage, cls, ptr(out)))
        nd = len(self.shape)
        return {"parity": int(out[0]), "step": int(out[1]), "offset": int(out[2]),
                "size": int(out[3]), "starts": tuple(out[4:4 + nd]),
                "sizes": tuple(out[4 + nd:4 + 2 * nd]), "strides": tuple(out[4 + 2 * nd:])}

    def close(self):
        if self.handle:
            lib().latec_plan_destroy(self.handle)
            self.handle = None


_PLANS: dict = {}


def get_plan(shape, channels, levels, stride, block) -> Plan:
    key = (tuple(int(n) for n in shape), int(channels), int(levels), int(stride), int(block))
    p = _PLANS.get(key)
    if p is None:
        p = Plan(*key)
        _PLANS[key] = p
    return p


@atexit.register
def _close_plans():
    for p in _PLANS.values():
        try:
            p.close()
        except Exception:
            pass
    _PLANS.clear()


class InterpPredictor:
    """predictor.py:159-317 behind the same constructor and attributes."""

    tunable = True
    fast_eb_ratio = 0.9

    def __init__(self, order: str = "cubic", levels: int = 4, anchor_stride: int = 16,
                 anchor_block: int = 4, center=None, end_mode=None):
        if order not in ("linear", "cubic"):
            raise ValueError("order must be 'linear' or 'cubic'")
        if end_mode is not None and not 0 <= end_mode <= END_MODE_MAX:
            raise ValueError(f"end_mode must be in [0, {END_MODE_MAX}]")
        self.order = order
        self.levels = levels
        self.anchor_stride = anchor_stride
        self.anchor_block = anchor_block
        self.center = center
        self.end_mode = end_mode
        self.stream_flag = FLAG_INTERP | (FLAG_CUBIC if order == "cubic" else 0)
        self.checkpoint_hash = b"\0" * 16
        self._cache: dict = {}

    def _schedule(self, shape):
        """|known| before a stage -> stage index (predictor.py:219-250)."""
        key = tuple(int(n) for n in shape)
        hit = self._cache.get(key)
        if hit is not None:
            return hit
        counts = stage_counts(key, self.levels, self.anchor_stride, self.anchor_block)
        sched, covered = {}, 0
        for i, n in enumerate(counts):
            if i and n:
                sched[covered] = i
            covered += n
        self._cache[key] = sched
        return sched

    def stage_masks(self, shape, levels, anchor_stride, anchor_block):
        return LazyStageMasks(shape, levels, anchor_stride, anchor_block)

    def resolved(self, ndim: int):
        center = self.center if self.center is not None else default_interp_center(ndim)
        end = self.end_mode if self.end_mode is not None else default_interp_end_mode(ndim)
        return center, end

    def predict(self, recon, known, pos=None):
        """(C, |pos|) float32 predictions of the stage that follows ``known``."""
        from .quantizer import opts_for

        shape = recon.shape[1:]
        stage = self._schedule(shape).get(int(np.count_nonzero(known)))
        if stage is None:
            raise ValueError("known mask does not match the interp schedule")
        if pos is None:
            raise ValueError("pos is required: the GPU predictor predicts stage points")
        plan = get_plan(shape, recon.shape[0], self.levels, self.anchor_stride,
                        self.anchor_block)
        n = plan.counts[stage]
        if int(np.count_nonzero(pos)) != n:
            raise ValueError("pos does not match the interp schedule")
        center, end = self.resolved(len(shape))
        o = opts_for(1 << 15, cubic=self.order == "cubic", center=center, end_mode=end)
        drec = D.to_device(np.ascontiguousarray(recon, np.float32).reshape(-1))
        out = D.empty(recon.shape[0] * n, np.float32)
        check(lib().latec_interp_predict(plan.handle, stage, ptr(drec), ptr(out),
                                         ctypes.cast(ctypes.pointer(o), ctypes.c_void_p),
                                         ptr(None)))
        return D.to_host(out).reshape(recon.shape[0], n)
