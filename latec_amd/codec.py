"""Closed-loop plain codec (the API of latec/codec.py).

``compress``/``decompress`` keep the reference's signatures, tune sweep,
header and stats keys (codec.py:45-322).  With the package's
``InterpPredictor`` the whole dyadic schedule runs as fused GPU stage
kernels (predict + quantize + reconstruct in place, one launch per stage),
the per-stage level fit and tables run on the GPU, and the host packs the
stage records (multi-threaded stack rANS / literal zstd) into the same
LATEC001 v4 stream.  Any other predictor object goes through the generic
loop: prediction in Python, quantization and fit on the GPU.
"""

from __future__ import annotations

import ctypes
import os
import time
from concurrent.futures import ThreadPoolExecutor
from typing import NamedTuple

import numpy as np

from . import _device as D
from . import rans as R
from ._lib import Fit, check, lib, ptr
from .bitstream import (DTYPE_CODES, DTYPE_IDS, END_MODE_SHIFT, FLAG_CUBIC, FLAG_INTERP,
                        FLAG_RANS, VERSION, VERSION_MULTILANE, Header, pack_record,
                        pack_stage, parse_record, read_stream, unpack_record_into, unpack_stage,
                        write_stream)
from .levels import mask_tiles, stage_ebs, stage_masks
from .predictor import (END_MODE_MAX, InterpPredictor, default_interp_center,
                        default_interp_end_mode, get_plan)
from .quantizer import dequantize, opts_for, quantize

_THREADS = int(os.environ.get("LATEC_HOST_THREADS", min(16, os.cpu_count() or 1)))


def _optr(o):
    return ctypes.cast(ctypes.pointer(o), ctypes.c_void_p)


# ---------------------------------------------------------------------------
# GPU fast path
# ---------------------------------------------------------------------------

class GpuEncodeResult:
    """Device-side result of the fused stage kernels for one encode."""

    def __init__(self, plan, d_recon, d_codes, d_hist, d_rank, d_val, d_cnt, opts):
        self.plan, self.d_recon, self.d_codes, self.d_hist = plan, d_recon, d_codes, d_hist
        self.d_rank, self.d_val, self.d_cnt, self.opts = d_rank, d_val, d_cnt, opts


class Workspace:
    """Reusable device buffers for one plan + radius (sized to the field)."""

    def __init__(self, plan, radius):
        C, N, ns = plan.channels, plan.npoints, plan.n_stages
        cb = 2 if radius <= 32768 else 4
        self.radius = radius
        self.recon = D.empty(C * N, np.float32)
        self.codes = D.empty(C * N, np.uint16 if cb == 2 else np.uint32)
        self.hist = D.empty(ns * lib().latec_hist_words(radius), np.uint32)
        self.rank = D.empty(C * N, np.uint32)
        self.val = D.empty(C * N, np.float32)
        self.cnt = D.empty(ns, np.uint32)


_WS: dict = {}


def workspace(plan, radius) -> Workspace:
    key = (id(plan), radius)
    ws = _WS.get(key)
    if ws is None:
        _WS.clear()  # one field at a time: release the previous one
        ws = Workspace(plan, radius)
        _WS[key] = ws
    return ws


def encode_device(d_x, plan, ebs, o, graph: bool = False, ws: Workspace = None):
    """Run the fused encode over all stages of ``plan`` on device input
    ``d_x`` ((C*N,) float32).  Returns a GpuEncodeResult (asynchronous)."""
    ws = ws or workspace(plan, o.radius)
    ws.recon.zero_()
    ws.hist.zero_()
    ws.cnt.zero_()
    ebs_arr = np.ascontiguousarray(ebs, np.float64)
    fn = lib().latec_interp_encode_graph if graph else None
    if graph:
        check(fn(plan.handle, ptr(d_x), ptr(ws.recon), ptr(ws.codes), ptr(ws.hist),
                 ptr(ws.rank), ptr(ws.val), ptr(ws.cnt), ptr(ebs_arr), _optr(o), ptr(None)))
    else:
        check(lib().latec_interp_encode(plan.handle, 0, plan.n_stages, ptr(d_x), ptr(ws.recon),
                                        ptr(ws.codes), ptr(ws.hist), ptr(ws.rank), ptr(ws.val),
                                        ptr(ws.cnt), ptr(ebs_arr), _optr(o), ptr(None)))
    return GpuEncodeResult(plan, ws.recon, ws.codes, ws.hist, ws.rank, ws.val, ws.cnt, o)


def fit_device(res: GpuEncodeResult, precision: int = 24):
    """Per-stage (kexp, level, shape, owx, bits) fitted on the GPU."""
    plan, o = res.plan, res.opts
    fits = (Fit * plan.n_stages)()
    check(lib().latec_fit_const(plan.handle, 0, plan.n_stages, ptr(res.d_hist),
                                ptr(res.d_codes), o.radius, o.code_bytes, precision,
                                ctypes.cast(fits, ctypes.c_void_p), ptr(None)))
    return fits


def pack_device(res: GpuEncodeResult, fits, zstd_level: int, threads: int = _THREADS):
    """Copy codes/outliers to the host and pack every stage record:
    (payload, record lengths, outliers per stage)."""
    return _pack_device(res, fits, zstd_level, threads, None)[:3]


def pack_device_framed(res: GpuEncodeResult, fits, zstd_level: int, threads: int = _THREADS):
    """pack_device plus the outer zstd frame of the payload at zstd_level
    (write_stream's body), started while the last large stages finish:
    (payload, record lengths, outliers per stage, frame)."""
    return _pack_device(res, fits, zstd_level, threads, zstd_level)


def _pack_device(res, fits, zstd_level, threads, frame_level):
    plan, o = res.plan, res.opts
    C, ns = plan.channels, plan.n_stages
    cb = o.code_bytes
    codes = D.to_host_staged(res.d_codes, np.uint16 if cb == 2 else np.uint32)
    cnt = D.to_host(res.d_cnt, np.uint32)
    code_off = np.array([C * plan.offsets[s] for s in range(ns)], np.int64)
    n_codes = np.array([C * plan.counts[s] for s in range(ns)], np.int64)
    n_out = cnt.astype(np.int64)
    outl = np.zeros(max(int(n_out.sum()), 1), np.float32)
    out_off = np.zeros(ns, np.int64)
    at = 0
    for s in range(ns):
        out_off[s] = at
        k = int(n_out[s])
        if k:
            lo = int(code_off[s])
            rk = D.to_host(res.d_rank[lo:lo + k], np.uint32)
            vl = D.to_host(res.d_val[lo:lo + k])
            outl[at:at + k] = vl[np.argsort(rk, kind="stable")]
            at += k
    rows, row_off, pos = [], np.zeros(ns, np.int64), 0
    for s in range(ns):
        f = fits[s]
        row = (R.cdf_row(f.kexp, f.shape, f.owx, f.level) if n_codes[s]
               else np.zeros((2 << 1) + 1, np.uint32))
        rows.append(row)
        row_off[s] = pos
        pos += len(row)
    cdf_rows = np.ascontiguousarray(np.concatenate(rows), np.uint32)
    cap = int(64 * ns + 4 * (n_codes.sum() + 2 * ns) + 4 * n_out.sum() + (1 << 16))
    payload = np.empty(cap, np.uint8)
    rec_len = np.zeros(ns, np.int64)
    args = (ns, ptr(codes), cb, ptr(code_off), ptr(n_codes), ptr(outl), ptr(out_off),
            ptr(n_out), o.radius, int(zstd_level), ctypes.cast(fits, ctypes.c_void_p),
            ptr(cdf_rows), ptr(row_off), ptr(payload), cap, ptr(rec_len), int(threads))
    if frame_level is None:
        size = check(lib().latec_pack_stages(*args))
        return payload[:size].tobytes(), rec_len, n_out, None
    fcap = int(lib().latec_zstd_bound(cap))
    frame = np.empty(fcap, np.uint8)
    flen = ctypes.c_int64()
    size = check(lib().latec_pack_stages_framed(*args, int(frame_level), ptr(frame), fcap,
                                                ctypes.byref(flen)))
    return payload[:size].tobytes(), rec_len, n_out, frame[:flen.value].tobytes()


def decode_device(payload: bytes, plan, ebs, o, threads: int = _THREADS):
    """Parse and entropy-decode every stage record (host threads), then run
    the fused decode kernels; returns the device reconstruction."""
    C, ns = plan.channels, plan.n_stages
    arr = np.frombuffer(payload, np.uint8)
    offs, n_out, off = [], [], 0
    for s in range(ns):
        offs.append(off)
        _, hlen, _, _, _, _, _, length = parse_record(arr, off)
        n_out.append((length - 16 - hlen) // 4)
        off += length
    if off != len(payload):
        raise ValueError("trailing bytes in LATEC payload")
    cb = o.code_bytes
    n_all = C * plan.npoints
    # codes land in pinned memory (one H2D copy at the pinned rate); the
    # outlier ranks and values are packed per stage and copied stage by stage
    t_codes, codes = D.staging(n_all, np.uint16 if cb == 2 else np.uint32)
    out_at = np.concatenate([[0], np.cumsum(n_out)]).astype(np.int64)
    rank = np.empty(max(int(out_at[-1]), 1), np.uint32)
    val = np.empty(max(int(out_at[-1]), 1), np.float32)

    def one(s):
        lo, a, b = C * plan.offsets[s], int(out_at[s]), int(out_at[s + 1])
        unpack_record_into(arr, offs[s], C * plan.counts[s], o.radius, codes[lo:],
                           rank[a:b], val[a:b])

    order = sorted(range(ns), key=lambda s: -plan.counts[s])
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(one, order))
    ws = workspace(plan, o.radius)
    ws.codes[:n_all].copy_(t_codes)
    t_rank, t_val = D.torch().from_numpy(rank.view(np.int32)), D.torch().from_numpy(val)
    for s in range(ns):
        a, b = int(out_at[s]), int(out_at[s + 1])
        if b > a:
            lo = C * plan.offsets[s]
            ws.rank[lo:lo + b - a].copy_(t_rank[a:b])
            ws.val[lo:lo + b - a].copy_(t_val[a:b])
    ws.cnt.copy_(D.to_device(np.asarray(n_out, np.uint32)))
    ws.recon.zero_()
    ebs_arr = np.ascontiguousarray(ebs, np.float64)
    check(lib().latec_interp_decode(plan.handle, 0, ns, ptr(ws.codes), ptr(ws.rank),
                                    ptr(ws.val), ptr(ws.cnt), ptr(ws.recon), ptr(ebs_arr),
                                    _optr(o), ptr(None)))
    return ws.recon


def _interp_opts(predictor, ndim, radius, round_output):
    center, end = predictor.resolved(ndim)
    return opts_for(radius, round_output, cubic=predictor.order == "cubic", center=center,
                    end_mode=end)


# ---------------------------------------------------------------------------
# generic predictor loop (codec.py:384-510)
# ---------------------------------------------------------------------------

def _gather_stage(field, pos, n):
    out = np.empty((field.shape[0], n), field.dtype)
    flat_field = field.reshape(field.shape[0], -1)
    at = 0
    for flat in mask_tiles(pos, field.shape[1:]):
        out[:, at:at + len(flat)] = flat_field[:, flat]
        at += len(flat)
    return out


def _scatter_stage(recon, pos, values):
    flat_recon = recon.reshape(recon.shape[0], -1)
    at = 0
    for flat in mask_tiles(pos, recon.shape[1:]):
        flat_recon[:, flat] = values[:, at:at + len(flat)]
        at += len(flat)


def _stage_error_stats(stats, s, values, pred, recon):
    pe = values - pred
    re = np.abs(values - recon)
    stats["stage_pred_sae"][s] += float(np.abs(pe).sum(dtype=np.float64))
    stats["stage_pred_sse"][s] += float(np.square(pe, dtype=np.float64).sum())
    stats["stage_recon_sae"][s] += float(re.sum(dtype=np.float64))
    stats["stage_recon_sse"][s] += float(np.square(re, dtype=np.float64).sum())
    stats["stage_recon_max"][s] = max(stats["stage_recon_max"][s], float(re.max(initial=0.0)))


def _compress_region(field, masks, ebs, predictor, radius, round_output, stats, zstd_level=9):
    c = field.shape[0]
    recon = np.zeros_like(field)
    known = np.zeros(field.shape[1:], bool)
    parts = []
    use_rans = getattr(predictor, "provides_scale", False)
    for s, pos in enumerate(masks):
        n = int(np.count_nonzero(pos))
        if n == 0:
            if use_rans:
                parts.append(pack_stage(np.zeros(0, np.uint32), np.zeros(0, np.float32),
                                        rans_levels=np.zeros(0, np.uint8),
                                        rans_tables=R.build_laplace_tables(ebs[s], radius)))
            else:
                parts.append(pack_stage(np.zeros(0, np.uint32), np.zeros(0, np.float32),
                                        radius=radius, zstd_level=zstd_level))
            continue
        eb = ebs[s]
        scale = None
        if s == 0:
            pred = np.zeros((c, n), np.float32)
            if use_rans:
                scale = np.full((c, n), eb, np.float32)
        else:
            t0 = time.time()
            if use_rans:
                pred, scale = predictor.predict(field_recon_view(recon), known, pos, eb=eb)
            else:
                pred = predictor.predict(recon, known, pos)
            stats["predict_s"] += time.time() - t0
        t0 = time.time()
        values = _gather_stage(field, pos, n)
        codes, outliers = quantize(values, pred, eb, radius, round_output=round_output)
        stage_recon = dequantize(pred, codes, outliers, eb, radius).reshape(c, n)
        _scatter_stage(recon, pos, stage_recon)
        _stage_error_stats(stats, s, values, np.asarray(pred, np.float32).reshape(c, n),
                           stage_recon)
        stats["quantize_s"] += time.time() - t0
        known |= pos
        stats["outliers"] += len(outliers)
        stats["stage_outliers"][s] += len(outliers)
        stats["stage_codes"][s] += n * c
        t0 = time.time()
        if use_rans:
            tables = R.build_laplace_tables(eb, radius)
            levels64 = R.scale_to_level(scale, eb).reshape(-1)
            kexp = R.choose_kexp(codes, radius)
            shape = R.choose_shape(codes, levels64, tables, kexp)
            stats["stage_model_bits"][s] += R.model_bits(codes, levels64, tables, kexp=kexp,
                                                         shape=shape)
            part = pack_stage(codes, outliers, rans_levels=levels64, rans_tables=tables)
        else:
            part = pack_stage(codes, outliers, radius=radius, zstd_level=zstd_level)
        stats["stage_payload_bytes"][s] += len(part)
        parts.append(part)
        stats["entropy_s"] += time.time() - t0
    return b"".join(parts), recon


def field_recon_view(recon):
    return recon


def _decompress_region(payload, masks, ebs, header, predictor):
    c = header.channels
    region = tuple(header.spatial)
    recon = np.zeros((c, *region), np.float32)
    known = np.zeros(region, bool)
    off = 0
    use_rans = bool(header.flags & FLAG_RANS)
    for s, pos in enumerate(masks):
        n = int(np.count_nonzero(pos))
        if n == 0:
            if use_rans:
                _, _, off = unpack_stage(payload, off, rans_levels=np.zeros(0, np.uint8),
                                         rans_tables=R.build_laplace_tables(ebs[s],
                                                                            header.radius))
            else:
                _, _, off = unpack_stage(payload, off, radius=header.radius)
            continue
        if s == 0:
            pred = np.zeros((c, n), np.float32)
            scale = np.full((c, n), ebs[s], np.float32)
        else:
            if use_rans:
                pred, scale = predictor.predict(recon, known, pos, eb=ebs[s])
            elif getattr(predictor, "provides_scale", False):
                pred, _ = predictor.predict(recon, known, pos, eb=ebs[s])
            else:
                pred = predictor.predict(recon, known, pos)
        if use_rans:
            tables = R.build_laplace_tables(ebs[s], header.radius)
            levels64 = R.scale_to_level(scale, ebs[s]).reshape(-1)
            codes, outliers, off = unpack_stage(payload, off, rans_levels=levels64,
                                                rans_tables=tables)
        else:
            codes, outliers, off = unpack_stage(payload, off, radius=header.radius)
        _scatter_stage(recon, pos, dequantize(pred, codes, outliers, ebs[s],
                                              header.radius).reshape(c, n))
        known |= pos
    return recon


def _finalize(canvas, header):
    """codec.py:513-524."""
    out = np.moveaxis(canvas, 0, -1)
    dtype = DTYPE_CODES[header.src_dtype]
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        out = np.clip(np.rint(out), info.min, info.max)
    out = out.astype(dtype, copy=False)
    if header.channels == 1:
        out = out[..., 0]
    return out


def _fast_path(predictor) -> bool:
    return isinstance(predictor, InterpPredictor) and not getattr(predictor, "provides_scale",
                                                                 False)


class _Input:
    """The array handed to ``compress`` as a (C, *S) float32 field.

    An image is 2-D, or 3-D with a last axis of 1 or 3 channels (channel
    last, SZ's convention); any other array is a one-channel field of its
    own rank.  vmin / vmax are the float32 field's extremes, vmax pushed to
    vmin + 1 for a constant field (codec.py:75-97)."""

    def __init__(self, img, upload: bool = False):
        img = np.asarray(img)
        image = img.ndim <= 2 or (img.ndim == 3 and img.shape[-1] in (1, 3))
        chan_last = img if image and img.ndim == 3 else img[..., None]
        self.spatial = tuple(chan_last.shape[:-1])
        self.channels = int(chan_last.shape[-1])
        self.src_dtype = DTYPE_IDS[np.dtype(img.dtype)]
        f32 = np.asarray(chan_last, np.float32)
        self.field = np.ascontiguousarray(np.moveaxis(f32, -1, 0))
        self.d_field = None
        if upload:  # the range from the float32 values in HBM (latec_minmax)
            self.d_field = D.to_device(self.field.reshape(-1))
            self.vmin, self.vmax = map(float, D.minmax(self.d_field, [(0, self.field.size)])[0])
        else:
            self.vmin, self.vmax = float(f32.min()), float(f32.max())
        if self.vmax <= self.vmin:
            self.vmax = self.vmin + 1.0
        self.round_output = bool(np.issubdtype(img.dtype, np.integer))
        self.nbytes = img.nbytes


def _new_stats(n_stages: int) -> dict:
    """The per-encode stats dict (keys of codec.py:120-135)."""
    st = {"predict_s": 0.0, "quantize_s": 0.0, "entropy_s": 0.0, "outliers": 0}
    for key, zero in (("stage_codes", 0), ("stage_outliers", 0), ("stage_payload_bytes", 0),
                      ("stage_model_bits", 0.0), ("stage_pred_sae", 0.0),
                      ("stage_pred_sse", 0.0), ("stage_recon_sae", 0.0),
                      ("stage_recon_sse", 0.0), ("stage_recon_max", 0.0)):
        st[key] = [zero] * n_stages
    return st


class _Candidate(NamedTuple):
    size: int
    sse: float
    ratio: float
    center: int
    end: int
    stream: bytes
    recon: np.ndarray
    stats: dict


def _tune_space(predictor, tune: str, eb_ratio):
    """What the encoder-side search tries (codec.py:171-199): (ratios, centres,
    sweep).  A fixed eb_ratio, or fast mode, or a predictor without knobs gives
    one ratio and one centre; a sweep tries ratios 1.0 .. 0.7 and the three
    centres, base centre first."""
    tunable = getattr(predictor, "tunable", False)
    sweep = bool(tunable and tune != "fast" and eb_ratio is None)
    if eb_ratio is not None:
        ratios = [eb_ratio]
    elif sweep:
        ratios = [1.0, 0.9, 0.8, 0.7]
    else:
        ratios = [getattr(predictor, "fast_eb_ratio", 1.0)]
    base = getattr(predictor, "center", 0)
    centres = [base] + ([c for c in (0, 1, 2) if c != base]
                        if sweep and hasattr(predictor, "center") else [])
    return ratios, centres, sweep


def _pick(cands, slack: float, rd: bool) -> _Candidate:
    """Smallest stream, ties by reconstruction SSE then (ratio, centre, end);
    in rate-distortion mode the lowest SSE among the streams within ``slack``
    of the smallest (codec.py:244-251)."""
    if rd:
        limit = min(c.size for c in cands) * slack
        return min((c for c in cands if c.size <= limit),
                   key=lambda c: (c.sse, c.size, c.ratio, c.center, c.end))
    return min(cands, key=lambda c: (c.size, c.sse, c.ratio, c.center, c.end))


def compress(img, eb, predictor, levels=4, anchor_stride=16, anchor_block=1, radius=1 << 15,
             zstd_level=9, eb_ratio=None, tune="fast", tune_size_slack=1.05, verbose=False,
             *, stream_version=VERSION):
    """latec.compress (codec.py:45-285): (stream, stats).

    The search: encode the base candidate; when sweeping, encode every end
    mode at the first ratio and base centre and keep the end mode of the
    smallest (size, SSE) stream so far; then every ratio x centre at that end
    mode.  A (ratio, centre, end) triple is encoded at most once.  The chosen
    candidate's centre and end mode are left on the predictor.
    ``stream_version=5`` (an addition) writes the multi-lane rANS version with
    the same codes, outliers and model fit; 4 is the reference's layout."""
    if stream_version not in (VERSION, VERSION_MULTILANE):
        raise ValueError(f"unsupported version {stream_version}")
    fast = _fast_path(predictor)
    inp = _Input(img, upload=fast)
    for name, want in (("levels", levels), ("anchor_stride", anchor_stride),
                       ("anchor_block", anchor_block)):
        got = getattr(predictor, name, want)
        if got != want:
            raise ValueError(f"predictor {name}={got} != compress {name}={want}")
    shape = inp.field.shape[1:]
    masks = getattr(predictor, "stage_masks", stage_masks)(shape, levels, anchor_stride,
                                                           anchor_block)
    use_rans = getattr(predictor, "provides_scale", False)
    flags = getattr(predictor, "stream_flag", 0) | (FLAG_RANS if use_rans else 0)
    if stream_version == VERSION_MULTILANE and not fast:
        raise ValueError("stream version 5 needs the package's InterpPredictor")
    if tune not in ("fast", "size", "rd"):
        raise ValueError("tune must be 'fast', 'size', or 'rd'")
    # rank-aware defaults resolved once the rank is known (codec.py:183-187)
    if hasattr(predictor, "center") and predictor.center is None:
        predictor.center = default_interp_center(len(shape))
    if hasattr(predictor, "end_mode") and predictor.end_mode is None:
        predictor.end_mode = default_interp_end_mode(len(shape))
    ratios, centres, sweep = _tune_space(predictor, tune, eb_ratio)
    d_field = inp.d_field
    agg = getattr(predictor, "agg_level", None)

    def header(center, ratio, end):
        return Header(channels=inp.channels, src_dtype=inp.src_dtype, spatial=inp.spatial,
                      eb=float(eb), levels=levels, anchor_stride=anchor_stride,
                      anchor_block=anchor_block, radius=radius,
                      max_radius=getattr(predictor, "max_radius", 0),
                      agg_level=-1 if agg is None else int(agg), vmin=inp.vmin, vmax=inp.vmax,
                      ckpt_hash=getattr(predictor, "checkpoint_hash", b"\0" * 16),
                      flags=flags | (int(end) << END_MODE_SHIFT), interp_center=center,
                      eb_ratio=ratio, version=stream_version)

    tried: dict = {}

    def encode(ratio, center, end):
        key = (ratio, center, end)
        if key in tried:
            return
        if hasattr(predictor, "center"):
            predictor.center = center
        if hasattr(predictor, "end_mode"):
            predictor.end_mode = end
        st = _new_stats(len(masks))
        ebs = stage_ebs(shape, levels, anchor_stride, anchor_block, eb, ratio)
        if fast:
            payload, recon, frame = _compress_region_gpu_impl(
                inp.field, d_field, ebs, predictor, radius, inp.round_output, st, zstd_level,
                stream_version)
        else:
            payload, recon = _compress_region(inp.field, masks, ebs, predictor, radius,
                                              inp.round_output, st, zstd_level)
            frame = None
        t0 = time.time()
        stream = write_stream(header(center, ratio, end), payload, zstd_level, frame=frame)
        st["entropy_s"] += time.time() - t0
        st["raw_payload_bytes"] = len(payload)
        st["compressed_bytes"] = len(stream)
        st["recon_sse"] = float(sum(st["stage_recon_sse"]))
        tried[key] = _Candidate(len(stream), st["recon_sse"], ratio, center, end, stream,
                                recon, st)

    base_end = getattr(predictor, "end_mode", 0)
    encode(ratios[0], centres[0], base_end)
    end = base_end
    if sweep and hasattr(predictor, "end_mode"):
        for e in range(END_MODE_MAX + 1):
            encode(ratios[0], centres[0], e)
        end = min(tried.values(), key=lambda c: (c.size, c.sse)).end
    for ratio in ratios:
        for center in centres:
            encode(ratio, center, end)

    cands = list(tried.values())
    tunable = getattr(predictor, "tunable", False)
    rd = tune == "rd" and eb_ratio is None and tunable and len(cands) > 1
    best = _pick(cands, tune_size_slack, rd)
    stats = best.stats
    stats["tune_candidates"] = len(cands)
    for k in ("predict_s", "quantize_s", "entropy_s"):
        stats["search_" + k] = sum(c.stats[k] for c in cands)
    if hasattr(predictor, "center"):
        predictor.center = best.center
    if hasattr(predictor, "end_mode"):
        predictor.end_mode = best.end
    stats.update(eb_ratio=best.ratio, interp_center=best.center, interp_end_mode=best.end,
                 tune=tune,
                 tune_size_slack=tune_size_slack if tune == "rd" and eb_ratio is None
                 and tunable else 1.0)
    if verbose:
        print(f"tuned: eb_ratio={best.ratio} center={best.center} end_mode={best.end} "
              f"compressed={best.size} bytes raw={stats['raw_payload_bytes']} bytes "
              f"sse={stats['recon_sse']:.6g}")
    stats["recon"] = _finalize(best.recon, header(best.center, best.ratio, best.end))
    stats["original_bytes"] = inp.nbytes
    stats["ratio"] = inp.nbytes / best.size
    return best.stream, stats


def stage_stats_device(res: "GpuEncodeResult", d_x) -> np.ndarray:
    """(n_stages, 5) float64 error moments of a device encode: pred_sae,
    pred_sse, recon_sae, recon_sse, recon_max (latec_interp_stats)."""
    plan = res.plan
    words = int(lib().latec_interp_stats_words(plan.handle))
    scratch = D.empty(words, np.float64)
    out = D.empty(plan.n_stages * 5, np.float64)
    check(lib().latec_interp_stats(plan.handle, ptr(d_x), ptr(res.d_recon), _optr(res.opts),
                                   ptr(out), ptr(scratch), words, ptr(None)))
    return D.to_host(out, np.float64).reshape(plan.n_stages, 5)


def _compress_region_gpu_impl(field, d_field, ebs, predictor, radius, round_output, stats,
                              zstd_level, stream_version=VERSION):
    C = field.shape[0]
    region = field.shape[1:]
    plan = get_plan(region, C, predictor.levels, predictor.anchor_stride,
                    predictor.anchor_block)
    o = _interp_opts(predictor, len(region), radius, round_output)
    t0 = time.time()
    res = encode_device(d_field, plan, ebs, o)
    fits = fit_device(res)
    mom = stage_stats_device(res, d_field)
    stats["quantize_s"] += time.time() - t0
    t0 = time.time()
    frame = None
    if stream_version == VERSION_MULTILANE:
        from . import mlrans

        payload, rec_len, n_out = mlrans.pack_payload(res, fits)
    else:
        payload, rec_len, n_out, frame = pack_device_framed(res, fits, zstd_level)
    stats["entropy_s"] += time.time() - t0
    recon = D.to_host_pinned(res.d_recon).reshape(C, *region)
    for s in range(plan.n_stages):
        stats["stage_codes"][s] += C * plan.counts[s]
        stats["stage_outliers"][s] += int(n_out[s])
        stats["stage_payload_bytes"][s] += int(rec_len[s])
        stats["stage_model_bits"][s] += float(fits[s].bits)
        if plan.counts[s]:
            stats["stage_pred_sae"][s] += float(mom[s, 0])
            stats["stage_pred_sse"][s] += float(mom[s, 1])
            stats["stage_recon_sae"][s] += float(mom[s, 2])
            stats["stage_recon_sse"][s] += float(mom[s, 3])
            stats["stage_recon_max"][s] = max(stats["stage_recon_max"][s], float(mom[s, 4]))
    stats["outliers"] += int(n_out.sum())
    return payload, recon, frame


def decode_device_v5(payload: bytes, plan, ebs, o):
    """decode_device for a version-5 payload: lanes decoded on the GPU."""
    from . import mlrans

    ws = workspace(plan, o.radius)
    mlrans.decode_payload(payload, plan, ws)
    ws.recon.zero_()
    ebs_arr = np.ascontiguousarray(ebs, np.float64)
    check(lib().latec_interp_decode(plan.handle, 0, plan.n_stages, ptr(ws.codes), ptr(ws.rank),
                                    ptr(ws.val), ptr(ws.cnt), ptr(ws.recon), ptr(ebs_arr),
                                    _optr(o), ptr(None)))
    return ws.recon


def decompress(stream: bytes, predictor_factory=None) -> np.ndarray:
    """codec.py:288-322."""
    header, payload = read_stream(stream)
    if predictor_factory is None:
        if header.flags & FLAG_INTERP:
            def predictor_factory(hdr):
                return InterpPredictor("cubic" if hdr.flags & FLAG_CUBIC else "linear",
                                       hdr.levels, hdr.anchor_stride, hdr.anchor_block)
        else:
            raise ValueError("stream needs a predictor_factory")
    predictor = predictor_factory(header)
    if hasattr(predictor, "center"):
        predictor.center = header.interp_center
    if hasattr(predictor, "end_mode"):
        predictor.end_mode = (header.flags >> END_MODE_SHIFT) & END_MODE_MAX
    spatial = tuple(header.spatial)
    ebs = stage_ebs(spatial, header.levels, header.anchor_stride, header.anchor_block,
                    header.eb, header.eb_ratio)
    if _fast_path(predictor) and not header.flags & FLAG_RANS:
        plan = get_plan(spatial, header.channels, header.levels, header.anchor_stride,
                        header.anchor_block)
        round_output = np.issubdtype(DTYPE_CODES[header.src_dtype], np.integer)
        o = _interp_opts(predictor, len(spatial), header.radius, round_output)
        if header.version == VERSION_MULTILANE:
            d_recon = decode_device_v5(payload, plan, ebs, o)
        else:
            d_recon = decode_device(payload, plan, ebs, o)
        canvas = D.to_host_pinned(d_recon).reshape(header.channels, *spatial)
        return _finalize(canvas, header)
    if header.version == VERSION_MULTILANE:
        raise ValueError("stream version 5 needs the package's InterpPredictor")
    make_masks = getattr(predictor, "stage_masks", stage_masks)
    masks = make_masks(spatial, header.levels, header.anchor_stride, header.anchor_block)
    canvas = _decompress_region(payload, masks, ebs, header, predictor)
    return _finalize(canvas, header)
