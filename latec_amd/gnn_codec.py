"""Checkpoint-backed GNN tensor codec on the GPU (the API of latec/gnn_codec.py).

``GNNCodec(checkpoint).compress(x, error_bound, ...) -> bytes`` and
``.uncompress(stream) -> torch.Tensor`` keep the reference's signatures,
defaults, validation messages and its LATECGNN v14 stream
(``"<8sII"`` prefix + sorted compact JSON meta + zstd payload,
gnn_codec.py:229-246).  The payload is the reference's record sequence: the
global anchor pass (per-point scale levels at eb0), then for every chunk in
raster order one per-point-level rANS stage record per local stage >= 1
(``_compress_chunked``, gnn_codec.py:1048-1336).

What differs, MI355X-first:

* the automatic chunk budget is ``2**25`` points, so a 64^4 tensor is one
  chunk whose embedding field (bf16) lives in HBM; explicit ``chunk_size``
  keeps the extended-block partition with halo frames;
* every per-stage step runs in liblatec kernels (geometry, message pass,
  pools, head, gate rate proxy, quantizer, scale levels);
* the meta gains ``"numerics"`` (``"bf16-mfma"`` or ``"fp32-mfma"``): the
  decoder refuses streams whose predictions it cannot replay (the
  reference's own torch numerics, or another mode).
"""

from __future__ import annotations

import base64
import ctypes
import json
import struct
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path
from typing import Any

import numpy as np

from . import _device as D
from . import rans as R
from ._lib import GnnDecodeStageIO, GnnStageIO, InterpOpts, check, lib, ptr, stream_ptr
from .bitstream import pack_stage, unpack_stage, zstd_compress, zstd_decompress
from .gnn_predictor import ChunkedGNNPredictor, load_model
from .levels import _slice_count, stage_ebs, stage_layouts, stage_strides
from .predictor import END_EXTRAP, END_QUAD, default_interp_center, get_plan
from .quantizer import dequantize, opts_for, quantize, quantize_device

_MAGIC = b"LATECGNN"
_VERSION = 14
_PREFIX = "<8sII"
_PREFIX_SIZE = struct.calcsize(_PREFIX)
_ANCHOR_BLOCK = 1
_AUTO_CHUNK_THRESHOLD = 1 << 21      # the reference's point budget (auto levels)
_HBM_CHUNK_THRESHOLD = 1 << 25       # one chunk per tensor up to this many points
_RANK_LEVELS = {2: 8, 4: 5, 5: 4}
_GNN_EB_COARSE_FACTOR = 0.9
_GNN_EB_COARSE_SWEEP = (1.0, 0.9, 0.65, 0.41)
_M_TILE = 1 << 16                     # recorded in meta for the reference's reader
_GATE_T = np.arange(2, 14, dtype=np.float64)
_GATE_SHIFTS = (0, 2, 4, 6)
_GATE_DIR_SHIFT = 10
_GATE_END_MODE = END_QUAD | END_EXTRAP
INTERP_KINDS = {"interp-cubic": 1, "interp-linear": 2, "interp-linear-last": 3}
NUMERICS_TAG = {"bf16": "bf16-mfma", "fp32": "fp32-mfma"}
_F32_U = float(np.finfo(np.float32).eps) / 2


def _auto_levels(shape, agg_level: int = 2) -> int:
    """gnn_codec.py:69-95: schedule depth from the rank, capped so the anchor
    stride stays <= half the longest axis."""
    ndim = len(shape)
    if ndim == 3:
        rank_levels = 7 if agg_level == 1 else 6
    elif ndim in _RANK_LEVELS:
        rank_levels = _RANK_LEVELS[ndim]
    else:
        rank_levels = (_AUTO_CHUNK_THRESHOLD.bit_length() - 1) // ndim
    max_axis = int(max(shape))
    size_cap = max(1, (max_axis.bit_length() - 1) - 1)
    return max(1, min(rank_levels, size_cap))


def _per_step_eb_ratio(coarse_factor: float, levels: int) -> float:
    """gnn_codec.py:111-122: coarse_factor ** (1 / (levels - 1))."""
    if levels <= 1:
        return 1.0
    return float(coarse_factor) ** (1.0 / (levels - 1))


def roundtrip_slack(vmin: float, vmax: float) -> float:
    """Float32 normalize/denormalize slack u*(3S + max|v|) (gnn_codec.py:193-226)."""
    span = float(vmax) - float(vmin)
    if span <= 0:
        return 0.0
    return _F32_U * (3.0 * span + max(abs(float(vmin)), abs(float(vmax))))


def _as_numpy(x: Any) -> np.ndarray:
    if isinstance(x, np.ndarray):
        return x
    if hasattr(x, "detach") and hasattr(x, "cpu"):
        return x.detach().cpu().numpy()
    if hasattr(x, "numpy"):
        return x.numpy()
    return np.asarray(x)


def _restore_dtype(values: np.ndarray, dtype: np.dtype) -> np.ndarray:
    if dtype.kind in "iu":
        info = np.iinfo(dtype)
        values = np.clip(np.rint(values), info.min, info.max)
    elif dtype.kind == "b":
        values = values >= 0.5
    return values.astype(dtype, copy=False)


def _write_stream(meta: dict, payload: bytes, zstd_level: int) -> bytes:
    header = json.dumps(meta, sort_keys=True, separators=(",", ":")).encode("utf-8")
    return struct.pack(_PREFIX, _MAGIC, _VERSION, len(header)) + header + zstd_compress(
        payload, zstd_level)


def _read_stream(stream: bytes):
    if len(stream) < _PREFIX_SIZE:
        raise ValueError("not a LATEC GNN stream")
    magic, version, header_len = struct.unpack_from(_PREFIX, stream, 0)
    if magic != _MAGIC:
        raise ValueError(f"not a LATEC GNN stream (bad magic {magic!r})")
    if version != _VERSION:
        raise ValueError(f"unsupported LATEC GNN stream version {version}")
    off = _PREFIX_SIZE
    meta = json.loads(stream[off:off + header_len].decode("utf-8"))
    return meta, zstd_decompress(stream[off + header_len:])


def _pack_gates(gates) -> str:
    raw = np.asarray(gates, dtype=np.uint16).tobytes()
    return base64.b64encode(zstd_compress(raw, 19)).decode("ascii")


def _unpack_gates(blob: str) -> list:
    return np.frombuffer(zstd_decompress(base64.b64decode(blob)), dtype=np.uint16).tolist()


def _chunk_stage_ebs(ndim, levels, stride, eb, ratio):
    """gnn_codec.py:262-268: stage_ebs on a tiny same-rank shape."""
    return stage_ebs((2 * stride,) * ndim, levels, stride, 1, eb, ratio)


def _anchor_axes(shape, stride: int, block: int = 1):
    return [np.arange(n)[(np.arange(n) % stride) < block] for n in shape]


def _anchor_flat_indices(axes, shape) -> np.ndarray:
    strides = np.cumprod((1,) + tuple(shape)[:0:-1])[::-1].astype(np.int64)
    idx = np.zeros(1, np.int64)
    for axis, st in zip(axes, strides):
        idx = (idx[:, None] + axis.astype(np.int64) * st).reshape(-1)
    return idx


def _auto_chunk_edges(shape, stride: int, target: int = _AUTO_CHUNK_THRESHOLD):
    """Largest near-isotropic chunk within ``target`` points, short axes whole,
    edges multiples of the stride (gnn_codec.py:286-314)."""
    fixed, remaining = 1, len(shape)
    full_axes: set = set()
    for axis in sorted(range(len(shape)), key=shape.__getitem__):
        free_edge = (target / fixed) ** (1.0 / remaining)
        if shape[axis] > free_edge:
            break
        full_axes.add(axis)
        fixed *= shape[axis]
        remaining -= 1
        if not remaining:
            break
    free_edge = (target / fixed) ** (1.0 / remaining) if remaining else 0.0
    return tuple(
        max(stride, -(-n // stride) * stride) if axis in full_axes
        else max(stride, int(free_edge) // stride * stride)
        for axis, n in enumerate(shape))


def _normalize_block(block: np.ndarray, norm) -> np.ndarray:
    vmin, span = norm
    out = block.astype(np.float32, order="C")
    out -= vmin
    out /= span
    return out


def _gate_scale_thresholds(eb: float):
    """63 float32 boundaries of the gate's 64-level scale snapping and the grid
    (gnn_codec.py:880-917)."""
    lo_f, hi_f = float(eb) / R.SCALE_LO_DIV, float(eb) * R.SCALE_HI_MULT
    log_lo = np.log(lo_f)
    log_step = (np.log(hi_f) - log_lo) / 63.0
    targets = np.arange(1, 64)
    lo_bits = np.zeros(63, dtype=np.uint32)
    hi_bits = np.full(63, np.asarray(np.float32(hi_f)).view(np.uint32), dtype=np.uint32)
    while np.any(lo_bits < hi_bits):
        mid = lo_bits + (hi_bits - lo_bits) // 2
        values = mid.view(np.float32).astype(np.float64)
        level = np.rint((np.log(np.clip(values, lo_f, hi_f)) - log_lo) / log_step)
        upper = level >= targets
        hi_bits = np.where(upper, mid, hi_bits).astype(np.uint32)
        lo_bits = np.where(upper, lo_bits, mid + 1).astype(np.uint32)
    grid = np.exp(log_lo + log_step * np.arange(64, dtype=np.float64))
    return np.ascontiguousarray(lo_bits.view(np.float32)), np.ascontiguousarray(grid)


def gate_costs_device(model, x, pred, scale, cands, n: int, eb: float, radius: int, tabs):
    """The gate's rate proxy on the device (gnn_codec.py:777-797, 920-971): a
    (1 + 4 * len(cands), 13) float64 array of bit sums, row 0 the model's
    prediction, then each interpolation candidate at scale shifts 0, 2, 4, 6,
    in 13 buckets of the scale against eb * 2^T.  ``tabs`` is
    _gate_scale_thresholds(eb)."""
    thr, grid = tabs
    W = np.zeros((1 + 4 * len(cands)) * 13, np.float64)
    ipa = (ctypes.c_void_p * len(cands))(*[ptr(c).value for c in cands])
    check(lib().latec_gnn_gate_costs(model.handle, ptr(x), ptr(pred), ptr(scale),
                                     ctypes.cast(ipa, ctypes.c_void_p), len(cands), int(n),
                                     float(eb), int(radius), ptr(thr), ptr(grid), ptr(W),
                                     stream_ptr()))
    return W.reshape(-1, 13)


def _gate_select(W: np.ndarray):
    """Branch-free choice of (kind, dir, T, shift) from the per-bucket rate
    sums W (1 + 12 candidates, 13 buckets), gnn_codec.py:920-971."""
    cums = np.cumsum(W, axis=-1)
    base, ci = cums[0], cums[1:]
    tot = base[-1]
    low = tot - base[:-1] + ci[:, :-1]
    high = base[:-1] + (ci[:, -1:] - ci[:, :-1])
    costs = np.concatenate([low, high])
    fi = int(np.argmin(costs.reshape(-1)))
    if not costs.reshape(-1)[fi] < tot:
        return 0, 0, 0, 0
    ncol, nsh = W.shape[1] - 1, len(_GATE_SHIFTS)
    n_side = (W.shape[0] - 1)
    row = fi // ncol
    side_row = row % n_side
    return (side_row // nsh + 1, int(row >= n_side), int(_GATE_T[fi % ncol]),
            _GATE_SHIFTS[side_row % nsh])


class _ChunkSchedule:
    """One chunk's coded layouts and the positions of its coded rows in the
    full local stage (the column split of ``_split_rows``, gnn_codec.py:704-725)."""

    def __init__(self, cshape, levels, stride, low_axes):
        self.cshape = cshape
        full = stage_layouts(cshape, levels, stride)
        lay = stage_layouts(cshape, levels, stride, low_axes) if low_axes else full
        self.full_counts = [l.size for l in full]
        self.counts = [l.size for l in lay]
        self.keep = [None] * len(full)
        if low_axes:
            for s, (L, F) in enumerate(zip(lay, full)):
                if L.size == F.size:
                    continue
                src = np.arange(F.size, dtype=np.int64)
                out = np.empty(L.size, np.int64)
                for cl, fcl in zip(L.classes, F.classes):
                    if cl.size == 0:
                        continue
                    v = np.lib.stride_tricks.as_strided(
                        src[fcl.offset:], fcl.sizes, [8 * x for x in fcl.strides])
                    for j, (b, fb) in enumerate(zip(cl.starts, fcl.starts)):
                        if b != fb or cl.sizes[j] != fcl.sizes[j]:
                            v = np.take(v, np.arange((b - fb) // cl.step,
                                                     (b - fb) // cl.step + cl.sizes[j]), axis=j)
                    dst = np.lib.stride_tricks.as_strided(
                        out[cl.offset:], cl.sizes, [8 * x for x in cl.strides], writeable=True)
                    dst[...] = v
                self.keep[s] = out
        self._dkeep = {}

    def device_keep(self, s):
        """keep[s] on the device (None when the whole stage is coded), uploaded
        once."""
        if self.keep[s] is None:
            return None
        d = self._dkeep.get(s)
        if d is None:
            d = self._dkeep[s] = D.to_device(self.keep[s])
        return d


_SCHEDULES: dict = {}


def _chunk_schedule(cshape, levels, stride, low_axes) -> _ChunkSchedule:
    """_ChunkSchedule, kept per (chunk shape, levels, stride, low axes): a
    chunked tensor repeats a few chunk shapes, and every call on a tensor of
    the same shape repeats them all."""
    key = (tuple(int(n) for n in cshape), int(levels), int(stride), tuple(low_axes))
    sched = _SCHEDULES.get(key)
    if sched is None:
        if len(_SCHEDULES) >= 256:
            _SCHEDULES.clear()
        sched = _SCHEDULES[key] = _ChunkSchedule(cshape, levels, stride, low_axes)
    return sched


class GNNCompressorCodec:
    """GNN-backed tensor codec (gnn_codec.py:1509-1919) on liblatec kernels.

    The tensor is min-max normalized to [0, 1] internally and ``error_bound``
    is relative to its range."""

    def __init__(self, checkpoint_path, device: str | None = None, *,
                 strict_checkpoint: bool = True, numerics: str = "bf16"):
        self.checkpoint_path = Path(checkpoint_path)
        if not self.checkpoint_path.exists():
            raise FileNotFoundError(f"GNN checkpoint not found: {self.checkpoint_path}")
        if numerics not in NUMERICS_TAG:
            raise ValueError(f"numerics must be one of {sorted(NUMERICS_TAG)}")
        self.device = device or "cuda"
        self.strict_checkpoint = bool(strict_checkpoint)
        self.numerics = numerics
        self.checkpoint_hash = load_model(self.checkpoint_path)[3]
        self._predictors: dict = {}
        self.last_stats: dict = {}
        self.last_decoded = None  # the last uncompress's normalized reconstruction (device)
        # count the model's GEMM flops and time its stage passes (HIP events)
        # into last_stats (model_flops, model_s; syncs)
        self.count_flops = False
        # each stage after the model in one library call (latec_gnn_code_stage;
        # the decoder: latec_gnn_decode_stage_begin / _end around the rANS
        # decode); False runs the same steps as separate calls from Python
        self.native_stages = True

    @staticmethod
    def _chunk_edges(shape, anchor_stride: int, cs):
        if cs is None:
            if int(np.prod(shape)) <= _HBM_CHUNK_THRESHOLD:
                return tuple(-(-n // anchor_stride) * anchor_stride for n in shape)
            return _auto_chunk_edges(shape, anchor_stride, _HBM_CHUNK_THRESHOLD)
        edges = (int(cs),) * len(shape) if np.isscalar(cs) else tuple(int(e) for e in cs)
        if len(edges) != len(shape):
            raise ValueError("chunk_size must be scalar or one entry per axis")
        for e in edges:
            if e < anchor_stride or e % anchor_stride:
                raise ValueError("chunk_size must be a positive multiple of anchor_stride")
        return edges

    def _predictor(self, levels: int, numerics: str) -> ChunkedGNNPredictor:
        """The predictor of (levels, numerics), kept for the last four keys
        (fields of different ranks and sizes take different levels; each
        predictor holds its model handles and the last tensor's frame)."""
        key = (int(levels), numerics)
        p = self._predictors.pop(key, None)
        if p is None:
            p = ChunkedGNNPredictor(self.checkpoint_path, 0.0, 1.0, levels=levels,
                                    anchor_stride=1 << levels, numerics=numerics)
            while len(self._predictors) >= 4:
                self._predictors.pop(next(iter(self._predictors)))
        self._predictors[key] = p  # most recent last
        return p

    # -- encode -----------------------------------------------------------------
    def compress(self, x: Any, error_bound: float = 1e-2, *, levels: int | str = "auto",
                 radius: int = 1 << 15, zstd_level: int = 9, eb_ratio: float | None = None,
                 tune: str = "fast", chunk_size=None, fp16: bool = True,
                 compile: bool | str = "auto", gate: bool = True, gate_fine: bool = False,
                 predictor: str = "gnn", numerics: str | None = None) -> bytes:
        """Compress any-rank numeric array/tensor; the reference's arguments.
        ``fp16``/``compile`` select torch float paths in the reference and are
        recorded but have no effect here; ``numerics`` (default: the codec's)
        picks the kernels' mode."""
        if tune not in ("fast", "size"):
            raise ValueError("tune must be 'fast' or 'size'")
        if isinstance(levels, str) and levels != "auto":
            raise ValueError("levels must be a positive int or 'auto'")
        if isinstance(compile, str) and compile != "auto":
            raise ValueError("compile must be a bool or 'auto'")
        if not isinstance(levels, str) and int(levels) < 1:
            raise ValueError("levels must be >= 1")
        numerics = numerics or self.numerics
        if numerics not in NUMERICS_TAG:
            raise ValueError(f"numerics must be one of {sorted(NUMERICS_TAG)}")
        radius = int(radius)
        gate = bool(gate)
        gate_fine = bool(gate_fine) and gate
        arr = np.asarray(_as_numpy(x))
        if arr.size == 0:
            raise ValueError("cannot compress an empty tensor")
        if arr.dtype.kind not in "biuf":
            raise TypeError(f"unsupported dtype {arr.dtype}; expected numeric data")
        dtype = np.dtype(arr.dtype)
        original_shape = tuple(int(n) for n in arr.shape)
        shape = original_shape if original_shape else (1,)
        values = arr.reshape(shape)
        vmin = float(np.float32(values.min()))
        vmax = float(np.float32(values.max()))
        if vmax <= vmin:
            vmax = vmin + 1.0
        norm = (vmin, vmax - vmin)
        round_output = (vmax - vmin, vmin) if dtype.kind in "bi" else False
        eb = float(error_bound)
        if eb <= 0:
            raise ValueError("error_bound must be > 0")
        if predictor != "gnn" and predictor not in INTERP_KINDS:
            raise ValueError(f"predictor must be 'gnn' or one of {sorted(INTERP_KINDS)}")
        interp_kind = INTERP_KINDS.get(predictor, 0)
        if interp_kind:
            gate = gate_fine = False
        coarse = ([float(eb_ratio)] if eb_ratio is not None else
                  (list(_GNN_EB_COARSE_SWEEP) if tune == "size" else [_GNN_EB_COARSE_FACTOR]))
        if isinstance(levels, str):
            levels = _auto_levels(shape, load_model(self.checkpoint_path)[1])
        levels = int(levels)
        stride = 1 << levels
        ratios = sorted({_per_step_eb_ratio(c, levels) for c in coarse})
        edges = self._chunk_edges(shape, stride, chunk_size)
        xn = D.to_device(np.ascontiguousarray(_normalize_block(values, norm)).reshape(-1))
        best = None
        for ratio in ratios:
            payload, gates, fine_gates = self._encode(
                values, xn, norm, round_output, eb, ratio, edges, levels, radius, gate,
                gate_fine, interp_kind, numerics, int(zstd_level))
            if gates is not None and not any(gates):
                gates = None
            if fine_gates is not None and not any(fine_gates):
                fine_gates = None
            meta = {
                "shape": list(original_shape), "dtype": dtype.str, "error_bound": eb,
                "levels": levels, "radius": radius, "vmin": vmin, "vmax": vmax,
                "eb_ratio": ratio, "checkpoint_hash": self.checkpoint_hash.hex(),
                "chunks": list(edges), "m_tile": _M_TILE, "fp16": bool(fp16),
                "compiled": False, "predictor": predictor,
                "numerics": NUMERICS_TAG[numerics],
            }
            if gates is not None:
                meta["gates"] = _pack_gates(gates)
            if fine_gates is not None:
                meta["fine_gates"] = _pack_gates(fine_gates)
            stream = _write_stream(meta, payload, int(zstd_level))
            if best is None or len(stream) < len(best):
                best = stream
        return best

    def _interp_candidates(self, recon, sls, single, plan, s, kinds, keep, n):
        """Gate candidates (gnn_codec.py:533-682) from the interpolation
        kernel on the chunk's block: {kind: device f32 (n,)}."""
        T = D.torch()
        ndim = len(plan.shape)
        block = recon if single else recon[(slice(None), *sls)].contiguous()
        out = {}
        M = plan.counts[s]
        for k in kinds:
            o = InterpOpts()
            o.cubic = 1 if k == 1 else 0
            o.center = default_interp_center(ndim) if k == 1 else (1 if k == 2 else 2)
            o.end_mode = _GATE_END_MODE
            o.radius = 1 << 15
            o.code_bytes = 2
            buf = T.empty(M, dtype=T.float32, device=D.device())
            check(lib().latec_interp_predict(plan.handle, s, ptr(block), ptr(buf), _optp(o),
                                             stream_ptr()))
            out[k] = buf if keep is None else buf[keep]
        return out

    @staticmethod
    def _code_stage(pred_obj, plan, s, n, dkeep, p_full, sc_full, xn, recon, block, o,
                    cand_opts, gate, interp_kind, reuse, choice, eb, radius, gate_tab, hist):
        """latec_gnn_code_stage: the stage after the model (gate, quantizer,
        reconstruction, levels) in one library call; (codes, outlier values,
        levels, outlier count, gate word)."""
        io = GnnStageIO()
        io.M, io.n = int(plan.counts[s]), int(n)
        io.d_keep = None if dkeep is None else dkeep.data_ptr()
        io.d_pred, io.d_scale = p_full[0].data_ptr(), sc_full[0].data_ptr()
        io.d_gidx = pred_obj.stage_gidx(s).data_ptr()
        io.d_x, io.d_recon = xn.data_ptr(), recon.data_ptr()
        io.d_block = None if block is None else block.data_ptr()
        io.opts = ctypes.addressof(o)
        io.cand_opts = ctypes.addressof(cand_opts)
        io.gate, io.interp_kind, io.reuse_choice = gate, int(interp_kind), int(reuse)
        word = np.zeros(1, np.uint32)
        io.choice, io.gate_word = choice.ctypes.data, word.ctypes.data
        io.eb, io.radius = eb, radius
        thr = grid = None
        if gate_tab is not None:
            thr, grid = gate_tab
            io.gate_thr, io.gate_grid = thr.ctypes.data, grid.ctypes.data
        lthr = np.ascontiguousarray(R.scale_thresholds(eb), np.float32)
        io.level_thr = lthr.ctypes.data
        io.d_hist = hist.data_ptr()
        codes = np.empty(n, np.uint16 if o.code_bytes == 2 else np.uint32)
        levels = np.empty(n, np.uint8)
        outl = np.empty(max(n, 1), np.float32)
        n_out = np.zeros(1, np.int64)
        io.codes, io.levels = codes.ctypes.data, levels.ctypes.data
        io.outliers, io.n_out = outl.ctypes.data, n_out.ctypes.data
        check(lib().latec_gnn_code_stage(pred_obj.model.handle, plan.handle, s,
                                         ctypes.byref(io), stream_ptr()))
        k = int(n_out[0])
        return codes, outl[:k].copy(), levels, k, int(word[0])

    @staticmethod
    def _decode_stage_begin(pred_obj, plan, s, n, dkeep, p_full, sc_full, block, cand_opts,
                            interp_kind, word, eb):
        """latec_gnn_decode_stage_begin: the stage's prediction with the gate
        applied; returns its scale levels (host)."""
        io = GnnDecodeStageIO()
        io.M, io.n = int(plan.counts[s]), int(n)
        io.d_keep = None if dkeep is None else dkeep.data_ptr()
        io.d_pred, io.d_scale = p_full[0].data_ptr(), sc_full[0].data_ptr()
        io.d_gidx = pred_obj.stage_gidx(s).data_ptr()
        io.d_block = None if block is None else block.data_ptr()
        io.cand_opts = ctypes.addressof(cand_opts)
        io.interp_kind, io.gate_word, io.eb = int(interp_kind), int(word), eb
        lthr = np.ascontiguousarray(R.scale_thresholds(eb), np.float32)
        io.level_thr = lthr.ctypes.data
        levels = np.empty(n, np.uint8)
        io.levels = levels.ctypes.data
        check(lib().latec_gnn_decode_stage_begin(pred_obj.model.handle, plan.handle, s,
                                                 ctypes.byref(io), stream_ptr()))
        return levels

    def _encode(self, values, xn, norm, round_output, eb, ratio, edges, levels, radius,
                gate, gate_fine, interp_kind, numerics, zstd_level):
        T = D.torch()
        shape = values.shape
        ndim = len(shape)
        stride = 1 << levels
        ebs = _chunk_stage_ebs(ndim, levels, stride, eb, ratio)
        pred_obj = self._predictor(levels, numerics)
        pred_obj.begin(shape, edges, channels=1)
        n_chunks = pred_obj.n_chunks
        # global anchor pass (host quantizer on device kernels)
        axes = _anchor_axes(shape, stride)
        avals = _normalize_block(values[np.ix_(*axes)], norm).reshape(-1)
        zeros = np.zeros(avals.size, np.float32)
        a_codes, a_out = quantize(avals, zeros, ebs[0], radius, round_output)
        a_recon = dequantize(zeros, a_codes, a_out, ebs[0], radius)
        tables = [R.build_laplace_tables(e, radius) for e in ebs]
        a_levels = R.scale_to_level(np.full(avals.size, ebs[0], np.float32), ebs[0])
        parts = [pack_stage(a_codes, a_out, rans_levels=a_levels, rans_tables=tables[0])]
        recon = T.zeros((1, *shape), dtype=T.float32, device=D.device())
        aidx = D.to_device(_anchor_flat_indices(axes, shape))
        recon.view(-1)[aidx] = D.to_device(a_recon)
        strides_s = stage_strides(ndim, levels, stride)
        is_finest = [st == 1 for st in strides_s]
        gate_tabs = {}
        gates, fine_gates = ([] if gate else None), ([] if gate and gate_fine else None)
        o = opts_for(radius, round_output)
        cand_opts = _candidate_opts(ndim)
        hist = D.zeros(lib().latec_hist_words(o.radius), np.uint32)
        empty = np.zeros(0, np.uint32), np.zeros(0, np.float32), np.zeros(0, np.uint8)
        pool = ThreadPoolExecutor(max_workers=8)
        futures = []
        stats = {"outliers": 0, "stage_codes": 0, "gate_words": 0}
        events = []  # (start, end) HIP events around the model's stage passes (count_flops)
        for ci in range(n_chunks):
            sls = pred_obj.chunk_slices(ci)
            cshape = tuple(sl.stop - sl.start for sl in sls)
            sched = _chunk_schedule(cshape, levels, stride, pred_obj.low_axes(ci))
            plan = get_plan(cshape, 1, levels, stride, 1)
            single = n_chunks == 1
            pred_obj.start_wave([ci], recon)
            fine_desc = None
            for s in range(1, len(sched.full_counts)):
                if sched.full_counts[s] == 0:
                    futures.append(pool.submit(pack_stage, *empty[:2], rans_levels=empty[2],
                                               rans_tables=tables[s]))
                    continue
                dkeep = sched.device_keep(s)
                if self.count_flops:
                    e0, e1 = T.cuda.Event(enable_timing=True), T.cuda.Event(enable_timing=True)
                    e0.record()
                p_full, sc_full = pred_obj.predict_wave_stage(s, recon, ebs[s])
                if self.count_flops:
                    e1.record()
                    events.append((e0, e1))
                    stats["model_flops"] = stats.get("model_flops", 0) + pred_obj.stage_flops(s)
                n = int(sched.counts[s])
                if n == 0:
                    futures.append(pool.submit(pack_stage, *empty[:2], rans_levels=empty[2],
                                               rans_tables=tables[s]))
                    continue
                if self.native_stages:
                    fine = bool(gate and gate_fine and is_finest[s])
                    reuse = fine and fine_desc is not None
                    if gate and not reuse and ebs[s] not in gate_tabs:
                        gate_tabs[ebs[s]] = _gate_scale_thresholds(ebs[s])
                    if fine and not reuse:
                        fine_desc = np.zeros(4, np.int32)
                    choice = fine_desc if fine else np.zeros(4, np.int32)
                    block = None
                    if gate or interp_kind:
                        block = recon if single else recon[(slice(None), *sls)].contiguous()
                    h_codes, h_val, h_lv, k, word = self._code_stage(
                        pred_obj, plan, s, n, dkeep, p_full, sc_full, xn, recon, block, o,
                        cand_opts, int(bool(gate)), interp_kind, reuse, choice, float(ebs[s]),
                        radius, gate_tabs.get(ebs[s]), hist)
                    if gate and not reuse:
                        (fine_gates if fine else gates).append(word)
                    stats["outliers"] += k
                    stats["stage_codes"] += n
                    futures.append(pool.submit(pack_stage, h_codes, h_val,
                                               rans_levels=h_lv, rans_tables=tables[s]))
                    continue
                gidx = pred_obj.stage_gidx(s)
                if dkeep is not None:
                    p, sc, gidx = p_full[0][dkeep], sc_full[0][dkeep], gidx[dkeep]
                else:
                    p, sc = p_full[0].clone(), sc_full[0].clone()
                x_st = xn[gidx]
                if interp_kind:
                    ip = self._interp_candidates(recon, sls, single, plan, s, (interp_kind,),
                                                 dkeep, n)[interp_kind]
                    p = ip
                    sc = T.full((n,), float(ebs[s]), dtype=T.float32, device=D.device())
                elif gate:
                    cands = self._interp_candidates(recon, sls, single, plan, s, (1, 2, 3),
                                                    dkeep, n)
                    if gate_fine and is_finest[s] and fine_desc is not None:
                        gk, gd, gt, gs = fine_desc
                    else:
                        if ebs[s] not in gate_tabs:
                            gate_tabs[ebs[s]] = _gate_scale_thresholds(ebs[s])
                        W = gate_costs_device(pred_obj.model, x_st, p, sc,
                                              (cands[1], cands[2], cands[3]), n,
                                              float(ebs[s]), radius, gate_tabs[ebs[s]])
                        gk, gd, gt, gs = _gate_select(W)
                        word = (gd << _GATE_DIR_SHIFT) + gk * 256 + gt * 16 + gs
                        if gate_fine and is_finest[s]:
                            fine_desc = (gk, gd, gt, gs)
                            fine_gates.append(word)
                        else:
                            gates.append(word)
                    if gt > 0:
                        check(lib().latec_gnn_gate_apply(ptr(p), ptr(sc), ptr(cands[gk]), n,
                                                         float(ebs[s]) * 2.0 ** gt,
                                                         float(2.0 ** -gs), gd, stream_ptr()))
                codes, rank, val, k, _ = quantize_device(x_st, p, n, ebs[s], o)
                rec = T.empty(n, dtype=T.float32, device=D.device())
                d_cnt = D.to_device(np.array([k], np.uint32))
                check(lib().latec_dequantize(ptr(p), ptr(codes), n, ptr(rank), ptr(val),
                                             ptr(d_cnt), float(ebs[s]), _optp(o), ptr(rec),
                                             stream_ptr()))
                recon.view(-1)[gidx] = rec
                lv = T.empty(n, dtype=T.uint8, device=D.device())
                R.scale_to_level_device(sc, n, float(ebs[s]), lv)
                h_codes = D.to_host(codes, np.uint16 if o.code_bytes == 2 else np.uint32)
                h_val = D.to_host(val[:k]) if k else np.zeros(0, np.float32)
                h_lv = D.to_host(lv, np.uint8)
                stats["outliers"] += k
                stats["stage_codes"] += n
                futures.append(pool.submit(pack_stage, h_codes, h_val,
                                           rans_levels=h_lv, rans_tables=tables[s]))
            pred_obj.finish_wave(recon)
        parts.extend(f.result() for f in futures)
        pool.shutdown()
        T.cuda.current_stream().synchronize()
        if self.count_flops:
            stats["model_s"] = sum(a.elapsed_time(b) for a, b in events) * 1e-3
        stats["gate_words"] = len(gates or []) + len(fine_gates or [])
        stats["gates_fired"] = sum(1 for g in (gates or []) + (fine_gates or []) if g)
        # the normalized reconstruction the decoder must reproduce bit for bit
        stats["recon"] = recon
        self.last_stats = stats
        return b"".join(parts), gates, fine_gates

    # -- decode -----------------------------------------------------------------
    def uncompress(self, stream):
        """Decompress bytes from ``compress``; returns a torch tensor of the
        original shape and dtype."""
        T = D.torch()
        meta, payload = _read_stream(bytes(stream))
        if self.strict_checkpoint and meta["checkpoint_hash"] != self.checkpoint_hash.hex():
            raise ValueError("checkpoint hash differs from the stream metadata")
        tag = meta.get("numerics")
        inv = {v: k for k, v in NUMERICS_TAG.items()}
        if tag not in inv:
            raise ValueError(
                "stream predictions come from another numerics mode "
                f"({tag!r}); this decoder replays {sorted(inv)}")
        numerics = inv[tag]
        original_shape = tuple(int(n) for n in meta["shape"])
        shape = original_shape or (1,)
        dtype = np.dtype(meta["dtype"])
        vmin, vmax = float(meta["vmin"]), float(meta["vmax"])
        if vmax <= vmin:
            vmax = vmin + 1.0
        edges = tuple(int(e) for e in meta["chunks"])
        levels = int(meta["levels"])
        stride = 1 << levels
        radius = int(meta["radius"])
        ndim = len(shape)
        ebs = _chunk_stage_ebs(ndim, levels, stride, float(meta["error_bound"]),
                               float(meta["eb_ratio"]))
        gates = _unpack_gates(meta["gates"]) if "gates" in meta else None
        fine_gates = _unpack_gates(meta["fine_gates"]) if "fine_gates" in meta else None
        interp_kind = INTERP_KINDS.get(meta.get("predictor", "gnn"), 0)
        pred_obj = self._predictor(levels, numerics)
        pred_obj.begin(shape, edges, channels=1)
        axes = _anchor_axes(shape, stride)
        n_anchor = int(np.prod([len(a) for a in axes]))
        tables = [R.build_laplace_tables(e, radius) for e in ebs]
        a_levels = R.scale_to_level(np.full(n_anchor, ebs[0], np.float32), ebs[0])
        a_codes, a_out, off = unpack_stage(payload, 0, rans_levels=a_levels, rans_tables=tables[0],
                                           radius=radius)
        a_recon = dequantize(np.zeros(n_anchor, np.float32), a_codes, a_out, ebs[0], radius)
        recon = T.zeros((1, *shape), dtype=T.float32, device=D.device())
        recon.view(-1)[D.to_device(_anchor_flat_indices(axes, shape))] = D.to_device(a_recon)
        is_finest = [st == 1 for st in stage_strides(ndim, levels, stride)]
        o = opts_for(radius)
        cand_opts = _candidate_opts(ndim)
        gi = fgi = 0
        for ci in range(pred_obj.n_chunks):
            sls = pred_obj.chunk_slices(ci)
            cshape = tuple(sl.stop - sl.start for sl in sls)
            sched = _chunk_schedule(cshape, levels, stride, pred_obj.low_axes(ci))
            plan = get_plan(cshape, 1, levels, stride, 1)
            single = pred_obj.n_chunks == 1
            pred_obj.start_wave([ci], recon)
            fine_desc = None
            for s in range(1, len(sched.full_counts)):
                if sched.full_counts[s] == 0:
                    _, _, off = unpack_stage(payload, off, rans_levels=np.zeros(0, np.uint8),
                                             rans_tables=tables[s], radius=radius)
                    continue
                dkeep = sched.device_keep(s)
                p_full, sc_full = pred_obj.predict_wave_stage(s, recon, ebs[s])
                n = int(sched.counts[s])
                if n == 0:
                    _, _, off = unpack_stage(payload, off, rans_levels=np.zeros(0, np.uint8),
                                             rans_tables=tables[s], radius=radius)
                    continue
                word = None
                if not interp_kind:
                    if fine_gates is not None and is_finest[s]:
                        if fine_desc is None:
                            fine_desc = int(fine_gates[fgi])
                            fgi += 1
                        word = fine_desc
                    elif gates is not None:
                        word = int(gates[gi])
                        gi += 1
                if self.native_stages:
                    gated = bool(interp_kind) or (word is not None and (word >> 4) & 15 > 0)
                    block = None
                    if gated:
                        block = recon if single else recon[(slice(None), *sls)].contiguous()
                    h_lv = self._decode_stage_begin(pred_obj, plan, s, n, dkeep, p_full, sc_full,
                                                    block, cand_opts, interp_kind, word or 0,
                                                    float(ebs[s]))
                    cdt = np.uint16 if o.code_bytes == 2 else np.uint32
                    codes, outl, off = unpack_stage(payload, off, rans_levels=h_lv,
                                                    rans_tables=tables[s], radius=radius,
                                                    code_dtype=cdt)
                    if len(codes) != n:
                        raise ValueError("stage code count mismatch")
                    hc = np.ascontiguousarray(codes, cdt)
                    ov = np.ascontiguousarray(outl, np.float32)
                    check(lib().latec_gnn_decode_stage_end(
                        pred_obj.model.handle, ptr(hc), ptr(ov) if len(ov) else None, len(ov),
                        float(ebs[s]), _optp(o), ptr(recon), stream_ptr()))
                    continue
                gidx = pred_obj.stage_gidx(s)
                if dkeep is not None:
                    p, sc, gidx = p_full[0][dkeep], sc_full[0][dkeep], gidx[dkeep]
                else:
                    p, sc = p_full[0].clone(), sc_full[0].clone()
                if interp_kind:
                    p = self._interp_candidates(recon, sls, single, plan, s, (interp_kind,),
                                                dkeep, n)[interp_kind]
                    sc = T.full((n,), float(ebs[s]), dtype=T.float32, device=D.device())
                else:
                    if word is not None:
                        gk, gt, gs, gd = (word >> 8) & 3, (word >> 4) & 15, word & 15, \
                            (word >> _GATE_DIR_SHIFT) & 1
                        if gt > 0:
                            ip = self._interp_candidates(recon, sls, single, plan, s, (gk,),
                                                         dkeep, n)[gk]
                            check(lib().latec_gnn_gate_apply(ptr(p), ptr(sc), ptr(ip), n,
                                                             float(ebs[s]) * 2.0 ** gt,
                                                             float(2.0 ** -gs), gd, stream_ptr()))
                lv = T.empty(n, dtype=T.uint8, device=D.device())
                R.scale_to_level_device(sc, n, float(ebs[s]), lv)
                h_lv = D.to_host(lv, np.uint8)
                codes, outl, off = unpack_stage(payload, off, rans_levels=h_lv,
                                                rans_tables=tables[s], radius=radius)
                if len(codes) != n:
                    raise ValueError("stage code count mismatch")
                zeros = np.flatnonzero(codes == 0).astype(np.uint32)
                d_codes, d_rank, d_val, d_cnt = _stage_upload(codes, zeros, outl, o.code_bytes)
                rec = T.empty(n, dtype=T.float32, device=D.device())
                check(lib().latec_dequantize(ptr(p), ptr(d_codes), n, ptr(d_rank), ptr(d_val),
                                             ptr(d_cnt), float(ebs[s]), _optp(o), ptr(rec),
                                             stream_ptr()))
                recon.view(-1)[gidx] = rec
            pred_obj.finish_wave(recon)
        if off != len(payload):
            raise ValueError("trailing bytes in LATEC GNN payload")
        if gates is not None and gi != len(gates):
            raise ValueError("gate list length does not match the stream")
        self.last_decoded = recon  # normalized, as the encoder's stats["recon"]
        values = D.to_host_pinned(recon[0])  # a new array: one copy at the pinned rate
        values *= vmax - vmin
        values += vmin
        out = _restore_dtype(values.reshape(original_shape), dtype)
        return T.as_tensor(out)

    decompress = uncompress


def _optp(o):
    return ctypes.cast(ctypes.pointer(o), ctypes.c_void_p)


def _candidate_opts(ndim: int):
    """Options of the gate's three interpolation candidates (kinds 1..3), as
    _interp_candidates sets them."""
    arr = (InterpOpts * 3)()
    for k in (1, 2, 3):
        o = arr[k - 1]
        o.cubic = 1 if k == 1 else 0
        o.center = default_interp_center(ndim) if k == 1 else (1 if k == 2 else 2)
        o.end_mode = _GATE_END_MODE
        o.radius = 1 << 15
        o.code_bytes = 2
    return arr


def _stage_upload(codes, zeros, outl, code_bytes):
    """A decoded stage's codes, outlier ranks, outlier values and outlier
    count as device views of one host-to-device copy from the thread's
    pinned staging buffer (the caller's next device-to-host read orders the
    buffer's reuse after the copy)."""
    T = D.torch()
    al = lambda b: (b + 255) & ~255  # noqa: E731
    n, k = len(codes), len(zeros)
    o1 = al(n * code_bytes)
    o2 = o1 + al(4 * max(k, 1))
    o3 = o2 + al(4 * max(k, 1))
    tot = o3 + 16
    hv, ha = D.staging(tot, np.uint8)
    ha[:n * code_bytes] = codes.astype(np.uint16 if code_bytes == 2 else np.uint32).view(np.uint8)
    ha[o1:o1 + 4 * k] = zeros.view(np.uint8)
    ha[o2:o2 + 4 * k] = np.ascontiguousarray(outl, np.float32).view(np.uint8)
    ha[o3:o3 + 4] = np.array([k], np.uint32).view(np.uint8)
    dbuf = T.empty(tot, dtype=T.uint8, device=D.device())
    dbuf.copy_(hv, non_blocking=True)
    return (dbuf[:n * code_bytes], dbuf[o1:o1 + 4 * max(k, 1)], dbuf[o2:o2 + 4 * max(k, 1)],
            dbuf[o3:o3 + 4])


GNNCodec = GNNCompressorCodec
