"""The chunked GNN predictor on the GPU (the API of latec/gnn_predictor.py).

``ChunkedGNNPredictor`` keeps the reference's per-tensor wave protocol
(``begin`` / ``start_wave`` / ``predict_wave_stage`` / ``finish_wave``,
gnn_predictor.py:1369-1789) and its chunk partition (extended blocks that own
their internal high faces, ``chunk_slices``/``low_axes`` :1595-1614).  Every
per-stage step runs in liblatec: the neighbour geometry (``latec_gnn_geometry``),
the message pass and line pool (``latec_gnn_embed``), the axis pool and head
(``latec_gnn_head``) and the finalize of the revealed points
(``latec_gnn_finalize``).  The embedding field of a chunk is one dense array
over the chunk grown by a halo of ``anchor_stride`` cells (none when the
chunk is the whole tensor); halo cells hold value embeddings of decoded
neighbours (``value_halo_embed`` :1357-1366).

Numerics: ``numerics="bf16"`` (BF16 MFMA message/mix GEMMs, FP32 accumulate,
FP32 pools and head) or ``"fp32"``.  Encoder and decoder run the same kernels
on the same inputs, so their predictions agree bit for bit.
"""

from __future__ import annotations

import ctypes
import itertools
import os

import numpy as np

from . import _device as D
from . import gnn_model as GM
from ._lib import GnnFrame, check, lib, ptr, stream_ptr

NUMERICS = {"fp32": 0, "bf16": 1}

_MODEL_CACHE: dict = {}


def half_directions(ndim: int, agg_level: int | None = None):
    """One representative per line, first non-zero +1, at most ``agg_level``
    non-zeros (gnn_predictor.py:117-136)."""
    out = []
    for d in itertools.product((-1, 0, 1), repeat=ndim):
        nz = [x for x in d if x != 0]
        if nz and nz[0] > 0 and (agg_level is None or len(nz) <= agg_level):
            out.append(d)
    return out


class GnnModel:
    """A loaded model on the device for one rank and numerics mode."""

    def __init__(self, d: int, agg: int, params: dict, ndim: int, numerics: str = "bf16"):
        if numerics not in NUMERICS:
            raise ValueError(f"numerics must be one of {sorted(NUMERICS)}")
        self.d, self.agg, self.ndim, self.numerics = int(d), int(agg), int(ndim), numerics
        flat = GM.pack(params, d, agg)
        h = ctypes.c_void_p()
        check(lib().latec_gnn_create(self.d, self.agg, self.ndim, NUMERICS[numerics],
                                     ptr(flat), flat.size, ctypes.byref(h)))
        self.handle = h
        self.L = lib().latec_gnn_n_lines(h)

    @property
    def dtype(self):
        T = D.torch()
        return T.bfloat16 if self.numerics == "bf16" else T.float32

    def close(self):
        if self.handle:
            lib().latec_gnn_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_model(checkpoint_path):
    """(d, agg_level, params, checkpoint_hash) cached per file revision."""
    st = os.stat(checkpoint_path)
    key = (os.path.realpath(checkpoint_path), st.st_mtime_ns, st.st_size)
    hit = _MODEL_CACHE.get(key)
    if hit is None:
        for k in [k for k in _MODEL_CACHE if k[0] == key[0]]:
            del _MODEL_CACHE[k]
        d, agg, params = GM.load_checkpoint(checkpoint_path)
        hit = (d, agg, params, GM.file_sha256(checkpoint_path)[:16])
        _MODEL_CACHE[key] = hit
    return hit


class StageGeom:
    """Device geometry of one stage: frame index, tensor index and steps."""

    def __init__(self, model: GnnModel, plan, stage: int, frame: GnnFrame):
        T = D.torch()
        self.M = int(plan.counts[stage])
        n = max(self.M, 1)
        dev = D.device()
        self.qidx = T.empty(n, dtype=T.int64, device=dev)
        self.gidx = T.empty(n, dtype=T.int64, device=dev)
        self.steps = T.empty(n * model.L, dtype=T.int32, device=dev)
        if self.M:
            check(lib().latec_gnn_geometry(model.handle, plan.handle, stage,
                                           ctypes.byref(frame), ptr(self.qidx), ptr(self.gidx),
                                           ptr(self.steps), stream_ptr()))


class ChunkedGNNPredictor:
    """Chunk-by-chunk GNN predictor with the reference's wave protocol.

    Per tensor (encoder and decoder alike)::

        begin(shape, chunk_edges, channels)
        for each chunk (raster order):
            start_wave([ci], recon)
            for each non-empty local stage s >= 1, in order:
                pred, scale = predict_wave_stage(s, recon, eb)
                ... caller quantizes and writes recon ...
            finish_wave(recon)

    ``recon`` is a device float32 tensor (C, *S) holding the reconstruction
    so far; ``pred``/``scale`` are device float32 (1, M) in the order of the
    chunk's local stage (flat C-order of the stage points).
    """

    provides_scale = True

    def __init__(self, checkpoint_path, vmin: float, vmax: float, device: str = "cuda",
                 levels: int = 4, anchor_stride: int = 16, anchor_block: int = 1,
                 numerics: str = "bf16"):
        if int(anchor_block) != 1:
            raise ValueError("the chunked GNN schedule uses anchor_block 1")
        if int(anchor_stride) != 1 << int(levels):
            raise ValueError("anchor_stride must be 2**levels")
        self.vmin, self.vmax = float(vmin), float(vmax)
        span = self.vmax - self.vmin
        self.span = span if span > 0 else 1.0
        self.levels, self.anchor_stride, self.anchor_block = int(levels), int(anchor_stride), 1
        self.numerics = numerics
        self.d, self.agg_level, self.params, self.checkpoint_hash = load_model(checkpoint_path)
        self._models: dict = {}
        self._geom_cache: dict = {}

    # -- per-tensor lifecycle ---------------------------------------------------
    def model_for(self, ndim: int) -> GnnModel:
        m = self._models.get(ndim)
        if m is None:
            m = GnnModel(self.d, self.agg_level, self.params, ndim, self.numerics)
            self._models[ndim] = m
        return m

    def begin(self, shape, chunk_edges, channels: int = 1):
        self.shape = tuple(int(n) for n in shape)
        self.edges = tuple(int(e) for e in chunk_edges)
        if len(self.edges) != len(self.shape):
            raise ValueError("chunk_edges must have one entry per axis")
        for e in self.edges:
            if e < self.anchor_stride or e % self.anchor_stride:
                raise ValueError("chunk edges must be positive multiples of anchor_stride")
        if int(channels) != 1:
            raise ValueError("the GNN codec codes one channel")
        self.grid = tuple(-(-n // e) for n, e in zip(self.shape, self.edges))
        self.n_chunks = int(np.prod(self.grid))
        self.C = 1
        self.coded = np.zeros(self.n_chunks, bool)
        self.model = self.model_for(len(self.shape))
        T = D.torch()
        self._coded_dev = T.zeros(self.n_chunks, dtype=T.uint8, device=D.device())
        self._wave = None

    def chunk_slices(self, ci: int):
        """Extended-block partition (gnn_predictor.py:1595-1607): a chunk grows
        by one cell on each internal high face and owns that plane."""
        cidx = np.unravel_index(ci, self.grid)
        sls = []
        for i, e, n in zip(cidx, self.edges, self.shape):
            hi = min((i + 1) * e, n)
            if (i + 1) * e < n:
                hi = (i + 1) * e + 1
            sls.append(slice(i * e, hi))
        return tuple(sls)

    def low_axes(self, ci: int) -> tuple:
        """Axes whose local coordinate-0 plane was decoded by the low neighbour
        (gnn_predictor.py:1609-1614)."""
        cidx = np.unravel_index(ci, self.grid)
        return tuple(a for a, i in enumerate(cidx) if i > 0)

    def _frame(self, origin, cshape, pad: int) -> GnnFrame:
        fr = GnnFrame()
        fr.pad = pad
        fr.has_halo = 1 if pad else 0
        for j, n in enumerate(self.shape):
            fr.origin[j] = int(origin[j])
            fr.tshape[j] = n
            fr.edges[j] = self.edges[j]
            fr.grid[j] = self.grid[j]
        fr.d_coded = ptr(self._coded_dev).value if pad else None
        return fr

    def start_wave(self, chunk_ids, recon):
        from .predictor import get_plan

        if len(chunk_ids) != 1:
            raise ValueError("waves hold one chunk")
        T = D.torch()
        ci = int(chunk_ids[0])
        sls = self.chunk_slices(ci)
        origin = tuple(sl.start for sl in sls)
        cshape = tuple(sl.stop - sl.start for sl in sls)
        pad = self.anchor_stride if self.n_chunks > 1 else 0
        plan = get_plan(cshape, 1, self.levels, self.anchor_stride, 1)
        fshape = tuple(n + 2 * pad for n in cshape)
        m = self.model
        n_rows = int(np.prod(fshape))
        E = T.zeros(n_rows * m.ndim * m.d, dtype=m.dtype, device=D.device())
        frame = self._frame(origin, cshape, pad)
        key = (cshape, pad)
        geoms = self._geom_cache.get(key) if not pad else None
        if geoms is None:
            geoms = [StageGeom(m, plan, s, frame) for s in range(plan.n_stages)]
            if not pad:
                self._geom_cache = {key: geoms}
        if pad:
            check(lib().latec_gnn_halo_frame(m.handle, plan.handle, ctypes.byref(frame), ptr(E),
                                             ptr(recon), 0.0 + self.vmin, self.vmax, stream_ptr()))
        self._wave = {
            "ci": ci, "plan": plan, "E": E, "geoms": geoms, "fshape": np.asarray(fshape, np.int64),
            "chain": [0] + [s for s in range(1, plan.n_stages) if plan.counts[s]],
            "pos": 0, "ctx": None,
        }
        return self._wave

    def _embed(self, geom: StageGeom):
        T = D.torch()
        m, w = self.model, self._wave
        ctx = T.empty(max(geom.M, 1) * m.ndim * m.d, dtype=m.dtype, device=D.device())
        if geom.M:
            check(lib().latec_gnn_embed(m.handle, ptr(w["E"]), ptr(w["fshape"]),
                                        ptr(geom.steps), ptr(geom.qidx), geom.M, ptr(ctx),
                                        stream_ptr()))
        return ctx

    def stage_flops(self, s: int) -> int:
        """Multiply-add flops of the model's GEMMs for stage s of the current
        chunk: per live (line, point) pair and axis the message MLP (two-sided
        2*(2d*h) + 2*h*d, one-sided 2*d*h + 2*h*d) and the attention key
        (2*d*d); per point and axis of the previous stage the mix MLP
        (2*2d*h + 2*h*d)."""
        T = D.torch()
        m, w = self.model, self._wave
        g = w["geoms"][s]
        if not g.M:
            return 0
        st = g.steps[: g.M * m.L].to(T.int64)
        sp, sn = st & 0xFFFF, st >> 16
        nb = int(((sp > 0) & (sn > 0)).sum())
        nd = int(((sp > 0) ^ (sn > 0)).sum())
        d, h, K = m.d, 2 * m.d, m.ndim
        j = w["chain"].index(s)
        prev = w["geoms"][w["chain"][j - 1]].M if j > 0 else 0
        return (K * (nb * (2 * 2 * d * h + 2 * h * d) + nd * (2 * d * h + 2 * h * d)
                     + (nb + nd) * 2 * d * d) + prev * K * (2 * 2 * d * h + 2 * h * d))

    def stage_gidx(self, s: int):
        """Tensor flat indices of the current chunk's stage-s points (device)."""
        return self._wave["geoms"][s].gidx[: self._wave["geoms"][s].M]

    def predict_wave_stage(self, s: int, recon, eb: float, want_head=False):
        T = D.torch()
        w, m = self._wave, self.model
        j = w["pos"] + 1
        if j >= len(w["chain"]) or w["chain"][j] != s:
            raise ValueError(f"stage {s} out of order for this wave")
        prev = w["chain"][j - 1]
        gp, gh = w["geoms"][prev], w["geoms"][s]
        if gp.M:
            ctx_p = w["ctx"] if w["ctx"] is not None else self._embed(gp)
            check(lib().latec_gnn_finalize(m.handle, ptr(w["E"]), ptr(ctx_p), ptr(gp.qidx),
                                           ptr(gp.gidx), ptr(recon), gp.M, self.vmin, self.vmax,
                                           stream_ptr()))
        ctx = self._embed(gh)
        pred = T.empty(gh.M, dtype=T.float32, device=D.device())
        scale = T.empty(gh.M, dtype=T.float32, device=D.device())
        mu = lb = None
        if want_head:
            mu = T.empty(gh.M, dtype=T.float32, device=D.device())
            lb = T.empty(gh.M, dtype=T.float32, device=D.device())
        check(lib().latec_gnn_head(m.handle, ptr(ctx), gh.M, float(eb) / self.span,
                                   self.vmin, self.vmax, ptr(pred), ptr(scale), ptr(mu), ptr(lb),
                                   stream_ptr()))
        w["ctx"] = ctx
        w["pos"] = j
        if want_head:
            return pred[None], scale[None], mu, lb
        return pred[None], scale[None]

    def finish_wave(self, recon):
        w = self._wave
        if w["pos"] != len(w["chain"]) - 1:
            raise ValueError("finish_wave before all non-empty stages predicted")
        self.coded[w["ci"]] = True
        self._coded_dev[w["ci"]] = 1
        self._wave = None

    def clear_runtime_cache(self):
        self._geom_cache.clear()
