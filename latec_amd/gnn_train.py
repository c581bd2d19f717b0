"""GNN training on the GPU with hand-written forward/backward kernels.

The trainer follows the reference's ``run_stages`` (scripts/train_gnn.py:373-459):
over the codec's own progressive schedule, at stage i the points revealed at
stage i-1 are finalized into the embedding field, then stage i is predicted
and scored by the discretized-Laplace code length in bits
(``discretized_laplace_nll``, :195-230).  Teacher forcing reveals the truth
plus uniform noise in [-eb, eb]; ``closed_loop=True`` reveals the model's own
linear-quantized reconstruction instead (``qz``, :79-85).

Every model op is a ``torch.autograd.Function`` whose forward and backward
run liblatec kernels (``latec_train_*``): row gathers / scatter-adds, the
finalize row write, split-K FP32 GEMMs for every Linear and its two
gradients, column sums for biases, exact GELU, the segmented attention pool
and the Laplace head loss.  torch supplies the parameter tensors, index
bookkeeping, additions of partial pre-activations and the AdamW update
(FP32 master weights).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _device as D
from . import gnn_model as GM
from ._lib import check, lib, ptr

_WORK = {"t": None}


def _torch():
    import torch

    return torch


def _work(n_floats: int):
    T = _torch()
    w = _WORK["t"]
    if w is None or w.numel() < n_floats:
        w = T.empty(max(n_floats, 1 << 20), dtype=T.float32, device=D.device())
        _WORK["t"] = w
    return w


# GEMM arithmetic of every Linear, forward and both gradients: "fp32" or
# "bf16" (operands rounded to bf16, FP32 accumulation on MFMA tiles; the FP32
# master weights stay in the optimizer).  Set with set_precision or by
# GnnTrainer(precision=...).
_PRECISION = {"gemm": "fp32"}


def set_precision(mode: str) -> None:
    if mode not in ("fp32", "bf16"):
        raise ValueError("precision must be 'fp32' or 'bf16'")
    _PRECISION["gemm"] = mode


def _gemm(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias=None, accumulate=False, prec=None):
    need = max(1, M * N) * 256 if M * N <= (1 << 20) else (1 << 20)
    w = _work(need)
    prec = prec or _PRECISION["gemm"]
    fn = lib().latec_train_gemm_bf16 if prec == "bf16" else lib().latec_train_gemm
    check(fn(M, N, K, ptr(A), sam, sak, ptr(B), sbk, sbn, ptr(C), ldc, ptr(bias),
             int(accumulate), ptr(w), w.numel(), ptr(None)))


def _colsum(X, out, accumulate=False):
    M, N = X.shape
    w = _work(1024 * N + 1024)
    check(lib().latec_train_colsum(M, N, ptr(X), ptr(out), int(accumulate), ptr(w), w.numel(),
                                   ptr(None)))


# ---- autograd functions over the kernels ------------------------------------

def _functions():
    T = _torch()

    class Linear(T.autograd.Function):
        """y = x W^T + b; W may be a column slice of a parameter."""

        @staticmethod
        def forward(ctx, x, W, b):
            x = x.contiguous()
            n, k = x.shape
            out = W.shape[0]
            y = T.empty(n, out, dtype=T.float32, device=x.device)
            ctx.prec = _PRECISION["gemm"]  # the backward's GEMMs use the forward's mode
            _gemm(n, out, k, x, k, 1, W, W.stride(1), W.stride(0), y, out, b, prec=ctx.prec)
            ctx.save_for_backward(x, W)
            ctx.has_b = b is not None
            return y

        @staticmethod
        def backward(ctx, dy):
            x, W = ctx.saved_tensors
            dy = dy.contiguous()
            n, k = x.shape
            out = W.shape[0]
            dx = dW = db = None
            if ctx.needs_input_grad[0]:
                dx = T.empty(n, k, dtype=T.float32, device=dy.device)
                _gemm(n, k, out, dy, out, 1, W, W.stride(0), W.stride(1), dx, k, prec=ctx.prec)
            if ctx.needs_input_grad[1]:
                dW = T.empty(out, k, dtype=T.float32, device=dy.device)
                _gemm(out, k, n, dy, 1, out, x, k, 1, dW, k, prec=ctx.prec)
            if ctx.has_b and ctx.needs_input_grad[2]:
                db = T.empty(out, dtype=T.float32, device=dy.device)
                _colsum(dy, db)
            return dx, dW, db

    class Gelu(T.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            x = x.contiguous()
            y = T.empty_like(x)
            check(lib().latec_train_gelu(ptr(x), ptr(y), x.numel(), ptr(None)))
            ctx.save_for_backward(x)
            return y

        @staticmethod
        def backward(ctx, dy):
            (x,) = ctx.saved_tensors
            dy = dy.contiguous()
            dx = T.empty_like(x)
            check(lib().latec_train_gelu_bwd(ptr(x), ptr(dy), ptr(dx), x.numel(), ptr(None)))
            return dx

    class Gather(T.autograd.Function):
        """out[r] = src[idx[r]] over rows; backward scatter-adds."""

        @staticmethod
        def forward(ctx, src, idx):
            src = src.contiguous()
            w = src.shape[1]
            out = T.empty(idx.numel(), w, dtype=T.float32, device=src.device)
            check(lib().latec_train_gather(ptr(src), w, ptr(idx), idx.numel(), ptr(out), ptr(None)))
            ctx.save_for_backward(idx)
            ctx.n_src = src.shape[0]
            return out

        @staticmethod
        def backward(ctx, dout):
            (idx,) = ctx.saved_tensors
            dout = dout.contiguous()
            w = dout.shape[1]
            dsrc = T.empty(ctx.n_src, w, dtype=T.float32, device=dout.device)
            acc = T.empty(ctx.n_src * w + 1, dtype=T.int64, device=dout.device)
            check(lib().latec_train_scatter_add(ptr(dout), w, ptr(idx), idx.numel(), ptr(dsrc),
                                                ctx.n_src, ptr(acc), ptr(None)))
            return dsrc, None

    class WriteRows(T.autograd.Function):
        """E[idx] = vals in place (each row is written once per schedule)."""

        @staticmethod
        def forward(ctx, E, idx, vals):
            vals = vals.contiguous()
            w = E.shape[1]
            check(lib().latec_train_write_rows(ptr(E), w, ptr(idx), idx.numel(), ptr(vals),
                                               ptr(None)))
            ctx.mark_dirty(E)
            ctx.save_for_backward(idx)
            return E

        @staticmethod
        def backward(ctx, dE):
            (idx,) = ctx.saved_tensors
            dE = dE.contiguous()
            w = dE.shape[1]
            dvals = T.empty(idx.numel(), w, dtype=T.float32, device=dE.device)
            check(lib().latec_train_gather(ptr(dE), w, ptr(idx), idx.numel(), ptr(dvals),
                                           ptr(None)))
            dprev = dE.clone()
            check(lib().latec_train_write_rows(ptr(dprev), w, ptr(idx), idx.numel(), ptr(None),
                                               ptr(None)))
            return dprev, None, dvals

    class Attn(T.autograd.Function):
        """AttnPool over CSR segments + null token; s already scaled."""

        @staticmethod
        def forward(ctx, s, v, null_k, null_v, off, scale):
            s = s.contiguous().reshape(-1)
            v = v.contiguous()
            d = v.shape[1]
            n_items = off.numel() - 1
            out = T.empty(n_items, d, dtype=T.float32, device=v.device)
            w_rows = T.empty(max(s.numel(), 1), dtype=T.float32, device=v.device)
            w_null = T.empty(max(n_items, 1), dtype=T.float32, device=v.device)
            null_s = float(null_k.detach()) * scale
            check(lib().latec_train_attn(n_items, ptr(off), ptr(s), ptr(v), d, null_s,
                                         ptr(null_v.contiguous()), ptr(out), ptr(w_rows),
                                         ptr(w_null), ptr(None)))
            ctx.save_for_backward(v, null_v, w_rows, w_null, off)
            ctx.scale = scale
            ctx.n_s = s.numel()
            return out

        @staticmethod
        def backward(ctx, dout):
            v, null_v, w_rows, w_null, off = ctx.saved_tensors
            dout = dout.contiguous()
            d = v.shape[1]
            ds = T.empty(max(ctx.n_s, 1), dtype=T.float32, device=v.device)
            dv = T.empty_like(v)
            n_items = off.numel() - 1
            dnull = T.zeros(d + 1, dtype=T.float32, device=v.device)
            work = T.empty(n_items * (d + 1) + 1024 * (d + 1) + 16, dtype=T.float32,
                           device=v.device)
            check(lib().latec_train_attn_bwd(n_items, ptr(off), ptr(v), d,
                                             ptr(null_v.contiguous()), ptr(w_rows), ptr(w_null),
                                             ptr(dout), ptr(ds), ptr(dv), ptr(dnull), ptr(work),
                                             work.numel(), ptr(None)))
            return (ds[: ctx.n_s], dv, (dnull[0] * ctx.scale).reshape(()), dnull[1:].clone(),
                    None, None)

    class HeadLoss(T.autograd.Function):
        """Per-point bits of the head outputs (M, 2) at targets t, bounds eb."""

        @staticmethod
        def forward(ctx, o, t, eb, mse_w):
            o = o.contiguous()
            M = o.shape[0]
            bits = T.empty(M, dtype=T.float32, device=o.device)
            dO = T.empty(M, 2, dtype=T.float32, device=o.device)
            mu = T.empty(M, dtype=T.float32, device=o.device)
            check(lib().latec_train_nll(M, ptr(o), ptr(t.contiguous()), ptr(eb.contiguous()),
                                        float(mse_w), ptr(bits), ptr(dO), ptr(mu), ptr(None)))
            ctx.save_for_backward(dO)
            ctx.mu = mu
            return bits

        @staticmethod
        def backward(ctx, dbits):
            (dO,) = ctx.saved_tensors
            return dO * dbits.reshape(-1, 1), None, None, None

    return Linear, Gelu, Gather, WriteRows, Attn, HeadLoss


_F = {}


def F():
    if not _F:
        names = ("Linear", "Gelu", "Gather", "WriteRows", "Attn", "HeadLoss")
        _F.update(zip(names, _functions()))
    return _F


# ---- stage geometry as index tensors ----------------------------------------

@dataclass
class StageIdx:
    M: int
    q: object            # (M,) frame flat index in one field
    # message rows, item-major (m, k, live l)
    R: int
    item_off: object     # (M*K + 1,) CSR over rows
    bi_rows: object      # rows computed by the two-sided MLP
    bi_in: object        # E rows (frame*K + k) of the negative neighbour
    bi_ip: object
    bi_sn: object        # state index rows (negative / positive side)
    bi_sp: object
    bi_z: object         # (nb, 2) [logd_neg, logd_pos]
    dir_rows: object
    dir_src: object
    dir_st: object
    dir_z: object        # (nd, 2) [sign, logd]


def _stage_index(steps, qidx, M, L, K, dirs, fstr, agg, lognnz):
    """Index tensors of one stage from the geometry (steps (M, L))."""
    T = _torch()
    dev = D.device()
    st = steps[: M * L].reshape(M, L).to(T.int64)
    sp, sn = st & 0xFFFF, st >> 16
    live = (sp > 0) | (sn > 0)
    cnt = live.sum(1)
    start = T.cumsum(cnt, 0) - cnt
    m_idx, l_idx = live.nonzero(as_tuple=True)
    j = T.arange(m_idx.numel(), device=dev) - start[m_idx]
    k_ar = T.arange(K, device=dev)
    # row of (pair, k): start[m]*K + k*cnt[m] + j
    rows = (start[m_idx][:, None] * K + k_ar[None, :] * cnt[m_idx][:, None] + j[:, None])
    item_off = T.zeros(M * K + 1, dtype=T.int64, device=dev)
    item_off[1:] = (start[:, None] * K + (k_ar[None, :] + 1) * cnt[:, None]).reshape(-1)
    R = int(cnt.sum()) * K
    off = T.as_tensor([int(np.dot(d, fstr)) for d in dirs], dtype=T.int64, device=dev)
    q = qidx[:M]
    spl, snl = sp[m_idx, l_idx], sn[m_idx, l_idx]
    both = (spl > 0) & (snl > 0)
    qm, ol = q[m_idx], off[l_idx]
    lnz = T.as_tensor(lognnz, dtype=T.float32, device=dev)[l_idx]
    # DirState index per (line, axis, side) (gnn_predictor.py:657-666)
    dirs_t = T.as_tensor(np.asarray(dirs), dtype=T.int64, device=dev)  # (L, K)
    nnz = (dirs_t != 0).sum(1, keepdim=True)
    idx_pos = T.where(dirs_t > 0, 2 * agg - nnz + 1, T.where(dirs_t < 0, nnz - 1, T.full_like(dirs_t, agg)))
    idx_neg = T.where(dirs_t < 0, 2 * agg - nnz + 1, T.where(dirs_t > 0, nnz - 1, T.full_like(dirs_t, agg)))
    idx_pos, idx_neg = idx_pos.clamp(0, 2 * agg), idx_neg.clamp(0, 2 * agg)

    def erow(center, k):  # E row of a frame cell for axis k
        return center[:, None] * K + k_ar[None, :]

    ld_p = T.log2(spl.clamp(min=1).float()) + lnz
    ld_n = T.log2(snl.clamp(min=1).float()) + lnz
    bsel = both.nonzero(as_tuple=True)[0]
    dsel = (~both).nonzero(as_tuple=True)[0]
    bi_rows = rows[bsel].reshape(-1)
    bi_in = erow(qm[bsel] - snl[bsel] * ol[bsel], k_ar).reshape(-1)
    bi_ip = erow(qm[bsel] + spl[bsel] * ol[bsel], k_ar).reshape(-1)
    bi_sn = idx_neg[l_idx[bsel]].reshape(-1)
    bi_sp = idx_pos[l_idx[bsel]].reshape(-1)
    bi_z = T.stack([ld_n[bsel], ld_p[bsel]], 1)[:, None, :].expand(-1, K, 2).reshape(-1, 2)
    pos = spl[dsel] > 0
    t_side = T.where(pos, spl[dsel], snl[dsel])
    ctr = qm[dsel] + T.where(pos, t_side, -t_side) * ol[dsel]
    dir_rows = rows[dsel].reshape(-1)
    dir_src = erow(ctr, k_ar).reshape(-1)
    dir_st = T.where(pos[:, None], idx_pos[l_idx[dsel]], idx_neg[l_idx[dsel]]).reshape(-1)
    sign = T.where(pos, 1.0, -1.0).float()
    ld_s = T.where(pos, ld_p[dsel], ld_n[dsel])
    dir_z = T.stack([sign, ld_s], 1)[:, None, :].expand(-1, K, 2).reshape(-1, 2)
    return StageIdx(M, q, R, item_off, bi_rows, bi_in, bi_ip, bi_sn, bi_sp,
                    bi_z.contiguous(), dir_rows, dir_src, dir_st, dir_z.contiguous())


def _batch(si: StageIdx, B: int, N: int, K: int) -> StageIdx:
    """Replicate one field's stage indices over a batch of B fields laid out
    field after field (frame rows + b*N, message rows + b*R)."""
    if B == 1:
        return si
    T = _torch()
    dev = D.device()
    b = T.arange(B, device=dev)

    def rep(x, step):
        return (x[None, :] + b[:, None] * step).reshape(-1)

    def tile(x):
        return x.repeat((B,) + (1,) * (x.dim() - 1))

    R, NK = si.R, N * K
    off = (si.item_off[:-1][None, :] + b[:, None] * R).reshape(-1)
    off = T.cat([off, T.as_tensor([B * R], dtype=T.int64, device=dev)])
    return StageIdx(si.M * B, rep(si.q, N), R * B, off,
                    rep(si.bi_rows, R), rep(si.bi_in, NK), rep(si.bi_ip, NK), tile(si.bi_sn),
                    tile(si.bi_sp), tile(si.bi_z), rep(si.dir_rows, R), rep(si.dir_src, NK),
                    tile(si.dir_st), tile(si.dir_z))


class GeometryCache:
    """Per (shape, levels) stage indices of one field (single chunk)."""

    def __init__(self, d, agg):
        self.d, self.agg = d, agg
        self._c = {}

    def get(self, shape, levels):
        key = (tuple(shape), levels)
        hit = self._c.get(key)
        if hit is None:
            from .gnn_predictor import GnnModel, StageGeom, half_directions
            from .predictor import get_plan
            from ._lib import GnnFrame

            ndim = len(shape)
            stride = 1 << levels
            model = GnnModel(self.d, self.agg, GM.init_params(self.d, self.agg, 0), ndim, "fp32")
            plan = get_plan(shape, 1, levels, stride, 1)
            fr = GnnFrame()
            for j, n in enumerate(shape):
                fr.origin[j], fr.tshape[j], fr.edges[j], fr.grid[j] = 0, n, stride * 10 ** 6, 1
            dirs = half_directions(ndim, self.agg)
            nnz = np.asarray([(np.asarray(d) != 0).sum() for d in dirs])
            lognnz = (0.5 * np.log2(nnz)).astype(np.float32)
            fstr = np.cumprod((1,) + tuple(shape)[:0:-1])[::-1]
            stages = []
            for s in range(plan.n_stages):
                if plan.counts[s] == 0:
                    continue
                g = StageGeom(model, plan, s, fr)
                stages.append(_stage_index(g.steps, g.qidx, g.M, len(dirs), ndim, dirs, fstr,
                                           self.agg, lognnz))
            model.close()
            hit = stages
            self._c[key] = hit
        return hit


# ---- the model ----------------------------------------------------------------

class GnnTrainer:
    """FP32 master parameters + the training forward over a schedule.
    ``precision``: the GEMM mode of run_stages, "fp32" or "bf16" (BF16 MFMA
    operands, FP32 accumulation)."""

    def __init__(self, d: int = 64, agg: int = 2, seed: int = 0, params: dict | None = None,
                 precision: str = "fp32"):
        T = _torch()
        if precision not in ("fp32", "bf16"):
            raise ValueError("precision must be 'fp32' or 'bf16'")
        self.precision = precision
        self.d, self.agg = d, agg
        P = params if params is not None else GM.init_params(d, agg, seed)
        self.params = {n: T.nn.Parameter(T.from_numpy(np.asarray(P[n], np.float32).copy())
                                         .to(D.device()))
                       for n, _ in GM.param_shapes(d, agg)}
        self.geoms = GeometryCache(d, agg)

    def parameters(self):
        return list(self.params.values())

    def numpy_params(self) -> dict:
        return {n: p.detach().cpu().numpy() for n, p in self.params.items()}

    def save(self, path, extra=None):
        GM.save_checkpoint(path, self.d, self.agg, self.numpy_params(), extra)

    # -- pieces
    def _mlp2(self, prefix, x):
        f = F()
        P = self.params
        h = f["Gelu"].apply(f["Linear"].apply(x, P[prefix + ".0.weight"], P[prefix + ".0.bias"]))
        return f["Linear"].apply(h, P[prefix + ".2.weight"], P[prefix + ".2.bias"])

    def _embed(self, E2, si: StageIdx, K: int):
        """Line messages + line pool -> ctx (M*K, d)."""
        T = _torch()
        f, P, d = F(), self.params, self.d
        ds = max(1, d // 4)
        de = d + ds
        scale = d ** -0.5
        table = P["dir_state.table"]
        parts, order = [], []
        if si.bi_rows.numel():
            W = P["bidir.net.0.weight"]
            Sn = f["Linear"].apply(table, W[:, d:de], None)
            Sp = f["Linear"].apply(table, W[:, de + d:2 * de], None)
            h = (f["Linear"].apply(f["Gather"].apply(E2, si.bi_in), W[:, :d], P["bidir.net.0.bias"])
                 + f["Linear"].apply(f["Gather"].apply(E2, si.bi_ip), W[:, de:de + d], None)
                 + f["Gather"].apply(Sn, si.bi_sn) + f["Gather"].apply(Sp, si.bi_sp)
                 + f["Linear"].apply(si.bi_z, W[:, 2 * de:2 * de + 2], None))
            parts.append(f["Linear"].apply(f["Gelu"].apply(h), P["bidir.net.2.weight"],
                                           P["bidir.net.2.bias"]))
            order.append(si.bi_rows)
        if si.dir_rows.numel():
            W = P["dir.net.0.weight"]
            S = f["Linear"].apply(table, W[:, d:de], None)
            h = (f["Linear"].apply(f["Gather"].apply(E2, si.dir_src), W[:, :d], P["dir.net.0.bias"])
                 + f["Gather"].apply(S, si.dir_st)
                 + f["Linear"].apply(si.dir_z, W[:, de:de + 2], None))
            parts.append(f["Linear"].apply(f["Gelu"].apply(h), P["dir.net.2.weight"],
                                           P["dir.net.2.bias"]))
            order.append(si.dir_rows)
        n_items = si.item_off.numel() - 1
        if si.R == 0:
            msg = T.zeros(0, d, dtype=T.float32, device=D.device())
            s = T.zeros(0, dtype=T.float32, device=D.device())
            v = msg
        else:
            cat = T.cat(parts, 0) if len(parts) > 1 else parts[0]
            inv = T.empty(si.R, dtype=T.int64, device=D.device())
            inv[T.cat(order)] = T.arange(si.R, device=D.device())
            msg = f["Gather"].apply(cat, inv)
            s = self._mlp2("line_pool.wk", msg).reshape(-1) * scale
            v = f["Linear"].apply(msg, P["line_pool.wv.weight"], P["line_pool.wv.bias"])
        assert n_items == si.M * K
        return f["Attn"].apply(s, v, P["line_pool.null_k"], P["line_pool.null_v"], si.item_off,
                               scale)

    def _head(self, ctx, M, K, eb_pts):
        T = _torch()
        f, P, d = F(), self.params, self.d
        scale = d ** -0.5
        s = self._mlp2("axis_pool.wk", ctx).reshape(-1) * scale
        v = f["Linear"].apply(ctx, P["axis_pool.wv.weight"], P["axis_pool.wv.bias"])
        off = T.arange(0, M * K + 1, K, dtype=T.int64, device=D.device())
        pooled = f["Attn"].apply(s, v, P["axis_pool.null_k"], P["axis_pool.null_v"], off, scale)
        log_eb = T.log2(eb_pts.clamp_min(1.17549435e-38)).reshape(-1, 1)
        W = P["head.net.0.weight"]
        h = (f["Linear"].apply(pooled, W[:, :d], P["head.net.0.bias"])
             + f["Linear"].apply(log_eb, W[:, d:d + 1], None))
        return f["Linear"].apply(f["Gelu"].apply(h), P["head.net.2.weight"], P["head.net.2.bias"])

    def _finalize(self, E2, ctx, si: StageIdx, vals, K):
        T = _torch()
        f, P, d = F(), self.params, self.d
        ve = self._mlp2("init.net", vals.reshape(-1, 1))
        rep = T.arange(si.M, device=D.device()).repeat_interleave(K)
        W = P["mix.net.0.weight"]
        h = (f["Linear"].apply(ctx, W[:, :d], P["mix.net.0.bias"])
             + f["Linear"].apply(f["Gather"].apply(ve, rep), W[:, d:2 * d], None))
        fin = f["Linear"].apply(f["Gelu"].apply(h), P["mix.net.2.weight"], P["mix.net.2.bias"])
        rows = (si.q[:, None] * K + T.arange(K, device=D.device())[None, :]).reshape(-1)
        return f["WriteRows"].apply(E2, rows, fin)

    def run_stages(self, x, shape, levels, eb, noise=True, closed_loop=False, mse_weight=0.0,
                   gen=None):
        """Sum of bits over every predicted point of the batch x (B, N) in
        [0, 1] at bounds eb ((B,) or scalar), and the point count.  Returns
        (bits_sum, n_points, per-stage (mu, target) for diagnostics)."""
        prev = _PRECISION["gemm"]
        _PRECISION["gemm"] = self.precision
        try:
            return self._run_stages(x, shape, levels, eb, noise, closed_loop, mse_weight, gen)
        finally:
            _PRECISION["gemm"] = prev

    def _run_stages(self, x, shape, levels, eb, noise, closed_loop, mse_weight, gen):
        T = _torch()
        B, N = x.shape
        K = len(shape)
        stages = [_batch(si, B, N, K) for si in self.geoms.get(shape, levels)]
        eb = T.as_tensor(eb, dtype=T.float32, device=x.device).reshape(-1)
        eb = eb.expand(B) if eb.numel() == 1 else eb
        xf = x.reshape(-1)
        ebf = eb[:, None].expand(B, N).reshape(-1)
        known = T.full_like(xf, 0.5)

        def reveal(q, pred=None):
            if closed_loop and pred is not None:
                w = 2 * ebf[q]
                known[q] = pred + w * T.round((xf[q] - pred) / w)
            elif noise:
                u = T.rand(q.numel(), generator=gen, device=x.device) * 2 - 1
                known[q] = (xf[q] + u * ebf[q]).clamp(0, 1)
            else:
                known[q] = xf[q]

        a0 = stages[0].q
        if closed_loop:
            w = 2 * ebf[a0]
            known[a0] = w * T.round(xf[a0] / w)
        else:
            reveal(a0)
        E2 = T.zeros(B * N * K, self.d, dtype=T.float32, device=x.device)
        bits = T.zeros((), dtype=T.float64, device=x.device)
        npts, ctx_prev, diag = 0, None, []
        for i in range(1, len(stages)):
            gp, gh = stages[i - 1], stages[i]
            if ctx_prev is None:
                ctx_prev = self._embed(E2, gp, K)
            E2 = self._finalize(E2, ctx_prev, gp, known[gp.q].detach(), K)
            ctx = self._embed(E2, gh, K)
            out = self._head(ctx, gh.M, K, ebf[gh.q])
            tgt = xf[gh.q]
            b = F()["HeadLoss"].apply(out, tgt, ebf[gh.q], mse_weight)
            bits = bits + b.double().sum()
            npts += gh.M
            mu = T.sigmoid(out[:, 0].detach())
            diag.append((mu, tgt))
            reveal(gh.q, mu.float() if closed_loop else None)
            ctx_prev = ctx
        return bits, npts, diag


def laplace_bits_reference(mu, log_b, target, eb):
    """Plain torch form of the loss (for tests of the loss kernel)."""
    T = _torch()
    eb = T.as_tensor(eb, dtype=target.dtype, device=target.device)
    b = T.exp2(log_b).clamp_min(T.finfo(target.dtype).tiny)
    a = (target - mu).abs()
    rho, e = a / b, eb / b
    tail_mass = T.log1p(-T.exp((-2.0 * e).clamp_max(-T.finfo(target.dtype).eps)))
    t = rho - e
    tm = 32.0 * math.log(2.0)
    t = T.where(t > tm, tm + T.log1p((t - tm).clamp_min(0.0)), t)
    logp_tail = -math.log(2.0) - t + tail_mass
    rc = T.minimum(rho, e)
    cross = 1.0 - 0.5 * T.exp(-(e - rc)) - 0.5 * T.exp(-(e + rc))
    logp = T.where(a >= eb, logp_tail, T.log(cross.clamp_min(1e-30)))
    return -logp / math.log(2.0)


# The training configuration of the benchmark's learned-predictor workload:
# batches of 32 fields (16 x 2-D 256^2 with correlation 2..32, 16 x 4-D 16^4
# with correlation 1..8), turbulence fraction 0.5, 3 discontinuities, stride 8
# (levels 3), eb 1e-3, targets from the codec's own closed-loop reconstruction
# (train_gnn.py:413-414, 446-450), BF16 GEMMs with FP32 master weights.
C5 = {"batch2d": 16, "shape2d": (256, 256), "corr2d": (2.0, 32.0), "batch4d": 16,
      "shape4d": (16, 16, 16, 16), "corr4d": (1.0, 8.0), "levels": 3,
      "eb_range": (1e-3, 1e-3), "turbulence_frac": 0.5, "n_jumps": 3, "closed_loop": True}


def train(trainer: GnnTrainer, steps: int, *, lr=2e-3, warmup=50, batch2d=8, shape2d=(128, 128),
          corr2d=(2.0, 32.0), batch4d=2, shape4d=(16, 16, 16, 16), corr4d=(1.0, 8.0),
          levels=3, eb_range=(1e-4, 1e-2), turbulence_frac=0.5, n_jumps=3, seed=0,
          log=None, mse_weight=0.0, closed_loop=False):
    """AdamW (weight decay 1e-3) with linear warm-up and cosine decay
    (train_gnn.py:937-950) on synthetic 2-D and 4-D batches, eb drawn
    log-uniformly from ``eb_range`` per sample; targets teacher-forced (the
    truth plus U(-eb, eb) noise) or, with ``closed_loop``, each stage's
    context the codec's own quantized reconstruction (``C5`` holds the
    benchmark's configuration)."""
    from .synthetic import _gen, training_batch

    T = _torch()
    opt = T.optim.AdamW(trainer.parameters(), lr=lr, weight_decay=1e-3)

    def lr_factor(step):
        if step < warmup:
            return (step + 1) / warmup
        return 0.5 * (1 + math.cos(math.pi * min(1.0, (step - warmup) / max(1, steps - warmup))))

    sched = T.optim.lr_scheduler.LambdaLR(opt, lr_factor)
    gen = _gen(seed, D.device())
    hist = []
    for step in range(steps):
        opt.zero_grad(set_to_none=True)
        total, tot_pts = 0.0, 0
        for batch, shape, corr in ((batch2d, shape2d, corr2d), (batch4d, shape4d, corr4d)):
            if not batch:
                continue
            x = training_batch(batch, shape, corr, gen, D.device(), turbulence_frac, n_jumps)
            lo, hi = math.log(eb_range[0]), math.log(eb_range[1])
            eb = T.exp(lo + (hi - lo) * T.rand(batch, generator=gen, device=D.device()))
            bits, n, _ = trainer.run_stages(x, shape, levels, eb, noise=True, gen=gen,
                                            mse_weight=mse_weight, closed_loop=closed_loop)
            (bits / n).backward()
            total += float(bits.detach()) / n
            tot_pts += n
        T.nn.utils.clip_grad_norm_(trainer.parameters(), 1.0)
        opt.step()
        sched.step()
        hist.append(total)
        if log and (step % 10 == 0 or step == steps - 1):
            log(f"step {step} bits/pt {total:.4f} lr {sched.get_last_lr()[0]:.2e}")
    return hist
