"""FP64 CPU restatement of the LATEC GNN predictor, its geometry and its loss.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py`` (the FP64 run that the BF16/FP32 GPU predictions are measured
against).  The product package (``latec_amd``) never imports this module.

Everything here is plain torch float64 on the CPU, written for clarity:

* geometry by exhaustive scan (the method of the reference's own geometry test,
  ``tests/test_gnn_predictor.py:34-65``): for every query point and every
  half-direction, the nearest known sample on each side within the radius;
* the model as pure functions of a parameter dict whose keys are the
  reference checkpoint's ``state_dict`` keys (``latec/gnn_predictor.py:593-933``);
* ``stage_forward`` (``latec/gnn_predictor.py:984-1000``) and the whole-field
  stage loop of the trainer (``scripts/train_gnn.py:373-459``);
* the discretized-Laplace NLL in bits (``scripts/train_gnn.py:195-230``).

Pinned by: the scan (geometry); golden forward vectors written by
``tests/golden/make_gnn_golden.py`` from the reference module (model); the
Laplace CDF from scipy (loss).
"""

from __future__ import annotations

import itertools
import math

import numpy as np
import torch

F64 = torch.float64


# ---- geometry --------------------------------------------------------------

def half_directions(ndim: int, agg_level: int | None = None):
    """One representative per lattice line: vectors in {-1,0,1}^ndim whose first
    non-zero entry is +1, with at most ``agg_level`` non-zeros
    (gnn_predictor.py:117-136), in itertools.product order."""
    out = []
    for d in itertools.product((-1, 0, 1), repeat=ndim):
        nz = [x for x in d if x != 0]
        if nz and nz[0] > 0 and (agg_level is None or len(nz) <= agg_level):
            out.append(d)
    return out


def line_consts(dirs):
    """(cos (L, ndim), lognnz (L,)): unit direction and 0.5*log2(nnz)
    (gnn_predictor.py:245-249)."""
    v = torch.tensor(dirs, dtype=F64)
    nnz = (v != 0).sum(1).to(F64)
    return v / torch.sqrt(nnz)[:, None], 0.5 * torch.log2(nnz)


def scan_steps(known: np.ndarray, coords: np.ndarray, dirs, max_radius: int,
               usable=None):
    """Nearest known sample along +d and -d for each query, by stepping.

    known: bool array over the grid (known before the stage); coords (M, ndim);
    returns (sp, sn) int64 (L, M), 0 where no sample lies within
    ``min(max_radius, max(shape))`` steps inside the grid.  ``usable`` (bool
    array, optional) further restricts which known cells count."""
    shape = np.asarray(known.shape)
    limit = min(int(max_radius), int(shape.max()))
    ok = known if usable is None else (known & usable)
    M = coords.shape[0]
    sp = np.zeros((len(dirs), M), np.int64)
    sn = np.zeros((len(dirs), M), np.int64)
    for li, d in enumerate(dirs):
        for out, sgn in ((sp, 1), (sn, -1)):
            dv = sgn * np.asarray(d)
            found = np.zeros(M, bool)
            for t in range(1, limit + 1):
                nb = coords + t * dv
                inb = np.all((nb >= 0) & (nb < shape), axis=1)
                hit = np.zeros(M, bool)
                hit[inb] = ok[tuple(nb[inb].T)]
                new = hit & ~found
                out[li, new] = t
                found |= new
                if found.all():
                    break
    return sp, sn


def frame_scan_steps(known, usable, coords, dirs, limit):
    """Chunk-frame geometry (gnn_predictor.py:139-176 with the validity of
    :1241-1275): along +d and -d the first cell within ``limit`` steps that
    is ``known`` (the schedule's periodic pattern before the stage, halo
    included); the step counts only when that cell is ``usable`` (interior,
    or a decoded halo cell), else the line is dead (0).  The frame must hold
    every cell within ``limit`` steps of a query.  Returns (sp, sn) (L, M)."""
    M = coords.shape[0]
    sp = np.zeros((len(dirs), M), np.int64)
    sn = np.zeros((len(dirs), M), np.int64)
    for li, d in enumerate(dirs):
        for out, sgn in ((sp, 1), (sn, -1)):
            dv = sgn * np.asarray(d)
            done = np.zeros(M, bool)
            for t in range(1, int(limit) + 1):
                nb = tuple((coords + t * dv).T)
                hit = known[nb] & ~done
                out[li, hit & usable[nb]] = t
                done |= hit
                if done.all():
                    break
    return sp, sn


def stage_geometry(known, coords, shape, dirs, max_radius):
    """Per-stage geometry dict: query flat index, steps, line offsets."""
    sp, sn = scan_steps(known, coords, dirs, max_radius)
    strides = np.cumprod((1,) + tuple(shape)[:0:-1])[::-1].astype(np.int64)
    q = coords @ strides if len(coords) else np.zeros(0, np.int64)
    off = np.asarray([np.asarray(d) @ strides for d in dirs], np.int64)
    return {"q": q, "sp": sp, "sn": sn, "off": off, "M": len(q)}


# ---- model -----------------------------------------------------------------

def gelu(x):
    """Exact (erf) GELU, torch.nn.GELU's default."""
    return 0.5 * x * (1.0 + torch.erf(x / math.sqrt(2.0)))


def linear(P, name, x):
    return x @ P[name + ".weight"].T + P[name + ".bias"]


def mlp2(P, prefix, x):
    """Linear -> GELU -> Linear (``_mlp`` with three sizes, gnn_predictor.py:583-590)."""
    return linear(P, prefix + ".2", gelu(linear(P, prefix + ".0", x)))


def dims(P):
    d = P["init.net.2.weight"].shape[0]
    n_states, ds = P["dir_state.table"].shape
    return d, (n_states - 1) // 2, ds


def state_index(cos, sign, agg):
    """DirState._index (gnn_predictor.py:657-666): bucket of the signed cosine."""
    v = sign * cos
    nnz = torch.round(1.0 / (v * v))  # inf where v == 0, unused there
    idx = torch.where(v > 0, 2 * agg - nnz + 1,
                      torch.where(v < 0, nnz - 1, torch.full_like(v, float(agg))))
    return idx.clamp(0, 2 * agg).long()


def attn_pool(P, prefix, msgs, valid):
    """AttnPool (gnn_predictor.py:720-747): msgs (L, ..., d), valid broadcastable
    to (L, ...) -> (..., d)."""
    d = msgs.shape[-1]
    scale = d ** -0.5
    v = linear(P, prefix + ".wv", msgs)
    s = mlp2(P, prefix + ".wk", msgs)[..., 0] * scale
    s = torch.where(valid, s, torch.full_like(s, -math.inf))
    sn = (P[prefix + ".null_k"] * scale).expand(1, *s.shape[1:])
    s = torch.cat([s, sn], 0)
    v = torch.cat([v, P[prefix + ".null_v"].expand(1, *v.shape[1:])], 0)
    w = torch.softmax(s, 0)
    return (w[..., None] * v).sum(0)


def line_messages(P, E, geom, dirs):
    """Per-line messages (L, B, M, K, d) and validity (L, M)
    (GNN._line_messages, gnn_predictor.py:818-878), dense form."""
    d, agg, ds = dims(P)
    B, N, K, _ = E.shape
    L, M = geom["sp"].shape
    cos, lognnz = line_consts(dirs)
    sp = torch.from_numpy(geom["sp"])
    sn = torch.from_numpy(geom["sn"])
    q = torch.from_numpy(geom["q"])[None, :]
    off = torch.from_numpy(geom["off"])[:, None]
    vp, vn = sp > 0, sn > 0
    ip = torch.where(vp, q + sp * off, torch.zeros_like(sp))
    in_ = torch.where(vn, q - sn * off, torch.zeros_like(sn))
    logdp = torch.log2(sp.clamp(min=1).to(F64)) + lognnz[:, None]
    logdn = torch.log2(sn.clamp(min=1).to(F64)) + lognnz[:, None]
    table = P["dir_state.table"]
    idx_p = state_index(cos, 1.0, agg)   # (L, K)
    idx_n = state_index(cos, -1.0, agg)
    st_p = table[idx_p][:, None, :, :]   # (L, 1, K, ds)
    st_n = table[idx_n][:, None, :, :]
    Ep = E[:, ip.reshape(-1)].reshape(B, L, M, K, d).permute(1, 0, 2, 3, 4)
    En = E[:, in_.reshape(-1)].reshape(B, L, M, K, d).permute(1, 0, 2, 3, 4)
    ep = torch.cat([Ep, st_p[:, None].expand(L, B, M, K, ds)], -1)
    en = torch.cat([En, st_n[:, None].expand(L, B, M, K, ds)], -1)
    de = d + ds
    lp = logdp[:, None, :, None, None].expand(L, B, M, K, 1)
    ln = logdn[:, None, :, None, None].expand(L, B, M, K, 1)
    W = P["bidir.net.0.weight"]
    h1 = (en @ W[:, :de].T + ep @ W[:, de:2 * de].T + P["bidir.net.0.bias"]
          + torch.cat([ln, lp], -1) @ W[:, 2 * de:].T)
    bi = linear(P, "bidir.net.2", gelu(h1))
    Wd = P["dir.net.0.weight"]
    one = torch.ones_like(lp)

    def one_sided(e, sign, logd):
        h = e @ Wd[:, :de].T + P["dir.net.0.bias"] + torch.cat([sign, logd], -1) @ Wd[:, de:].T
        return linear(P, "dir.net.2", gelu(h))

    dp_msg = one_sided(ep, one, lp)
    dn_msg = one_sided(en, -one, ln)
    both = (vp & vn)[:, None, :, None, None]
    only_p = (vp & ~vn)[:, None, :, None, None]
    only_n = (vn & ~vp)[:, None, :, None, None]
    msg = torch.where(both, bi, torch.where(only_p, dp_msg,
                      torch.where(only_n, dn_msg, torch.zeros_like(bi))))
    return msg, vp | vn


def embed(P, E, geom, dirs):
    """Per-axis contexts (B, M, K, d): line messages pooled by ``line_pool``
    (GNN._embed_block / embed, gnn_predictor.py:880-906)."""
    msg, valid = line_messages(P, E, geom, dirs)
    return attn_pool(P, "line_pool", msg, valid[:, None, :, None])


def finalize(P, ctx, vals):
    """GNN.finalize (gnn_predictor.py:908-920): mix(ctx, init(val))."""
    ve = mlp2(P, "init.net", vals[..., None])[:, :, None, :].expand_as(ctx)
    return mlp2(P, "mix.net", torch.cat([ctx, ve], -1))


def head_of(P, ctx, eb):
    """axis_pool over the K axes, then PredHead (gnn_predictor.py:749-789, 922-931).
    eb: python float or (B,) tensor (normalized).  Returns (mu, log_b) (B, M)."""
    B, M, K, d = ctx.shape
    pooled = attn_pool(P, "axis_pool", ctx.permute(2, 0, 1, 3),
                       torch.ones(K, 1, 1, dtype=torch.bool))
    eb = torch.as_tensor(eb, dtype=F64).reshape(-1).expand(B)
    log_eb = torch.log2(eb.clamp_min(torch.finfo(F64).tiny))
    cond = log_eb.view(B, 1, 1).expand(B, M, 1)
    out = mlp2(P, "head.net", torch.cat([pooled, cond], -1))
    mu = torch.sigmoid(out[..., 0])
    log_b = log_eb.view(B, 1) + out[..., 1].clamp(-4.0, 12.0)
    return mu, log_b


def stage_forward(P, E, gp, gh, fvals, dirs, finalize_ctx=None, eb=0.01):
    """stage_forward (gnn_predictor.py:984-1000): finalize the previous stage's
    points into E, then predict this stage.  Returns ((mu, log_b), E, head_ctx)."""
    if gp is not None and gp["M"]:
        ctx = finalize_ctx if finalize_ctx is not None else embed(P, E, gp, dirs)
        E = E.clone()
        E[:, torch.from_numpy(gp["q"])] = finalize(P, ctx, fvals)
    head_ctx = embed(P, E, gh, dirs)
    return head_of(P, head_ctx, eb), E, head_ctx


def schedule_geoms(shape, levels, stride, block, max_radius, agg):
    """Whole-field geometry per non-empty stage, in stage order (the codec's
    ``build_stage_geoms``, gnn_predictor.py:539-580), from the C oracle's
    stage masks."""
    from oracle import oracle as O

    dirs = half_directions(len(shape), agg)
    cum = np.zeros(shape, bool)
    geoms = []
    for mask, _, _, _ in O.stage_plan(shape, levels, stride, block):
        if mask.any():
            coords = np.stack(np.nonzero(mask), 1)
            geoms.append(stage_geometry(cum, coords, shape, dirs, max_radius))
        cum |= mask
    return geoms, dirs


def run_schedule(P, x, shape, levels, stride, eb, max_radius=None, agg=None, block=1):
    """Teacher-free forward over a whole schedule with the true values revealed
    (the trainer's ``run_stages`` with zero noise, train_gnn.py:373-459).
    x: (B, N) float64 in [0, 1].  Returns [(mu, log_b)] per predicted stage."""
    d, agg0, _ = dims(P)
    agg = agg0 if agg is None else agg
    geoms, dirs = schedule_geoms(shape, levels, stride, block,
                                 stride if max_radius is None else max_radius, agg)
    B = x.shape[0]
    E = torch.zeros(B, x.shape[1], len(shape), d, dtype=F64)
    out, ctx = [], None
    for i in range(1, len(geoms)):
        gp, gh = geoms[i - 1], geoms[i]
        (mu, lb), E, ctx = stage_forward(P, E, gp, gh, x[:, gp["q"]], dirs, ctx, eb)
        out.append((mu, lb))
    return out


# ---- loss ------------------------------------------------------------------

def laplace_nll_bits(mu, log_b, target, eb):
    """Discretized Laplace code length in bits (train_gnn.py:195-230), float64,
    with the reference's 32-bit logarithmic tail compression."""
    eb = torch.as_tensor(eb, dtype=F64).reshape(-1, 1)
    b = torch.exp2(log_b).clamp_min(torch.finfo(F64).tiny)
    abs_r = (target - mu).abs()
    rho, e = abs_r / b, eb / b
    tail_mass = torch.log1p(-torch.exp((-2.0 * e).clamp_max(-torch.finfo(F64).eps)))
    t = rho - e
    t_max = 32.0 * math.log(2.0)
    t = torch.where(t > t_max, t_max + torch.log1p((t - t_max).clamp_min(0.0)), t)
    logp_tail = -math.log(2.0) - t + tail_mass
    rc = torch.minimum(rho, e)
    cross = 1.0 - 0.5 * torch.exp(-(e - rc)) - 0.5 * torch.exp(-(e + rc))
    logp_cross = torch.log(cross.clamp_min(1e-30))
    logp = torch.where(abs_r >= eb, logp_tail, logp_cross)
    return -logp / math.log(2.0)


# ---- gate rate proxy ---------------------------------------------------------

SCALE_LO_DIV, SCALE_HI_MULT = 16.0, 4096.0
GATE_T = np.arange(2, 14, dtype=np.float64)
GATE_SHIFTS = (0, 2, 4, 6)


def gate_scale_tables(eb):
    """The gate's scale snapping (gnn_codec.py:880-917): level =
    rint((log(clip(b)) - log lo) / step) on the 64-level grid [eb/16, 4096 eb],
    as 63 float32 thresholds, and the grid exp(log lo + step * level)."""
    lo_f, hi_f = eb / SCALE_LO_DIV, eb * SCALE_HI_MULT
    log_lo = np.log(lo_f)
    step = (np.log(hi_f) - log_lo) / 63.0
    thr = np.empty(63, np.float32)
    hi_bits = int(np.asarray(np.float32(hi_f)).view(np.uint32))
    for t in range(1, 64):  # first float32 whose level reaches t (bisection)
        lo, hi = 0, hi_bits
        while lo < hi:
            mid = (lo + hi) // 2
            v = float(np.asarray(mid, np.uint32).view(np.float32))
            lev = np.rint((np.log(np.clip(v, lo_f, hi_f)) - log_lo) / step)
            if lev >= t:
                hi = mid
            else:
                lo = mid + 1
        thr[t - 1] = np.asarray(lo, np.uint32).view(np.float32)
    return thr, np.exp(log_lo + step * np.arange(64, dtype=np.float64))


def laplace_bits(absr, b, eb, radius):
    """_laplace_bits_t (gnn_codec.py:777-797), float64 numpy."""
    thr, grid = gate_scale_tables(eb)
    bs = grid[np.searchsorted(thr, np.asarray(b, np.float32), side="right")]
    k = np.rint(np.asarray(absr, np.float64) / (2 * eb))
    with np.errstate(divide="ignore", over="ignore"):
        p = np.where(k == 0, -np.expm1(-eb / bs),
                     0.5 * np.exp(-((2 * k - 1) * eb / bs)) * -np.expm1(-2 * eb / bs))
        bits = np.minimum(-np.log2(p), 24.0)
    return np.where(k >= radius, 56.0, bits)


def gate_costs(x, pred, scale, ips, eb, radius=1 << 15):
    """Per-bucket bit sums W (1 + 4 * len(ips), 13): row 0 the model, then each
    candidate at each scale shift; buckets by scale against eb * 2^T."""
    x = np.asarray(x, np.float64)
    scale = np.asarray(scale, np.float32)
    bucket = np.searchsorted(eb * np.exp2(GATE_T), scale.astype(np.float64), side="left")
    rows = [laplace_bits(np.abs(x - pred), scale, eb, radius)]
    for ip in ips:
        for sh in GATE_SHIFTS:
            rows.append(laplace_bits(np.abs(x - ip), scale * np.float32(2.0 ** -sh), eb, radius))
    W = np.zeros((len(rows), 13))
    for r, bits in enumerate(rows):
        np.add.at(W[r], bucket, bits)
    return W
