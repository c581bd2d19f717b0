"""GNN predictor and codec on the GPU against the FP64 oracle."""

import numpy as np
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]


def _params(d, agg, seed):
    from latec_amd import gnn_model as GM

    P = GM.init_params(d, agg, seed)
    rng = np.random.default_rng(seed + 1)
    # move the head away from its flat start so mu and log_b vary per point
    P["head.net.2.weight"] = (rng.standard_normal(P["head.net.2.weight"].shape) * 0.5).astype(np.float32)
    P["head.net.2.bias"][1] = 1.0
    P["line_pool.null_v"] = (rng.standard_normal(P["line_pool.null_v"].shape) * 0.3).astype(np.float32)
    return P


@pytest.fixture(params=[1, 0], ids=["table", "scan"])
def geo_table(request):
    """Geometry through the per-stage lookup tables (1) or the cell scan (0)."""
    from latec_amd import _lib

    L = _lib.lib()
    base = L.latec_tune_get(b"gnn_geo_table")
    assert L.latec_tune_set(b"gnn_geo_table", request.param) == 0
    yield request.param
    L.latec_tune_set(b"gnn_geo_table", base)


def _ckpt(tmp_path, d=32, agg=2, seed=0, name="m.pt"):
    from latec_amd import gnn_model as GM

    path = tmp_path / name
    GM.save_checkpoint(path, d, agg, _params(d, agg, seed))
    return path


def _field(shape, seed=0):
    rng = np.random.default_rng(seed)
    x = np.cumsum(np.cumsum(rng.standard_normal(shape), 0), -1)
    x = (x - x.min()) / (x.max() - x.min())
    return x.astype(np.float32)


@pytest.mark.parametrize("shape,levels,d,agg", [
    ((20, 24), 2, 32, 2),
    ((9, 8, 10), 2, 32, 2),
    ((6, 5, 6, 4), 1, 32, 2),
    ((17, 19), 3, 64, 1),
])
@pytest.mark.parametrize("numerics", ["fp32", "bf16"])
def test_forward_matches_oracle(tmp_path, shape, levels, d, agg, numerics):
    """Teacher-forced forward over a whole schedule (true values revealed):
    GPU mu/log_b against the FP64 oracle (fp32: 5e-5; bf16: RMS <= 1e-2)."""
    from latec_amd.gnn_predictor import ChunkedGNNPredictor
    from oracle import gnn_oracle as G

    path = _ckpt(tmp_path, d, agg, 3)
    x = _field(shape, 1)
    stride = 1 << levels
    pr = ChunkedGNNPredictor(path, 0.0, 1.0, levels=levels, anchor_stride=stride,
                             numerics=numerics)
    edges = tuple(-(-n // stride) * stride for n in shape)
    pr.begin(shape, edges)
    recon = torch.from_numpy(x).cuda()[None]
    w = pr.start_wave([0], recon)
    eb = 0.01
    got = []
    for s in w["chain"][1:]:
        _, _, mu, lb = pr.predict_wave_stage(s, recon, eb, want_head=True)
        got.append((mu.cpu().numpy().astype(np.float64), lb.cpu().numpy().astype(np.float64)))
    pr.finish_wave(recon)
    P = {k: torch.from_numpy(np.asarray(v, np.float64)) for k, v in _params(d, agg, 3).items()}
    ref = G.run_schedule(P, torch.from_numpy(x.reshape(1, -1)).double(), shape, levels, stride,
                         eb, stride, agg)
    assert len(ref) == len(got)
    for i, ((mu, lb), (rm, rl)) in enumerate(zip(got, ref)):
        rm, rl = rm.numpy()[0], rl.numpy()[0]
        if numerics == "fp32":
            np.testing.assert_allclose(mu, rm, atol=5e-5, rtol=0, err_msg=f"mu stage {i}")
            np.testing.assert_allclose(lb, rl, atol=5e-4, rtol=0, err_msg=f"log_b stage {i}")
        else:
            assert np.sqrt(np.mean((mu - rm) ** 2)) <= 1e-2, i
            assert np.sqrt(np.mean((lb - rl) ** 2)) <= 1e-1, i


@pytest.mark.parametrize("d,ndim", [(64, 4), (32, 3), (64, 2)])
@pytest.mark.parametrize("numerics", ["fp32", "bf16"])
def test_head_matches_oracle_on_same_context(d, ndim, numerics):
    """latec_gnn_head (axis pool + PredHead) against the FP64 oracle's head_of
    on one context tensor (stored as the mode's type): FP32 mu within 5e-5;
    bf16 (GEMMs on split-bf16 MFMA, the polynomial GELU) mu within 1e-3 and
    log_b within 1e-2, with M not a multiple of the 16-point tile."""
    from latec_amd import _lib
    from latec_amd.gnn_predictor import GnnModel
    from oracle import gnn_oracle as G

    P = _params(d, 2, 7)
    m = GnnModel(d, 2, P, ndim, numerics)
    M, eb, vmin, vmax = 1237, 3e-3, -2.0, 5.0
    rng = np.random.default_rng(11)
    ctx = torch.from_numpy(rng.standard_normal((M, ndim, d)).astype(np.float32) * 1.5).cuda()
    ctx = ctx.to(m.dtype)
    out = [torch.empty(M, dtype=torch.float32, device="cuda") for _ in range(4)]
    _lib.check(_lib.lib().latec_gnn_head(m.handle, ctx.data_ptr(), M, eb, vmin, vmax,
                                         *[o.data_ptr() for o in out], None))
    torch.cuda.synchronize()
    pred, scale, mu, lb = (o.cpu().numpy().astype(np.float64) for o in out)
    Pd = {k: torch.from_numpy(np.asarray(v, np.float64)) for k, v in P.items()}
    rmu, rlb = G.head_of(Pd, ctx.float().cpu().double()[None], eb)
    rmu, rlb = rmu.numpy()[0], rlb.numpy()[0]
    tol_mu, tol_lb = (5e-5, 5e-4) if numerics == "fp32" else (1e-3, 1e-2)
    np.testing.assert_allclose(mu, rmu, atol=tol_mu, rtol=0)
    np.testing.assert_allclose(lb, rlb, atol=tol_lb, rtol=0)
    span = vmax - vmin
    np.testing.assert_allclose(pred, np.clip(mu * span + vmin, vmin, vmax), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(scale, np.exp2(lb) * span, rtol=1e-5)
    assert np.std(rmu) > 1e-2 and np.std(rlb) > 1e-2  # the head is not flat on this input
    m.close()


def test_geometry_matches_scan(tmp_path, geo_table):
    """Stage geometry of the kernel (lookup tables or cell scan) == the
    oracle's exhaustive scan."""
    from latec_amd.gnn_predictor import ChunkedGNNPredictor
    from oracle import gnn_oracle as G

    path = _ckpt(tmp_path, 32, 2, 0)
    for shape, levels in (((13, 11), 2), ((9, 10, 7), 2), ((5, 6, 5, 4), 1)):
        stride = 1 << levels
        pr = ChunkedGNNPredictor(path, 0.0, 1.0, levels=levels, anchor_stride=stride)
        pr.begin(shape, tuple(-(-n // stride) * stride for n in shape))
        w = pr.start_wave([0], torch.zeros((1, *shape), device="cuda"))
        geoms, dirs = G.schedule_geoms(shape, levels, stride, 1, stride, 2)
        plan = w["plan"]
        gi = 0
        for s in range(plan.n_stages):
            g = w["geoms"][s]
            if not g.M:
                continue
            ref = geoms[gi]
            gi += 1
            L = len(dirs)
            steps = g.steps[: g.M * L].cpu().numpy().reshape(g.M, L)
            assert np.array_equal(g.qidx[: g.M].cpu().numpy(), ref["q"])
            assert np.array_equal(steps & 0xFFFF, ref["sp"].T), (shape, s)
            assert np.array_equal(steps >> 16, ref["sn"].T), (shape, s)
        assert gi == len(geoms)


@pytest.mark.parametrize("numerics", ["bf16", "fp32"])
@pytest.mark.parametrize("shape", [(64, 64), (24, 20, 28), (16, 16, 16, 16)])
def test_codec_roundtrip(tmp_path, shape, numerics):
    from latec_amd.gnn_codec import GNNCodec, roundtrip_slack

    path = _ckpt(tmp_path, 32, 2, 5)
    x = _field(shape, 2) * 7.0 - 3.0
    codec = GNNCodec(path, numerics=numerics)
    eb = 1e-3
    stream = codec.compress(x, eb)
    y = codec.uncompress(stream).numpy()
    assert y.shape == x.shape and y.dtype == x.dtype
    S = float(x.max() - x.min())
    assert np.abs(y - x).max() <= eb * S + roundtrip_slack(x.min(), x.max())
    assert codec.compress(x, eb) == stream  # deterministic


def test_codec_options(tmp_path):
    """Gate off / fine gate / tune size / chunked / integer input / ablation."""
    from latec_amd.gnn_codec import GNNCodec, roundtrip_slack

    path = _ckpt(tmp_path, 32, 2, 7)
    x = _field((40, 48), 3)
    codec = GNNCodec(path)
    S = float(x.max() - x.min())
    tol = 2e-3 * S + roundtrip_slack(x.min(), x.max())
    for kw in ({"gate": False}, {"gate_fine": True}, {"tune": "size"},
               {"chunk_size": 16}, {"predictor": "interp-cubic"}, {"levels": 2}):
        stream = codec.compress(x, 2e-3, **kw)
        y = codec.uncompress(stream).numpy()
        assert np.abs(y - x).max() <= tol, kw
    xi = (x * 255).astype(np.uint8)
    y = codec.uncompress(codec.compress(xi, 1e-2)).numpy()
    assert y.dtype == np.uint8
    assert np.abs(y.astype(np.int32) - xi).max() <= np.ceil(1e-2 * (int(xi.max()) - int(xi.min())))
    x64 = x.astype(np.float64)
    y = codec.uncompress(codec.compress(x64, 1e-3)).numpy()
    assert y.dtype == np.float64


def test_codec_rejects_foreign_streams(tmp_path):
    from latec_amd.gnn_codec import GNNCodec

    a = GNNCodec(_ckpt(tmp_path, 32, 2, 1, "a.pt"))
    b = GNNCodec(_ckpt(tmp_path, 32, 2, 2, "b.pt"))
    x = _field((32, 32), 4)
    s = a.compress(x, 1e-2)
    with pytest.raises(ValueError, match="checkpoint hash differs"):
        b.uncompress(s)
    with pytest.raises(ValueError, match="bad magic"):
        a.uncompress(b"XXXXXXXX" + s[8:])


def _chunk_frame_masks(pr, cshape, origin, pad, stage_tile, s):
    """(known, usable) over chunk ci's padded frame, restated from the
    reference's chunk geometry: known = the cell's stage in the schedule's
    periodic pattern is before s (gnn_predictor.py:1011-1060, the pattern of
    _period_prefixes); usable = interior, or a halo cell inside the tensor in
    a coded chunk or on the anchor lattice (:1241-1275, 1338-1356)."""
    nd = len(cshape)
    fshape = tuple(n + 2 * pad for n in cshape)
    idx = np.indices(fshape).reshape(nd, -1).T - pad
    P = stage_tile.shape[0]
    known = stage_tile[tuple((idx % P).T)] < s
    inside = np.all((idx >= 0) & (idx < np.asarray(cshape)), axis=1)
    g = idx + np.asarray(origin)
    intensor = np.all((g >= 0) & (g < np.asarray(pr.shape)), axis=1)
    halo = ~inside & intensor
    cid = np.zeros(len(g), np.int64)
    cid[halo] = np.ravel_multi_index(tuple((g[halo] // np.asarray(pr.edges)).T), pr.grid)
    usable = inside | (halo & (pr.coded[cid] | np.all(g % pr.anchor_stride == 0, axis=1)))
    return known.reshape(fshape), usable.reshape(fshape)


def _stage_tile(ndim, levels, stride):
    """Stage index of every residue mod the stride (the schedule's period tile),
    from the oracle's masks on a grid two periods wide."""
    from oracle import oracle as O

    tile = np.full((2 * stride,) * ndim, -1, np.int64)
    for s, (mask, _, _, _) in enumerate(O.stage_plan(tile.shape, levels, stride, 1)):
        tile[mask] = s
    assert (tile >= 0).all()
    return tile[(slice(0, stride),) * ndim]


@pytest.mark.parametrize("shape,levels,edges", [
    ((37, 41), 2, (16, 12)),
    ((21, 18, 17), 2, (8, 8, 8)),
    ((9, 10, 8, 9), 1, (4, 4, 4, 4)),
])
def test_multichunk_halo_geometry_matches_scan(tmp_path, shape, levels, edges, geo_table):
    """Every chunk of an explicit chunk grid (extended blocks, halo frames):
    the kernel's frame indices, tensor indices and steps (lookup tables or
    cell scan) equal an exhaustive scan of the padded frame with the earlier
    chunks coded."""
    from latec_amd.gnn_predictor import ChunkedGNNPredictor
    from oracle import gnn_oracle as G
    from oracle import oracle as O

    path = _ckpt(tmp_path, 32, 2, 0)
    stride = 1 << levels
    pr = ChunkedGNNPredictor(path, 0.0, 1.0, levels=levels, anchor_stride=stride)
    pr.begin(shape, edges)
    assert pr.n_chunks > 1
    dirs = G.half_directions(len(shape), 2)
    tile = _stage_tile(len(shape), levels, stride)
    recon = torch.zeros((1, *shape), device="cuda")
    for ci in range(pr.n_chunks):
        pr.coded[:] = False
        pr.coded[:ci] = True
        pr._coded_dev.zero_()
        pr._coded_dev[:ci] = 1
        sls = pr.chunk_slices(ci)
        origin = tuple(sl.start for sl in sls)
        cshape = tuple(sl.stop - sl.start for sl in sls)
        pad = stride
        w = pr.start_wave([ci], recon)
        fshape = tuple(n + 2 * pad for n in cshape)
        fstr = np.cumprod((1,) + fshape[:0:-1])[::-1]
        for s, (mask, _, _, _) in enumerate(O.stage_plan(cshape, levels, stride, 1)):
            g = w["geoms"][s]
            assert g.M == int(mask.sum()), (ci, s)
            if g.M:
                coords = np.stack(np.nonzero(mask), 1)
                known, usable = _chunk_frame_masks(pr, cshape, origin, pad, tile, s)
                sp, sn = G.frame_scan_steps(known, usable, coords + pad, dirs, stride)
                steps = g.steps[: g.M * len(dirs)].cpu().numpy().reshape(g.M, len(dirs))
                assert np.array_equal(g.qidx[: g.M].cpu().numpy(), (coords + pad) @ fstr), (ci, s)
                gref = np.ravel_multi_index(tuple((coords + np.asarray(origin)).T), shape)
                assert np.array_equal(g.gidx[: g.M].cpu().numpy(), gref), (ci, s)
                assert np.array_equal(steps & 0xFFFF, sp.T), (shape, ci, s)
                assert np.array_equal(steps >> 16, sn.T), (shape, ci, s)
        pr._wave = None


def _host_halo_cells(pr, origin, cshape, pad):
    """Frame rows and tensor indices of the usable halo cells, enumerated on
    the host: outside the chunk, inside the tensor, in a coded chunk or on
    the anchor lattice (gnn_predictor.py:1338-1356)."""
    nd = len(pr.shape)
    fshape = tuple(n + 2 * pad for n in cshape)
    idx = np.indices(fshape).reshape(nd, -1).T - pad
    inside = np.all((idx >= 0) & (idx < np.asarray(cshape)), axis=1)
    g = idx + np.asarray(origin)
    intensor = np.all((g >= 0) & (g < np.asarray(pr.shape)), axis=1)
    keep = ~inside & intensor
    idx, g = idx[keep], g[keep]
    cid = np.ravel_multi_index(tuple((g // np.asarray(pr.edges)).T), pr.grid)
    ok = pr.coded[cid] | np.all(g % pr.anchor_stride == 0, axis=1)
    idx, g = idx[ok], g[ok]
    return (np.ravel_multi_index(tuple((idx + pad).T), fshape).astype(np.int64),
            np.ravel_multi_index(tuple(g.T), pr.shape).astype(np.int64))


@pytest.mark.parametrize("numerics", ["bf16", "fp32"])
@pytest.mark.parametrize("shape,levels,edges", [
    ((37, 41), 2, (16, 12)),
    ((21, 18, 17), 2, (8, 8, 8)),
    ((9, 10, 8, 9), 1, (4, 4, 4, 4)),
])
def test_halo_frame_equals_listed_halo_rows(tmp_path, shape, levels, edges, numerics):
    """The device halo fill of every chunk frame equals the halo rows of the
    host-enumerated usable cells written through latec_gnn_halo, and leaves
    every other row zero."""
    from latec_amd._lib import check, lib, ptr
    from latec_amd.gnn_predictor import ChunkedGNNPredictor

    path = _ckpt(tmp_path, 32, 2, 0)
    stride = 1 << levels
    pr = ChunkedGNNPredictor(path, 0.0, 1.0, levels=levels, anchor_stride=stride,
                             numerics=numerics)
    pr.begin(shape, edges)
    recon = torch.from_numpy(_field(shape, 4)).cuda()[None]
    m = pr.model
    for ci in range(pr.n_chunks):
        pr.coded[:] = False
        pr.coded[:ci] = True
        pr._coded_dev.zero_()
        pr._coded_dev[:ci] = 1
        sls = pr.chunk_slices(ci)
        origin = tuple(sl.start for sl in sls)
        cshape = tuple(sl.stop - sl.start for sl in sls)
        w = pr.start_wave([ci], recon)
        ref = torch.zeros_like(w["E"])
        frows, gflat = _host_halo_cells(pr, origin, cshape, stride)
        if len(frows):
            dr, dg = torch.from_numpy(frows).cuda(), torch.from_numpy(gflat).cuda()
            check(lib().latec_gnn_halo(m.handle, ptr(ref), ptr(dr), ptr(dg), ptr(recon),
                                       len(frows), 0.0, 1.0, ptr(None)))
        torch.cuda.synchronize()
        assert torch.equal(w["E"], ref), (shape, ci)
        pr._wave = None


def _gate_costs_flat(W):
    """The selection's candidate costs (every (row, T) split of the 12
    candidate rows) followed by the no-gate total, as _gate_select ranks them."""
    cums = np.cumsum(W, axis=-1)
    base, ci = cums[0], cums[1:]
    tot = base[-1]
    low = tot - base[:-1] + ci[:, :-1]
    high = base[:-1] + (ci[:, -1:] - ci[:, :-1])
    return np.concatenate([np.concatenate([low, high]).reshape(-1), [tot]])


@pytest.mark.parametrize("shape", [(40, 36, 44), (16, 18, 16, 14)])
def test_gate_costs_match_oracle_on_real_stages(monkeypatch, shape):
    """Every gate evaluation of a real encode (committed checkpoint): the
    device per-bucket bit sums equal the oracle's gate_costs on the same
    stage values, model prediction, scale and interpolation candidates, and
    the gate chosen from them is the one chosen from the oracle's sums."""
    import latec_amd.gnn_codec as GC
    from latec_amd.gnn_codec import GNNCodec
    from oracle import gnn_oracle as G

    seen = []
    real = GC.gate_costs_device

    def spy(model, x, pred, scale, cands, n, eb, radius, tabs):
        W = real(model, x, pred, scale, cands, n, eb, radius, tabs)
        seen.append((x[:n].cpu().numpy(), pred[:n].cpu().numpy(), scale[:n].cpu().numpy(),
                     [c[:n].cpu().numpy() for c in cands], eb, radius, W.copy()))
        return W

    monkeypatch.setattr(GC, "gate_costs_device", spy)
    ckpt = GC.__file__.replace("gnn_codec.py", "checkpoints/gnn_d64_agg2.pt")
    x = _field(shape, 6)
    codec = GNNCodec(ckpt)
    # the Python stage loop reads the sums back through gate_costs_device (the
    # native loop runs the same kernels and codes the same bytes,
    # test_native_stage_loop_equals_python_stage_loop)
    codec.native_stages = False
    stream = codec.compress(x, 1e-3)
    assert len(seen) >= 3
    for i, (xs, p, sc, ips, eb, radius, W) in enumerate(seen):
        ref = G.gate_costs(xs, p, sc, ips, eb, radius)
        np.testing.assert_allclose(W, ref, rtol=1e-9, atol=1e-6, err_msg=f"gate {i}")
        if GC._gate_select(W) != GC._gate_select(ref):
            # a near-tie: the device's choice must cost the same under the
            # oracle's own sums, to the sums' rounding
            cw, cr = _gate_costs_flat(W), _gate_costs_flat(ref)
            fw, fr = int(np.argmin(cw)), int(np.argmin(cr))
            assert abs(cr[fw] - cr[fr]) <= 1e-6 + 1e-9 * abs(cr[fr]), (i, cr[fw], cr[fr])
    y = codec.uncompress(stream).numpy()
    assert np.abs(y - x).max() <= 1e-3 * float(x.max() - x.min()) + GC.roundtrip_slack(
        x.min(), x.max())


@pytest.mark.parametrize("numerics", ["bf16", "fp32"])
def test_row_launch_shapes_give_identical_streams(tmp_path, numerics):
    """Every launch shape of the message / mix rows (knob gnn_rows: 16-row
    tiles per wave group, waves per workgroup), both line-pool kernels (knob
    gnn_pool: line by line, loads issued ahead for 8 or 4 lines) and the
    geometry with or without its lookup tables (knob gnn_geo_table) code the
    same stream and decode it to the same values: a row's result does not
    depend on the tile it shares, both pools add the lines in the same order
    and both geometry paths find the same cells."""
    from latec_amd import _lib
    from latec_amd.gnn_codec import GNNCodec

    ckpt = _ckpt(tmp_path, 64, 2, 5)
    x = _field((18, 20, 17, 19), 9)
    codec = GNNCodec(ckpt, numerics=numerics)
    L = _lib.lib()
    knobs = (b"gnn_rows", b"gnn_pool", b"gnn_geo_table")
    base = [L.latec_tune_get(k) for k in knobs]
    streams, decoded = [], []
    try:
        for vals in [(v, 2, 1) for v in range(8)] + [(0, 0, 1), (0, 1, 1), (3, 2, 0)]:
            for k, v in zip(knobs, vals):
                assert L.latec_tune_set(k, v) == 0
            streams.append(codec.compress(x, 1e-3))
            decoded.append(codec.uncompress(streams[0]).numpy())
    finally:
        for k, v in zip(knobs, base):
            L.latec_tune_set(k, v)
    assert all(s == streams[0] for s in streams[1:])
    assert all(np.array_equal(d, decoded[0]) for d in decoded[1:])
    assert L.latec_tune_set(b"gnn_rows", 8) != 0
    assert L.latec_tune_set(b"gnn_pool", 3) != 0
    assert L.latec_tune_set(b"gnn_geo_table", 2) != 0


@pytest.mark.parametrize("case", [
    dict(shape=(18, 20, 17, 19), levels=2, chunk=None, fine=False, pred="gnn", num="bf16",
         radius=1 << 15),
    dict(shape=(40, 36, 44), levels=2, chunk=16, fine=True, pred="gnn", num="bf16", radius=8),
    dict(shape=(64, 60), levels=3, chunk=None, fine=False, pred="gnn", num="fp32", radius=16),
    dict(shape=(24, 20, 28), levels=2, chunk=8, fine=False, pred="interp-cubic", num="bf16",
         radius=8),
    dict(shape=(30, 26), levels=2, chunk=8, fine=False, pred="gnn", num="bf16", radius=1 << 15,
         gate=False),
], ids=["4d", "3d-chunks-fine-r8", "2d-fp32-r16", "3d-interp-r8", "2d-chunks-nogate"])
def test_native_stage_loop_equals_python_stage_loop(tmp_path, case):
    """latec_gnn_code_stage (gate, quantizer, reconstruction, levels and the
    outliers by position in one library call per stage) codes the same bytes
    as the same steps called one by one from Python; small radii make many
    outliers, explicit chunk sizes take the chunk's column split."""
    from latec_amd import gnn_codec as GC
    from latec_amd.gnn_codec import GNNCodec

    ckpt = _ckpt(tmp_path, 32, 2, 4)
    x = _field(case["shape"], 6)
    codec = GNNCodec(ckpt, numerics=case["num"])
    kw = dict(levels=case["levels"], chunk_size=case["chunk"], gate_fine=case["fine"],
              predictor=case["pred"], radius=case["radius"], gate=case.get("gate", True))
    codec.native_stages = False
    ref = codec.compress(x, 1e-3, **kw)
    ref_stats = dict(codec.last_stats)
    codec.native_stages = True
    got = codec.compress(x, 1e-3, **kw)
    assert got == ref
    assert codec.last_stats["outliers"] == ref_stats["outliers"]
    assert codec.last_stats["stage_codes"] == ref_stats["stage_codes"]
    if case["radius"] < 64:
        assert ref_stats["outliers"] > 2
    y = codec.uncompress(got).numpy()
    assert np.abs(y - x).max() <= 1e-3 * float(x.max() - x.min()) + GC.roundtrip_slack(
        x.min(), x.max())
    # the decoder's latec_gnn_decode_stage_begin / _end against the same
    # steps called one by one from Python
    rec_native = codec.last_decoded.clone()
    codec.native_stages = False
    y_ref = codec.uncompress(got).numpy()
    codec.native_stages = True
    assert np.array_equal(y.view(np.uint32), y_ref.view(np.uint32))
    assert torch.equal(rec_native, codec.last_decoded)


@pytest.mark.gpu
def test_native_decode_stage_rejects_outlier_count_mismatch(tmp_path, monkeypatch):
    """A stream whose outlier list is one value short fails with the
    outlier count error of latec_gnn_decode_stage_end (a host check before
    any copy or kernel); the next decode of a good stream still succeeds."""
    from latec_amd import gnn_codec as GC
    from latec_amd.gnn_codec import GNNCodec

    ckpt = _ckpt(tmp_path, 32, 2, 4)
    x = _field((40, 36), 6)
    codec = GNNCodec(ckpt, numerics="fp32")
    blob = codec.compress(x, 1e-3, levels=3, radius=8)
    assert codec.last_stats["outliers"] > 2
    good = codec.uncompress(blob).numpy()
    real = GC.unpack_stage

    def short(*a, **k):
        codes, outl, off = real(*a, **k)
        return codes, (outl[:-1] if len(outl) else outl), off

    monkeypatch.setattr(GC, "unpack_stage", short)
    with pytest.raises(ValueError, match="outlier count mismatch"):
        codec.uncompress(blob)
    monkeypatch.setattr(GC, "unpack_stage", real)
    again = codec.uncompress(blob).numpy()
    assert np.array_equal(good.view(np.uint32), again.view(np.uint32))


@pytest.mark.gpu
def test_stream_ptr_is_torch_current_stream():
    """The library calls' stream is torch's current stream, also inside a
    `torch.cuda.stream` block."""
    from latec_amd._lib import stream_ptr

    torch.cuda.init()
    assert (stream_ptr().value or 0) == torch.cuda.current_stream().cuda_stream
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        assert (stream_ptr().value or 0) == s.cuda_stream != 0
    assert (stream_ptr().value or 0) == torch.cuda.current_stream().cuda_stream
