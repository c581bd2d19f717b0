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


def test_geometry_matches_scan(tmp_path):
    """Stage geometry of the kernel == the oracle's exhaustive scan."""
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
