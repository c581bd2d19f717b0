"""Batched uneven workload: one persistent launch per schedule step over many
fields; every field's stream must equal its own single-field ``compress``."""

import numpy as np
import pytest

from tests.conftest import has_gpu


def test_batch_fields_generator_mix():
    from latec_amd.synthetic import batch_fields

    fs = batch_fields(n=12, seed=4, scale=0.1)
    assert len(fs) == 12
    assert [f.ndim for f in fs] == [2, 3] * 6
    assert {f.dtype for f in fs} <= {np.dtype(np.float32), np.dtype(np.float64)}
    for f in fs:
        assert f.min() >= 0 and f.max() <= 1 and f.max() > f.min()
    again = batch_fields(n=12, seed=4, scale=0.1)
    assert all(np.array_equal(a, b) for a, b in zip(fs, again))


def _mixed(seed):
    rng = np.random.RandomState(seed)
    out = []
    for i, shape in enumerate([(70, 90), (33, 40, 29), (128, 65), (17, 18, 19), (40, 41, 42),
                               (200, 31), (9, 12, 10, 11), (64, 64, 64)]):
        x = np.cumsum(np.cumsum(rng.randn(*shape), 0), -1)
        x += (0.3 if i % 3 == 0 else 0.01) * rng.randn(*shape)
        x = (x - x.min()) / (x.max() - x.min())
        out.append(x.astype(np.float64 if i % 4 == 1 else np.float32))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("version", [5, 4])
def test_gpu_batch_streams_equal_single_field_streams(version):
    import latec_amd
    from latec_amd.batch import compress_batch, decompress_batch
    from latec_amd.predictor import InterpPredictor

    fields = _mixed(3)
    print("case batch v%d, %d fields" % (version, len(fields)), flush=True)
    streams = compress_batch(fields, 1e-3, levels=3, anchor_stride=8, stream_version=version)
    for x, s in zip(fields, streams):
        x32 = x.astype(np.float32)
        vmin, vmax = float(x32.min()), float(x32.max())
        eb = 1e-3 * (vmax - vmin) if vmax > vmin else 1e-3
        ref, _ = latec_amd.compress(x32, eb, InterpPredictor("cubic", 3, 8, 1), levels=3,
                                    anchor_stride=8, anchor_block=1, stream_version=version)
        assert s == ref
    if version == 5:
        outs = decompress_batch(streams)
        for x, y, s in zip(fields, outs, streams):
            assert np.array_equal(y.view(np.uint32), latec_amd.decompress(s).view(np.uint32))


@pytest.mark.gpu
def test_gpu_batch_graph_and_direct_launches_agree():
    from latec_amd.batch import compress_batch

    fields = _mixed(5)[:5]
    a = compress_batch(fields, 1e-3, levels=3, anchor_stride=8, use_graph=True)
    b = compress_batch(fields, 1e-3, levels=3, anchor_stride=8, use_graph=False)
    assert a == b


@pytest.mark.gpu
def test_gpu_batch_gnn_selection_keeps_bound_and_smaller_stream():
    """Per-field GNN-vs-interpolation choice in the batch: every kept stream is
    the smaller of the field's interpolation stream and its LATECGNN stream
    (whose stages each pick GNN or interpolation), and the batch decodes
    within the bound."""
    import os

    from latec_amd.batch import compress_batch, decompress_batch
    from latec_amd.gnn_codec import GNNCodec, roundtrip_slack

    ckpt = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "latec_amd", "checkpoints", "gnn_d64_agg2.pt")
    codec = GNNCodec(ckpt)
    rng = np.random.RandomState(12)
    fields = []
    for i, shape in enumerate([(96, 80), (24, 26, 28), (12, 12, 12, 12), (130, 64)]):
        x = np.cumsum(np.cumsum(rng.randn(*shape), 0), -1)
        if i == 2:  # a discontinuity: the learned predictor's kind of field
            x[..., shape[-1] // 2:] += 5.0 * x.std()
        x += 0.01 * rng.randn(*shape)
        fields.append(((x - x.min()) / (x.max() - x.min())).astype(np.float32))
    print("case batch gnn selection", flush=True)
    plain = compress_batch(fields, 1e-3)
    timings = {}
    mixed = compress_batch(fields, 1e-3, gnn=codec, timings=timings)
    assert timings["gnn_streams"] == sum(m != p for m, p in zip(mixed, plain))
    for f, (m, p) in enumerate(zip(mixed, plain)):
        if m != p:
            assert m[:8] == b"LATECGNN" and len(m) < len(p)
            assert m == codec.compress(fields[f], 1e-3)
    out = decompress_batch(mixed, gnn=codec)
    for x, y, m in zip(fields, out, mixed):
        eb = 1e-3 * float(x.max() - x.min())
        # a LATECGNN stream holds the bound up to its float32 normalization slack
        slack = roundtrip_slack(float(x.min()), float(x.max())) if m[:8] == b"LATECGNN" else 0.0
        assert np.abs(np.asarray(y, np.float64) - x).max() <= eb + slack
    with pytest.raises(ValueError):
        decompress_batch(mixed if timings["gnn_streams"] else [b"LATECGNN" + mixed[0][8:]])
