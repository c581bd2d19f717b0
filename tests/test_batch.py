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
