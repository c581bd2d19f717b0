"""The benchmark configurations at their own sizes on the GPU.

C3: the GNN codec on the 64^4 anisotropic field with the committed d64 /
agg-2 checkpoint at levels 3 (one chunk, and the reference's partition into
16 chunks of 32^4 with halo context).  C4: the batched workload of 256
uneven fields.  Each against the properties that hold at full size (the
bound, encoder == decoder, determinism, equality with the single-field
path) and, where it runs in seconds, against the oracle.
"""

import os

import numpy as np
import pytest

from tests.conftest import has_gpu

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CKPT = os.path.join(ROOT, "latec_amd", "checkpoints", "gnn_d64_agg2.pt")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def c3_field():
    from latec_amd.synthetic import config3_field

    return config3_field(3, (64, 64, 64, 64), device="cuda").cpu().numpy()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("chunk_size", [None, 32], ids=["one-chunk", "16-chunks"])
def test_c3_gnn_bound_identity_and_gain(c3_field, chunk_size):
    """64^4, levels 3, eb = 1e-3 * range, BF16 kernels: every value within
    the bound, the decoder's reconstruction equal to the encoder's bit for
    bit, the stream deterministic, and fewer bits per value than plain cubic
    interpolation of the same field at the same absolute bound."""
    import torch

    import latec_amd
    from latec_amd import InterpPredictor
    from latec_amd.gnn_codec import GNNCodec, roundtrip_slack

    x = c3_field
    print("case C3", chunk_size, flush=True)
    codec = GNNCodec(CKPT)
    s1 = codec.compress(x, 1e-3, levels=3, chunk_size=chunk_size)
    enc = codec.last_stats["recon"].clone()
    y = codec.uncompress(s1).numpy()
    assert torch.equal(codec.last_decoded, enc)
    S = float(x.max() - x.min())
    assert float(np.abs(y - x).max()) <= 1e-3 * S + roundtrip_slack(x.min(), x.max())
    assert codec.compress(x, 1e-3, levels=3, chunk_size=chunk_size) == s1
    s_int, _ = latec_amd.compress(x, 1e-3 * S, InterpPredictor("cubic", 3, 8, 1), levels=3,
                                  anchor_stride=8, anchor_block=1)
    bpv, bpv_int = 8 * len(s1) / x.size, 8 * len(s_int) / x.size
    print(f"C3 {chunk_size}: GNN {bpv:.4f} bits/value, interpolation {bpv_int:.4f}")
    assert bpv < bpv_int


def test_c3_fp32_forward_matches_fp64_oracle_on_real_schedule(c3_field):
    """The committed checkpoint's FP32 kernels against the FP64 oracle over
    the real levels-3 schedule on a 16^4 crop of the C3 field (teacher
    forced: every stage sees the true values of the earlier ones)."""
    import torch

    from latec_amd import gnn_model as GM
    from latec_amd.gnn_predictor import ChunkedGNNPredictor
    from oracle import gnn_oracle as G

    x = np.ascontiguousarray(c3_field[8:24, 16:32, 24:40, 0:16])
    x = ((x - x.min()) / (x.max() - x.min())).astype(np.float32)
    shape, levels = x.shape, 3
    stride = 1 << levels
    pr = ChunkedGNNPredictor(CKPT, 0.0, 1.0, levels=levels, anchor_stride=stride,
                             numerics="fp32")
    pr.begin(shape, shape)
    recon = torch.from_numpy(x).cuda()[None]
    w = pr.start_wave([0], recon)
    eb = 1e-3
    got = [pr.predict_wave_stage(s, recon, eb, want_head=True)[2:] for s in w["chain"][1:]]
    pr.finish_wave(recon)
    d, agg, P = GM.load_checkpoint(CKPT)
    Pd = {k: torch.from_numpy(np.asarray(v, np.float64)) for k, v in P.items()}
    ref = G.run_schedule(Pd, torch.from_numpy(x.reshape(1, -1)).double(), shape, levels,
                         stride, eb, stride, agg)
    assert len(ref) == len(got) == 21
    for i, ((mu, lb), (rm, rl)) in enumerate(zip(got, ref)):
        np.testing.assert_allclose(mu.cpu().numpy(), rm.numpy()[0], atol=5e-5, rtol=0,
                                   err_msg=f"mu stage {i}")
        np.testing.assert_allclose(lb.cpu().numpy(), rl.numpy()[0], atol=5e-4, rtol=0,
                                   err_msg=f"log_b stage {i}")


@pytest.mark.timeout(900)
def test_c4_256_uneven_fields():
    """The 256-field workload (half 2-D sides U[64,1024], half 3-D sides
    U[32,256], 70 % float32 / 30 % float64, smooth / turbulent /
    discontinuous, relative eb 1e-3, cubic L4 stride 16, version 5): every
    batched stream equals its single-field stream, the smallest eight equal
    the oracle's streams, and the batched decode returns every field within
    its bound (float64 fields within it after their float32 cast)."""
    import latec_amd
    from latec_amd import InterpPredictor
    from latec_amd.batch import compress_batch, decompress_batch
    from latec_amd.synthetic import batch_fields
    from oracle import oracle as O

    fields = batch_fields(n=256, seed=4, device="cuda")
    fields = [f.cpu().numpy() if hasattr(f, "cpu") else f for f in fields]
    assert len(fields) == 256 and sum(f.ndim == 3 for f in fields) == 128
    print("case C4 batch", sum(f.size for f in fields), "values", flush=True)
    streams = compress_batch(fields, 1e-3, levels=4, anchor_stride=16, stream_version=5)
    order = np.argsort([f.size for f in fields])
    for i, (x, s) in enumerate(zip(fields, streams)):
        x32 = x.astype(np.float32)
        vmin, vmax = float(x32.min()), float(x32.max())
        eb = 1e-3 * (vmax - vmin)
        ref, _ = latec_amd.compress(x32, eb, InterpPredictor("cubic", 4, 16, 1), levels=4,
                                    anchor_stride=16, anchor_block=1, stream_version=5)
        assert s == ref, i
        if i in order[:8]:
            assert s == O.compress_v5(x32, eb, "cubic", 4, 16, 1, None, 1)[0], i
    outs = decompress_batch(streams)
    for i, (x, y) in enumerate(zip(fields, outs)):
        x32 = x.astype(np.float32)
        eb = 1e-3 * float(x32.max() - x32.min())
        assert float(np.abs(np.asarray(y, np.float32) - x32).max()) <= eb, i
