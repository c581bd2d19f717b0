"""GPU parity of the plain interpolation path against the oracle.

Bit-exact bar: codes, outliers, reconstruction and the whole LATEC001
stream must equal the oracle's on the same input; the GPU-built tables and
fits must equal the golden vectors of the reference's rans.py.
"""

import ctypes
import hashlib
import json
import os
import struct

import numpy as np
import pytest

from tests.conftest import has_gpu

pytestmark = pytest.mark.gpu

if has_gpu():
    import torch

    import latec_amd
    from latec_amd import _device as D
    from latec_amd import codec, rans
    from latec_amd._lib import Fit, check, lib, ptr
    from latec_amd.predictor import InterpPredictor, get_plan
    from latec_amd.quantizer import opts_for
from oracle import oracle as O

G = os.path.join(os.path.dirname(__file__), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def smooth(shape, seed=0, rough=0.0):
    rng = np.random.RandomState(seed)
    x = rng.randn(*shape)
    for ax in range(len(shape)):
        x = np.cumsum(x, axis=ax)
    x = x / np.abs(x).max()
    if rough:
        x = x + rough * rng.randn(*shape)
    return x.astype(np.float32)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip("no GPU")


def test_gpu_tables_match_reference_golden():
    gold = json.load(open(os.path.join(G, "rans_golden_libm.json")))
    for key, v in gold["tables"].items():
        if key.startswith("p"):
            continue
        k, s, o = map(int, key.split(","))
        if k <= 10 and s % 4 != 1 and o != 3:
            continue  # a representative subset of the small windows
        assert sha(rans.build_cdfs(k, s, 24, o)) == v["cdf"], key
        assert sha(rans._neg_log2_weights(k, s, 64, 24, o)) == v["nl2"], key
    for k, h in gold["level_weights_f64"].items():
        assert sha(rans._level_weights(int(k))) == h, k


def test_gpu_fit_matches_reference_golden():
    gold = json.load(open(os.path.join(G, "rans_golden_libm.json")))
    fv = np.load(os.path.join(G, "fit_vectors.npz"))
    R = 1 << 15
    for f in gold["const_fits"]:
        tiles = fv[f"tiles{f['i']}"]
        kexp = f["kexp"]
        parts = [rans._widen(np.repeat(np.arange(t.size), t), R, kexp) for t in tiles]
        codes = np.concatenate(parts).astype(np.uint16)
        assert len(codes) == f["n"]
        fit = rans.fit_codes_device(D.to_device(codes), len(codes), R, 2)
        assert (fit.kexp, fit.level, fit.shape, fit.owx) == \
            (kexp, f["level"], f["shape"], f["owx"]), f
        assert abs(fit.bits - f["bits"]) <= 3e-7 * max(1.0, f["bits"]), f


def test_gpu_scale_to_level_exact():
    """Device levels equal the host NumPy scale_to_level on every value, with
    no tolerance: on the golden scales (and the golden levels themselves where
    this machine's NumPy float32 log rounds as the machine that wrote them
    did), and on every threshold and its float32 neighbours."""
    fv = np.load(os.path.join(G, "fit_vectors.npz"))
    rng = np.random.RandomState(5)
    for i, eb in enumerate(fv["scale_ebs"]):
        eb = float(eb)
        host = rans.scale_to_level_host(fv["scales"], eb)
        assert np.array_equal(rans.scale_to_level(fv["scales"], eb), host)
        if np.array_equal(host, fv["scale_levels"][i]):
            assert np.array_equal(rans.scale_to_level(fv["scales"], eb), fv["scale_levels"][i])
        t = rans.scale_thresholds(eb)
        v = np.concatenate([t, np.nextafter(t, 0, dtype=np.float32),
                            np.nextafter(t, np.inf, dtype=np.float32),
                            (eb * np.exp(rng.uniform(-4, 9.5, 100000))).astype(np.float32),
                            np.array([0, -1, np.inf, eb / 16, eb * 4096], np.float32)])
        assert np.array_equal(rans.scale_to_level(v, eb), rans.scale_to_level_host(v, eb))


def _oracle_stage_codes(payload, counts, C):
    out, off = [], 0
    for n in counts:
        c, o, off = O.unpack_stage(payload, off, C * n)
        out.append((c, o))
    return out


CASES = [
    # shape, C, levels, stride, block, order, center, end, eb_rel, rough
    ((64, 48), 1, 4, 16, 4, "cubic", 0, 1, 1e-3, 0.0),
    ((64, 48), 3, 4, 16, 4, "cubic", 0, 1, 1e-2, 0.01),
    ((70, 90), 2, 3, 8, 1, "linear", 0, 0, 3e-3, 0.0),
    ((40, 33, 17), 1, 3, 8, 1, "cubic", 0, 3, 1e-3, 0.02),
    ((40, 33, 17), 1, 3, 8, 1, "cubic", 1, 2, 1e-3, 0.0),
    ((40, 33, 17), 1, 3, 8, 1, "linear", 2, 1, 1e-3, 0.0),
    ((31, 29, 27), 1, 4, 16, 2, "cubic", 0, 1, 5e-4, 0.05),
    # large enough for the level-fused slab kernel, with partial tiles and segments
    ((70, 45, 133), 1, 3, 8, 1, "cubic", 0, 1, 1e-3, 0.05),
    ((67, 33, 129), 1, 3, 8, 1, "linear", 2, 3, 1e-3, 0.02),
    # the line lengths of the 512^3 schedule (cubic L5 stride 32): level grids
    # of 64, 128, 256 and 512 points along the last axis, so the finest level
    # has 256 pairs (one line per workgroup) and 128 quads per line
    ((33, 40, 512), 1, 5, 32, 1, "cubic", 0, 1, 1e-3, 0.02),
    # 257 pairs at the finest level: lines longer than a workgroup (wave runs)
    ((20, 18, 513), 1, 5, 32, 1, "cubic", 0, 1, 1e-3, 0.02),
    # line quads with 8 and 16 quads per line, and with 1 and 2
    ((36, 30, 64), 1, 3, 8, 1, "cubic", 0, 3, 1e-3, 0.05),
    ((40, 36, 8), 1, 3, 8, 1, "linear", 2, 1, 1e-3, 0.02),
    ((12, 11, 10, 9), 1, 2, 4, 1, "cubic", 1, 1, 1e-3, 0.0),
    ((9, 8, 7, 6, 5), 1, 2, 4, 1, "cubic", 0, 3, 1e-2, 0.0),
    ((300,), 1, 5, 32, 1, "cubic", 0, 1, 1e-4, 0.001),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-C{c[1]}-b{c[4]}-{c[5]}-c{c[6]}-e{c[7]}")
def test_gpu_encode_matches_oracle_bit_exact(case):
    shape, C, levels, stride, block, order, center, end, rel, rough = case
    print("case", case, flush=True)
    x = np.stack([smooth(shape, seed=c, rough=rough) for c in range(C)])
    eb = rel * float(x.max() - x.min())
    ebs = O.stage_ebs_ref(len(shape), levels, stride, eb, 0.9)
    ref_payload, ref_recon, _ = O.compress_payload(x, ebs, levels, stride, block, order, center,
                                                   end)
    plan = get_plan(shape, C, levels, stride, block)
    o = opts_for(1 << 15, cubic=order == "cubic", center=center, end_mode=end)
    res = codec.encode_device(D.to_device(x.reshape(-1)), plan, ebs, o)
    torch.cuda.synchronize()
    recon = D.to_host(res.d_recon).reshape(x.shape)
    assert np.array_equal(recon.view(np.uint32), ref_recon.view(np.uint32))
    ref = _oracle_stage_codes(ref_payload, plan.counts, C)
    codes = D.to_host(res.d_codes, np.uint16)
    for s in range(plan.n_stages):
        lo, n = C * plan.offsets[s], C * plan.counts[s]
        assert np.array_equal(codes[lo:lo + n].astype(np.uint32), ref[s][0]), f"stage {s}"
    fits = codec.fit_device(res)
    payload, _, _ = codec.pack_device(res, fits, 9)
    assert payload == ref_payload


def test_gpu_stream_equals_oracle_and_decodes():
    x = smooth((96, 80, 40), seed=3, rough=0.01)
    eb = 1e-3 * float(x.max() - x.min())
    pred = InterpPredictor("cubic", 3, 8, 1)
    stream, stats = latec_amd.compress(x, eb, pred, levels=3, anchor_stride=8, anchor_block=1)
    ref_stream, ref_recon = O.compress(x, eb, "cubic", 3, 8, 1, None, 1)
    assert stream == ref_stream
    dec = latec_amd.decompress(stream)
    assert np.array_equal(dec, ref_recon[0])
    assert np.array_equal(dec, stats["recon"])
    assert np.abs(dec - x).max() <= eb
    assert np.array_equal(O.decompress(stream)[0], dec)


def test_gpu_config1_image_stream():
    """512x512 GRF image, InterpPredictor('cubic', 4, 16) (anchor_block 4)."""
    from latec_amd.synthetic import config1_field

    x = config1_field(device="cpu")
    pred = InterpPredictor("cubic", levels=4, anchor_stride=16)
    stream, stats = latec_amd.compress(x, 1e-3, pred, levels=4, anchor_stride=16,
                                       anchor_block=4)
    ref_stream, _ = O.compress(x, 1e-3, "cubic", 4, 16, 4, None, 1)
    assert stream == ref_stream
    dec = latec_amd.decompress(stream)
    assert np.abs(dec - x).max() <= 1e-3
    print("C1 bits/value", 8 * len(stream) / x.size)


def test_gpu_uint8_rgb_round_output():
    rng = np.random.RandomState(4)
    img = np.clip(smooth((60, 50, 3), seed=2) * 100 + 128 + rng.randn(60, 50, 3) * 3, 0,
                  255).astype(np.uint8)
    pred = InterpPredictor("cubic", 3, 8, 1)
    stream, _ = latec_amd.compress(img, 2.0, pred, levels=3, anchor_stride=8, anchor_block=1)
    ref_stream, _ = O.compress(img, 2.0, "cubic", 3, 8, 1, None, 1)
    assert stream == ref_stream
    dec = latec_amd.decompress(stream)
    assert dec.dtype == np.uint8
    assert np.abs(dec.astype(int) - img.astype(int)).max() <= 2


def test_gpu_nonfinite_values_become_outliers():
    x = smooth((40, 36), seed=5)
    x[5, 7] = np.nan
    x[20, 20] = np.inf
    pred = InterpPredictor("cubic", 3, 8, 1)
    stream, stats = latec_amd.compress(x, 1e-3, pred, levels=3, anchor_stride=8, anchor_block=1)
    ref_stream, _ = O.compress(x, 1e-3, "cubic", 3, 8, 1, None, 1)
    assert stream == ref_stream
    dec = latec_amd.decompress(stream)
    assert np.isnan(dec[5, 7]) and np.isinf(dec[20, 20])


def test_gpu_determinism():
    x = smooth((50, 40, 30), seed=7, rough=0.05)
    pred = InterpPredictor("cubic", 3, 8, 1)
    s1, _ = latec_amd.compress(x, 1e-3, pred, levels=3, anchor_stride=8, anchor_block=1)
    s2, _ = latec_amd.compress(x, 1e-3, pred, levels=3, anchor_stride=8, anchor_block=1)
    assert s1 == s2


def test_gpu_tune_size_sweep_never_worse():
    x = smooth((48, 40, 24), seed=8, rough=0.01)
    p1 = InterpPredictor("cubic", 3, 8, 1)
    fast, _ = latec_amd.compress(x, 1e-3, p1, levels=3, anchor_stride=8, anchor_block=1)
    p2 = InterpPredictor("cubic", 3, 8, 1)
    size, st = latec_amd.compress(x, 1e-3, p2, levels=3, anchor_stride=8, anchor_block=1,
                                  tune="size")
    assert len(size) <= len(fast)
    assert st["tune_candidates"] > 1
    assert np.abs(latec_amd.decompress(size) - x).max() <= 1e-3


class NearestPredictor:
    """Foreign predictor (restates the reference's tests/helpers.py:6-27
    test double): nearest known value by a distance transform + box blur."""

    def predict(self, recon, known, pos=None):
        from scipy import ndimage

        idx = ndimage.distance_transform_edt(~known, return_distances=False, return_indices=True)
        filled = recon[(slice(None),) + tuple(idx)]
        out = np.stack([ndimage.uniform_filter(ch, 3) for ch in filled]).astype(np.float32)
        return out[:, pos] if pos is not None else out


def test_gpu_generic_predictor_path():
    img = (smooth((40, 30), seed=9) * 80 + 100).astype(np.uint8)
    stream, stats = latec_amd.compress(img, 2.0, NearestPredictor(), levels=3, anchor_stride=8)
    with pytest.raises(ValueError, match="predictor_factory"):
        latec_amd.decompress(stream)
    dec = latec_amd.decompress(stream, lambda h: NearestPredictor())
    assert np.abs(dec.astype(int) - img.astype(int)).max() <= 2
    assert np.array_equal(dec, stats["recon"])


def test_gpu_interp_predict_standalone():
    """InterpPredictor.predict == the oracle's per-stage prediction."""
    shape = (70, 90)
    recon = np.random.RandomState(0).rand(2, *shape).astype(np.float32) * 50
    for order in ("linear", "cubic"):
        for center in (0, 1, 2):
            pred = InterpPredictor(order, 3, 8, 1, center=center, end_mode=3)
            known = np.zeros(shape, bool)
            for mask, s, w, axis in O.stage_plan(shape, 3, 8, 1):
                if w and mask.any():
                    got = pred.predict(recon, known, mask)
                    ref = O.predict(recon, mask, s, w, axis, order, center, 3)
                    assert np.array_equal(got, ref)
                known |= mask


@pytest.mark.parametrize("case", [c for c in CASES if c[0] in ((70, 45, 133), (67, 33, 129))],
                         ids=lambda c: f"{c[0]}-{c[5]}")
def test_gpu_slab_kernel_bit_exact(case):
    """The level-fused slab kernel (tuning knob "slab") on levels with partial
    tiles and segments: codes, outliers, recon and stream equal the oracle's."""
    from latec_amd import autotune

    old = autotune.get()
    autotune.set_knobs(slab=1)
    try:
        test_gpu_encode_matches_oracle_bit_exact(case)
        shape, C, levels, stride, block, order, center, end, rel, rough = case
        x = smooth(shape, seed=0, rough=rough)
        eb = rel * float(x.max() - x.min())
        pred = InterpPredictor(order, levels, stride, 1)
        pred.center, pred.end_mode = center, end
        stream, _ = latec_amd.compress(x, eb, pred, levels=levels, anchor_stride=stride,
                                       anchor_block=1)
        ref_stream, ref_recon = O.compress(x, eb, order, levels, stride, 1, center, end)
        assert stream == ref_stream
        assert np.array_equal(latec_amd.decompress(stream), ref_recon[0])
    finally:
        autotune.set_knobs(**old)


@pytest.mark.parametrize("knobs", [(0, 0, 1, 3), (0, 1 << 24, 1, 3), (1, 1 << 24, 1, 3),
                                   (1, 0, 1, 3), (1, 0, 0, 3), (1, 1 << 24, 0, 3), (1, 0, 1, 0),
                                   (1, 1 << 24, 1, 0), (1, 0, 1, 12), (1, 1 << 24, 1, 12)],
                         ids=["march", "march-short", "quads-short", "quads", "pair-waves",
                              "pair-waves-short", "pair-lines", "pair-lines-short",
                              "quads-phase2", "quads-phase2-short"])
@pytest.mark.parametrize("case", [c for c in CASES if len(c[0]) == 3 and c[4] == 1],
                         ids=lambda c: f"{c[0]}-{c[5]}-c{c[6]}e{c[7]}")
def test_gpu_level_kernel_launch_shapes_bit_exact(case, knobs):
    """The per-class march kernels (tuning knob "pair" off), the pair kernels
    with wave runs instead of line blocks ("pair_lines" off) or with pairs
    instead of quads ("pair_quads" off, or quads in phase 2 only), and each
    with one point per thread
    ("min_threads" above every level's size) on every 3-D case: the same
    codes, outliers, recon and stream as the oracle's."""
    from latec_amd import autotune

    old = autotune.get()
    autotune.set_knobs(pair=knobs[0], min_threads=knobs[1], pair_lines=knobs[2],
                       pair_quads=knobs[3], slab=0)
    shape, C, levels, stride, block, order, center, end, rel, rough = case
    quad_ok = knobs[0] == 1 and knobs[2] == 1 and _has_quad_level(shape, levels, stride)
    try:
        check(lib().latec_tune_set(b"quad_launches", 0))
        test_gpu_encode_matches_oracle_bit_exact(case)
        # the quad kernels really ran where the knobs and the line lengths ask for them
        enc_quads = int(lib().latec_tune_get(b"quad_launches"))
        assert (enc_quads > 0) == (quad_ok and bool(knobs[3] & 5)), enc_quads
        x = smooth(shape, seed=0, rough=rough)
        eb = rel * float(x.max() - x.min())
        pred = InterpPredictor(order, levels, stride, 1)
        pred.center, pred.end_mode = center, end
        stream, _ = latec_amd.compress(x, eb, pred, levels=levels, anchor_stride=stride,
                                       anchor_block=1)
        ref_stream, ref_recon = O.compress(x, eb, order, levels, stride, 1, center, end)
        assert stream == ref_stream
        check(lib().latec_tune_set(b"quad_launches", 0))
        assert np.array_equal(latec_amd.decompress(stream), ref_recon[0])
        dec_quads = int(lib().latec_tune_get(b"quad_launches"))
        assert (dec_quads > 0) == (quad_ok and bool(knobs[3] & 10)), dec_quads
    finally:
        autotune.set_knobs(**old)


def _has_quad_level(shape, levels, stride):
    """Whether a level above the first (the ones the pair and quad kernels
    run) has lines along the last axis that the quad kernels take: a length
    that is a multiple of 4, at most 1024.  A level at lattice step s codes
    the grid x[::s], whose last axis has ceil(n / s) points."""
    steps = [max(stride >> k, 1) for k in range(2, levels + 1)]
    lens = [-(-shape[-1] // s) for s in steps]
    return any(n % 4 == 0 and n <= 1024 for n in lens)


def _config2_against_oracle(xd, shape):
    """Encode the device field xd (cubic L5 stride 32, block 1, eb = 1e-3 *
    range, fast mode) with the stage kernels and with the oracle on its host
    copy: reconstruction bits, every stage's codes, the outliers (through the
    payload bytes) and the payload itself must be equal.  Returns what the
    callers check further."""
    from latec_amd.levels import stage_ebs

    x = D.to_host(xd).reshape(1, *shape)
    eb = 1e-3 * (float(x.max()) - float(x.min()))
    plan = get_plan(shape, 1, 5, 32, 1)
    pred = InterpPredictor("cubic", 5, 32, 1)
    o = codec._interp_opts(pred, 3, 1 << 15, False)
    ebs = stage_ebs(shape, 5, 32, 1, eb, 0.9)
    ref_ebs = O.stage_ebs_ref(3, 5, 32, eb, 0.9)
    assert list(ebs) == list(ref_ebs)
    print("oracle encode", shape, flush=True)
    ref_payload, ref_recon, _ = O.compress_payload(x, ref_ebs, 5, 32, 1, "cubic", o.center,
                                                   o.end_mode)
    print("GPU encode", shape, flush=True)
    res = codec.encode_device(xd, plan, ebs, o)
    torch.cuda.synchronize()
    recon = D.to_host(res.d_recon).reshape(x.shape)
    assert np.array_equal(recon.view(np.uint32), ref_recon.view(np.uint32))
    del recon, ref_recon
    fits = codec.fit_device(res)
    payload, rec_len, n_out, frame = codec.pack_device_framed(res, fits, 9)
    assert payload == ref_payload
    return x, eb, plan, ebs, o, res, payload, rec_len, n_out, frame


def test_gpu_config2_slab_matches_oracle():
    """The first 64 planes of the 512^3 field (64 x 512 x 512, cubic L5
    stride 32): codes of every stage, reconstruction, payload and the whole
    stream through the public API equal the oracle's."""
    from latec_amd.synthetic import config2_field

    shape = (64, 512, 512)
    xd = config2_field(device=D.device())[:64].contiguous().reshape(-1)
    x, eb, plan, ebs, o, res, payload, _, _, _ = _config2_against_oracle(xd, shape)
    ref = _oracle_stage_codes(payload, plan.counts, 1)
    codes = D.to_host(res.d_codes, np.uint16)
    for s in range(plan.n_stages):
        lo, n = plan.offsets[s], plan.counts[s]
        assert np.array_equal(codes[lo:lo + n].astype(np.uint32), ref[s][0]), f"stage {s}"
    pred = InterpPredictor("cubic", 5, 32, 1)
    stream, stats = latec_amd.compress(x[0], eb, pred, levels=5, anchor_stride=32,
                                       anchor_block=1)
    ref_stream, ref_recon = O.compress(x[0], eb, "cubic", 5, 32, 1, None, 1)
    assert stream == ref_stream
    dec = latec_amd.decompress(stream)
    assert np.array_equal(dec.view(np.uint32), ref_recon[0].view(np.uint32))
    assert float(np.abs(dec - x[0]).max()) <= eb


@pytest.mark.timeout(900)
def test_gpu_config2_full_size_matches_oracle():
    """512^3, cubic, levels 5, stride 32, eb = 1e-3 * range (the benchmarked
    field): reconstruction and payload bytes equal the oracle's (about a
    minute of oracle time), the framed pack writes the same payload and the
    one-shot zstd-9 frame of it, every value is within the bound and the
    decoder reproduces the encoder's reconstruction."""
    from latec_amd.bitstream import zstd_compress
    from latec_amd.synthetic import config2_field

    shape = (512, 512, 512)
    xd = config2_field(device=D.device()).reshape(-1).contiguous()
    x, eb, plan, ebs, o, res, payload, rec_len, n_out, frame = \
        _config2_against_oracle(xd, shape)
    del x
    assert frame == zstd_compress(payload, 9)
    p2, rec_len2, _ = codec.pack_device(res, codec.fit_device(res), 9)
    assert p2 == payload and np.array_equal(rec_len2, rec_len)
    enc_recon = res.d_recon.clone()
    assert (enc_recon - xd).abs().max().item() <= eb
    d_recon = codec.decode_device(payload, plan, ebs, o)
    assert torch.equal(d_recon, enc_recon)
    print("C2 payload bytes", len(payload), "bits/value", 8 * len(payload) / xd.numel(),
          "outliers", int(n_out.sum()))


def test_gpu_3d_outliers_and_nonfinite_match_oracle():
    """Rough 3-D field with many outliers and non-finite values through the
    level kernels: stream and decode bit-identical to the oracle's."""
    x = smooth((66, 70, 130), seed=11, rough=0.3)
    x[3, 5, 7] = np.nan
    x[40, 41, 99] = np.inf
    x[1, 1, 1] = -np.inf
    pred = InterpPredictor("cubic", 4, 16, 1)
    stream, _ = latec_amd.compress(x, 1e-3, pred, levels=4, anchor_stride=16, anchor_block=1)
    ref_stream, ref_recon = O.compress(x, 1e-3, "cubic", 4, 16, 1, None, 1)
    assert stream == ref_stream
    dec = latec_amd.decompress(stream)
    assert np.array_equal(dec.view(np.uint32), ref_recon[0].view(np.uint32))


def _reference_moments(x, recon, shape, levels, stride, block, order, center, end,
                       tile=1 << 20):
    """Per-stage moments as codec.py:353-381 computes them (float32
    differences a 2^20 tile at a time, float64 sums), on the oracle's
    reconstruction with the oracle's per-stage predictions from it."""
    rows = []
    C = x.shape[0]
    for s, (mask, st, w, ax) in enumerate(O.stage_plan(shape, levels, stride, block)):
        n = int(mask.sum())
        if not n:
            rows.append((0.0, 0.0, 0.0, 0.0, 0.0))
            continue
        vals = x.reshape(C, -1)[:, mask.reshape(-1)]
        rec = recon.reshape(C, -1)[:, mask.reshape(-1)]
        pred = (np.zeros((C, n), np.float32) if s == 0
                else O.predict(recon, mask, st, w, ax, order, center, end))
        m = [0.0, 0.0, 0.0, 0.0, 0.0]
        for a in range(0, n, tile):
            v = vals[:, a:a + tile]
            pe = v - pred[:, a:a + tile]
            re = np.abs(v - rec[:, a:a + tile])
            m[0] += float(np.abs(pe).sum(dtype=np.float64))
            m[1] += float(np.square(pe, dtype=np.float64).sum())
            m[2] += float(re.sum(dtype=np.float64))
            m[3] += float(np.square(re, dtype=np.float64).sum())
            m[4] = max(m[4], float(re.max(initial=0.0)))
        rows.append(tuple(m))
    return np.array(rows)


@pytest.mark.parametrize("case", [
    ((96, 80, 40), 1, 3, 8, 1, "cubic", 0, 1, 1e-3, 0.02),
    ((64, 48), 3, 4, 16, 4, "cubic", 0, 1, 1e-2, 0.01),
    ((33, 40, 512), 1, 5, 32, 1, "cubic", 0, 1, 1e-3, 0.02),
    ((12, 11, 10, 9), 1, 2, 4, 1, "linear", 2, 3, 1e-3, 0.0),
], ids=lambda c: f"{c[0]}-C{c[1]}-b{c[4]}")
def test_gpu_stage_stats_match_reference_moments(case):
    """stats['stage_pred_*'] / ['stage_recon_*'] of the GPU path equal the
    reference's per-stage moments on the oracle's encode to float64 rounding
    of the summation order (the max exactly)."""
    shape, C, levels, stride, block, order, center, end, rel, rough = case
    x = np.stack([smooth(shape, seed=c, rough=rough) for c in range(C)])
    eb = rel * float(x.max() - x.min())
    img = x[0] if C == 1 else np.moveaxis(x, 0, -1)
    pred = InterpPredictor(order, levels, stride, block, center=center, end_mode=end)
    stream, stats = latec_amd.compress(img, eb, pred, levels=levels, anchor_stride=stride,
                                       anchor_block=block)
    ref_stream, ref_recon = O.compress(img, eb, order, levels, stride, block, center, end)
    assert stream == ref_stream
    ref = _reference_moments(x, ref_recon, shape, levels, stride, block, order, center, end)
    got = np.array([stats[k] for k in ("stage_pred_sae", "stage_pred_sse", "stage_recon_sae",
                                       "stage_recon_sse", "stage_recon_max")]).T
    assert got.shape == ref.shape
    np.testing.assert_allclose(got[:, :4], ref[:, :4], rtol=1e-12, atol=0)
    assert np.array_equal(got[:, 4], ref[:, 4])
    assert stats["recon_sse"] == float(sum(stats["stage_recon_sse"]))
    assert all(v > 0 for v, n in zip(stats["stage_pred_sae"], stats["stage_codes"]) if n)


@pytest.mark.parametrize("tune", ["size", "rd"])
def test_gpu_tune_sweep_picks_the_oracle_candidate(tune):
    """tune='size' / 'rd' (codec.py:171-254): the candidates (every end mode
    at the first ratio, then ratio x centre at the kept end mode) are encoded
    by the oracle, the winner chosen by the reference's keys, and the GPU
    path's stream must be that candidate's oracle stream."""
    from latec_amd.predictor import END_MODE_MAX, default_interp_center, default_interp_end_mode

    x = smooth((48, 40, 24), seed=8, rough=0.01)
    eb = 1e-3 * float(x.max() - x.min())
    stream, st = latec_amd.compress(x, eb, InterpPredictor("cubic", 3, 8, 1), levels=3,
                                    anchor_stride=8, anchor_block=1, tune=tune)
    c0, e0 = default_interp_center(3), default_interp_end_mode(3)
    ratios = [1.0, 0.9, 0.8, 0.7]
    centres = [c0] + [c for c in (0, 1, 2) if c != c0]
    cands = {}

    def run(r, c, e):
        if (r, c, e) not in cands:
            s, rec = O.compress(x, eb, "cubic", 3, 8, 1, c, e, eb_ratio=r)
            cands[(r, c, e)] = (len(s), float(np.square(x - rec[0], dtype=np.float64).sum()), s)

    run(ratios[0], c0, e0)
    for e in range(END_MODE_MAX + 1):
        run(ratios[0], c0, e)
    end = min(cands.items(), key=lambda kv: (kv[1][0], kv[1][1]))[0][2]
    for r in ratios:
        for c in centres:
            run(r, c, end)
    keyed = [(size, sse, r, c, e, s) for (r, c, e), (size, sse, s) in cands.items()]
    if tune == "rd":
        lim = min(k[0] for k in keyed) * 1.05
        best = min((k for k in keyed if k[0] <= lim), key=lambda k: (k[1], k[0], k[2], k[3], k[4]))
    else:
        best = min(keyed, key=lambda k: k[:5])
    assert st["tune_candidates"] == len(cands)
    assert (st["eb_ratio"], st["interp_center"], st["interp_end_mode"]) == best[2:5]
    assert stream == best[5]
    assert np.abs(latec_amd.decompress(stream) - x).max() <= eb


def test_gpu_choose_shape_matches_reference_golden():
    """choose_shape (rans.py:204-243, per-point levels) on the GPU-built level
    weights picks the reference's shape on every recorded stage
    (tests/golden/make_shape_golden.py)."""
    fv = np.load(os.path.join(G, "shape_vectors.npz"))
    meta = fv["meta"]
    assert len(meta) >= 50 and len({int(m[4]) for m in meta}) >= 5
    for i, n, eb, kexp, shape in meta:
        i = int(i)
        codes = fv[f"codes{i}"].astype(np.uint32)
        levels = fv[f"levels{i}"]
        assert len(codes) == int(n)
        tables = rans.build_laplace_tables(float(eb), 1 << 15)
        assert rans.choose_kexp(codes, 1 << 15) == int(kexp)
        assert rans.choose_shape(codes, levels, tables, int(kexp)) == int(shape), i


@pytest.mark.parametrize("case", [c for c in CASES if len(c[0]) == 3 and c[4] == 1],
                         ids=lambda c: f"{c[0]}-{c[5]}-c{c[6]}e{c[7]}")
def test_gpu_pair_waves_variant_bit_exact(case):
    """The finest level's encode phases compiled for eight waves per SIMD
    (tuning knob "pair_waves"): the same codes, outliers, recon and payload as
    the oracle's."""
    from latec_amd import autotune

    old = autotune.get()
    autotune.set_knobs(pair=1, pair_lines=1, pair_quads=2, slab=0, pair_waves=1)
    try:
        test_gpu_encode_matches_oracle_bit_exact(case)
    finally:
        autotune.set_knobs(**old)
