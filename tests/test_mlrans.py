"""Multi-lane rANS stream version 5 (this project's addition beside version 4).

CPU tests pin the oracle's specification of the lane coder by what the
mathematics fixes: exact round trips, the lane split, and the code length
against the model's cross-entropy (a dropped renormalization word, a wrong
slot shift or a swapped frequency/start breaks one of them).  GPU tests
check the library's coder against that specification byte for byte.
"""

import struct

import numpy as np
import pytest

from oracle import oracle as O


def _cdf(kexp=6, level=30, shape=0, owx=0):
    return O.build_cdfs(kexp, shape, 24, owx)[level]


def _laplace_syms(n, kexp, scale, seed):
    rng = np.random.RandomState(seed)
    q = np.clip(np.rint(rng.laplace(0, scale, n)), -(1 << kexp) + 1, (1 << kexp) - 1)
    return (q + (1 << kexp)).astype(np.uint32)


@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 12289, 100000])
def test_oracle_lane_coder_round_trip(n):
    cdf = _cdf()
    syms = _laplace_syms(n, 6, 3.0, n)
    words = O.mlrans_encode_syms(syms, cdf)
    L = O.mlrans_lanes(n)
    assert L == max(1, -(-n // 4096))
    assert words[0] == L and words[1] == 0
    assert words[1 + L] == len(words) - 2 - L
    assert np.array_equal(O.mlrans_decode_syms(words, n, cdf), syms)


def test_oracle_lane_of_symbol_is_index_mod_lanes():
    """Changing symbol i alters only lane i % L's words."""
    n, cdf = 3 * 4096 + 5, _cdf()
    syms = _laplace_syms(n, 6, 3.0, 1)
    a = O.mlrans_encode_syms(syms, cdf)
    s2 = syms.copy()
    s2[4097] = syms[4097] + 1 if syms[4097] < 100 else syms[4097] - 1
    b = O.mlrans_encode_syms(s2, cdf)
    L = O.mlrans_lanes(n)
    lane = 4097 % L

    def lane_words(w, l):
        return w[2 + L + w[1 + l]: 2 + L + w[2 + l]]

    for l in range(L):
        same = np.array_equal(lane_words(a, l), lane_words(b, l))
        assert same == (l != lane)


def test_oracle_code_length_meets_cross_entropy():
    """Coded bits = sum of -log2(f/2^24) + at most 64 bits of final state per
    lane + the lane table (range ANS loses < 1e-4 bits/symbol here)."""
    cdf = _cdf(kexp=8, level=34)
    syms = _laplace_syms(200000, 8, 6.0, 7)
    words = O.mlrans_encode_syms(syms, cdf)
    f = (cdf[1:].astype(np.float64) - cdf[:-1])[syms]
    ideal = float(np.sum(24 - np.log2(f)))
    L = O.mlrans_lanes(len(syms))
    data_bits = 32 * (len(words) - 2 - L)
    assert ideal <= data_bits <= ideal + 64 * L + 1e-4 * len(syms)
    assert data_bits >= ideal + 31 * L - 32 * L  # final states carry >= 31 bits each


def test_oracle_decode_rejects_corruption():
    cdf = _cdf()
    syms = _laplace_syms(9000, 6, 3.0, 3)
    w = O.mlrans_encode_syms(syms, cdf)
    bad = w.copy()
    bad[0] += 1
    with pytest.raises(ValueError):
        O.mlrans_decode_syms(bad, len(syms), cdf)
    with pytest.raises(ValueError):
        O.mlrans_decode_syms(w[:-1], len(syms), cdf)
    bad = w.copy()
    bad[-1] ^= 0x5A5A5A5A  # a flipped data word: decodes to other symbols or fails the end check
    try:
        out = O.mlrans_decode_syms(bad, len(syms), cdf)
        assert not np.array_equal(out, syms)
    except ValueError:
        pass


def test_oracle_v5_stream_same_reconstruction_as_v4():
    rng = np.random.RandomState(2)
    x = np.cumsum(np.cumsum(rng.randn(40, 33, 17), 0), 2).astype(np.float32)
    x += 0.05 * rng.randn(*x.shape).astype(np.float32)
    eb = 1e-3 * float(x.max() - x.min())
    s4, r4 = O.compress(x, eb, "cubic", 3, 8, 1)
    s5, r5 = O.compress_v5(x, eb, "cubic", 3, 8, 1)
    assert np.array_equal(r4, r5)
    assert struct.unpack_from("<H", s5, 8)[0] == 5
    assert np.array_equal(O.decompress_v5(s5), r4)
    # the v5 record blob is the lane stream of the stage's own fit
    assert abs(len(s5) - len(s4)) < 0.05 * len(s4) + 2048


# ---------------------------------------------------------------------------
# GPU: the library's lane coder against the specification
# ---------------------------------------------------------------------------

@pytest.mark.gpu
def test_gpu_v5_stream_equals_oracle_and_decodes():
    import latec_amd
    from latec_amd.predictor import InterpPredictor

    rng = np.random.RandomState(5)
    x = np.cumsum(np.cumsum(rng.randn(96, 80, 40), 0), 2).astype(np.float32)
    x += 0.02 * rng.randn(*x.shape).astype(np.float32)
    eb = 1e-3 * float(x.max() - x.min())
    pred = InterpPredictor("cubic", 3, 8, 1)
    print("case v5 96x80x40", flush=True)
    stream, stats = latec_amd.compress(x, eb, pred, levels=3, anchor_stride=8, anchor_block=1,
                                       stream_version=5)
    ref, ref_recon = O.compress_v5(x, eb, "cubic", 3, 8, 1)
    assert stream == ref
    dec = latec_amd.decompress(stream)
    assert np.array_equal(dec, ref_recon[0])
    assert np.abs(dec - x).max() <= eb


@pytest.mark.gpu
def test_gpu_v5_outliers_nonfinite_and_wide_window():
    """Outliers, non-finite values, a wide coder window (kexp > 12: CDF rows
    read from global memory), and a 1 x n grid."""
    import latec_amd
    from latec_amd.predictor import InterpPredictor

    rng = np.random.RandomState(9)
    x = rng.randn(70, 90).astype(np.float32) * 50
    x[3, 4] = np.nan
    x[60, 7] = np.inf
    pred = InterpPredictor("cubic", 3, 8, 1)
    print("case v5 70x90 noise", flush=True)
    s, _ = latec_amd.compress(x, 1e-4, pred, levels=3, anchor_stride=8, anchor_block=1,
                              stream_version=5)
    ref, ref_recon = O.compress_v5(x, 1e-4, "cubic", 3, 8, 1)
    assert s == ref
    dec = latec_amd.decompress(s)
    assert np.array_equal(dec.view(np.uint32), ref_recon[0].view(np.uint32))
    y = np.cumsum(rng.randn(5000)).astype(np.float32)[None, :]  # a 1 x 5000 image
    pred = InterpPredictor("linear", 5, 32, 1)
    print("case v5 1x5000", flush=True)
    s, _ = latec_amd.compress(y, 1e-3, pred, levels=5, anchor_stride=32, anchor_block=1,
                              stream_version=5)
    ref, ref_recon = O.compress_v5(y, 1e-3, "linear", 5, 32, 1)
    assert s == ref
    assert np.array_equal(latec_amd.decompress(s), ref_recon[0])


@pytest.mark.gpu
def test_gpu_v5_rejects_corrupt_blob():
    import latec_amd
    from latec_amd.predictor import InterpPredictor

    rng = np.random.RandomState(1)
    x = np.cumsum(rng.randn(64, 64), 1).astype(np.float32)
    pred = InterpPredictor("cubic", 3, 8, 1)
    s, _ = latec_amd.compress(x, 1e-3, pred, levels=3, anchor_stride=8, anchor_block=1,
                              stream_version=5)
    bad = bytearray(s)
    # the last stage's first lane offset table entry after [L]: make it point past the blob
    hsize = 80 + 1 + 4 * 2
    off = hsize
    import latec_amd.bitstream as B

    recs = []
    for _ in range(1 + 3 * 3):
        rec = B.parse_record(bytes(s[hsize:]), off - hsize)
        recs.append((off, rec))
        off += rec[7]
    big_off, big = max(recs, key=lambda r: r[1][0])
    struct.pack_into("<I", bad, big_off + 12 + 8, 0x7FFFFFF0)
    print("case v5 corrupt", flush=True)
    with pytest.raises(ValueError):
        latec_amd.decompress(bytes(bad))


def test_host_blob_check_rejects_wrapping_lane_offsets():
    """A crafted version-5 blob whose lane offsets wrap a 32-bit sum (a lane
    starting at word 0xFFFFFFFE) is rejected on the host, before any device
    read: by check_blob directly, and by decode_payload while it parses the
    records (it raises before it touches the workspace)."""
    from latec_amd import mlrans
    from latec_amd.predictor import get_plan

    cdf = _cdf()
    n = 3 * 4096 + 5  # four lanes
    syms = _laplace_syms(n, 6, 3.0, 11)
    blob = O.mlrans_encode_syms(syms, cdf)
    L = O.mlrans_lanes(n)
    assert L == 4 and blob[0] == L
    raw = blob.astype("<u4").tobytes()
    mlrans.check_blob(raw, 0, len(raw), n)  # the genuine blob passes
    for lane, value in ((1, 0xFFFFFFFE), (2, 0xFFFFFFFF), (0, 3), (L, int(blob[L]) + 1)):
        bad = blob.copy()
        bad[1 + lane] = value
        with pytest.raises(ValueError, match="multi-lane"):
            mlrans.check_blob(bad.astype("<u4").tobytes(), 0, len(raw), n)
    # wrong lane count, truncation, misalignment
    bad = blob.copy()
    bad[0] = L + 1
    with pytest.raises(ValueError):
        mlrans.check_blob(bad.astype("<u4").tobytes(), 0, len(raw), n)
    with pytest.raises(ValueError):
        mlrans.check_blob(raw[:-4], 0, len(raw) - 4, n)
    with pytest.raises(ValueError):
        mlrans.check_blob(b"\0\0" + raw, 2, len(raw), n)
    # through decode_payload: a 1-D plan whose stages are all rANS records
    plan = get_plan((64,), 1, 2, 4, 1)
    recs = []
    for s in range(plan.n_stages):
        k = plan.counts[s]
        codes = np.full(k, 1 << 15, np.uint32)
        recs.append(O.pack_stage_v5(codes, np.zeros(0, np.float32)))
    s_last = max(range(plan.n_stages), key=lambda s: plan.counts[s])
    rec = bytearray(recs[s_last])
    struct.pack_into("<I", rec, 12 + 4, 0xFFFFFFFE)  # lane 0's end offset
    recs[s_last] = bytes(rec)
    with pytest.raises(ValueError, match="multi-lane"):
        mlrans.decode_payload(b"".join(recs), plan, None)
