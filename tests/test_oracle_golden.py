"""Pin the C oracle against the vectors written from the reference's own
latec/rans.py (tests/golden/make_golden.py, NumPy on its libm path)."""

import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

G = os.path.join(os.path.dirname(__file__), "golden")
GOLD = json.load(open(os.path.join(G, "rans_golden_libm.json")))
FIT = np.load(os.path.join(G, "fit_vectors.npz"))
SMALL = np.load(os.path.join(G, "cdf_small.npz"))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("kexp", range(1, 11))
def test_level_weights_bit_exact(kexp):
    assert sha(O.level_weights(kexp)) == GOLD["level_weights_f64"][str(kexp)]


def test_small_weight_tables_bit_exact():
    for k in (1, 2, 3, 4):
        assert np.array_equal(O.level_weights(k), SMALL[f"lw{k}"])
    assert np.array_equal(O.model_weights(1, 1, 3), SMALL["mw1,1,3"])
    assert np.array_equal(O.model_weights(3, 7, 4), SMALL["mw3,7,4"])


@pytest.mark.parametrize("kexp", range(1, 11))
def test_cdf_and_nl2_tables(kexp):
    bad = []
    for key, v in GOLD["tables"].items():
        if key.startswith("p"):
            continue
        k, s, o = map(int, key.split(","))
        if k != kexp:
            continue
        if sha(O.build_cdfs(k, s, 24, o)) != v["cdf"]:
            bad.append(("cdf", key))
        if sha(O.neg_log2_weights(k, s, o)) != v["nl2"]:
            bad.append(("nl2", key))
    assert not bad


def test_wide_tables():
    """kexp 11..15: rows of more than 8192 symbols are summed in NumPy's
    8192-element buffer chunks."""
    for key, v in GOLD["tables"].items():
        if key.startswith("p"):
            continue
        k, s, o = map(int, key.split(","))
        if k >= 11:
            assert sha(O.build_cdfs(k, s, 24, o)) == v["cdf"], key
            assert sha(O.neg_log2_weights(k, s, o)) == v["nl2"], key


def test_other_precisions():
    for key, v in GOLD["tables"].items():
        if key.startswith("p"):
            p, k, _, _ = map(int, key[1:].split(","))
            assert sha(O.build_cdfs(k, 0, p, 0)) == v["cdf"], key


def test_full_small_tables():
    for name in ("1,0,0", "3,0,0", "3,7,4", "5,12,1"):
        k, s, o = map(int, name.split(","))
        assert np.array_equal(O.build_cdfs(k, s, 24, o), SMALL[name])


def _hist(tiles):
    h = np.zeros(tiles.shape[1], np.float32)
    for t in tiles:
        h = (h.astype(np.float64) + t).astype(np.float32)
    return h


def test_const_level_fit_choices():
    """(level, shape, owx) match the reference on every recorded histogram;
    the returned bit count matches its float32 BLAS value to 3e-7."""
    for f in GOLD["const_fits"]:
        h = _hist(FIT[f"tiles{f['i']}"])
        lv, sh, ow, bits = O.choose_const_level_hist(h, f["n"], f["kexp"])
        assert (lv, sh, ow) == (f["level"], f["shape"], f["owx"]), f
        assert abs(bits - f["bits"]) <= 3e-7 * max(1.0, abs(f["bits"])), f


def test_scale_to_level():
    for i, eb in enumerate(FIT["scale_ebs"]):
        assert np.array_equal(O.scale_to_level(FIT["scales"], eb), FIT["scale_levels"][i])


def test_outlier_weight_kats():
    for a, b, e in GOLD["outlier_weight_kats"]:
        assert O.choose_outlier_weight(a, b) == e
    # reference KAT (tests/test_rans.py:349-353)
    assert O.choose_outlier_weight(0, 100) == 0


def test_tile_histogram_accumulation_is_float32_per_tile():
    """rans.py:413-420 accumulates float32 per 2**22 codes: a bin past 2**24
    keeps the per-tile roundings, which a single total would not."""
    n = (1 << 22) * 5 + 3
    codes = np.full(n, 1 << 15, np.uint32)
    h = O.const_hist(codes, 1 << 15, 1)
    ref = np.float32(0)
    for a in range(0, n, 1 << 22):
        ref = np.float32(np.float64(ref) + min(1 << 22, n - a))
    assert h[2] == ref


def test_scale_thresholds_reproduce_host_levels():
    """The bisected thresholds the device compares against: the number of
    thresholds at or below a scale equals the host NumPy level of that scale
    (on the golden scales, on random scales over the whole clipped range and
    past it, and on every threshold and its float32 neighbours); on this
    machine the host levels also equal the golden ones."""
    from latec_amd import rans as R

    rng = np.random.RandomState(2)
    for i, eb in enumerate(FIT["scale_ebs"]):
        eb = float(eb)
        t = R.scale_thresholds(eb)
        assert t.dtype == np.float32 and len(t) == 63 and np.all(np.diff(t) >= 0)
        v = np.concatenate([FIT["scales"].astype(np.float32), t,
                            np.nextafter(t, 0, dtype=np.float32),
                            np.nextafter(t, np.inf, dtype=np.float32),
                            (eb * np.exp(rng.uniform(-4, 9.5, 200000))).astype(np.float32)])
        assert np.array_equal(np.searchsorted(t, v, side="right"), R.scale_to_level_host(v, eb))
        assert np.array_equal(R.scale_to_level_host(FIT["scales"], eb), FIT["scale_levels"][i])
