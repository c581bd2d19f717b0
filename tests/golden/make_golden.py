"""Write golden vectors for the functions only an import of the reference pins.

Run once in the build container (where /root/reference exists):

    NPY_DISABLE_CPU_FEATURES="$AVX512" python tests/golden/make_golden.py
    python tests/golden/make_golden.py

with AVX512="AVX512F AVX512CD AVX512_KNL AVX512_KNM AVX512_SKX AVX512_CLX
AVX512_CNL AVX512_ICL".  NumPy's float64 exp on AVX-512 hosts is its own SIMD
routine, which differs from libm's in the last bit for ~5% of arguments, and
the reference's PMF tie-breaking between mirror-image bins turns those bits
into different integer tables.  The first run (NumPy on its libm path, the
configuration of any host without AVX-512) writes ``rans_golden_libm.json``,
the one the oracle is pinned to; the second writes
``rans_golden_avx512.json``, kept to document the host dependence.

It loads ``latec/rans.py`` from the reference by file path (the module needs
only NumPy at import; ``constriction`` is imported lazily inside the coder
functions, which are never called here) and records, as small data files:

* sha256 of the integer CDF tables ``_build_cdfs_cached(kexp, shape, 64, p,
  owx)`` and of the float32 ``-log2`` weight tables ``_neg_log2_weights`` --
  NumPy SIMD exp/log and pairwise row sums make these depend on NumPy itself;
* ``choose_const_level`` results on recorded stage code histograms (float32
  BLAS sgemv inside);
* ``scale_to_level`` outputs (float32 log under NEP 50 promotion).

Nothing here is imported by the tests; they read the JSON/NPZ files only.
"""

from __future__ import annotations

import hashlib
import importlib.util
import json
import os
import sys

import numpy as np

REF = "/root/reference/latec/rans.py"
OUT = os.path.dirname(os.path.abspath(__file__))


def load_rans():
    spec = importlib.util.spec_from_file_location("_latec_ref_rans", REF)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod  # dataclasses resolve their module by name
    spec.loader.exec_module(mod)
    return mod


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def table_cases():
    cases = []
    for kexp in range(1, 11):
        for shape in range(21):
            for owx in (0, 3, 7):
                cases.append((kexp, shape, owx))
    for kexp in (11, 12, 13, 14, 15):
        for shape in (0, 1, 6, 16):
            for owx in (0, 2):
                cases.append((kexp, shape, owx))
    return cases


def stage_codes(rng, n, scale, radius, out_rate):
    """Laplace residual codes as a stage of the plain codec would emit them."""
    q = np.rint(rng.laplace(0.0, scale, n)).astype(np.int64)
    q = np.clip(q, -(radius - 1), radius - 1)
    codes = (q + radius).astype(np.uint32)
    if out_rate:
        codes[rng.rand(n) < out_rate] = 0
    return codes


def main():
    rans = load_rans()
    tables = {}
    for kexp, shape, owx in table_cases():
        cdf = rans._build_cdfs_cached(kexp, shape, 64, 24, owx)
        nl2 = rans._neg_log2_weights(kexp, shape, 64, 24, owx)
        tables[f"{kexp},{shape},{owx}"] = {"cdf": sha(cdf), "nl2": sha(nl2)}
        rans._build_cdfs_cached.cache_clear()
        rans._neg_log2_weights.cache_clear()
    # other precisions used by the reference's own tests (test_rans.py:26-62)
    for prec in (10, 14, 15):
        for kexp in range(1, 8):
            cdf = rans._build_cdfs_cached(kexp, 0, 64, prec, 0)
            tables[f"p{prec},{kexp},0,0"] = {"cdf": sha(cdf)}
    lw = {str(k): sha(rans._level_weights(k, 64)) for k in range(1, 11)}
    small = {
        f"{k},{s},{o}": rans._build_cdfs_cached(k, s, 64, 24, o)
        for k, s, o in ((1, 0, 0), (3, 0, 0), (3, 7, 4), (5, 12, 1))
    }
    small.update({f"lw{k}": rans._level_weights(k, 64) for k in (1, 2, 3, 4)})
    small.update({f"mw{k},{s},{o}": rans._with_outlier_weight(
        rans._shaped_weights(k, s, 64), o) for k, s, o in ((1, 1, 3), (3, 7, 4))})
    from numpy._core._multiarray_umath import __cpu_features__

    if not __cpu_features__.get("AVX512F"):
        np.savez_compressed(os.path.join(OUT, "cdf_small.npz"), **small)

    # const-level fit on recorded histograms
    rng = np.random.RandomState(1234)
    radius = 1 << 15
    fits = []
    arrays = {}
    specs = [(n, sc, orate) for n in (37, 1000, 5000, 70000)
             for sc in (0.05, 0.6, 3.0, 40.0, 900.0) for orate in (0.0, 0.002, 0.05)]
    specs += [(5_000_000, 2.0, 0.0001), (4_300_000, 0.3, 0.0)]  # > one 2**22 tile
    for i, (n, sc, orate) in enumerate(specs):
        codes = stage_codes(rng, n, sc, radius, orate)
        kexp = rans.choose_kexp(codes, radius)
        tab = rans.const_tables(radius)
        level, shape, owx, bits = rans.choose_const_level(codes, tab, kexp)
        # store per-tile symbol counts (the reference accumulates float32 per tile)
        tiles = []
        for a in range(0, n, 1 << 22):
            syms = rans._narrow(codes[a:a + (1 << 22)], radius, kexp)
            tiles.append(np.bincount(syms, minlength=2 << kexp).astype(np.int64))
        arrays[f"tiles{i}"] = np.stack(tiles)
        fits.append({"i": i, "n": n, "kexp": int(kexp), "level": int(level),
                     "shape": int(shape), "owx": int(owx), "bits": float(bits)})

    # scale_to_level
    ebs = np.array([1e-3, 0.37, 2.0, 1e-6, 5e-4 * 0.59049], np.float64)
    scales = np.concatenate([
        np.exp(rng.uniform(np.log(1e-9), np.log(1e4), 4000)),
        np.array([0.0, 1e-30, 1e30, np.inf]),
    ]).astype(np.float32)
    levels = np.stack([rans.scale_to_level(scales, eb) for eb in ebs])
    arrays["scales"] = scales
    arrays["scale_ebs"] = ebs
    arrays["scale_levels"] = levels
    if not __cpu_features__.get("AVX512F"):
        np.savez_compressed(os.path.join(OUT, "fit_vectors.npz"), **arrays)

    out = {
        "source": "latec/rans.py (reference), NumPy " + np.__version__,
        "tables": tables,
        "level_weights_f64": lw,
        "const_fits": fits,
        "outlier_weight_kats": [
            [int(a), int(b), int(rans.choose_outlier_weight(a, b))]
            for a, b in ((0, 10), (1, 10), (1, 100), (1, 1000), (1, 8192), (3, 7),
                         (1, 3), (5, 10), (9, 10), (1, 70), (2, 513))
        ],
    }
    from numpy._core._multiarray_umath import __cpu_features__

    tag = "avx512" if __cpu_features__.get("AVX512F") else "libm"
    out["numpy_exp_path"] = tag
    if tag == "avx512":  # the host-dependent part only
        out = {k: out[k] for k in ("source", "tables", "level_weights_f64", "numpy_exp_path")}
    with open(os.path.join(OUT, f"rans_golden_{tag}.json"), "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print("wrote", len(tables), "table hashes,", len(fits), "fits")


if __name__ == "__main__":
    main()
