"""Write golden vectors of the reference's ``choose_shape`` (latec/rans.py:204-243).

Run once in the build container (where /root/reference exists), on NumPy's
libm exp path like make_golden.py's reference configuration:

    NPY_DISABLE_CPU_FEATURES="$AVX512" python tests/golden/make_shape_golden.py

It loads ``latec/rans.py`` by file path (NumPy only at import) and records,
for stage-like code arrays with per-point scale levels, the shape id the
reference picks: its float64 ``-log2`` sum over a strided sample of the
mixture weights depends on NumPy's exp/log2 and pairwise summation, so only
the reference itself pins it.  Output: ``shape_vectors.npz`` (inputs and
chosen ids).  Nothing here is imported by the tests.
"""

from __future__ import annotations

import os

import numpy as np

from make_golden import load_rans

OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    rans = load_rans()
    rng = np.random.RandomState(4321)
    radius = 1 << 15
    arrays, meta = {}, []
    i = 0
    for n in (50, 2000, 30000):  # 30000: above 1024 / 0.05, so the strided sample
        for base in (0.4, 3.0, 60.0):
            for spread in (0.0, 1.0, 3.0):
                for eb in (1e-3, 0.25):
                    # per-point Laplace scales around base * eb, spread in log2
                    sc = base * eb * np.exp2(spread * rng.standard_normal(n))
                    # residuals drawn from a mixture: the model's own scale or 6x wider
                    wide = rng.rand(n) < 0.15
                    r = rng.laplace(0.0, sc * np.where(wide, 6.0, 1.0)) / (2 * eb)
                    q = np.clip(np.rint(r), -(radius - 1), radius - 1).astype(np.int64)
                    codes = (q + radius).astype(np.uint32)
                    codes[rng.rand(n) < 0.001] = 0
                    levels = rans.scale_to_level(sc.astype(np.float32), eb)
                    kexp = rans.choose_kexp(codes, radius)
                    tables = rans.build_laplace_tables(eb, radius)
                    shape = rans.choose_shape(codes, levels, tables, kexp)
                    arrays[f"codes{i}"] = codes.astype(np.uint16)  # < 2 * radius
                    arrays[f"levels{i}"] = levels
                    meta.append((i, n, eb, kexp, shape))
                    i += 1
    arrays["meta"] = np.array([(a, b, c, d, e) for a, b, c, d, e in meta], np.float64)
    np.savez_compressed(os.path.join(OUT, "shape_vectors.npz"), **arrays)
    print("wrote", i, "cases; shapes chosen:", sorted({int(m[4]) for m in meta}))


if __name__ == "__main__":
    main()
