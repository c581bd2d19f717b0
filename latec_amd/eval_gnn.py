"""Bits/value of the GNN codec against plain interpolation on one field.

    python -m latec_amd.eval_gnn --ckpt ckpt.pt --field c3 --shape 32,32,32,32

Both codecs run on the same GPU at the same relative bound (eb = rel * range):
``GNNCodec.compress`` (BF16 or FP32 kernels, gate on) and ``latec_amd.compress``
with ``InterpPredictor('cubic')`` over the same dyadic schedule depth.  Every
decoded value is checked against the bound.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def field(kind: str, shape, seed: int):
    from .synthetic import config1_field, config2_field, config3_field, mixed_field

    if kind == "c1":
        return config1_field(seed, shape)
    if kind == "c2":
        return config2_field(seed, shape)
    if kind == "c3":
        return config3_field(seed, shape)
    return mixed_field(shape, kind, seed)


def compare(ckpt, x: np.ndarray, rel_eb: float = 1e-3, numerics: str = "bf16",
            levels=None) -> dict:
    import torch

    from . import InterpPredictor, compress, decompress
    from .gnn_codec import GNNCodec, _auto_levels, roundtrip_slack
    from .gnn_predictor import load_model

    codec = GNNCodec(ckpt, numerics=numerics)
    lv = levels or _auto_levels(x.shape, load_model(ckpt)[1])
    t0 = time.perf_counter()
    s_gnn = codec.compress(x, rel_eb, levels=lv)
    torch.cuda.synchronize()
    t_gnn = time.perf_counter() - t0
    t0 = time.perf_counter()
    y = codec.uncompress(s_gnn).numpy()
    t_gnn_dec = time.perf_counter() - t0
    S = float(x.max() - x.min())
    eb_abs = rel_eb * S
    err_gnn = float(np.abs(y - x).max())
    assert err_gnn <= eb_abs + roundtrip_slack(x.min(), x.max()), (err_gnn, eb_abs)
    stride = 1 << lv
    p = InterpPredictor("cubic", lv, stride, 1)
    t0 = time.perf_counter()
    s_int, _ = compress(x, eb_abs, p, levels=lv, anchor_stride=stride, anchor_block=1)
    t_int = time.perf_counter() - t0
    yi = decompress(s_int)
    err_int = float(np.abs(yi - x).max())
    assert err_int <= eb_abs
    n = x.size
    return {"shape": list(x.shape), "rel_eb": rel_eb, "levels": lv, "numerics": numerics,
            "gnn_bits_per_value": 8 * len(s_gnn) / n, "interp_bits_per_value": 8 * len(s_int) / n,
            "gnn_bytes": len(s_gnn), "interp_bytes": len(s_int),
            "gnn_compress_s": t_gnn, "gnn_decompress_s": t_gnn_dec, "interp_compress_s": t_int,
            "gnn_max_err": err_gnn, "interp_max_err": err_int, "eb_abs": eb_abs,
            "gate_stats": {k: v for k, v in codec.last_stats.items() if k != "recon"}}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--ckpt", required=True)
    ap.add_argument("--field", default="c3")
    ap.add_argument("--shape", default="32,32,32,32")
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--eb", type=float, default=1e-3)
    ap.add_argument("--numerics", default="bf16")
    ap.add_argument("--levels", type=int, default=None)
    a = ap.parse_args(argv)
    shape = tuple(int(s) for s in a.shape.split(","))
    x = field(a.field, shape, a.seed)
    print(json.dumps(compare(a.ckpt, x, a.eb, a.numerics, a.levels)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
