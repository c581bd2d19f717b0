"""Round-trip evaluation of a raw tensor (scripts/eval_tensor.py).

    python -m latec_amd.eval_tensor field.npy --eb 1e-3 --rel
    python -m latec_amd.eval_tensor field.npy --eb 1e-3 --rel --predictor gnn \\
        --gnn-checkpoint latec_amd/checkpoints/gnn_d64_agg2.pt

The CLI's flags, the tensor read from ``.npy`` / ``.pt``, and one report line
per codec (``report``, eval_tensor.py:44-73): bytes, ratio, bits/value,
PSNR over the value range, max error and PASS/FAIL against the bound, and
compress/decompress seconds.  With ``--predictor gnn`` the interpolation
codec runs beside it as the baseline (on a centred crop of at most
``LATEC_BASELINE_MAXVOX`` values, each edge a multiple of the anchor
stride).  SZ3 needs an external package that is not part of this one: its
lines report it skipped.
"""

from __future__ import annotations

import argparse
import copy
import os
import sys
import time

import numpy as np

from .cli import add_common, build_predictor, load_tensor, run_compress

DEFAULT_CKPT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "checkpoints",
                            "gnn_d64_agg2.pt")


def cap_crop(arr: np.ndarray, stride: int, cap: int) -> np.ndarray:
    """Centred crop of at most ``cap`` values with every edge a multiple of
    ``stride`` (eval_tensor.py:29-41); the array itself when it fits."""
    if arr.size <= cap:
        return arr
    edge = max(int(cap ** (1.0 / arr.ndim)), stride)
    edge = max(edge - edge % stride, stride)
    sl = tuple(slice((d - min(d, edge)) // 2, (d - min(d, edge)) // 2 + min(d, edge))
               for d in arr.shape)
    return np.ascontiguousarray(arr[sl])


def report_line(label, arr, rec, nbytes, eb, t_comp=None, t_dec=None, *, slack=0.0):
    """(line, bound held) for one codec's result on ``arr``."""
    a = arr.astype(np.float64)
    r = np.asarray(rec).reshape(arr.shape).astype(np.float64)
    max_err = float(np.abs(a - r).max())
    mse = float(np.mean((a - r) ** 2))
    peak = 255.0 if arr.dtype == np.uint8 else max(float(arr.max()) - float(arr.min()), 1e-12)
    psnr = 10 * np.log10(peak ** 2 / mse) if mse > 0 else float("inf")
    bpv = 8 * nbytes / arr.size
    ratio = arr.nbytes / nbytes
    t = f"  {t_comp:.1f}s/{t_dec:.1f}s" if t_comp is not None else ""
    ok = max_err <= eb + slack
    line = (f"  [{label:6s}] {nbytes:>10} B  ratio {ratio:7.2f}  {bpv:7.3f} bpv  "
            f"PSNR {psnr:6.2f} dB  maxerr {max_err:.3g} {'PASS' if ok else 'FAIL'}{t}")
    return line, ok


def report(label, arr, rec, nbytes, eb, t_comp=None, t_dec=None, *, slack=0.0) -> bool:
    line, ok = report_line(label, arr, rec, nbytes, eb, t_comp, t_dec, slack=slack)
    print(line)
    return ok


def sz3_skipped(tag: str) -> None:
    print(f"  [{tag:6s}] skipped: SZ3 is not available")


def parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("input", help="tensor file (.npy or .pt/.pth)")
    add_common(ap)
    ap._option_string_actions["--predictor"].choices = ("gnn", "interp", "interp-linear")
    ap.add_argument("--gnn-checkpoint", default=DEFAULT_CKPT,
                    help="GNN checkpoint (.pt) for --predictor gnn")
    ap.add_argument("--chunk-size", type=int, default=None,
                    help="gnn only: force the chunk edge (a multiple of 2 ** levels)")
    ap.add_argument("--fp16", action="store_true",
                    help="gnn only: recorded in the stream; the kernels' numerics come "
                         "from --numerics")
    ap.add_argument("--numerics", choices=("bf16", "fp32"), default="bf16",
                    help="gnn only: BF16 MFMA or FP32 model kernels")
    ap.add_argument("--normalize", action="store_true",
                    help="min-max scale the tensor to [0,1] before compressing")
    ap.add_argument("--compile", action="store_true", help="gnn only: accepted, no effect")
    return ap


def main(argv=None) -> bool:
    args = parser().parse_args(argv)
    _argv = sys.argv[1:] if argv is None else list(argv)
    levels_explicit = any(a == "--levels" or a.startswith("--levels=") for a in _argv)
    arr = load_tensor(args.input)
    orig_bytes = arr.nbytes
    if arr.dtype == np.float64:
        print("note: float64 input cast to float32 (float32-precision reconstruction)")
        arr = arr.astype(np.float32)
    if args.normalize:
        lo, hi = float(arr.min()), float(arr.max())
        arr = (arr.astype(np.float32) - lo) / max(hi - lo, 1e-12)
        print(f"normalized [{lo:.4g},{hi:.4g}] -> [0,1]")
    eb = args.eb
    if args.rel:
        eb = args.eb * max(float(arr.max()) - float(arr.min()), 1e-12)
    slack = 0.0
    if args.predictor == "gnn":
        from .gnn_codec import GNNCodec, roundtrip_slack

        span = max(float(arr.max()) - float(arr.min()), 1e-12)
        codec = GNNCodec(args.gnn_checkpoint, numerics=args.numerics)
        t0 = time.time()
        stream = codec.compress(arr, error_bound=eb / span,
                                levels=args.levels if levels_explicit else "auto",
                                radius=args.radius, zstd_level=args.zstd_level,
                                eb_ratio=args.eb_ratio,
                                tune=args.tune if args.tune in ("fast", "size") else "fast",
                                chunk_size=args.chunk_size, fp16=args.fp16,
                                compile=args.compile)
        t_comp = time.time() - t0
        outliers = int(codec.last_stats.get("outliers", 0))
        t0 = time.time()
        rec = codec.uncompress(stream).numpy()
        t_dec = time.time() - t0
        slack = roundtrip_slack(float(arr.min()), float(arr.max()))
    else:
        from .codec import decompress

        t0 = time.time()
        stream, stats = run_compress(arr, args)
        t_comp = time.time() - t0
        outliers = int(stats["outliers"])
        t0 = time.time()
        rec = decompress(stream, lambda hdr: build_predictor(args, hdr))
        t_dec = time.time() - t0
    print(f"tensor {args.input} {arr.shape} {arr.dtype}, eb={eb} (orig {orig_bytes} B)")
    ok = report(args.predictor, arr, rec, len(stream), eb, t_comp, t_dec, slack=slack)
    print(f"  ({args.predictor}: outliers {outliers} = {100 * outliers / arr.size:.3f}%)")
    if args.predictor == "gnn":
        from .codec import decompress

        cap = int(os.environ.get("LATEC_BASELINE_MAXVOX", 1 << 24))
        sub = cap_crop(arr, args.anchor_stride, cap) if cap else arr
        cropped = sub.shape != arr.shape
        bargs = copy.copy(args)
        bargs.predictor = "interp"
        bargs.rel = False
        bargs.eb = eb
        t0 = time.time()
        b_stream, _ = run_compress(sub, bargs)
        b_tc = time.time() - t0
        t0 = time.time()
        b_rec = decompress(b_stream, lambda hdr: build_predictor(bargs, hdr))
        report(f"interp{list(sub.shape) if cropped else ''}", sub, b_rec, len(b_stream), eb,
               b_tc, time.time() - t0)
        sz3_skipped("sz3")
    else:
        sz3_skipped("sz3")
    return ok


if __name__ == "__main__":
    sys.exit(0 if main() else 1)
