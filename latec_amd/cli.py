"""LATEC command line: compress / decompress / eval (latec/cli.py:23-247).

    python -m latec_amd.cli compress photo.png photo.msz --eb 2
    python -m latec_amd.cli decompress photo.msz rec.png
    python -m latec_amd.cli eval photo.png --eb 2 --levels 3

The reference's commands, flags and report lines.  Inputs are images read
with PIL (grayscale or RGB, alpha dropped with a warning) and, as an
addition, raw tensors (``.npy``, or a single tensor in ``.pt``/``.pth`` read
with ``weights_only=True``); ``decompress`` writes an image for image
outputs (``.png`` and friends) and a ``.npy`` array otherwise.  The
``zstd-raw`` baseline uses the library's libzstd; the SZ3 baseline needs an
external binary that is not part of this package, so it is reported skipped.
"""

from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

TENSOR_EXT = (".npy", ".pt", ".pth")


def load_image(path: str) -> np.ndarray:
    from PIL import Image

    im = Image.open(path)
    if im.mode in ("L", "I;16"):
        return np.asarray(im.convert("L"))
    if im.mode != "RGB":
        if "A" in im.mode:
            print(f"warning: dropping alpha channel of {im.mode} image", file=sys.stderr)
        im = im.convert("RGB")
    return np.asarray(im)


def load_tensor(path: str) -> np.ndarray:
    if path.endswith(".npy"):
        return np.load(path)
    if path.endswith((".pt", ".pth")):
        import torch

        obj = torch.load(path, map_location="cpu", weights_only=True)
        if not isinstance(obj, torch.Tensor):
            raise ValueError(f"{path} holds {type(obj).__name__}, expected a single tensor")
        return obj.detach().cpu().numpy()
    raise ValueError(f"unsupported extension: {path} (use .npy/.pt/.pth)")


def load_input(path: str) -> np.ndarray:
    return load_tensor(path) if path.endswith(TENSOR_EXT) else load_image(path)


def save_output(path: str, arr: np.ndarray) -> None:
    if path.endswith(".npy"):
        np.save(path, arr)
        return
    from PIL import Image

    Image.fromarray(arr).save(path)


def build_predictor(args, header):
    """Decompress-side predictor; every parameter comes from the stream header."""
    from .bitstream import FLAG_CUBIC, FLAG_INTERP
    from .predictor import InterpPredictor

    if not header.flags & FLAG_INTERP:
        raise ValueError("stream needs a predictor factory the CLI can't build")
    return InterpPredictor("cubic" if header.flags & FLAG_CUBIC else "linear", header.levels,
                           header.anchor_stride, header.anchor_block)


def add_common(ap) -> None:
    ap.add_argument("--eb", type=float, required=True,
                    help="error bound in original data units (e.g. 0-255 for uint8)")
    ap.add_argument("--rel", action="store_true",
                    help="interpret --eb as relative to the data range")
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--anchor-stride", type=int, default=16)
    ap.add_argument("--anchor-block", type=int, default=1)
    ap.add_argument("--radius", type=int, default=1 << 15)
    ap.add_argument("--zstd-level", type=int, default=9)
    ap.add_argument("--eb-ratio", type=float, default=None,
                    help="per-level eb decay; omit to use --tune")
    ap.add_argument("--tune", choices=("fast", "size", "rd"), default="fast",
                    help="auto search when --eb-ratio is omitted")
    ap.add_argument("--tune-size-slack", type=float, default=1.05,
                    help="for --tune rd, may spend up to this size factor for lower "
                         "reconstruction SSE")
    ap.add_argument("--predictor", choices=("interp", "interp-linear"), default="interp",
                    help="interp = SZ-style cubic interpolation (default); "
                         "interp-linear = its linear variant")
    ap.add_argument("-v", "--verbose", action="store_true")


def absolute_eb(img: np.ndarray, args) -> float:
    if not args.rel:
        return float(args.eb)
    span = float(img.max()) - float(img.min())
    return float(args.eb) * (span if span > 0 else 1.0)


def run_compress(img: np.ndarray, args):
    from .codec import compress
    from .predictor import InterpPredictor

    order = "linear" if args.predictor == "interp-linear" else "cubic"
    predictor = InterpPredictor(order, args.levels, args.anchor_stride, args.anchor_block)
    return compress(img, absolute_eb(img, args), predictor, levels=args.levels,
                    anchor_stride=args.anchor_stride, anchor_block=args.anchor_block,
                    radius=args.radius, zstd_level=args.zstd_level, eb_ratio=args.eb_ratio,
                    tune=args.tune, tune_size_slack=args.tune_size_slack,
                    verbose=args.verbose)


def _plane(img: np.ndarray) -> int:
    """Pixels per channel plane of an image (bits per pixel divide by it)."""
    return int(img.shape[0] * (img.shape[1] if img.ndim > 1 else 1))


def cmd_compress(args) -> None:
    img = load_input(args.input)
    t0 = time.time()
    stream, stats = run_compress(img, args)
    Path(args.output).write_bytes(stream)
    bpp = 8 * len(stream) / _plane(img)
    print(f"{args.input}: {stats['original_bytes']} -> {len(stream)} bytes "
          f"(ratio {stats['ratio']:.2f}, {bpp:.3f} bpp) in {time.time() - t0:.1f}s")
    print(f"  predict {stats['predict_s']:.1f}s | quantize {stats['quantize_s']:.1f}s "
          f"| entropy {stats['entropy_s']:.1f}s | outliers {stats['outliers']}")


def cmd_decompress(args) -> None:
    from .codec import decompress

    stream = Path(args.input).read_bytes()
    t0 = time.time()
    out = decompress(stream, lambda hdr: build_predictor(args, hdr))
    save_output(args.output, out)
    print(f"{args.input} -> {args.output} {out.shape} in {time.time() - t0:.1f}s")


def zstd_raw_bytes(arr: np.ndarray, level: int) -> int:
    from .bitstream import zstd_compress

    return len(zstd_compress(np.ascontiguousarray(arr).tobytes(), level))


def cmd_eval(args) -> None:
    from .codec import decompress

    img = load_input(args.input)
    eb = args.eb if not args.rel else args.eb * max(float(img.max()) - float(img.min()), 1.0)
    t0 = time.time()
    stream, stats = run_compress(img, args)
    t_comp = time.time() - t0
    t0 = time.time()
    rec = decompress(stream, lambda hdr: build_predictor(args, hdr))
    t_dec = time.time() - t0
    rec2d = rec if rec.ndim == img.ndim else rec[..., None]
    diff = img.astype(np.float64) - rec2d.astype(np.float64)
    max_err = float(np.abs(diff).max())
    mse = float(np.mean(diff ** 2))
    peak = 255.0 if img.dtype == np.uint8 else float(img.max()) - float(img.min())
    psnr = 10 * np.log10(peak ** 2 / mse) if mse > 0 else float("inf")
    bpp = 8 * len(stream) / _plane(img)
    zstd_raw = zstd_raw_bytes(img, args.zstd_level)
    bound_ok = max_err <= eb
    print(f"image {args.input} {img.shape} {img.dtype}, eb={eb}")
    print(f"  compressed:  {len(stream)} bytes, ratio {stats['ratio']:.2f}, {bpp:.3f} bpp")
    print(f"  zstd-raw:    {zstd_raw} bytes, ratio {img.nbytes / zstd_raw:.2f}")
    print("  sz3:         (sz3 binary not found, baseline skipped)")
    print(f"  PSNR:        {psnr:.2f} dB")
    print(f"  max abs err: {max_err} <= eb: {'PASS' if bound_ok else 'FAIL'}")
    print(f"  outliers:    {stats['outliers']} "
          f"({100 * stats['outliers'] / img.size:.3f}% of values)")
    print(f"  time:        compress {t_comp:.1f}s, decompress {t_dec:.1f}s "
          f"(predict {stats['predict_s']:.1f}s)")
    if not bound_ok:
        raise SystemExit(1)


def parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="latec", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("compress", help="compress an image to a .msz stream")
    p.add_argument("input")
    p.add_argument("output")
    add_common(p)
    p.set_defaults(fn=cmd_compress)
    p = sub.add_parser("decompress", help="decompress a .msz stream to an image")
    p.add_argument("input")
    p.add_argument("output")
    p.set_defaults(fn=cmd_decompress)
    p = sub.add_parser("eval", help="in-memory roundtrip with metrics")
    p.add_argument("input")
    add_common(p)
    p.set_defaults(fn=cmd_eval)
    return ap


def main(argv=None) -> None:
    args = parser().parse_args(argv)
    args.fn(args)


if __name__ == "__main__":
    main()
