"""One GNN round trip on a centred n-D subset of a tensor, with the report of
scripts/bench_gnn_subset.py:411-510.

    python -m latec_amd.bench_subset field.npy --eb 1e-4 --normalize \\
        --levels 3 --subset-edge 64 --chunk-size 32

The subset edge is capped per axis by the tensor and floored to a multiple of
2 ** levels.  Reported: PSNR, max error (PASS/FAIL against the bound plus the
float32 slack of the reference, 4 * eps * max|x|), bits/value and ratio,
compress and decompress wall time with the voxel rate, peak host RSS during
the round trip, and peak device memory that torch allocated (every device
buffer of the codec is a torch tensor).  Exit status 1 when the bound fails.
"""

from __future__ import annotations

import argparse
import os
import threading
import time

import numpy as np

from .cli import load_tensor
from .eval_tensor import DEFAULT_CKPT


def centred_subset(arr: np.ndarray, edge: int, stride: int) -> np.ndarray:
    """Centred crop, per axis min(edge, n) floored to a multiple of ``stride``
    (never below it), materialised contiguous."""
    lens = []
    for d in arr.shape:
        e = min(edge, d)
        e -= e % stride
        lens.append(max(min(e if e else stride, d), 1))
    sl = tuple(slice((d - e) // 2, (d - e) // 2 + e) for d, e in zip(arr.shape, lens))
    return np.ascontiguousarray(arr[sl])


class HostPeak:
    """Peak resident set size of this process, sampled while active."""

    def __init__(self, interval: float = 0.02):
        import psutil

        self.proc, self.interval = psutil.Process(os.getpid()), interval
        self.base = self.peak = self.proc.memory_info().rss
        self._stop = threading.Event()

    def _run(self):
        while not self._stop.is_set():
            self.peak = max(self.peak, self.proc.memory_info().rss)
            time.sleep(self.interval)

    def __enter__(self):
        self._t = threading.Thread(target=self._run, daemon=True)
        self._t.start()
        return self

    def __exit__(self, *exc):
        self._stop.set()
        self._t.join()
        self.peak = max(self.peak, self.proc.memory_info().rss)


def run(arr, eb, *, levels, chunk_size=None, numerics="bf16", ckpt=DEFAULT_CKPT,
        tune="fast", gate=True, gate_fine=False) -> dict:
    import torch

    from .gnn_codec import GNNCodec

    codec = GNNCodec(ckpt, numerics=numerics)
    torch.cuda.synchronize()
    base_alloc = torch.cuda.memory_allocated()
    base_resv = torch.cuda.memory_reserved()
    torch.cuda.reset_peak_memory_stats()
    with HostPeak() as host:
        t0 = time.perf_counter()
        stream = codec.compress(arr, error_bound=eb / max(float(arr.max()) - float(arr.min()),
                                                          1e-12),
                                levels=levels, chunk_size=chunk_size, tune=tune, gate=gate,
                                gate_fine=gate_fine)
        torch.cuda.synchronize()
        t_comp = time.perf_counter() - t0
        t0 = time.perf_counter()
        rec = codec.uncompress(stream).numpy().reshape(arr.shape)
        torch.cuda.synchronize()
        t_dec = time.perf_counter() - t0
    a, r = arr.astype(np.float64), rec.astype(np.float64)
    max_err = float(np.abs(a - r).max())
    slack = 4 * np.finfo(np.float32).eps * max(float(np.abs(arr).max()), 1.0)
    mse = float(np.mean((a - r) ** 2))
    peak = max(float(arr.max()) - float(arr.min()), 1e-12)
    return {"psnr": 10 * np.log10(peak ** 2 / mse) if mse > 0 else float("inf"),
            "max_err": max_err, "error_ok": max_err <= eb + slack, "bytes": len(stream),
            "bpv": 8 * len(stream) / arr.size, "ratio": arr.nbytes / len(stream),
            "t_comp": t_comp, "t_dec": t_dec, "mvox_s": arr.size / t_comp / 1e6,
            "host_peak_mib": host.peak / 2 ** 20,
            "host_increase_mib": (host.peak - host.base) / 2 ** 20,
            "gpu_peak_alloc_mib": torch.cuda.max_memory_allocated() / 2 ** 20,
            "gpu_peak_resv_mib": torch.cuda.max_memory_reserved() / 2 ** 20,
            "gpu_increase_alloc_mib": (torch.cuda.max_memory_allocated() - base_alloc) / 2 ** 20,
            "gpu_increase_resv_mib": (torch.cuda.max_memory_reserved() - base_resv) / 2 ** 20}


def format_report(args, sub, eb, m, label="") -> str:
    w = 24
    tag = f" [{label}]" if label else ""
    lines = [f"\n=== GNN subset benchmark{tag} ===",
             f"{'input':<{w}} {args.input} {tuple(sub.shape)} ({sub.size} voxels, "
             f"{sub.nbytes} B)",
             f"{'eb':<{w}} {eb:g}   device {args.device}",
             f"{'levels/stride':<{w}} {args.levels}/{1 << args.levels}   chunk "
             f"{args.chunk_size} tune {args.tune} numerics {args.numerics} gate "
             f"{not args.no_gate} gate_fine {args.gate_fine}",
             "-" * 60,
             f"{'PSNR':<{w}} {m['psnr']:8.3f} dB",
             f"{'max error':<{w}} {m['max_err']:.4g}  ({'PASS' if m['error_ok'] else 'FAIL'} "
             f"vs eb {eb:g})",
             f"{'bits/value':<{w}} {m['bpv']:8.4f} bpv   ratio {m['ratio']:.2f}x  "
             f"({m['bytes']} B)",
             f"{'compress':<{w}} {m['t_comp']:8.3f} s   ({m['mvox_s']:.2f} Mvox/s)",
             f"{'decompress':<{w}} {m['t_dec']:8.3f} s",
             f"{'peak host RAM':<{w}} {m['host_peak_mib']:8.1f} MiB  "
             f"(+{m['host_increase_mib']:.1f} MiB during roundtrip)",
             f"{'peak GPU alloc/resv':<{w}} {m['gpu_peak_alloc_mib']:8.1f} / "
             f"{m['gpu_peak_resv_mib']:.1f} MiB  (+{m['gpu_increase_alloc_mib']:.1f} / "
             f"+{m['gpu_increase_resv_mib']:.1f} MiB during roundtrip)",
             f"{'GPU util':<{w}} (no utilisation samples)"]
    return "\n".join(lines)


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("input", help="tensor file (.npy or .pt/.pth)")
    ap.add_argument("--eb", type=float, required=True)
    ap.add_argument("--rel", action="store_true")
    ap.add_argument("--normalize", action="store_true")
    ap.add_argument("--levels", type=int, default=3)
    ap.add_argument("--subset-edge", type=int, default=64)
    ap.add_argument("--chunk-size", type=int, default=None)
    ap.add_argument("--tune", choices=("fast", "size"), default="fast")
    ap.add_argument("--numerics", choices=("bf16", "fp32"), default="bf16")
    ap.add_argument("--no-gate", action="store_true")
    ap.add_argument("--gate-fine", action="store_true")
    ap.add_argument("--gnn-checkpoint", default=DEFAULT_CKPT)
    ap.add_argument("--device", default="cuda:0")
    ap.add_argument("--label", default="")
    args = ap.parse_args(argv)
    arr = load_tensor(args.input)
    sub = centred_subset(arr, args.subset_edge, 1 << args.levels).astype(np.float32)
    if args.normalize:
        lo, hi = float(sub.min()), float(sub.max())
        sub = (sub - lo) / max(hi - lo, 1e-12)
    eb = args.eb * (max(float(sub.max()) - float(sub.min()), 1e-12) if args.rel else 1.0)
    m = run(sub, eb, levels=args.levels, chunk_size=args.chunk_size, numerics=args.numerics,
            ckpt=args.gnn_checkpoint, tune=args.tune, gate=not args.no_gate,
            gate_fine=args.gate_fine)
    print(format_report(args, sub, eb, m, args.label))
    if not m["error_ok"]:
        raise SystemExit(1)


if __name__ == "__main__":
    main()
