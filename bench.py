"""Benchmark: LATEC plain interpolation codec on one MI355X.

Workload (one "step" = one full compress of the field, inputs resident in
HBM): 512^3 float32 field = 0.5 GRF(l=24) + 0.5 k^-5/3 turbulence,
InterpPredictor cubic, levels 5, anchor stride 32, anchor_block 1,
eb = 1e-3 * (max - min), tune fast (eb_ratio 0.9, centre 0, END_QUAD),
LATEC001 v4 stream (GPU fused stage kernels + GPU fit/tables + host rANS +
zstd-9 frame).  Decompress is timed the same way.

Prints one JSON line: value = compress GB/s of input (end to end, 4 B per
value), plus decompress GB/s, bits/value, the HBM roofline fraction of the
fused stage kernels (HIP events over the timed region) and the CPU oracle
timed on a bounded sample of the same field.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "encode_c2_512cubed_hbm_traffic.md")


def profiled_traffic(shape):
    """HBM bytes of one encode (all stage launches) from the committed
    rocprofv3 FETCH_SIZE + WRITE_SIZE summary of the 512^3 workload, or None
    for another shape or when the summary is absent."""
    if tuple(shape) != (512, 512, 512) or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("traffic_bytes_per_encode:"):
            return int(line.split(":")[1])
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--shape", type=str, default="512,512,512")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-planes", type=int, default=64)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain 1 GiB device copy (counter calibration)")
    ap.add_argument("--skip-decode", action="store_true")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this "
              "build (the codec path runs on a single MI355X).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from latec_amd import _device as D
    from latec_amd import codec
    from latec_amd.bitstream import Header, read_stream, write_stream
    from latec_amd.levels import stage_ebs
    from latec_amd.predictor import InterpPredictor, get_plan
    from latec_amd.synthetic import config2_field

    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    shape = tuple(int(s) for s in args.shape.split(","))
    levels, stride = 5, 32
    dev = D.device()
    print(f"[bench] generating {shape} field on {dev}", file=sys.stderr, flush=True)
    xd = config2_field(shape=shape, device=dev).reshape(-1).contiguous()
    vmin, vmax = float(xd.min()), float(xd.max())
    eb = 1e-3 * (vmax - vmin)
    n_values = xd.numel()
    in_bytes = 4 * n_values
    pred = InterpPredictor("cubic", levels, stride, 1)
    pred.center, pred.end_mode = pred.resolved(len(shape))
    plan = get_plan(shape, 1, levels, stride, 1)
    ebs = stage_ebs(shape, levels, stride, 1, eb, pred.fast_eb_ratio)
    o = codec._interp_opts(pred, len(shape), 1 << 15, False)
    header = Header(channels=1, src_dtype=1, spatial=shape, eb=eb, levels=levels,
                    anchor_stride=stride, anchor_block=1, radius=1 << 15, max_radius=0,
                    agg_level=-1, vmin=vmin, vmax=vmax,
                    flags=pred.stream_flag | (pred.end_mode << 9), interp_center=pred.center,
                    eb_ratio=pred.fast_eb_ratio)
    ws = codec.workspace(plan, o.radius)

    def compress_once():
        res = codec.encode_device(xd, plan, ebs, o, graph=True, ws=ws)
        fits = codec.fit_device(res)
        payload, _, n_out = codec.pack_device(res, fits, 9)
        return write_stream(header, payload, 9), int(n_out.sum())

    def decompress_once(stream):
        h, payload = read_stream(stream)
        d = codec.decode_device(payload, plan, ebs, o)
        torch.cuda.synchronize()
        return d

    for _ in range(args.warmup):
        stream, n_out = compress_once()
        if not args.skip_decode:
            decompress_once(stream)
    torch.cuda.synchronize()

    # end-to-end compress
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stream, n_out = compress_once()
    torch.cuda.synchronize()
    t_comp = (time.perf_counter() - t0) / args.steps
    print(f"[bench] compress {t_comp * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(0 if args.skip_decode else args.steps):
        recon = decompress_once(stream)
    torch.cuda.synchronize()
    t_dec = (time.perf_counter() - t0) / args.steps
    if args.skip_decode:
        recon, t_dec = ws.recon, float("nan")
    max_err = float((recon - xd).abs().max())
    print(f"[bench] decompress {t_dec * 1e3:.1f} ms/step, max err {max_err:.3e} <= {eb:.3e}",
          file=sys.stderr, flush=True)
    assert max_err <= eb

    # kernel time: HIP events around the fused stage kernels (graph replay)
    stream_t = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    k_enc, k_dec = [], []
    import ctypes

    from latec_amd._lib import check, lib, ptr
    ebs_arr = np.ascontiguousarray(ebs, np.float64)
    optr = ctypes.cast(ctypes.pointer(o), ctypes.c_void_p)
    for _ in range(args.steps):
        ws.hist.zero_()
        ws.cnt.zero_()
        ev[0].record(stream_t)
        check(lib().latec_interp_encode_graph(plan.handle, ptr(xd), ptr(ws.recon), ptr(ws.codes),
                                              ptr(ws.hist), ptr(ws.rank), ptr(ws.val),
                                              ptr(ws.cnt), ptr(ebs_arr), optr, ptr(None)))
        ev[1].record(stream_t)
        torch.cuda.synchronize()
        k_enc.append(ev[0].elapsed_time(ev[1]) * 1e-3)
    # decode kernels on the decoded codes of the last stream
    h, payload = read_stream(stream)
    if not args.skip_decode:
        codec.decode_device(payload, plan, ebs, o)
    for _ in range(0 if args.skip_decode else args.steps):
        ev[2].record(stream_t)
        check(lib().latec_interp_decode(plan.handle, 0, plan.n_stages, ptr(ws.codes),
                                        ptr(ws.rank), ptr(ws.val), ptr(ws.cnt), ptr(ws.recon),
                                        ptr(ebs_arr), optr, ptr(None)))
        ev[3].record(stream_t)
        torch.cuda.synchronize()
        k_dec.append(ev[2].elapsed_time(ev[3]) * 1e-3)
    t_kenc = float(np.mean(k_enc))
    t_kdec = float(np.mean(k_dec)) if k_dec else float("nan")
    enc_bytes = 10 * n_values + 4 * n_out
    dec_bytes = 6 * n_values + 4 * n_out
    ach_enc = enc_bytes / t_kenc
    ach_dec = dec_bytes / t_kdec
    print(f"[bench] encode kernels {t_kenc * 1e3:.3f} ms ({ach_enc / 1e9:.0f} GB/s), "
          f"decode kernels {t_kdec * 1e3:.3f} ms ({ach_dec / 1e9:.0f} GB/s)",
          file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        planes = min(args.cpu_planes, shape[0])
        sample = D.to_host(xd.reshape(shape)[:planes].contiguous())
        s_eb = 1e-3 * float(sample.max() - sample.min())
        s_ebs = O.stage_ebs_ref(len(shape), levels, stride, s_eb, 0.9)
        t0 = time.perf_counter()
        O.compress_payload(sample[None], s_ebs, levels, stride, 1, "cubic", 0, 1)
        dt = time.perf_counter() - t0
        cpu = {"value": sample.nbytes / dt / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"first {planes} planes of the same field ({sample.size} values), "
                         "one compress (payload, without the outer zstd frame), C oracle"}
        print(f"[bench] cpu oracle {dt:.1f} s on {sample.size} values", file=sys.stderr,
              flush=True)

    line = {
        "metric": "compress/decompress GB/s + bits/value at fixed eb: GNN-BF16 vs plain interp",
        "value": in_bytes / t_comp / 1e9,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_comp * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain interpolation, 512^3 f32 0.5 GRF(l=24) + 0.5 k^-5/3 "
                               "turbulence, cubic L5 stride 32 block 1, eb=1e-3*range, "
                               "LATEC001 v4 stream",
                   "shape": list(shape), "eb": eb},
        "decompress_GBps": in_bytes / t_dec / 1e9,
        "bits_per_value": 8 * len(stream) / n_values,
        "stream_bytes": len(stream),
        "max_abs_err": max_err,
        "kernel_ms": {"encode": t_kenc * 1e3, "decode": t_kdec * 1e3},
        "roofline": {"bound": "hbm", "achieved": ach_enc / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": ach_enc / HBM_PEAK,
                     "traffic": profiled_traffic(shape), "algorithmic_bytes": enc_bytes,
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT),
                     "kernel": "one encode: encode_stage_kernel x26 + lattice_kernel x8 "
                               "(10 B/value + 4 B/outlier)",
                     "decode": {"achieved": ach_dec / 1e9, "frac": ach_dec / HBM_PEAK}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
