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
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="threads of the all-cores CPU baseline (the box's CPU share is 16)")
    ap.add_argument("--no-api", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain 1 GiB device copy (counter calibration)")
    ap.add_argument("--skip-decode", action="store_true")
    ap.add_argument("--no-gnn", action="store_true")
    ap.add_argument("--gnn-ckpt", default=os.path.join(ROOT, "latec_amd", "checkpoints",
                                                       "gnn_d64_agg2.pt"))
    ap.add_argument("--gnn-shape", default="64,64,64,64")
    ap.add_argument("--gnn-levels", type=int, default=3)
    ap.add_argument("--gnn-steps", type=int, default=2)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--knob", action="append", default=[],
                    help="name=value launch-shape knob (latec_tune_set), skips the autotuner")
    ap.add_argument("--batch-n", type=int, default=256)
    ap.add_argument("--batch-scale", type=float, default=1.0)
    return ap.parse_args()


def batch_leg(args):
    """Batched uneven workload: ``--batch-n`` fields (half 2-D sides
    U[64,1024], half 3-D sides U[32,256], 70 % f32 / 30 % f64, smooth /
    turbulent / discontinuous), relative eb 1e-3, cubic levels 4 stride 16,
    stream version 5: one persistent launch per schedule step in a hipGraph.
    Compared with coding the same fields one ``compress`` call at a time."""
    import numpy as np
    import torch

    import latec_amd
    from latec_amd.batch import compress_batch, decompress_batch
    from latec_amd.predictor import InterpPredictor
    from latec_amd.synthetic import batch_fields

    def note(msg):
        print(f"[bench] batch: {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    fields = batch_fields(args.batch_n, seed=4, scale=args.batch_scale, device="cuda")
    n_values = sum(f.size for f in fields)
    in_bytes = sum(f.astype(np.float32).nbytes for f in fields)
    note(f"generated {len(fields)} fields, {n_values / 1e6:.1f} M values "
         f"({time.perf_counter() - t0:.1f} s)")
    kw = dict(levels=4, anchor_stride=16)
    t0 = time.perf_counter()
    # one untimed step of the whole batch (plans, tables, buffers, kernels)
    decompress_batch(compress_batch(fields, 1e-3, **kw))
    torch.cuda.synchronize()
    note(f"warm-up {time.perf_counter() - t0:.1f} s")
    tm = {}
    t0 = time.perf_counter()
    streams = compress_batch(fields, 1e-3, timings=tm, **kw)
    t_c = time.perf_counter() - t0
    note(f"compress {t_c:.2f} s {tm}")
    td = {}
    t0 = time.perf_counter()
    outs = decompress_batch(streams, timings=td)
    t_d = time.perf_counter() - t0
    note(f"decompress {t_d:.2f} s {td}")
    worst = 0.0
    for x, y, s in zip(fields, outs, streams):
        x32 = x.astype(np.float32)
        eb = 1e-3 * (float(x32.max()) - float(x32.min()))
        err = float(np.abs(y - x32).max())
        assert err <= eb
        worst = max(worst, err / eb)
    # the same fields one compress call at a time (first 32)
    k = min(32, len(fields))
    t0 = time.perf_counter()
    for x in fields[:k]:
        x32 = x.astype(np.float32)
        eb = 1e-3 * (float(x32.max()) - float(x32.min()))
        s, _ = latec_amd.compress(x32, eb, InterpPredictor("cubic", 4, 16, 1), anchor_block=1,
                                  stream_version=5, **kw)
    t_one = time.perf_counter() - t0
    in_k = sum(f.astype(np.float32).nbytes for f in fields[:k])
    # per-field GNN-vs-interpolation choice (the LATECGNN stream picks GNN or
    # interpolation per stage) on up to 16 fields of at most 4 M values
    sel = None
    if not args.no_gnn:
        from latec_amd.gnn_codec import GNNCodec, roundtrip_slack

        idx = [i for i, f in enumerate(fields) if f.size <= (4 << 20)][:16]
        sub = [fields[i] for i in idx]
        codec = GNNCodec(args.gnn_ckpt)
        compress_batch(sub, 1e-3, gnn=codec, **kw)  # untimed: model handles, tables
        ts = {}
        t0 = time.perf_counter()
        mixed = compress_batch(sub, 1e-3, gnn=codec, timings=ts, **kw)
        t_sel = time.perf_counter() - t0
        plain_bits = 8 * sum(len(streams[i]) for i in idx)
        t0 = time.perf_counter()
        outs2 = decompress_batch(mixed, gnn=codec)
        t_sel_dec = time.perf_counter() - t0
        for x, y, m in zip(sub, outs2, mixed):
            x32 = x.astype(np.float32)
            eb = 1e-3 * (float(x32.max()) - float(x32.min()))
            slack = (roundtrip_slack(float(x32.min()), float(x32.max()))
                     if m[:8] == b"LATECGNN" else 0.0)
            assert float(np.abs(np.asarray(y, np.float64) - x32).max()) <= eb + slack
        n_sub = sum(f.size for f in sub)
        sel = {"fields": len(sub), "values": int(n_sub), "gnn_streams": ts["gnn_streams"],
               "bits_per_value_interp": plain_bits / n_sub,
               "bits_per_value_selected": 8 * sum(len(m) for m in mixed) / n_sub,
               "compress_s": t_sel, "gnn_s": ts["gnn_s"], "decompress_s": t_sel_dec}
        note(f"GNN selection on {len(sub)} fields: {ts['gnn_streams']} LATECGNN, "
             f"{sel['bits_per_value_interp']:.4f} -> {sel['bits_per_value_selected']:.4f} "
             f"bits/value ({t_sel:.1f} s)")
    out = {"workload": f"{len(fields)} fields (half 2-D sides U[64,1024], half 3-D sides "
                       f"U[32,256], x{args.batch_scale:g}), 70% f32 / 30% f64, relative eb 1e-3, "
                       "cubic L4 stride 16, stream version 5",
           "fields": len(fields), "values": int(n_values),
           "compress_GBps": in_bytes / t_c / 1e9, "decompress_GBps": in_bytes / t_d / 1e9,
           "compress_s": t_c, "decompress_s": t_d, "kernels_s": tm["kernels_s"],
           "fit_s": tm["fit_s"], "entropy_s": tm["entropy_s"], "setup_s": tm["setup_s"],
           "decode_kernels_s": td["kernels_s"],
           "bits_per_value": 8 * sum(len(s) for s in streams) / n_values,
           "max_err_over_eb": worst,
           "one_at_a_time_GBps_first32": in_k / t_one / 1e9,
           "gnn_selection": sel}
    print(f"[bench] batch: {len(fields)} fields, {n_values / 1e6:.1f} M values, compress "
          f"{t_c:.2f} s ({out['compress_GBps']:.2f} GB/s, kernels {tm['kernels_s'] * 1e3:.1f} ms), "
          f"decompress {t_d:.2f} s; one at a time {out['one_at_a_time_GBps_first32']:.2f} GB/s",
          file=sys.stderr, flush=True)
    return out


def host_cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_all_cores(xd, shape, levels, stride, threads):
    """The C oracle on every slab of the field at once: the field cut into
    ``threads`` slabs along axis 0, each compressed (payload, no outer frame)
    on its own thread (the oracle releases the GIL), so the rate is what the
    host's cores give for the same algorithm on independent slabs."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    from latec_amd import _device as D
    from oracle import oracle as O

    nproc = os.cpu_count() or 1
    t = max(1, min(threads, nproc, shape[0]))
    bounds = [shape[0] * i // t for i in range(t + 1)]
    host = D.to_host(xd.reshape(shape))
    slabs = [np.ascontiguousarray(host[a:b])[None] for a, b in zip(bounds, bounds[1:])]
    del host

    def one(sl):
        eb = 1e-3 * float(sl.max() - sl.min())
        O.compress_payload(sl, O.stage_ebs_ref(len(shape), levels, stride, eb, 0.9), levels,
                           stride, 1, "cubic", 0, 1)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=t) as ex:
        list(ex.map(one, slabs))
    dt = time.perf_counter() - t0
    nbytes = sum(sl.nbytes for sl in slabs)
    print(f"[bench] cpu oracle, {t} threads: {dt:.1f} s on {nbytes // 4} values",
          file=sys.stderr, flush=True)
    return {"value": nbytes / dt / 1e9, "unit": "GB/s", "cores": t, "nproc": nproc,
            "cpu_model": host_cpu_model(),
            "sample": f"the whole field as {t} slabs of ~{shape[0] // t} planes, one C-oracle "
                      "compress (payload, without the outer zstd frame) per slab, all slabs at "
                      "once on separate threads"}


def api_leg(xd, shape, levels, stride, eb, steps):
    """latec_amd.compress / decompress on a host array (the reference's call):
    input upload, encode, stage moments, fit, pack, zstd frame, reconstruction
    back to the host, and the decode from stream bytes to a host array."""
    import numpy as np

    import latec_amd
    from latec_amd import InterpPredictor
    from latec_amd import _device as D

    x = D.to_host(xd.reshape(shape))
    pred = InterpPredictor("cubic", levels, stride, 1)
    stream, st = latec_amd.compress(x, eb, pred, levels=levels, anchor_stride=stride,
                                    anchor_block=1)
    latec_amd.decompress(stream)
    tc, td = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        stream, st = latec_amd.compress(x, eb, pred, levels=levels, anchor_stride=stride,
                                        anchor_block=1)
        tc.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        y = latec_amd.decompress(stream)
        td.append(time.perf_counter() - t0)
    assert float(np.abs(y - x).max()) <= eb
    out = {"compress_s": float(np.median(tc)), "decompress_s": float(np.median(td)),
           "compress_GBps": x.nbytes / float(np.median(tc)) / 1e9,
           "decompress_GBps": x.nbytes / float(np.median(td)) / 1e9,
           "bits_per_value": 8 * len(stream) / x.size,
           "stats_keys": sorted(k for k in st if k.startswith("stage_"))}
    print(f"[bench] api: compress {out['compress_s']:.3f} s, decompress "
          f"{out['decompress_s']:.3f} s", file=sys.stderr, flush=True)
    return out


def gnn_leg(args):
    """GNN codec on the 4-D anisotropic field (per-axis correlation U[1,8],
    50% turbulence, 3 hyperplane jumps), levels 3, eb = 1e-3 * range: BF16 and
    FP32 kernels against plain cubic interpolation on the same GPU, the BF16 /
    FP32 prediction RMS against an FP64 run of the same network (oracle, on a
    16^4 crop), the reference's 16-chunk partition of the same tensor, and the
    model GEMM throughput."""
    import numpy as np
    import torch

    from latec_amd import InterpPredictor, compress, decompress
    from latec_amd.gnn_codec import GNNCodec, roundtrip_slack
    from latec_amd.synthetic import config3_field

    shape = tuple(int(v) for v in args.gnn_shape.split(","))
    lv = args.gnn_levels
    x = config3_field(3, shape, device="cuda").cpu().numpy()
    rel = 1e-3
    S = float(x.max() - x.min())
    out = {"workload": f"GNN codec, {'x'.join(map(str, shape))} f32 4-D anisotropic GRF + "
                       "turbulence + 3 jumps, levels 3, eb=1e-3*range, one chunk",
           "checkpoint": os.path.relpath(args.gnn_ckpt, ROOT)}
    res = {}
    for numerics in ("bf16", "fp32"):
        codec = GNNCodec(args.gnn_ckpt, numerics=numerics)
        stream = codec.compress(x, rel, levels=lv)  # warm-up (geometry, tables)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.gnn_steps):
            stream = codec.compress(x, rel, levels=lv)
        torch.cuda.synchronize()
        tc = (time.perf_counter() - t0) / args.gnn_steps
        stats = dict(codec.last_stats)
        codec.uncompress(stream)  # warm-up (decoder staging), as for compress
        t0 = time.perf_counter()
        for _ in range(args.gnn_steps):
            y = codec.uncompress(stream).numpy()
        td = (time.perf_counter() - t0) / args.gnn_steps
        err = float(np.abs(y - x).max())
        assert err <= rel * S + roundtrip_slack(x.min(), x.max())
        codec.count_flops = True  # model flops and model time (events), one more run
        codec.compress(x, rel, levels=lv)
        flops = codec.last_stats.get("model_flops", 0)
        stats["model_s"] = codec.last_stats["model_s"]
        res[numerics] = {"bits_per_value": 8 * len(stream) / x.size,
                         "compress_GBps": x.nbytes / tc / 1e9, "decompress_GBps": x.nbytes / td / 1e9,
                         "compress_s": tc, "decompress_s": td, "max_abs_err": err,
                         "model_s": stats["model_s"], "model_gemm_flops": flops,
                         "model_TFLOPs": flops / stats["model_s"] / 1e12,
                         "gates_fired": stats.get("gates_fired"), "outliers": stats["outliers"]}
        print(f"[bench] gnn {numerics}: {res[numerics]['bits_per_value']:.4f} bits/value, "
              f"compress {tc:.2f} s, decompress {td:.2f} s, model {stats['model_s']:.3f} s",
              file=sys.stderr, flush=True)
    # the reference's default partition of this tensor: 16 chunks of 32^4 with
    # halo context (its point budget 2^21; here the automatic budget keeps the
    # tensor as one chunk)
    codec = GNNCodec(args.gnn_ckpt, numerics="bf16")
    codec.uncompress(codec.compress(x, rel, levels=lv, chunk_size=32))  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s_ch = codec.compress(x, rel, levels=lv, chunk_size=32)
    torch.cuda.synchronize()
    t_ch = time.perf_counter() - t0
    t0 = time.perf_counter()
    y = codec.uncompress(s_ch).numpy()
    td_ch = time.perf_counter() - t0
    assert float(np.abs(y - x).max()) <= rel * S + roundtrip_slack(x.min(), x.max())
    out["bf16_16_chunks"] = {"chunk_edge": 32, "bits_per_value": 8 * len(s_ch) / x.size,
                             "compress_s": t_ch, "decompress_s": td_ch,
                             "compress_GBps": x.nbytes / t_ch / 1e9,
                             "decompress_GBps": x.nbytes / td_ch / 1e9}
    print(f"[bench] gnn bf16, 16 chunks: {8 * len(s_ch) / x.size:.4f} bits/value, compress "
          f"{t_ch:.2f} s, decompress {td_ch:.2f} s", file=sys.stderr, flush=True)
    stride = 1 << lv
    t0 = time.perf_counter()
    s_int, _ = compress(x, rel * S, InterpPredictor("cubic", lv, stride, 1), levels=lv,
                        anchor_stride=stride, anchor_block=1)
    t_int = time.perf_counter() - t0
    assert float(np.abs(decompress(s_int) - x).max()) <= rel * S
    out.update({"bf16": res["bf16"], "fp32": res["fp32"],
                "interp_bits_per_value": 8 * len(s_int) / x.size, "interp_compress_s": t_int,
                "bf16_gain_bits_per_value": 8 * len(s_int) / x.size - res["bf16"]["bits_per_value"]})
    # prediction RMS against the FP64 oracle on a crop (teacher-forced schedule)
    from latec_amd.gnn_predictor import ChunkedGNNPredictor
    from oracle import gnn_oracle as G

    crop = tuple(min(16, n) for n in shape)
    xc = x[tuple(slice(0, c) for c in crop)]
    xc = ((xc - xc.min()) / (xc.max() - xc.min())).astype(np.float32)
    from latec_amd import gnn_model as GM

    d, agg, P = GM.load_checkpoint(args.gnn_ckpt)
    Pd = {k: torch.from_numpy(np.asarray(v, np.float64)) for k, v in P.items()}
    ref = G.run_schedule(Pd, torch.from_numpy(xc.reshape(1, -1)).double(), crop, lv, stride,
                         rel, stride, agg)
    ref_mu = np.concatenate([m.numpy()[0] for m, _ in ref])
    for numerics in ("bf16", "fp32"):
        pr = ChunkedGNNPredictor(args.gnn_ckpt, 0.0, 1.0, levels=lv, anchor_stride=stride,
                                 numerics=numerics)
        pr.begin(crop, tuple(-(-n // stride) * stride for n in crop))
        rc = torch.from_numpy(xc).cuda()[None]
        w = pr.start_wave([0], rc)
        mus = []
        for st in w["chain"][1:]:
            _, _, mu, _ = pr.predict_wave_stage(st, rc, rel, want_head=True)
            mus.append(mu.cpu().numpy().astype(np.float64))
        pr.finish_wave(rc)
        out[f"rms_{numerics}_vs_fp64"] = float(np.sqrt(np.mean((np.concatenate(mus) - ref_mu) ** 2)))
    out["rms_crop"] = list(crop)
    return out


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
    tuned = None
    if args.knob:
        from latec_amd import autotune

        autotune.set_knobs(**{k: int(v) for k, v in (kv.split("=") for kv in args.knob)})
        print(f"[bench] knobs: {autotune.get()}", file=sys.stderr, flush=True)
    elif not args.no_autotune and len(shape) == 3:
        # launch shapes picked on this very field (outputs identical for all)
        from latec_amd import autotune

        tuned = autotune.tune(xd, shape, levels, stride, ebs, o, reps=2)
        print(f"[bench] autotune: {tuned['knobs']} ({tuned['ms']:.3f} ms, "
              f"{len(tuned['table'])} candidates)", file=sys.stderr, flush=True)

    def compress_once():
        res = codec.encode_device(xd, plan, ebs, o, graph=True, ws=ws)
        fits = codec.fit_device(res)
        payload, _, n_out, frame = codec.pack_device_framed(res, fits, 9)
        return write_stream(header, payload, 9, frame=frame), int(n_out.sum())

    def decompress_once(stream):
        h, payload = read_stream(stream)
        d = codec.decode_device(payload, plan, ebs, o)
        torch.cuda.synchronize()
        return d

    from latec_amd import mlrans

    header5 = Header(**{**header.__dict__, "version": mlrans.VERSION_MULTILANE})

    def compress_once_v5():
        res = codec.encode_device(xd, plan, ebs, o, graph=True, ws=ws)
        fits = codec.fit_device(res)
        payload, _, _ = mlrans.pack_payload(res, fits)
        return write_stream(header5, payload)

    def decompress_once_v5(stream):
        h, payload = read_stream(stream)
        d = codec.decode_device_v5(payload, plan, ebs, o)
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

    # stream version 5: the same encode, multi-lane rANS on the GPU, no outer zstd
    v5 = None
    if not args.skip_decode:
        recon4 = recon.clone()
        s5 = compress_once_v5()
        decompress_once_v5(s5)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s5 = compress_once_v5()
        torch.cuda.synchronize()
        t_c5 = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r5 = decompress_once_v5(s5)
        torch.cuda.synchronize()
        t_d5 = (time.perf_counter() - t0) / args.steps
        assert torch.equal(r5, recon4), "version-5 decode differs from version 4's"
        v5 = {"compress_GBps": in_bytes / t_c5 / 1e9, "decompress_GBps": in_bytes / t_d5 / 1e9,
              "ms_per_step": t_c5 * 1e3, "decompress_ms": t_d5 * 1e3,
              "bits_per_value": 8 * len(s5) / n_values, "stream_bytes": len(s5)}
        print(f"[bench] v5 multi-lane: compress {t_c5 * 1e3:.1f} ms, decompress "
              f"{t_d5 * 1e3:.1f} ms, {v5['bits_per_value']:.4f} bits/value",
              file=sys.stderr, flush=True)

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
    # which level kernels one encode runs (launch counters of a direct encode)
    check(lib().latec_tune_set(b"sheet_launches", 0))
    codec.encode_device(xd, plan, ebs, o, graph=False, ws=ws)
    torch.cuda.synchronize()
    n_sheet_levels = int(lib().latec_tune_get(b"sheet_launches")) // 2
    n_pyr = max(0, levels - 1)  # levels above the first
    enc_desc = (f"one encode, one hipGraph: x pyramid gather, anchors encode_stage_kernel, "
                f"first level march_kernel x3, {n_pyr - n_sheet_levels} level(s) as "
                f"pair_kernel phases 1-3, {n_sheet_levels} level(s) as sheet_kernel "
                f"(even sheets, odd sheets); 10 B/value + 4 B/outlier")
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
        one = {"value": sample.nbytes / dt / 1e9, "unit": "GB/s", "cores": 1,
               "sample": f"first {planes} planes of the same field ({sample.size} values), "
                         "one compress (payload, without the outer zstd frame), C oracle"}
        print(f"[bench] cpu oracle {dt:.1f} s on {sample.size} values", file=sys.stderr,
              flush=True)
        cpu = {**cpu_all_cores(xd, shape, levels, stride, args.cpu_threads), "kind": "port",
               "one_core": one}

    api = None
    if not args.no_api:
        api = api_leg(xd, shape, levels, stride, eb, max(2, min(args.steps, 3)))
        api["pipeline_compress_s"] = t_comp
        api["pipeline_decompress_s"] = t_dec

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
                     "kernel": enc_desc,
                     "decode": {"achieved": ach_dec / 1e9, "frac": ach_dec / HBM_PEAK}},
        "cpu_baseline": cpu,
        "autotune": None if tuned is None else {"knobs": tuned["knobs"], "ms": tuned["ms"],
                                                "candidates": len(tuned["table"]),
                                                "table": tuned["table"]},
        "stream_v5": v5,
        "api": api,
    }
    if not args.no_gnn and os.path.exists(args.gnn_ckpt):
        line["gnn"] = gnn_leg(args)
    if not args.no_batch:
        line["batch"] = batch_leg(args)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
