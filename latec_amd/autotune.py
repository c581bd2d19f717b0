"""Autotuner for the plain path's launch shapes.

The 3-D level kernels have eleven launch-shape knobs, set process-wide with
``latec_tune_set``:

* ``march_blocks`` -- workgroups per parity class (with 4 waves each, this is
  the waves-per-CU dial: 4096 blocks put 64 waves per CU in flight on the
  256 CUs; fewer blocks make each thread walk more of its class);
* ``march_hist_peel`` -- how many distinct codes per wave the histogram
  counts with one atomic each before falling back to per-lane atomics;
* ``slab`` -- whether large levels run as one level-fused launch (tiles of
  the level marched plane by plane through LDS) or as three class launches;
* ``pair`` -- whether the levels above the first run as line-pair launches
  (one thread codes the two classes of a line along axis 2 together, so x is
  read and recon written as whole lines) or as class launches;
* ``pair_lines`` -- whether the pair launches give each workgroup whole
  lines (every lane owns its pair, neighbours exchanged through LDS) or wave
  runs of 61 owned pairs with three recomputing halo lanes;
* ``pair_quads`` -- whether the line blocks of phases 2 and 3 give each
  thread four points of a line (lines of a multiple of 4, at most 1024) or
  two: bit 0 for the encoder (the one the tuner times), bit 1 for the
  decoder, bits 2 and 3 the same for phase 2 alone;
* ``min_threads`` -- a level launch shortens each thread's march (8, 4, 2, 1
  points) until it has at least this many threads: small levels are
  latency-bound, so they trade march length for parallelism;
* ``pair_waves`` -- the waves-per-SIMD dial of the finest level's encode
  phases 2 and 3: 0 the compiler's own register budget (77 / 103 VGPRs, six
  and four waves per SIMD), 1 the variant compiled for at least eight waves
  per SIMD (64 VGPRs, with spills);
* ``sheet`` -- whether the 3-D levels above the first whose lines have an
  even length of at most 512 run as two sheet launches (even sheets, then odd
  sheets, each marching along the rows of whole lines) instead of the three
  pair phases: bit 0 for the encoder, bit 1 for the decoder;
* ``sheet_blocks`` -- a sheet launch halves its march segment (16, 8, 4 odd
  rows per lane) while it would have fewer workgroups than this;
* ``sheet_min_points`` -- the smallest level (points of its compact grid) that
  runs as sheet launches.

The slab kernel and histogram peeling are measured slower than their
alternatives on every workload tried so far, so the default candidates leave
them at 0 (pass them explicitly to time them).

Neither changes a single code, outlier or reconstructed value (they only
decide which thread does which point and how counts are summed).
``tune`` times the whole encode (the captured hipGraph, HIP events on the
launch stream) for every candidate on the caller's field, checks that the
outputs are bit-identical, keeps the fastest and caches the choice per
(device, shape, levels, stride) in a JSON file (``$LATEC_TUNE_CACHE``, else
``~/.cache/latec_amd/tune.json``).
"""

from __future__ import annotations

import json
import os

import numpy as np

from ._lib import check, lib, ptr

KNOBS = ("march_blocks", "march_hist_peel", "slab", "pair", "pair_lines", "pair_quads",
         "min_threads", "pair_waves", "sheet", "sheet_blocks", "sheet_min_points")
CANDIDATES = {"march_blocks": (1024, 2048, 4096),
              "march_hist_peel": (0,),
              "slab": (0,),
              "pair": (0, 1),
              "pair_lines": (0, 1),
              "pair_quads": (2, 3),
              "min_threads": (1 << 18, 1 << 19),
              "pair_waves": (0, 1),
              "sheet": (0, 3),
              "sheet_blocks": (1024,),
              "sheet_min_points": (1 << 25,)}


def get() -> dict:
    return {k: int(lib().latec_tune_get(k.encode())) for k in KNOBS}


def set_knobs(**kw) -> None:
    for k, v in kw.items():
        if k not in KNOBS:
            raise ValueError(f"unknown tuning knob {k}")
        check(lib().latec_tune_set(k.encode(), int(v)))


def _cache_path() -> str:
    return os.environ.get("LATEC_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "latec_amd",
                                       "tune.json"))


def _key(shape, levels, stride) -> str:
    import torch

    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
    return f"{dev}|{'x'.join(map(str, shape))}|L{levels}|s{stride}"


def load(shape, levels, stride) -> dict | None:
    """Apply and return the cached choice for this workload, if any."""
    try:
        with open(_cache_path()) as f:
            got = json.load(f).get(_key(shape, levels, stride))
    except (OSError, ValueError):
        return None
    if got:
        set_knobs(**got["knobs"])
        return got
    return None


def _save(key, entry) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        try:
            with open(path) as f:
                data = json.load(f)
        except (OSError, ValueError):
            data = {}
        data[key] = entry
        with open(path, "w") as f:
            json.dump(data, f, indent=1)
    except OSError:
        pass  # a read-only home: the choice still applies to this process


def tune(d_x, shape, levels, stride, ebs, opts, reps=3, candidates=None, save=True) -> dict:
    """Time every candidate setting on the device field ``d_x`` ((N,) float32,
    already resident) and keep the fastest.  Returns {"knobs", "ms", "table"}."""
    import itertools

    import torch

    from . import codec
    from .predictor import get_plan

    cand = candidates or CANDIDATES
    plan = get_plan(shape, 1, levels, stride, 1)
    ws = codec.workspace(plan, opts.radius)
    start = get()
    table, ref, best = [], None, None
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ebs_arr = np.ascontiguousarray(ebs, np.float64)
    stream = torch.cuda.current_stream()
    try:
        # a knob the candidates leave out keeps its current value
        for values in itertools.product(*[cand.get(k, (start[k],)) for k in KNOBS]):
            knobs = dict(zip(KNOBS, values))
            set_knobs(**knobs)
            codec.encode_device(d_x, plan, ebs, opts, graph=True, ws=ws)  # capture + warm
            torch.cuda.synchronize()
            ms = []
            for _ in range(reps):
                ws.hist.zero_()
                ws.cnt.zero_()
                ev0.record(stream)
                check(lib().latec_interp_encode_graph(
                    plan.handle, ptr(d_x), ptr(ws.recon), ptr(ws.codes), ptr(ws.hist),
                    ptr(ws.rank), ptr(ws.val), ptr(ws.cnt), ptr(ebs_arr), codec._optr(opts),
                    ptr(None)))
                ev1.record(stream)
                torch.cuda.synchronize()
                ms.append(ev0.elapsed_time(ev1))
            # identical outputs for every launch shape: codes, reconstruction
            # bits and outlier counts equal the first candidate's exactly
            got = (ws.codes, ws.recon.view(torch.int32), ws.cnt)
            if ref is None:
                ref = tuple(t.clone() for t in got)
            elif not all(torch.equal(a, b) for a, b in zip(got, ref)):
                raise RuntimeError(f"launch shape {knobs} changed the encode output")
            t = float(np.median(ms))
            table.append({"knobs": knobs, "ms": t})
            if best is None or t < best["ms"]:
                best = {"knobs": knobs, "ms": t}
    except Exception:
        set_knobs(**start)
        raise
    set_knobs(**best["knobs"])
    entry = {"knobs": best["knobs"], "ms": best["ms"], "table": table}
    if save:
        _save(_key(shape, levels, stride), entry)
    return entry
