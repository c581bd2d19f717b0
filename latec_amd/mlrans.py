"""LATEC001 version 5: multi-lane rANS stage blobs coded on the GPU.

Version 5 is a stream version added beside version 4 (the reference's
layout, latec/bitstream.py:33-344, which ``compress`` writes by default).
It keeps version 4's header, stage records and stage-constant model: the
per-stage fit (kexp window, Laplace scale level, mixture shape, outlier
weight; latec/rans.py:395-430) and the 2^24 CDF rows of the same tables.
What changes:

* the record blob is the multi-lane rANS word stream of liblatec's
  ``latec_mlrans_encode`` (``[L][L + 1 lane offsets][lane words]``, symbol i
  in lane i % L, L = clamp(ceil(n / 4096), 1, 65536)), so every stage codes
  in parallel on the GPU instead of one serial host coder per stage;
* the payload follows the header as is, without the outer zstd frame (the
  rANS words do not compress, and zstd-9 over them is serial host work).

Decoding runs the lanes on the GPU straight into the device code array, the
outlier ranks are recovered there (``latec_zero_ranks``), and the fused
decode kernels follow.
"""

from __future__ import annotations

import struct

import numpy as np

from . import _device as D
from . import rans as R
from ._lib import check, lib, ptr
from .bitstream import _pack_hlen, parse_record

VERSION_MULTILANE = 5


def _stage_arrays(plan):
    C, ns = plan.channels, plan.n_stages
    n = np.array([C * plan.counts[s] for s in range(ns)], np.int64)
    off = np.array([C * plan.offsets[s] for s in range(ns)], np.int64)
    return n, off


def _rows(params, n):
    """Device CDF rows for the stages with codes, and each stage's row offset."""
    rows, row_off, at = [], np.zeros(len(n), np.int64), 0
    for s, (kexp, shape, owx, level) in enumerate(params):
        row_off[s] = at
        if n[s]:
            r = R.cdf_row(kexp, shape, owx, level)
            rows.append(r)
            at += len(r)
    host = np.concatenate(rows) if rows else np.zeros(1, np.uint32)
    return D.to_device(np.ascontiguousarray(host, np.uint32)), row_off


def sorted_outliers(res, plan, n_out):
    """Stage-order outlier values (sorted by rank) and per-stage offsets."""
    ns = plan.n_stages
    _, code_off = _stage_arrays(plan)
    outl = np.zeros(max(int(n_out.sum()), 1), np.float32)
    out_off = np.zeros(ns, np.int64)
    at = 0
    for s in range(ns):
        out_off[s] = at
        k = int(n_out[s])
        if k:
            lo = int(code_off[s])
            rk = D.to_host(res.d_rank[lo:lo + k], np.uint32)
            vl = D.to_host(res.d_val[lo:lo + k])
            outl[at:at + k] = vl[np.argsort(rk, kind="stable")]
            at += k
    return outl, out_off


def pack_payload(res, fits):
    """Version-5 payload of a device encode result (``GpuEncodeResult``) and
    its per-stage fits: (payload bytes, record lengths, outlier counts)."""
    plan, o = res.plan, res.opts
    ns = plan.n_stages
    n, code_off = _stage_arrays(plan)
    params = [(int(f.kexp) if n[s] else 1, int(f.shape), int(f.owx), int(f.level))
              for s, f in enumerate(fits)]
    kexp = np.array([p[0] for p in params], np.int32)
    d_rows, row_off = _rows(params, n)
    sw = lib().latec_mlrans_scratch_words(ns, ptr(n))
    scratch = D.empty(int(sw), np.uint32)
    lanes = sum(int(lib().latec_mlrans_lanes(int(k))) for k in n)
    cap = 4 * (int(n.sum()) + 3 * lanes + 4 * ns + 16)
    out = D.empty(cap, np.uint8)
    blob_off = np.zeros(ns, np.int64)
    blob_len = np.zeros(ns, np.int64)
    check(lib().latec_mlrans_encode(ns, ptr(res.d_codes), o.code_bytes, ptr(code_off), ptr(n),
                                    ptr(kexp), o.radius, ptr(d_rows), ptr(row_off),
                                    ptr(scratch), int(sw), ptr(out), cap, ptr(blob_off),
                                    ptr(blob_len), ptr(None)))
    total = int((blob_off + blob_len).max(initial=0))
    # blobs through the pinned staging buffer, then each record written once
    # into the payload buffer
    blobs = D.to_host_staged(out[:total], np.uint8) if total else np.zeros(0, np.uint8)
    n_out = D.to_host(res.d_cnt, np.uint32).astype(np.int64)
    outl, out_off = sorted_outliers(res, plan, n_out)
    rec_len = 16 + blob_len + 4 * n_out
    payload = np.empty(int(rec_len.sum()), np.uint8)
    at = 0
    for s in range(ns):
        kx, shape, owx, level = params[s]
        hl, k = int(blob_len[s]), int(n_out[s])
        payload[at:at + 12] = np.frombuffer(
            struct.pack("<IQ", int(n[s]), _pack_hlen(hl, kx, shape, level, owx)), np.uint8)
        payload[at + 12:at + 12 + hl] = blobs[blob_off[s]:blob_off[s] + hl]
        payload[at + 12 + hl:at + 16 + hl] = np.frombuffer(struct.pack("<I", k), np.uint8)
        payload[at + 16 + hl:at + 16 + hl + 4 * k] = \
            outl[out_off[s]:out_off[s] + k].view(np.uint8)
        at += int(rec_len[s])
    return payload.tobytes(), rec_len, n_out


def check_blob(payload, at: int, nbytes: int, n_codes: int) -> None:
    """Validate a version-5 blob header on the host before any device read:
    ``[L][L + 1 lane word offsets][words]`` with L the lane count of an
    ``n_codes``-symbol stage, offsets starting at 0, at least two words (the
    coder state) per lane, and the last offset ending the blob exactly.
    Raises ValueError otherwise."""
    if not n_codes:
        if nbytes:
            raise ValueError("multi-lane blob of an empty stage is not empty")
        return
    L = int(lib().latec_mlrans_lanes(int(n_codes)))
    if at % 4 or nbytes % 4 or nbytes < 4 * (2 + 3 * L) or at + nbytes > len(payload):
        raise ValueError("multi-lane blob is truncated or misaligned")
    head = np.frombuffer(payload, np.uint32, count=L + 2, offset=at).astype(np.int64)
    offs = head[1:]
    if (head[0] != L or offs[0] != 0 or np.any(np.diff(offs) < 2)
            or 2 + L + offs[-1] != nbytes // 4):
        raise ValueError("corrupt multi-lane rANS blob header")


def decode_payload(payload: bytes, plan, ws):
    """Entropy-decode a version-5 payload into the workspace's device codes,
    outlier ranks, values and counts (``codec.Workspace``)."""
    ns = plan.n_stages
    n, code_off = _stage_arrays(plan)
    params, blob_off, blob_len, outs = [], np.zeros(ns, np.int64), np.zeros(ns, np.int64), []
    off = 0
    for s in range(ns):
        n_codes, hlen, kexp, shape, level, owx, literal, rlen = parse_record(payload, off)
        if n_codes != n[s]:
            raise ValueError("stage code count mismatch")
        if literal:
            raise ValueError("version-5 stage record has the literal bit set")
        if n_codes and not kexp:
            raise ValueError("rANS stage record has no coder window")
        params.append((int(kexp) if n_codes else 1, int(shape), int(owx), int(level)))
        blob_off[s] = off + 12
        blob_len[s] = hlen
        check_blob(payload, off + 12, hlen, int(n_codes))
        (k,) = struct.unpack_from("<I", payload, off + 12 + hlen)
        outs.append(np.frombuffer(payload, np.float32, count=k, offset=off + 16 + hlen))
        off += rlen
    if off != len(payload):
        raise ValueError("trailing bytes in LATEC payload")
    kexp = np.array([p[0] for p in params], np.int32)
    d_rows, row_off = _rows(params, n)
    d_payload = D.to_device(np.frombuffer(payload, np.uint8)) if len(payload) else \
        D.empty(4, np.uint8)
    lanes = [int(lib().latec_mlrans_lanes(int(k))) for k in n]
    sw = 16 * ns + 2 * sum((L + 255) // 256 for L in lanes) + 64
    sw = max(sw, 8 * ns + 2 * sum((int(k) + 4095) // 4096 for k in n) + 64)
    scratch = D.empty(sw, np.uint32)
    cb = ws.codes.element_size()
    check(lib().latec_mlrans_decode(ns, ptr(d_payload), len(payload), ptr(blob_off),
                                    ptr(blob_len), ptr(n), ptr(kexp), plan_radius(ws),
                                    ptr(d_rows), ptr(row_off), ptr(ws.codes), cb,
                                    ptr(code_off), ptr(scratch), sw, ptr(None)))
    check(lib().latec_zero_ranks(ns, ptr(ws.codes), cb, ptr(code_off), ptr(n), ptr(ws.rank),
                                 ptr(ws.cnt), ptr(scratch), sw, ptr(None)))
    cnt = D.to_host(ws.cnt, np.uint32)
    for s in range(ns):
        if int(cnt[s]) != len(outs[s]):
            raise ValueError("outlier count mismatch")
        if len(outs[s]):
            lo = int(code_off[s])
            ws.val[lo:lo + len(outs[s])].copy_(D.to_device(np.ascontiguousarray(outs[s])))


def plan_radius(ws) -> int:
    return int(ws.radius)
