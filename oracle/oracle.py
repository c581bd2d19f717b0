"""ctypes wrapper over the C oracle (``latec_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker and the timed CPU
baseline.  The product package (``latec_amd``) never imports this module.

The wrapper mirrors the reference's plain-path functions one for one (each
docstring cites the reference file:line restated by the C function it calls).
"""

from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liblatec_oracle.so")

MAGIC = b"LATEC001"
VERSION = 4
HEADER_FMT = "<8sHHIIBBBBIHhddd16sd"
FLAG_INTERP = 1 << 3
FLAG_CUBIC = 1 << 4
END_MODE_SHIFT = 9


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (no FMA contraction); returns the .so path."""
    src = os.path.join(_HERE, "latec_oracle.c")
    if force or not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _SO


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        p = C.c_void_p
        i64, i32, f64 = C.c_int64, C.c_int, C.c_double
        sig = {
            "orc_n_stages": (i32, [i32, i32]),
            "orc_stage_mask": (i32, [p, i32, i32, i32, i32, i32, p, p, p, p]),
            "orc_stage_ebs": (None, [i32, i32, i32, f64, f64, p]),
            "orc_predict": (None, [p, i32, p, i32, p, i32, i32, i32, i32, i32, i32, p, i64]),
            "orc_quantize": (i64, [p, p, i64, f64, i64, i32, f64, f64, p, p]),
            "orc_dequantize": (i32, [p, p, p, i64, i64, f64, i64, p]),
            "orc_level_weights": (None, [i32, p]),
            "orc_model_weights": (None, [i32, i32, i32, p]),
            "orc_neg_log2_weights": (None, [i32, i32, i32, i32, p]),
            "orc_build_cdfs": (i32, [i32, i32, i32, i32, p]),
            "orc_choose_kexp": (i32, [p, i64, i64]),
            "orc_choose_outlier_weight": (i32, [i64, i64]),
            "orc_const_hist": (None, [p, i64, i64, i32, p]),
            "orc_choose_const_level": (C.c_float, [p, i64, i32, i32, p, p, p]),
            "orc_scale_to_level": (None, [p, i64, f64, p]),
            "orc_rans_encode": (i64, [p, i64, p, i64, p, i64]),
            "orc_rans_encode_levels": (i64, [p, p, i64, p, i64, p, i64]),
            "orc_rans_decode": (i32, [p, i64, p, i64, p, i64]),
            "orc_pack_stage": (i64, [p, i64, p, i64, i64, i32, i32, p, i64]),
            "orc_unpack_stage": (i64, [p, i64, i64, i32, p, i64, p, i64, p, p]),
            "orc_compress_region": (i64, [p, i32, p, i32, i32, i32, i32, p, i32, i32, i32, i64, i32, i32, p, i64, p, p]),
            "orc_decompress_region": (i32, [p, i64, i32, p, i32, i32, i32, i32, p, i32, i32, i32, i64, p]),
            "orc_zstd_compress": (i64, [p, i64, i32, p, i64]),
            "orc_zstd_bound": (i64, [i64]),
            "orc_zstd_content_size": (i64, [p, i64]),
            "orc_zstd_decompress": (i64, [p, i64, p, i64]),
            "orc_mlrans_lanes": (i64, [i64]),
            "orc_mlrans_encode": (i64, [p, i64, p, i64, p, i64]),
            "orc_mlrans_decode": (i32, [p, i64, i64, p, i64, p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# ---- schedule (levels.py) ------------------------------------------------

def stage_plan(shape, levels, stride, block=1):
    """Yield (mask bool, stride, weight, axis) per stage; levels.py:100-147."""
    shape64 = np.asarray(shape, np.int64)
    n = int(np.prod(shape64))
    covered = np.zeros(n, np.uint8)
    w, a = C.c_int(), C.c_int()
    for st in range(lib().orc_n_stages(len(shape), levels)):
        mask = np.empty(n, np.uint8)
        s = lib().orc_stage_mask(_p(shape64), len(shape), levels, stride, block, st,
                                 _p(covered), _p(mask), C.byref(w), C.byref(a))
        yield mask.reshape(shape).astype(bool), s, w.value, a.value


def stage_ebs(ndim, levels, stride, eb, ratio):
    out = np.empty(lib().orc_n_stages(ndim, levels), np.float64)
    lib().orc_stage_ebs(ndim, levels, stride, float(eb), float(ratio), _p(out))
    return out


# ---- interp (predictor.py) -----------------------------------------------

def predict(recon, mask, s, weight, axis, order="cubic", center=0, end_mode=1):
    """Compact (C, |mask|) float32 prediction; predictor.py:255-317."""
    recon = np.ascontiguousarray(recon, np.float32)
    shape64 = np.asarray(recon.shape[1:], np.int64)
    m = np.ascontiguousarray(mask.reshape(-1), np.uint8)
    n = int(m.sum())
    out = np.empty((recon.shape[0], n), np.float32)
    lib().orc_predict(_p(recon), recon.shape[0], _p(shape64), len(shape64), _p(m),
                      s, weight, axis, order == "cubic", center, end_mode, _p(out), n)
    return out


# ---- quantizer (quantizer.py) --------------------------------------------

def quantize(x, pred, eb, radius=1 << 15, round_output=False):
    """(codes u32, outliers f32); quantizer.py:28-87."""
    if eb <= 0:
        raise ValueError(f"error bound must be > 0, got {eb}")
    x = np.ascontiguousarray(x, np.float32).ravel()
    pred = np.ascontiguousarray(pred, np.float32).ravel()
    codes = np.empty(len(x), np.uint32)
    outl = np.empty(len(x), np.float32)
    mode, span, off = 0, 0.0, 0.0
    if round_output is True:
        mode = 1
    elif round_output:
        mode, (span, off) = 2, round_output
    no = lib().orc_quantize(_p(x), _p(pred), len(x), float(eb), int(radius), mode,
                            float(span), float(off), _p(codes), _p(outl))
    return codes, outl[:no].copy()


def dequantize(pred, codes, outliers, eb, radius=1 << 15):
    pred = np.ascontiguousarray(pred, np.float32).ravel()
    codes = np.ascontiguousarray(codes, np.uint32).ravel()
    outliers = np.ascontiguousarray(outliers, np.float32).ravel()
    out = np.empty(len(pred), np.float32)
    if lib().orc_dequantize(_p(pred), _p(codes), _p(outliers), len(outliers), len(pred),
                            float(eb), int(radius), _p(out)):
        raise ValueError("outlier count mismatch")
    return out


# ---- tables and fit (rans.py) --------------------------------------------

def level_weights(kexp):
    out = np.empty((64, 2 << kexp), np.float64)
    lib().orc_level_weights(kexp, _p(out))
    return out


def model_weights(kexp, shape=0, owx=0):
    out = np.empty((64, 2 << kexp), np.float64)
    lib().orc_model_weights(kexp, shape, owx, _p(out))
    return out


def neg_log2_weights(kexp, shape=0, owx=0, precision=24):
    out = np.empty((64, 2 << kexp), np.float32)
    lib().orc_neg_log2_weights(kexp, shape, owx, precision, _p(out))
    return out


def build_cdfs(kexp, shape=0, precision=24, owx=0):
    out = np.empty((64, (2 << kexp) + 1), np.uint32)
    if lib().orc_build_cdfs(kexp, shape, precision, owx, _p(out)):
        raise ValueError("could not normalize PMF at requested precision")
    return out


def choose_kexp(codes, radius):
    codes = np.ascontiguousarray(codes, np.uint32).ravel()
    return lib().orc_choose_kexp(_p(codes), len(codes), int(radius))


def choose_outlier_weight(n_out, n_codes):
    return lib().orc_choose_outlier_weight(int(n_out), int(n_codes))


def const_hist(codes, radius, kexp):
    codes = np.ascontiguousarray(codes, np.uint32).ravel()
    hist = np.empty(2 << kexp, np.float32)
    lib().orc_const_hist(_p(codes), len(codes), int(radius), kexp, _p(hist))
    return hist


def choose_const_level_hist(hist, n_codes, kexp, precision=24):
    hist = np.ascontiguousarray(hist, np.float32)
    lv, sh, ow = C.c_int(), C.c_int(), C.c_int()
    bits = lib().orc_choose_const_level(_p(hist), int(n_codes), kexp, precision,
                                        C.byref(lv), C.byref(sh), C.byref(ow))
    return lv.value, sh.value, ow.value, float(bits)


def choose_const_level(codes, radius, kexp, precision=24):
    """(level, shape, owx, bits); rans.py:395-430."""
    codes = np.ascontiguousarray(codes, np.uint32).ravel()
    return choose_const_level_hist(const_hist(codes, radius, kexp), len(codes), kexp,
                                   precision)


def scale_to_level(scale, eb):
    scale = np.ascontiguousarray(scale, np.float32).ravel()
    out = np.empty(len(scale), np.uint8)
    lib().orc_scale_to_level(_p(scale), len(scale), float(eb), _p(out))
    return out


def rans_encode_syms(syms, cdf):
    syms = np.ascontiguousarray(syms, np.uint32)
    cdf = np.ascontiguousarray(cdf, np.uint32)
    cap = 2 * len(syms) + 8
    words = np.empty(cap, np.uint32)
    nw = lib().orc_rans_encode(_p(syms), len(syms), _p(cdf), len(cdf) - 1, _p(words), cap)
    if nw < 0:
        raise ValueError("rANS capacity")
    return words[:nw].copy()


def rans_encode_levels(syms, levels, cdfs):
    """Per-point levels: cdfs is (64, A + 1)."""
    syms = np.ascontiguousarray(syms, np.uint32)
    levels = np.ascontiguousarray(levels, np.uint8)
    cdfs = np.ascontiguousarray(cdfs, np.uint32)
    cap = 2 * len(syms) + 8
    words = np.empty(cap, np.uint32)
    nw = lib().orc_rans_encode_levels(_p(syms), _p(levels), len(syms), _p(cdfs),
                                      cdfs.shape[1] - 1, _p(words), cap)
    if nw < 0:
        raise ValueError("rANS capacity")
    return words[:nw].copy()


def rans_decode_syms(words, n, cdf):
    words = np.ascontiguousarray(words, np.uint32)
    cdf = np.ascontiguousarray(cdf, np.uint32)
    out = np.empty(n, np.uint32)
    if lib().orc_rans_decode(_p(words), len(words), _p(cdf), len(cdf) - 1, _p(out), n):
        raise ValueError("trailing data in rANS payload")
    return out


# ---- stage records (bitstream.py) ----------------------------------------

def pack_stage(codes, outliers, radius=1 << 15, zstd_level=9, precision=24):
    codes = np.ascontiguousarray(codes, np.uint32).ravel()
    outliers = np.ascontiguousarray(outliers, np.float32).ravel()
    cap = 64 + 8 * len(codes) + 4 * len(outliers) + 1024
    out = np.empty(cap, np.uint8)
    r = lib().orc_pack_stage(_p(codes), len(codes), _p(outliers), len(outliers),
                             int(radius), zstd_level, precision, _p(out), cap)
    if r < 0:
        raise ValueError("pack_stage failed")
    return out[:r].tobytes()


_UNPACK_ERR = {
    1: "rANS stage record has no coder window",
    2: "stage code count mismatch",
    3: "rANS payload is not aligned to 32-bit words",
    4: "trailing data in rANS payload",
    5: "malformed stage record",
}


def unpack_stage(buf, off, n_codes, radius=1 << 15, precision=24):
    arr = np.frombuffer(buf, np.uint8)[off:]
    arr = np.ascontiguousarray(arr)
    codes = np.empty(max(n_codes, 1), np.uint32)
    outl = np.empty(max(n_codes, 1), np.float32)
    no, err = C.c_int64(), C.c_int()
    r = lib().orc_unpack_stage(_p(arr), len(arr), int(radius), precision, _p(codes),
                               n_codes, _p(outl), n_codes, C.byref(no), C.byref(err))
    if r < 0:
        raise ValueError(_UNPACK_ERR.get(err.value, "bad stage record"))
    return codes[:n_codes].copy(), outl[: no.value].copy(), off + r


# ---- whole codec (codec.py) ------------------------------------------------

def _zstd(data: bytes, level: int) -> bytes:
    src = np.frombuffer(data, np.uint8)
    cap = lib().orc_zstd_bound(len(src))
    dst = np.empty(cap, np.uint8)
    r = lib().orc_zstd_compress(_p(np.ascontiguousarray(src)), len(src), level, _p(dst), cap)
    return dst[:r].tobytes()


def _unzstd(data: bytes) -> bytes:
    src = np.ascontiguousarray(np.frombuffer(data, np.uint8))
    n = lib().orc_zstd_content_size(_p(src), len(src))
    dst = np.empty(max(n, 1), np.uint8)
    r = lib().orc_zstd_decompress(_p(src), len(src), _p(dst), n)
    if r != n:
        raise ValueError("bad zstd frame")
    return dst[:n].tobytes()


def compress_payload(field, eb_list, levels, stride, block, order, center, end_mode,
                     radius=1 << 15, round_mode=0, zstd_level=9):
    """_compress_region over a (C, *S) float32 field: (payload, recon, bytes/stage)."""
    field = np.ascontiguousarray(field, np.float32)
    shape64 = np.asarray(field.shape[1:], np.int64)
    ebs = np.ascontiguousarray(eb_list, np.float64)
    cap = 64 * len(ebs) + 8 * field.size + (1 << 16)
    payload = np.empty(cap, np.uint8)
    recon = np.empty_like(field)
    sb = np.zeros(len(ebs), np.int64)
    r = lib().orc_compress_region(_p(field), field.shape[0], _p(shape64), len(shape64),
                                  levels, stride, block, _p(ebs), order == "cubic",
                                  center, end_mode, int(radius), round_mode, zstd_level,
                                  _p(payload), cap, _p(recon), _p(sb))
    if r < 0:
        raise ValueError("oracle compress failed")
    return payload[:r].tobytes(), recon, sb


def header_bytes(channels, src_dtype, spatial, eb, levels, stride, block, radius,
                 vmin, vmax, flags, center, ratio):
    fixed = struct.pack(HEADER_FMT, MAGIC, VERSION, flags, channels, src_dtype, levels,
                        stride, block, center, radius, 0, -1, float(eb), vmin, vmax,
                        b"\0" * 16, float(ratio))
    return fixed + struct.pack(f"<B{len(spatial)}I", len(spatial), *spatial)


def compress(img, eb, order="cubic", levels=4, stride=16, block=1, center=None,
             end_mode=1, eb_ratio=0.9, radius=1 << 15, zstd_level=9):
    """One fast-mode encode of ``latec.compress(img, eb, InterpPredictor(...))``
    (codec.py:45-285 with tune='fast'): returns (stream, recon (C, *S))."""
    is_image = img.ndim <= 2 or (img.ndim == 3 and img.shape[-1] in (1, 3))
    src_dtype = {np.dtype(np.uint8): 0, np.dtype(np.float32): 1}[np.dtype(img.dtype)]
    if is_image:
        if img.ndim == 2:
            img = img[..., None]
        spatial, c = img.shape[:-1], img.shape[-1]
        fimg = np.asarray(img, np.float32)
    else:
        spatial, c = img.shape, 1
        fimg = np.asarray(img, np.float32)[..., None]
    vmin, vmax = float(fimg.min()), float(fimg.max())
    if vmax <= vmin:
        vmax = vmin + 1.0
    field = np.ascontiguousarray(np.moveaxis(fimg, -1, 0))
    if center is None:
        center = 1 if len(spatial) >= 4 else 0
    ebs = stage_ebs_ref(len(spatial), levels, stride, eb, eb_ratio)
    round_mode = 1 if src_dtype == 0 else 0
    payload, recon, _ = compress_payload(field, ebs, levels, stride, block, order, center,
                                         end_mode, radius, round_mode, zstd_level)
    flags = FLAG_INTERP | (FLAG_CUBIC if order == "cubic" else 0) | (end_mode << END_MODE_SHIFT)
    head = header_bytes(c, src_dtype, tuple(spatial), eb, levels, stride, block, radius,
                        vmin, vmax, flags, center, eb_ratio)
    return head + _zstd(payload, zstd_level), recon


def stage_ebs_ref(ndim, levels, stride, eb, ratio):
    """levels.py:360-386 evaluated with the same Python/NumPy scalar ops."""
    per = 2 * ndim - 1
    strides = [stride] + [max(stride >> k, 1) for k in range(1, levels + 1) for _ in range(per)]
    finest = min(strides)
    return [eb * ratio ** np.log2(s / finest) for s in strides]


def decompress(stream):
    """decompress(stream) for interp streams (codec.py:288-322): (C, *S) recon."""
    hsize = struct.calcsize(HEADER_FMT)
    fields = struct.unpack_from(HEADER_FMT, stream, 0)
    magic, version, flags, channels, src_dtype, levels, stride, block, center, radius = fields[:10]
    eb, ratio = fields[12], fields[16]
    if magic != MAGIC:
        raise ValueError(f"not a LATEC stream (bad magic {magic!r})")
    (ndim,) = struct.unpack_from("<B", stream, hsize)
    spatial = struct.unpack_from(f"<{ndim}I", stream, hsize + 1)
    payload = _unzstd(stream[hsize + 1 + 4 * ndim:])
    ebs = np.asarray(stage_ebs_ref(ndim, levels, stride, eb, ratio), np.float64)
    shape64 = np.asarray(spatial, np.int64)
    recon = np.empty((channels, *spatial), np.float32)
    pl = np.frombuffer(payload, np.uint8).copy()
    rc = lib().orc_decompress_region(_p(pl), len(pl), channels, _p(shape64), ndim, levels,
                                     stride, block, _p(ebs),
                                     bool(flags & FLAG_CUBIC), center,
                                     (flags >> END_MODE_SHIFT) & 3, int(radius), _p(recon))
    if rc:
        raise ValueError(f"bad stage record {-rc - 1}")
    return recon


# ---- multi-lane rANS stream version 5 (this project's own format) ----------
#
# Version 5 has no reference counterpart: the functions below are its CPU
# specification (latec_oracle.c, orc_mlrans_*), used to check the GPU coder.
# Records keep version 4's framing and model fit; the blob is the multi-lane
# word stream; the payload is stored without the outer zstd frame.

VERSION_MULTILANE = 5
_HLEN_BITS, _KEXP_SHIFT, _SHAPE_SHIFT, _LEVEL_SHIFT, _OWX_SHIFT = 45, 45, 49, 54, 60


def mlrans_lanes(n):
    return int(lib().orc_mlrans_lanes(int(n)))


def mlrans_encode_syms(syms, cdf):
    syms = np.ascontiguousarray(syms, np.uint32)
    cdf = np.ascontiguousarray(cdf, np.uint32)
    cap = len(syms) + 3 * mlrans_lanes(len(syms)) + 8
    out = np.empty(cap, np.uint32)
    nw = lib().orc_mlrans_encode(_p(syms), len(syms), _p(cdf), len(cdf) - 1, _p(out), cap)
    if nw < 0:
        raise ValueError("symbol outside the CDF row")
    return out[:nw].copy()


def mlrans_decode_syms(words, n, cdf):
    words = np.ascontiguousarray(words, np.uint32)
    cdf = np.ascontiguousarray(cdf, np.uint32)
    out = np.empty(max(int(n), 1), np.uint32)
    if lib().orc_mlrans_decode(_p(words), len(words), int(n), _p(cdf), len(cdf) - 1, _p(out)):
        raise ValueError("corrupt multi-lane rANS blob")
    return out[: int(n)].copy()


def _narrow(codes, radius, kexp):
    codes = np.asarray(codes, np.int64)
    return np.where(codes == 0, 0, codes - int(radius) + (1 << kexp)).astype(np.uint32)


def _widen(syms, radius, kexp):
    syms = np.asarray(syms, np.int64)
    return np.where(syms == 0, 0, syms - (1 << kexp) + int(radius)).astype(np.uint32)


def pack_stage_v5(codes, outliers, radius=1 << 15, precision=24):
    """Version-5 stage record: version 4's fit (kexp, level, shape, owx), the
    multi-lane blob of the narrowed codes under that CDF row."""
    codes = np.ascontiguousarray(codes, np.uint32).ravel()
    outliers = np.ascontiguousarray(outliers, np.float32).ravel()
    n = len(codes)
    if n:
        kexp = choose_kexp(codes, radius)
        level, shape, owx, _ = choose_const_level(codes, radius, kexp, precision)
        cdf = build_cdfs(kexp, shape, precision, owx)[level]
        blob = mlrans_encode_syms(_narrow(codes, radius, kexp), cdf).astype("<u4").tobytes()
    else:
        kexp, level, shape, owx, blob = 1, 0, 0, 0, b""
    framed = (len(blob) | (kexp << _KEXP_SHIFT) | (shape << _SHAPE_SHIFT)
              | (level << _LEVEL_SHIFT) | (owx << _OWX_SHIFT))
    return (struct.pack("<IQ", n, framed) + blob + struct.pack("<I", len(outliers))
            + outliers.tobytes())


def unpack_stage_v5(buf, off, n_codes, radius=1 << 15, precision=24):
    n, framed = struct.unpack_from("<IQ", buf, off)
    if n != n_codes:
        raise ValueError("stage code count mismatch")
    hlen = framed & ((1 << _HLEN_BITS) - 1)
    kexp = (framed >> _KEXP_SHIFT) & 15
    shape = (framed >> _SHAPE_SHIFT) & 31
    level = (framed >> _LEVEL_SHIFT) & 63
    owx = (framed >> _OWX_SHIFT) & 7
    o = off + 12
    if n:
        words = np.frombuffer(buf, "<u4", count=hlen // 4, offset=o)
        cdf = build_cdfs(kexp, shape, precision, owx)[level]
        codes = _widen(mlrans_decode_syms(words, n, cdf), radius, kexp)
    else:
        codes = np.zeros(0, np.uint32)
    o += hlen
    (n_out,) = struct.unpack_from("<I", buf, o)
    o += 4
    outl = np.frombuffer(buf, np.float32, count=n_out, offset=o).copy()
    return codes, outl, o + 4 * n_out


def _split_v4_payload(payload, counts, channels, radius):
    """(codes, outliers) per stage of a version-4 payload."""
    out, off = [], 0
    for n in counts:
        c, o, off = unpack_stage(payload, off, channels * n, radius)
        out.append((c, o))
    return out


def compress_v5(img, eb, order="cubic", levels=4, stride=16, block=1, center=None,
                end_mode=1, eb_ratio=0.9, radius=1 << 15):
    """The version-5 stream of one fast-mode encode: the same codes and
    outliers as ``compress``, version-5 records, no outer zstd frame."""
    is_image = img.ndim <= 2 or (img.ndim == 3 and img.shape[-1] in (1, 3))
    src_dtype = {np.dtype(np.uint8): 0, np.dtype(np.float32): 1}[np.dtype(img.dtype)]
    if is_image:
        if img.ndim == 2:
            img = img[..., None]
        spatial, c = img.shape[:-1], img.shape[-1]
        fimg = np.asarray(img, np.float32)
    else:
        spatial, c = img.shape, 1
        fimg = np.asarray(img, np.float32)[..., None]
    vmin, vmax = float(fimg.min()), float(fimg.max())
    if vmax <= vmin:
        vmax = vmin + 1.0
    field = np.ascontiguousarray(np.moveaxis(fimg, -1, 0))
    if center is None:
        center = 1 if len(spatial) >= 4 else 0
    ebs = stage_ebs_ref(len(spatial), levels, stride, eb, eb_ratio)
    round_mode = 1 if src_dtype == 0 else 0
    payload, recon, _ = compress_payload(field, ebs, levels, stride, block, order, center,
                                         end_mode, radius, round_mode, 1)
    counts = [int(m.sum()) for m, *_ in stage_plan(spatial, levels, stride, block)]
    parts = [pack_stage_v5(cd, ol, radius)
             for cd, ol in _split_v4_payload(payload, counts, c, radius)]
    flags = FLAG_INTERP | (FLAG_CUBIC if order == "cubic" else 0) | (end_mode << END_MODE_SHIFT)
    head = header_bytes(c, src_dtype, tuple(spatial), eb, levels, stride, block, radius,
                        vmin, vmax, flags, center, eb_ratio)
    head = head[:8] + struct.pack("<H", VERSION_MULTILANE) + head[10:]
    return head + b"".join(parts), recon


def decompress_v5(stream):
    """Decode a version-5 stream: records are re-expressed as version-4
    records and decoded by ``decompress``'s region decoder."""
    hsize = struct.calcsize(HEADER_FMT)
    fields = struct.unpack_from(HEADER_FMT, stream, 0)
    if fields[1] != VERSION_MULTILANE:
        raise ValueError(f"unsupported version {fields[1]}")
    channels, levels, stride, block, radius = fields[3], fields[5], fields[6], fields[7], fields[9]
    (ndim,) = struct.unpack_from("<B", stream, hsize)
    spatial = struct.unpack_from(f"<{ndim}I", stream, hsize + 1)
    off = hsize + 1 + 4 * ndim
    v4 = []
    for m, *_ in stage_plan(spatial, levels, stride, block):
        codes, outl, off = unpack_stage_v5(stream, off, channels * int(m.sum()), radius)
        v4.append(pack_stage(codes, outl, radius, 1))
    if off != len(stream):
        raise ValueError("trailing bytes in LATEC payload")
    head = stream[:8] + struct.pack("<H", VERSION) + stream[10:hsize + 1 + 4 * ndim]
    return decompress(head + _zstd(b"".join(v4), 1))
