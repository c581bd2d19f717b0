"""LATEC001 v4 container (the API and byte layout of latec/bitstream.py).

File = 80-byte little-endian header ``"<8sHHIIBBBBIHhddd16sd"``, the spatial
shape ``[ndim u8][ndim x u32]``, then one zstd frame of the payload.  The
payload is one record per stage: ``[n_codes u32][framed u64][blob][n_out
u32][n_out x f32]`` with ``framed = hlen | kexp<<45 | shape<<49 | level<<54 |
owx<<60 | literal<<63`` (bitstream.py:33-52, 178-285).
"""

from __future__ import annotations

import ctypes
import struct
from dataclasses import dataclass

import numpy as np

from . import _device as D
from . import rans as R
from ._lib import Fit, check, lib, ptr
from .quantizer import DEFAULT_RADIUS

MAGIC = b"LATEC001"
VERSION = 4
# Version 5 (this package's addition): multi-lane rANS blobs coded on the GPU
# and the payload stored without the outer zstd frame (latec_amd.mlrans).
VERSION_MULTILANE = 5

FLAG_MOCK = 1 << 0
FLAG_GRAY = 1 << 1
FLAG_INTERP = 1 << 3
FLAG_CUBIC = 1 << 4
FLAG_NOTILE = 1 << 5
FLAG_RANS = 1 << 6
FLAG_END_QUAD = 1 << 9
FLAG_END_EXTRAP = 1 << 10
END_MODE_SHIFT = 9

_HEADER_FMT = "<8sHHIIBBBBIHhddd16sd"
_HEADER_SIZE = struct.calcsize(_HEADER_FMT)

DTYPE_CODES = {0: np.uint8, 1: np.float32}
DTYPE_IDS = {np.dtype(np.uint8): 0, np.dtype(np.float32): 1}

_HLEN_BITS, _KEXP_BITS, _SHAPE_BITS, _LEVEL_BITS = 45, 4, 5, 6
_KEXP_SHIFT = _HLEN_BITS
_SHAPE_SHIFT = _KEXP_SHIFT + _KEXP_BITS
_LEVEL_SHIFT = _SHAPE_SHIFT + _SHAPE_BITS
_OWX_SHIFT = _LEVEL_SHIFT + _LEVEL_BITS
_LIT_SHIFT = _OWX_SHIFT + R.OUTLIER_BITS
_LITERAL_MIN = 1 << 12
_LITERAL_SCOUT_SLACK = 0.85


@dataclass
class Header:
    channels: int
    src_dtype: int
    spatial: tuple
    eb: float
    levels: int
    anchor_stride: int
    anchor_block: int
    radius: int
    max_radius: int
    agg_level: int
    vmin: float
    vmax: float
    ckpt_hash: bytes = b"\0" * 16
    flags: int = 0
    interp_center: int = 0
    eb_ratio: float = 1.0
    version: int = VERSION

    def pack(self) -> bytes:
        fixed = struct.pack(_HEADER_FMT, MAGIC, self.version, self.flags, self.channels,
                            self.src_dtype, self.levels, self.anchor_stride, self.anchor_block,
                            self.interp_center, self.radius, self.max_radius, self.agg_level,
                            self.eb, self.vmin, self.vmax, self.ckpt_hash, self.eb_ratio)
        if not self.spatial or len(self.spatial) > 255:
            raise ValueError("spatial shape must contain 1..255 dimensions")
        return fixed + struct.pack(f"<B{len(self.spatial)}I", len(self.spatial), *self.spatial)

    @classmethod
    def unpack(cls, buf: bytes) -> "Header":
        (magic, version, flags, channels, src_dtype, levels, anchor_stride, anchor_block,
         interp_center, radius, max_radius, agg_level, eb, vmin, vmax, ckpt_hash,
         eb_ratio) = struct.unpack_from(_HEADER_FMT, buf, 0)
        if magic != MAGIC:
            raise ValueError(f"not a LATEC stream (bad magic {magic!r})")
        if version not in (VERSION, VERSION_MULTILANE):
            raise ValueError(f"unsupported version {version}")
        (ndim,) = struct.unpack_from("<B", buf, _HEADER_SIZE)
        if not ndim:
            raise ValueError("stream spatial shape is empty")
        spatial = struct.unpack_from(f"<{ndim}I", buf, _HEADER_SIZE + 1)
        return cls(channels=channels, src_dtype=src_dtype, spatial=spatial, eb=eb,
                   levels=levels, anchor_stride=anchor_stride, anchor_block=anchor_block,
                   radius=radius, max_radius=max_radius, agg_level=agg_level, vmin=vmin,
                   vmax=vmax, ckpt_hash=ckpt_hash, flags=flags, interp_center=interp_center,
                   eb_ratio=eb_ratio, version=version)


def zstd_compress(data: bytes, level: int) -> bytes:
    src = np.frombuffer(data, np.uint8)
    cap = lib().latec_zstd_bound(len(src))
    dst = np.empty(max(cap, 1), np.uint8)
    n = check(lib().latec_zstd_compress(ptr(src) if len(src) else None, len(src), int(level),
                                        ptr(dst), cap))
    return dst[:n].tobytes()


def zstd_decompress(data: bytes) -> bytes:
    src = np.frombuffer(data, np.uint8)
    n = check(lib().latec_zstd_content_size(ptr(src), len(src)))
    dst = np.empty(max(n, 1), np.uint8)
    got = check(lib().latec_zstd_decompress(ptr(src), len(src), ptr(dst), n))
    if got != n:
        raise ValueError("truncated zstd frame")
    return dst[:n].tobytes()


def write_stream(header: Header, payload: bytes, zstd_level: int = 9, *,
                 frame: bytes | None = None) -> bytes:
    """bitstream.py:155-157.  ``frame``: the payload's zstd frame at
    zstd_level when the caller already has it (latec_pack_stages_framed)."""
    if header.version == VERSION_MULTILANE:
        return header.pack() + payload
    return header.pack() + (zstd_compress(payload, zstd_level) if frame is None else frame)


def read_stream(data: bytes):
    header = Header.unpack(data)
    off = _HEADER_SIZE + 1 + 4 * len(header.spatial)
    if header.version == VERSION_MULTILANE:
        return header, bytes(data[off:])
    return header, zstd_decompress(data[off:])


def _pack_hlen(hlen, kexp, shape, level=0, owx=0, literal=False) -> int:
    """bitstream.py:260-285."""
    if hlen >= 1 << _HLEN_BITS:
        raise ValueError("stage entropy blob is too large to frame")
    if not 0 <= int(kexp) < 1 << _KEXP_BITS:
        raise ValueError("rANS coder window does not fit the stage frame")
    if not 0 <= int(shape) < 1 << _SHAPE_BITS:
        raise ValueError("rANS shape id does not fit the stage frame")
    if not 0 <= int(level) < 1 << _LEVEL_BITS:
        raise ValueError("rANS scale level does not fit the stage frame")
    if not 0 <= int(owx) < 1 << R.OUTLIER_BITS:
        raise ValueError("rANS outlier weight does not fit the stage frame")
    return (hlen | (int(kexp) << _KEXP_SHIFT) | (int(shape) << _SHAPE_SHIFT)
            | (int(level) << _LEVEL_SHIFT) | (int(owx) << _OWX_SHIFT)
            | (int(bool(literal)) << _LIT_SHIFT))


def _fit_ptr(f: Fit):
    return ctypes.cast(ctypes.pointer(f), ctypes.c_void_p)


def pack_record(codes: np.ndarray, outliers: np.ndarray, fit: Fit, radius: int,
                zstd_level: int) -> bytes:
    """One const-level record from host codes and a fit (native packer)."""
    cb = 2 if radius <= 32768 else 4
    codes = np.ascontiguousarray(codes, np.uint16 if cb == 2 else np.uint32)
    outliers = np.ascontiguousarray(outliers, np.float32)
    n = len(codes)
    row = R.cdf_row(fit.kexp, fit.shape, fit.owx, fit.level) if n else \
        np.zeros(5, np.uint32)
    cap = 16 + 4 * len(outliers) + 4 * (n + 2) + 64
    out = np.empty(cap, np.uint8)
    r = check(lib().latec_pack_stage(ptr(codes) if n else None, cb, n,
                                     ptr(outliers) if len(outliers) else None, len(outliers),
                                     int(radius), int(zstd_level), _fit_ptr(fit), ptr(row),
                                     ptr(out), cap))
    return out[:r].tobytes()


def pack_stage(codes, outliers, *, rans_levels=None, rans_tables=None,
               radius: int = DEFAULT_RADIUS, zstd_level: int = 9) -> bytes:
    """bitstream.py:178-243: const-level fit (GPU) + literal/rANS back end, or
    per-point levels when ``rans_levels`` is given."""
    out = np.ascontiguousarray(outliers, np.float32).ravel()
    if rans_levels is None:
        blob_codes = np.asarray(codes, np.uint32).ravel()
        tables = rans_tables if rans_tables is not None else R.const_tables(radius)
        R_ = tables.radius
        cb = 2 if R_ <= 32768 else 4
        if len(blob_codes):
            host = blob_codes.astype(np.uint16) if cb == 2 else blob_codes
            fit = R.fit_codes_device(D.to_device(host), len(blob_codes), R_, cb,
                                     precision=tables.precision)
        else:
            fit = Fit()
            fit.kexp = 1
        return pack_record(blob_codes, out, fit, R_, zstd_level)
    if rans_tables is None:
        raise ValueError("rans_tables are required with rans_levels")
    blob_codes = R.code_array(codes)  # u16 codes stay u16
    kexp = R.choose_kexp(blob_codes, rans_tables.radius)
    shape = R.choose_shape(blob_codes, rans_levels, rans_tables, kexp)
    hblob = R.rans_encode(blob_codes, rans_levels, rans_tables, kexp=kexp, shape=shape)
    return (struct.pack("<IQ", len(blob_codes), _pack_hlen(len(hblob), kexp, shape))
            + hblob + struct.pack("<I", len(out)) + out.tobytes())


def parse_record(buf, off: int):
    """(n_codes, hlen, kexp, shape, level, owx, literal, record length)."""
    n_codes, framed = struct.unpack_from("<IQ", buf, off)
    hlen = framed & ((1 << _HLEN_BITS) - 1)
    kexp = (framed >> _KEXP_SHIFT) & ((1 << _KEXP_BITS) - 1)
    shape = (framed >> _SHAPE_SHIFT) & ((1 << _SHAPE_BITS) - 1)
    level = (framed >> _LEVEL_SHIFT) & ((1 << _LEVEL_BITS) - 1)
    owx = (framed >> _OWX_SHIFT) & ((1 << R.OUTLIER_BITS) - 1)
    literal = bool((framed >> _LIT_SHIFT) & 1)
    (n_out,) = struct.unpack_from("<I", buf, off + 12 + hlen)
    return n_codes, hlen, kexp, shape, level, owx, literal, 12 + hlen + 4 + 4 * n_out


def unpack_record(buf: bytes, off: int, n_expect: int, radius: int):
    """Native unpack of a const-level record: (codes, out_rank, out_val, next_off)."""
    n_codes, hlen, kexp, shape, level, owx, literal, _ = parse_record(buf, off)
    if not kexp:
        raise ValueError("rANS stage record has no coder window")
    if n_codes != n_expect:
        raise ValueError("stage code count mismatch")
    cb = 2 if radius <= 32768 else 4
    row = R.cdf_row(kexp, shape, owx, level) if (n_codes and not literal) else \
        np.zeros(5, np.uint32)
    arr = np.frombuffer(buf, np.uint8, offset=off)
    codes = np.empty(max(n_codes, 1), np.uint16 if cb == 2 else np.uint32)
    rank = np.empty(max(n_codes, 1), np.uint32)
    val = np.empty(max(n_codes, 1), np.float32)
    n_out = ctypes.c_int64()
    used = check(lib().latec_unpack_stage(ptr(arr), len(arr), int(radius), ptr(row), n_expect,
                                          ptr(codes), cb, ptr(rank), ptr(val), n_codes,
                                          ctypes.byref(n_out)))
    k = n_out.value
    return codes[:n_codes], rank[:k].copy(), val[:k].copy(), off + used


def unpack_record_into(arr: np.ndarray, off: int, n_expect: int, radius: int, codes: np.ndarray,
                       rank: np.ndarray, val: np.ndarray) -> int:
    """unpack_record writing into caller arrays: ``codes`` holds n_expect
    codes, ``rank``/``val`` exactly the record's outlier count (``arr`` is the
    payload as uint8).  Returns the next offset."""
    n_codes, hlen, kexp, shape, level, owx, literal, length = parse_record(arr, off)
    if not kexp:
        raise ValueError("rANS stage record has no coder window")
    if n_codes != n_expect:
        raise ValueError("stage code count mismatch")
    n_out = (length - 16 - hlen) // 4
    if len(codes) < n_codes or len(rank) != n_out or len(val) != n_out:
        raise ValueError("unpack buffers do not match the record")
    cb = 2 if radius <= 32768 else 4
    if codes.itemsize != cb:
        raise ValueError("code buffer width does not match the radius")
    row = R.cdf_row(kexp, shape, owx, level) if (n_codes and not literal) else \
        np.zeros(5, np.uint32)
    got = ctypes.c_int64()
    used = check(lib().latec_unpack_stage(ptr(arr[off:]), len(arr) - off, int(radius), ptr(row),
                                          n_expect, ptr(codes), cb, ptr(rank), ptr(val), n_out,
                                          ctypes.byref(got)))
    if got.value != n_out or used != length:
        raise ValueError("stage record length mismatch")
    return off + used


def unpack_stage(buf, off, *, rans_levels=None, rans_tables=None, radius=DEFAULT_RADIUS,
                 code_dtype=np.uint32):
    """bitstream.py:288-344: (codes u32, outliers f32, next offset); with
    ``rans_levels`` the codes come as ``code_dtype`` (u32, or u16)."""
    n_codes, hlen, kexp, shape, level, owx, literal, _ = parse_record(buf, off)
    if rans_levels is None:
        tables = rans_tables if rans_tables is not None else R.const_tables(radius)
        codes, _rank, val, nxt = unpack_record(buf, off, n_codes, tables.radius)
        return codes.astype(np.uint32), val, nxt
    if rans_tables is None:
        raise ValueError("rans_tables are required with rans_levels")
    if not kexp:
        raise ValueError("rANS stage record has no coder window")
    o = off + 12
    codes = R.rans_decode(np.frombuffer(buf, np.uint8, count=hlen, offset=o), rans_levels,
                          rans_tables, kexp=kexp, shape=shape, code_dtype=code_dtype)
    if len(codes) != n_codes:
        raise ValueError("stage code count mismatch")
    o += hlen
    (n_out,) = struct.unpack_from("<I", buf, o)
    o += 4
    outliers = np.frombuffer(buf, np.float32, count=n_out, offset=o).copy()
    return codes, outliers, o + 4 * n_out
