"""Laplace-mixture rANS tables, fits and coders (the API of latec/rans.py).

Tables are built on the GPU (liblatec ``latec_level_weights`` /
``latec_build_cdfs``: the same float64 arithmetic, bit for bit), the
stage-constant fit runs on the GPU, and the stack rANS coder is native host
code (``latec_rans_encode``/``decode``).  Public names and semantics follow
latec/rans.py; each function cites the lines it mirrors.
"""

from __future__ import annotations

import threading
from dataclasses import dataclass
from functools import lru_cache

import numpy as np

from . import _device as D
from ._lib import Fit, check, lib, ptr, stream_ptr

SCALE_LO_DIV = 16.0
SCALE_HI_MULT = 4096.0
N_LEVELS = 64
DEFAULT_PRECISION = 24

_MIX_DELTAS = (4, 8, 16, 24)
_MIX_WEIGHTS = (0.02, 0.05, 0.12, 0.25, 0.45)
SHAPES = ((0, 0.0),) + tuple((d, w) for d in _MIX_DELTAS for w in _MIX_WEIGHTS)
N_SHAPES = len(SHAPES)
SHAPE_BITS = 5
OUTLIER_WEIGHTS = (0.0, 2.0**-13, 2.0**-10, 2.0**-8, 2.0**-6, 2.0**-4, 2.0**-2, 0.5)
OUTLIER_BITS = 3
_WIDE_KEXP = 13
_SELECT_SAMPLE = 0.05
_SELECT_MIN = 1024


@dataclass(frozen=True)
class RansTables:
    """rans.py:17-37: eb-relative tables; ``cdfs`` is fetched lazily."""

    precision: int
    eb: float
    radius: int
    scale_grid: np.ndarray

    @property
    def total(self) -> int:
        return 1 << self.precision

    @property
    def alphabet(self) -> int:
        return 2 * self.radius

    def cdfs_for(self, kexp: int, shape: int = 0, owx: int = 0) -> np.ndarray:
        return build_cdfs(int(kexp), int(shape), self.precision, int(owx))


def full_kexp(radius: int) -> int:
    return max(1, int(int(radius) - 1).bit_length())


def code_array(codes) -> np.ndarray:
    """codes as a flat contiguous u16 or u32 array (other dtypes widened to u32)."""
    a = np.asarray(codes).ravel()
    if a.dtype not in (np.uint16, np.uint32):
        a = a.astype(np.uint32)
    return np.ascontiguousarray(a)


def choose_kexp(codes, radius: int) -> int:
    """rans.py:70-80 (latec_code_kexp: one pass over the codes)."""
    a = code_array(codes)
    return check(lib().latec_code_kexp(ptr(a), a.itemsize, len(a), int(radius)))


def _narrow(codes, radius: int, kexp: int) -> np.ndarray:
    """rans.py:83-89."""
    span = (1 << int(kexp)) - 1
    q = np.asarray(codes).astype(np.int64) - int(radius)
    syms = np.where(np.asarray(codes) == 0, 0, q + span + 1)
    if len(syms) and (int(syms.min()) < 0 or int(syms.max()) > 2 * span + 1):
        raise ValueError("code outside the narrowed rANS alphabet")
    return syms


def _widen(syms, radius: int, kexp: int) -> np.ndarray:
    """Narrowed symbols -> codes (0 stays 0): sym - span - 1 + radius in
    uint32 arithmetic (every nonzero symbol of a valid stream widens to a
    code >= 1)."""
    syms = np.asarray(syms)
    off = np.uint32((int(radius) - (1 << int(kexp))) % (1 << 32))
    codes = np.add(syms, off, dtype=np.uint32, casting="unsafe")
    codes[syms == 0] = 0
    return codes


def choose_outlier_weight(n_outliers: int, n_codes: int) -> int:
    """rans.py:162-169."""
    if n_codes <= 0 or n_outliers <= 0:
        return 0
    rate = n_outliers / n_codes
    return int(np.argmin([abs(np.log(max(w, 1e-12)) - np.log(rate)) for w in OUTLIER_WEIGHTS]))


@lru_cache(maxsize=None)
def _shape_ids(kexp: int):
    if kexp < _WIDE_KEXP:
        return tuple(range(N_SHAPES))
    return (0,) + tuple(i for i, (_d, w) in enumerate(SHAPES) if w == 0.02)


# ---- GPU-built tables -------------------------------------------------------

_table_lock = threading.Lock()


@lru_cache(maxsize=32)
def _level_weights(kexp: int, n_levels: int = N_LEVELS) -> np.ndarray:
    """rans.py:247-267, built on the GPU."""
    out = D.empty(N_LEVELS * (2 << kexp), np.float64)
    with _table_lock:
        check(lib().latec_level_weights(int(kexp), ptr(out), stream_ptr()))
    D.torch().cuda.current_stream().synchronize()
    return D.to_host(out).reshape(N_LEVELS, 2 << kexp)


@lru_cache(maxsize=64)
def _model_weights(kexp: int, shape: int, owx: int) -> np.ndarray:
    out = D.empty(N_LEVELS * (2 << kexp), np.float64)
    with _table_lock:
        check(lib().latec_model_weights(int(kexp), int(shape), int(owx), ptr(out), stream_ptr()))
    D.torch().cuda.current_stream().synchronize()
    return D.to_host(out).reshape(N_LEVELS, 2 << kexp)


@lru_cache(maxsize=16)
def _neg_log2_weights(kexp: int, shape: int, n_levels: int, precision: int, owx: int):
    out = D.empty(N_LEVELS * (2 << kexp), np.float32)
    with _table_lock:
        check(lib().latec_neg_log2_weights(int(kexp), int(shape), int(owx), int(precision),
                                           ptr(out), stream_ptr()))
    D.torch().cuda.current_stream().synchronize()
    return D.to_host(out).reshape(N_LEVELS, 2 << kexp)


_cdf_cache: dict = {}


def build_cdfs(kexp: int, shape: int = 0, precision: int = DEFAULT_PRECISION, owx: int = 0,
               levels=None) -> np.ndarray:
    """_build_cdfs_cached rows (rans.py:305-329) built on the GPU; ``levels``
    selects rows (default all 64).  Returns (rows, A+1) uint32."""
    rows = range(N_LEVELS) if levels is None else levels
    A = 2 << int(kexp)
    out = np.empty((len(rows), A + 1), np.uint32)
    missing = [r for r in rows if (kexp, shape, precision, owx, r) not in _cdf_cache]
    if missing:
        lo, hi = min(missing), max(missing) + 1
        d = D.empty((hi - lo) * (A + 1), np.uint32)
        with _table_lock:
            check(lib().latec_build_cdfs(int(kexp), int(shape), int(precision), int(owx), lo, hi,
                                         ptr(d), stream_ptr()))
        D.torch().cuda.current_stream().synchronize()
        h = D.to_host(d, np.uint32).reshape(hi - lo, A + 1)
        for r in range(lo, hi):
            _cdf_cache[(kexp, shape, precision, owx, r)] = h[r - lo].copy()
    for i, r in enumerate(rows):
        out[i] = _cdf_cache[(kexp, shape, precision, owx, r)]
    return out


def cdf_row(kexp: int, shape: int, owx: int, level: int, precision: int = DEFAULT_PRECISION):
    return build_cdfs(kexp, shape, precision, owx, [level])[0]


def build_laplace_tables(eb: float, radius: int, *, n_levels: int = 64,
                         precision: int = DEFAULT_PRECISION) -> RansTables:
    """rans.py:281-301 (the CDFs themselves are fetched per window/shape)."""
    eb = float(eb)
    if eb <= 0:
        raise ValueError("eb must be > 0")
    if n_levels != N_LEVELS:
        raise ValueError("only 64 scale levels are supported")
    lo, hi = eb / SCALE_LO_DIV, eb * SCALE_HI_MULT
    grid = np.exp(np.linspace(np.log(lo), np.log(hi), n_levels)).astype(np.float32)
    return RansTables(precision=int(precision), eb=eb, radius=int(radius), scale_grid=grid)


@lru_cache(maxsize=4)
def const_tables(radius: int) -> RansTables:
    return build_laplace_tables(1.0, int(radius))


def scale_to_level_host(scale, eb: float) -> np.ndarray:
    """rans.py:270-278 on the host with NumPy: the definition the device levels
    reproduce (its float32 log is NumPy's, whose rounding depends on the
    machine's SIMD code path)."""
    if eb <= 0:
        raise ValueError("eb must be > 0")
    scale = np.asarray(scale, np.float32)
    lo, hi = float(eb) / SCALE_LO_DIV, float(eb) * SCALE_HI_MULT
    t = (np.log(np.clip(scale, lo, hi)) - np.log(lo)) / (np.log(hi) - np.log(lo))
    return np.rint(t * (N_LEVELS - 1)).clip(0, N_LEVELS - 1).astype(np.uint8)


@lru_cache(maxsize=1024)
def _thresholds(eb: float) -> bytes:
    # Non-negative float32 values order like their bit patterns as unsigned
    # integers, so the first pattern at or above each level is found by
    # bisection over [0, bits(f32(eb * 4096))] (gnn_codec.py:827-861 does the
    # same search).  All 63 levels move together; 32 halvings close a 2^31 span.
    want = np.arange(1, N_LEVELS, dtype=np.int64)
    lo = np.zeros(N_LEVELS - 1, np.int64)
    hi = np.full(N_LEVELS - 1, int(np.float32(float(eb) * SCALE_HI_MULT).view(np.uint32)),
                 np.int64)
    for _ in range(33):
        mid = (lo + hi) // 2
        at = scale_to_level_host(mid.astype(np.uint32).view(np.float32), eb) >= want
        hi = np.where(at, mid, hi)
        lo = np.where(at, lo, mid + 1)
    assert np.all(lo == hi)
    return lo.astype(np.uint32).view(np.float32).tobytes()


def scale_thresholds(eb: float) -> np.ndarray:
    """The 63 float32 thresholds of ``latec_scale_to_level`` for this eb: entry k
    is the smallest float32 that scale_to_level_host puts at level k + 1."""
    if eb <= 0:
        raise ValueError("eb must be > 0")
    return np.frombuffer(_thresholds(float(eb)), np.float32)


def scale_to_level_device(d_scale, n: int, eb: float, d_levels, stream=None) -> None:
    """Device levels (u8) of n device float32 scales: equal to
    scale_to_level_host on every value (asynchronous on the stream)."""
    thr = np.ascontiguousarray(scale_thresholds(eb))
    check(lib().latec_scale_to_level(ptr(d_scale), int(n), ptr(thr), ptr(d_levels),
                                     stream_ptr(stream)))


def scale_to_level(scale, eb: float, n_levels: int = 64) -> np.ndarray:
    """rans.py:270-278 evaluated on the GPU."""
    if eb <= 0:
        raise ValueError("eb must be > 0")
    if n_levels != N_LEVELS:
        raise ValueError("only 64 scale levels are supported")
    s = np.ascontiguousarray(scale, np.float32).ravel()
    ds = D.to_device(s)
    out = D.empty(len(s), np.uint8)
    scale_to_level_device(ds, len(s), eb, out)
    D.torch().cuda.current_stream().synchronize()
    return D.to_host(out).reshape(np.shape(scale))


# ---- stage-constant fit ------------------------------------------------------

def fit_codes_device(d_codes, n: int, radius: int, code_bytes: int, d_hist=None,
                     precision: int = DEFAULT_PRECISION) -> Fit:
    """choose_kexp + choose_const_level on device codes (rans.py:70-80, 395-430)."""
    if d_hist is None:
        d_hist = D.zeros(lib().latec_hist_words(radius), np.uint32)
        check(lib().latec_code_hist(ptr(d_codes), int(n), code_bytes, int(radius), ptr(d_hist),
                                    stream_ptr()))
    f = Fit()
    check(lib().latec_fit_codes(ptr(d_hist), ptr(d_codes), int(n), int(radius), code_bytes,
                                precision, ptr_fit(f), stream_ptr()))
    return f


def ptr_fit(f: Fit):
    import ctypes

    return ctypes.cast(ctypes.pointer(f), ctypes.c_void_p)


def choose_const_level(codes, tables: RansTables, kexp: int):
    """(level, shape, owx, bits) -- rans.py:395-430 (GPU fit).  The window
    is recomputed from the codes; ``kexp`` must agree with it."""
    codes = np.ascontiguousarray(codes, np.uint32).ravel()
    if len(codes) == 0:
        return 0, 0, 0, 0.0
    cb = 2 if tables.radius <= 32768 else 4
    host = codes.astype(np.uint16) if cb == 2 else codes
    f = fit_codes_device(D.to_device(host), len(codes), tables.radius, cb,
                         precision=tables.precision)
    if f.kexp != int(kexp):
        raise ValueError("kexp does not match the codes' window")
    return int(f.level), int(f.shape), int(f.owx), float(f.bits)


# ---- coders ------------------------------------------------------------------

def _encode_syms(syms: np.ndarray, cdf: np.ndarray) -> bytes:
    syms = np.ascontiguousarray(syms, np.uint32)
    cap = len(syms) + 2
    words = np.empty(cap, np.uint32)
    nw = check(lib().latec_rans_encode(ptr(syms), len(syms), ptr(cdf), len(cdf) - 1,
                                       ptr(words), cap))
    return words[:nw].astype("<u4", copy=False).tobytes()


def rans_encode_const(codes, level: int, tables: RansTables, *, kexp=None, shape: int = 0,
                      owx: int = 0) -> bytes:
    """rans.py:466-498."""
    codes = np.asarray(codes, np.uint32).ravel()
    if len(codes) == 0:
        return b""
    if int(codes.max(initial=0)) >= tables.alphabet:
        raise ValueError("code exceeds table alphabet")
    if not 0 <= int(level) < N_LEVELS:
        raise ValueError("scale level exceeds table count")
    kexp = full_kexp(tables.radius) if kexp is None else int(kexp)
    syms = _narrow(codes, tables.radius, kexp)
    return _encode_syms(syms, cdf_row(kexp, shape, owx, level, tables.precision))


def rans_decode_const(blob: bytes, n_codes: int, level: int, tables: RansTables, *, kexp=None,
                      shape: int = 0, owx: int = 0) -> np.ndarray:
    """rans.py:501-535."""
    n_codes = int(n_codes)
    if n_codes == 0:
        if blob:
            raise ValueError("non-empty rANS blob for an empty stage")
        return np.empty(0, np.uint32)
    if len(blob) % 4:
        raise ValueError("rANS payload is not aligned to 32-bit words")
    if not 0 <= int(level) < N_LEVELS:
        raise ValueError("scale level exceeds table count")
    kexp = full_kexp(tables.radius) if kexp is None else int(kexp)
    cdf = cdf_row(kexp, shape, owx, level, tables.precision)
    words = np.frombuffer(blob, "<u4").astype(np.uint32)
    out = np.empty(n_codes, np.uint32)
    check(lib().latec_rans_decode(ptr(words), len(words), ptr(cdf), len(cdf) - 1, ptr(out),
                                  n_codes))
    return _widen(out, tables.radius, kexp)


def choose_shape(codes, levels64, tables: RansTables, kexp: int, *,
                 sample: float = _SELECT_SAMPLE) -> int:
    """rans.py:204-243 on the GPU-built level weights."""
    codes = np.asarray(codes).ravel()
    if len(codes) == 0:
        return 0
    levels = np.asarray(levels64, np.uint8).ravel()
    if codes.shape != levels.shape:
        raise ValueError("codes and levels must have the same flattened length")
    # the stride sample first (narrowing is elementwise, so this is the
    # reference's sample of the narrowed symbols)
    if sample < 1.0:
        step = max(1, len(codes) // max(_SELECT_MIN, int(len(codes) * sample)))
        if step > 1:
            codes, levels = codes[::step], levels[::step]
    syms = _narrow(codes, tables.radius, kexp)
    levels = levels.astype(np.int64)
    weights = _level_weights(int(kexp))
    floor = 1.0 / tables.total
    idx = np.arange(N_LEVELS)
    p0 = weights[levels, syms]
    broad_of = {}  # level offset -> gathered broad weights (shared by shapes)
    best_id, best_cost = 0, np.inf
    for shape_id in _shape_ids(int(kexp)):
        d, w = SHAPES[shape_id]
        p = p0
        if w:
            if d not in broad_of:
                broad_of[d] = weights[np.minimum(idx + d, N_LEVELS - 1)[levels], syms]
            p = (1.0 - w) * p + w * broad_of[d]
        cost = -np.log2(np.maximum(p, floor)).sum()
        if cost < best_cost:
            best_id, best_cost = shape_id, float(cost)
    return best_id


def rans_encode(codes, levels64, tables: RansTables, *, kexp=None, shape: int = 0) -> bytes:
    """Per-point levels (rans.py:546-578)."""
    codes = code_array(codes)
    levels = np.ascontiguousarray(levels64, np.uint8).ravel()
    if codes.shape != levels.shape:
        raise ValueError("codes and levels must have the same flattened length")
    if len(codes) == 0:
        return b""
    if int(codes.max(initial=0)) >= tables.alphabet:
        raise ValueError("code exceeds table alphabet")
    kexp = full_kexp(tables.radius) if kexp is None else int(kexp)
    cdfs = np.ascontiguousarray(build_cdfs(kexp, shape, tables.precision, 0))
    cap = len(codes) + 2
    words = np.empty(cap, np.uint32)
    # narrowing (rans.py:83-89) and the level coder in one pass; levels
    # >= 64 and codes outside the narrowed alphabet are rejected there
    nw = check(lib().latec_rans_encode_codes_levels(
        ptr(codes), codes.itemsize, ptr(levels), len(codes), int(tables.radius), kexp,
        ptr(cdfs), cdfs.shape[1] - 1, ptr(words), cap))
    return words[:nw].astype("<u4", copy=False).tobytes()


def rans_decode(blob, levels64, tables: RansTables, *, kexp=None, shape: int = 0,
                code_dtype=np.uint32):
    """rans.py:581-616: the stage's codes (``code_dtype`` u32, or u16 when
    every code fits).  ``blob`` is bytes or a uint8 array of the words."""
    levels = np.ascontiguousarray(levels64, np.uint8).ravel()
    if len(levels) == 0:
        if len(blob):
            raise ValueError("non-empty rANS blob for an empty stage")
        return np.empty(0, code_dtype)
    if len(blob) % 4:
        raise ValueError("rANS payload is not aligned to 32-bit words")
    kexp = full_kexp(tables.radius) if kexp is None else int(kexp)
    raw = np.frombuffer(blob, np.uint8)
    if raw.ctypes.data % 4:
        raw = raw.copy()
    words = raw.view("<u4")
    cdfs = np.ascontiguousarray(build_cdfs(kexp, shape, tables.precision, 0))
    out = np.empty(len(levels), code_dtype)
    # decode, level check and widening (0 stays 0) in one call
    check(lib().latec_rans_decode_codes_levels(ptr(words), len(words), ptr(levels), ptr(cdfs),
                                               cdfs.shape[1] - 1, int(tables.radius), kexp,
                                               ptr(out), out.itemsize, len(levels)))
    return out


def model_bits(codes, levels64, tables: RansTables, *, kexp=None, shape: int = 0) -> float:
    """rans.py:332-357."""
    codes = np.asarray(codes, np.uint32).ravel()
    levels = np.asarray(levels64, np.uint8).ravel()
    if codes.shape != levels.shape:
        raise ValueError("codes and levels must have the same flattened length")
    if len(codes) == 0:
        return 0.0
    kexp = full_kexp(tables.radius) if kexp is None else int(kexp)
    cdfs = build_cdfs(kexp, shape, tables.precision, 0)
    syms = _narrow(codes, tables.radius, kexp)
    probs = (cdfs[levels, syms + 1].astype(np.int64) - cdfs[levels, syms].astype(np.int64))
    return float((-np.log2(probs / tables.total)).sum())


def literal_dtype(kexp: int) -> np.dtype:
    return np.dtype(np.uint8 if int(kexp) <= 7 else np.uint16)


def narrow_literal(codes, radius: int, kexp: int) -> np.ndarray:
    return _narrow(codes, radius, kexp).astype(literal_dtype(kexp))


def widen_literal(buf, n_codes: int, radius: int, kexp: int) -> np.ndarray:
    syms = np.frombuffer(buf, literal_dtype(kexp), count=int(n_codes))
    return _widen(syms, radius, kexp)
