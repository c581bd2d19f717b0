"""The C-ABI library loads and exports every symbol include/latec.h
declares; its host-side coder matches the oracle word for word."""

import os
import re
import struct

import numpy as np
import pytest

from latec_amd import _lib
from oracle import oracle as O

HDR = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "latec.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"\b(latec_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported_and_bound():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, n


def test_rans_words_match_oracle():
    rng = np.random.RandomState(3)
    for kexp, level in ((1, 0), (4, 20), (7, 33), (10, 50)):
        cdf = O.build_cdfs(kexp, 3, 24, 2)[level]
        pmf = np.diff(cdf.astype(np.int64))
        syms = rng.choice(len(pmf), 30000, p=pmf / pmf.sum()).astype(np.uint32)
        ref = O.rans_encode_syms(syms, cdf)
        words = np.empty(len(syms) + 2, np.uint32)
        nw = _lib.check(_lib.lib().latec_rans_encode(_lib.ptr(syms), len(syms), _lib.ptr(cdf),
                                                     len(cdf) - 1, _lib.ptr(words), len(words)))
        assert np.array_equal(words[:nw], ref)
        out = np.empty(len(syms), np.uint32)
        _lib.check(_lib.lib().latec_rans_decode(_lib.ptr(ref), len(ref), _lib.ptr(cdf),
                                                len(cdf) - 1, _lib.ptr(out), len(out)))
        assert np.array_equal(out, syms)


@pytest.mark.parametrize("kexp,n,nlev", [(1, 7, 3), (4, 5000, 64), (7, 40000, 12), (10, 20000, 40)])
def test_rans_levels_words_match_oracle(kexp, n, nlev):
    """Per-point levels (the GNN stages' coder): the library's words equal the
    oracle's level loop, decode back, and with one model for every level equal
    the single-model coder on the symbols taken level by level (stable)."""
    rng = np.random.RandomState(kexp + n)
    cdfs = O.build_cdfs(kexp, 2, 24, 1)
    A = cdfs.shape[1] - 1
    levels = rng.randint(0, 64, n).astype(np.uint8) % nlev * (64 // nlev)
    syms = np.empty(n, np.uint32)
    for l in np.unique(levels):
        pmf = np.diff(cdfs[l].astype(np.int64))
        sel = levels == l
        syms[sel] = rng.choice(A, int(sel.sum()), p=pmf / pmf.sum())
    ref = O.rans_encode_levels(syms, levels, cdfs)
    words = np.empty(n + 2, np.uint32)
    nw = _lib.check(_lib.lib().latec_rans_encode_levels(
        _lib.ptr(syms), _lib.ptr(levels), n, _lib.ptr(cdfs), A, _lib.ptr(words), len(words)))
    assert np.array_equal(words[:nw], ref)
    out = np.empty(n, np.uint32)
    _lib.check(_lib.lib().latec_rans_decode_levels(_lib.ptr(ref), len(ref), _lib.ptr(levels),
                                                   _lib.ptr(cdfs), A, _lib.ptr(out), n))
    assert np.array_equal(out, syms)
    one = np.ascontiguousarray(np.repeat(cdfs[20:21], 64, axis=0))
    ok = syms < A
    same = O.rans_encode_levels(syms, levels, one)
    order = np.argsort(levels, kind="stable")
    assert np.array_equal(same, O.rans_encode_syms(syms[order], one[0])), "grouping order"
    nw = _lib.check(_lib.lib().latec_rans_encode_levels(
        _lib.ptr(syms), _lib.ptr(levels), n, _lib.ptr(one), A, _lib.ptr(words), len(words)))
    assert ok.all() and np.array_equal(words[:nw], same)


@pytest.mark.parametrize("kexp,radius,n", [(1, 512, 9), (4, 512, 6000), (9, 512, 30000),
                                            (7, 32768, 20000), (12, 1 << 17, 5000)])
def test_rans_codes_levels_narrow_then_code(kexp, radius, n):
    """latec_rans_encode_codes_levels (the quantizer's codes narrowed inside
    the coder) writes the oracle's words for the narrowed symbols
    (rans.py:83-89: code 0 -> 0, else code - radius + 2^kexp), from u16 and
    u32 codes; latec_code_kexp is choose_kexp's bit length; a code outside the
    narrowed alphabet and a level >= 64 are rejected."""
    L = _lib.lib()
    rng = np.random.RandomState(kexp * 7 + n)
    cdfs = O.build_cdfs(kexp, 1, 24, 0)
    A = cdfs.shape[1] - 1
    span = (1 << kexp) - 1
    levels = rng.randint(0, 64, n).astype(np.uint8)
    syms = rng.randint(0, A, n).astype(np.uint32)
    codes = np.where(syms == 0, 0, syms.astype(np.int64) - span - 1 + radius).astype(np.uint32)
    assert codes.min() >= 0 and (codes[syms > 0] > 0).all()
    ref = O.rans_encode_levels(syms, levels, cdfs)
    kinds = [codes] + ([codes.astype(np.uint16)] if radius <= 32768 else [])
    for c in kinds:
        words = np.empty(n + 2, np.uint32)
        nw = _lib.check(L.latec_rans_encode_codes_levels(
            _lib.ptr(c), c.itemsize, _lib.ptr(levels), n, radius, kexp, _lib.ptr(cdfs), A,
            _lib.ptr(words), len(words)))
        assert np.array_equal(words[:nw], ref)
        small = np.empty(4, np.uint32)  # short output array: the coder's own buffer
        assert L.latec_rans_encode_codes_levels(
            _lib.ptr(c), c.itemsize, _lib.ptr(levels), n, radius, kexp, _lib.ptr(cdfs), A,
            _lib.ptr(small), len(small)) == _lib.LATEC_E_CAPACITY or len(ref) <= 4
        nz = c[c != 0].astype(np.int64)
        want = max(1, int(np.abs(nz - radius).max()).bit_length()) if len(nz) else 1
        assert L.latec_code_kexp(_lib.ptr(c), c.itemsize, n, radius) == want
        # and back: the widening decoder returns the codes in the same width
        back = np.full(n, 7, c.dtype)
        _lib.check(L.latec_rans_decode_codes_levels(
            _lib.ptr(ref), len(ref), _lib.ptr(levels), _lib.ptr(cdfs), A, radius, kexp,
            _lib.ptr(back), c.itemsize, n))
        assert np.array_equal(back, c)
    if radius + span > 0xffff:  # a u16 code that does not fit is a format error
        back = np.empty(n, np.uint16)
        rc = L.latec_rans_decode_codes_levels(_lib.ptr(ref), len(ref), _lib.ptr(levels),
                                              _lib.ptr(cdfs), A, radius, kexp, _lib.ptr(back),
                                              2, n)
        assert rc == _lib.LATEC_E_FORMAT and "16 bits" in _lib.last_error()
    for bad in (radius + span + 1, radius - span - 2):  # one past each end
        if bad <= 0:
            continue
        c = codes.copy()
        c[n // 2] = bad
        words = np.empty(n + 2, np.uint32)
        rc = L.latec_rans_encode_codes_levels(_lib.ptr(c), 4, _lib.ptr(levels), n, radius, kexp,
                                              _lib.ptr(cdfs), A, _lib.ptr(words), len(words))
        assert rc == _lib.LATEC_E_ARG and "narrowed" in _lib.last_error()
    lv = levels.copy()
    lv[-1] = 64
    rc = L.latec_rans_encode_codes_levels(_lib.ptr(codes), 4, _lib.ptr(lv), n, radius, kexp,
                                          _lib.ptr(cdfs), A, _lib.ptr(words), len(words))
    assert rc == _lib.LATEC_E_ARG and "table count" in _lib.last_error()


def test_rans_trailing_data_rejected():
    cdf = O.build_cdfs(2, 0)[10]
    w = O.rans_encode_syms(np.array([1, 2, 3], np.uint32), cdf)
    bad = np.concatenate([[7], w]).astype(np.uint32)
    out = np.empty(3, np.uint32)
    rc = _lib.lib().latec_rans_decode(_lib.ptr(bad), len(bad), _lib.ptr(cdf), len(cdf) - 1,
                                      _lib.ptr(out), 3)
    assert rc == _lib.LATEC_E_FORMAT
    assert "trailing" in _lib.last_error()


def test_zstd_frame_matches_oracle():
    data = np.random.RandomState(0).randint(0, 7, 100000).astype(np.uint8).tobytes()
    cap = _lib.lib().latec_zstd_bound(len(data))
    dst = np.empty(cap, np.uint8)
    src = np.frombuffer(data, np.uint8)
    n = _lib.lib().latec_zstd_compress(_lib.ptr(src), len(src), 9, _lib.ptr(dst), cap)
    assert dst[:n].tobytes() == O._zstd(data, 9)


def test_unpack_stage_of_oracle_records():
    """Records written by the oracle unpack with the native parser."""
    R = 1 << 15
    rng = np.random.RandomState(1)
    for n, sc in ((300, 1.5), (6000, 0.3), (9000, 40.0)):
        codes = (np.rint(rng.laplace(0, sc, n)) + R).astype(np.uint32)
        codes[::53] = 0
        outl = rng.randn(int((codes == 0).sum())).astype(np.float32)
        rec = np.frombuffer(O.pack_stage(codes, outl), np.uint8)
        _, framed = struct.unpack_from("<IQ", rec.tobytes(), 0)
        kexp, shape, level, owx = ((framed >> 45) & 15, (framed >> 49) & 31,
                                   (framed >> 54) & 63, (framed >> 60) & 7)
        row = O.build_cdfs(int(kexp), int(shape), 24, int(owx))[int(level)]
        c = np.empty(n, np.uint16)
        rk = np.empty(n, np.uint32)
        vl = np.empty(n, np.float32)
        import ctypes
        no = ctypes.c_int64()
        used = _lib.check(_lib.lib().latec_unpack_stage(_lib.ptr(rec), len(rec), R, _lib.ptr(row),
                                                        n, _lib.ptr(c), 2, _lib.ptr(rk),
                                                        _lib.ptr(vl), n, ctypes.byref(no)))
        assert used == len(rec)
        assert np.array_equal(c.astype(np.uint32), codes)
        assert np.array_equal(rk[:no.value], np.flatnonzero(codes == 0))
        assert np.array_equal(vl[:no.value], outl)


@pytest.mark.parametrize("kind", ["runs", "iid", "wide"])
def test_pack_stages_large_stage_equals_oracle_record(kind):
    """A stage above a million codes takes the concurrent path (rANS, the
    literal scout and the full-level literal frame side by side, each zstd
    frame fed in runs of blocks so it can stop early); its record must be
    the oracle's serial scout-then-full record byte for byte."""
    import ctypes

    R = 1 << 15
    n = 1_200_000
    rng = np.random.RandomState(5)
    if kind == "runs":  # literal back end wins
        codes = np.repeat(np.arange(n // 400) % 7 + R - 3, 400).astype(np.uint32)
    elif kind == "iid":  # scout passes, the full frame loses to rANS
        codes = (np.rint(rng.laplace(0, 3.0, n)) + R).astype(np.uint32)
    else:  # 16-bit literal symbols
        codes = (np.rint(rng.laplace(0, 60.0, n)) + R).astype(np.uint32)
    codes[::997] = 0
    outl = rng.randn(int((codes == 0).sum())).astype(np.float32)
    want = O.pack_stage(codes, outl)
    kexp = O.choose_kexp(codes, R)
    level, shape, owx, bits = O.choose_const_level(codes, R, kexp)
    fit = (_lib.Fit * 1)()
    fit[0].kexp, fit[0].level, fit[0].shape, fit[0].owx, fit[0].bits = kexp, level, shape, owx, bits
    row = np.ascontiguousarray(O.build_cdfs(kexp, shape, 24, owx)[level], np.uint32)
    c16 = codes.astype(np.uint16)
    cap = 64 + 8 * n + 4 * len(outl) + 1024
    out = np.empty(cap, np.uint8)
    rec_len = np.zeros(1, np.int64)
    z = lambda v: np.array([v], np.int64)
    size = _lib.check(_lib.lib().latec_pack_stages(
        1, _lib.ptr(c16), 2, _lib.ptr(z(0)), _lib.ptr(z(n)), _lib.ptr(outl), _lib.ptr(z(0)),
        _lib.ptr(z(len(outl))), R, 9, ctypes.cast(fit, ctypes.c_void_p), _lib.ptr(row),
        _lib.ptr(z(0)), _lib.ptr(out), cap, _lib.ptr(rec_len), 4))
    assert out[:size].tobytes() == want
    if kind == "runs":
        assert struct.unpack_from("<IQ", want, 0)[1] >> 63 == 1


@pytest.mark.parametrize("kinds,threads", [(("small", "iid", "small2"), 4),
                                           (("runs", "small", "iid"), 4),
                                           (("iid", "runs"), 1)])
def test_pack_stages_framed_equals_pack_then_zstd(kinds, threads):
    """The framed pack (outer zstd frame started on the large stages' rANS
    records, redone when a literal frame wins) writes the payload and frame
    that packing first and framing the payload after would."""
    import ctypes

    R = 1 << 15
    rng = np.random.RandomState(7)
    stages = []
    for kind in kinds:
        if kind == "runs":
            n = 1_200_000
            c = np.repeat(np.arange(n // 400) % 7 + R - 3, 400).astype(np.uint32)
        elif kind == "iid":
            c = (np.rint(rng.laplace(0, 3.0, 1_100_000)) + R).astype(np.uint32)
        else:
            c = (np.rint(rng.laplace(0, 1.5 if kind == "small" else 9.0, 9000)) + R).astype(
                np.uint32)
        c[::991] = 0
        stages.append(c)
    ns = len(stages)
    n_codes = np.array([len(c) for c in stages], np.int64)
    code_off = np.concatenate([[0], np.cumsum(n_codes)[:-1]]).astype(np.int64)
    codes = np.concatenate(stages).astype(np.uint16)
    outs = [rng.randn(int((c == 0).sum())).astype(np.float32) for c in stages]
    n_out = np.array([len(v) for v in outs], np.int64)
    out_off = np.concatenate([[0], np.cumsum(n_out)[:-1]]).astype(np.int64)
    outl = np.concatenate(outs)
    fits = (_lib.Fit * ns)()
    rows = []
    for i, c in enumerate(stages):
        kexp = O.choose_kexp(c, R)
        level, shape, owx, bits = O.choose_const_level(c, R, kexp)
        f = fits[i]
        f.kexp, f.level, f.shape, f.owx, f.bits = kexp, level, shape, owx, bits
        rows.append(np.asarray(O.build_cdfs(kexp, shape, 24, owx)[level], np.uint32))
    row_off = np.concatenate([[0], np.cumsum([len(r) for r in rows])[:-1]]).astype(np.int64)
    cdf = np.ascontiguousarray(np.concatenate(rows), np.uint32)
    cap = int(64 * ns + 4 * (n_codes.sum() + 2 * ns) + 4 * n_out.sum() + (1 << 16))
    P = _lib.ptr
    common = (ns, P(codes), 2, P(code_off), P(n_codes), P(outl), P(out_off), P(n_out), R, 9,
              ctypes.cast(fits, ctypes.c_void_p), P(cdf), P(row_off))
    want = np.empty(cap, np.uint8)
    want_len = np.zeros(ns, np.int64)
    size = _lib.check(_lib.lib().latec_pack_stages(*common, P(want), cap, P(want_len), threads))
    want_frame = O._zstd(want[:size].tobytes(), 9)
    got = np.empty(cap, np.uint8)
    got_len = np.zeros(ns, np.int64)
    fcap = int(_lib.lib().latec_zstd_bound(cap))
    frame = np.empty(fcap, np.uint8)
    flen = ctypes.c_int64()
    gsize = _lib.check(_lib.lib().latec_pack_stages_framed(
        *common, P(got), cap, P(got_len), threads, 9, P(frame), fcap, ctypes.byref(flen)))
    assert gsize == size and got[:size].tobytes() == want[:size].tobytes()
    assert np.array_equal(got_len, want_len)
    assert frame[:flen.value].tobytes() == want_frame


# Hand-derived known answers for the stack rANS coder (the design of
# constriction 0.4's AnsCoder<u32, u64> over Categorical(probs, perfect=False)
# at 24-bit precision, latec/rans.py:466-535, 629-657).  Model: the table row
# cdf = [0, 3, 8, 2^24] (pmf 3, 5, 2^24 - 8).  Leaky quantization: scale =
# f32((2^24 - 3) / 2^24) = 1 - 3 * 2^-24; left[k] = floor(f32(cum_k * scale)) + k
# -> left[1] = floor(f32(3 - 9 * 2^-24) = 3 - 2^-21) + 1 = 3,
#    left[2] = floor(8 - 3 * 2^-21) + 2 = 9, left[3] = 2^24
# (probabilities 3, 6, 2^24 - 9: not the table's 3, 5).  Symbols are pushed
# last to first; a push of s: if state >> 40 >= p(s), emit the low word and
# state >>= 32; then state = (state // p) << 24 | (left[s] + state % p).  The
# seal writes the final state's low word, then its high word if nonzero.
#   [1, 0, 2]: 0 -push 2-> 9 -push 0-> 0x3000000 -push 1-> 2^47 + 3
#              -> words [0x3, 0x8000]
#   [0, 0, 0, 2]: 9 -> 0x3000000 -> 2^48 -push 0: 2^48 >> 40 = 256 >= 3, emit
#              0x0, state 2^16 -> (21845 << 24) | 1 = 0x5555000001
#              -> words [0x0, 0x55000001, 0x55]
# A wrong quantizer (round instead of floor, no +k), push order, word order,
# renormalization test or seal changes at least one of these words.
RANS_KAT_CDF = np.array([0, 3, 8, 1 << 24], np.uint32)
RANS_KATS = [([1, 0, 2], [0x3, 0x8000]),
             ([0, 0, 0, 2], [0x0, 0x55000001, 0x55]),
             # p(2) > state: each push only adds left[2] = 9 -> 9, 18, 27
             ([2, 2, 2], [0x1B])]


@pytest.mark.parametrize("syms,words", RANS_KATS)
def test_rans_known_answers(syms, words):
    syms = np.array(syms, np.uint32)
    cdf = RANS_KAT_CDF
    assert list(O.rans_encode_syms(syms, cdf)) == words
    out = np.empty(len(syms) + 4, np.uint32)
    nw = _lib.check(_lib.lib().latec_rans_encode(_lib.ptr(syms), len(syms), _lib.ptr(cdf),
                                                 len(cdf) - 1, _lib.ptr(out), len(out)))
    assert list(out[:nw]) == words
    assert np.array_equal(O.rans_decode_syms(np.array(words, np.uint32), len(syms), cdf), syms)
    dec = np.empty(len(syms), np.uint32)
    w = np.array(words, np.uint32)
    assert _lib.lib().latec_rans_decode(_lib.ptr(w), len(w), _lib.ptr(cdf), len(cdf) - 1,
                                        _lib.ptr(dec), len(syms)) == 0
    assert np.array_equal(dec, syms)
