"""The no-allocation multi-lane encode (latec_mlrans_encode_noalloc) writes
the same blobs as latec_mlrans_encode with a caller workspace, on two calls
in a row (the second smaller, so the fixed workspace holds the first call's
contents), and refuses stages larger than its fixed workspace before it runs
anything."""

import numpy as np
import pytest

from tests.conftest import has_gpu

pytestmark = pytest.mark.gpu

if has_gpu():
    import torch

    from latec_amd import _device as D
    from latec_amd import codec, mlrans
    from latec_amd._lib import LATEC_E_CAPACITY, check, lib, ptr
    from latec_amd.predictor import get_plan
    from latec_amd.quantizer import opts_for

@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip("no GPU")

def _field(shape, seed):
    rng = np.random.RandomState(seed)
    x = rng.randn(*shape)
    for ax in range(len(shape)):
        x = np.cumsum(x, axis=ax)
    x = x / np.abs(x).max() + 0.01 * rng.randn(*shape)
    return x.astype(np.float32)

def _stage_inputs(shape, seed):
    """Device codes and the per-stage model arguments of one plain encode."""
    x = _field(shape, seed)
    plan = get_plan(shape, 1, 3, 8, 1)
    o = opts_for(1 << 15, cubic=True, center=0, end_mode=1)
    ebs = [1e-3 * float(x.max() - x.min())] * plan.n_stages
    res = codec.encode_device(D.to_device(x.reshape(-1)), plan, ebs, o)
    fits = codec.fit_device(res)
    n, code_off = mlrans._stage_arrays(plan)
    params = [(int(f.kexp) if n[s] else 1, int(f.shape), int(f.owx), int(f.level))
              for s, f in enumerate(fits)]
    kexp = np.array([p[0] for p in params], np.int32)
    d_rows, row_off = mlrans._rows(params, n)
    return res, n, code_off, kexp, d_rows, row_off

def _encode(noalloc, res, n, code_off, kexp, d_rows, row_off):
    ns = len(n)
    lanes = sum(int(lib().latec_mlrans_lanes(int(k))) for k in n)
    cap = 4 * (int(n.sum()) + 3 * lanes + 4 * ns + 16)
    out = D.empty(cap, np.uint8)
    out.fill_(0xA5)
    blob_off = np.zeros(ns, np.int64)
    blob_len = np.zeros(ns, np.int64)
    if noalloc:
        check(lib().latec_mlrans_encode_noalloc(
            ns, ptr(res.d_codes), 2, ptr(code_off), ptr(n), ptr(kexp), 1 << 15, ptr(d_rows),
            ptr(row_off), ptr(out), cap, ptr(blob_off), ptr(blob_len), ptr(None)))
    else:
        sw = int(lib().latec_mlrans_scratch_words(ns, ptr(n)))
        scratch = D.empty(sw, np.uint32)
        check(lib().latec_mlrans_encode(
            ns, ptr(res.d_codes), 2, ptr(code_off), ptr(n), ptr(kexp), 1 << 15, ptr(d_rows),
            ptr(row_off), ptr(scratch), sw, ptr(out), cap, ptr(blob_off), ptr(blob_len),
            ptr(None)))
    total = int((blob_off + blob_len).max(initial=0))
    return blob_off, blob_len, D.to_host(out[:total], np.uint8)

def test_noalloc_encode_equals_workspace_encode_twice():
    for shape, seed in (((96, 80, 64), 1), ((40, 36, 30), 2)):
        print("case noalloc", shape, flush=True)
        args = _stage_inputs(shape, seed)
        ref = _encode(False, *args)
        got = _encode(True, *args)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        assert np.array_equal(got[2], ref[2])
        assert int(args[1].sum()) == int(np.prod(shape))

def test_noalloc_encode_refuses_stages_beyond_its_workspace():
    words = int(lib().latec_mlrans_noalloc_words())
    assert words == 1 << 23
    n = np.array([words], np.int64)
    assert int(lib().latec_mlrans_scratch_words(1, ptr(n))) > words
    code_off = np.zeros(1, np.int64)
    kexp = np.array([6], np.int32)
    row_off = np.zeros(1, np.int64)
    blob_off = np.zeros(1, np.int64)
    blob_len = np.zeros(1, np.int64)
    # refused on the host before any device access: the device pointers are never read
    rc = lib().latec_mlrans_encode_noalloc(1, ptr(None), 2, ptr(code_off), ptr(n), ptr(kexp),
                                           1 << 15, ptr(None), ptr(row_off), ptr(None), 0,
                                           ptr(blob_off), ptr(blob_len), ptr(None))
    assert rc == LATEC_E_CAPACITY
    assert b"fixed" in lib().latec_last_error()
