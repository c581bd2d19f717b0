"""GPU parity of the sheet kernels against the oracle.

With the "sheet" knob on, a 3-D level above the first whose lines (last
axis) have an even length of at most 512 points runs as two launches: the
even sheets (i even), then the odd sheets.  Every case here compares the
codes of every stage, the reconstruction bits and the payload with the
oracle's, then the whole stream and its decode (decode through the sheet
kernels too).  The cases cover odd and even sheet counts and row counts,
short last segments (segments of 32 and of 4 odd rows), lines of 4 to 512
points (runs of pairs across line ends, idle lanes), every end rule along
i, j and k, the three centre rules, and outliers (spikes, NaN and
infinities); linear cases check that those levels keep the pair kernels.
"""

import numpy as np
import pytest

from tests.conftest import has_gpu

pytestmark = pytest.mark.gpu

if has_gpu():
    import torch

    import latec_amd
    from latec_amd import _device as D
    from latec_amd import autotune, codec
    from latec_amd._lib import check, lib
    from latec_amd.predictor import InterpPredictor, get_plan
    from latec_amd.quantizer import opts_for
from oracle import oracle as O


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip("no GPU")


def field(shape, seed=0, rough=0.0, spikes=0):
    rng = np.random.RandomState(seed)
    x = rng.randn(*shape)
    for ax in range(len(shape)):
        x = np.cumsum(x, axis=ax)
    x = x / np.abs(x).max()
    if rough:
        x = x + rough * rng.randn(*shape)
    x = x.astype(np.float32)
    if spikes:
        idx = rng.choice(x.size, spikes, replace=False)
        x.flat[idx] = (rng.randn(spikes) * 1e5).astype(np.float32)
        x.flat[idx[0]] = np.nan
        x.flat[idx[1]] = np.inf
        x.flat[idx[2]] = -np.inf
    return x


CASES = [
    # shape, levels, stride, order, center, end, eb_rel, rough, spikes
    # the 512^3 schedule's line lengths: 512, 256, 128 and 64 points
    ((33, 40, 512), 5, 32, "cubic", 0, 1, 1e-3, 0.02, 0),
    # 65 pairs per line: three lines per workgroup, 61 idle lanes
    ((45, 70, 130), 3, 8, "cubic", 0, 1, 1e-3, 0.05, 0),
    # odd row count (131): one more even row than odd rows
    ((16, 131, 66), 4, 16, "cubic", 1, 3, 1e-3, 0.1, 0),
    ((33, 9, 256), 3, 8, "cubic", 2, 0, 1e-3, 0.3, 40),
    ((50, 66, 12), 2, 4, "linear", 0, 2, 1e-3, 0.02, 0),
    ((36, 30, 64), 3, 8, "cubic", 0, 3, 1e-3, 0.05, 12),
    ((40, 36, 8), 3, 8, "linear", 2, 1, 1e-3, 0.02, 0),
    ((7, 5, 4), 2, 4, "cubic", 0, 2, 1e-3, 0.0, 0),
    ((66, 70, 130), 4, 16, "cubic", 0, 1, 1e-3, 0.3, 25),
]


def sheet_levels(shape, levels, order="cubic"):
    """Levels (2 .. levels) that run as sheet launches (cubic interpolation
    only): the level at lattice step 2^(levels - k) codes the grid x[::step], of
    ceil(n / step) points per axis; its last axis must be even and at most 512,
    the others at least 2."""
    if order != "cubic":
        return []
    out = []
    for k in range(2, levels + 1):
        step = 1 << (levels - k)
        ns = [-(-n // step) for n in shape]
        if ns[2] % 2 == 0 and ns[2] <= 512 and ns[0] >= 2 and ns[1] >= 2:
            out.append(k)
    return out


def encode_against_oracle(x, levels, stride, order, center, end, eb):
    shape = x.shape
    xs = x.reshape(1, *shape)
    ebs = O.stage_ebs_ref(3, levels, stride, eb, 0.9)
    ref_payload, ref_recon, _ = O.compress_payload(xs, ebs, levels, stride, 1, order, center,
                                                   end)
    plan = get_plan(shape, 1, levels, stride, 1)
    o = opts_for(1 << 15, cubic=order == "cubic", center=center, end_mode=end)
    res = codec.encode_device(D.to_device(xs.reshape(-1)), plan, ebs, o)
    torch.cuda.synchronize()
    recon = D.to_host(res.d_recon).reshape(xs.shape)
    assert np.array_equal(recon.view(np.uint32), ref_recon.view(np.uint32))
    codes = D.to_host(res.d_codes, np.uint16)
    off = 0
    for s in range(plan.n_stages):
        c, _, off = O.unpack_stage(ref_payload, off, plan.counts[s])
        lo, n = plan.offsets[s], plan.counts[s]
        assert np.array_equal(codes[lo:lo + n].astype(np.uint32), c), f"stage {s}"
    fits = codec.fit_device(res)
    payload, _, _ = codec.pack_device(res, fits, 9)
    assert payload == ref_payload


@pytest.mark.parametrize("blocks", [1, 1 << 20], ids=["seg32", "seg4"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-L{c[1]}-{c[3]}-c{c[4]}e{c[5]}")
def test_gpu_sheet_kernels_bit_exact(case, blocks):
    shape, levels, stride, order, center, end, rel, rough, spikes = case
    print("case", case, "sheet_blocks", blocks, flush=True)
    x = field(shape, seed=sum(shape), rough=rough, spikes=spikes)
    fin = x[np.isfinite(x)]
    eb = rel * float(fin.max() - fin.min())
    nsheet = 2 * len(sheet_levels(shape, levels, order))
    old = autotune.get()
    autotune.set_knobs(sheet=3, sheet_blocks=blocks, sheet_min_points=0, slab=0)
    try:
        check(lib().latec_tune_set(b"sheet_launches", 0))
        encode_against_oracle(x, levels, stride, order, center, end, eb)
        assert int(lib().latec_tune_get(b"sheet_launches")) == nsheet
        pred = InterpPredictor(order, levels, stride, 1)
        pred.center, pred.end_mode = center, end
        stream, _ = latec_amd.compress(x, eb, pred, levels=levels, anchor_stride=stride,
                                       anchor_block=1)
        ref_stream, ref_recon = O.compress(x, eb, order, levels, stride, 1, center, end)
        assert stream == ref_stream
        check(lib().latec_tune_set(b"sheet_launches", 0))
        dec = latec_amd.decompress(stream)
        assert np.array_equal(dec.view(np.uint32), ref_recon[0].view(np.uint32))
        assert int(lib().latec_tune_get(b"sheet_launches")) == nsheet
    finally:
        autotune.set_knobs(**old)
