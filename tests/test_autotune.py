"""Autotuner knobs: validation on the host, and on the GPU that every launch
shape gives the same encode (the tuner checks it) and keeps oracle parity."""

import numpy as np
import pytest

from latec_amd import autotune
from latec_amd._lib import lib


def test_knobs_round_trip_and_validation():
    start = autotune.get()
    try:
        autotune.set_knobs(march_blocks=1000, march_hist_peel=3, slab=0, pair=0,
                           min_threads=12345, pair_lines=0, pair_quads=0, pair_waves=1,
                           sheet=2, sheet_blocks=77, sheet_min_points=5)
        assert autotune.get() == {"march_blocks": 1000, "march_hist_peel": 3, "slab": 0,
                                  "pair": 0, "pair_lines": 0, "pair_quads": 0,
                                  "min_threads": 12345, "pair_waves": 1, "sheet": 2,
                                  "sheet_blocks": 77, "sheet_min_points": 5}
        with pytest.raises(ValueError):
            autotune.set_knobs(pair_waves=2)
        with pytest.raises(ValueError):
            autotune.set_knobs(sheet=4)
        with pytest.raises(ValueError):
            autotune.set_knobs(sheet_blocks=0)
        with pytest.raises(ValueError):
            autotune.set_knobs(pair=2)
        with pytest.raises(ValueError):
            autotune.set_knobs(min_threads=-1)
        with pytest.raises(ValueError):
            autotune.set_knobs(march_blocks=4)
        with pytest.raises(ValueError):
            autotune.set_knobs(march_hist_peel=9)
        with pytest.raises(ValueError):
            autotune.set_knobs(bogus=1)
        assert lib().latec_tune_get(b"bogus") < 0
    finally:
        autotune.set_knobs(**start)


@pytest.mark.gpu
def test_gpu_autotune_keeps_outputs_and_parity(tmp_path, monkeypatch):
    import latec_amd
    from latec_amd import _device as D
    from latec_amd import codec
    from latec_amd.levels import stage_ebs
    from latec_amd.predictor import InterpPredictor
    from oracle import oracle as O

    monkeypatch.setenv("LATEC_TUNE_CACHE", str(tmp_path / "tune.json"))
    rng = np.random.RandomState(3)
    shape = (72, 48, 136)
    x = np.cumsum(np.cumsum(rng.randn(*shape), 0), 2).astype(np.float32)
    x += 0.02 * rng.randn(*shape).astype(np.float32)
    eb = 1e-3 * float(x.max() - x.min())
    pred = InterpPredictor("cubic", 3, 8, 1)
    o = codec._interp_opts(pred, 3, 1 << 15, False)
    ebs = stage_ebs(shape, 3, 8, 1, eb, 0.9)
    start = autotune.get()
    try:
        print("case autotune 72x48x136", flush=True)
        got = autotune.tune(D.to_device(x.reshape(-1)), shape, 3, 8, ebs, o, reps=1,
                            candidates={"march_blocks": (8, 1024), "march_hist_peel": (0, 2),
                                        "slab": (0, 1), "pair": (0, 1)})
        assert len(got["table"]) == 16 and autotune.get() == got["knobs"]
        assert autotune.load(shape, 3, 8)["knobs"] == got["knobs"]
        for blocks, slab, pair, mt in ((8, 0, 0, 1 << 19), (1024, 1, 0, 1 << 19),
                                       (8, 0, 1, 0), (1024, 0, 1, 1 << 24)):
            autotune.set_knobs(march_blocks=blocks, march_hist_peel=0, slab=slab, pair=pair,
                               min_threads=mt)
            s, _ = latec_amd.compress(x, eb, InterpPredictor("cubic", 3, 8, 1), levels=3,
                                      anchor_stride=8, anchor_block=1)
            ref, _ = O.compress(x, eb, "cubic", 3, 8, 1)
            assert s == ref
    finally:
        autotune.set_knobs(**start)
