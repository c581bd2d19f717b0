"""Schedule: the package's Python levels and the native plan (liblatec,
no GPU needed) against brute-force masks (reference tests/test_levels.py,
tests/test_gnn_codec_chunked.py:447-496)."""

import numpy as np
import pytest

from latec_amd import levels as LV
from latec_amd.predictor import Plan
from oracle import oracle as O

CASES = [((33,), 3, 8, 1), ((20, 17), 3, 8, 1), ((16, 16), 4, 16, 4), ((9, 10, 11), 2, 4, 1),
         ((12, 9, 7), 3, 8, 2), ((6, 5, 4, 5), 2, 4, 1), ((8, 8), 5, 16, 1),
         ((5, 6, 3, 4, 3), 1, 2, 1), ((64, 48), 4, 16, 4), ((40, 33, 17), 3, 8, 1)]


@pytest.mark.parametrize("shape,levels,stride,block", CASES)
def test_python_masks_match_oracle(shape, levels, stride, block):
    ours = LV.stage_masks(shape, levels, stride, block)
    ref = [m for m, _, _, _ in O.stage_plan(shape, levels, stride, block)]
    assert len(ours) == len(ref) == LV.n_stages(len(shape), levels)
    for a, b in zip(ours, ref):
        assert np.array_equal(a, b)
    assert LV.stage_counts(shape, levels, stride, block) == [int(m.sum()) for m in ref]


@pytest.mark.parametrize("shape,levels,stride,block", CASES)
def test_native_plan_matches_masks(shape, levels, stride, block):
    """Every class point lands on its mask, at its flatnonzero rank."""
    p = Plan(shape, 1, levels, stride, block)
    ref = [m for m, _, _, _ in O.stage_plan(shape, levels, stride, block)]
    assert p.counts == [int(m.sum()) for m in ref]
    for s, mask in enumerate(ref):
        flat = np.flatnonzero(mask.reshape(-1))
        rank_of = {int(f): i for i, f in enumerate(flat)}
        seen = np.zeros(len(flat), int)
        for c in range(p.n_classes[s]):
            ci = p.class_info(s, c)
            for m in np.ndindex(*ci["sizes"]):
                if s == 0 and block > 1:
                    coord = [(mm // block) * stride + mm % block for mm in m]
                else:
                    coord = [a + ci["step"] * mm for a, mm in zip(ci["starts"], m)]
                f = int(np.ravel_multi_index(coord, shape))
                r = ci["offset"] + sum(mm * st for mm, st in zip(m, ci["strides"]))
                if s > 0 and block > 1 and all(x % stride < block for x in coord):
                    assert f not in rank_of  # anchor, excluded from this stage
                    continue
                if s == 0 or block == 1:
                    assert rank_of[f] == r
                seen[rank_of[f]] += 1
        assert (seen == 1).all()
    p.close()


@pytest.mark.parametrize("shape,levels,stride", [((20, 17), 3, 8), ((9, 10, 11), 2, 4),
                                                  ((6, 5, 4, 5), 2, 4), ((12, 9, 7), 3, 8)])
@pytest.mark.parametrize("low_axes", [(), (0,), (1, 2)])
def test_python_layouts_match_masks(shape, levels, stride, low_axes):
    low = tuple(a for a in low_axes if a < len(shape))
    lay = LV.stage_layouts(shape, levels, stride, low)
    for st, mask in zip(lay, LV.stage_masks(shape, levels, stride, 1)):
        m = mask.copy()
        for a in low:
            idx = [slice(None)] * len(shape)
            idx[a] = 0
            m[tuple(idx)] = False
        flat = np.flatnonzero(m.reshape(-1))
        arr = np.full(len(flat), -1, np.int64)
        for cl in st.classes:
            grid = np.arange(int(np.prod(shape))).reshape(shape)[cl.slices()]
            sv = np.lib.stride_tricks.as_strided(arr[cl.offset:], shape=cl.sizes,
                                                 strides=[8 * x for x in cl.strides])
            sv[...] = grid
        assert np.array_equal(arr, flat)


def test_plan_validation_messages():
    with pytest.raises(ValueError, match="power of two"):
        Plan((8, 8), 1, 3, 6, 1)
    with pytest.raises(ValueError, match="too small"):
        Plan((8, 8), 1, 2, 16, 1)
    with pytest.raises(ValueError, match="anchor_block"):
        Plan((8, 8), 1, 3, 8, 9)
    with pytest.raises(ValueError, match="too small"):
        LV.stage_counts((8, 8), 2, 16)


def test_c2_schedule_counts():
    p = Plan((512, 512, 512), 1, 5, 32, 1)
    assert p.counts == [4096] * 4 + [12288, 4096] + [32768] * 3 + [98304, 32768] + \
        [262144] * 3 + [786432, 262144] + [2097152] * 3 + [6291456, 2097152] + \
        [16777216] * 3 + [50331648, 16777216]
    p.close()
