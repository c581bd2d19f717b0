"""Dyadic stage schedule (the API of latec/levels.py).

Stage 0 reveals the anchors (``anchor_block**ndim`` blocks on the
``anchor_stride`` lattice).  Each dyadic level at stride ``s`` then reveals,
in order, the points that are odd multiples of ``s`` on exactly one axis (one
sub-stage per axis) and then, fused per weight, the points odd on ``w >= 2``
axes (levels.py:1-27, 100-147).  ``stage_layouts`` describes every stage as
parity classes whose grid positions and stage-array ranks are affine in the
class index (levels.py:389-574); the GPU kernels use the native copy of the
same layouts (liblatec's plan).
"""

from __future__ import annotations

import itertools

import numpy as np


def _check_plan(shape, levels, anchor_stride, anchor_block):
    """Validation with the reference's messages (levels.py:54-76)."""
    shape = tuple(int(n) for n in shape)
    if len(shape) < 1:
        raise ValueError("shape must have at least one axis")
    if levels < 1:
        raise ValueError("levels must be >= 1")
    if anchor_stride < 2 or anchor_stride & (anchor_stride - 1):
        raise ValueError("anchor_stride must be a power of two >= 2")
    if not 1 <= anchor_block <= anchor_stride:
        raise ValueError("anchor_block must be in [1, anchor_stride]")
    if (1 << levels) < anchor_stride:
        raise ValueError(
            f"levels={levels} too small for anchor_stride={anchor_stride}: need levels >= "
            f"log2(anchor_stride) = {anchor_stride.bit_length() - 1} to densify to stride 1"
        )
    return shape


def n_stages(ndim: int, levels: int) -> int:
    return 1 + levels * (2 * ndim - 1)


def _lattice_sel(n: int, start: int, step: int) -> np.ndarray:
    sel = np.zeros(n, bool)
    sel[start::step] = True
    return sel


def iter_stage_plan(shape, levels, anchor_stride, anchor_block=1):
    """(mask, stride, candidate axes) per stage, one mask at a time."""
    shape = _check_plan(shape, levels, anchor_stride, anchor_block)
    ndim = len(shape)
    covered = np.zeros(shape, bool)
    anchor_axes = [np.arange(n) % anchor_stride < anchor_block for n in shape]
    anchor = np.ones(shape, bool)
    for j, sel in enumerate(anchor_axes):
        anchor &= sel.reshape([-1 if k == j else 1 for k in range(ndim)])
    covered |= anchor
    yield anchor, anchor_stride, ()
    for k in range(1, levels + 1):
        s = max(anchor_stride >> k, 1)
        odd = [_lattice_sel(n, s, 2 * s) for n in shape]
        even = [_lattice_sel(n, 0, 2 * s) for n in shape]

        def combo(axes):
            m = np.ones(shape, bool)
            for j in range(ndim):
                sel = odd[j] if j in axes else even[j]
                m &= sel.reshape([-1 if q == j else 1 for q in range(ndim)])
            return m

        for axis in range(ndim):
            mask = combo((axis,)) & ~covered
            if k == levels and ndim == 1:
                mask |= ~covered
            covered |= mask
            yield mask, s, (axis,)
        for w in range(2, ndim + 1):
            mask = np.zeros(shape, bool)
            for axes in itertools.combinations(range(ndim), w):
                mask |= combo(axes)
            mask &= ~covered
            if k == levels and w == ndim:
                mask |= ~covered
            covered |= mask
            yield mask, s, tuple(range(ndim))


def stage_plan(shape, levels, anchor_stride, anchor_block=1):
    return list(iter_stage_plan(shape, levels, anchor_stride, anchor_block))


def stage_masks(shape, levels, anchor_stride, anchor_block=1):
    return [m for m, _, _ in iter_stage_plan(shape, levels, anchor_stride, anchor_block)]


class LazyStageMasks:
    """Sized, re-iterable stream of stage masks (levels.py:150-177)."""

    def __init__(self, shape, levels, anchor_stride, anchor_block=1):
        self.shape = _check_plan(shape, levels, anchor_stride, anchor_block)
        self.levels = int(levels)
        self.anchor_stride = int(anchor_stride)
        self.anchor_block = int(anchor_block)

    def __len__(self):
        return n_stages(len(self.shape), self.levels)

    def __iter__(self):
        for mask, _, _ in iter_stage_plan(self.shape, self.levels, self.anchor_stride,
                                          self.anchor_block):
            yield mask


_MASK_TILE = 1 << 20


def mask_tiles(pos, shape, tile: int = _MASK_TILE):
    """Flat indices of a mask (C order) in tiles; pos=None means all points."""
    if pos is None:
        n = int(np.prod(shape))
        for a in range(0, n, tile):
            yield np.arange(a, min(a + tile, n), dtype=np.int64)
        return
    idx = np.flatnonzero(pos.reshape(-1))
    for a in range(0, len(idx), tile):
        yield idx[a:a + tile]


def _slice_count(n: int, start: int, step: int) -> int:
    return 0 if n <= start else (n - start + step - 1) // step


def stage_strides(ndim: int, levels: int, anchor_stride: int):
    if ndim < 1:
        raise ValueError("shape must have at least one axis")
    _check_plan((1,) * ndim, levels, anchor_stride, 1)
    per = 2 * ndim - 1
    return [anchor_stride] + [max(anchor_stride >> k, 1) for k in range(1, levels + 1)
                              for _ in range(per)]


def stage_counts(shape, levels, anchor_stride, anchor_block=1):
    """Points per stage (levels.py:221-282); closed form for anchor_block 1."""
    shape = _check_plan(shape, levels, anchor_stride, anchor_block)
    ndim = len(shape)
    if anchor_block != 1:
        return [int(m.sum()) for m, _, _ in
                iter_stage_plan(shape, levels, anchor_stride, anchor_block)]
    counts = [int(np.prod([_slice_count(n, 0, anchor_stride) for n in shape]))]
    prev = None
    for k in range(1, levels + 1):
        s = max(anchor_stride >> k, 1)
        if s == prev:
            counts += [0] * (2 * ndim - 1)
            continue
        prev = s
        for axis in range(ndim):
            counts.append(int(np.prod([_slice_count(n, s if j == axis else 0, 2 * s)
                                       for j, n in enumerate(shape)])))
        for w in range(2, ndim + 1):
            counts.append(sum(int(np.prod([_slice_count(n, s if j in axes else 0, 2 * s)
                                           for j, n in enumerate(shape)]))
                              for axes in itertools.combinations(range(ndim), w)))
    counts[-1] += int(np.prod(shape)) - sum(counts)
    return counts


def point_levels(coords, levels, anchor_stride, anchor_block=1):
    """Dyadic level revealing each point (levels.py:329-357)."""
    coords = [np.asarray(c, np.int64) for c in coords]
    out = np.full(coords[0].shape, levels, np.int8)
    anchor = np.ones(coords[0].shape, bool)
    for c in coords:
        anchor &= (c % anchor_stride) < anchor_block
    for k in range(levels - 1, 0, -1):
        s = max(anchor_stride >> k, 1)
        on = np.ones(coords[0].shape, bool)
        for c in coords:
            on &= (c % s) == 0
        out[on] = k
    out[anchor] = 0
    return out


def stage_ebs(shape, levels, anchor_stride, anchor_block, eb, eb_ratio):
    """Per-stage bound eb * ratio**log2(stride / finest) (levels.py:360-386),
    evaluated with the same Python/NumPy scalar operations."""
    strides = stage_strides(len(shape), levels, anchor_stride)
    if not 1 <= anchor_block <= anchor_stride:
        raise ValueError("anchor_block must be in [1, anchor_stride]")
    finest = min(strides)
    return [eb * eb_ratio ** np.log2(stride / finest) for stride in strides]


class ClassLayout:
    """One parity class: grid slices (starts/step/sizes) and its affine view
    (offset/strides) of the stage array (levels.py:410-448)."""

    __slots__ = ("parity", "axes", "starts", "step", "sizes", "strides", "offset")

    def __init__(self, parity, starts, step, sizes, strides, offset):
        self.parity = tuple(parity)
        self.axes = tuple(j for j, p in enumerate(parity) if p)
        self.starts, self.step, self.sizes = tuple(starts), step, tuple(sizes)
        self.strides, self.offset = tuple(strides), offset

    @property
    def size(self) -> int:
        return int(np.prod(self.sizes)) if self.sizes else 1

    def slices(self, origins=None):
        o = origins if origins is not None else (0,) * len(self.starts)
        return tuple(slice(a + b, a + b + self.step * n, self.step)
                     for a, b, n in zip(o, self.starts, self.sizes))


class StageLayout:
    __slots__ = ("stride", "weight", "axes", "classes", "size")

    def __init__(self, stride, weight, axes, classes, size):
        self.stride, self.weight, self.axes = stride, weight, axes
        self.classes, self.size = classes, size


def _row_major(sizes):
    st = [1] * len(sizes)
    for j in range(len(sizes) - 2, -1, -1):
        st[j] = st[j + 1] * sizes[j + 1]
    return st


def _fused(shape, s, w, low_axes):
    ndim = len(shape)
    cnt = []
    for j, n in enumerate(shape):
        base = 2 * s if j in low_axes else 0
        cnt.append((_slice_count(n, base, 2 * s), _slice_count(n, s, 2 * s)))
    F = [[0] * (w + 1) for _ in range(ndim + 1)]
    G = [[0] * (w + 1) for _ in range(ndim + 1)]
    F[ndim][0] = 1
    for j in range(ndim - 1, -1, -1):
        for r in range(w + 1):
            odd = F[j + 1][r - 1] if r else 0
            F[j][r] = cnt[j][0] * F[j + 1][r] + cnt[j][1] * odd
            G[j][r] = F[j + 1][r] + odd
    out = []
    for combo in itertools.combinations(range(ndim), w):
        parity = tuple(int(j in combo) for j in range(ndim))
        starts, sizes, strides = [], [], []
        offset, r = 0, w
        for j, p in enumerate(parity):
            strides.append(G[j][r])
            starts.append(s if p else (2 * s if j in low_axes else 0))
            sizes.append(cnt[j][p])
            if j in low_axes:
                offset += (G[j][r] if not p else 0) - (F[j + 1][r] if not p else 0)
            elif p:
                offset += F[j + 1][r]
            r -= p
        out.append(ClassLayout(parity, starts, 2 * s, sizes, strides, offset))
    return out, F[0][w]


def stage_layouts(shape, levels, anchor_stride, low_axes=()):
    """The schedule as parity-class layouts (anchor_block 1); ``low_axes``
    drops each listed axis's coordinate-0 hyperplane (levels.py:515-574)."""
    shape = _check_plan(shape, levels, anchor_stride, 1)
    ndim = len(shape)
    low_axes = tuple(sorted(set(low_axes)))
    a_starts = [anchor_stride if j in low_axes else 0 for j in range(ndim)]
    a_sizes = [_slice_count(n, b, anchor_stride) for n, b in zip(shape, a_starts)]
    anc = ClassLayout((0,) * ndim, a_starts, anchor_stride, a_sizes, _row_major(a_sizes), 0)
    out = [StageLayout(anchor_stride, 0, (), [anc], anc.size)]
    prev = None
    for k in range(1, levels + 1):
        s = max(anchor_stride >> k, 1)
        if s == prev:
            out += [StageLayout(s, w, (), [], 0) for w in [1] * ndim + list(range(2, ndim + 1))]
            continue
        prev = s
        for axis in range(ndim):
            parity = tuple(int(j == axis) for j in range(ndim))
            starts = [s if j == axis else (2 * s if j in low_axes else 0) for j in range(ndim)]
            sizes = [_slice_count(n, b, 2 * s) for n, b in zip(shape, starts)]
            cl = ClassLayout(parity, starts, 2 * s, sizes, _row_major(sizes), 0)
            out.append(StageLayout(s, 1, (axis,), [cl], cl.size))
        for w in range(2, ndim + 1):
            classes, size = _fused(shape, s, w, low_axes)
            out.append(StageLayout(s, w, tuple(range(ndim)), classes, size))
    return out
