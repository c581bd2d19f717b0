"""Host-side logic of the GNN codec (CPU): the reference's known answers for
schedule depth, chunking, bounds, gate choice and stream framing."""

import numpy as np
import pytest

from latec_amd import gnn_codec as C
from latec_amd import gnn_model as GM
from oracle import gnn_oracle as G


@pytest.mark.parametrize("shape,expected", [
    ((64, 64, 64, 64), 5), ((128, 128, 128, 128), 5), ((256, 256, 256), 6),
    ((2048, 2048), 8), ((4, 4, 4, 100000), 5), ((256, 256, 256, 256, 256), 4),
    ((512, 512), 8), ((64, 64), 5), ((32, 32), 4), ((16, 16, 16, 16), 3),
    ((4,), 1), ((1,), 1),
])
def test_auto_levels(shape, expected):
    """tests/test_gnn_codec.py:66-90 of the reference."""
    assert C._auto_levels(shape) == expected


def test_auto_levels_3d_aggregation():
    assert C._auto_levels((256, 256, 256), 1) == 7
    assert C._auto_levels((256, 256, 256), 2) == 6


def test_per_step_eb_ratio():
    for coarse in (0.41, 0.65, 0.25):
        for levels in (2, 5, 7, 9):
            assert C._per_step_eb_ratio(coarse, levels) ** (levels - 1) == pytest.approx(coarse)
    assert C._per_step_eb_ratio(0.41, 1) == 1.0
    assert C._GNN_EB_COARSE_FACTOR == pytest.approx(0.9)


def test_auto_chunk_edges():
    """Reference KAT (_auto_chunk_edges((64,)*4, 32) == (32,)*4) and the
    budget invariants; the MI355X default keeps 64^4 as one chunk."""
    assert C._auto_chunk_edges((64, 64, 64, 64), 32) == (32, 32, 32, 32)
    assert C._auto_chunk_edges((64, 64, 64, 64), 8) == (32, 32, 32, 32)
    big = (1 << 12, 1 << 12)
    e = C._auto_chunk_edges(big, 16)
    assert np.prod([min(a, n) for a, n in zip(e, big)]) <= 1 << 21
    assert np.prod([min(a + 16, n) for a, n in zip(e, big)]) > 1 << 21
    assert C.GNNCodec._chunk_edges((64,) * 4, 8, None) == (64,) * 4
    assert C.GNNCodec._chunk_edges((20, 23), 4, None) == (20, 24)
    with pytest.raises(ValueError, match="multiple of anchor_stride"):
        C.GNNCodec._chunk_edges((64, 64), 8, 12)


def test_roundtrip_slack():
    assert C.roundtrip_slack(0.0, 1.0) == pytest.approx(C._F32_U * 4.0, rel=1e-9)
    assert C.roundtrip_slack(1000.0, 1001.0) > C.roundtrip_slack(0.0, 1.0)
    assert C.roundtrip_slack(5.0, 5.0) == 0.0
    assert C.roundtrip_slack(0.14, 1.0) < 1e-6


def test_gate_tables_match_oracle_bisection():
    for eb in (1e-3, 3.7e-5, 0.02):
        thr, grid = C._gate_scale_thresholds(eb)
        t2, g2 = G.gate_scale_tables(eb)
        assert np.array_equal(thr, t2)
        np.testing.assert_allclose(grid, g2, rtol=1e-15)


def test_gate_rate_proxy_outlier_and_cap():
    """Reference KAT: an outlier costs 56 bits, a regular code at most 24."""
    eb, radius = 1e-3, 8
    reg = G.laplace_bits([2 * eb * (radius - 1)], [eb], eb, radius)
    out = G.laplace_bits([2 * eb * radius], [eb], eb, radius)
    assert reg[0] <= 24.0 and out[0] == 56.0


@pytest.mark.parametrize("winner", [1, 2, 3])
def test_gate_selects_each_candidate(winner):
    """tests/test_gnn_codec_chunked.py:177-198 of the reference."""
    eb, n = 1e-3, 32
    x = np.zeros(n)
    cands = [np.full(n, 5 * eb) for _ in range(3)]
    cands[winner - 1] = np.zeros(n)
    W = G.gate_costs(x, np.full(n, 10 * eb), np.full(n, eb, np.float32), cands, eb)
    kind, _d, T, _s = C._gate_select(W)
    assert kind == winner and T > 0


def test_gate_selects_high_side():
    """tests/test_gnn_codec_chunked.py:201-225 of the reference."""
    eb, n = 1e-3, 64
    scale = np.concatenate([np.full(n // 2, eb), np.full(n // 2, eb * 4096)]).astype(np.float32)
    gnn = np.concatenate([np.zeros(n // 2), np.full(n // 2, 50 * eb)])
    cubic = np.concatenate([np.full(n // 2, 50 * eb), np.zeros(n // 2)])
    other = np.full(n, 50 * eb)
    W = G.gate_costs(np.zeros(n), gnn, scale, [cubic, other, other], eb)
    kind, direction, T, _ = C._gate_select(W)
    assert kind == 1 and direction == 1 and T > 0


def test_gate_off_when_model_wins():
    eb, n = 1e-3, 40
    W = G.gate_costs(np.zeros(n), np.zeros(n), np.full(n, eb, np.float32),
                     [np.full(n, 9 * eb)] * 3, eb)
    assert C._gate_select(W) == (0, 0, 0, 0)


def test_pack_gates_roundtrip():
    gates = [0, 0, 256, 1 << 10 | 3 << 8 | 13 << 4 | 6, 0, 4095, 0, 0, 0]
    assert C._unpack_gates(C._pack_gates(gates)) == gates


def test_stream_prefix_and_errors():
    s = C._write_stream({"b": 1, "a": [1, 2]}, b"payload", 3)
    assert s[:8] == b"LATECGNN"
    meta, payload = C._read_stream(s)
    assert meta == {"a": [1, 2], "b": 1} and payload == b"payload"
    assert b'{"a":[1,2],"b":1}' in s  # sorted compact JSON
    with pytest.raises(ValueError, match="bad magic"):
        C._read_stream(b"XXXXXXXX" + s[8:])
    bad = bytearray(s)
    bad[8] = 13
    with pytest.raises(ValueError, match="unsupported LATEC GNN stream version 13"):
        C._read_stream(bytes(bad))
    with pytest.raises(ValueError, match="not a LATEC GNN stream"):
        C._read_stream(b"LAT")


@pytest.mark.parametrize("cshape,low", [((9, 9), (0,)), ((9, 17), (0, 1)), ((5, 9, 9), (1, 2)),
                                        ((5, 5, 3, 5), (0, 3))])
def test_chunk_split_rows_match_masks(cshape, low):
    """Coded rows of a chunk with inherited low faces = the stage points whose
    coordinate on every low axis is > 0, in the full stage's order."""
    levels, stride = 2, 4
    sched = C._ChunkSchedule(cshape, levels, stride, low)
    from oracle import oracle as O

    for s, (mask, _, _, _) in enumerate(O.stage_plan(cshape, levels, stride, 1)):
        pts = np.stack(np.nonzero(mask), 1)
        keep = np.all(pts[:, list(low)] > 0, axis=1)
        assert sched.full_counts[s] == len(pts)
        assert sched.counts[s] == int(keep.sum())
        want = np.flatnonzero(keep)
        got = np.arange(len(pts)) if sched.keep[s] is None else sched.keep[s]
        assert np.array_equal(got, want), s


def test_checkpoint_format(tmp_path):
    import torch

    d, agg = 32, 2
    P = GM.init_params(d, agg, 0)
    path = tmp_path / "m.pt"
    GM.save_checkpoint(path, d, agg, P)
    d2, a2, Q = GM.load_checkpoint(path)
    assert (d2, a2) == (d, agg)
    assert all(np.array_equal(P[k], Q[k]) for k in P)
    assert GM.pack(P, d, agg).size == sum(int(np.prod(s)) for _, s in GM.param_shapes(d, agg))
    from latec_amd._lib import lib

    assert lib().latec_gnn_param_count(d, agg) == GM.pack(P, d, agg).size
    assert lib().latec_gnn_param_count(64, 1) == 102710  # v7-d64-1agg
    for version in (None, 6):
        ck = {"d": d, "agg_level": agg, "state_dict": {}}
        if version is not None:
            ck["version"] = version
        torch.save(ck, tmp_path / "old.pt")
        with pytest.raises(ValueError, match="format v7"):
            GM.load_checkpoint(tmp_path / "old.pt")
