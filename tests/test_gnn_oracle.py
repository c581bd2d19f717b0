"""The FP64 GNN oracle against the reference's golden vectors, brute force and
closed forms (CPU only)."""

import math
import os

import numpy as np
import pytest
import torch

from oracle import gnn_oracle as G

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "gnn_golden.npz")


def golden_cases():
    z = np.load(GOLDEN)
    names = sorted({k.split("/")[0] for k in z.files})
    out = []
    for n in names:
        meta = z[f"{n}/meta"]
        levels, stride, d, agg, mr, batch = (int(v) for v in meta[:6])
        shape = tuple(int(v) for v in meta[6:])
        P = {k[len(n) + 3:]: torch.from_numpy(z[k]).double()
             for k in z.files if k.startswith(f"{n}/p/")}
        stages = int(z[f"{n}/n_stages"])
        ref = [(z[f"{n}/q{i}"], z[f"{n}/mu{i}"], z[f"{n}/lb{i}"]) for i in range(1, stages)]
        out.append((n, shape, levels, stride, d, agg, mr, float(z[f"{n}/eb"]),
                    z[f"{n}/x"], P, z[f"{n}/q0"], ref))
    return out


def test_half_direction_counts():
    assert len(G.half_directions(2)) == 4
    assert len(G.half_directions(3)) == 13
    assert len(G.half_directions(4, 2)) == 16
    assert len(G.half_directions(2, 1)) == 2
    for nd in range(1, 5):
        assert len(G.half_directions(nd)) == (3 ** nd - 1) // 2


def test_state_index_closed_form():
    """Signed cosines +-1/sqrt(j) and 0 map to 2L+1 ordered states."""
    agg = 3
    cos = torch.tensor([1.0, 1 / math.sqrt(2), 1 / math.sqrt(3), 0.0,
                        -1 / math.sqrt(3), -1 / math.sqrt(2), -1.0], dtype=torch.float64)
    assert G.state_index(cos, 1.0, agg).tolist() == [6, 5, 4, 3, 2, 1, 0]
    assert G.state_index(cos, -1.0, agg).tolist() == [0, 1, 2, 3, 4, 5, 6]


def test_scan_geometry_small_cases():
    """Scan against hand-checked answers on a 1-D line and a 2-D cross."""
    known = np.zeros(10, bool)
    known[[0, 4, 8]] = True
    sp, sn = G.scan_steps(known, np.array([[2], [5], [9]]), [(1,)], 64)
    assert sp.tolist() == [[2, 3, 0]] and sn.tolist() == [[2, 1, 1]]
    sp, sn = G.scan_steps(known, np.array([[2], [5], [9]]), [(1,)], 2)
    assert sp.tolist() == [[2, 0, 0]] and sn.tolist() == [[2, 1, 1]]
    k2 = np.zeros((5, 5), bool)
    k2[0, 0] = k2[4, 4] = k2[2, 0] = True
    dirs = G.half_directions(2)  # (0,1), (1,-1), (1,0), (1,1)
    sp, sn = G.scan_steps(k2, np.array([[2, 2]]), dirs, 64)
    assert dirs == [(0, 1), (1, -1), (1, 0), (1, 1)]
    assert sp[:, 0].tolist() == [0, 0, 0, 2]
    assert sn[:, 0].tolist() == [2, 0, 0, 2]


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_forward_matches_reference_golden(case):
    """The FP64 oracle reproduces the reference's float32 forward over a whole
    schedule (the reference runs in float32: observed differences are ~1e-7 on mu and
    ~1e-6 on log_b, so the tolerances are 1e-6 and 1e-5)."""
    name, shape, levels, stride, d, agg, mr, eb, x, P, q0, ref = case
    geoms, dirs = G.schedule_geoms(shape, levels, stride, 1, mr, agg)
    assert np.array_equal(geoms[0]["q"], q0)
    assert len(geoms) - 1 == len(ref)
    out = G.run_schedule(P, torch.from_numpy(x).double(), shape, levels, stride, eb, mr, agg)
    for i, ((mu, lb), (q, mu_r, lb_r)) in enumerate(zip(out, ref)):
        assert np.array_equal(geoms[i + 1]["q"], q)
        np.testing.assert_allclose(mu.numpy(), mu_r, atol=1e-6, rtol=0, err_msg=f"mu {i}")
        np.testing.assert_allclose(lb.numpy(), lb_r, atol=1e-5, rtol=0, err_msg=f"lb {i}")


def test_forward_sensitive_to_a_planted_mistake():
    """A dropped term (the bidir log-distance columns zeroed) leaves the golden
    tolerance, so the comparison above can fail."""
    name, shape, levels, stride, d, agg, mr, eb, x, P, q0, ref = golden_cases()[0]
    P = dict(P)
    W = P["bidir.net.0.weight"].clone()
    W[:, -2:] = 0
    P["bidir.net.0.weight"] = W
    out = G.run_schedule(P, torch.from_numpy(x).double(), shape, levels, stride, eb, mr, agg)
    worst_mu = max(float(np.abs(mu.numpy() - r[1]).max()) for (mu, _), r in zip(out, ref))
    worst_lb = max(float(np.abs(lb.numpy() - r[2]).max()) for (_, lb), r in zip(out, ref))
    assert worst_mu > 1e-5 and worst_lb > 1e-4


def test_laplace_nll_matches_cdf():
    """Bits = -log2 of the Laplace(mu, b) mass of the cell [t - eb, t + eb]."""
    from scipy.stats import laplace

    rng = np.random.default_rng(0)
    n = 2000
    mu = rng.random(n)
    b = np.exp2(rng.uniform(-12, -3, n))
    eb = 1e-3
    t = mu + rng.laplace(0, 1, n) * b * rng.uniform(0.1, 5, n)
    bits = G.laplace_nll_bits(torch.from_numpy(mu[None]), torch.from_numpy(np.log2(b)[None]),
                              torch.from_numpy(t[None]), eb).numpy()[0]
    mass = laplace.cdf(t + eb, mu, b) - laplace.cdf(t - eb, mu, b)
    ok = mass > 1e-9  # inside the uncompressed range (< ~30 bits)
    np.testing.assert_allclose(bits[ok], -np.log2(mass[ok]), rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("shape,levels", [((13, 11), 2), ((9, 10, 7), 2), ((5, 6, 5, 4), 1)])
def test_frame_scan_equals_whole_field_scan_on_one_chunk(shape, levels):
    """The chunk-frame rule (first pattern-known cell, live only if usable)
    reduces to the whole-field scan when the frame is the tensor plus a halo
    that nothing may use: a line that leaves the tensor never re-enters it."""
    from oracle import oracle as O

    stride = 1 << levels
    geoms, dirs = G.schedule_geoms(shape, levels, stride, 1, stride, 2)
    tile = np.full((2 * stride,) * len(shape), -1)
    for s, (m, _, _, _) in enumerate(O.stage_plan(tile.shape, levels, stride, 1)):
        tile[m] = s
    tile = tile[(slice(0, stride),) * len(shape)]
    fshape = tuple(n + 2 * stride for n in shape)
    idx = np.indices(fshape).reshape(len(shape), -1).T - stride
    inside = np.all((idx >= 0) & (idx < np.asarray(shape)), 1).reshape(fshape)
    gi = 0
    for s, (m, _, _, _) in enumerate(O.stage_plan(shape, levels, stride, 1)):
        if not m.any():
            continue
        known = (tile[tuple((idx % stride).T)] < s).reshape(fshape)
        coords = np.stack(np.nonzero(m), 1)
        sp, sn = G.frame_scan_steps(known, inside, coords + stride, dirs, stride)
        assert np.array_equal(sp, geoms[gi]["sp"]) and np.array_equal(sn, geoms[gi]["sn"])
        gi += 1
    assert gi == len(geoms)
