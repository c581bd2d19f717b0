"""Hand-written backward kernels of the GNN trainer: against finite
differences, torch autograd on float64 references and the FP64 oracle."""

import math

import numpy as np
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-12))


def test_linear_gelu_gather_attn_grads():
    """Each Function's forward/backward against the same ops in float64 torch."""
    from latec_amd.gnn_train import F

    f = F()
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn(300, 37, device="cuda", generator=g)
    W = torch.randn(45, 80, device="cuda", generator=g)[:, 5:42]  # a column slice
    b = torch.randn(45, device="cuda", generator=g)
    dy = torch.randn(300, 45, device="cuda", generator=g)
    xr, Wr, br = (t.detach().double().requires_grad_() for t in (x, W, b))
    x_, W_, b_ = (t.detach().clone().requires_grad_() for t in (x, W, b))
    y = f["Linear"].apply(x_, W_, b_)
    yr = xr @ Wr.T + br
    assert _rel(y.detach().cpu(), yr.detach().cpu()) < 1e-5
    y.backward(dy)
    yr.backward(dy.double())
    for a, r in ((x_, xr), (W_, Wr), (b_, br)):
        assert _rel(a.grad.cpu(), r.grad.cpu()) < 1e-5
    # gelu
    z = torch.randn(1000, device="cuda", generator=g) * 3
    z_, zr = z.clone().requires_grad_(), z.double().requires_grad_()
    gz = f["Gelu"].apply(z_)
    gr = torch.nn.functional.gelu(zr)
    assert _rel(gz.detach().cpu(), gr.detach().cpu()) < 1e-6
    gz.sum().backward()
    gr.sum().backward()
    assert _rel(z_.grad.cpu(), zr.grad.cpu()) < 1e-5
    # gather with repeats (scatter-add backward)
    src = torch.randn(50, 16, device="cuda", generator=g)
    idx = torch.randint(0, 50, (400,), device="cuda", generator=g)
    s_, sr = src.clone().requires_grad_(), src.double().requires_grad_()
    out = f["Gather"].apply(s_, idx)
    outr = sr[idx]
    w = torch.randn(400, 16, device="cuda", generator=g)
    (out * w).sum().backward()
    (outr * w.double()).sum().backward()
    assert _rel(s_.grad.cpu(), sr.grad.cpu()) < 1e-5
    # segmented attention pool with empty segments
    counts = torch.randint(0, 5, (60,), device="cuda", generator=g)
    off = torch.zeros(61, dtype=torch.int64, device="cuda")
    off[1:] = torch.cumsum(counts, 0)
    R, d = int(off[-1]), 32
    s = torch.randn(R, device="cuda", generator=g)
    v = torch.randn(R, d, device="cuda", generator=g)
    nk = torch.tensor(0.3, device="cuda")
    nv = torch.randn(d, device="cuda", generator=g)
    scale = d ** -0.5
    args = [t.clone().requires_grad_() for t in (s, v, nk, nv)]
    out = f["Attn"].apply(args[0], args[1], args[2], args[3], off, scale)
    ref_args = [t.double().requires_grad_() for t in (s, v, nk, nv)]
    outs = []
    for i in range(60):
        a, bnd = int(off[i]), int(off[i + 1])
        sc = torch.cat([ref_args[0][a:bnd], (ref_args[2] * scale).reshape(1)])
        vv = torch.cat([ref_args[1][a:bnd], ref_args[3][None]])
        outs.append((torch.softmax(sc, 0)[:, None] * vv).sum(0))
    outr = torch.stack(outs)
    assert _rel(out.detach().cpu(), outr.detach().cpu()) < 1e-5
    wo = torch.randn(60, d, device="cuda", generator=g)
    (out * wo).sum().backward()
    (outr * wo.double()).sum().backward()
    for a, r in zip(args, ref_args):
        assert _rel(a.grad.cpu(), r.grad.cpu()) < 1e-4


def test_head_loss_kernel_matches_reference_and_fd():
    from latec_amd.gnn_train import F, laplace_bits_reference

    g = torch.Generator(device="cuda").manual_seed(1)
    M = 4000
    o = torch.randn(M, 2, device="cuda", generator=g) * torch.tensor([2.0, 4.0], device="cuda")
    o[:, 1] += 3.0
    t = torch.rand(M, device="cuda", generator=g)
    eb = torch.full((M,), 1e-3, device="cuda")
    o_ = o.clone().requires_grad_()
    bits = F()["HeadLoss"].apply(o_, t, eb, 0.0)
    od = o.double().requires_grad_()
    mu = torch.sigmoid(od[:, 0])
    lb = math.log2(1e-3) + od[:, 1].clamp(-4, 12)
    ref = laplace_bits_reference(mu, lb, t.double(), 1e-3)
    assert _rel(bits.detach().cpu(), ref.detach().cpu()) < 1e-4
    bits.sum().backward()
    ref.sum().backward()
    ok = (od[:, 1] > -4) & (od[:, 1] < 12)
    gk, gr = o_.grad[ok].cpu(), od.grad[ok].cpu()
    assert _rel(gk, gr) < 2e-3
    # central differences of the float64 reference at a few points
    for i in (0, 17, 1234):
        for c in range(2):
            h = 1e-4
            op, om = od.detach().clone(), od.detach().clone()
            op[i, c] += h
            om[i, c] -= h

            def L(z):
                return laplace_bits_reference(torch.sigmoid(z[i:i + 1, 0]),
                                              math.log2(1e-3) + z[i:i + 1, 1].clamp(-4, 12),
                                              t[i:i + 1].double(), 1e-3).sum()
            fd = float((L(op) - L(om)) / (2 * h))
            assert abs(float(o_.grad[i, c]) - fd) <= 2e-3 * max(1.0, abs(fd)), (i, c)


def _small_params(d, agg, seed):
    from latec_amd import gnn_model as GM

    P = GM.init_params(d, agg, seed)
    rng = np.random.default_rng(seed)
    P["head.net.2.weight"] = (rng.standard_normal(P["head.net.2.weight"].shape) * 0.3).astype(np.float32)
    P["head.net.2.bias"][1] = 2.0
    P["line_pool.null_v"] = (rng.standard_normal(P["line_pool.null_v"].shape) * 0.2
                             ).astype(np.float32)
    P["axis_pool.null_v"] = (rng.standard_normal(P["axis_pool.null_v"].shape) * 0.2
                             ).astype(np.float32)
    return P


@pytest.mark.parametrize("shape,levels", [((12, 10), 2), ((6, 5, 6, 4), 1)])
def test_model_gradient_matches_oracle(shape, levels):
    """Loss and every parameter gradient of the trainer (hand-written kernels)
    against torch autograd through the FP64 oracle, same schedule, true values
    revealed."""
    from latec_amd.gnn_train import GnnTrainer
    from oracle import gnn_oracle as G

    d, agg, eb = 32, 2, 2e-3
    P = _small_params(d, agg, 4)
    tr = GnnTrainer(d, agg, params=P)
    rng = np.random.default_rng(0)
    x = np.cumsum(rng.standard_normal(shape), 0)
    x = ((x - x.min()) / (x.max() - x.min())).astype(np.float32).reshape(1, -1)
    bits, n, _ = tr.run_stages(torch.from_numpy(x).cuda(), shape, levels, eb, noise=False)
    bits.backward()
    Pd = {k: torch.from_numpy(np.asarray(v, np.float64)).requires_grad_() for k, v in P.items()}
    stride = 1 << levels
    out = G.run_schedule(Pd, torch.from_numpy(x).double(), shape, levels, stride, eb, stride, agg)
    geoms, _ = G.schedule_geoms(shape, levels, stride, 1, stride, agg)
    xt = torch.from_numpy(x).double()
    ref = sum(G.laplace_nll_bits(mu, lb, xt[:, g["q"]], eb).sum()
              for (mu, lb), g in zip(out, geoms[1:]))
    ref.backward()
    assert n == sum(g["M"] for g in geoms[1:])
    assert abs(float(bits.detach()) - float(ref.detach())) <= 1e-4 * abs(float(ref.detach()))
    for name, p in tr.params.items():
        gk = p.grad.detach().cpu().numpy() if p.grad is not None else np.zeros(p.shape)
        gr = Pd[name].grad.numpy() if Pd[name].grad is not None else np.zeros(p.shape)
        scale = max(np.abs(gr).max(), 1e-3)
        assert np.abs(gk - gr).max() <= 2e-3 * scale, name


def test_model_gradient_finite_differences():
    """Central differences of the trainer's own loss on a few weights."""
    from latec_amd.gnn_train import GnnTrainer

    d, agg, eb, shape, levels = 32, 2, 2e-3, (12, 12), 2
    P = _small_params(d, agg, 5)
    rng = np.random.default_rng(1)
    x = np.cumsum(rng.standard_normal(shape), 1)
    x = ((x - x.min()) / (x.max() - x.min())).astype(np.float32).reshape(1, -1)
    xt = torch.from_numpy(x).cuda()
    tr = GnnTrainer(d, agg, params=P)
    bits, _, _ = tr.run_stages(xt, shape, levels, eb, noise=False)
    bits.backward()
    names = ["bidir.net.0.weight", "dir.net.0.weight", "mix.net.2.weight",
             "line_pool.wk.0.weight", "axis_pool.wv.weight", "head.net.0.weight",
             "init.net.0.weight", "dir_state.table", "line_pool.null_k"]
    for name in names:
        gt = tr.params[name].grad.detach().cpu().numpy()
        ij = np.unravel_index(int(np.argmax(np.abs(gt))), gt.shape)
        g = float(gt[ij])
        h = 3e-2
        vals = []
        for sgn in (1, -1):
            Q = {k: v.copy() for k, v in P.items()}
            Q[name][ij] += sgn * h
            t2 = GnnTrainer(d, agg, params=Q)
            with torch.no_grad():
                b2, _, _ = t2.run_stages(xt, shape, levels, eb, noise=False)
            vals.append(float(b2))
        fd = (vals[0] - vals[1]) / (2 * h)
        # the loss is float32 per point: its rounding puts ~3e-3 of noise on fd
        assert abs(g - fd) <= 0.05 * abs(fd) + 4e-3, (name, g, fd)


def test_training_reduces_loss():
    from latec_amd.gnn_train import GnnTrainer, train

    tr = GnnTrainer(32, 2, seed=0)
    hist = train(tr, 40, lr=3e-3, warmup=5, batch2d=4, shape2d=(64, 64), batch4d=1,
                 shape4d=(8, 8, 8, 8), levels=3, seed=0)
    assert np.mean(hist[-5:]) < np.mean(hist[:5]) - 0.5


def test_bf16_gemm_equals_bf16_rounded_product():
    """latec_train_gemm_bf16 on generic strides (plain, transposed operands,
    split-K, accumulate with bias): equal to the float64 product of the
    bf16-rounded operands up to FP32 accumulation."""
    from latec_amd.gnn_train import _gemm

    g = torch.Generator(device="cuda").manual_seed(3)
    for M, N, K in ((300, 45, 37), (64, 64, 64), (17, 130, 9000), (1000, 8, 256)):
        A = torch.randn(M, K, device="cuda", generator=g)
        Bt = torch.randn(N, K, device="cuda", generator=g)  # B = Bt^T, k fastest
        bias = torch.randn(N, device="cuda", generator=g)
        C = torch.randn(M, N, device="cuda", generator=g)
        C0 = C.clone()
        _gemm(M, N, K, A, K, 1, Bt, 1, K, C, N, bias, accumulate=True, prec="bf16")
        ref = (A.bfloat16().double() @ Bt.bfloat16().double().T + bias.double()
               + C0.double())
        scale = (A.bfloat16().double().abs() @ Bt.bfloat16().double().abs().T).max()
        assert float((C.double() - ref).abs().max() / scale) < 1e-5, (M, N, K)
        # transposed A (m fastest) into a strided output
        At = A.T.contiguous()
        D2 = torch.empty(M, N + 3, device="cuda")
        _gemm(M, N, K, At, 1, M, Bt, 1, K, D2, N + 3, prec="bf16")
        ref = A.bfloat16().double() @ Bt.bfloat16().double().T
        assert float((D2[:, :N].double() - ref).abs().max() / scale) < 1e-5, (M, N, K)


def test_scatter_add_is_deterministic_and_accurate():
    """The fixed-point scatter-add: bit-identical on repeated calls (heavy
    duplication, mixed magnitudes) and within float32 rounding of the float64
    sum."""
    from latec_amd.gnn_train import F

    f = F()
    g = torch.Generator(device="cuda").manual_seed(7)
    src = torch.randn(40, 24, device="cuda", generator=g)
    idx = torch.randint(0, 40, (200000,), device="cuda", generator=g)
    w = torch.randn(200000, 24, device="cuda", generator=g) * torch.exp(
        torch.randn(200000, 1, device="cuda", generator=g) * 3)
    grads = []
    for _ in range(3):
        s_ = src.clone().requires_grad_()
        (f["Gather"].apply(s_, idx) * w).sum().backward()
        grads.append(s_.grad.clone())
    assert all(torch.equal(grads[0], x) for x in grads[1:])
    ref = torch.zeros(40, 24, dtype=torch.float64, device="cuda").index_add_(
        0, idx, w.double())
    assert _rel(grads[0].cpu(), ref.cpu()) < 1e-6


@pytest.mark.timeout(600)
def test_c5_training_step_bf16_closed_loop_is_deterministic():
    """One training step at the C5 configuration (32 fields: 16 x 256^2 and
    16 x 16^4, levels 3, eb 1e-3, closed-loop targets, BF16 GEMMs, FP32 master
    weights, d 64 / agg 2): a finite loss, every parameter updated, and the
    step reproduced bit for bit from the same start."""
    from latec_amd.gnn_train import C5, GnnTrainer, train

    outs = []
    for _ in range(2):
        tr = GnnTrainer(64, 2, seed=1, precision="bf16")
        p0 = {k: v.detach().clone() for k, v in tr.params.items()}
        hist = train(tr, 1, lr=1e-3, warmup=1, seed=5, **C5)
        assert np.isfinite(hist[0]) and 0 < hist[0] < 30
        moved = [k for k, v in tr.params.items() if not torch.equal(v.detach(), p0[k])]
        assert len(moved) == len(p0)
        outs.append((hist[0], {k: v.detach().clone() for k, v in tr.params.items()}))
    assert outs[0][0] == outs[1][0]
    assert all(torch.equal(outs[0][1][k], outs[1][1][k]) for k in outs[0][1])


@pytest.mark.parametrize("shape,levels", [((24, 20), 2), ((8, 8, 8, 8), 1)])
def test_model_gradient_bf16_matches_oracle(shape, levels):
    """BF16 GEMMs: the trainer's loss and every parameter gradient against
    torch autograd through the FP64 oracle (reduced width d 32), within the
    bf16 operand rounding."""
    from latec_amd.gnn_train import GnnTrainer
    from oracle import gnn_oracle as G

    d, agg, eb = 32, 2, 2e-3
    P = _small_params(d, agg, 4)
    tr = GnnTrainer(d, agg, params=P, precision="bf16")
    rng = np.random.default_rng(0)
    x = np.cumsum(rng.standard_normal(shape), 0)
    x = ((x - x.min()) / (x.max() - x.min())).astype(np.float32).reshape(1, -1)
    bits, n, _ = tr.run_stages(torch.from_numpy(x).cuda(), shape, levels, eb, noise=False)
    bits.backward()
    Pd = {k: torch.from_numpy(np.asarray(v, np.float64)).requires_grad_() for k, v in P.items()}
    stride = 1 << levels
    out = G.run_schedule(Pd, torch.from_numpy(x).double(), shape, levels, stride, eb, stride, agg)
    geoms, _ = G.schedule_geoms(shape, levels, stride, 1, stride, agg)
    xt = torch.from_numpy(x).double()
    ref = sum(G.laplace_nll_bits(mu, lb, xt[:, g["q"]], eb).sum()
              for (mu, lb), g in zip(out, geoms[1:]))
    ref.backward()
    assert abs(float(bits.detach()) - float(ref.detach())) <= 2e-2 * abs(float(ref.detach()))
    # bf16 operands (8-bit mantissas) through two softmax pools: each gradient
    # must point the FP64 gradient's way (cosine >= 0.99; a sign, index or
    # transposition mistake falls far below) with a norm within 25 % (bf16
    # scores change the softmax sharpness, hence the size more than the direction)
    # (the FP32 mode's test holds them to 2e-3 elementwise)
    for name, p in tr.params.items():
        gk = (p.grad.detach().cpu().numpy() if p.grad is not None
              else np.zeros(p.shape)).astype(np.float64).ravel()
        gr = (Pd[name].grad.numpy() if Pd[name].grad is not None
              else np.zeros(p.shape)).ravel()
        nr = np.linalg.norm(gr)
        if nr < 1e-6:
            assert np.linalg.norm(gk) < 1e-4, name
            continue
        cos = float(gk @ gr / (np.linalg.norm(gk) * nr))
        assert cos >= 0.99 and np.linalg.norm(gk - gr) <= 0.25 * nr, (name, cos)
