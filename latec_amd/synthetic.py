"""Seeded synthetic fields for the benchmark configurations.

Conventions (fixed here, model: latec/synthetic_data.py:56-173, 285-287):
  * Gaussian random field with correlation length l: white noise filtered by
    exp(-2 pi^2 f^2 l^2) in Fourier space (a Gaussian kernel of std l),
    anisotropic when l is given per axis;
  * power-law turbulence: random phases, energy spectrum E(k) ~ k^-5/3, i.e.
    amplitude ~ k^-(5/3 + ndim - 1)/2, zero mean;
  * hyperplane jumps: a random unit normal through an interior point, a step
    of amplitude ~ N(0, 1) in standardized units;
  * components are standardized before mixing; the result is min-max scaled
    to [0, 1].
Generation runs in float64: with torch FFTs on a GPU device, with NumPy's
FFT on the host (this torch build's CPU float64 inverse FFT corrupts the
heap, so it is never used).  A field's exact bits depend on the backend, so
parity tests always hand the same array to both sides.
"""

from __future__ import annotations

import numpy as np


def _torch():
    import torch

    return torch


def _freq_grid(shape, device):
    t = _torch()
    axes = [t.fft.fftfreq(n, dtype=t.float64, device=device) for n in shape[:-1]]
    axes.append(t.fft.rfftfreq(shape[-1], dtype=t.float64, device=device))
    return t.meshgrid(*axes, indexing="ij")


def _standardize(x):
    return (x - x.mean()) / x.std().clamp_min(1e-30)


def grf(shape, corr, gen, device="cpu"):
    """Gaussian random field, std-l Gaussian smoothing (l scalar or per axis)."""
    t = _torch()
    corr = np.broadcast_to(np.asarray(corr, np.float64), (len(shape),))
    noise = t.randn(shape, dtype=t.float64, generator=gen, device=gen.device).to(device)
    f = t.fft.rfftn(noise)
    h = t.zeros_like(f.real)
    for fk, lk in zip(_freq_grid(shape, device), corr):
        h = h + fk * fk * float(lk) ** 2
    f = f * t.exp(-2.0 * np.pi**2 * h)
    return _standardize(t.fft.irfftn(f, s=shape))


def turbulence(shape, gen, device="cpu", slope=5.0 / 3.0):
    """Random-phase field with E(k) ~ k^-slope."""
    t = _torch()
    ndim = len(shape)
    k2 = None
    for fk, n in zip(_freq_grid(shape, device), shape):
        k = fk * n
        k2 = k * k if k2 is None else k2 + k * k
    kk = t.sqrt(k2)
    amp = t.where(kk > 0, kk.clamp_min(1e-12) ** (-(slope + ndim - 1) / 2.0), t.zeros_like(kk))
    phase = t.rand(kk.shape, dtype=t.float64, generator=gen, device=gen.device).to(device)
    spec = amp * t.exp(2j * np.pi * phase)
    return _standardize(t.fft.irfftn(spec, s=shape))


def jumps(shape, n_jumps, gen, device="cpu"):
    t = _torch()
    out = t.zeros(shape, dtype=t.float64, device=device)
    coords = t.meshgrid(*[t.arange(n, dtype=t.float64, device=device) for n in shape],
                        indexing="ij")
    for _ in range(n_jumps):
        nrm = t.randn(len(shape), dtype=t.float64, generator=gen, device=gen.device).to(device)
        nrm = nrm / nrm.norm()
        q = t.rand(len(shape), dtype=t.float64, generator=gen, device=gen.device).to(device)
        pt = [(0.25 + 0.5 * float(q[j])) * (shape[j] - 1) for j in range(len(shape))]
        amp = float(t.randn((), dtype=t.float64, generator=gen, device=gen.device))
        side = sum((coords[j] - pt[j]) * float(nrm[j]) for j in range(len(shape)))
        out = out + amp * (side > 0).to(t.float64)
    return out


def minmax(x):
    lo, hi = x.min(), x.max()
    return (x - lo) / (hi - lo if float(hi - lo) > 0 else 1.0)


def _gen(seed, device):
    t = _torch()
    g = t.Generator(device=device)
    g.manual_seed(int(seed))
    return g


# ---- host (NumPy) backend ----------------------------------------------------

def _np_freq(shape):
    axes = [np.fft.fftfreq(n) for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1])]
    return np.meshgrid(*axes, indexing="ij", sparse=True)


def _np_std(x):
    return (x - x.mean()) / max(float(x.std()), 1e-30)


def np_grf(shape, corr, rng):
    corr = np.broadcast_to(np.asarray(corr, np.float64), (len(shape),))
    f = np.fft.rfftn(rng.standard_normal(shape))
    h = sum(fk * fk * float(lk) ** 2 for fk, lk in zip(_np_freq(shape), corr))
    return _np_std(np.fft.irfftn(f * np.exp(-2.0 * np.pi ** 2 * h), s=shape))


def np_turbulence(shape, rng, slope=5.0 / 3.0):
    k2 = sum((fk * n) ** 2 for fk, n in zip(_np_freq(shape), shape))
    kk = np.sqrt(k2)
    with np.errstate(divide="ignore"):
        amp = np.where(kk > 0, np.maximum(kk, 1e-12) ** (-(slope + len(shape) - 1) / 2.0), 0.0)
    spec = amp * np.exp(2j * np.pi * rng.random_sample(kk.shape))
    return _np_std(np.fft.irfftn(spec, s=shape))


def np_jumps(shape, n_jumps, rng):
    out = np.zeros(shape)
    coords = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij",
                         sparse=True)
    for _ in range(n_jumps):
        nrm = rng.standard_normal(len(shape))
        nrm /= np.linalg.norm(nrm)
        pt = (0.25 + 0.5 * rng.random_sample(len(shape))) * (np.asarray(shape) - 1)
        amp = rng.standard_normal()
        side = sum((coords[j] - pt[j]) * nrm[j] for j in range(len(shape)))
        out = out + amp * (side > 0)
    return out


def _np_minmax(x):
    lo, hi = x.min(), x.max()
    return ((x - lo) / (hi - lo if hi > lo else 1.0)).astype(np.float32)


def _host(device) -> bool:
    return str(device) == "cpu"


def config1_field(seed=1, shape=(512, 512), device="cpu"):
    """2-D GRF, correlation length 16 px, min-max scaled to [0, 1] (float32);
    a NumPy array on the host, a tensor on a GPU device."""
    if _host(device):
        return _np_minmax(np_grf(shape, 16.0, np.random.RandomState(seed)))
    return minmax(grf(shape, 16.0, _gen(seed, device), device)).to(_torch().float32)


def config2_field(seed=2, shape=(512, 512, 512), device="cpu"):
    """0.5 GRF(l=24) + 0.5 k^-5/3 turbulence, min-max scaled (float32)."""
    if _host(device):
        rng = np.random.RandomState(seed)
        return _np_minmax(0.5 * np_grf(shape, 24.0, rng) + 0.5 * np_turbulence(shape, rng))
    g = _gen(seed, device)
    x = 0.5 * grf(shape, 24.0, g, device) + 0.5 * turbulence(shape, g, device)
    return minmax(x).to(_torch().float32)


def config3_field(seed=3, shape=(64, 64, 64, 64), device="cpu"):
    """4-D anisotropic GRF (per-axis l ~ U[1, 8]) + 50% turbulence + 3 jumps."""
    if _host(device):
        rng = np.random.RandomState(seed)
        corr = 1.0 + 7.0 * rng.random_sample(len(shape))
        x = 0.5 * np_grf(shape, corr, rng) + 0.5 * np_turbulence(shape, rng)
        return _np_minmax(_np_std(x + np_jumps(shape, 3, rng)))
    g = _gen(seed, device)
    t = _torch()
    corr = (1.0 + 7.0 * t.rand(len(shape), dtype=t.float64, generator=g, device=g.device)).cpu()
    x = 0.5 * grf(shape, corr.numpy(), g, device) + 0.5 * turbulence(shape, g, device)
    x = _standardize(x + jumps(shape, 3, g, device))
    return minmax(x).to(t.float32)


def mixed_field(shape, kind, seed, device="cpu"):
    """'smooth' | 'turbulent' | 'discontinuous' fields for batched workloads."""
    if _host(device):
        rng = np.random.RandomState(seed)
        if kind == "smooth":
            return _np_minmax(np_grf(shape, 8.0, rng))
        if kind == "turbulent":
            return _np_minmax(np_turbulence(shape, rng))
        return _np_minmax(_np_std(np_grf(shape, 4.0, rng) + np_jumps(shape, 3, rng)))
    g = _gen(seed, device)
    if kind == "smooth":
        x = grf(shape, 8.0, g, device)
    elif kind == "turbulent":
        x = turbulence(shape, g, device)
    else:
        x = _standardize(grf(shape, 4.0, g, device) + jumps(shape, 3, g, device))
    return minmax(x).to(_torch().float32)


def training_field(shape, corr_range, gen, device, turbulence_frac=0.5, n_jumps=3):
    """One training sample (the trainer's generator family,
    latec/synthetic_data.py:176-332 as a model): per-axis correlation lengths
    uniform in ``corr_range``, a ``turbulence_frac`` share of k^-5/3
    turbulence, up to ``n_jumps`` hyperplane discontinuities; [0, 1] float32."""
    t = _torch()
    lo, hi = corr_range
    corr = (lo + (hi - lo) * t.rand(len(shape), dtype=t.float64, generator=gen,
                                    device=gen.device)).cpu().numpy()
    x = (1.0 - turbulence_frac) * grf(shape, corr, gen, device)
    if turbulence_frac > 0:
        x = x + turbulence_frac * turbulence(shape, gen, device)
    k = int(t.randint(0, n_jumps + 1, (1,), generator=gen, device=gen.device).item()) if n_jumps else 0
    if k:
        x = _standardize(x) + jumps(shape, k, gen, device)
    return minmax(x).to(t.float32)


def training_batch(batch, shape, corr_range, gen, device, turbulence_frac=0.5, n_jumps=3):
    """(batch, prod(shape)) float32 on ``device``."""
    t = _torch()
    return t.stack([training_field(shape, corr_range, gen, device, turbulence_frac,
                                   n_jumps).reshape(-1) for _ in range(batch)])


def batch_fields(n=256, seed=4, scale=1.0, device="cpu"):
    """The batched uneven workload: ``n`` fields, half 2-D with sides uniform
    in [64, 1024], half 3-D with sides uniform in [32, 256] (each axis drawn
    on its own, times ``scale``), 70 % float32 / 30 % float64, generator
    smooth / turbulent / discontinuous uniformly.  Returns a list of host
    arrays (NumPy)."""
    rng = np.random.RandomState(seed)
    out = []
    for i in range(n):
        if i % 2 == 0:
            shape = tuple(max(8, int(rng.randint(64, 1025) * scale)) for _ in range(2))
        else:
            shape = tuple(max(8, int(rng.randint(32, 257) * scale)) for _ in range(3))
        kind = ("smooth", "turbulent", "discontinuous")[rng.randint(3)]
        f64 = rng.rand() < 0.3
        # generated on the next power-of-two box and cropped: a few FFT sizes
        # serve every field (each new FFT length costs a kernel build)
        box = tuple(1 << (n - 1).bit_length() for n in shape)
        x = mixed_field(box, kind, seed * 100003 + i, device)
        x = x.cpu().numpy() if hasattr(x, "cpu") else np.asarray(x)
        x = np.ascontiguousarray(x[tuple(slice(0, n) for n in shape)], np.float64)
        x = (x - x.min()) / max(x.max() - x.min(), 1e-30)
        out.append(x.astype(np.float64 if f64 else np.float32))
    return out
