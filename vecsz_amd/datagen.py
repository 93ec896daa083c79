"""Seeded synthetic fields for the workloads this library is measured on.

All fields are float32, C-order numpy arrays whose LAST axis is x (the
fastest-varying dimension the archive lists first).

* ``sinusoids_2d`` -- sum of 8 random-phase plane waves (wavelength U(50, 900)
  px, amplitude U(0, 1), random direction) plus N(0, 1e-3^2) noise; the 2D
  1800x3600 (y, x) field compressed at abs eb 1e-4.
* ``grf`` -- Gaussian random field synthesised by FFT whose shell-integrated
  energy spectrum is E(k) ~ k^-5/3 (so the d-dimensional power spectrum is
  P(k) ~ k^-(5/3 + d - 1)), min/max scaled to [-50, 50]; the 3D 512^3 and
  1024x512x512 fields and the batched mix.
* ``walk_spikes_1d`` -- random walk with N(0, 1e-2^2) steps plus 0.1% seeded
  spikes of magnitude U(10, 1000) and random sign; the 2^28-element 1D stream.

Generation uses numpy on the host by default; ``device="cuda"`` generates the
same kind of field with torch on the GPU (fast for 512 MiB+ fields; the bits
differ from the numpy path, so parity tests always compare on the very array
they compressed).
"""
from __future__ import annotations

import math

import numpy as np


def sinusoids_2d(ny: int = 1800, nx: int = 3600, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    y = np.arange(ny, dtype=np.float64)[:, None]
    x = np.arange(nx, dtype=np.float64)[None, :]
    f = np.zeros((ny, nx), dtype=np.float64)
    for _ in range(8):
        lam = rng.uniform(50.0, 900.0)
        amp = rng.uniform(0.0, 1.0)
        phase = rng.uniform(0.0, 2 * math.pi)
        theta = rng.uniform(0.0, 2 * math.pi)
        kx, ky = math.cos(theta) * 2 * math.pi / lam, math.sin(theta) * 2 * math.pi / lam
        f += amp * np.sin(kx * x + ky * y + phase)
    f += rng.normal(0.0, 1e-3, size=(ny, nx))
    return np.ascontiguousarray(f.astype(np.float32))


def _grf_numpy(shape, seed, lo, hi):
    rng = np.random.default_rng(seed)
    d = len(shape)
    white = rng.standard_normal(size=shape, dtype=np.float32)
    spec = np.fft.rfftn(white)
    ks = [np.fft.fftfreq(n) for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1])]
    k2 = np.zeros(spec.shape, dtype=np.float64)
    for i, kv in enumerate(ks):
        sh = [1] * d
        sh[i] = kv.size
        k2 = k2 + (kv.reshape(sh) ** 2)
    slope = (5.0 / 3.0 + d - 1) / 2.0  # amplitude exponent: sqrt of P(k)
    with np.errstate(divide="ignore"):
        amp = np.where(k2 > 0, k2 ** (-slope / 2.0), 0.0)
    spec *= amp
    f = np.fft.irfftn(spec, s=shape, axes=tuple(range(d)))
    del spec, amp, k2
    fmin, fmax = float(f.min()), float(f.max())
    f = (f - fmin) * ((hi - lo) / (fmax - fmin)) + lo
    return np.ascontiguousarray(f.astype(np.float32))


def _grf_torch(shape, seed, lo, hi):
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    d = len(shape)
    white = torch.randn(shape, generator=g, device="cuda", dtype=torch.float32)
    spec = torch.fft.rfftn(white)
    del white
    ks = [torch.fft.fftfreq(n, device="cuda") for n in shape[:-1]] + [torch.fft.rfftfreq(shape[-1], device="cuda")]
    k2 = torch.zeros(spec.shape, device="cuda", dtype=torch.float32)
    for i, kv in enumerate(ks):
        sh = [1] * d
        sh[i] = kv.numel()
        k2 += kv.reshape(sh) ** 2
    slope = (5.0 / 3.0 + d - 1) / 2.0
    amp = torch.where(k2 > 0, k2.clamp_min(1e-30) ** (-slope / 2.0), torch.zeros_like(k2))
    del k2
    spec *= amp
    del amp
    f = torch.fft.irfftn(spec, s=shape)
    del spec
    fmin, fmax = f.min(), f.max()
    f = (f - fmin) * ((hi - lo) / (fmax - fmin)) + lo
    return f.contiguous()


def grf(shape, seed: int = 0, lo: float = -50.0, hi: float = 50.0, device: str = "cpu", pow2: bool = False):
    """Gaussian random field, E(k) ~ k^-5/3, scaled to [lo, hi].

    ``device="cpu"`` returns a numpy array; ``device="cuda"`` returns a torch
    tensor on the GPU.  ``pow2=True`` synthesises on the next power-of-two grid
    and crops (a window of a GRF is a GRF): FFT sizes then come from a small
    set, which keeps FFT plan creation out of the way for many odd shapes."""
    shape = tuple(int(s) for s in shape)
    if pow2:
        big = tuple(1 << max(0, (s - 1).bit_length()) for s in shape)
        f = grf(big, seed, lo, hi, device)
        f = f[tuple(slice(0, s) for s in shape)]
        fmin, fmax = f.min(), f.max()
        f = (f - fmin) * ((hi - lo) / (fmax - fmin)) + lo
        return f.contiguous() if device != "cpu" else np.ascontiguousarray(f.astype(np.float32))
    if device == "cpu":
        return _grf_numpy(shape, seed, lo, hi)
    return _grf_torch(shape, seed, lo, hi)


def walk_spikes_1d(n: int = 1 << 28, seed: int = 0, sigma: float = 1e-2, spike_frac: float = 1e-3,
                   device: str = "cpu"):
    """Random walk with N(0, sigma^2) steps plus seeded spikes."""
    if device == "cpu":
        rng = np.random.default_rng(seed)
        x = np.cumsum(rng.normal(0.0, sigma, size=n))
        k = int(n * spike_frac)
        pos = rng.choice(n, size=k, replace=False)
        mag = rng.uniform(10.0, 1000.0, size=k) * rng.choice([-1.0, 1.0], size=k)
        x[pos] += mag
        return np.ascontiguousarray(x.astype(np.float32))
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    steps = torch.randn(n, generator=g, device="cuda", dtype=torch.float64) * sigma
    x = torch.cumsum(steps, 0)
    del steps
    k = int(n * spike_frac)
    pos = torch.randperm(n, generator=g, device="cuda")[:k]
    mag = (torch.rand(k, generator=g, device="cuda", dtype=torch.float64) * 990.0 + 10.0)
    sign = torch.where(torch.rand(k, generator=g, device="cuda") < 0.5, -1.0, 1.0).to(torch.float64)
    x[pos] += mag * sign
    return x.to(torch.float32).contiguous()


# Largest fields of the batched mix: 2D 4096 x 4096, 3D 256 x 384 x 384.
_MIX_MAX_2D = 4096 * 4096
_MIX_MAX_3D = 256 * 384 * 384


def _loguniform_mean(lo: float, hi: float) -> float:
    return (hi - lo) / math.log(hi / lo)


def mix_lower_fraction(n_fields: int = 96, target_bytes: float = 8 * 2**30) -> float:
    """f such that element counts drawn log-uniformly on [f * max, max] (half
    the fields 2D, half 3D, maxima as above) have an expected total of
    target_bytes of f32.  The stated ranges (2D from 64^2, 3D from 32^3)
    average about 16 MiB a field, 1.6 GiB for 96; a mix of ~8 GiB needs the
    lower ends raised, and this raises both by the same fraction of their
    maximum."""
    want = target_bytes / 4.0 / n_fields  # mean elements per field
    lo, hi = 1e-9, 1.0 - 1e-12
    for _ in range(200):
        f = 0.5 * (lo + hi)
        mean = 0.5 * (_loguniform_mean(f * _MIX_MAX_2D, _MIX_MAX_2D) + _loguniform_mean(f * _MIX_MAX_3D, _MIX_MAX_3D))
        if mean < want:
            lo = f
        else:
            hi = f
    return 0.5 * (lo + hi)


def batch_mix(n_fields: int = 96, seed: int = 0, target_bytes: float = 8 * 2**30, tol: float = 0.10):
    """Shapes (C order) and r2r error bounds of the ~8 GiB batched mix.

    Each field is 2D or 3D with equal odds; its element count is drawn
    log-uniformly on [f * max, max] with f = mix_lower_fraction(n_fields,
    target_bytes) (about 0.66 for 96 fields and 8 GiB) and split into axes
    with a log-uniform aspect, each axis clamped to the stated per-axis range
    (2D 64..4096, 3D z 32..256, y and x 32..384); the error bound is drawn from
    {1e-2, 1e-3, 1e-4}.  Draws are repeated with successive seeds until the
    total is within `tol` of target_bytes.  Returns ([(shape, eb)], the seed
    that was accepted)."""
    f = mix_lower_fraction(n_fields, target_bytes)
    for attempt in range(1000):
        rng = np.random.default_rng(seed + attempt)
        out = []
        total = 0
        for _ in range(n_fields):
            eb = float(rng.choice([1e-2, 1e-3, 1e-4]))
            if rng.random() < 0.5:
                n = np.exp(rng.uniform(np.log(f * _MIX_MAX_2D), np.log(_MIX_MAX_2D)))
                y = int(np.clip(round(np.sqrt(n) * np.exp(rng.uniform(np.log(0.8), np.log(1.25)))), 64, 4096))
                x = int(np.clip(round(n / y), 64, 4096))
                shape = (y, x)
            else:
                n = np.exp(rng.uniform(np.log(f * _MIX_MAX_3D), np.log(_MIX_MAX_3D)))
                z = int(np.clip(round(np.cbrt(n) * np.exp(rng.uniform(np.log(0.7), np.log(1.0)))), 32, 256))
                y = int(np.clip(round(np.sqrt(n / z) * np.exp(rng.uniform(np.log(0.85), np.log(1.15)))), 32, 384))
                x = int(np.clip(round(n / (z * y)), 32, 384))
                shape = (z, y, x)
            out.append((shape, eb))
            total += 4 * int(np.prod(shape))
        if abs(total - target_bytes) <= tol * target_bytes:
            return out, seed + attempt
    raise RuntimeError("no draw within tolerance")


def batch_mix_shapes(nfields: int = 96, seed: int = 0):
    """Shapes (C order) of the batched mix at its stated ranges (about
    1.6 GiB for 96 fields; batch_mix draws the ~8 GiB one): half 2D, half 3D (seeded);
    element count log-uniform over 64x64 .. 4096x4096 (2D) and
    32^3 .. 256x384x384 (3D), split into axes with log-uniform aspect and
    clamped to those per-axis ranges.  Returns [(shape, eb)] with the r2r
    error bound drawn from {1e-2, 1e-3, 1e-4}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nfields):
        eb = float(rng.choice([1e-2, 1e-3, 1e-4]))
        if rng.random() < 0.5:
            n = np.exp(rng.uniform(np.log(64 * 64), np.log(4096 * 4096)))
            y = int(np.clip(round(np.sqrt(n) * np.exp(rng.uniform(np.log(0.5), np.log(2.0)))), 64, 4096))
            x = int(np.clip(round(n / y), 64, 4096))
            out.append(((y, x), eb))
        else:
            n = np.exp(rng.uniform(np.log(32 ** 3), np.log(256 * 384 * 384)))
            z = int(np.clip(round(np.cbrt(n) * np.exp(rng.uniform(np.log(0.5), np.log(2.0)))), 32, 256))
            y = int(np.clip(round(np.sqrt(n / z) * np.exp(rng.uniform(np.log(0.7), np.log(1.4)))), 32, 384))
            x = int(np.clip(round(n / (z * y)), 32, 384))
            out.append(((z, y, x), eb))
    return out
