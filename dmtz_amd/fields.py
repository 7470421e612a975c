"""Synthetic scalar fields and SZ-style decompressed fields for the correction workloads.

Fields are float32 numpy arrays indexed [z, y, x] (x fastest in memory, the layout of
main_3d.cu:1945-1947). Generators are seeded (Philox counter-based streams) and computed in
float64 before the final rounding to float32.

* ``gaussian_mixture`` — sum of isotropic Gaussian bumps plus white noise.
* ``fbm`` — Gaussian random field with power spectrum ~ k^-beta, normalised to [0, 1].
* ``quantize`` — the decompressed field g = 2*eps*round(f / (2*eps)), kept inside [f-eps, f+eps].
* ``eps_from_rel`` — the absolute bound and range of main_3d.cu:6301-6309.
"""
import os

import numpy as np


def _rng(seed, stream=0):
    return np.random.Generator(np.random.Philox(key=int(seed) * 1000003 + int(stream)))


def gaussian_mixture(shape, n_bumps=8, seed=0, noise=0.01, sigma=(3.0, 10.0), amp=(-1.0, 1.0)):
    """f = sum_k a_k exp(-|x - c_k|^2 / (2 s_k^2)) + noise * N(0, 1); centres uniform in the grid."""
    D, H, W = shape
    r = _rng(seed, 1)
    c = r.uniform(0.0, 1.0, size=(n_bumps, 3)) * np.array([W, H, D], np.float64)
    s = r.uniform(sigma[0], sigma[1], size=n_bumps)
    a = r.uniform(amp[0], amp[1], size=n_bumps)
    x = np.arange(W, dtype=np.float64)
    y = np.arange(H, dtype=np.float64)
    z = np.arange(D, dtype=np.float64)
    f = np.zeros(shape, np.float64)
    for k in range(n_bumps):
        inv = 1.0 / (2.0 * s[k] * s[k])
        ex = np.exp(-((x - c[k, 0]) ** 2) * inv)
        ey = np.exp(-((y - c[k, 1]) ** 2) * inv)
        ez = a[k] * np.exp(-((z - c[k, 2]) ** 2) * inv)
        f += ez[:, None, None] * ey[None, :, None] * ex[None, None, :]
    if noise:
        f += noise * _rng(seed, 2).standard_normal(shape)
    return f.astype(np.float32)


def fbm(shape, seed=0, beta=3.0, device=None):
    """Gaussian random field with P(k) ~ k^-beta (DC removed), normalised to [0, 1].

    ``device`` selects a torch device for the FFT of large grids (None: numpy on the host)."""
    D, H, W = shape
    if device is None:
        n = _rng(seed, 3).standard_normal(shape)
        try:  # scipy's FFT runs on all host cores; numpy's on one
            import scipy.fft as fft

            kw = dict(workers=os.cpu_count() or 1)
        except ImportError:
            fft, kw = np.fft, {}
        spec = fft.rfftn(n, **kw)
        kz = np.fft.fftfreq(D)[:, None, None]
        ky = np.fft.fftfreq(H)[None, :, None]
        kx = np.fft.rfftfreq(W)[None, None, :]
        k = np.sqrt(kz * kz + ky * ky + kx * kx)
        k[0, 0, 0] = 1.0
        amp = k ** (-beta / 2.0)
        amp[0, 0, 0] = 0.0
        out = fft.irfftn(spec * amp, s=shape, axes=(0, 1, 2), **kw)
        lo, hi = out.min(), out.max()
        return ((out - lo) / (hi - lo)).astype(np.float32)
    import torch

    g = torch.Generator(device="cpu")
    g.manual_seed(int(seed) * 1000003 + 3)
    n = torch.randn(shape, generator=g, dtype=torch.float32).to(device)
    spec = torch.fft.rfftn(n)
    kz = torch.fft.fftfreq(D, device=device)[:, None, None]
    ky = torch.fft.fftfreq(H, device=device)[None, :, None]
    kx = torch.fft.rfftfreq(W, device=device)[None, None, :]
    k = torch.sqrt(kz * kz + ky * ky + kx * kx)
    k[0, 0, 0] = 1.0
    amp = k ** (-beta / 2.0)
    amp[0, 0, 0] = 0.0
    out = torch.fft.irfftn(spec * amp, s=shape)
    lo, hi = out.min(), out.max()
    return ((out - lo) / (hi - lo)).float()


def eps_from_rel(f, rel):
    """(eps, range) with range = max f - min f and eps = rel * range, main_3d.cu:6301-6309."""
    lo = float(np.min(f)) if isinstance(f, np.ndarray) else float(f.min())
    hi = float(np.max(f)) if isinstance(f, np.ndarray) else float(f.max())
    rng = hi - lo
    return rel * rng, rng


def quantize(f, eps):
    """SZ-style decompressed field: g = 2 eps round(f / (2 eps)), clamped into [f - eps, f + eps].

    The clamp works on the float32-rounded values so that |g - f| <= eps holds in double
    arithmetic for every stored value. f may also be a torch tensor (the result is then one on
    the same device)."""
    if not isinstance(f, np.ndarray):
        import torch

        f64 = f.double()
        g = (2.0 * eps * torch.round(f64 / (2.0 * eps))).to(f.dtype)
        gd = g.double()
        g = torch.where(gd - f64 > eps, torch.nextafter(g, torch.full_like(g, -float("inf"))), g)
        g = torch.where(f64 - gd > eps, torch.nextafter(g, torch.full_like(g, float("inf"))), g)
        return g.contiguous()
    f64 = f.astype(np.float64)
    q = 2.0 * eps * np.round(f64 / (2.0 * eps))
    g = q.astype(f.dtype)
    gd = g.astype(np.float64)
    hi = gd - f64 > eps
    lo = f64 - gd > eps
    if hi.any():
        g[hi] = np.nextafter(g[hi], np.array(-np.inf, f.dtype))
    if lo.any():
        g[lo] = np.nextafter(g[lo], np.array(np.inf, f.dtype))
    return g
