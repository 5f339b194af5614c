"""Seeded hydraulic-conductivity fields for the synthetic GWF workloads.

Log-normal K: ln K = ln(K_mean) + sigma * G, with G a standard Gaussian
field -- independent per cell, or correlated with an exponential covariance
exp(-r / L) (r in cell units) synthesised spectrally by circulant embedding
on a padded periodic grid.
"""
from __future__ import annotations

import numpy as np


def iid_lognormal(shape, ln_mean, sigma, rng):
    """ln K ~ N(ln_mean, sigma^2) independently per cell (float64)."""
    return np.exp(ln_mean + sigma * rng.standard_normal(shape))


def _next_fast(n):
    """Smallest 2^a 3^b 5^c >= n (FFT-friendly length)."""
    best = None
    p2 = 1
    while p2 < 2 * n:
        p3 = p2
        while p3 < 2 * n:
            p5 = p3
            while p5 < 2 * n:
                if p5 >= n and (best is None or p5 < best):
                    best = p5
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return best


def gaussian_field_exponential(shape, corr_len, rng, pad=None,
                               dtype=np.float32):
    """Zero-mean, unit-variance Gaussian field with covariance
    exp(-r / corr_len) on a grid of `shape` (cell units).

    The periodic embedding grid is shape + pad (pad = 3 * corr_len by
    default, so the wrap-around correlation is below exp(-3)); the spectrum
    is the FFT of the embedded covariance with negative values clipped.
    """
    if pad is None:
        pad = int(np.ceil(3 * corr_len))
    ext = tuple(_next_fast(s + pad) for s in shape)
    axes = [np.minimum(np.arange(e), e - np.arange(e)).astype(dtype)
            for e in ext]
    r2 = np.zeros(ext, dtype=dtype)
    for ax, a in enumerate(axes):
        sh = [1] * len(ext)
        sh[ax] = -1
        r2 += (a.reshape(sh) ** 2)
    cov = np.exp(-np.sqrt(r2) / dtype(corr_len)).astype(dtype)
    del r2
    spec = np.fft.rfftn(cov).real
    del cov
    np.maximum(spec, 0, out=spec)
    amp = np.sqrt(spec).astype(dtype)
    del spec
    noise = rng.standard_normal(ext, dtype=dtype)
    f = np.fft.rfftn(noise)
    del noise
    f *= amp
    del amp
    g = np.fft.irfftn(f, s=ext)
    del f
    g = g[tuple(slice(0, s) for s in shape)]
    g = (g - g.mean()) / g.std()
    return np.ascontiguousarray(g, dtype=np.float64)


def correlated_lognormal(shape, ln_mean, sigma, corr_len, rng):
    """ln K = ln_mean + sigma * G, G exponential-covariance Gaussian."""
    g = gaussian_field_exponential(shape, corr_len, rng)
    return np.exp(ln_mean + sigma * g)
