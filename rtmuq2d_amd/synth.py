"""Seeded synthetic inputs in the reference's formats.

The reference ships no velocity models or seismograms (its Marmousi2 inputs
are absent from the tree), so every workload is generated here: layered,
anomalous and folded/faulted velocity models, Gaussian smoothing, Gaussian
random fields and the Ricker wavelet.  Models are [Nx][Nz] float32 with z
fastest, the layout of SRC/rtm_iso2d.c:70-71.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import ndimage


def ricker(f0: float, dt: float, n: int) -> np.ndarray:
    """w(t) = (1 - 2 pi^2 f0^2 tau^2) exp(-pi^2 f0^2 tau^2), tau = t - 1/f0."""
    t = np.arange(n, dtype=np.float64) * dt - 1.0 / f0
    a = (math.pi * f0 * t) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)


def smooth(v: np.ndarray, sigma: float) -> np.ndarray:
    return ndimage.gaussian_filter(v.astype(np.float64), sigma, mode="nearest")


def grf(shape, sigma: float, seed: int) -> np.ndarray:
    """Unit-variance Gaussian random field: white noise smoothed with a
    Gaussian of `sigma` grid points, renormalised."""
    g = np.random.default_rng(seed).standard_normal(shape)
    g = ndimage.gaussian_filter(g, sigma, mode="wrap")
    g -= g.mean()
    return g / g.std()


def layered_anomaly(nx: int, nz: int, seed: int,
                    speeds=(1500.0, 2000.0, 2500.0, 3000.0, 3500.0),
                    bump: float = 300.0, sigma: float = 15.0) -> np.ndarray:
    """Flat layers at seeded random depths plus a Gaussian velocity anomaly."""
    rng = np.random.default_rng(seed)
    depths = np.sort(rng.uniform(0.1 * nz, 0.9 * nz, len(speeds) - 1))
    z = np.arange(nz)[None, :].repeat(nx, 0).astype(np.float64)
    v = np.full((nx, nz), speeds[0])
    for d, s in zip(depths, speeds[1:]):
        v[z >= d] = s
    cx, cz = rng.uniform(0.25 * nx, 0.75 * nx), rng.uniform(0.3 * nz, 0.8 * nz)
    x = np.arange(nx)[:, None].astype(np.float64)
    v += bump * np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2 * sigma ** 2))
    return v.astype(np.float32)


def folded_faulted(nx: int, nz: int, h: float, seed: int, n_layers: int = 12,
                   vmin: float = 1500.0, vmax: float = 4700.0, water_m: float = 450.0,
                   fold_amp_m: float = 150.0, fold_len_m: float = 3000.0,
                   throw_m: float = 200.0) -> np.ndarray:
    """Layers between vmin and vmax below a water layer, folded sinusoidally
    and cut by two normal faults."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx, dtype=np.float64)[:, None] * h
    z = np.arange(nz, dtype=np.float64)[None, :] * h
    water = water_m
    zmax = nz * h
    # fold: shift depths by a sinusoid of x
    zf = z - fold_amp_m * np.sin(2 * math.pi * x / fold_len_m + rng.uniform(0, 2 * math.pi))
    # two normal faults (dipping 60 deg); hanging wall dropped by throw_m
    for xf in rng.uniform(0.25, 0.75, 2) * nx * h:
        hanging = x > xf + (z - water) / math.tan(math.radians(60.0))
        zf = np.where(hanging & (z > water), zf - throw_m, zf)
    bounds = np.sort(rng.uniform(water, zmax, n_layers - 1))
    speeds = np.linspace(vmin + 200.0, vmax, n_layers)
    v = np.full((nx, nz), speeds[0])
    for b, s in zip(bounds, speeds[1:]):
        v = np.where(zf >= b, s, v)
    v = np.where(z < water, 1500.0, v)
    return np.broadcast_to(v, (nx, nz)).astype(np.float32).copy()


def migration_velocity(v_true: np.ndarray, seed: int, rel: float = 0.02,
                       sigma_smooth: float = 10.0, sigma_field: float = 20.0) -> np.ndarray:
    """gauss(v_true, 10) * (1 + rel * g), g a unit-variance GRF (config c0)."""
    g = grf(v_true.shape, sigma_field, seed)
    return (smooth(v_true, sigma_smooth) * (1.0 + rel * g)).astype(np.float32)


def ensemble_velocity(v_true: np.ndarray, seed: int, rel: float = 0.03,
                      sigma_smooth: float = 10.0, corr_pts: float = 20.0) -> np.ndarray:
    """gauss(v_true, 10) * exp(rel * g_k) (config c2)."""
    g = grf(v_true.shape, corr_pts, seed)
    return (smooth(v_true, sigma_smooth) * np.exp(rel * g)).astype(np.float32)
