"""Seeded synthetic inputs for the benchmark and tests.

* ``radial``: reference ``experiments/data_prep/generate_trajectories.py:14-23``
  (angle step pi*(3-sqrt(5)) for golden angle, pi/spokes otherwise; radius
  ``linspace(-0.5, 0.5, R, endpoint=False)``; float32 [spokes*R, 2] with
  column 0 = cos(angle)*radius = slow image axis).
* ``spiral``: variable-density interleaved spiral, radius ~ s^(1/(1+1.5)) so
  the sample density falls off as r^-1.5, scaled to |k| <= 0.5.
* ``phantom``: sum of ellipses with magnitudes in [0, 1], clipped to [0, 1],
  times a smooth random phase exp(i * quadratic polynomial).
* ``coil_maps``: Gaussian bumps at evenly spaced angles on a circle of radius
  0.35 FOV, sigma 0.3 FOV, smooth linear phase per coil, normalised so
  sum_c |S_c|^2 = 1 per pixel.
* ``add_noise``: complex Gaussian noise for a given SNR in dB relative to
  ||y|| / sqrt(2 * numel) per real component.
"""

from __future__ import annotations

import math

import numpy as np


def radial(spokes: int, readout: int, golden: bool = True, offset: float = 0.0) -> np.ndarray:
    step = np.pi * (3.0 - np.sqrt(5.0)) if golden else np.pi / spokes
    angles = offset + np.arange(spokes, dtype=np.float64) * step
    radius = np.linspace(-0.5, 0.5, readout, endpoint=False, dtype=np.float64)
    x = np.cos(angles)[:, None] * radius[None, :]
    y = np.sin(angles)[:, None] * radius[None, :]
    return np.stack((x, y), axis=-1).reshape(-1, 2).astype(np.float32)


def spiral(interleaves: int = 16, samples: int = 8192, turns: float = 32.0, density_power: float = 1.5) -> np.ndarray:
    """Archimedean-style interleaved spiral with density ~ r^-density_power."""

    s = np.arange(samples, dtype=np.float64) / samples
    r = 0.5 * s ** (1.0 / (1.0 + density_power)) * (1.0 - 1e-6)
    theta = 2.0 * np.pi * turns * s ** (1.0 / (1.0 + density_power))
    out = []
    for i in range(interleaves):
        phi = theta + 2.0 * np.pi * i / interleaves
        out.append(np.stack((r * np.cos(phi), r * np.sin(phi)), axis=-1))
    return np.concatenate(out).astype(np.float32)


def _grid(n0: int, n1: int):
    u = (np.arange(n0) + 0.5) / n0 - 0.5
    v = (np.arange(n1) + 0.5) / n1 - 0.5
    return np.meshgrid(u, v, indexing="ij")


def phantom(n0: int, n1: int | None = None, ellipses: int = 12, seed: int = 0,
            motion: float = 0.0, frame: int = 0, frames: int = 1) -> np.ndarray:
    """complex64 [n0, n1]; ``motion`` > 0 moves ellipse centres sinusoidally over frames."""

    n1 = n0 if n1 is None else n1
    rng = np.random.default_rng(seed)
    u, v = _grid(n0, n1)
    mag = np.zeros((n0, n1))
    t = 2.0 * np.pi * frame / max(frames, 1)
    for _ in range(ellipses):
        cu, cv = rng.uniform(-0.25, 0.25, size=2)
        au, av = rng.uniform(0.05, 0.4, size=2) * 0.5
        ang = rng.uniform(0.0, np.pi)
        amp = rng.uniform(0.0, 1.0)
        ph = rng.uniform(0.0, 2.0 * np.pi)
        if motion:
            cu += motion * np.sin(t + ph)
            cv += motion * np.cos(t + ph)
            amp *= 1.0 + 0.2 * np.sin(t + 2 * ph)
        du, dv = u - cu, v - cv
        ru = (du * np.cos(ang) + dv * np.sin(ang)) / au
        rv = (-du * np.sin(ang) + dv * np.cos(ang)) / av
        mag += amp * (ru * ru + rv * rv <= 1.0)
    mag = np.clip(mag, 0.0, 1.0)
    c = rng.uniform(-np.pi, np.pi, size=6)
    phase = c[0] + c[1] * u + c[2] * v + c[3] * u * u + c[4] * u * v + c[5] * v * v
    return (mag * np.exp(1j * phase)).astype(np.complex64)


def coil_maps(coils: int, n0: int, n1: int | None = None, radius: float = 0.35, sigma: float = 0.3) -> np.ndarray:
    """complex64 [coils, n0, n1] with sum_c |S_c|^2 = 1."""

    n1 = n0 if n1 is None else n1
    u, v = _grid(n0, n1)
    maps = np.zeros((coils, n0, n1), dtype=np.complex128)
    for c in range(coils):
        a = 2.0 * np.pi * c / coils
        cu, cv = radius * np.cos(a), radius * np.sin(a)
        g = np.exp(-((u - cu) ** 2 + (v - cv) ** 2) / (2.0 * sigma * sigma))
        maps[c] = g * np.exp(1j * (a + np.pi * (u * np.cos(a) + v * np.sin(a))))
    norm = np.sqrt(np.sum(np.abs(maps) ** 2, axis=0))
    return (maps / norm).astype(np.complex64)


def add_noise(samples: np.ndarray, snr_db: float, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    sigma = np.linalg.norm(samples) / math.sqrt(2.0 * samples.size) * 10.0 ** (-snr_db / 20.0)
    noise = rng.standard_normal(samples.shape) + 1j * rng.standard_normal(samples.shape)
    return (samples + sigma * noise).astype(np.complex64)


__all__ = ["radial", "spiral", "phantom", "coil_maps", "add_noise"]
