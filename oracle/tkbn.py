"""Kaiser-Bessel interpolation coefficients and apodization (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference does not compute these itself: it calls the third-party
package torchkbnufft 1.5.2 (pinned at reference ``requirements-gpu.txt:5`` and
``experiments/baselines/baselines.lock.json:12-18``), which is absent from
``/root/reference`` and from this image.  Call sites:
``src/export_torchkbnufft_gpu_case.py:537-542`` (``calc_tensor_spmatrix`` with
``numpoints=6``, ``grid_size=2*im_size``) and
``experiments/production/reference_sparse_solver.py:23,157-194``
(``fft_and_scale`` / ``ifft_and_scale`` with ``scaling_coef``).

This module restates torchkbnufft's published algorithm:

* ``build_numpy_spmatrix``: per dimension, with gamma = 2*pi/K, t = omega/gamma
  (float32, as torchkbnufft receives float32 omega), koff = floor(t - J/2),
  distance u_j = t - koff - j for j = 1..J, coefficient
  I0(alpha*sqrt(1-(u/(J/2))^2))/I0(alpha) for |u| < J/2, alpha = 2.34*J,
  times the centring phase exp(i*gamma*(N-1)/2*u); grid index mod(koff+j, K).
  Dimensions combine as an outer product, the product is conjugated and
  multiplied by the fftshift phase exp(i*omega.n_shift), n_shift = N//2.
  Column index = k0*K1 + k1 (dimension 0 is the slow axis).
* ``compute_scaling_coefs`` / ``kaiser_bessel_ft``: apodization
  1/prod_d FT_KB((n_d - (N_d-1)/2)/K_d) with the closed form of the order-0,
  one-dimensional KB transform J*sinh(w)/(w*I0(alpha)),
  w = sqrt(alpha^2 - (pi*J*f)^2).

Parity: coefficient values are "parity unpinned" against torchkbnufft itself
(it cannot be imported here).  The restatement is pinned instead by what the
mathematics fixes: the resulting gridding NUFFT must reproduce the FP64 direct
NUDFT ``X(omega) = sum_n x[n] exp(-i omega.(n - N/2))`` (``oracle/nudft.py``)
to KB-width-6 accuracy, which fails for a wrong phase sign, a wrong centring
phase, a wrong shift, transposed dimensions or a wrong apodization; plus the
exact structural facts (36 entries per row, support 6 x 6, per-row columns
unique).
"""

from __future__ import annotations

import math

import numpy as np

KB_WIDTH_FACTOR = 2.34  # torchkbnufft default kbwidth; alpha = kbwidth * numpoints


def i0_series(x: np.ndarray) -> np.ndarray:
    """Modified Bessel function I0 by its power series sum_k (x^2/4)^k/(k!)^2.

    Converges to double precision for the arguments used here (|x| <= 15);
    the native packer evaluates the same series so both agree bitwise.
    """

    x = np.asarray(x, dtype=np.float64)
    quarter = 0.25 * x * x
    term = np.ones_like(x)
    total = np.ones_like(x)
    for k in range(1, 64):
        term = term * quarter / float(k * k)
        total = total + term
    return total


def _dimension_coefficients(
    omega32: np.ndarray, im_size: int, grid_size: int, numpoints: int, alpha: float
) -> tuple[np.ndarray, np.ndarray]:
    """One dimension of torchkbnufft ``build_numpy_spmatrix`` (``interp_coeff``).

    Returns (grid indices [M, J] int64, complex coefficients [M, J]) before the
    outer product, conjugation and fftshift phase.
    """

    gam = 2.0 * math.pi / grid_size
    t = omega32.astype(np.float32) / np.float32(gam)
    koff = np.floor(t - np.float32(numpoints / 2.0)).astype(np.float32)
    interp_dist = (t - koff).astype(np.float32)
    j = np.arange(1, numpoints + 1, dtype=np.float64)[None, :]
    kern_in = interp_dist.astype(np.float64)[:, None] - j
    half = numpoints / 2.0
    inside = np.abs(kern_in) < half
    arg = np.sqrt(np.clip(1.0 - (kern_in / half) ** 2, 0.0, None))
    coef = np.where(inside, i0_series(alpha * arg) / i0_series(np.float64(alpha)), 0.0)
    phase = np.exp(1j * gam * (im_size - 1) / 2.0 * kern_in)
    index = np.mod(koff.astype(np.float64)[:, None] + j, grid_size).astype(np.int64)
    return index, phase * coef


def kb_interpolation(
    omega: np.ndarray,
    im_size: tuple[int, int],
    grid_size: tuple[int, int],
    numpoints: int = 6,
    kbwidth: float = KB_WIDTH_FACTOR,
) -> tuple[np.ndarray, np.ndarray]:
    """Complex interpolation matrix G (samples x grid) in row-sorted ELL form.

    ``omega``: [2, M] radians in [-pi, pi) (float32 as torchkbnufft sees it).
    Returns ``cols`` [M, J*J] int64 (ascending within each row) and ``vals``
    [M, J*J] complex128 (the values torchkbnufft computes before its float32
    cast; callers round to float32 as the reference matrices are float32).
    """

    omega32 = np.asarray(omega, dtype=np.float32)
    if omega32.ndim != 2 or omega32.shape[0] != 2:
        raise ValueError("omega must have shape [2, M]")
    alpha = kbwidth * numpoints
    idx0, c0 = _dimension_coefficients(omega32[0], im_size[0], grid_size[0], numpoints, alpha)
    idx1, c1 = _dimension_coefficients(omega32[1], im_size[1], grid_size[1], numpoints, alpha)
    m = omega32.shape[1]
    cols = (idx0[:, :, None] * grid_size[1] + idx1[:, None, :]).reshape(m, -1)
    vals = (c0[:, :, None] * c1[:, None, :]).reshape(m, -1)
    n_shift = np.asarray([im_size[0] // 2, im_size[1] // 2], dtype=np.float64)
    shift = omega32.astype(np.float64).T @ n_shift
    vals = np.conj(vals) * np.exp(1j * shift)[:, None]
    order = np.argsort(cols, axis=1, kind="stable")
    cols = np.take_along_axis(cols, order, axis=1)
    vals = np.take_along_axis(vals, order, axis=1)
    return cols, vals


def kb_interpolation_f32(
    omega: np.ndarray, im_size: tuple[int, int], grid_size: tuple[int, int], numpoints: int = 6
) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(cols, real float32, imag float32): the two float32 component matrices."""

    cols, vals = kb_interpolation(omega, im_size, grid_size, numpoints)
    return cols, vals.real.astype(np.float32), vals.imag.astype(np.float32)


def kaiser_bessel_ft(f: np.ndarray, numpoints: int, alpha: float) -> np.ndarray:
    """Fourier transform of the order-0 1-D KB kernel at normalised frequency f."""

    f = np.asarray(f, dtype=np.float64)
    w = np.sqrt(alpha * alpha - (math.pi * numpoints * f) ** 2 + 0j)
    value = numpoints * np.sinh(w) / (w * i0_series(np.float64(alpha)))
    return np.real(value)


def apodization(
    im_size: tuple[int, int], grid_size: tuple[int, int], numpoints: int = 6,
    kbwidth: float = KB_WIDTH_FACTOR,
) -> np.ndarray:
    """torchkbnufft ``scaling_coef`` (real): [N0, N1] float64."""

    alpha = kbwidth * numpoints
    n0 = np.arange(im_size[0]) - (im_size[0] - 1) / 2.0
    n1 = np.arange(im_size[1]) - (im_size[1] - 1) / 2.0
    s0 = 1.0 / kaiser_bessel_ft(n0 / grid_size[0], numpoints, alpha)
    s1 = 1.0 / kaiser_bessel_ft(n1 / grid_size[1], numpoints, alpha)
    return s0[:, None] * s1[None, :]
