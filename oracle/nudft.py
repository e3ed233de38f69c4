"""FP64 direct NUDFT with SENSE, and CG-SENSE on it (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The forward model is the one torchkbnufft's gridding approximates (reference
call sites ``experiments/production/reference_sparse_solver.py:584-588``,
``KbNufft(im_size, grid_size, numpoints=6)`` with ``smaps``):

    y_c(omega_m) = sum_n S_c[n] x[n] exp(-i omega_m . (n - N/2))

with n = (n0, n1), n0 the slow image axis.  It is evaluated exactly in
float64 as two separable exponential GEMMs per coil.  ``scale`` divides the
operator (the native path works with the unitary grid normalisation
1/sqrt(K0*K1), see ``native_scale``).

CG follows the reference update order of ``cg_sense``
(``experiments/production/reference_sparse_solver.py:240-332``) without the
optional reliable-update / scheduled-reference steps, with float64 scalars.

Parity: the NUDFT is the definition of the forward model; it is pinned by
the closed forms checked in ``tests/test_oracle_math.py`` (single-pixel
images give pure exponentials, adjointness to 1e-12, a sample at omega=0
returns the image sum).
"""

from __future__ import annotations

import numpy as np


def native_scale(grid_size: tuple[int, int]) -> float:
    """Operator divisor of the native path: sqrt(K0*K1) (512 for a 512^2 grid).

    The reference divides measurements and the forward interpolation by 512
    and scales the inverse FFT by 512/512^2
    (``nufft_fp16x2_cg_bench.cu:230-236, 329-352``), i.e. it works with
    A / sqrt(K0*K1) and its exact adjoint.
    """

    return float(np.sqrt(grid_size[0] * grid_size[1]))


def _exponentials(omega: np.ndarray, im_size: tuple[int, int]):
    om = np.asarray(omega, dtype=np.float64)
    n0 = np.arange(im_size[0], dtype=np.float64) - im_size[0] // 2
    n1 = np.arange(im_size[1], dtype=np.float64) - im_size[1] // 2
    e0 = np.exp(-1j * np.outer(om[0], n0))  # [M, N0]
    e1 = np.exp(-1j * np.outer(om[1], n1))  # [M, N1]
    return e0, e1


def nudft_forward(
    image: np.ndarray, smaps: np.ndarray, omega: np.ndarray, scale: float = 1.0,
    chunk: int = 16384,
) -> np.ndarray:
    """[N0,N1] image, [C,N0,N1] maps -> [C, M] samples (complex128)."""

    image = np.asarray(image, dtype=np.complex128)
    smaps = np.asarray(smaps, dtype=np.complex128)
    m = omega.shape[1]
    out = np.zeros((smaps.shape[0], m), dtype=np.complex128)
    for start in range(0, m, chunk):
        stop = min(m, start + chunk)
        e0, e1 = _exponentials(omega[:, start:stop], image.shape)
        for c in range(smaps.shape[0]):
            z = smaps[c] * image
            t = z @ e1.T  # [N0, m]
            out[c, start:stop] = np.einsum("mn,nm->m", e0, t)
    return out / scale


def nudft_adjoint(
    samples: np.ndarray, smaps: np.ndarray, omega: np.ndarray, scale: float = 1.0,
    chunk: int = 16384,
) -> np.ndarray:
    """Exact adjoint of ``nudft_forward``: [C, M] -> [N0, N1]."""

    samples = np.asarray(samples, dtype=np.complex128)
    smaps = np.asarray(smaps, dtype=np.complex128)
    im_size = smaps.shape[1:]
    out = np.zeros(im_size, dtype=np.complex128)
    for start in range(0, samples.shape[1], chunk):
        stop = min(samples.shape[1], start + chunk)
        e0, e1 = _exponentials(omega[:, start:stop], im_size)
        for c in range(smaps.shape[0]):
            weighted = samples[c, start:stop, None] * np.conj(e1)  # [m, N1]
            out += np.conj(smaps[c]) * (np.conj(e0).T @ weighted)
    return out / scale


def cg_sense(normal, rhs: np.ndarray, iterations: int):
    """Reference CG update order (reference_sparse_solver.py:240-332).

    ``normal(p)`` must already include the regulariser.  Returns
    (x, residual history [iterations+1], status dict).  Scalars are float64.
    """

    b = np.asarray(rhs)
    x = np.zeros_like(b)
    status = {"finite": True, "positive_curvature": True, "completed_iterations": 0}
    if not np.all(np.isfinite(b)):
        status.update(finite=False, failure="nonfinite_rhs")
        return x, [], status
    r = b.copy()
    p = r.copy()
    rs_old = float(np.vdot(r.ravel(), r.ravel()).real)
    residuals = [float(np.sqrt(max(rs_old, 0.0)))]
    for iteration in range(iterations):
        ap = normal(p)
        denominator = float(np.vdot(p.ravel(), ap.ravel()).real)
        if not np.isfinite(denominator) or not np.all(np.isfinite(ap)):
            status.update(finite=False, failure=f"nonfinite_iteration_{iteration}")
            break
        if denominator <= 0.0:
            status.update(positive_curvature=False,
                          failure=f"nonpositive_curvature_iteration_{iteration}")
            break
        alpha = rs_old / max(denominator, 1e-30)
        x = x + alpha * p
        r = r - alpha * ap
        rs_new = float(np.vdot(r.ravel(), r.ravel()).real)
        residuals.append(float(np.sqrt(max(rs_new, 0.0))))
        beta = rs_new / max(rs_old, 1e-30)
        p = r + beta * p
        rs_old = rs_new
        status["completed_iterations"] = iteration + 1
    return x, residuals, status


def nudft_cg_sense(
    samples: np.ndarray, smaps: np.ndarray, omega: np.ndarray, grid_size: tuple[int, int],
    iterations: int = 10, lam: float = 1e-4,
):
    """FP64 CG-SENSE on the exact NUDFT: (E^H E + lam) x = E^H y.

    ``samples`` are raw (torchkbnufft-scale) measurements; ``lam`` is the raw
    Tikhonov weight (reference default 1e-4).  The solve runs in the native
    normalisation (E/sqrt(K0*K1)) with lam/(K0*K1), which is the same system.
    """

    s = native_scale(grid_size)
    y = np.asarray(samples, dtype=np.complex128) / s
    lam_n = lam / (s * s)
    rhs = nudft_adjoint(y, smaps, omega, scale=s)

    def normal(p):
        return nudft_adjoint(nudft_forward(p, smaps, omega, scale=s), smaps, omega, scale=s) + lam_n * p

    return cg_sense(normal, rhs, iterations)


def magnitude_nrmse(candidate: np.ndarray, reference: np.ndarray) -> float:
    """||(|c| - |r|)|| / ||r|| (reference ``validate_native_cg.py:158-160``)."""

    ref = np.abs(reference)
    return float(np.linalg.norm(np.abs(candidate) - ref) / max(np.linalg.norm(ref), 1e-30))


def complex_rel_l2(candidate: np.ndarray, reference: np.ndarray) -> float:
    return float(np.linalg.norm(candidate - reference) / max(np.linalg.norm(reference), 1e-30))
