"""CPU CG-SENSE restatement used as the timed CPU baseline (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``); ``bench.py`` times it
in its ``cpu_baseline`` leg.

The reference's FP32 reconstruction (``reference_sparse_solver.py:131-197``
with matrix/panel mode fp32, CG order ``:240-332``) as a CPU program: the
float32 complex KB matrix in scipy CSR (forward) and its transpose (adjoint),
``scipy.fft`` in complex64 with the apodization, zero-pad and crop of
``fft_and_scale``/``ifft_and_scale``, SENSE expand/combine, native
normalisation 1/sqrt(K0*K1), float64 CG scalars.

Parity: its operator is the fp32 arithmetic model of ``oracle.solver``
(pinned against the exact NUDFT); ``tests/test_oracle_cpu_recon.py`` checks it
against ``KbOperator`` and the FP64 NUDFT CG.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np
import scipy.fft
import scipy.sparse

from .nudft import cg_sense

_C_LIBRARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtsn_oracle.so")


def c_library() -> ctypes.CDLL:
    """The oracle's C part (``make -C oracle``): the multi-core CSR product."""

    if not os.path.exists(_C_LIBRARY):
        raise RuntimeError(f"{_C_LIBRARY} is missing: build it with `make -C oracle`")
    lib = ctypes.CDLL(_C_LIBRARY)
    fn = lib.tsn_oracle_csr_c64
    fn.restype = None
    fn.argtypes = [ctypes.c_int64] + [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    return lib


class _CsrC64:
    """complex64 CSR for the C product: int64 row pointers, int32 columns."""

    def __init__(self, m: scipy.sparse.csr_matrix, threads: int):
        self.rows = m.shape[0]
        self.row_ptr = np.ascontiguousarray(m.indptr, dtype=np.int64)
        self.cols = np.ascontiguousarray(m.indices, dtype=np.int32)
        self.vals = np.ascontiguousarray(m.data, dtype=np.complex64)
        self.threads = threads
        self.fn = c_library().tsn_oracle_csr_c64

    def __matmul__(self, panel: np.ndarray) -> np.ndarray:
        panel = np.ascontiguousarray(panel, dtype=np.complex64)
        coils = panel.shape[1]
        if coils > 64:
            raise ValueError("the C product takes at most 64 panel columns")
        out = np.empty((self.rows, coils), np.complex64)
        self.fn(self.rows, self.row_ptr.ctypes.data, self.cols.ctypes.data, self.vals.ctypes.data,
                panel.ctypes.data, coils, out.ctypes.data, self.threads)
        return out


class CpuSense:
    """FP32 CG-SENSE on the CPU.  ``backend`` "scipy": scipy CSR products (one
    thread); "omp": the C CSR product on ``workers`` threads.  Both run the
    FFTs through scipy.fft on ``workers`` threads and give the same sums (each
    output row added in entry order)."""

    def __init__(self, cols, re, im, smaps, apod, image_size, grid_size, workers: int = 1,
                 backend: str = "scipy"):
        m, w = cols.shape
        self.g = scipy.sparse.csr_matrix(
            ((re.astype(np.complex64) + 1j * im.astype(np.complex64)).ravel(),
             cols.ravel().astype(np.int64), np.arange(0, m * w + 1, w)),
            shape=(m, grid_size[0] * grid_size[1]))
        self.gh = self.g.conj().T.tocsr()
        if backend == "omp":
            self.g, self.gh = _CsrC64(self.g, workers), _CsrC64(self.gh, workers)
        elif backend != "scipy":
            raise ValueError(f"unknown backend {backend!r}")
        self.smaps = np.asarray(smaps, np.complex64)
        self.apod = np.asarray(apod, np.float32)
        self.image_size = tuple(image_size)
        self.grid_size = tuple(grid_size)
        self.norm = np.float32(1.0 / np.sqrt(grid_size[0] * grid_size[1]))
        self.workers = workers

    def forward(self, x: np.ndarray) -> np.ndarray:
        c = self.smaps.shape[0]
        grid = np.zeros((c,) + self.grid_size, np.complex64)
        grid[:, :self.image_size[0], :self.image_size[1]] = self.smaps * (x * self.apod)[None]
        panel = scipy.fft.fft2(grid, workers=self.workers).reshape(c, -1).T     # [G, C]
        return np.asarray(self.g @ panel).T * self.norm                          # [C, M]

    def adjoint(self, y: np.ndarray) -> np.ndarray:
        c = self.smaps.shape[0]
        grid = np.asarray(self.gh @ np.ascontiguousarray(y.T)).T.reshape((c,) + self.grid_size)
        k = self.grid_size[0] * self.grid_size[1]
        img = scipy.fft.ifft2(grid, workers=self.workers) * np.float32(k)
        img = img[:, :self.image_size[0], :self.image_size[1]] * self.apod[None]
        return (np.sum(np.conj(self.smaps) * img, axis=0) * self.norm).astype(np.complex64)

    def cg(self, measurement: np.ndarray, iterations: int, operator_scale: float, lam: float = 1e-4):
        k = self.grid_size[0] * self.grid_size[1]
        inv = np.float32(1.0 / (operator_scale * operator_scale))
        lam_n = np.float32(lam / k / (operator_scale * operator_scale))
        y = (np.asarray(measurement, np.complex64) * self.norm).astype(np.complex64)
        rhs = self.adjoint(y) * inv
        return cg_sense(lambda p: (self.adjoint(self.forward(p)) * inv + lam_n * p).astype(np.complex64),
                        rhs.astype(np.complex64), iterations)
