"""Arithmetic models of the interpolation and the full NUFFT-SENSE operator (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``apply_pack``: float64 product of a pack's FP16 tiles (expanded through the
  pinned ``layout.expand_tiles``) with FP16 hi/res panels, complex combine as
  the reference's fused kernel (``nufft_spmma_bench.cu:377-389, 605-638``):
  re = A_re.Re -/+ A_im.Im, im = A_re.Im +/- A_im.Re (conjugate for the
  adjoint), summed over the group's tiles, hi + residual.  This is the
  ``reference.bin`` semantics (``src/export_gpu_case.py:221-230``) extended to
  complex packs.
* ``apply_csr``: float64 product of the float32 KB matrix (plain FP32 path).
* ``KbOperator``: the reference's component arithmetic model
  (``experiments/production/reference_sparse_solver.py:66-197``:
  ``dense_components`` fp32/fp16/split2 panels, ``sparse_components`` fp32/
  fp16 matrix, ``fft_and_scale``/``ifft_and_scale`` with the apodization,
  SENSE expand/combine) in float64, in the native normalisation
  1/sqrt(K0*K1) of ``nufft_fp16x2_cg_bench.cu``.

Parity: ``apply_pack``/``apply_csr`` are products of given arrays in float64
(closed form); ``KbOperator`` in fp32 mode is pinned against the exact NUDFT
(``oracle/nudft.py``) to KB-width-6 accuracy by ``tests/test_oracle_math.py``.
"""

from __future__ import annotations

import numpy as np

from . import layout


def _f16(bits: np.ndarray) -> np.ndarray:
    return np.asarray(bits, dtype=np.uint16).view(np.float16).astype(np.float64)


def expand_pack(arrays: dict, values: str = "fp16") -> tuple[np.ndarray, np.ndarray]:
    """(A_re, A_im) dense float64 [T, 16, 32] from pack arrays (FP16 or BF16 tile values)."""

    t = len(arrays["tile_col_ids"]) // 32

    def dense(comp, meta):
        bits = layout.expand_tiles(np.asarray(comp, np.uint16).reshape(t, 16, 16),
                                   np.asarray(meta, np.uint32).reshape(t, 16))
        vals = layout.bf16_decode(bits) if values == "bf16" else bits.view(np.float16)
        return vals.astype(np.float64)

    return dense(arrays["real_comp"], arrays["real_meta"]), dense(arrays["imag_comp"], arrays["imag_meta"])


def apply_pack(arrays: dict, n_out: int, panel_hi: np.ndarray, panel_res: np.ndarray | None, coils: int,
               conjugate: bool = False, scale: float = 1.0, values: str = "fp16") -> np.ndarray:
    """Packed interpolation in float64 -> complex128 [coils, n_out].

    panels: [blocks, n_cols, 16] real values (channels 0..7 Re, 8..15 Im of
    the block's eight coils); rows absent from the pack stay zero.  values:
    the tiles' format (FP16 reference packs or BF16 packs)."""

    a_re, a_im = expand_pack(arrays, values)
    cols = np.asarray(arrays["tile_col_ids"], np.int64).reshape(-1, 32)
    offsets = np.asarray(arrays["group_offsets"], np.int64)
    rows = np.asarray(arrays["group_row_ids"], np.int64).reshape(-1, 16)
    panel = np.asarray(panel_hi, np.float64)
    if panel_res is not None:
        panel = panel + np.asarray(panel_res, np.float64)
    out = np.zeros((coils, n_out), dtype=np.complex128)
    group_of_tile = np.repeat(np.arange(len(offsets) - 1), np.diff(offsets))
    for b in range(panel.shape[0]):
        gathered = panel[b][cols]                               # [T, 32, 16]
        d_re = np.einsum("trk,tkn->trn", a_re, gathered)        # [T, 16, 16]
        d_im = np.einsum("trk,tkn->trn", a_im, gathered)
        g_re = np.zeros((len(offsets) - 1, 16, 16))
        g_im = np.zeros_like(g_re)
        np.add.at(g_re, group_of_tile, d_re)
        np.add.at(g_im, group_of_tile, d_im)
        for j in range(8):
            c = 8 * b + j
            if c >= coils:
                break
            sign = 1.0 if conjugate else -1.0
            re = g_re[:, :, j] + sign * g_im[:, :, j + 8]
            im = g_re[:, :, j + 8] - sign * g_im[:, :, j]
            valid = rows >= 0
            out[c, rows[valid]] = scale * (re[valid] + 1j * im[valid])
    return out


def apply_csr(cols: np.ndarray, re: np.ndarray, im: np.ndarray, x: np.ndarray) -> np.ndarray:
    """G x in float64: cols/re/im [M, J*J], x complex [n_grid, coils] -> [M, coils]."""

    vals = re.astype(np.float64) + 1j * im.astype(np.float64)
    return np.einsum("mk,mkc->mc", vals, x[cols])


def apply_csr_adjoint(cols: np.ndarray, re: np.ndarray, im: np.ndarray, y: np.ndarray, n_grid: int) -> np.ndarray:
    vals = re.astype(np.float64) + 1j * im.astype(np.float64)
    out = np.zeros((n_grid, y.shape[1]), dtype=np.complex128)
    contrib = np.conj(vals)[:, :, None] * y[:, None, :]      # [M, J*J, C]
    np.add.at(out, cols.ravel(), contrib.reshape(-1, y.shape[1]))
    return out


def components(values: np.ndarray, mode: str) -> list[np.ndarray]:
    """dense_components / sparse_components value rounding (float32 base);
    "bf16" / "bf16split" are the library's BF16 variant (round to nearest even)."""

    v = np.asarray(values, dtype=np.float32)
    if mode == "fp32":
        return [v.astype(np.float64)]
    if mode in ("bf16", "bf16split"):
        hb = layout.bf16_decode(layout.bf16_bits(v))
        if mode == "bf16":
            return [hb.astype(np.float64)]
        lb = layout.bf16_decode(layout.bf16_bits(v - hb))
        return [hb.astype(np.float64), lb.astype(np.float64)]
    hi = v.astype(np.float16).astype(np.float32)
    if mode == "fp16":
        return [hi.astype(np.float64)]
    if mode == "split2":
        lo = (v - hi).astype(np.float16).astype(np.float32)
        return [hi.astype(np.float64), lo.astype(np.float64)]
    raise ValueError(mode)


class KbOperator:
    """A = G F (apod S x) / sqrt(K0 K1) and its adjoint, in a chosen arithmetic model."""

    def __init__(self, cols, re, im, smaps, apod, image_size, grid_size, matrix_mode="fp32", panel_mode="fp32"):
        self.cols = np.asarray(cols, np.int64)
        self.m_re = components(re, matrix_mode)
        self.m_im = components(im, matrix_mode)
        self.smaps = np.asarray(smaps, np.complex128)
        self.apod = np.asarray(apod, np.float64)
        self.image_size = tuple(image_size)
        self.grid_size = tuple(grid_size)
        self.panel_mode = panel_mode
        self.norm = 1.0 / np.sqrt(grid_size[0] * grid_size[1])
        self.coils = self.smaps.shape[0]

    def _product(self, mats_re, mats_im, pr_list, pi_list, fn):
        """sum over matrix components x panel components, reference order."""

        def s(mats, panels):
            total = 0.0
            for m in mats:
                for p in panels:
                    total = total + fn(m, p)
            return total
        return s(mats_re, pr_list), s(mats_im, pi_list), s(mats_re, pi_list), s(mats_im, pr_list)

    def forward(self, image: np.ndarray) -> np.ndarray:
        """complex [N0, N1] -> complex [C, M]."""

        k0, k1 = self.grid_size
        coil = self.smaps * np.asarray(image, np.complex128)[None] * self.apod[None]
        grid = np.zeros((self.coils, k0, k1), dtype=np.complex128)
        grid[:, :self.image_size[0], :self.image_size[1]] = coil
        panel = np.fft.fft2(grid).reshape(self.coils, -1).T          # [G, C]
        pr = components(panel.real, self.panel_mode)
        pi = components(panel.imag, self.panel_mode)
        fn = lambda m, p: np.einsum("mk,mkc->mc", m, p[self.cols])  # noqa: E731
        rr, ii, ri, ir = self._product(self.m_re, self.m_im, pr, pi, fn)
        return ((rr - ii) + 1j * (ri + ir)).T * self.norm

    def adjoint(self, samples: np.ndarray) -> np.ndarray:
        """complex [C, M] -> complex [N0, N1] (exact adjoint in the fp32 model)."""

        k0, k1 = self.grid_size
        y = np.asarray(samples, np.complex128).T                     # [M, C]
        pr = components(y.real, self.panel_mode)
        pi = components(y.imag, self.panel_mode)

        def fn(m, p):
            out = np.zeros((k0 * k1, p.shape[1]))
            np.add.at(out, self.cols.ravel(), (m[:, :, None] * p[:, None, :]).reshape(-1, p.shape[1]))
            return out
        rr, ii, ri, ir = self._product(self.m_re, self.m_im, pr, pi, fn)
        grid = ((rr + ii) + 1j * (ri - ir)).T.reshape(self.coils, k0, k1)
        img = np.fft.ifft2(grid) * (k0 * k1)
        img = img[:, :self.image_size[0], :self.image_size[1]] * self.apod[None]
        return np.sum(np.conj(self.smaps) * img, axis=0) * self.norm

    def normal(self, image: np.ndarray, inv_scale_sq: float = 1.0, lam: float = 0.0) -> np.ndarray:
        return self.adjoint(self.forward(image)) * inv_scale_sq + lam * image


def cg_native(op: KbOperator, measurement: np.ndarray, iterations: int, operator_scale: float, lam: float = 1e-4):
    """CG with the native conventions of ``Plan.cg`` (raw measurement, native scale)."""

    from .nudft import cg_sense
    k = op.grid_size[0] * op.grid_size[1]
    inv = 1.0 / (operator_scale * operator_scale)
    lam_n = lam / k * inv
    y = np.asarray(measurement, np.complex128) / np.sqrt(k)
    rhs = op.adjoint(y) * inv
    return cg_sense(lambda p: op.normal(p, inv, lam_n), rhs, iterations)
