"""Device directions, interpolation and FFT endpoints (torch tensors as buffers).

Mirrors the reference's operator entry points:

* ``Direction.load(real_dir, imag_dir)``  <- ``load_direction``
  (``experiments/production/src/nufft_integrated_bench.cu:135-226``)
* ``interp(...)``                         <- ``launch_fused_interpolation`` +
  ``launch_fused_interpolation_residual`` (``:431-466``)
* ``split_panel(...)``                    <- ``integrated_complex_to_split2_half_rows`` (``:339-359``)
* ``fft_*`` / ``ifft_adjoint``            <- ``jkf_launch_custom_fft_f30_*`` /
  ``jkf_launch_custom_ifft_f31_adjoint`` (``nufft_custom_fft_f30_api.h:6-34``)

All compute runs in libtsn.so on the current HIP device and the current torch
stream; torch only supplies memory.  Shapes and dtypes are checked on the host
before every launch.
"""

from __future__ import annotations

import ctypes as C
import os
import sys

import torch

from . import _lib
from ._lib import OUT_COIL_MAJOR, OUT_ROW_MAJOR, OUT_SPLIT_IL, PackHeader, check, lib
from .pack import HEADER_FIELDS, HostPack, KbMatrix

LAYOUTS = {"coil_major": OUT_COIL_MAJOR, "row_major": OUT_ROW_MAJOR, "split_il": OUT_SPLIT_IL}


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor | None):
    return None if t is None else C.c_void_p(t.data_ptr())


def _require(t: torch.Tensor, dtype: torch.dtype, numel: int, name: str) -> None:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != dtype or not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous {dtype}, got {t.dtype}")
    if t.numel() != numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")


def coil_blocks(coils: int) -> int:
    return (coils + 7) // 8


class Direction:
    """A packed direction (FP16 2:4 tiles) or an FP32 CSR direction on the device."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        h = PackHeader()
        check(lib().tsn_direction_header(self._h, C.byref(h)))
        self.header = {name: int(getattr(h, name)) for name in HEADER_FIELDS}

    def __del__(self):
        # At interpreter exit the process is ending and module globals may
        # already be gone: the memory is left to the process teardown.
        h, self._h = getattr(self, "_h", None), None
        try:
            if h and not sys.is_finalizing():
                lib().tsn_direction_free(h)
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def rows(self) -> int:
        return self.header["output_size"]

    @property
    def cols(self) -> int:
        return self.header["n_cols"]

    @property
    def device_bytes(self) -> int:
        return int(lib().tsn_direction_bytes(self._h))

    @property
    def unit_tiles(self) -> int:
        """Tile budget of one regular work unit (0 for CSR directions)."""
        return int(lib().tsn_direction_unit_tiles(self._h))

    def autotune(self, coils: int = 8) -> tuple[int, float]:
        """Time the interpolation over the candidate unit sizes, pipeline depths
        and (9..16 coils) waves per entry, and keep the fastest; returns
        (unit_tiles, microseconds per launch)."""

        best, us = C.c_uint32(), C.c_float()
        check(lib().tsn_direction_autotune(self._h, int(coils), _stream(), C.byref(best), C.byref(us)))
        return int(best.value), float(us.value)

    @property
    def coil_waves(self) -> int:
        """Waves per work entry of interleaved-panel launches of 9..16 coils
        (1: one wave for both coil blocks, 2: one per block)."""
        return int(lib().tsn_direction_coil_waves(self._h))

    @coil_waves.setter
    def coil_waves(self, waves: int) -> None:
        check(lib().tsn_direction_set_coil_waves(self._h, int(waves)))

    @property
    def pipe_depth(self) -> int:
        """Tiles in flight ahead of the multiply in interleaved-panel launches
        (0: the kernel's default; 1..3 where the launch offers it)."""
        return int(lib().tsn_direction_pipe_depth(self._h))

    @pipe_depth.setter
    def pipe_depth(self, depth: int) -> None:
        check(lib().tsn_direction_set_pipe_depth(self._h, int(depth)))

    def schedule(self, unit_tiles: int = 0) -> None:
        """Re-cut the work units (runs of whole groups of at most unit_tiles
        tiles; 0 = default).  Results are bitwise independent of the cut."""
        check(lib().tsn_direction_schedule(self._h, int(unit_tiles)))

    @classmethod
    def upload(cls, pack: HostPack) -> "Direction":
        out = C.c_void_p()
        check(lib().tsn_direction_upload(pack.handle, C.byref(out)))
        return cls(out)

    @classmethod
    def load(cls, real_dir: str | os.PathLike, imag_dir: str | os.PathLike) -> "Direction":
        out = C.c_void_p()
        check(lib().tsn_direction_load(os.fsencode(real_dir), os.fsencode(imag_dir), C.byref(out)))
        return cls(out)

    @classmethod
    def csr(cls, kb: KbMatrix, view: str) -> "Direction":
        out = C.c_void_p()
        check(lib().tsn_direction_upload_csr(kb.handle, {"G": 0, "G_T": 1}[view], C.byref(out)))
        return cls(out)


def interp(direction: Direction, panel_hi: torch.Tensor, panel_res: torch.Tensor | None,
           out: torch.Tensor, coils: int = 8, conjugate: bool = False, accumulate: bool = False,
           scale: float = 1.0, layout: str = "coil_major", dense: bool = False) -> torch.Tensor:
    """Packed interpolation of FP16 N16 panels ([blocks, n_cols, 16] float16).

    ``out``: complex64 [coils, rows] (coil_major) or [rows, coils] (row_major).
    ``panel_res=None`` is the hi-only single pass.  ``dense=True`` runs the
    dense-MFMA control on the same tiles."""

    nb = coil_blocks(coils)
    _require(panel_hi, torch.float16, nb * direction.cols * 16, "panel_hi")
    if panel_res is not None:
        _require(panel_res, torch.float16, nb * direction.cols * 16, "panel_res")
    _require(out, torch.complex64, coils * direction.rows, "out")
    fn = lib().tsn_interp_dense if dense else lib().tsn_interp
    check(fn(direction.handle, _ptr(panel_hi), _ptr(panel_res), _ptr(out), coils, int(conjugate),
             int(accumulate), float(scale), LAYOUTS[layout], _stream()))
    return out


def interp_f32(direction: Direction, panel: torch.Tensor, out: torch.Tensor, coils: int = 8,
               conjugate: bool = False, accumulate: bool = False, scale: float = 1.0,
               layout: str = "coil_major") -> torch.Tensor:
    """Plain FP32 CSR interpolation of complex64 [n_cols, coils] rows."""

    _require(panel, torch.complex64, coils * direction.cols, "panel")
    _require(out, torch.complex64, coils * direction.rows, "out")
    check(lib().tsn_interp_f32(direction.handle, _ptr(panel), _ptr(out), coils, int(conjugate), int(accumulate),
                               float(scale), LAYOUTS[layout], _stream()))
    return out


def split_panel(samples: torch.Tensor, coils: int, scale: float = 1.0):
    """complex64 [coils, rows] -> (hi, res) float16 [blocks, rows, 16]."""

    rows = samples.numel() // coils
    _require(samples, torch.complex64, coils * rows, "samples")
    nb = coil_blocks(coils)
    hi = torch.empty((nb, rows, 16), dtype=torch.float16, device=samples.device)
    res = torch.empty_like(hi)
    check(lib().tsn_split_panel(_ptr(samples), rows, coils, float(scale), _ptr(hi), _ptr(res), _stream()))
    return hi, res


def split_panel_il(samples: torch.Tensor, coils: int, scale: float = 1.0, values: str = "fp16") -> torch.Tensor:
    """complex64 [coils, rows] -> interleaved int32 [blocks, rows, 16] (hi | res << 16),
    FP16 halves or (values="bf16") bfloat16 halves."""

    rows = samples.numel() // coils
    _require(samples, torch.complex64, coils * rows, "samples")
    panel = torch.empty((coil_blocks(coils), rows, 16), dtype=torch.int32, device=samples.device)
    fmt = {"fp16": _lib.VALUES_FP16, "bf16": _lib.VALUES_BF16}[values]
    check(lib().tsn_split_panel_il(_ptr(samples), rows, coils, float(scale), fmt, _ptr(panel), _stream()))
    return panel


def interp_il(direction: Direction, panel: torch.Tensor, out: torch.Tensor, coils: int = 8,
              conjugate: bool = False, accumulate: bool = False, scale: float = 1.0,
              layout: str = "coil_major") -> torch.Tensor:
    """Packed interpolation of an interleaved split panel (int32 [blocks, n_cols, 16]).
    ``layout="split_il"`` writes ``out`` as such a panel of the output rows
    (int32 [blocks, rows, 16]); the other layouts write complex64."""

    _require(panel, torch.int32, coil_blocks(coils) * direction.cols * 16, "panel")
    if layout == "split_il":
        _require(out, torch.int32, coil_blocks(coils) * direction.rows * 16, "out")
    else:
        _require(out, torch.complex64, coils * direction.rows, "out")
    check(lib().tsn_interp_il(direction.handle, _ptr(panel), _ptr(out), coils, int(conjugate), int(accumulate),
                              float(scale), LAYOUTS[layout], _stream()))
    return out


DC_MODES = {"materialized": _lib.DC_MATERIALIZED, "fused": _lib.DC_FUSED}


def interp_dc(direction: Direction, panel: torch.Tensor, measurements: torch.Tensor, coils: int = 8,
              density: torch.Tensor | None = None, scale: float = 1.0, measurement_scale: float = 1.0,
              mode: str = "fused", layout: str = "coil_major", workspace: torch.Tensor | None = None,
              out: torch.Tensor | None = None) -> torch.Tensor:
    """Data-consistency residual panel of the forward direction (the adjoint's input):
    split((scale * G panel - measurement_scale * y) * density), int32 [blocks, rows, 16].

    measurements: complex64 [coils, rows] (``layout="coil_major"``, the reference's)
    or [rows, coils] (``"row_major"``); density float32 [rows] or None (ones).
    ``mode="materialized"`` interpolates into ``workspace`` (complex64
    [rows, coils], allocated when None) and forms the residual in a second pass;
    ``"fused"`` forms it in the interpolation epilogue.  Both give the same bits
    (reference ``JKF_DATA_CONSISTENCY_MODE`` 1 and 2)."""

    rows = direction.rows
    _require(panel, torch.int32, coil_blocks(coils) * direction.cols * 16, "panel")
    _require(measurements, torch.complex64, coils * rows, "measurements")
    if density is not None:
        _require(density, torch.float32, rows, "density")
    if mode == "materialized" and workspace is None:
        workspace = torch.empty((rows, coils), dtype=torch.complex64, device=panel.device)
    if workspace is not None:
        _require(workspace, torch.complex64, coils * rows, "workspace")
    if out is None:
        out = torch.empty((coil_blocks(coils), rows, 16), dtype=torch.int32, device=panel.device)
    _require(out, torch.int32, coil_blocks(coils) * rows * 16, "out")
    check(lib().tsn_interp_dc(direction.handle, _ptr(panel), _ptr(measurements), LAYOUTS[layout],
                              _ptr(density) if density is not None else None, coils, float(scale),
                              float(measurement_scale), DC_MODES[mode],
                              _ptr(workspace) if workspace is not None else None, _ptr(out), _stream()))
    return out


# ------------------------------------------------------ reference FFT ABI

_twiddles_ready: set[int] = set()


def _fft_init(device: torch.device) -> None:
    idx = device.index if device.index is not None else torch.cuda.current_device()
    if idx not in _twiddles_ready:
        status = lib().jkf_initialize_custom_fft_twiddles()
        if status != 0:
            raise RuntimeError(f"jkf_initialize_custom_fft_twiddles failed with hipError {status}")
        _twiddles_ready.add(idx)


def fft_forward(image: torch.Tensor, scaling: torch.Tensor, workspace: torch.Tensor,
                smaps: torch.Tensor | None = None, split: bool = True, debug: torch.Tensor | None = None):
    """Reference-shape forward endpoint (256^2 image, 512^2 grid, 8 coils).

    image: complex64 [8,256,256] (or [256,256] with smaps [8,256,256]);
    returns hi [, res] float16 [512*512, 16]."""

    _fft_init(image.device)
    _require(image, torch.complex64, 65536 * (1 if smaps is not None else 8), "image")
    _require(scaling, torch.float32, 65536, "scaling")
    _require(workspace, torch.complex64, 8 * 512 * 512, "workspace")
    hi = torch.empty((512 * 512, 16), dtype=torch.float16, device=image.device)
    res = torch.empty_like(hi) if split else None
    L = lib()
    if smaps is not None:
        _require(smaps, torch.complex64, 8 * 65536, "smaps")
        if split:
            L.jkf_launch_custom_fft_f30_forward_sense_split2(_ptr(image), _ptr(smaps), _ptr(scaling), _ptr(workspace),
                                                             _ptr(hi), _ptr(res), _stream())
        else:
            L.jkf_launch_custom_fft_f30_forward_sense(_ptr(image), _ptr(smaps), _ptr(scaling), _ptr(workspace),
                                                      _ptr(hi), _stream())
    elif debug is not None:
        _require(debug, torch.complex64, 8 * 512 * 512, "debug")
        L.jkf_launch_custom_fft_f30_forward_split2_debug(_ptr(image), _ptr(scaling), _ptr(workspace), _ptr(hi),
                                                         _ptr(res), _ptr(debug), _stream())
    elif split:
        L.jkf_launch_custom_fft_f30_forward_split2(_ptr(image), _ptr(scaling), _ptr(workspace), _ptr(hi), _ptr(res),
                                                   _stream())
    else:
        L.jkf_launch_custom_fft_f30_forward(_ptr(image), _ptr(scaling), _ptr(workspace), _ptr(hi), _stream())
    return (hi, res) if split else hi


def ifft_adjoint(grid: torch.Tensor, workspace: torch.Tensor, scaling: torch.Tensor) -> torch.Tensor:
    """Reference-shape inverse endpoint: crop(IDFT2(grid)) * scaling / 512^2 -> [8,256,256]."""

    _fft_init(grid.device)
    _require(grid, torch.complex64, 8 * 512 * 512, "grid")
    _require(workspace, torch.complex64, 8 * 512 * 512, "workspace")
    _require(scaling, torch.float32, 65536, "scaling")
    out = torch.empty((8, 256, 256), dtype=torch.complex64, device=grid.device)
    lib().jkf_launch_custom_ifft_f31_adjoint(_ptr(grid), _ptr(workspace), _ptr(scaling), None, _ptr(out), _stream())
    return out
