"""Host-side pack directories and KB preparation (native packer bindings).

Mirrors the reference's one-time preparation interface:

* pack directory format ``{G,G_T}/{real,imag}/`` with ``header.bin``,
  ``group_offsets.bin``, ``group_row_ids.bin``, ``tile_col_ids.bin``,
  ``tile_a_comp.bin``, ``tile_pair_meta.bin`` (reference
  ``src/export_torchkbnufft_gpu_case.py:387-436``; reader
  ``experiments/production/src/nufft_integrated_bench.cu:135-226``);
* the production packer ``export_torchkbnufft_gpu_case.py --view {G,G_T}
  --policy morton --optimized-pack --seed 20260823`` (reference
  ``experiments/packing/run_fast_planner_matrix.py:97-124``), here one call
  per view on a ``KbMatrix`` built natively (no torchkbnufft).
"""

from __future__ import annotations

import ctypes as C
import os
import sys
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import PackHeader, check, lib

DEFAULT_SEED = 20260823
VIEWS = {"G": 0, "G_T": 1}
POLICIES = {"morton": 0, "trajectory": 1, "overlap": 2}
HEADER_FIELDS = [name for name, _ in PackHeader._fields_]


VALUE_FORMATS = {"fp16": _lib.VALUES_FP16, "bf16": _lib.VALUES_BF16}


def _header_dict(h: PackHeader) -> dict:
    return {name: int(getattr(h, name)) for name in HEADER_FIELDS}


class HostPack:
    """One packed direction (real + imag components, shared structure) in host memory."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    def __del__(self):
        # At interpreter exit the process is ending and module globals may
        # already be gone: the memory is left to the process teardown.
        h, self._h = getattr(self, "_h", None), None
        try:
            if h and not sys.is_finalizing():
                lib().tsn_host_pack_free(h)
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @classmethod
    def read(cls, real_dir: str | os.PathLike, imag_dir: str | os.PathLike) -> "HostPack":
        out = C.c_void_p()
        check(lib().tsn_host_pack_read(os.fsencode(real_dir), os.fsencode(imag_dir), C.byref(out)))
        return cls(out)

    def write(self, real_dir: str | os.PathLike, imag_dir: str | os.PathLike) -> None:
        check(lib().tsn_host_pack_write(self._h, os.fsencode(real_dir), os.fsencode(imag_dir)))

    @classmethod
    def from_arrays(cls, header: dict, arrays: dict) -> "HostPack":
        h = PackHeader(**{k: int(header[k]) for k in HEADER_FIELDS})
        keep = [np.ascontiguousarray(arrays[k], dtype=d) for k, d in (
            ("group_offsets", np.uint32), ("group_row_ids", np.int32), ("tile_col_ids", np.int32),
            ("real_comp", np.uint16), ("real_meta", np.uint32), ("imag_comp", np.uint16),
            ("imag_meta", np.uint32))]
        out = C.c_void_p()
        check(lib().tsn_host_pack_create(C.byref(h), *[k.ctypes.data for k in keep], C.byref(out)))
        return cls(out)

    @property
    def header(self) -> dict:
        h = PackHeader()
        check(lib().tsn_host_pack_header(self._h, C.byref(h)))
        return _header_dict(h)

    def arrays(self) -> dict:
        """Copies of the pack arrays (tile_a_comp as uint16 FP16 bits)."""

        hd = self.header
        ptrs = [C.c_void_p() for _ in range(7)]
        check(lib().tsn_host_pack_arrays(self._h, *[C.byref(p) for p in ptrs]))
        spec = [("group_offsets", np.uint32, hd["groups"] + 1), ("group_row_ids", np.int32, hd["groups"] * 16),
                ("tile_col_ids", np.int32, hd["tiles"] * 32), ("real_comp", np.uint16, hd["tiles"] * 256),
                ("real_meta", np.uint32, hd["tiles"] * 16), ("imag_comp", np.uint16, hd["tiles"] * 256),
                ("imag_meta", np.uint32, hd["tiles"] * 16)]
        out = {}
        for p, (name, dtype, count) in zip(ptrs, spec):
            buf = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(p.value)
            out[name] = np.frombuffer(buf, dtype=dtype, count=count).copy()
        return out


class KbMatrix:
    """Complex KB interpolation matrix G (float32 components, J*J entries per row)."""

    def __init__(self, handle: C.c_void_p, image_size, grid_size, numpoints):
        self._h = handle
        self.image_size = tuple(image_size)
        self.grid_size = tuple(grid_size)
        self.numpoints = numpoints

    def __del__(self):
        # At interpreter exit the process is ending and module globals may
        # already be gone: the memory is left to the process teardown.
        h, self._h = getattr(self, "_h", None), None
        try:
            if h and not sys.is_finalizing():
                lib().tsn_kb_free(h)
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @classmethod
    def build(cls, trajectory: np.ndarray, image_size=(256, 256), grid_size=(512, 512),
              numpoints: int = 6, threads: int = 0) -> "KbMatrix":
        """trajectory: [M, 2] float32 cycles/pixel in [-0.5, 0.5); column 0 = slow image axis."""

        traj = np.ascontiguousarray(trajectory, dtype=np.float32).reshape(-1, 2)
        out = C.c_void_p()
        check(lib().tsn_kb_build(traj.ctypes.data, traj.shape[0], image_size[0], image_size[1],
                                 grid_size[0], grid_size[1], numpoints, threads, C.byref(out)))
        return cls(out, image_size, grid_size, numpoints)

    def arrays(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        rows, width = C.c_uint32(), C.c_uint32()
        cols, re, im = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().tsn_kb_arrays(self._h, C.byref(rows), C.byref(width), C.byref(cols), C.byref(re), C.byref(im)))
        n = rows.value * width.value
        shape = (rows.value, width.value)

        def grab(ptr, dtype):
            buf = (C.c_char * (n * 4)).from_address(ptr.value)
            return np.frombuffer(buf, dtype=dtype, count=n).reshape(shape).copy()

        return grab(cols, np.int32), grab(re, np.float32), grab(im, np.float32)

    def pack(self, view: str = "G", policy: str = "morton", seed: int = DEFAULT_SEED,
             threads: int = 0, values: str = "fp16") -> HostPack:
        """Pack one view with the production planner.  values "fp16" is the
        reference format; "bf16" keeps BF16-rounded coefficients (same planner)."""

        out = C.c_void_p()
        check(lib().tsn_kb_pack_values(self._h, VIEWS[view], POLICIES[policy], seed, threads, VALUE_FORMATS[values],
                                       C.byref(out)))
        return HostPack(out)


def apodization(image_size=(256, 256), grid_size=(512, 512), numpoints: int = 6) -> np.ndarray:
    """torchkbnufft ``scaling_coef`` (real), float32 [image0, image1]."""

    out = np.zeros(image_size, dtype=np.float32)
    check(lib().tsn_kb_apodization(image_size[0], image_size[1], grid_size[0], grid_size[1], numpoints,
                                   out.ctypes.data))
    return out


@dataclass
class TrajectoryPacks:
    """Everything prepared once per trajectory: G / G_T packs, KB matrix, apodization."""

    kb: KbMatrix
    forward: HostPack
    adjoint: HostPack
    apod: np.ndarray
    values: str = "fp16"
    seed: int = DEFAULT_SEED

    def with_values(self, values: str, threads: int = 0) -> "TrajectoryPacks":
        """The same trajectory packed with another tile value format."""

        if values == self.values:
            return self
        return TrajectoryPacks(self.kb, self.kb.pack("G", "morton", self.seed, threads, values),
                               self.kb.pack("G_T", "morton", self.seed, threads, values), self.apod, values,
                               self.seed)

    def write(self, root: str | os.PathLike) -> None:
        """Write ``root/{G,G_T}/{real,imag}`` in the reference layout."""

        for view, pack in (("G", self.forward), ("G_T", self.adjoint)):
            pack.write(os.path.join(root, view, "real"), os.path.join(root, view, "imag"))


def prepare_trajectory(trajectory: np.ndarray, image_size=(256, 256), grid_size=(512, 512),
                       numpoints: int = 6, seed: int = DEFAULT_SEED, threads: int = 0,
                       values: str = "fp16") -> TrajectoryPacks:
    kb = KbMatrix.build(trajectory, image_size, grid_size, numpoints, threads)
    return TrajectoryPacks(kb, kb.pack("G", "morton", seed, threads, values),
                           kb.pack("G_T", "morton", seed, threads, values),
                           apodization(image_size, grid_size, numpoints), values, seed)


__all__ = ["HostPack", "KbMatrix", "TrajectoryPacks", "apodization", "prepare_trajectory", "DEFAULT_SEED"]
_ = _lib
