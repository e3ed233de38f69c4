"""Reconstruction plan: A, A^H, A^H A and CG-SENSE on one MI355X.

Mirrors the reference's solver interface (``launch_adjoint_image``,
``launch_normal``, ``launch_cg`` and the ``nufft_fp16x2_cg`` CLI,
``experiments/production/src/nufft_fp16x2_cg_bench.cu:203-298``):
``cg(measurement, iterations, operator_scale, lam)`` takes the raw
measurement (torchkbnufft scale), the native operator scale (reference CLI
argument ``OPERATOR_SCALE`` = raw scale / sqrt(K0*K1)) and the raw Tikhonov
weight (1e-4).  A status of 0 (non-positive curvature) or a non-finite image
raises ``TsnError`` with code 3, the reference's exit status 3.
"""

from __future__ import annotations

import ctypes as C
import sys
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import (PREC_DENSE16, PREC_FP32, PREC_HI16, PREC_SPLIT16, PREC_SPLITBF16, TSN_ERR_CG, BatchProblem,
                   PlanConfig, TsnError, check, lib)
from .device import Direction, _ptr, _require, _stream
from .pack import TrajectoryPacks, prepare_trajectory

PRECISIONS = {"split16": PREC_SPLIT16, "dense16": PREC_DENSE16, "fp32": PREC_FP32, "hi16": PREC_HI16,
              "splitbf16": PREC_SPLITBF16}


@dataclass
class CgResult:
    image: torch.Tensor           # complex64 [image0, image1]
    residuals: torch.Tensor       # float64 [iterations + 1], sqrt(rho)
    status: torch.Tensor          # int32 [1], 1 = positive curvature throughout


class Plan:
    """Device-resident operator for one trajectory, coil set and precision."""

    def __init__(self, forward: Direction, adjoint: Direction, smaps: torch.Tensor, apod: torch.Tensor,
                 image_size, grid_size, precision: str = "split16"):
        self.image_size = tuple(image_size)
        self.grid_size = tuple(grid_size)
        self.coils = int(smaps.shape[0])
        self.precision = precision
        self.samples = forward.rows
        _require(smaps, torch.complex64, self.coils * image_size[0] * image_size[1], "smaps")
        _require(apod, torch.float32, image_size[0] * image_size[1], "apodization")
        cfg = PlanConfig(image_size[0], image_size[1], grid_size[0], grid_size[1], self.coils, PRECISIONS[precision])
        out = C.c_void_p()
        check(lib().tsn_plan_create(C.byref(cfg), forward.handle, adjoint.handle, _ptr(smaps), _ptr(apod), C.byref(out)))
        self._h = out
        self._dirs = (forward, adjoint)  # keep the directions alive
        self._keep_call = None  # tensors a captured graph reads and writes

    def __del__(self):
        # At interpreter exit the process is ending and module globals may
        # already be gone: the memory is left to the process teardown.
        h, self._h = getattr(self, "_h", None), None
        try:
            if h and not sys.is_finalizing():
                lib().tsn_plan_free(h)
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        return int(lib().tsn_plan_bytes(self._h))

    # ------------------------------------------------------------ operators
    def _image(self, x: torch.Tensor, name: str) -> None:
        _require(x, torch.complex64, self.image_size[0] * self.image_size[1], name)

    def forward(self, image: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """A x / sqrt(K0*K1): complex64 [image0, image1] -> [coils, samples]."""

        image = image.contiguous()
        self._image(image, "image")
        out = torch.empty((self.coils, self.samples), dtype=torch.complex64, device=image.device) if out is None else out
        _require(out, torch.complex64, self.coils * self.samples, "samples")
        check(lib().tsn_forward(self._h, _ptr(image), _ptr(out), _stream()))
        return out

    def adjoint(self, samples: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        samples = samples.contiguous()
        _require(samples, torch.complex64, self.coils * self.samples, "samples")
        out = torch.empty(self.image_size, dtype=torch.complex64, device=samples.device) if out is None else out
        self._image(out, "out")
        check(lib().tsn_adjoint(self._h, _ptr(samples), _ptr(out), _stream()))
        return out

    def normal(self, image: torch.Tensor, inv_scale_sq: float = 1.0, scaled_lambda: float = 0.0,
               out: torch.Tensor | None = None) -> torch.Tensor:
        image = image.contiguous()
        self._image(image, "image")
        out = torch.empty_like(image) if out is None else out
        self._image(out, "out")
        check(lib().tsn_normal(self._h, _ptr(image), _ptr(out), float(inv_scale_sq), float(scaled_lambda), _stream()))
        return out

    def data_consistency(self, image: torch.Tensor, measurement: torch.Tensor, density: torch.Tensor | None = None,
                         measurement_scale: float | None = None, out: torch.Tensor | None = None) -> torch.Tensor:
        """Gradient of the weighted data term, A^H W (A x - measurement_scale * y).

        A is ``forward``'s operator; y is complex64 [coils, samples]; density
        float32 [samples] or None (W = I).  ``measurement_scale`` defaults to
        1/sqrt(K0*K1), which brings raw torchkbnufft-scale data to A's scale.
        The split-panel precisions form W (A x - y) in the forward
        interpolation's epilogue."""

        image = image.contiguous()
        self._image(image, "image")
        measurement = measurement.contiguous()
        _require(measurement, torch.complex64, self.coils * self.samples, "measurement")
        if density is not None:
            density = density.contiguous()
            _require(density, torch.float32, self.samples, "density")
        if measurement_scale is None:
            measurement_scale = 1.0 / float(np.sqrt(self.grid_size[0] * self.grid_size[1]))
        out = torch.empty_like(image) if out is None else out
        self._image(out, "out")
        check(lib().tsn_data_consistency(self._h, _ptr(image), _ptr(measurement),
                                         _ptr(density) if density is not None else None, float(measurement_scale),
                                         _ptr(out), _stream()))
        return out

    def operator_scale(self, measurement: torch.Tensor) -> float:
        """Truth-free native operator scale sqrt(max |A^H (y / sqrt(K0 K1))|).

        With it the CG right-hand side has unit peak magnitude, which keeps the
        FP16 panels of the split paths in range (the reference calibrates the
        same quantity offline, ``run_cg_quality_matrix.py:15-25, 82-84``)."""

        k = float(np.sqrt(self.grid_size[0] * self.grid_size[1]))
        b = self.adjoint(measurement / k)
        peak = float(torch.max(torch.abs(b)).item())
        if not np.isfinite(peak) or peak <= 0.0:
            raise TsnError(TSN_ERR_CG, f"adjoint of the measurement has peak {peak}")
        return float(np.sqrt(peak))

    # ------------------------------------------------------------ CG
    def _cg_buffers(self, measurement, iterations, out):
        if not measurement.is_contiguous():
            raise ValueError("measurement must be contiguous (the solve keeps reading it)")
        _require(measurement, torch.complex64, self.coils * self.samples, "measurement")
        dev = measurement.device
        if out is None:
            out = CgResult(torch.empty(self.image_size, dtype=torch.complex64, device=dev),
                           torch.empty(iterations + 1, dtype=torch.float64, device=dev),
                           torch.empty(1, dtype=torch.int32, device=dev))
        self._image(out.image, "x")
        _require(out.residuals, torch.float64, iterations + 1, "residuals")
        _require(out.status, torch.int32, 1, "status")
        return out

    def cg(self, measurement: torch.Tensor, iterations: int = 10, operator_scale: float = 1.0, lam: float = 1e-4,
           out: CgResult | None = None) -> CgResult:
        """Enqueue CG-SENSE on the current stream (no host synchronisation)."""

        out = self._cg_buffers(measurement, iterations, out)
        check(lib().tsn_cg(self._h, _ptr(measurement), iterations, float(operator_scale), float(lam), _ptr(out.image),
                           _ptr(out.residuals), _ptr(out.status), _stream()))
        return out

    def capture_cg(self, measurement: torch.Tensor, iterations: int = 10, operator_scale: float = 1.0,
                   lam: float = 1e-4, out: CgResult | None = None) -> CgResult:
        """Capture the whole solve into one hipGraph bound to these buffers."""

        out = self._cg_buffers(measurement, iterations, out)
        stream = torch.cuda.current_stream()
        if stream.cuda_stream == 0:
            raise ValueError("graph capture needs a non-default stream (use torch.cuda.stream(...))")
        check(lib().tsn_cg_graph_capture(self._h, _ptr(measurement), iterations, float(operator_scale), float(lam),
                                         _ptr(out.image), _ptr(out.residuals), _ptr(out.status), _stream()))
        # the graph holds raw device addresses: keep its buffers alive until
        # the next capture, so the caching allocator cannot hand them out
        self._keep_call = (measurement, out)
        return out

    def replay(self) -> None:
        check(lib().tsn_cg_graph_launch(self._h, _stream()))


class Batch:
    """Many CG-SENSE problems of one image / grid shape solved in lockstep
    (one launch per step for all of them); results equal each problem's own
    ``Plan.cg`` bit for bit.  Problems may have their own packs and coil
    counts (slice batches) or share packs (frame series)."""

    def __init__(self, problems, apod: torch.Tensor, image_size, grid_size, precision: str = "split16"):
        """problems: sequence of (forward Direction, adjoint Direction, smaps complex64 [C, N0, N1])."""

        self.image_size = tuple(image_size)
        self.grid_size = tuple(grid_size)
        self.precision = precision
        self.count = len(problems)
        self.tuning = None  # the autotuner's choice (build_batch(autotune=True))
        arr = (BatchProblem * self.count)()
        self._keep = []
        self.coils, self.samples = [], []
        for i, (fwd, adj, smaps) in enumerate(problems):
            smaps = smaps.contiguous()
            c = int(smaps.shape[0])
            _require(smaps, torch.complex64, c * image_size[0] * image_size[1], "smaps")
            arr[i] = BatchProblem(fwd.handle, adj.handle, _ptr(smaps), c)
            self._keep += [fwd, adj, smaps]
            self.coils.append(c)
            self.samples.append(fwd.rows)
        _require(apod, torch.float32, image_size[0] * image_size[1], "apodization")
        cfg = PlanConfig(image_size[0], image_size[1], grid_size[0], grid_size[1], max(self.coils),
                         PRECISIONS[precision])
        out = C.c_void_p()
        check(lib().tsn_batch_create(C.byref(cfg), arr, self.count, _ptr(apod), C.byref(out)))
        self._h = out

    def __del__(self):
        # At interpreter exit the process is ending and module globals may
        # already be gone: the memory is left to the process teardown.
        h, self._h = getattr(self, "_h", None), None
        try:
            if h and not sys.is_finalizing():
                lib().tsn_batch_free(h)
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        return int(lib().tsn_batch_bytes(self._h))

    def _args(self, measurements, iterations, scales, outs):
        if len(measurements) != self.count or len(scales) != self.count:
            raise ValueError("one measurement and one operator scale per problem")
        dev = measurements[0].device
        if outs is None:
            outs = [CgResult(torch.empty(self.image_size, dtype=torch.complex64, device=dev),
                             torch.empty(iterations + 1, dtype=torch.float64, device=dev), None)
                    for _ in range(self.count)]
        status = torch.empty(self.count, dtype=torch.int32, device=dev)
        for i, (m, o) in enumerate(zip(measurements, outs)):
            if not m.is_contiguous():
                raise ValueError("measurements must be contiguous (the solve keeps reading them)")
            _require(m, torch.complex64, self.coils[i] * self.samples[i], f"measurement {i}")
            _require(o.image, torch.complex64, self.image_size[0] * self.image_size[1], f"x {i}")
            _require(o.residuals, torch.float64, iterations + 1, f"residuals {i}")
            o.status = status[i:i + 1]
        ptrs = lambda ts: (C.c_void_p * self.count)(*[t.data_ptr() for t in ts])  # noqa: E731
        return (outs, status, ptrs(measurements), (C.c_float * self.count)(*map(float, scales)),
                ptrs([o.image for o in outs]), ptrs([o.residuals for o in outs]))

    def autotune(self) -> dict:
        """Time the batched interpolation steps over the member group (frames
        sharing a pack per work entry), the pipeline depth and the waves per
        entry of 9..16-coil problems, and keep the fastest (results unchanged
        bit for bit)."""

        group, waves, depth, us = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_float()
        check(lib().tsn_batch_autotune(self._h, _stream(), C.byref(group), C.byref(waves), C.byref(depth),
                                       C.byref(us)))
        return {"member_group": int(group.value), "coil_waves": int(waves.value), "pipe_depth": int(depth.value),
                "us": float(us.value)}

    def cg(self, measurements, iterations: int = 10, operator_scales=None, lam: float = 1e-4, outs=None):
        """Enqueue all solves on the current stream; returns one CgResult per problem."""

        scales = operator_scales if operator_scales is not None else [1.0] * self.count
        outs, status, m, sc, xo, ro = self._args(measurements, iterations, scales, outs)
        check(lib().tsn_batch_cg(self._h, m, iterations, sc, float(lam), xo, ro, _ptr(status), _stream()))
        self._keep_call = (measurements, outs, status)
        return outs

    def capture_cg(self, measurements, iterations: int = 10, operator_scales=None, lam: float = 1e-4, outs=None):
        scales = operator_scales if operator_scales is not None else [1.0] * self.count
        if torch.cuda.current_stream().cuda_stream == 0:
            raise ValueError("graph capture needs a non-default stream (use torch.cuda.stream(...))")
        outs, status, m, sc, xo, ro = self._args(measurements, iterations, scales, outs)
        check(lib().tsn_batch_cg_graph_capture(self._h, m, iterations, sc, float(lam), xo, ro, _ptr(status),
                                               _stream()))
        self._keep_call = (measurements, outs, status)
        return outs

    def replay(self) -> None:
        check(lib().tsn_batch_cg_graph_launch(self._h, _stream()))


def build_batch(cases, precision: str = "split16", shared: bool = False, autotune: bool = False) -> Batch:
    """cases: sequence of (TrajectoryPacks, smaps tensor).  With ``shared`` the
    directions of the first case serve every problem (one packed trajectory
    for a frame series); ``autotune`` tunes each direction's schedule, then
    the batch's member group and coil waves."""

    values = "bf16" if precision == "splitbf16" else "fp16"
    problems, dirs = [], {}
    for packs, smaps in cases:
        key = id(cases[0][0]) if shared else id(packs)
        if key not in dirs:
            p = (cases[0][0] if shared else packs).with_values(values)
            dirs[key] = (Direction.upload(p.forward), Direction.upload(p.adjoint))
            if autotune:
                for d in dirs[key]:
                    d.autotune(int(smaps.shape[0]))
        problems.append((*dirs[key], smaps))
    packs0, smaps0 = cases[0]
    apod = torch.from_numpy(np.ascontiguousarray(packs0.apod)).to(smaps0.device)
    batch = Batch(problems, apod, packs0.kb.image_size, packs0.kb.grid_size, precision)
    if autotune:
        batch.tuning = batch.autotune()
    return batch


def check_cg(result: CgResult) -> None:
    """Host-side failure check after a solve (reference exit status 3)."""

    status = int(result.status.item())
    nonfinite = int((~torch.isfinite(torch.view_as_real(result.image))).sum().item())
    if status != 1 or nonfinite:
        raise TsnError(TSN_ERR_CG, f"CG correctness failure curvature={status} nonfinite={nonfinite}")


def build_plan(packs: TrajectoryPacks, smaps: torch.Tensor, precision: str = "split16",
               autotune: bool = False) -> Plan:
    """Upload the directions a precision needs and create the plan; with
    ``autotune`` each packed direction keeps its fastest work schedule for
    this coil count (results are unchanged bit for bit)."""

    image_size, grid_size = packs.kb.image_size, packs.kb.grid_size
    if precision == "fp32":
        fwd, adj = Direction.csr(packs.kb, "G"), Direction.csr(packs.kb, "G_T")
    else:
        packs = packs.with_values("bf16" if precision == "splitbf16" else "fp16")
        fwd, adj = Direction.upload(packs.forward), Direction.upload(packs.adjoint)
        if autotune:
            for d in (fwd, adj):
                d.autotune(int(smaps.shape[0]))
    apod = torch.from_numpy(np.ascontiguousarray(packs.apod)).to(smaps.device)
    return Plan(fwd, adj, smaps, apod, image_size, grid_size, precision)


__all__ = ["Plan", "Batch", "CgResult", "build_plan", "build_batch", "check_cg", "prepare_trajectory", "PRECISIONS"]
