"""Seeded synthetic workloads of the batched reconstructions.

* ``uneven_slices``: 48 independent 256^2 / 512^2 slices, spoke counts drawn
  uniformly from 64..402 (x 512 samples), coil counts from {4, 8, 12, 16},
  each slice with its own golden-angle offset, phantom, noise and pack.
* ``dynamic_frames``: 24 frames of a 384^2 image on a 768^2 grid, 16 coils,
  one shared golden-angle radial trajectory of 600 spokes x 768 samples,
  ellipses with seeded sinusoidal motion and intensity change, 35 dB SNR.

Measurements come from each case's FP32 plan (the plain path) plus complex
Gaussian noise, in the raw (torchkbnufft) scale the solvers take.  The
reference has no batching (its shapes are constexpr,
experiments/production/src/nufft_integrated_bench.cu:111-113); these are the
build's batched configurations.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import pack, synth
from .recon import build_plan

SLICE_IMAGE, SLICE_GRID, SLICE_READOUT = 256, 512, 512
SLICE_SNR_DB, SLICE_SEED = 40.0, 20260903
FRAME_IMAGE, FRAME_GRID, FRAME_COILS, FRAME_SPOKES, FRAME_READOUT = 384, 768, 16, 600, 768
FRAME_SNR_DB, FRAME_MOTION = 35.0, 0.05


@dataclass
class Workload:
    """Cases (TrajectoryPacks, smaps on the device), raw measurements on the
    device, native operator scales, and the truth images (host)."""

    cases: list = field(default_factory=list)
    measurements: list = field(default_factory=list)
    scales: list = field(default_factory=list)
    truth: list = field(default_factory=list)
    spokes: list = field(default_factory=list)
    coils: list = field(default_factory=list)


def uneven_slices(count: int = 48, seed: int = SLICE_SEED, device: str = "cuda") -> Workload:
    rng = np.random.default_rng(seed)
    spokes = rng.integers(64, 403, count)
    coils = rng.choice([4, 8, 12, 16], count)
    w = Workload(spokes=[int(s) for s in spokes], coils=[int(c) for c in coils])
    for i in range(count):
        traj = synth.radial(int(spokes[i]), SLICE_READOUT, golden=True, offset=float(rng.uniform(0, np.pi)))
        packs = pack.prepare_trajectory(traj, (SLICE_IMAGE, SLICE_IMAGE), (SLICE_GRID, SLICE_GRID))
        maps = torch.from_numpy(synth.coil_maps(int(coils[i]), SLICE_IMAGE)).to(device)
        x = synth.phantom(SLICE_IMAGE, seed=100 + i)
        fp32 = build_plan(packs, maps, "fp32")
        # the plan's forward is normalised by 1/sqrt(K0 K1); measurements are raw
        y = fp32.forward(torch.from_numpy(x).to(device)).cpu().numpy() * SLICE_GRID
        y = torch.from_numpy(synth.add_noise(y, SLICE_SNR_DB, seed=200 + i)).to(device)
        w.cases.append((packs, maps))
        w.measurements.append(y)
        w.scales.append(fp32.operator_scale(y))
        w.truth.append(x)
    return w


def dynamic_frames(frames: int = 24, device: str = "cuda") -> Workload:
    traj = synth.radial(FRAME_SPOKES, FRAME_READOUT, golden=True)
    packs = pack.prepare_trajectory(traj, (FRAME_IMAGE, FRAME_IMAGE), (FRAME_GRID, FRAME_GRID))
    maps = torch.from_numpy(synth.coil_maps(FRAME_COILS, FRAME_IMAGE)).to(device)
    fp32 = build_plan(packs, maps, "fp32")
    w = Workload(spokes=[FRAME_SPOKES] * frames, coils=[FRAME_COILS] * frames)
    for f in range(frames):
        x = synth.phantom(FRAME_IMAGE, FRAME_IMAGE, seed=300, motion=FRAME_MOTION, frame=f, frames=frames)
        y = fp32.forward(torch.from_numpy(x).to(device)).cpu().numpy() * FRAME_GRID
        w.cases.append((packs, maps))
        w.measurements.append(torch.from_numpy(synth.add_noise(y, FRAME_SNR_DB, seed=400 + f)).to(device))
        w.truth.append(x)
    # one operator scale for the series (the first frame's), as a fixed protocol would
    w.scales = [fp32.operator_scale(w.measurements[0])] * frames
    return w


def magnitude_nrmse(a, b) -> float:
    a, b = np.abs(np.asarray(a)), np.abs(np.asarray(b))
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


__all__ = ["Workload", "uneven_slices", "dynamic_frames", "magnitude_nrmse"]
