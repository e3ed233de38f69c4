"""Shared helpers of the GPU tests (inputs, packs, comparison)."""

from __future__ import annotations

import functools

import numpy as np
import pytest

torch = pytest.importorskip("torch")


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@functools.lru_cache(maxsize=None)
def small_case(image: int = 32, spokes: int = 24, readout: int | None = None, golden: bool = True):
    """(trajectory, TrajectoryPacks) of a small golden-angle radial case."""

    from trajsparsenufft_amd import pack, synth
    readout = readout or 2 * image
    traj = synth.radial(spokes, readout, golden)
    packs = pack.prepare_trajectory(traj, (image, image), (2 * image, 2 * image))
    return traj, packs


def random_panels(n_cols: int, coils: int, seed: int = 0, scale: float = 1.0):
    """FP16 hi/res panels of random float32 data: float16 [blocks, n_cols, 16] each."""

    rng = np.random.default_rng(seed)
    blocks = (coils + 7) // 8
    v = (rng.standard_normal((blocks, n_cols, 16)) * scale).astype(np.float32)
    for b in range(blocks):  # channels of absent coils are zero
        for j in range(8):
            if 8 * b + j >= coils:
                v[b, :, j] = 0
                v[b, :, j + 8] = 0
    hi = v.astype(np.float16)
    res = (v - hi.astype(np.float32)).astype(np.float16)
    return v, hi, res


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()
