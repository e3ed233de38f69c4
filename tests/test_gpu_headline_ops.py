"""Plan operators at the headline grid (256^2 image, 512^2 grid: the radix-8
FFT passes) against the arithmetic model ``oracle.solver.KbOperator`` with the
same roundings; the small-case operator tests run the radix-1/2 passes only."""

import numpy as np
import pytest

from gpu_util import dev, need_gpu, rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def headline_case():
    need_gpu()
    from trajsparsenufft_amd import pack, synth
    traj = synth.radial(48, 512, True)
    packs = pack.prepare_trajectory(traj, (256, 256), (512, 512))
    x = synth.phantom(256, 256, seed=3)
    smaps = synth.coil_maps(8, 256)
    rng = np.random.default_rng(8)
    m = traj.shape[0]
    ys = (rng.standard_normal((8, m)) + 1j * rng.standard_normal((8, m))).astype(np.complex64)
    return packs, x, smaps, ys


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-6), ("split16", 2e-5)])
def test_headline_grid_operators_match_model(headline_case, precision, tol):
    from oracle.solver import KbOperator
    from trajsparsenufft_amd.recon import build_plan
    packs, x, smaps, ys = headline_case
    cols, re, im = packs.kb.arrays()
    modes = ("fp32", "fp32") if precision == "fp32" else ("fp16", "split2")
    model = KbOperator(cols, re, im, smaps, packs.apod, (256, 256), (512, 512), *modes)
    plan = build_plan(packs, dev(smaps), precision)
    assert plan.samples == ys.shape[1]
    f = rel_l2(plan.forward(dev(x)).cpu().numpy(), model.forward(x))
    a = rel_l2(plan.adjoint(dev(ys)).cpu().numpy(), model.adjoint(ys))
    print(f"\n[headline-ops] {precision}: forward {f:.3e} adjoint {a:.3e}")
    assert f <= tol and a <= tol
