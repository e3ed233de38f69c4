"""Batched CG-SENSE: every problem of a batch (uneven slices, or frames
sharing one packed trajectory) reproduces its own single-plan solve bit for
bit, eager and as one captured graph."""

import numpy as np
import pytest

from gpu_util import dev, need_gpu, small_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    need_gpu()


def _problem(image, spokes, coils, seed):
    from oracle import nudft
    from trajsparsenufft_amd import synth
    traj, packs = small_case(image, spokes)
    x = synth.phantom(image, image, seed=seed)
    smaps = synth.coil_maps(coils, image)
    omega = (traj.T * 2.0 * np.pi).astype(np.float32)
    y = synth.add_noise(nudft.nudft_forward(x, smaps, omega), 40.0, seed=seed).astype(np.complex64)
    return packs, smaps, y


def _single(packs, smaps, y, precision, iterations, scale):
    from trajsparsenufft_amd.recon import build_plan
    plan = build_plan(packs, dev(smaps), precision)
    res = plan.cg(dev(y), iterations, scale, 1e-4)
    torch.cuda.synchronize()
    return res, plan


def _bits(t):
    return t.cpu().numpy().view(np.uint32 if t.dtype == torch.complex64 else np.uint64)


@pytest.mark.parametrize("precision", ["split16", "splitbf16", "dense16"])
def test_uneven_slices_match_single_plans(precision):
    from trajsparsenufft_amd.recon import build_batch
    # coil counts of every channel layout: compact rows (4, 8, odd 5), two
    # 8-coil blocks (12, 16, odd 11)
    specs = [(32, 16, 4, 1), (32, 40, 12, 2), (32, 24, 8, 3), (32, 30, 16, 4), (32, 20, 5, 5), (32, 28, 11, 6)]
    cases = [_problem(*s) for s in specs]
    singles, scales = [], []
    for packs, smaps, y in cases:
        from trajsparsenufft_amd.recon import build_plan
        scale = build_plan(packs, dev(smaps), precision).operator_scale(dev(y))
        res, _ = _single(packs, smaps, y, precision, 6, scale)
        singles.append(res)
        scales.append(scale)
    batch = build_batch([(p, dev(s)) for p, s, _ in cases], precision)
    meas = [dev(y) for _, _, y in cases]
    outs = batch.cg(meas, 6, scales, 1e-4)
    torch.cuda.synchronize()
    for i, (o, ref) in enumerate(zip(outs, singles)):
        assert int(o.status.item()) == 1
        assert np.array_equal(_bits(o.image), _bits(ref.image)), f"problem {i}"
        assert np.array_equal(_bits(o.residuals), _bits(ref.residuals)), f"problem {i}"
    # the same batch as one captured graph, replayed twice
    gbatch = build_batch([(p, dev(s)) for p, s, _ in cases], precision)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        gouts = gbatch.capture_cg(meas, 6, scales, 1e-4)
        for _ in range(2):
            for o in gouts:
                o.image.zero_()
            gbatch.replay()
            stream.synchronize()
            for i, (o, ref) in enumerate(zip(gouts, singles)):
                assert np.array_equal(_bits(o.image), _bits(ref.image)), f"graph problem {i}"


def test_frames_sharing_a_pack_and_graph_replay():
    from trajsparsenufft_amd import synth
    from trajsparsenufft_amd.recon import build_batch, build_plan
    from oracle import nudft
    traj, packs = small_case(32, 24)
    smaps = synth.coil_maps(16, 32)
    omega = (traj.T * 2.0 * np.pi).astype(np.float32)
    frames = [synth.phantom(32, 32, seed=7, motion=0.05, frame=f, frames=5) for f in range(5)]
    ys = [synth.add_noise(nudft.nudft_forward(x, smaps, omega), 35.0, seed=f).astype(np.complex64)
          for f, x in enumerate(frames)]
    plan = build_plan(packs, dev(smaps), "split16")
    scale = plan.operator_scale(dev(ys[0]))
    singles = []
    for y in ys:
        singles.append(plan.cg(dev(y), 5, scale, 1e-4))
        torch.cuda.synchronize()
        singles[-1] = type(singles[-1])(singles[-1].image.clone(), singles[-1].residuals.clone(),
                                        singles[-1].status.clone())
    batch = build_batch([(packs, dev(smaps))] * len(ys), "split16", shared=True)
    meas = [dev(y) for y in ys]
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        outs = batch.capture_cg(meas, 5, [scale] * len(ys), 1e-4)
        for rep in range(3):
            for o in outs:
                o.image.zero_()
            batch.replay()
            stream.synchronize()
            for i, (o, ref) in enumerate(zip(outs, singles)):
                assert np.array_equal(_bits(o.image), _bits(ref.image)), f"replay {rep} frame {i}"
                assert np.array_equal(_bits(o.residuals), _bits(ref.residuals))


def test_batch_argument_errors():
    from trajsparsenufft_amd._lib import TsnError
    from trajsparsenufft_amd.recon import build_batch
    packs, smaps, y = _problem(32, 16, 4, 1)
    packs2, smaps2, _ = _problem(48, 16, 4, 1)
    with pytest.raises(TsnError) as err:
        build_batch([(packs, dev(smaps))], "fp32")
    assert err.value.code == 4
    with pytest.raises(TsnError) as err:  # grid sizes differ (maps sized for the batch image)
        from trajsparsenufft_amd.device import Direction
        from trajsparsenufft_amd.recon import Batch
        import torch as _t
        apod = _t.from_numpy(np.ascontiguousarray(packs.apod)).cuda()
        d1 = (Direction.upload(packs.forward), Direction.upload(packs.adjoint))
        d2 = (Direction.upload(packs2.forward), Direction.upload(packs2.adjoint))
        Batch([(*d1, dev(smaps)), (*d2, dev(smaps))], apod, (32, 32), (64, 64), "split16")
    assert err.value.code == 2
    batch = build_batch([(packs, dev(smaps))], "split16")
    with pytest.raises(TsnError) as err:
        batch.cg([dev(y)], 3, [-1.0], 1e-4)
    assert err.value.code == 4
