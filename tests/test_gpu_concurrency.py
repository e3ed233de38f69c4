"""Regression tests of shared state between blocks, launches and plans.

* The last CG kernel runs many blocks per problem (64 for a 256^2 plan, 256
  per problem in a batch); every block reads rho while the residual history
  is formed, so the history must never be rewritten in place.  Batches of
  256^2 slices must equal their per-slice plans bit for bit, image and
  history, eager and as a replayed graph.
* Plans sharing one direction hold their own long-group chunk slots and
  arrival counters: two plans on two streams give the results they give
  alone, bit for bit, launch after launch.
* A captured plan graph keeps the tensors it reads and writes alive.
* Dense-control tiles are expanded on the device on first use only.
"""

import gc

import numpy as np
import pytest

from gpu_util import dev, need_gpu

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    need_gpu()


def _bits(t):
    return t.cpu().numpy().view(np.uint32 if t.dtype == torch.complex64 else np.uint64)


def _clone(res):
    from trajsparsenufft_amd.recon import CgResult
    return CgResult(res.image.clone(), res.residuals.clone(), res.status.clone())


def _slice(image, spokes, coils, seed):
    """(packs, smaps, y): golden-angle radial slice, data from the FP32 plan's forward model + noise."""

    from trajsparsenufft_amd import pack, synth
    from trajsparsenufft_amd.recon import build_plan
    traj = synth.radial(spokes, 2 * image, True, offset=0.37 * seed)
    packs = pack.prepare_trajectory(traj, (image, image), (2 * image, 2 * image))
    x = synth.phantom(image, image, seed=seed)
    smaps = synth.coil_maps(coils, image)
    y = build_plan(packs, dev(smaps), "fp32").forward(dev(x)) * float(2 * image)
    torch.cuda.synchronize()
    y = synth.add_noise(y.cpu().numpy(), 40.0, seed=seed).astype(np.complex64)
    return packs, smaps, y


def _long_groups(host_pack) -> int:
    off = host_pack.arrays()["group_offsets"].astype(np.int64)
    return int(np.count_nonzero(np.diff(off) > 16))


def test_batch_of_256_slices_equals_per_slice_plans():
    from trajsparsenufft_amd.recon import build_batch, build_plan
    specs = [(256, 64, 4, 1), (256, 96, 8, 2), (256, 80, 12, 3), (256, 72, 16, 4)]
    cases = [_slice(*s) for s in specs]
    singles, scales = [], []
    for i, (packs, smaps, y) in enumerate(cases):
        print(f"[final-kernel] slice {i} spokes={specs[i][1]} coils={specs[i][2]}", flush=True)
        plan = build_plan(packs, dev(smaps), "split16")
        scale = plan.operator_scale(dev(y))
        singles.append(_clone(plan.cg(dev(y), 6, scale, 1e-4)))
        torch.cuda.synchronize()
        assert int(singles[-1].status.item()) == 1
        scales.append(scale)
    meas = [dev(y) for _, _, y in cases]
    batch = build_batch([(p, dev(s)) for p, s, _ in cases], "split16")
    outs = batch.cg(meas, 6, scales, 1e-4)
    torch.cuda.synchronize()
    for i, (o, ref) in enumerate(zip(outs, singles)):
        assert int(o.status.item()) == 1
        assert np.array_equal(_bits(o.image), _bits(ref.image)), f"slice {i} image"
        assert np.array_equal(_bits(o.residuals), _bits(ref.residuals)), f"slice {i} history"
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        gouts = batch.capture_cg(meas, 6, scales, 1e-4)
        for rep in range(3):
            for o in gouts:
                o.image.zero_()
                o.residuals.zero_()
            batch.replay()
            stream.synchronize()
            for i, (o, ref) in enumerate(zip(gouts, singles)):
                assert np.array_equal(_bits(o.image), _bits(ref.image)), f"replay {rep} slice {i}"
                assert np.array_equal(_bits(o.residuals), _bits(ref.residuals)), f"replay {rep} slice {i}"


def test_plan_history_is_sqrt_of_rho_at_any_length():
    """Solves of 1..4 iterations are prefixes of one another: the history of
    an n-step solve equals the first n + 1 entries of a longer one."""

    from trajsparsenufft_amd.recon import build_plan
    packs, smaps, y = _slice(128, 64, 8, 5)
    plan = build_plan(packs, dev(smaps), "split16")
    scale = plan.operator_scale(dev(y))
    long = _clone(plan.cg(dev(y), 4, scale, 1e-4))
    torch.cuda.synchronize()
    hist = long.residuals.cpu().numpy()
    assert np.all(np.diff(hist) < 0), hist
    for n in (1, 2, 3):
        short = plan.cg(dev(y), n, scale, 1e-4)
        torch.cuda.synchronize()
        assert np.array_equal(_bits(short.residuals), _bits(long.residuals)[: n + 1]), n


def test_plans_sharing_directions_on_two_streams():
    from trajsparsenufft_amd import synth
    from trajsparsenufft_amd.device import Direction
    from trajsparsenufft_amd.recon import Plan
    packs, smaps, y = _slice(256, 160, 8, 6)
    assert _long_groups(packs.adjoint) > 0, "the case needs long groups (chunked sums)"
    fwd, adj = Direction.upload(packs.forward), Direction.upload(packs.adjoint)
    apod = dev(np.ascontiguousarray(packs.apod))
    smaps2 = synth.coil_maps(8, 256, radius=0.3, sigma=0.25)
    plans = [Plan(fwd, adj, dev(smaps), apod, (256, 256), (512, 512)),
             Plan(fwd, adj, dev(np.ascontiguousarray(smaps2)), apod, (256, 256), (512, 512))]
    meas = [dev(y), dev(y[::-1].copy())]
    scales = [p.operator_scale(m) for p, m in zip(plans, meas)]
    alone = []
    for p, m, s in zip(plans, meas, scales):
        alone.append(_clone(p.cg(m, 5, s, 1e-4)))
        torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for rep in range(4):
        outs = []
        for p, m, s, st in zip(plans, meas, scales, streams):
            with torch.cuda.stream(st):
                outs.append(p.cg(m, 5, s, 1e-4))
        torch.cuda.synchronize()
        for i, (o, ref) in enumerate(zip(outs, alone)):
            assert np.array_equal(_bits(o.image), _bits(ref.image)), f"rep {rep} plan {i}"
            assert np.array_equal(_bits(o.residuals), _bits(ref.residuals)), f"rep {rep} plan {i}"


def test_captured_plan_keeps_its_buffers_alive():
    from trajsparsenufft_amd.recon import build_plan
    packs, smaps, y = _slice(64, 48, 8, 7)
    plan = build_plan(packs, dev(smaps), "split16")
    scale = plan.operator_scale(dev(y))
    ref = _clone(plan.cg(dev(y), 5, scale, 1e-4))
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        plan.capture_cg(dev(y), 5, scale, 1e-4)  # temporary measurement, result dropped
        gc.collect()
        # fresh allocations of the same sizes, filled with garbage: none may
        # alias the graph's buffers
        junk = [torch.full((y.size,), 7.0 + 7.0j, dtype=torch.complex64, device="cuda") for _ in range(4)]
        junk += [torch.full((64, 64), 3.0 + 3.0j, dtype=torch.complex64, device="cuda") for _ in range(4)]
        stream.synchronize()
        for _ in range(2):
            plan.replay()
            stream.synchronize()
    kept = plan._keep_call[1]
    assert np.array_equal(_bits(kept.image), _bits(ref.image))
    assert np.array_equal(_bits(kept.residuals), _bits(ref.residuals))
    assert all(bool((j == j.flatten()[0]).all()) for j in junk)


def test_dense_tiles_are_expanded_on_first_use():
    from trajsparsenufft_amd._lib import check, lib
    from trajsparsenufft_amd.device import Direction
    packs, _, _ = _slice(64, 24, 4, 8)
    d = Direction.upload(packs.adjoint)
    before = d.device_bytes
    check(lib().tsn_direction_expand_dense(d.handle))
    assert d.device_bytes == before + d.header["tiles"] * 2048
    check(lib().tsn_direction_expand_dense(d.handle))
    assert d.device_bytes == before + d.header["tiles"] * 2048
