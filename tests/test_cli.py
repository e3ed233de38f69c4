"""The tsn_cg command line, a drop-in for the reference binary nufft_fp16x2_cg
(``experiments/production/src/nufft_fp16x2_cg_bench.cu:279-518``): argv,
runtime files, dump files, one JSON line and exit codes 0/1/2/3."""

import json
import os
import subprocess

import numpy as np
import pytest

from trajsparsenufft_amd.build import CLI

pytestmark = pytest.mark.skipif(not CLI.exists(), reason="tsn_cg is built by trajsparsenufft_amd.build")


def _run(*args, env=None):
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=300,
                          env={**os.environ, **(env or {})})


def test_usage_and_argument_errors(tmp_path):
    assert _run().returncode == 2
    assert _run("a", "b", "c", "d", tmp_path, "1.0").returncode == 2          # fewer than 7 arguments
    assert _run("a", "b", "c", "d", tmp_path, "-1", "5").returncode == 2     # operator scale must be > 0
    assert _run("a", "b", "c", "d", tmp_path, "1", "5", "20", "0", "0", "0").returncode == 2  # cg_iterations
    missing = _run(tmp_path / "nope", tmp_path / "nope", tmp_path / "nope", tmp_path / "nope", tmp_path, "1", "5")
    assert missing.returncode == 1 and "cannot open" in missing.stderr


def _write_case(root, packs, smaps, apod, y):
    packs.write(root)
    rt = root / "runtime"
    rt.mkdir()
    np.ascontiguousarray(y, np.complex64).tofile(rt / "measurement.c64.bin")
    np.ascontiguousarray(smaps, np.complex64).tofile(rt / "sensitivity_maps.c64.bin")
    np.ascontiguousarray(apod, np.float32).tofile(rt / "scaling.f32.bin")
    return [root / "G" / "real", root / "G" / "imag", root / "G_T" / "real", root / "G_T" / "imag", rt]


@pytest.mark.gpu
def test_cli_reconstruction_matches_library(tmp_path):
    import torch
    from gpu_util import dev, need_gpu, small_case
    from oracle import nudft
    from trajsparsenufft_amd import synth
    from trajsparsenufft_amd.recon import build_plan, check_cg
    need_gpu()
    image, coils = 64, 4
    traj, packs = small_case(image, 48)
    x = synth.phantom(image, image, seed=2)
    smaps = synth.coil_maps(coils, image)
    omega = (traj.T * 2.0 * np.pi).astype(np.float32)
    y = synth.add_noise(nudft.nudft_forward(x, smaps, omega), 40.0, seed=2).astype(np.complex64)
    plan = build_plan(packs, dev(smaps), "split16")
    scale = plan.operator_scale(dev(y))
    res = plan.cg(dev(y), iterations=6, operator_scale=scale, lam=1e-4)
    torch.cuda.synchronize()
    check_cg(res)
    dirs = _write_case(tmp_path, packs, smaps, packs.apod, y)
    dump = tmp_path / "dump"
    dump.mkdir()
    scale_arg = f"{float(np.float32(scale)):.9g}"  # exact float32 round trip
    for graph in (0, 1):
        out = _run(*dirs, scale_arg, 1, 2, 0, graph, 6, dump)
        assert out.returncode == 0, out.stderr
        line = json.loads(out.stdout.strip().splitlines()[-1])
        assert line["variant"] == "fp16x2_cg10" and line["correct"] and line["graph"] == bool(graph)
        assert line["cg_iterations"] == 6 and line["nonfinite"] == 0 and line["us"] > 0
        rec = np.fromfile(dump / "reconstruction.c64.bin", np.complex64).reshape(image, image)
        # same library, same inputs
        assert np.array_equal(rec.view(np.uint32), res.image.cpu().numpy().view(np.uint32))
        r64 = np.fromfile(dump / "residuals.f64.bin", np.float64)
        np.testing.assert_array_equal(r64, res.residuals.cpu().numpy())
        np.testing.assert_array_equal(np.fromfile(dump / "residuals.f32.bin", np.float32), r64.astype(np.float32))
        assert line["residual0"] == pytest.approx(r64[0]) and line["residual_final"] == pytest.approx(r64[-1])
    # non-finite data: CG failure, exit 3 (reference :414-422)
    bad = y.copy()
    bad[0, 3] = np.nan
    np.ascontiguousarray(bad).tofile(dirs[-1] / "measurement.c64.bin")
    assert _run(*dirs, scale_arg, 0, 1).returncode == 3
    # incompatible pair (forward pack given as adjoint): exit 2
    assert _run(dirs[0], dirs[1], dirs[0], dirs[1], dirs[-1], "1.0", 0, 1).returncode == 2
