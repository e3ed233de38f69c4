"""Quality admission and a paired timing campaign on the GPU: a (case x frame)
matrix of split-FP16 solves admitted against FP32 solves (eager and Graph
bitwise, SSIM/PSNR gates), then tsn_cg under two precisions timed as
order-alternated process pairs."""

import json

import numpy as np
import pytest

from gpu_util import dev, need_gpu, small_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _bits(t):
    return t.cpu().numpy().view(np.uint32 if t.dtype == torch.complex64 else np.uint64)


def test_quality_admission_then_paired_campaign(tmp_path):
    need_gpu()
    from oracle import nudft
    from trajsparsenufft_amd import paired, synth
    from trajsparsenufft_amd.build import CLI
    from trajsparsenufft_amd.recon import build_plan
    from trajsparsenufft_amd.runtime import write_runtime_dir
    if not CLI.exists():
        pytest.skip("tsn_cg is built by trajsparsenufft_amd.build")
    cases = {"radial": small_case(64, 48), "golden_short": small_case(64, 32)}
    coils, frames, iters = 4, 2, 10
    smaps = synth.coil_maps(coils, 64)
    rows, scales = [], {}
    for name, (traj, packs) in cases.items():
        packs.write(tmp_path / "packed" / name)
        omega = (traj.T * 2.0 * np.pi).astype(np.float32)
        split = build_plan(packs, dev(smaps), "split16")
        fp32 = build_plan(packs, dev(smaps), "fp32")
        for frame in range(frames):
            x = synth.phantom(64, 64, seed=11 + frame)
            y = synth.add_noise(nudft.nudft_forward(x, smaps, omega), 40.0, seed=frame).astype(np.complex64)
            if frame == 0:
                write_runtime_dir(tmp_path / "runtime" / name, y, smaps, packs.apod, name=name)
            yd = dev(y)
            scale = split.operator_scale(yd)
            scales.setdefault(name, scale)
            eager = split.cg(yd, iters, scale, 1e-4)
            torch.cuda.synchronize()
            eager_img, eager_res = eager.image.clone(), eager.residuals.clone()
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                g = split.capture_cg(yd, iters, scale, 1e-4)
                split.replay()
                stream.synchronize()
            ref = fp32.cg(yd, iters, scale, 1e-4)
            torch.cuda.synchronize()
            rows.append(paired.quality_row(
                eager_img.cpu().numpy(), ref.image.cpu().numpy(), x,
                eager_graph_image_bitwise=np.array_equal(_bits(eager_img), _bits(g.image)),
                eager_graph_residuals_bitwise=np.array_equal(_bits(eager_res), _bits(g.residuals)),
                case=name, frame=frame, cg_iterations=iters))
    for r in rows:
        print(f"\n[admission] {r['case']} frame {r['frame']}: NRMSE vs FP32 {r['native_vs_fp32_magnitude_nrmse']:.2e} "
              f"SSIM {r['native_vs_fp32_ssim']:.6f} PSNR {r['native_vs_fp32_psnr_db']:.1f} dB")
    paired.validate_rows(rows, list(cases), frames, iters)
    command = paired.tsn_cg_command(CLI, tmp_path / "packed", tmp_path / "runtime", scales, warmup=2, iters=5,
                                    graph=True)
    summary = paired.run_paired({"split16": command, "dense16": command}, list(cases), tmp_path / "campaign",
                                pairs=2, audit=False,
                                variant_env={"split16": {"TSN_PRECISION": "split16"},
                                             "dense16": {"TSN_PRECISION": "dense16"}})
    print(f"\n[paired] dense16 / split16 geomean {summary['geomean_ratio']:.3f} "
          f"[{summary['ci95_low']:.3f}, {summary['ci95_high']:.3f}] {summary['decision']}")
    rows_out = [json.loads(l) for l in (tmp_path / "campaign" / "paired_rows.jsonl").read_text().splitlines()]
    assert len(rows_out) == 2 * len(cases) and all(r["ratio"] > 0 for r in rows_out)
    assert summary["decision"] in ("ACCEPT", "MEASURED")
