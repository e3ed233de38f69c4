#!/usr/bin/env python3
"""CG-SENSE reconstructions per second on one MI355X (gfx950).

Workload (the headline configuration): 256x256 image on a 512x512 grid
(oversampling 2), 8 coils, golden-angle radial trajectory of 402 spokes x 512
samples (205,824 samples per coil), Kaiser-Bessel width 6, 10 CG iterations,
lambda = 1e-4.  Synthetic data: a seeded 12-ellipse phantom with smooth phase,
Gaussian coil maps normalised to sum |S|^2 = 1, measurements from the exact
FP64 non-uniform DFT of the phantom plus complex Gaussian noise at 40 dB SNR.
The trajectory is packed once (native packer) and excluded from timing, as are
uploads.  A "step" is one complete device-resident reconstruction (RHS +
10 CG iterations) of the split-FP16 2:4-sparse-MFMA path, replayed as one
hipGraph; eager and the plain FP32 / dense-MFMA paths are timed beside it.

NRMSE figures: magnitude NRMSE against an FP64 CG on the exact NUDFT of the
same data (evaluated with torch float64 on the GPU; it is only the accuracy
reference, never timed) and against the plain FP32 reconstruction.

Roofline: the interpolation kernels (one complex interpolation application,
hi + residual, one direction) are launched on their own and timed with events
on the launching stream; algorithmic bytes per launch
B = nnz*8 + 2*n_in*2C*2 + n_out*C*8 (nnz = 36*M).

CPU baseline: the oracle's FP32 CSR + FFT CG-SENSE (oracle/cpu_recon.py, a
multi-core C CSR product) on the host cores, one full reconstruction (or a
bounded number of CG iterations, extrapolated), and at 128^2 / 4 coils the same
FP32 CG-SENSE beside the FP64 direct-NUDFT CG-SENSE.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

IMAGE, GRID, COILS, SPOKES, READOUT, CG_ITERS, LAMBDA, SNR_DB = 256, 512, 8, 402, 512, 10, 1e-4, 40.0
HBM_PEAK_GBS = 8000.0


def parse_args():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-fp64-reference", action="store_true")
    p.add_argument("--cpu-budget-s", type=float, default=25.0, help="bound on the CPU baseline sample")
    p.add_argument("--variants", default="split16,fp32,splitbf16,dense16,hi16")
    p.add_argument("--roofline-only", action="store_true",
                   help="only the standalone interpolation launches and a 64 MiB copy (counter runs)")
    p.add_argument("--no-batches", action="store_true",
                   help="skip the batched workloads (48 uneven slices; 24 frames at 768^2 / 16 coils)")
    p.add_argument("--batches-only", action="store_true", help="only the batched workloads")
    p.add_argument("--no-sweep", action="store_true", help="skip the spiral precision sweep")
    return p.parse_args()


def time_once(torch, fn, reps):
    """Average ms of fn over reps calls after one warm-up call (long launches)."""

    fn()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    stop.record()
    stop.synchronize()
    return start.elapsed_time(stop) / reps


def batch_workloads(torch, log, reps=3):
    """The two batched workloads, each one hipGraph:

    * slices: 48 independent 256^2 / 512^2 slices, spoke counts uniform in
      64..402 (x 512 samples), coils from {4, 8, 12, 16}, each with its own
      golden-angle offset and pack, schedules from the autotuner, 10 CG steps;
      batched reconstructions/s against a loop of per-slice graph replays.
    * frames: 24 frames of a 384^2 image on a 768^2 grid, 16 coils, sharing
      one packed golden-angle radial trajectory of 600 spokes x 768 samples,
      seeded sinusoidal motion, 35 dB SNR, 30 CG steps batched across frames;
      frames/s, the interpolation's HBM rate at this shape, and NRMSE of every
      frame against the plain FP32 path.

    Data come from the FP32 operator of each slice plus complex Gaussian
    noise (the FP64 NUDFT of 48 + 24 images would dominate the run)."""

    from trajsparsenufft_amd import device, workloads
    from trajsparsenufft_amd.device import Direction
    from trajsparsenufft_amd.recon import build_batch, build_plan
    out = {}
    # ---------------- slices
    n_slices = 48
    t0 = time.perf_counter()
    w = workloads.uneven_slices(n_slices)
    cases, meas, scales = w.cases, w.measurements, w.scales
    spokes, coils = w.spokes, w.coils
    plans = [build_plan(packs, maps, "split16", autotune=True) for packs, maps in cases]
    prep = time.perf_counter() - t0
    batch = build_batch(cases, "split16", autotune=True)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        outs = batch.capture_cg(meas, CG_ITERS, scales, LAMBDA)
        batch_ms = time_once(torch, batch.replay, reps)
        singles = [p.capture_cg(y, CG_ITERS, s, LAMBDA) for p, y, s in zip(plans, meas, scales)]

        def loop():
            for p in plans:
                p.replay()
        loop_ms = time_once(torch, loop, reps)
        stream.synchronize()
    same = all(np.array_equal(o.image.cpu().numpy().view(np.uint32), r.image.cpu().numpy().view(np.uint32))
               for o, r in zip(outs, singles))
    out["slices"] = {"workload": "48 slices 256x256 / 512x512, spokes U[64,402] x 512, coils {4,8,12,16}, "
                                 "own golden-angle offset and pack, autotuned schedules, 10 CG, split16",
                     "batched_recon_per_s": n_slices * 1000.0 / batch_ms, "batched_ms": batch_ms,
                     "loop_recon_per_s": n_slices * 1000.0 / loop_ms, "loop_ms": loop_ms,
                     "speedup_vs_loop": loop_ms / batch_ms, "bitwise_equal_to_per_slice": same,
                     "spokes": [int(v) for v in spokes], "coils": [int(v) for v in coils],
                     "batch_tuning": batch.tuning, "prep_seconds": prep}
    log(f"[bench] slices: batched {batch_ms:.2f} ms ({n_slices * 1000.0 / batch_ms:.0f} recon/s), per-slice loop "
        f"{loop_ms:.2f} ms, bitwise {same}")
    del batch, outs, singles, plans, cases, meas, w
    torch.cuda.empty_cache()
    # ---------------- frames
    nc, frames, iters = workloads.FRAME_COILS, 24, 30
    t0 = time.perf_counter()
    w = workloads.dynamic_frames(frames)
    packs, maps = w.cases[0]
    meas, truth, scale = w.measurements, w.truth, w.scales[0]
    fp32 = build_plan(packs, maps, "fp32")
    prep = time.perf_counter() - t0
    batch = build_batch(w.cases, "split16", shared=True, autotune=True)
    with torch.cuda.stream(stream):
        outs = batch.capture_cg(meas, iters, w.scales, LAMBDA)
        frames_ms = time_once(torch, batch.replay, reps)
        stream.synchronize()
    nrmse, truth_err = [], []
    for f in range(frames):
        ref = fp32.cg(meas[f], iters, scale, LAMBDA)
        torch.cuda.synchronize()
        img = outs[f].image.cpu().numpy()
        nrmse.append(magnitude_nrmse(img, ref.image.cpu().numpy()))
        truth_err.append(magnitude_nrmse(img, truth[f]))
    # interpolation HBM rate at this shape (16 coils: two coil blocks per launch)
    fh = packs.forward.header
    m, g, nnz = fh["output_size"], fh["n_cols"], fh["nnz"]
    fwd, adj = Direction.upload(packs.forward), Direction.upload(packs.adjoint)
    fwd.autotune(nc)
    adj.autotune(nc)
    prng = np.random.default_rng(6)

    def panel(rows):
        v = prng.standard_normal((nc, rows)) + 1j * prng.standard_normal((nc, rows))
        return device.split_panel_il(torch.from_numpy(v.astype(np.complex64)).cuda(), nc)
    gp, sp = panel(g), panel(m)
    out_s = torch.empty((2, m, 16), dtype=torch.int32, device="cuda")
    out_g = torch.zeros((g, nc), dtype=torch.complex64, device="cuda")
    tuned = {"forward": fwd.coil_waves, "adjoint": adj.coil_waves}
    interp = {}
    for waves in (1, 2):
        fwd.coil_waves = adj.coil_waves = waves
        f_ms = time_launches(torch, lambda: device.interp_il(fwd, gp, out_s, nc, False, False, 1.0, "split_il"), 20, 5)
        a_ms = time_launches(torch, lambda: device.interp_il(adj, sp, out_g, nc, True, False, 1.0, "row_major"), 20, 5)
        for name, ms, n_in, n_out in (("forward", f_ms, g, m), ("adjoint", a_ms, m, g)):
            b = nnz * 8 + 2 * n_in * 2 * nc * 2 + n_out * nc * 8
            row = {"us": ms * 1e3, "bytes": b, "GBs": b / (ms * 1e-3) / 1e9,
                   "frac_of_peak": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
            interp.setdefault(name, {})[f"coil_waves_{waves}"] = row
            if waves == tuned[name]:
                interp[name].update(row)
        log(f"[bench] frames interp, {waves} wave(s) per entry: forward {f_ms * 1e3:.1f} us, adjoint {a_ms * 1e3:.1f} us")
    interp["tuned_coil_waves"] = tuned
    out["frames"] = {"workload": "24 frames 384x384 / 768x768, 16 coils, shared golden-angle radial 600x768 pack, "
                                 "35 dB, 30 CG batched across frames, split16",
                     "frames_per_s": frames * 1000.0 / frames_ms, "batched_ms": frames_ms,
                     "nrmse_vs_fp32": {"max": max(nrmse), "mean": float(np.mean(nrmse))},
                     "nrmse_vs_truth": {"max": max(truth_err), "mean": float(np.mean(truth_err))},
                     "interp": interp, "batch_tuning": batch.tuning, "prep_seconds": prep}
    log(f"[bench] frames: {frames_ms:.1f} ms for {frames} frames ({frames * 1000.0 / frames_ms:.1f} frames/s), "
        f"NRMSE vs FP32 max {max(nrmse):.2e}, interp fwd {interp['forward']['GBs']:.0f} / adj "
        f"{interp['adjoint']['GBs']:.0f} GB/s")
    return out


def plain_128_gpu(torch, steps, warmup, log):
    """The CPU-runnable configuration on the GPU (128^2 / 256^2, 4 coils,
    golden-angle radial 96 x 256, 10 CG, 40 dB; data and FP64 CG reference as
    for the headline shape): split16 and FP32 graph latency and magnitude
    NRMSE against the FP64 NUDFT CG, beside the CPU baseline's timing of the
    same configuration."""

    from trajsparsenufft_amd import pack, synth
    from trajsparsenufft_amd.recon import build_plan, check_cg
    n, grid, coils = 128, 256, 4
    traj = synth.radial(96, 256, golden=True)
    x_true = synth.phantom(n, seed=0)
    smaps = synth.coil_maps(coils, n)
    ref = Fp64Nudft(torch, traj, smaps, n)
    y = synth.add_noise(ref.forward(torch.from_numpy(x_true).to(torch.complex128).cuda()).cpu().numpy(), SNR_DB,
                        seed=1)
    x64 = ref.cg(torch.from_numpy(y).to(torch.complex128).cuda(), CG_ITERS, LAMBDA, grid).cpu().numpy()
    del ref
    packs = pack.prepare_trajectory(traj, (n, n), (grid, grid))
    y_dev, smaps_dev = torch.from_numpy(y).cuda(), torch.from_numpy(smaps).cuda()
    plans = {v: build_plan(packs, smaps_dev, v) for v in ("split16", "fp32")}
    scale = plans["fp32"].operator_scale(y_dev)
    out = {"workload": "128x128 / 256x256 / 4 coils / golden-angle radial 96x256 / 10 CG"}
    for v, plan in plans.items():
        res = plan.cg(y_dev, CG_ITERS, scale, LAMBDA)
        torch.cuda.synchronize()
        check_cg(res)
        img = res.image.cpu().numpy()
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            plan.capture_cg(y_dev, CG_ITERS, scale, LAMBDA)
            ms = time_launches(torch, plan.replay, steps, warmup)
            stream.synchronize()
        out[v] = {"graph_ms": ms, "nrmse_vs_fp64": magnitude_nrmse(img, x64)}
    out["speedup_split16_vs_fp32"] = out["fp32"]["graph_ms"] / out["split16"]["graph_ms"]
    log(f"[bench] 128^2 plain config: split16 {out['split16']['graph_ms']:.3f} ms, FP32 {out['fp32']['graph_ms']:.3f} ms, "
        f"NRMSE vs FP64 {out['split16']['nrmse_vs_fp64']:.2e} / {out['fp32']['nrmse_vs_fp64']:.2e}")
    return out


def spiral_sweep(torch, variants, log):
    """Precision sweep at the headline shape with a variable-density spiral
    (16 interleaves x 8192 samples, density ~ r^-1.5; 256^2 / 512^2, 8 coils,
    40 dB, 10 CG): per variant the magnitude NRMSE against the FP64
    direct-NUDFT CG, the change of the error against the truth relative to the
    FP64 CG's, the adjointness defect |<Ax, y> - <x, A^H y>| / max of both,
    and the graph-replay latency."""

    from trajsparsenufft_amd import pack, synth
    from trajsparsenufft_amd.recon import build_plan, check_cg
    traj = synth.spiral(16, 8192)
    x = synth.phantom(IMAGE, seed=4)
    smaps = synth.coil_maps(COILS, IMAGE)
    ref = Fp64Nudft(torch, traj, smaps, IMAGE)
    y = ref.forward(torch.from_numpy(x).to(torch.complex128).cuda()).cpu().numpy()
    y = synth.add_noise(y, SNR_DB, seed=5)
    x64 = ref.cg(torch.from_numpy(y).to(torch.complex128).cuda(), CG_ITERS, LAMBDA, GRID).cpu().numpy()
    del ref
    packs = pack.prepare_trajectory(traj, (IMAGE, IMAGE), (GRID, GRID))
    smaps_dev = torch.from_numpy(smaps).cuda()
    y_dev = torch.from_numpy(y).cuda()
    rng = np.random.default_rng(6)
    xs = (rng.standard_normal((IMAGE, IMAGE)) + 1j * rng.standard_normal((IMAGE, IMAGE))).astype(np.complex64)
    ys = (rng.standard_normal(y.shape) + 1j * rng.standard_normal(y.shape)).astype(np.complex64)
    fp64_truth = magnitude_nrmse(x64, x)
    out = {"workload": "spiral 16 interleaves x 8192 (density r^-1.5), 256x256 / 512x512, 8 coils, 40 dB, 10 CG",
           "fp64_vs_truth": fp64_truth, "variants": {}}
    for v in variants:
        plan = build_plan(packs, smaps_dev, v)
        scale = plan.operator_scale(y_dev)
        res = plan.cg(y_dev, CG_ITERS, scale, LAMBDA)
        torch.cuda.synchronize()
        check_cg(res)
        img = res.image.cpu().numpy()
        left = np.vdot(plan.forward(torch.from_numpy(xs).cuda()).cpu().numpy().astype(np.complex128), ys)
        right = np.vdot(xs.astype(np.complex128), plan.adjoint(torch.from_numpy(ys).cuda()).cpu().numpy())
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            plan.capture_cg(y_dev, CG_ITERS, scale, LAMBDA)
            graph_ms = time_launches(torch, plan.replay, 10, 3)
            stream.synchronize()
        row = {"nrmse_vs_fp64": magnitude_nrmse(img, x64),
               "truth_delta": magnitude_nrmse(img, x) - fp64_truth,
               "adjointness": float(abs(left - right) / max(abs(left), abs(right))),
               "graph_ms": graph_ms}
        out["variants"][v] = row
        log(f"[bench] sweep {v}: NRMSE vs FP64 CG {row['nrmse_vs_fp64']:.2e}, truth delta {row['truth_delta']:+.2e}, "
            f"adjointness {row['adjointness']:.1e}, {graph_ms:.3f} ms")
    return out


class Fp64Nudft:
    """Exact NUDFT with SENSE in torch float64 on the GPU (data synthesis and
    the FP64 CG accuracy reference; not part of the measured path)."""

    def __init__(self, torch, traj, smaps, n):
        self.t = torch
        om = torch.from_numpy((traj.T * 2.0 * np.pi).astype(np.float32)).double().cuda()
        k = torch.arange(n, dtype=torch.float64, device="cuda") - n // 2
        self.e0 = torch.exp(-1j * om[0][:, None] * k[None, :])  # [M, n]
        self.e1 = torch.exp(-1j * om[1][:, None] * k[None, :])
        self.s = torch.from_numpy(smaps).to(torch.complex128).cuda()

    def forward(self, x):
        out = []
        for c in range(self.s.shape[0]):
            t = (self.s[c] * x) @ self.e1.T                      # [n0, M]
            out.append(self.t.sum(self.e0 * t.T, dim=1))
        return self.t.stack(out)

    def adjoint(self, y):
        acc = self.t.zeros_like(self.s[0])
        for c in range(self.s.shape[0]):
            acc += self.s[c].conj() * (self.e0.conj().T @ (y[c][:, None] * self.e1.conj()))
        return acc

    def cg(self, y, iterations, lam, grid):
        """(E^H E + lam) x = E^H y in the native normalisation, reference CG order."""

        t = self.t
        s = float(grid)  # sqrt(K0*K1) for a square grid
        b = self.adjoint(y / s) / s
        lam_n = lam / (s * s)
        x = t.zeros_like(b)
        r = b.clone()
        p = r.clone()
        rs = t.vdot(r.flatten(), r.flatten()).real
        for _ in range(iterations):
            ap = self.adjoint(self.forward(p) / s) / s + lam_n * p
            alpha = rs / t.vdot(p.flatten(), ap.flatten()).real
            x = x + alpha * p
            r = r - alpha * ap
            rs_new = t.vdot(r.flatten(), r.flatten()).real
            p = r + (rs_new / rs) * p
            rs = rs_new
        return x


def magnitude_nrmse(a, b) -> float:
    a, b = np.abs(np.asarray(a)), np.abs(np.asarray(b))
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def time_launches(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    stop.synchronize()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps  # ms


TRAFFIC_TABLE = "profiles/hbm_traffic_256x512_8coil_radial402.md"
COPY_BYTES = 64 << 20  # calibration copy of --roofline-only


def measured_traffic(kernel_substr: str):
    """HBM bytes per launch of a kernel from the committed summary of the
    rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of ``bench.py
    --roofline-only``: the table's last column, KiB per dispatch scaled by what
    each counter reads for the run's 64 MiB device-to-device copy.  None if the
    table or the kernel's row is absent."""

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), TRAFFIC_TABLE)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 7 and cells[0].strip("`") == kernel_substr:
                return float(cells[6])
    return None


def interp_roofline(torch, packs, steps, warmup, log):
    """Standalone launches of the packed interpolation exactly as the
    reconstruction runs it (interleaved split panel in; the forward writes the
    adjoint's split panel, the adjoint writes row-major grid rows), timed with
    events on the launch stream; algorithmic bytes per launch as defined for
    the unit of one complex interpolation application: nnz * (2 + 2 + 4) +
    2 * n_in * 2C * 2 + n_out * C * 8 (the split panel out is the same
    n_out * 16 channels * 4 bytes)."""

    from trajsparsenufft_amd import device
    from trajsparsenufft_amd.device import Direction
    fh = packs.forward.header
    m, g, nnz = fh["output_size"], fh["n_cols"], fh["nnz"]
    fwd, adj = Direction.upload(packs.forward), Direction.upload(packs.adjoint)
    rng = np.random.default_rng(5)

    def panel(rows):
        v = rng.standard_normal((COILS, rows)) + 1j * rng.standard_normal((COILS, rows))
        return device.split_panel_il(torch.from_numpy(v.astype(np.complex64)).cuda(), COILS)

    gp, sp = panel(g), panel(m)
    out_s = torch.empty((1, m, 16), dtype=torch.int32, device="cuda")
    out_g = torch.zeros((g, COILS), dtype=torch.complex64, device="cuda")
    fwd_ms = time_launches(torch, lambda: device.interp_il(fwd, gp, out_s, COILS, False, False, 1.0 / GRID,
                                                           "split_il"), steps, warmup)
    adj_ms = time_launches(torch, lambda: device.interp_il(adj, sp, out_g, COILS, True, False, 1.0, "row_major"),
                           steps, warmup)
    res = {}
    for name, ms, n_in, n_out, conj, om in (("forward", fwd_ms, g, m, "false", 3), ("adjoint", adj_ms, m, g, "true", 1)):
        b = nnz * 8 + 2 * n_in * 2 * COILS * 2 + n_out * COILS * 8
        kernel = f"interp_tiles_kernel<0, 1, true, {om}, {conj}, false, 1>"
        res[name] = {"kernel": kernel, "us": ms * 1e3, "bytes": b, "GBs": b / (ms * 1e-3) / 1e9,
                     "traffic": measured_traffic(kernel)}
        log(f"[bench] interp {name} {ms * 1e3:.1f} us ({res[name]['GBs']:.0f} GB/s algorithmic, "
            f"HBM traffic {res[name]['traffic']})")
    # the plain FP32 path's CSR launches as its reconstruction runs them
    # (row-major float2 rows in and out): B = nnz * 12 + n_in * C * 8 + n_out * C * 8
    cf, ca = Direction.csr(packs.kb, "G"), Direction.csr(packs.kb, "G_T")
    xg = torch.from_numpy((rng.standard_normal((g, COILS)) + 1j * rng.standard_normal((g, COILS))).astype(np.complex64)).cuda()
    xs = torch.from_numpy((rng.standard_normal((m, COILS)) + 1j * rng.standard_normal((m, COILS))).astype(np.complex64)).cuda()
    os_ = torch.empty((m, COILS), dtype=torch.complex64, device="cuda")
    og = torch.zeros((g, COILS), dtype=torch.complex64, device="cuda")
    f32_ms = time_launches(torch, lambda: device.interp_f32(cf, xg, os_, COILS, False, False, 1.0 / GRID, "row_major"),
                           steps, warmup)
    a32_ms = time_launches(torch, lambda: device.interp_f32(ca, xs, og, COILS, True, False, 1.0, "row_major"),
                           steps, warmup)
    for name, ms, n_in, n_out, conj in (("fp32_forward", f32_ms, g, m, "false"), ("fp32_adjoint", a32_ms, m, g, "true")):
        b = nnz * 12 + n_in * COILS * 8 + n_out * COILS * 8
        kernel = f"interp_csr_kernel<8, 1, {conj}, false>"
        res[name] = {"kernel": kernel, "us": ms * 1e3, "bytes": b, "GBs": b / (ms * 1e-3) / 1e9,
                     "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": measured_traffic(kernel)}
        log(f"[bench] FP32 CSR {name[5:]} {ms * 1e3:.1f} us ({res[name]['GBs']:.0f} GB/s algorithmic)")
    res["dominant"] = res["adjoint"] if adj_ms >= fwd_ms else res["forward"]
    res["traffic_source"] = (f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py --roofline-only, "
                             f"scaled by a {COPY_BYTES >> 20} MiB copy: {TRAFFIC_TABLE}")
    return res


def roofline_only(args) -> int:
    """Interpolation launches, the calibration copy and three split-FP16
    reconstructions (every FFT and CG kernel), for counter runs."""

    import torch
    from trajsparsenufft_amd import pack, synth
    from trajsparsenufft_amd.recon import build_plan
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    traj = synth.radial(SPOKES, READOUT, golden=True)
    packs = pack.prepare_trajectory(traj, (IMAGE, IMAGE), (GRID, GRID))
    roof = interp_roofline(torch, packs, args.steps, args.warmup, log)
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.steps):
        dst.copy_(src)
    plan = build_plan(packs, torch.from_numpy(synth.coil_maps(COILS, IMAGE)).cuda(), "split16")
    rng = np.random.default_rng(3)
    m = traj.shape[0]
    y = torch.from_numpy((rng.standard_normal((COILS, m)) + 1j * rng.standard_normal((COILS, m))).astype(np.complex64)).cuda()
    scale = plan.operator_scale(y)
    for _ in range(3):
        plan.cg(y, CG_ITERS, scale, LAMBDA)
    torch.cuda.synchronize()
    print(json.dumps({k: roof[k] for k in ("forward", "adjoint", "fp32_forward", "fp32_adjoint")}), flush=True)
    return 0


def host_cores() -> int:
    """Threads of the CPU baseline: the cores in this process's affinity,
    capped by OMP_NUM_THREADS when the host sets it (the host's share)."""

    cores = len(os.sched_getaffinity(0))
    try:
        cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    except ValueError:
        pass
    return max(1, cores)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(packs, smaps, y, scale, budget_s, log):
    """The oracle's CPU CG-SENSE, timed on the host cores (oracle/cpu_recon.py
    with the multi-core C CSR product, scipy.fft on the same threads):

    * FP32 CSR + FFT CG-SENSE at the benchmark shape (256^2 / 512^2, 8 coils,
      402 x 512 radial, 10 CG + the right-hand side): ``value``; if one
      reconstruction would exceed the budget, a bounded number of CG
      iterations is timed and scaled to the full solve;
    * at 128^2 / 256^2, 4 coils, 96 x 256 radial, 10 CG: the same FP32 CSR
      CG-SENSE and the FP64 direct-NUDFT CG-SENSE (BLAS on the same threads)."""

    import threadpoolctl
    from oracle import nudft
    from oracle.cpu_recon import CpuSense
    from trajsparsenufft_amd import pack, synth
    cores = host_cores()
    model = cpu_model()
    cols, re_, im_ = packs.kb.arrays()
    with threadpoolctl.threadpool_limits(cores):
        op = CpuSense(cols, re_, im_, smaps, packs.apod, (IMAGE, IMAGE), (GRID, GRID), workers=cores, backend="omp")
        t_apply = time.perf_counter()
        op.adjoint(op.forward(synth.phantom(IMAGE, seed=0)))
        apply_s = time.perf_counter() - t_apply
        applies_per_recon = 1 + 2 * CG_ITERS  # RHS adjoint + (forward + adjoint) per iteration
        iters = CG_ITERS if apply_s * 2 * (CG_ITERS + 1) <= budget_s else max(1, int(budget_s / (2 * apply_s)) - 1)
        t_cpu = time.perf_counter()
        op.cg(y, iters, scale)
        cpu_s = time.perf_counter() - t_cpu
        per_recon = cpu_s * applies_per_recon / (1 + 2 * iters)
        # the 128^2 plain configuration: FP32 CSR and the FP64 NUDFT oracle
        traj0 = synth.radial(96, 256, golden=True)
        packs0 = pack.prepare_trajectory(traj0, (128, 128), (256, 256))
        x0, maps0 = synth.phantom(128, seed=0), synth.coil_maps(4, 128)
        omega0 = (traj0.T * 2.0 * np.pi).astype(np.float32)
        y0 = synth.add_noise(nudft.nudft_forward(x0, maps0, omega0), SNR_DB, seed=1)
        c0, r0, i0 = packs0.kb.arrays()
        op0 = CpuSense(c0, r0, i0, maps0, packs0.apod, (128, 128), (256, 256), workers=cores, backend="omp")
        s0 = float(np.sqrt(np.abs(op0.adjoint(y0 / 256.0)).max()))
        t0 = time.perf_counter()
        x32, _, _ = op0.cg(y0, CG_ITERS, s0)
        fp32_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        x64, _, _ = nudft.nudft_cg_sense(y0, maps0, omega0, (256, 256), CG_ITERS, LAMBDA)
        fp64_s = time.perf_counter() - t0
    log(f"[bench] cpu baseline ({cores} threads, {model}): {per_recon:.2f} s per 256^2 reconstruction; "
        f"128^2: FP32 CSR {fp32_s:.2f} s, FP64 NUDFT {fp64_s:.2f} s")
    return {"value": 1.0 / per_recon, "unit": "recon/s", "cores": cores, "kind": "port",
            "sample": f"oracle FP32 CSR (C, OpenMP) + scipy.fft CG-SENSE at the benchmark shape on {cores} threads, "
                      f"{iters} CG iterations timed ({cpu_s:.1f} s) and scaled to 10 iterations + RHS",
            "cpu_model": model, "seconds_per_recon": per_recon,
            "plain_128": {"workload": "128x128 / 256x256 / 4 coils / golden-angle radial 96x256 / 10 CG",
                          "fp32_csr_seconds": fp32_s, "fp64_nudft_seconds": fp64_s,
                          "fp32_vs_fp64_nrmse": magnitude_nrmse(x32, x64)}}


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "independent slices/frames shard across GPUs without communication", file=sys.stderr)
        return 2
    if args.roofline_only:
        return roofline_only(args)
    import torch
    if args.batches_only:
        log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
        print(json.dumps(batch_workloads(torch, log)))
        return 0
    from trajsparsenufft_amd import device, pack, synth
    from trajsparsenufft_amd.recon import build_plan, check_cg

    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    torch.manual_seed(0)
    traj = synth.radial(SPOKES, READOUT, golden=True)
    x_true = synth.phantom(IMAGE, seed=0)
    smaps = synth.coil_maps(COILS, IMAGE)
    ref = Fp64Nudft(torch, traj, smaps, IMAGE)
    y_clean = ref.forward(torch.from_numpy(x_true).to(torch.complex128).cuda()).cpu().numpy()
    y = synth.add_noise(y_clean, SNR_DB, seed=1)
    log(f"[bench] data: {y.shape} samples, SNR {SNR_DB} dB")

    t0 = time.perf_counter()
    packs = pack.prepare_trajectory(traj, (IMAGE, IMAGE), (GRID, GRID))
    prep_s = time.perf_counter() - t0
    fh, ah = packs.forward.header, packs.adjoint.header
    log(f"[bench] packed in {prep_s:.2f}s: G {fh['tiles']} tiles / {fh['groups']} groups, "
        f"G_T {ah['tiles']} tiles / {ah['groups']} groups")

    y_dev = torch.from_numpy(y).cuda()
    smaps_dev = torch.from_numpy(smaps).cuda()
    variants = [v for v in args.variants.split(",") if v]
    plans = {v: build_plan(packs, smaps_dev, v) for v in variants}
    scale = plans["fp32" if "fp32" in plans else variants[0]].operator_scale(y_dev)
    log(f"[bench] operator scale (native) {scale:.6g}")

    images, timings = {}, {}
    for v, plan in plans.items():
        res = plan.cg(y_dev, CG_ITERS, scale, LAMBDA)
        torch.cuda.synchronize()
        check_cg(res)
        images[v] = res.image.cpu().numpy()
        eager_ms = time_launches(torch, lambda: plan.cg(y_dev, CG_ITERS, scale, LAMBDA, out=res),
                                 args.steps, args.warmup)
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            gout = plan.capture_cg(y_dev, CG_ITERS, scale, LAMBDA)
            graph_ms = time_launches(torch, plan.replay, args.steps, args.warmup)
            stream.synchronize()
        if not np.array_equal(gout.image.cpu().numpy().view(np.uint32), images[v].view(np.uint32)):
            raise RuntimeError(f"{v}: graph replay differs from eager")
        timings[v] = {"eager_ms": eager_ms, "graph_ms": graph_ms}
        log(f"[bench] {v}: eager {eager_ms:.3f} ms, graph {graph_ms:.3f} ms per reconstruction")

    nrmse = {}
    if not args.no_fp64_reference:
        x64 = ref.cg(torch.from_numpy(y).to(torch.complex128).cuda(), CG_ITERS, LAMBDA, GRID).cpu().numpy()
        nrmse = {f"{v}_vs_fp64": magnitude_nrmse(images[v], x64) for v in images}
        nrmse["fp64_vs_truth"] = magnitude_nrmse(x64, x_true)
    if "fp32" in images:
        for v in images:
            if v != "fp32":
                nrmse[f"{v}_vs_fp32"] = magnitude_nrmse(images[v], images["fp32"])
    log(f"[bench] nrmse {nrmse}")
    del ref

    # ---- roofline of the interpolation kernels (standalone launches)
    roof = interp_roofline(torch, packs, max(args.steps, 20), args.warmup, log)
    dominant = roof["dominant"]

    # ---- CPU baseline (the oracle's CPU reconstructions on the host cores)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(packs, smaps, y, scale, args.cpu_budget_s, log)
    head = timings["split16"]
    value = 1000.0 / head["graph_ms"]
    out = {
        "metric": "CG-SENSE recon/s 256²/8-coil/10-CG: split-FP16 smfmac vs FP32, NRMSE≤0.5%",
        "value": value,
        "unit": "recon/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["graph_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16x2",
        "data": "synthetic (seeded ellipse phantom, Gaussian coil maps, exact FP64 NUDFT + 40 dB noise)",
        "config": {"workload": "cg-sense 256x256 image / 512x512 grid / 8 coils / golden-angle radial 402x512 / "
                               "KB width 6 / 10 CG / lambda 1e-4",
                   "samples_per_coil": fh["output_size"], "precision": "fp16 hi+res panels x fp16 2:4 tiles (smfmac), fp32 accumulate, "
                                                         "fp32 FFT, fp64 CG scalars",
                   "mode": "hipGraph replay of the full reconstruction"},
        "recon_per_s": {v: {"eager": 1000.0 / t["eager_ms"], "graph": 1000.0 / t["graph_ms"]} for v, t in timings.items()},
        "latency_ms": timings,
        "speedup_split16_vs_fp32": (timings["fp32"]["graph_ms"] / head["graph_ms"]) if "fp32" in timings else None,
        "nrmse": nrmse,
        "roofline": {"bound": "hbm", "kernel": dominant["kernel"], "achieved": dominant["GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dominant["GBs"] / HBM_PEAK_GBS,
                     "traffic": dominant["traffic"], "algorithmic_bytes": dominant["bytes"],
                     "avg_launch_us": dominant["us"], "traffic_source": roof["traffic_source"]},
        "interp": {k: roof[k] for k in ("forward", "adjoint", "fp32_forward", "fp32_adjoint")} | {
            "pack_tiles": {"G": fh["tiles"], "G_T": ah["tiles"]}},
        "prep_seconds": prep_s,
        "cpu_baseline": cpu,
    }
    if not args.no_fp64_reference:  # (profiled runs leave it out: its kernels share the headline's names)
        out["plain_128_gpu"] = plain_128_gpu(torch, args.steps, args.warmup, log)
    if not args.no_sweep:
        del plans
        torch.cuda.empty_cache()
        out["sweep"] = spiral_sweep(torch, variants, log)
    if not args.no_batches:
        torch.cuda.empty_cache()
        out["batches"] = batch_workloads(torch, log)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
