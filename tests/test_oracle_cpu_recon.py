"""Pin the timed CPU baseline (scipy CSR + scipy.fft CG-SENSE) against the
float64 arithmetic model and the exact FP64 NUDFT CG."""

import numpy as np

from oracle import nudft, solver, tkbn
from oracle.cpu_recon import CpuSense


def _case(image=24, spokes=16, coils=4):
    ang = np.arange(spokes) * np.pi * (3 - np.sqrt(5))
    rad = np.linspace(-0.5, 0.5, 2 * image, endpoint=False)
    traj = np.stack((np.cos(ang)[:, None] * rad, np.sin(ang)[:, None] * rad), -1).reshape(-1, 2).astype(np.float32)
    omega = (traj.T * 2.0 * np.pi).astype(np.float32)
    k = 2 * image
    cols, re, im = tkbn.kb_interpolation_f32(omega, (image, image), (k, k))
    rng = np.random.default_rng(4)
    smaps = ((rng.standard_normal((coils, image, image)) + 1j * rng.standard_normal((coils, image, image))) / 3
             ).astype(np.complex64)
    apod = tkbn.apodization((image, image), (k, k))
    x = (rng.standard_normal((image, image)) + 1j * rng.standard_normal((image, image))).astype(np.complex64)
    return omega, cols, re, im, smaps, apod, x, (image, image), (k, k)


def test_cpu_operator_matches_fp64_model():
    omega, cols, re, im, smaps, apod, x, isz, gsz = _case()
    cpu = CpuSense(cols, re, im, smaps, apod, isz, gsz)
    model = solver.KbOperator(cols, re, im, smaps, apod, isz, gsz)
    y = cpu.forward(x)
    assert y.dtype == np.complex64
    assert nudft.complex_rel_l2(y, model.forward(x)) < 2e-6
    ys = y[::-1].copy()
    assert nudft.complex_rel_l2(cpu.adjoint(ys), model.adjoint(ys.astype(np.complex128))) < 2e-6


def test_cpu_cg_matches_model_and_exact_nudft_cg():
    omega, cols, re, im, smaps, apod, x, isz, gsz = _case()
    meas = nudft.nudft_forward(x, smaps, omega)  # raw (unnormalised) measurement
    cpu = CpuSense(cols, re, im, smaps, apod, isz, gsz)
    got, hist, status = cpu.cg(meas.astype(np.complex64), 8, 1.7)
    assert status["completed_iterations"] == 8 and hist[-1] < hist[0]
    model = solver.KbOperator(cols, re, im, smaps, apod, isz, gsz)
    emu, _, _ = solver.cg_native(model, meas, 8, 1.7)
    assert nudft.complex_rel_l2(got, emu) < 1e-5
    exact, _, _ = nudft.nudft_cg_sense(meas, smaps, omega, gsz, 8)
    assert nudft.magnitude_nrmse(got, exact) < 2e-4


def test_omp_backend_equals_scipy_backend():
    """The C CSR product (any thread count) adds each row in entry order, as
    scipy's CSR product does: the operators agree to float32 rounding."""

    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["make", "-s", "-C", os.path.join(root, "oracle")], check=True)
    omega, cols, re, im, smaps, apod, x, isz, gsz = _case()
    ref = CpuSense(cols, re, im, smaps, apod, isz, gsz)
    y = ref.forward(x)
    for threads in (1, 3, 8):
        omp = CpuSense(cols, re, im, smaps, apod, isz, gsz, workers=threads, backend="omp")
        assert nudft.complex_rel_l2(omp.forward(x), y) < 1e-6
        assert nudft.complex_rel_l2(omp.adjoint(y), ref.adjoint(y)) < 1e-6
