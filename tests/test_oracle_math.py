"""Pin the numerical oracles against what the mathematics fixes."""

import numpy as np
import pytest

from oracle import nudft, solver, tkbn


def _radial(spokes, readout):
    ang = np.arange(spokes) * np.pi * (3 - np.sqrt(5))
    rad = np.linspace(-0.5, 0.5, readout, endpoint=False)
    return np.stack((np.cos(ang)[:, None] * rad, np.sin(ang)[:, None] * rad), -1).reshape(-1, 2).astype(np.float32)


def test_nudft_closed_forms():
    rng = np.random.default_rng(0)
    n = 8
    omega = rng.uniform(-np.pi, np.pi, size=(2, 20)).astype(np.float32)
    smaps = np.ones((1, n, n))
    # single pixel at (a, b) -> pure exponential exp(-i (w0 (a - n/2) + w1 (b - n/2)))
    a, b = 3, 6
    img = np.zeros((n, n))
    img[a, b] = 1.0
    y = nudft.nudft_forward(img, smaps, omega)[0]
    expect = np.exp(-1j * (omega[0].astype(np.float64) * (a - n // 2) + omega[1].astype(np.float64) * (b - n // 2)))
    np.testing.assert_allclose(y, expect, rtol=1e-13, atol=1e-13)
    # omega = 0 returns the image sum
    img = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    y0 = nudft.nudft_forward(img, smaps, np.zeros((2, 1), np.float32))[0, 0]
    assert abs(y0 - img.sum()) < 1e-12
    # exact adjoint
    maps = rng.standard_normal((3, n, n)) + 1j * rng.standard_normal((3, n, n))
    ys = rng.standard_normal((3, 20)) + 1j * rng.standard_normal((3, 20))
    left = np.vdot(nudft.nudft_forward(img, maps, omega), ys)
    right = np.vdot(img, nudft.nudft_adjoint(ys, maps, omega))
    assert abs(left - right) / abs(left) < 1e-12


def test_cg_solves_small_spd_system_exactly():
    rng = np.random.default_rng(1)
    m = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    a = m.conj().T @ m + np.eye(6)
    x_true = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    x, hist, status = nudft.cg_sense(lambda p: a @ p, a @ x_true, 6)
    assert status["completed_iterations"] == 6
    np.testing.assert_allclose(x, x_true, rtol=1e-8, atol=1e-8)
    assert hist[-1] < 1e-8 * hist[0]


@pytest.mark.parametrize("image", [32, 48])
def test_kb_gridding_reproduces_fp64_nudft(image):
    """The torchkbnufft restatement (phase, shift, apodization, axis order) must
    reproduce the exact NUDFT to width-6 KB accuracy; any convention error is O(1)."""

    traj = _radial(20, 2 * image)
    omega = (traj.T * 2.0 * np.pi).astype(np.float32)
    k = 2 * image
    cols, re, im = tkbn.kb_interpolation_f32(omega, (image, image), (k, k))
    rng = np.random.default_rng(2)
    x = rng.standard_normal((image, image)) + 1j * rng.standard_normal((image, image))
    smaps = (rng.standard_normal((2, image, image)) + 1j * rng.standard_normal((2, image, image))) / 2
    apod = tkbn.apodization((image, image), (k, k))
    op = solver.KbOperator(cols, re, im, smaps, apod, (image, image), (k, k))
    y = op.forward(x)
    ref = nudft.nudft_forward(x, smaps, omega, scale=nudft.native_scale((k, k)))
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 5e-5
    # structure: 36 unique columns per row, all in range
    assert cols.shape[1] == 36 and cols.min() >= 0 and cols.max() < k * k
    assert all(len(np.unique(r)) == 36 for r in cols)
    # the fp32 model is an adjoint pair up to its float32 panel rounding
    ys = rng.standard_normal(y.shape) + 1j * rng.standard_normal(y.shape)
    left, right = np.vdot(op.forward(x), ys), np.vdot(x, op.adjoint(ys))
    assert abs(left - right) / abs(left) < 1e-6


def test_apodization_is_inverse_kb_transform():
    apod = tkbn.apodization((16, 16), (32, 32))
    assert apod.shape == (16, 16) and np.all(apod > 0)
    np.testing.assert_allclose(apod, apod.T, rtol=1e-14)
    np.testing.assert_allclose(apod, apod[::-1, ::-1], rtol=1e-14)  # symmetric about (N-1)/2
    # closed form of the 1-D KB transform at f = 0: J sinh(alpha)/(alpha I0(alpha))
    alpha = 2.34 * 6
    ft0 = tkbn.kaiser_bessel_ft(np.array([0.0]), 6, alpha)[0]
    assert abs(ft0 - 6 * np.sinh(alpha) / (alpha * np.i0(alpha))) < 1e-10 * ft0


def test_i0_series_matches_numpy():
    x = np.linspace(0, 15, 101)
    np.testing.assert_allclose(tkbn.i0_series(x), np.i0(x), rtol=1e-14)


def test_split2_components_bitwise():
    rng = np.random.default_rng(5)
    v = rng.standard_normal(1000).astype(np.float32)
    hi, lo = solver.components(v, "split2")
    np.testing.assert_array_equal(hi, v.astype(np.float16).astype(np.float64))
    np.testing.assert_array_equal(lo, (v - v.astype(np.float16).astype(np.float32)).astype(np.float16).astype(np.float64))
    # hi + lo keeps ~22 bits; residuals below the FP16 normal range keep 2^-24 absolute
    assert np.all(np.abs(hi + lo - v) <= 2.5e-7 * np.abs(v) + 2 ** -25)


def test_bf16_rounding_matches_torch_and_definition():
    """bfloat16 round-to-nearest-even of the BF16 variant, checked against
    torch's float32 -> bfloat16 conversion and the halfway/tie cases."""

    torch = pytest.importorskip("torch")
    from oracle import layout
    rng = np.random.default_rng(8)
    v = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 20000),
                        np.array([0.0, -0.0, 1.0, -2.5, 3.4e38, 1e-40, np.inf, -np.inf], np.float32)]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(layout.bf16_bits(v), ref)
    # ties: 1 + 2^-8 lies halfway between bf16 neighbours 1 and 1 + 2^-7 -> even (1.0)
    tie = np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8], np.float32)
    np.testing.assert_array_equal(layout.bf16_decode(layout.bf16_bits(tie)), [1.0, 1 + 2 ** -6])
    hi, lo = solver.components(v[:20000], "bf16split")
    assert np.all(np.abs(hi + lo - v[:20000]) <= 2.0 ** -16 * np.abs(v[:20000]))
