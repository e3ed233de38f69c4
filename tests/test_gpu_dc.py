"""Data consistency (reference ``JKF_DATA_CONSISTENCY_MODE`` 1 and 2):
r = (scale * G panel - measurement_scale * y) * density as the adjoint's split
panel, and the plan-level gradient A^H W (A x - y).

Gates: the fused epilogue equals the materialized pass bit for bit (the
reference's own check, ``nufft_integrated_bench.cu:1184-1204``, counts panel
bit mismatches); the residual matches the float64 product of the same tiles
and panels (2e-5, the interpolation gate); the plan gradient matches the
arithmetic model ``oracle.solver.KbOperator`` and the identity
A^H (A x - y) = A^H A x - A^H y.
"""

import numpy as np
import pytest

from gpu_util import dev, need_gpu, rel_l2, small_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    need_gpu()


def _forward_direction(values):
    from trajsparsenufft_amd.device import Direction
    _, packs = small_case()
    pack = packs.forward if values == "fp16" else packs.with_values("bf16").forward
    return Direction.upload(pack), pack


def _inputs(n_cols, rows, coils, seed):
    rng = np.random.default_rng(seed)
    grid = (rng.standard_normal((coils, n_cols)) + 1j * rng.standard_normal((coils, n_cols))).astype(np.complex64)
    y = (rng.standard_normal((coils, rows)) + 1j * rng.standard_normal((coils, rows))).astype(np.complex64)
    w = (0.75 + rng.random(rows) * 0.5).astype(np.float32)
    return grid, y, w


def _decode(panel, values):
    """interleaved int32 [blocks, rows, 16] -> (hi, res) float32 panels."""

    from oracle.layout import bf16_decode
    bits = panel.cpu().numpy().view(np.uint32)
    lo, hi = (bits & 0xFFFF).astype(np.uint16), (bits >> 16).astype(np.uint16)
    if values == "bf16":
        return bf16_decode(lo).astype(np.float32), bf16_decode(hi).astype(np.float32)
    return lo.view(np.float16).astype(np.float32), hi.view(np.float16).astype(np.float32)


def _panel_values(panel, coils, values):
    """split panel -> complex128 [coils, rows] of hi + res."""

    h, r = _decode(panel, values)
    v = h.astype(np.float64) + r
    out = np.zeros((coils, v.shape[1]), np.complex128)
    for c in range(coils):
        b, j = divmod(c, 8)
        out[c] = v[b, :, j] + 1j * v[b, :, j + 8]
    return out


@pytest.mark.parametrize("values", ["fp16", "bf16"])
@pytest.mark.parametrize("coils", [8, 4, 12, 16])
def test_fused_equals_materialized_and_float64(values, coils):
    from oracle.solver import apply_pack
    from trajsparsenufft_amd import device
    d, pack = _forward_direction(values)
    rows, n_cols = d.rows, d.cols
    grid, y, w = _inputs(n_cols, rows, coils, seed=coils)
    panel = device.split_panel_il(dev(grid), coils, values=values)
    scale, mscale = 0.03125, 0.5
    yd, wd = dev(y), dev(w)
    y_rows = dev(np.ascontiguousarray(y.T))
    runs = {}
    for mode in ("materialized", "fused"):
        for layout, meas in (("coil_major", yd), ("row_major", y_rows)):
            for dens in (wd, None):
                key = (mode, layout, dens is None)
                runs[key] = device.interp_dc(d, panel, meas, coils, density=dens, scale=scale,
                                             measurement_scale=mscale, mode=mode, layout=layout)
    torch.cuda.synchronize()
    base = {k: v.cpu().numpy() for k, v in runs.items()}
    for (mode, layout, ones), bits in base.items():
        ref = base[("materialized", "coil_major", ones)]
        assert np.array_equal(bits, ref), f"{mode} {layout} density={'ones' if ones else 'given'}"
    # float64: the same tiles times hi + res of the input panel
    ph, pr = _decode(panel, values)
    pred = apply_pack(pack.arrays(), rows, ph, pr, coils, False, scale, values=values)
    for ones in (False, True):
        wt = np.ones(rows) if ones else w.astype(np.float64)
        expect = (pred - mscale * y.astype(np.complex128)) * wt[None, :]
        got = _panel_values(runs[("fused", "coil_major", ones)], coils, values)
        assert rel_l2(got, expect) <= 2e-5


def test_dc_argument_errors():
    import ctypes as C
    from trajsparsenufft_amd import device
    from trajsparsenufft_amd._lib import TsnError, check, lib
    d, _ = _forward_direction("fp16")
    grid, y, _ = _inputs(d.cols, d.rows, 8, seed=1)
    panel = device.split_panel_il(dev(grid), 8)
    out = torch.empty((1, d.rows, 16), dtype=torch.int32, device="cuda")
    yd = dev(y)
    p, m, o = (C.c_void_p(t.data_ptr()) for t in (panel, yd, out))
    for mode, layout in ((1, 0), (3, 0), (2, 7)):  # materialized without workspace, unknown mode, unknown layout
        with pytest.raises(TsnError) as err:
            check(lib().tsn_interp_dc(d.handle, p, m, layout, None, 8, 1.0, 1.0, mode, None, o, None))
        assert err.value.code == 4


def _model(packs, smaps, precision):
    from oracle.solver import KbOperator
    cols, re, im = packs.kb.arrays()
    modes = {"fp32": ("fp32", "fp32"), "split16": ("fp16", "split2"), "dense16": ("fp16", "split2"),
             "hi16": ("fp16", "fp16"), "splitbf16": ("bf16", "bf16split")}[precision]
    return KbOperator(cols, re, im, smaps, packs.apod, packs.kb.image_size, packs.kb.grid_size, *modes)


@pytest.mark.parametrize("precision", ["fp32", "split16", "splitbf16", "dense16", "hi16"])
@pytest.mark.parametrize("coils", [8, 12])
def test_plan_gradient_matches_model(precision, coils):
    """A^H W (A x - s y) vs the model with the same roundings.  The residual
    is rounded to FP32 before its split here and kept in float64 by the
    model; one-16-bit-rounding variants (hi16, BF16) see that as 1e-4."""

    from trajsparsenufft_amd import synth
    from trajsparsenufft_amd.recon import build_plan
    _, packs = small_case()
    x = synth.phantom(32, 32, seed=coils)
    smaps = synth.coil_maps(coils, 32)
    plan = build_plan(packs, dev(smaps), precision)
    rng = np.random.default_rng(coils)
    m = plan.samples
    y = (rng.standard_normal((coils, m)) + 1j * rng.standard_normal((coils, m))).astype(np.complex64)
    w = (0.5 + rng.random(m)).astype(np.float32)
    ms = 0.25
    model = _model(packs.with_values("bf16") if precision == "splitbf16" else packs, smaps, precision)
    tol = 1e-4 if precision in ("splitbf16", "hi16") else 2e-5
    for wt in (w, None):
        got = plan.data_consistency(dev(x), dev(y), None if wt is None else dev(wt), ms).cpu().numpy()
        r = model.forward(x) - ms * y.astype(np.complex128)
        if wt is not None:
            r = r * wt.astype(np.float64)[None, :]
        assert rel_l2(got, model.adjoint(r)) <= tol


@pytest.mark.parametrize("precision", ["fp32", "split16"])
def test_plan_gradient_identity(precision):
    """With W = I: A^H (A x - s y) = A^H A x - A^H (s y), operator by operator."""

    from trajsparsenufft_amd import synth
    from trajsparsenufft_amd.recon import build_plan
    _, packs = small_case()
    x = synth.phantom(32, 32, seed=2)
    smaps = synth.coil_maps(8, 32)
    plan = build_plan(packs, dev(smaps), precision)
    rng = np.random.default_rng(3)
    y = (rng.standard_normal((8, plan.samples)) + 1j * rng.standard_normal((8, plan.samples))).astype(np.complex64)
    ms = 1.0 / 64.0  # 1/sqrt(K0 K1): the default scale for raw data
    g = plan.data_consistency(dev(x), dev(y)).cpu().numpy().astype(np.complex128)
    ref = plan.normal(dev(x), 1.0, 0.0).cpu().numpy().astype(np.complex128) - \
        plan.adjoint(dev((y * ms).astype(np.complex64))).cpu().numpy()
    assert rel_l2(g, ref) <= (1e-5 if precision == "fp32" else 2e-5)
