"""GPU parity of the interpolation kernels against the float64 oracle.

Reference gates: native vs arithmetic model <= 2e-5 relative
(``validate_fp16x2_native.py:251-252``); split panels bitwise equal to numpy
FP16 round-to-nearest (``:175-216``); fused vs unfused <= 2e-6
(``nufft_integrated_bench.cu:1347-1353``).
"""

import numpy as np
import pytest

from gpu_util import dev, need_gpu, random_panels, rel_l2, small_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 2e-5


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    need_gpu()


def _direction(view):
    from trajsparsenufft_amd.device import Direction
    _, packs = small_case()
    pack = packs.forward if view == "G" else packs.adjoint
    return Direction.upload(pack), pack.arrays(), pack.header


@pytest.mark.parametrize("view,conj", [("G", False), ("G_T", True), ("G", True), ("G_T", False)])
@pytest.mark.parametrize("coils", [8, 4, 12, 16])
def test_sparse_split_matches_oracle(view, conj, coils):
    from oracle.solver import apply_pack
    from trajsparsenufft_amd import device
    d, arrays, header = _direction(view)
    _, hi, res = random_panels(header["n_cols"], coils, seed=coils)
    out = torch.zeros((coils, header["output_size"]), dtype=torch.complex64, device="cuda")
    device.interp(d, dev(hi), dev(res), out, coils=coils, conjugate=conj, scale=0.5)
    ref = apply_pack(arrays, header["output_size"], hi, res, coils, conj, 0.5)
    assert rel_l2(out.cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize("view", ["G", "G_T"])
def test_layouts_accumulate_dense_and_hi_only(view):
    from oracle.solver import apply_pack
    from trajsparsenufft_amd import device
    coils = 8
    d, arrays, header = _direction(view)
    n_out = header["output_size"]
    _, hi, res = random_panels(header["n_cols"], coils, seed=5)
    ref = apply_pack(arrays, n_out, hi, res, coils, view == "G_T")
    conj = view == "G_T"
    # row-major output
    out_rm = torch.zeros((n_out, coils), dtype=torch.complex64, device="cuda")
    device.interp(d, dev(hi), dev(res), out_rm, coils=coils, conjugate=conj, layout="row_major")
    assert rel_l2(out_rm.cpu().numpy().T, ref) <= TOL
    # reference two-pass semantics: hi pass stores, residual pass accumulates
    out = torch.zeros((coils, n_out), dtype=torch.complex64, device="cuda")
    device.interp(d, dev(hi), None, out, coils=coils, conjugate=conj)
    ref_hi = apply_pack(arrays, n_out, hi, None, coils, conj)
    assert rel_l2(out.cpu().numpy(), ref_hi) <= TOL
    device.interp(d, dev(res), None, out, coils=coils, conjugate=conj, accumulate=True)
    assert rel_l2(out.cpu().numpy(), ref) <= TOL
    # dense-MFMA control on the same tiles
    out_d = torch.zeros((coils, n_out), dtype=torch.complex64, device="cuda")
    device.interp(d, dev(hi), dev(res), out_d, coils=coils, conjugate=conj, dense=True)
    assert rel_l2(out_d.cpu().numpy(), ref) <= TOL


def test_repeat_launches_are_bitwise_identical():
    from trajsparsenufft_amd import device
    d, _, header = _direction("G_T")
    _, hi, res = random_panels(header["n_cols"], 8, seed=9)
    h, r = dev(hi), dev(res)
    outs = []
    for _ in range(3):
        out = torch.zeros((8, header["output_size"]), dtype=torch.complex64, device="cuda")
        device.interp(d, h, r, out, coils=8, conjugate=True)
        outs.append(out.cpu().numpy().view(np.uint32))
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


@pytest.mark.parametrize("view", ["G", "G_T"])
@pytest.mark.parametrize("coils", [8, 4, 16, 3])
@pytest.mark.parametrize("spokes", [24, 96])
def test_fp32_csr_matches_oracle(view, coils, spokes):
    """FP32 CSR vs the float64 product.  With 96 spokes on a 32^2 image the
    G_T rows near the k-space centre hold up to 687 entries: rows above 64
    entries run as whole-wave chunks, rows above 256 as several chunks summed
    through partial slots."""

    from oracle.solver import apply_csr, apply_csr_adjoint
    from trajsparsenufft_amd import device
    _, packs = small_case(32, spokes)
    cols, re, im = packs.kb.arrays()
    d = device.Direction.csr(packs.kb, view)
    rng = np.random.default_rng(coils)
    x = (rng.standard_normal((d.cols, coils)) + 1j * rng.standard_normal((d.cols, coils))).astype(np.complex64)
    if view == "G":
        ref = 2.0 * apply_csr(cols, re, im, x.astype(np.complex128)).T
    else:
        ref = 2.0 * apply_csr_adjoint(cols, re, im, x.astype(np.complex128), d.rows).T
    out = torch.zeros((coils, d.rows), dtype=torch.complex64, device="cuda")
    device.interp_f32(d, dev(x), out, coils=coils, conjugate=view == "G_T", scale=2.0)
    assert rel_l2(out.cpu().numpy(), ref) <= 1e-6
    # row-major layout, accumulate, and bitwise repeatability
    rm = torch.zeros((d.rows, coils), dtype=torch.complex64, device="cuda")
    device.interp_f32(d, dev(x), rm, coils=coils, conjugate=view == "G_T", scale=2.0, layout="row_major")
    assert np.array_equal(np.ascontiguousarray(rm.cpu().numpy().T).view(np.uint32), out.cpu().numpy().view(np.uint32))
    device.interp_f32(d, dev(x), out, coils=coils, conjugate=view == "G_T", scale=2.0, accumulate=True)
    assert rel_l2(out.cpu().numpy(), 2.0 * ref) <= 1e-6
    again = torch.zeros_like(rm)
    device.interp_f32(d, dev(x), again, coils=coils, conjugate=view == "G_T", scale=2.0, layout="row_major")
    assert np.array_equal(again.cpu().numpy().view(np.uint32), rm.cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("coils", [8, 3, 16])
@pytest.mark.parametrize("layout", ["coil_major", "row_major"])
def test_fp32_csr_writes_every_row_over_garbage(coils, layout):
    """G_T of a sparse radial trajectory has empty rows (grid points no
    sample reaches), which the FP32 kernel covers with runs of empty rows:
    an output first filled with NaN comes back finite, zero on the empty
    rows and equal to the float64 product elsewhere."""

    from oracle.solver import apply_csr_adjoint
    from trajsparsenufft_amd import device
    _, packs = small_case(32, 12)
    cols, re, im = packs.kb.arrays()
    d = device.Direction.csr(packs.kb, "G_T")
    counts = np.bincount(cols.ravel(), minlength=d.rows)
    assert (counts == 0).sum() > 64  # runs of empty rows exist
    rng = np.random.default_rng(coils + 11)
    x = (rng.standard_normal((d.cols, coils)) + 1j * rng.standard_normal((d.cols, coils))).astype(np.complex64)
    ref = apply_csr_adjoint(cols, re, im, x.astype(np.complex128), d.rows).T  # [coil][row]
    shape = (coils, d.rows) if layout == "coil_major" else (d.rows, coils)
    out = torch.full(shape, float("nan"), dtype=torch.complex64, device="cuda")
    device.interp_f32(d, dev(x), out, coils=coils, conjugate=True, scale=1.0, layout=layout)
    got = out.cpu().numpy()
    got = got if layout == "coil_major" else np.ascontiguousarray(got.T)
    assert np.isfinite(got.view(np.float32)).all()
    assert np.array_equal(got[:, counts == 0], np.zeros_like(got[:, counts == 0]))
    assert rel_l2(got, ref) <= 1e-6


def test_packed_adjoint_writes_only_group_rows():
    """The packed G^T writes the rows of the pack's groups (the grid points
    some sample reaches) and leaves every other output row as it was, the
    pack format's convention (plans zero their grid once)."""

    from oracle.solver import apply_pack
    from trajsparsenufft_amd import device
    coils = 8
    _, packs = small_case(32, 12)
    pack = packs.adjoint
    arrays, header = pack.arrays(), pack.header
    d = device.Direction.upload(pack)
    n_out = header["output_size"]
    reached = np.zeros(n_out, dtype=bool)
    rows = np.asarray(arrays["group_row_ids"])
    reached[rows[rows >= 0]] = True
    assert 64 < (~reached).sum() < n_out  # a sparse trajectory leaves grid points unreached
    _, hi, res = random_panels(header["n_cols"], coils, seed=21)
    sentinel = np.complex64(3.0 - 7.0j)
    out = torch.full((coils, n_out), complex(sentinel), dtype=torch.complex64, device="cuda")
    device.interp(d, dev(hi), dev(res), out, coils=coils, conjugate=True)
    got = out.cpu().numpy()
    assert np.array_equal(got[:, ~reached], np.full_like(got[:, ~reached], sentinel))
    ref = apply_pack(arrays, n_out, hi, res, coils, True)
    assert rel_l2(got[:, reached], ref[:, reached]) <= TOL


@pytest.mark.parametrize("coils,rows", [(12, 1000), (5, 130), (16, 64), (3, 1)])
def test_split_panel_is_bitwise_numpy_rounding(coils, rows):
    """Both split kernels (separate hi/residual panels and the interleaved
    hi | res << 16 words, 64-row tiles) round like numpy, ragged row counts and
    partial coil blocks included."""

    from trajsparsenufft_amd import device
    rng = np.random.default_rng(3)
    v = (rng.standard_normal((coils, rows)) + 1j * rng.standard_normal((coils, rows))).astype(np.complex64)
    special = [0, 1e-9, -1e-9, 65504, 1e-5, 3.14159, -0.0, 6e-8, 1e4, 2.5e-3]
    v[0, :min(rows, 10)] = special[:min(rows, 10)]
    hi, res = device.split_panel(dev(v), coils, scale=0.25)
    hi, res = hi.cpu().numpy(), res.cpu().numpy()
    il = device.split_panel_il(dev(v), coils, scale=0.25).cpu().numpy().view(np.uint32)
    assert np.array_equal(il & 0xFFFF, hi.view(np.uint16)) and np.array_equal(il >> 16, res.view(np.uint16))
    s = v * np.float32(0.25)
    for b in range((coils + 7) // 8):
        for j in range(8):
            c = 8 * b + j
            for part, ch in ((s.real, j), (s.imag, j + 8)):
                if c < coils:
                    x = part[c].astype(np.float32)
                    h = x.astype(np.float16)
                    r = (x - h.astype(np.float32)).astype(np.float16)
                else:
                    h = r = np.zeros(rows, np.float16)
                assert np.array_equal(hi[b, :, ch].view(np.uint16), h.view(np.uint16))
                assert np.array_equal(res[b, :, ch].view(np.uint16), r.view(np.uint16))


def test_adjointness_of_packed_pair():
    """<G x, y> = <x, G^H y> through the two independent packs (FP16 tiles)."""

    from trajsparsenufft_amd import device
    from trajsparsenufft_amd.device import Direction
    _, packs = small_case()
    fwd, adj = Direction.upload(packs.forward), Direction.upload(packs.adjoint)
    rng = np.random.default_rng(11)
    _, xh, xr = random_panels(fwd.cols, 8, seed=1)
    _, yh, yr = random_panels(fwd.rows, 8, seed=2)
    gx = torch.zeros((8, fwd.rows), dtype=torch.complex64, device="cuda")
    ghy = torch.zeros((8, adj.rows), dtype=torch.complex64, device="cuda")
    device.interp(fwd, dev(xh), dev(xr), gx, coils=8)
    device.interp(adj, dev(yh), dev(yr), ghy, coils=8, conjugate=True)
    x = (xh.astype(np.float64) + xr.astype(np.float64))[0]
    y = (yh.astype(np.float64) + yr.astype(np.float64))[0]
    xc = x[:, :8] + 1j * x[:, 8:]
    yc = y[:, :8] + 1j * y[:, 8:]
    left = np.vdot(gx.cpu().numpy().T.astype(np.complex128), yc)
    right = np.vdot(xc, ghy.cpu().numpy().T.astype(np.complex128))
    assert abs(left - right) / max(abs(left), abs(right)) <= 1e-5
    _ = rng


@pytest.mark.parametrize("coils", [8, 16])
def test_long_groups_split_and_reduced_in_order(coils):
    """Dense radial: G_T groups near the k-space centre exceed 16 tiles, so they
    are split over several waves and reduced by the combine kernel."""

    from oracle.solver import apply_pack
    from trajsparsenufft_amd import device
    from trajsparsenufft_amd.device import Direction
    _, packs = small_case(32, 96)
    arrays, header = packs.adjoint.arrays(), packs.adjoint.header
    assert np.diff(arrays["group_offsets"]).max() > 16
    d = Direction.upload(packs.adjoint)
    _, hi, res = random_panels(header["n_cols"], coils, seed=21)
    out = torch.zeros((coils, header["output_size"]), dtype=torch.complex64, device="cuda")
    device.interp(d, dev(hi), dev(res), out, coils=coils, conjugate=True)
    ref = apply_pack(arrays, header["output_size"], hi, res, coils, True)
    assert rel_l2(out.cpu().numpy(), ref) <= TOL
    out2 = torch.zeros_like(out)
    device.interp(d, dev(hi), dev(res), out2, coils=coils, conjugate=True)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), out2.cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("view", ["G", "G_T"])
def test_schedule_cut_is_bitwise_invariant(view):
    """Re-cutting the work units (runs of whole groups per wave) changes only
    which wave streams a group, never the order of its tile sums."""

    from trajsparsenufft_amd import device
    d, _, header = _direction(view)
    _, hi, res = random_panels(header["n_cols"], 12, seed=31)
    h, r = dev(hi), dev(res)
    outs = []
    for unit in (0, 1, 3, 64):
        d.schedule(unit)
        assert unit == 0 or d.unit_tiles == min(unit, 16)  # units hold at most 16 tiles
        out = torch.zeros((12, header["output_size"]), dtype=torch.complex64, device="cuda")
        device.interp(d, h, r, out, coils=12, conjugate=view == "G_T")
        outs.append(out.cpu().numpy().view(np.uint32))
    d.schedule(0)
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


@pytest.mark.parametrize("view", ["G", "G_T"])
@pytest.mark.parametrize("coils", [8, 12])
def test_interleaved_panel_path_is_bitwise_the_split_path(view, coils):
    """tsn_split_panel_il rounds exactly like tsn_split_panel, and
    tsn_interp_il sums the same products in the same order as tsn_interp."""

    from trajsparsenufft_amd import device
    d, _, header = _direction(view)
    rng = np.random.default_rng(41)
    v = torch.from_numpy((rng.standard_normal((coils, header["n_cols"]))
                          + 1j * rng.standard_normal((coils, header["n_cols"]))).astype(np.complex64)).cuda()
    hi, res = device.split_panel(v, coils, scale=0.5)
    il = device.split_panel_il(v, coils, scale=0.5)
    ilw = il.cpu().numpy().view(np.uint32)
    assert np.array_equal(ilw & 0xFFFF, hi.cpu().numpy().view(np.uint16))
    assert np.array_equal(ilw >> 16, res.cpu().numpy().view(np.uint16))
    conj = view == "G_T"
    out = torch.zeros((coils, header["output_size"]), dtype=torch.complex64, device="cuda")
    out_il = torch.zeros_like(out)
    device.interp(d, hi, res, out, coils=coils, conjugate=conj, scale=0.25)
    device.interp_il(d, il, out_il, coils=coils, conjugate=conj, scale=0.25)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), out_il.cpu().numpy().view(np.uint32))
    device.interp_il(d, il, out_il, coils=coils, conjugate=conj, scale=0.25, accumulate=True)
    np.testing.assert_allclose(out_il.cpu().numpy(), 2 * out.cpu().numpy(), rtol=1e-6, atol=1e-6)
    # split-panel output (the plan's forward): the split of the float2 result, bit for bit
    blocks = (coils + 7) // 8
    out_sp = torch.zeros((blocks, header["output_size"], 16), dtype=torch.int32, device="cuda")
    device.interp_il(d, il, out_sp, coils=coils, conjugate=conj, scale=0.25, layout="split_il")
    ref_sp = device.split_panel_il(out.contiguous(), coils)
    assert np.array_equal(out_sp.cpu().numpy(), ref_sp.cpu().numpy())


@pytest.mark.parametrize("coils", [12, 16])
@pytest.mark.parametrize("layout", ["row_major", "split_il", "coil_major"])
def test_coil_waves_are_bitwise_invariant(coils, layout):
    """9..16 coils on interleaved panels: one wave for both coil blocks and
    one wave per block give the same bits, long groups (chunks reduced through
    slots by the last of both waves) included, and match the float64 product."""

    from oracle.solver import apply_pack
    from trajsparsenufft_amd import device
    from trajsparsenufft_amd.device import Direction
    _, packs = small_case(32, 96)
    arrays, header = packs.adjoint.arrays(), packs.adjoint.header
    assert np.diff(arrays["group_offsets"]).max() > 16
    d = Direction.upload(packs.adjoint)
    rng = np.random.default_rng(coils)
    v = (rng.standard_normal((coils, header["n_cols"])) + 1j * rng.standard_normal((coils, header["n_cols"])))
    il = device.split_panel_il(torch.from_numpy(v.astype(np.complex64)).cuda(), coils)
    rows, blocks = header["output_size"], (coils + 7) // 8
    outs = []
    for waves in (1, 2, 1):
        d.coil_waves = waves
        assert d.coil_waves == waves
        if layout == "split_il":
            out = torch.zeros((blocks, rows, 16), dtype=torch.int32, device="cuda")
        else:
            shape = (rows, coils) if layout == "row_major" else (coils, rows)
            out = torch.zeros(shape, dtype=torch.complex64, device="cuda")
        device.interp_il(d, il, out, coils=coils, conjugate=True, scale=0.5, layout=layout)
        outs.append(out.cpu().numpy())
    assert all(np.array_equal(outs[0].view(np.uint32), o.view(np.uint32)) for o in outs[1:])
    if layout != "split_il":
        w = il.cpu().numpy().view(np.uint32)
        hi = (w & 0xFFFF).astype(np.uint16).view(np.float16)
        res = (w >> 16).astype(np.uint16).view(np.float16)
        ref = apply_pack(arrays, rows, hi, res, coils, True, 0.5)
        got = outs[0].T if layout == "row_major" else outs[0]
        assert rel_l2(got, ref) <= TOL


@pytest.mark.parametrize("coils", [8, 16])
@pytest.mark.parametrize("role", ["forward", "adjoint"])
def test_pipe_depths_are_bitwise_invariant(coils, role):
    """Every pipeline depth (the default and the autotuner's alternatives) of
    the plan-style launches gives the same bits: forward into an interleaved
    split panel, adjoint (conjugate) into float2 rows, long groups included;
    the float2 rows also match the float64 product.  Depths outside 0..3 are
    refused."""

    from oracle.solver import apply_pack
    from trajsparsenufft_amd import device
    from trajsparsenufft_amd._lib import TsnError
    from trajsparsenufft_amd.device import Direction
    _, packs = small_case(32, 96)
    pk = packs.forward if role == "forward" else packs.adjoint
    arrays, header = pk.arrays(), pk.header
    d = Direction.upload(pk)
    rng = np.random.default_rng(coils + 7)
    v = (rng.standard_normal((coils, header["n_cols"])) + 1j * rng.standard_normal((coils, header["n_cols"])))
    il = device.split_panel_il(torch.from_numpy(v.astype(np.complex64)).cuda(), coils)
    rows, blocks = header["output_size"], (coils + 7) // 8
    conj = role == "adjoint"
    outs = []
    for depth in (0, 1, 2, 3, 0):
        d.pipe_depth = depth
        assert d.pipe_depth == depth
        if conj:
            out = torch.zeros((rows, coils), dtype=torch.complex64, device="cuda")
            device.interp_il(d, il, out, coils=coils, conjugate=True, scale=0.5, layout="row_major")
        else:
            out = torch.zeros((blocks, rows, 16), dtype=torch.int32, device="cuda")
            device.interp_il(d, il, out, coils=coils, conjugate=False, scale=0.5, layout="split_il")
        outs.append(out.cpu().numpy())
    assert all(np.array_equal(outs[0].view(np.uint32), o.view(np.uint32)) for o in outs[1:])
    if conj:
        w = il.cpu().numpy().view(np.uint32)
        hi = (w & 0xFFFF).astype(np.uint16).view(np.float16)
        res = (w >> 16).astype(np.uint16).view(np.float16)
        ref = apply_pack(arrays, rows, hi, res, coils, True, 0.5)
        assert rel_l2(outs[0].T, ref) <= TOL
    with pytest.raises(TsnError):
        d.pipe_depth = 4


@pytest.mark.parametrize("view", ["G", "G_T"])
@pytest.mark.parametrize("coils", [8, 16])
def test_bf16_tiles_and_panels_match_oracle(view, coils):
    """BF16 variant: BF16 tiles x BF16 hi + residual panels on the bf16 smfmac,
    against the float64 product of the same BF16 values; the split panel is
    bitwise the round-to-nearest-even split."""

    from oracle import layout
    from oracle.solver import apply_pack
    from trajsparsenufft_amd import device
    from trajsparsenufft_amd.device import Direction
    _, packs = small_case()
    bp = packs.with_values("bf16")
    hp = bp.forward if view == "G" else bp.adjoint
    d = Direction.upload(hp)
    header, arrays = hp.header, hp.arrays()
    rng = np.random.default_rng(coils)
    v = (rng.standard_normal((coils, header["n_cols"])) + 1j * rng.standard_normal((coils, header["n_cols"]))
         ).astype(np.complex64)
    il = device.split_panel_il(dev(v), coils, scale=0.5, values="bf16").cpu().numpy().view(np.uint32)
    nb = (coils + 7) // 8
    chan = np.zeros((nb, header["n_cols"], 16), np.float32)
    for c in range(coils):
        chan[c // 8, :, c % 8] = (v[c] * np.float32(0.5)).real
        chan[c // 8, :, c % 8 + 8] = (v[c] * np.float32(0.5)).imag
    hb = layout.bf16_bits(chan)
    lb = layout.bf16_bits(chan - layout.bf16_decode(hb))
    assert np.array_equal(il & 0xFFFF, hb) and np.array_equal(il >> 16, lb)
    conj = view == "G_T"
    out = torch.zeros((coils, header["output_size"]), dtype=torch.complex64, device="cuda")
    device.interp_il(d, torch.from_numpy(il.view(np.int32)).cuda(), out, coils=coils, conjugate=conj)
    ref = apply_pack(arrays, header["output_size"], layout.bf16_decode(hb), layout.bf16_decode(lb), coils, conj,
                     values="bf16")
    assert rel_l2(out.cpu().numpy(), ref) <= TOL
    # a BF16 direction refuses the FP16 entry points, and an FP16 one the BF16 panel path stays FP16
    from trajsparsenufft_amd._lib import TsnError
    with pytest.raises(TsnError):
        device.interp(d, dev(np.zeros((nb, header["n_cols"], 16), np.float16)),
                      dev(np.zeros((nb, header["n_cols"], 16), np.float16)), out, coils=coils)


def test_autotune_keeps_results_and_picks_a_candidate():
    from trajsparsenufft_amd import device
    d, _, header = _direction("G")
    _, hi, res = random_panels(header["n_cols"], 8, seed=51)
    out0 = torch.zeros((8, header["output_size"]), dtype=torch.complex64, device="cuda")
    device.interp(d, dev(hi), dev(res), out0, coils=8)
    best, us = d.autotune(8)
    assert best in (2, 3, 4, 5, 6, 8, 10, 12, 14, 16) and d.unit_tiles == best and us > 0
    out1 = torch.zeros_like(out0)
    device.interp(d, dev(hi), dev(res), out1, coils=8)
    assert np.array_equal(out0.cpu().numpy().view(np.uint32), out1.cpu().numpy().view(np.uint32))
