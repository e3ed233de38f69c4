"""tsn_interp_noalloc: tsn_interp with its long-group scratch in a static
device array (no allocation, no direction-held scratch).

It runs the same kernels as tsn_interp, so its output must equal tsn_interp's
bit for bit, and both must match the float64 product of the same FP16 tiles
and panels within 2e-5 (the reference's gate, validate_fp16x2_native.py:251-252).
Two calls in a row, the second on a smaller direction, check that nothing of
the first call's scratch leaks into the second.
"""

import ctypes as C

import numpy as np
import pytest

from gpu_util import dev, need_gpu, random_panels, rel_l2, small_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 2e-5


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    need_gpu()


def _noalloc(d, hi, res, out, coils, conj):
    from trajsparsenufft_amd._lib import lib
    return lib().tsn_interp_noalloc(d.handle, C.c_void_p(hi.data_ptr()), C.c_void_p(res.data_ptr()),
                                    C.c_void_p(out.data_ptr()), coils, int(conj), 0, C.c_float(0.75), 0,
                                    C.c_void_p(torch.cuda.current_stream().cuda_stream))


def _case(image, spokes, coils, seed):
    from oracle.solver import apply_pack
    from trajsparsenufft_amd import device
    from trajsparsenufft_amd._lib import check
    from trajsparsenufft_amd.device import Direction
    _, packs = small_case(image, spokes)
    arrays, header = packs.adjoint.arrays(), packs.adjoint.header
    assert np.diff(arrays["group_offsets"]).max() > 16  # long groups: the chunk scratch is used
    d = Direction.upload(packs.adjoint)
    _, hi, res = random_panels(header["n_cols"], coils, seed=seed)
    h, r = dev(hi), dev(res)
    rows = header["output_size"]
    want = torch.zeros((coils, rows), dtype=torch.complex64, device="cuda")
    device.interp(d, h, r, want, coils=coils, conjugate=True, scale=0.75)
    # rows absent from the pack are never written (zeroed beforehand, as the
    # reference requires of its output): both outputs start at zero
    got = torch.zeros((coils, rows), dtype=torch.complex64, device="cuda")
    print(f"tsn_interp_noalloc: image {image}, spokes {spokes}, coils {coils}", flush=True)
    check(_noalloc(d, h, r, got, coils, True))
    torch.cuda.synchronize()
    ref = apply_pack(arrays, rows, hi, res, coils, True, 0.75)
    got_np, want_np = got.cpu().numpy(), want.cpu().numpy()
    assert np.array_equal(got_np.view(np.uint32), want_np.view(np.uint32))
    assert rel_l2(got_np, ref) <= TOL
    # a second call on the same direction: the static counters were left at zero
    check(_noalloc(d, h, r, got, coils, True))
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy().view(np.uint32), want_np.view(np.uint32))
    return d


def test_noalloc_matches_tsn_interp_on_two_calls():
    first = _case(32, 96, 16, seed=41)   # 4,096 rows, 30 chunks, two coil blocks
    second = _case(16, 48, 12, seed=42)  # 1,024 rows, 12 chunks: smaller, after the first
    del first, second


def test_noalloc_rejects_directions_beyond_the_static_scratch():
    """One group of 1,537 x 16 tiles needs 1,537 chunk slots, one more than the
    static array holds: the call fails with TSN_ERR_ARGUMENT before any launch."""

    from trajsparsenufft_amd import pack
    from trajsparsenufft_amd._lib import TSN_ERR_ARGUMENT, lib
    from trajsparsenufft_amd.device import Direction
    tiles = 1537 * 16
    header = dict(magic=0x4E55464654323431, version=2, n_channels=16, output_size=16, n_cols=64,
                  active_rows=16, nnz=tiles * 16 * 16, groups=1, tiles=tiles, residual_nnz=0, m_tile=16,
                  k_tile=32)
    meta = np.full(tiles * 16, 0x44444444, dtype=np.uint32)  # every 4-bit code keeps positions 0 and 1
    arrays = dict(group_offsets=np.array([0, tiles], dtype=np.uint32),
                  group_row_ids=np.arange(16, dtype=np.int32),
                  tile_col_ids=np.tile(np.arange(32, dtype=np.int32), tiles),
                  real_comp=np.zeros(tiles * 256, dtype=np.uint16), real_meta=meta,
                  imag_comp=np.zeros(tiles * 256, dtype=np.uint16), imag_meta=meta)
    d = Direction.upload(pack.HostPack.from_arrays(header, arrays))
    hi = torch.zeros((1, 64, 16), dtype=torch.float16, device="cuda")
    out = torch.zeros((8, 16), dtype=torch.complex64, device="cuda")
    rc = _noalloc(d, hi, hi, out, 8, False)
    assert rc == TSN_ERR_ARGUMENT
    assert b"static scratch" in lib().tsn_last_error()
