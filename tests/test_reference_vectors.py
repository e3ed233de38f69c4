"""The reference's own vectors against this build's host code (CPU only).

* Trajectory generator (``experiments/data_prep/generate_trajectories.py:14-23``):
  ``synth.radial`` against the three arrays the reference's ``radial`` wrote.
* Pack directories the reference's exporter wrote (``src/export_gpu_case.py``)
  read unchanged by ``tsn_host_pack_read`` (the reader of
  ``nufft_integrated_bench.cu:135-226``): header and arrays byte for byte.
* The device format's tile limit is enforced before anything is uploaded.
"""

import numpy as np
import pytest

from refpack import CASES, header_fields, write_reference_dir


@pytest.mark.parametrize("name,spokes,readout,golden", [("radial_golden_8x16", 8, 16, True),
                                                        ("radial_uniform_8x16", 8, 16, False),
                                                        ("radial_golden_5x12", 5, 12, True)])
def test_radial_generator_bitwise_equals_reference_vectors(layout_golden, name, spokes, readout, golden):
    from trajsparsenufft_amd import synth
    got = synth.radial(spokes, readout, golden)
    want = layout_golden[name]
    assert got.dtype == want.dtype == np.float32
    assert got.shape == want.shape
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("case", CASES)
def test_reader_takes_reference_written_directories_unchanged(layout_golden, case, tmp_path):
    from trajsparsenufft_amd.pack import HostPack
    d = write_reference_dir(layout_golden, case, tmp_path / case)
    # the exporter writes one real component; the same directory serves as
    # both components of a complex direction
    pack = HostPack.read(d, d)
    assert pack.header == header_fields(layout_golden, case)
    assert pack.header["residual_nnz"] == 0
    arrays = pack.arrays()
    g = layout_golden
    np.testing.assert_array_equal(arrays["group_offsets"], g[f"{case}_group_offsets"])
    np.testing.assert_array_equal(arrays["group_row_ids"], g[f"{case}_group_row_ids"])
    np.testing.assert_array_equal(arrays["tile_col_ids"], g[f"{case}_tile_col_ids"])
    np.testing.assert_array_equal(arrays["real_comp"], g[f"{case}_tile_a_comp"].view(np.uint16))
    np.testing.assert_array_equal(arrays["imag_comp"], g[f"{case}_tile_a_comp"].view(np.uint16))
    np.testing.assert_array_equal(arrays["real_meta"], g[f"{case}_tile_pair_meta"])
    # and writing it back reproduces the reference's files byte for byte
    out_re, out_im = tmp_path / "w_re", tmp_path / "w_im"
    pack.write(out_re, out_im)
    for name in ("header", "group_offsets", "group_row_ids", "tile_col_ids", "tile_a_comp", "tile_pair_meta"):
        assert (out_re / f"{name}.bin").read_bytes() == (d / f"{name}.bin").read_bytes(), name
        assert (out_im / f"{name}.bin").read_bytes() == (d / f"{name}.bin").read_bytes(), name


def test_upload_rejects_more_tiles_than_the_device_format_holds():
    """Per-tile operand offsets are signed 32-bit: 2^21 sparse tiles at most.
    The check runs on the header before any device call, so it runs here."""

    from trajsparsenufft_amd import pack, synth
    from trajsparsenufft_amd._lib import TsnError
    from trajsparsenufft_amd.device import Direction
    small = pack.KbMatrix.build(synth.radial(8, 32, True), (16, 16), (32, 32)).pack("G")
    h = dict(small.header)
    a = small.arrays()
    tiles = (1 << 21) + 1
    h.update(tiles=tiles, groups=1, active_rows=16)
    meta = np.tile(a["real_meta"][:16], tiles)  # one legal tile's metadata, repeated
    arrays = {"group_offsets": np.array([0, tiles], np.uint32), "group_row_ids": np.zeros(16, np.int32),
              "tile_col_ids": np.zeros(tiles * 32, np.int32), "real_comp": np.zeros(tiles * 256, np.uint16),
              "real_meta": meta, "imag_comp": np.zeros(tiles * 256, np.uint16), "imag_meta": meta}
    big = pack.HostPack.from_arrays(h, arrays)
    del arrays, meta
    with pytest.raises(TsnError) as err:
        Direction.upload(big)
    assert err.value.code == 4
    assert "tiles" in str(err.value)
