"""CPU tests of the native library: exports, packer parity, pack I/O, error codes.

No GPU compute is launched here (host-only entry points).
"""

import os
import re
import struct
from pathlib import Path

import numpy as np
import pytest

from oracle import layout, tkbn

ROOT = Path(__file__).resolve().parents[1]


def _lib():
    from trajsparsenufft_amd import _lib
    return _lib


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "tsn.h").read_text()
    declared = set(re.findall(r"^\s*(?:[\w\s\*]+?)\b((?:tsn|jkf)_\w+)\s*\(", header, re.M))
    assert declared, "no declarations parsed"
    lib = _lib().lib()
    for name in declared:
        assert hasattr(lib, name), f"libtsn.so does not export {name}"
    assert declared <= set(_lib().exported_symbols())


def _radial(spokes, readout):
    from trajsparsenufft_amd import synth
    return synth.radial(spokes, readout, True)


@pytest.mark.parametrize("image,spokes,readout,values", [(32, 16, 64, "fp16"), (48, 20, 96, "fp16"),
                                                         (32, 16, 64, "bf16")])
def test_native_packer_bitwise_equals_oracle_planner(image, spokes, readout, values):
    from trajsparsenufft_amd import pack
    traj = _radial(spokes, readout)
    k = 2 * image
    kb = pack.KbMatrix.build(traj, (image, image), (k, k), threads=3)
    cols, re_, im_ = kb.arrays()
    coords = np.mod(traj.astype(np.float64) * k + k / 2.0, k)
    real = layout.sparse_rows_from_ell("R", cols, re_, k * k, coords)
    imag = layout.sparse_rows_from_ell("I", cols, im_, k * k, coords)
    views = {"G": (real, imag), "G_T": (layout.transpose_sparse(real, k), layout.transpose_sparse(imag, k))}
    for view, (vr, vi) in views.items():
        ref = layout.pack_components(vr, vi, fmt=values)
        hp = kb.pack(view, threads=2, values=values)
        got = hp.arrays()
        hdr = hp.header
        for comp in ("real", "imag"):
            r = ref[comp]
            assert r["header.bin"] == struct.pack("<Q11I", *[hdr[f] for f in pack.HEADER_FIELDS])
            np.testing.assert_array_equal(r["group_offsets.bin"], got["group_offsets"])
            np.testing.assert_array_equal(r["group_row_ids.bin"].ravel(), got["group_row_ids"])
            np.testing.assert_array_equal(r["tile_col_ids.bin"].ravel(), got["tile_col_ids"])
            np.testing.assert_array_equal(r["tile_a_comp.bin"].view(np.uint16).ravel(), got[comp + "_comp"])
            np.testing.assert_array_equal(r["tile_pair_meta.bin"].ravel(), got[comp + "_meta"])


def test_packer_is_thread_count_independent():
    from trajsparsenufft_amd import pack
    kb = pack.KbMatrix.build(_radial(24, 64), (32, 32), (64, 64))
    a = kb.pack("G_T", threads=1).arrays()
    b = kb.pack("G_T", threads=7).arrays()
    for key in a:
        np.testing.assert_array_equal(a[key], b[key])


def test_kb_coefficients_match_oracle_restatement():
    from trajsparsenufft_amd import pack
    traj = _radial(16, 64)
    kb = pack.KbMatrix.build(traj, (32, 32), (64, 64))
    cols, re_, im_ = kb.arrays()
    omega = (traj.T * 2.0 * np.pi).astype(np.float32)
    oc, ore, oim = tkbn.kb_interpolation_f32(omega, (32, 32), (64, 64))
    np.testing.assert_array_equal(cols, oc)
    np.testing.assert_allclose(re_, ore, rtol=0, atol=1e-7)
    np.testing.assert_allclose(im_, oim, rtol=0, atol=1e-7)
    assert np.mean(re_ == ore) > 0.999 and np.mean(im_ == oim) > 0.99
    apod = pack.apodization((32, 32), (64, 64))
    np.testing.assert_allclose(apod, tkbn.apodization((32, 32), (64, 64)), rtol=2e-7)


def test_pack_directory_round_trip_and_reader_checks(tmp_path):
    from trajsparsenufft_amd import pack
    from trajsparsenufft_amd._lib import TsnError
    kb = pack.KbMatrix.build(_radial(12, 64), (32, 32), (64, 64))
    hp = kb.pack("G")
    real, imag = tmp_path / "G" / "real", tmp_path / "G" / "imag"
    hp.write(real, imag)
    files = sorted(os.listdir(real))
    assert files == sorted(["header.bin", "group_offsets.bin", "group_row_ids.bin", "tile_col_ids.bin",
                            "tile_a_comp.bin", "tile_pair_meta.bin"])
    back = pack.HostPack.read(real, imag)
    a, b = hp.arrays(), back.arrays()
    for key in a:
        np.testing.assert_array_equal(a[key], b[key])
    assert back.header == hp.header
    header = (real / "header.bin").read_bytes()
    # corrupt magic -> status 1 (reference exit 1)
    (real / "header.bin").write_bytes(b"\0" * 8 + header[8:])
    with pytest.raises(TsnError) as err:
        pack.HostPack.read(real, imag)
    assert err.value.code == 1
    (real / "header.bin").write_bytes(header)
    # truncated array -> status 1
    arr = (real / "tile_col_ids.bin").read_bytes()
    (real / "tile_col_ids.bin").write_bytes(arr[:-4])
    with pytest.raises(TsnError) as err:
        pack.HostPack.read(real, imag)
    assert err.value.code == 1
    (real / "tile_col_ids.bin").write_bytes(arr)
    # differing real/imag structure -> status 2 (reference exit 2)
    rows = np.fromfile(imag / "group_row_ids.bin", dtype=np.int32)
    rows[[0, 1]] = rows[[1, 0]]
    rows.tofile(imag / "group_row_ids.bin")
    with pytest.raises(TsnError) as err:
        pack.HostPack.read(real, imag)
    assert err.value.code == 2
    # residual_nnz != 0 -> status 2
    fields = list(struct.unpack("<Q11I", header))
    fields[9] = 5
    for d in (real, imag):
        (d / "header.bin").write_bytes(struct.pack("<Q11I", *fields))
    with pytest.raises(TsnError) as err:
        pack.HostPack.read(real, imag)
    assert err.value.code == 2


def test_out_of_range_indices_are_rejected():
    from trajsparsenufft_amd import pack
    from trajsparsenufft_amd._lib import TsnError
    kb = pack.KbMatrix.build(_radial(12, 64), (32, 32), (64, 64))
    hp = kb.pack("G")
    arrays, header = hp.arrays(), hp.header
    bad = dict(arrays)
    bad["tile_col_ids"] = arrays["tile_col_ids"].copy()
    bad["tile_col_ids"][3] = header["n_cols"]
    with pytest.raises(TsnError):
        pack.HostPack.from_arrays(header, bad)
    bad = dict(arrays)
    bad["group_offsets"] = arrays["group_offsets"].copy()
    bad["group_offsets"][-1] += 1
    with pytest.raises(TsnError):
        pack.HostPack.from_arrays(header, bad)
    bad = dict(arrays)
    bad["real_meta"] = np.zeros_like(arrays["real_meta"])  # code 0x0: p0 == p1
    with pytest.raises(TsnError):
        pack.HostPack.from_arrays(header, bad)
    ok = pack.HostPack.from_arrays(header, arrays)
    assert ok.header == header


def test_invalid_trajectory_is_rejected():
    from trajsparsenufft_amd import pack
    from trajsparsenufft_amd._lib import TsnError
    traj = _radial(4, 16)
    traj[3, 0] = np.nan
    with pytest.raises(TsnError) as err:
        pack.KbMatrix.build(traj, (16, 16), (32, 32))
    assert err.value.code == 4


def test_dense_expansion_of_native_pack_round_trips():
    from trajsparsenufft_amd import pack
    kb = pack.KbMatrix.build(_radial(12, 64), (32, 32), (64, 64))
    arrays = kb.pack("G_T").arrays()
    t = len(arrays["tile_col_ids"]) // 32
    comp = arrays["real_comp"].view(np.float16).reshape(t, 16, 16)
    meta = arrays["real_meta"].reshape(t, 16)
    dense = layout.expand_tiles(comp, meta)
    comp2, _ = layout.compress_tiles_vectorized(dense.astype(np.float64))
    np.testing.assert_array_equal(comp2.view(np.uint16), comp.view(np.uint16))


def test_build_flags_kernels_with_scratch_memory():
    """The build parses the compiler's per-kernel resource remarks and names
    every kernel with private (scratch) memory, which it then refuses."""

    from trajsparsenufft_amd.build import scratch_kernels
    remarks = "\n".join([
        "a.hip:10:1: remark: Function Name: _Z4fastv [-Rpass-analysis=kernel-resource-usage]",
        "a.hip:10:1: remark:     VGPRs: 40 [-Rpass-analysis=kernel-resource-usage]",
        "a.hip:10:1: remark:     ScratchSize [bytes/lane]: 0 [-Rpass-analysis=kernel-resource-usage]",
        "a.hip:20:1: remark: Function Name: _Z5spillv [-Rpass-analysis=kernel-resource-usage]",
        "   20 | __global__ void spill() {",
        "a.hip:20:1: remark:     ScratchSize [bytes/lane]: 48 [-Rpass-analysis=kernel-resource-usage]",
    ])
    assert scratch_kernels(remarks) == ["_Z5spillv"]
    assert scratch_kernels("") == []
