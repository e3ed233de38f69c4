"""The production planner and pack writer against the reference's own exporter.

``tests/golden/production_golden.npz`` holds every file the reference's
production exporter (``src/export_torchkbnufft_gpu_case.py``:
``optimized_split24_pack`` :217-256, ``optimized_assign_columns`` /
``optimized_arrange_24`` :102-214, ``materialize_component_tiles`` :345-374,
``compress_tiles_vectorized`` :302-320, ``export_component`` :377-489) wrote
for a small golden-angle radial KB matrix, with that matrix as input (see
``tests/golden/make_production_golden.py``).  The oracle's restatement and the
native C++ packer must reproduce those files byte for byte: header, group
offsets and row ids, tile columns, FP16 values and pair metadata of the real
and the imaginary component.
"""

import json
import struct

import numpy as np
import pytest

from oracle import layout

RUNS = (("G", "morton"), ("G_T", "morton"), ("G_T", "overlap"))
NATIVE_POLICIES = ("morton", "overlap")


@pytest.fixture(scope="module")
def golden():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "production_golden.npz")
    with np.load(path, allow_pickle=False) as data:
        return {k: data[k] for k in data.files}


@pytest.fixture(scope="module")
def case(golden):
    from trajsparsenufft_amd import pack
    spec = json.loads(golden["case"].tobytes().decode())
    k = 2 * spec["image"]
    kb = pack.KbMatrix.build(golden["traj"], (spec["image"],) * 2, (k, k), numpoints=spec["numpoints"])
    return spec, kb


def _file(golden, view, policy, comp, name, dtype):
    return np.frombuffer(golden[f"{view}_{policy}_{comp}_{name}"].tobytes(), dtype=dtype)


def test_inputs_are_this_builds_kb_matrix(golden, case):
    _, kb = case
    cols, re_, im_ = kb.arrays()
    np.testing.assert_array_equal(cols, golden["kb_cols"])
    np.testing.assert_array_equal(re_.view(np.uint32), golden["kb_re"].view(np.uint32))
    np.testing.assert_array_equal(im_.view(np.uint32), golden["kb_im"].view(np.uint32))


def _oracle_views(golden, spec):
    k = 2 * spec["image"]
    coords = np.mod(golden["traj"].astype(np.float64) * k + k / 2.0, k)
    real = layout.sparse_rows_from_ell("R", golden["kb_cols"], golden["kb_re"], k * k, coords)
    imag = layout.sparse_rows_from_ell("I", golden["kb_cols"], golden["kb_im"], k * k, coords)
    return {"G": (real, imag), "G_T": (layout.transpose_sparse(real, k), layout.transpose_sparse(imag, k))}


@pytest.mark.parametrize("view,policy", RUNS)
def test_oracle_restatement_reproduces_reference_exporter(golden, case, view, policy):
    spec, _ = case
    vr, vi = _oracle_views(golden, spec)[view]
    ours = layout.pack_components(vr, vi, policy=policy, seed=spec["seed"])
    for comp in ("real", "imag"):
        o = ours[comp]
        assert o["header.bin"] == golden[f"{view}_{policy}_{comp}_header"].tobytes()
        np.testing.assert_array_equal(o["group_offsets.bin"], _file(golden, view, policy, comp, "group_offsets", np.uint32))
        np.testing.assert_array_equal(o["group_row_ids.bin"].ravel(), _file(golden, view, policy, comp, "group_row_ids", np.int32))
        np.testing.assert_array_equal(o["tile_col_ids.bin"].ravel(), _file(golden, view, policy, comp, "tile_col_ids", np.int32))
        np.testing.assert_array_equal(o["tile_a_comp.bin"].view(np.uint16).ravel(),
                                      _file(golden, view, policy, comp, "tile_a_comp", np.uint16))
        np.testing.assert_array_equal(o["tile_pair_meta.bin"].ravel(), _file(golden, view, policy, comp, "tile_pair_meta", np.uint32))


@pytest.mark.parametrize("threads", [1, 3])
@pytest.mark.parametrize("view,policy", [r for r in RUNS if r[1] in NATIVE_POLICIES])
def test_native_packer_reproduces_reference_exporter(golden, case, view, policy, threads, tmp_path):
    from trajsparsenufft_amd import pack
    spec, kb = case
    hp = kb.pack(view, policy=policy, seed=spec["seed"], threads=threads)
    got = hp.arrays()
    hdr = hp.header
    header = struct.pack("<Q11I", *[hdr[f] for f in pack.HEADER_FIELDS])
    for comp in ("real", "imag"):
        assert header == golden[f"{view}_{policy}_{comp}_header"].tobytes()
        np.testing.assert_array_equal(got["group_offsets"], _file(golden, view, policy, comp, "group_offsets", np.uint32))
        np.testing.assert_array_equal(got["group_row_ids"], _file(golden, view, policy, comp, "group_row_ids", np.int32))
        np.testing.assert_array_equal(got["tile_col_ids"], _file(golden, view, policy, comp, "tile_col_ids", np.int32))
        np.testing.assert_array_equal(got[comp + "_comp"], _file(golden, view, policy, comp, "tile_a_comp", np.uint16))
        np.testing.assert_array_equal(got[comp + "_meta"], _file(golden, view, policy, comp, "tile_pair_meta", np.uint32))
    # and the directories it writes are the reference exporter's files
    hp.write(tmp_path / "re", tmp_path / "im")
    for comp, d in (("real", tmp_path / "re"), ("imag", tmp_path / "im")):
        for name in ("header", "group_offsets", "group_row_ids", "tile_col_ids", "tile_a_comp", "tile_pair_meta"):
            assert (d / f"{name}.bin").read_bytes() == golden[f"{view}_{policy}_{comp}_{name}"].tobytes(), (comp, name)
