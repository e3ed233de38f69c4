"""Pin the layout oracle against the reference's golden vectors and invariants.

Mirrors the reference's structural tests (``tests/test_oracle.py:25-58``,
``tests/test_export_gpu_case.py:15-40``) and adds bit-exact comparisons with
fixtures written from the reference's own modules.
"""

import numpy as np
import pytest

from oracle import layout


def rows_from_csr(ptr, cols, vals):
    return tuple(cols[ptr[i]:ptr[i + 1]] for i in range(len(ptr) - 1)), \
        tuple(vals[ptr[i]:ptr[i + 1]] for i in range(len(ptr) - 1))


@pytest.fixture(scope="module")
def case(layout_golden):
    g = layout_golden
    traj = layout.generate_trajectory("radial", 24, 96)
    fwd = layout.build_interpolation(traj, 24, 4)
    adj = layout.transpose_sparse(fwd, 24)
    return g, traj, fwd, adj


def assert_tiles_equal(tiles, g, prefix):
    assert len(tiles) == len(g[prefix + "_row_ids"])
    np.testing.assert_array_equal(np.stack([t.row_ids for t in tiles]), g[prefix + "_row_ids"])
    np.testing.assert_array_equal(np.stack([t.col_ids for t in tiles]), g[prefix + "_col_ids"])
    np.testing.assert_array_equal(np.stack([t.values for t in tiles]), g[prefix + "_values"])
    np.testing.assert_array_equal(np.asarray([t.legal_24 for t in tiles]), g[prefix + "_legal"])


def test_trajectories_match_reference(case):
    g, traj, _, _ = case
    np.testing.assert_array_equal(traj, g["case_traj"])
    np.testing.assert_array_equal(layout.generate_trajectory("spiral", 24, 192), g["case_spiral_traj"])


def test_interpolation_and_transpose_match_reference(case):
    g, _, fwd, adj = case
    np.testing.assert_array_equal(np.concatenate(fwd.cols), g["fwd_cols"])
    np.testing.assert_array_equal(np.concatenate(fwd.vals), g["fwd_vals"])
    np.testing.assert_array_equal(adj.row_ids, g["adj_row_ids"])
    np.testing.assert_array_equal(np.concatenate(adj.cols), g["adj_cols"])
    np.testing.assert_array_equal(np.concatenate(adj.vals), g["adj_vals"])
    assert layout.sparse_checksum(fwd) == str(g["checksum_fwd"])
    # transpose preserves the coefficient multiset (reference test_oracle.py:25-29)
    assert layout.canonical_operator_checksum(fwd, False) == layout.canonical_operator_checksum(adj, True)


@pytest.mark.parametrize("policy", ["trajectory", "morton", "overlap", "packing256"])
def test_row_order_matches_reference(case, policy):
    g, _, fwd, adj = case
    assert layout.row_order(fwd, policy) == list(g["order_fwd_" + policy])
    assert layout.row_order(adj, policy) == list(g["order_adj_" + policy])


def test_pack_view_matches_reference_bitwise(case):
    g, _, fwd, adj = case
    assert_tiles_equal(layout.pack_view(fwd, "overlap", 32, 7), g, "pack_fwd_overlap_7")
    assert_tiles_equal(layout.pack_view(adj, "overlap", 32, 13), g, "pack_adj_overlap_13")
    assert_tiles_equal(layout.pack_view(fwd, "morton", 32, 11), g, "pack_fwd_morton_11")
    assert_tiles_equal(layout.pack_view(adj, "packing256", 32, 5), g, "pack_adj_packing256_5")


def test_pack_is_lossless_legal_and_adjoint(case):
    _, _, fwd, adj = case
    tiles = layout.pack_view(fwd, "overlap", 32, 7)
    assert layout.sparse_checksum(fwd) == layout.packed_checksum(tiles)
    assert all(layout.validate_24(t.values)[0] for t in tiles)
    rng = np.random.default_rng(9)
    dense = rng.standard_normal((fwd.n_cols, 2)) + 1j * rng.standard_normal((fwd.n_cols, 2))
    np.testing.assert_allclose(layout.apply_sparse(fwd, dense),
                               layout.apply_packed(fwd.output_size, tiles, dense), rtol=1e-13, atol=1e-13)
    tf = layout.pack_view(fwd, "overlap", 32, 11)
    ta = layout.pack_view(adj, "overlap", 32, 13)
    rng = np.random.default_rng(17)
    x = rng.standard_normal((fwd.n_cols, 2)) + 1j * rng.standard_normal((fwd.n_cols, 2))
    y = rng.standard_normal((fwd.output_size, 2)) + 1j * rng.standard_normal((fwd.output_size, 2))
    left = np.vdot(layout.apply_packed(fwd.output_size, tf, x), y)
    right = np.vdot(x, layout.apply_packed(adj.output_size, ta, y))
    assert abs(left - right) / max(abs(left), abs(right)) <= 1e-12


def test_compress_tile_matches_reference(layout_golden):
    g = layout_golden
    for t in range(g["compress_in"].shape[0]):
        comp, meta = layout.compress_tile(g["compress_in"][t])
        np.testing.assert_array_equal(comp.view(np.uint16), g["compress_comp"][t].view(np.uint16))
        np.testing.assert_array_equal(meta, g["compress_meta"][t])
    comp, meta = layout.compress_tiles_vectorized(g["compress_in"])
    np.testing.assert_array_equal(comp.view(np.uint16), g["compress_comp"].view(np.uint16))
    np.testing.assert_array_equal(meta, g["compress_meta"])


@pytest.mark.parametrize("mode", ["csr_residual", "split24"])
def test_export_case_matches_reference(layout_golden, mode):
    """Re-run the reference export (tests/test_export_gpu_case.py:19-33) in the oracle."""

    g = layout_golden
    traj = layout.generate_trajectory("spiral", 24, 192)
    view = layout.transpose_sparse(layout.build_interpolation(traj, 24, 4), 24)
    tiles = layout.pack_view(view, "overlap", 32, 23)
    assert layout.sparse_checksum(view) == layout.packed_checksum(tiles)
    offsets, row_ids, legal, res_ptr, res_cols, res_vals = \
        layout.split_legal_tiles_and_residual(view, tiles, mode)
    p = f"export_{mode}_"
    np.testing.assert_array_equal(offsets, g[p + "group_offsets"])
    np.testing.assert_array_equal(row_ids.ravel(), g[p + "group_row_ids"])
    col_ids = np.stack([t.col_ids for t in legal]).astype(np.int32)
    col_ids[col_ids < 0] = 0
    np.testing.assert_array_equal(col_ids.ravel(), g[p + "tile_col_ids"])
    comp, meta = zip(*(layout.compress_tile(t.values) for t in legal))
    np.testing.assert_array_equal(np.stack(comp).ravel().view(np.uint16),
                                  g[p + "tile_a_comp"].view(np.uint16))
    np.testing.assert_array_equal(np.stack(meta).ravel(), g[p + "tile_pair_meta"])
    np.testing.assert_array_equal(res_ptr, g[p + "residual_row_ptr"])
    np.testing.assert_array_equal(res_cols, g[p + "residual_cols"])
    np.testing.assert_array_equal(res_vals.view(np.uint16), g[p + "residual_vals"].view(np.uint16))
    header = layout.header_bytes(view.output_size, view.n_cols, view.n_active_rows, view.nnz,
                                 len(offsets) - 1, len(legal), len(res_cols))
    assert np.frombuffer(header, dtype=np.uint8).tolist() == g[p + "header"].tolist()
    np.testing.assert_array_equal(layout.decode_row_metadata(np.stack(meta)), g[p + "row_meta"])


def test_dense_expansion_round_trip(layout_golden):
    """decode -> expand -> re-compress is the identity (prepare_dense_control.py:59-69)."""

    g = layout_golden
    comp = g["export_split24_tile_a_comp"].reshape(-1, 16, 16)
    meta = g["export_split24_tile_pair_meta"].reshape(-1, 16)
    dense = layout.expand_tiles(comp, meta)
    assert int(np.count_nonzero(dense, axis=2).max()) <= 16
    comp2, meta2 = layout.compress_tiles_vectorized(dense.astype(np.float64))
    np.testing.assert_array_equal(comp2.view(np.uint16), comp.view(np.uint16))
    # metadata may differ only where a kept value is zero; the expansion is identical
    np.testing.assert_array_equal(layout.expand_tiles(comp2, meta2).view(np.uint16), dense.view(np.uint16))


def test_export_reference_bin_semantics(layout_golden):
    """``reference.bin`` = FP16 coefficients x FP16 dense input in float32 (export_gpu_case.py:221-230)."""

    g = layout_golden
    traj = layout.generate_trajectory("spiral", 24, 192)
    view = layout.transpose_sparse(layout.build_interpolation(traj, 24, 4), 24)
    dense_in = g["export_split24_dense_input"].reshape(view.n_cols, 16).astype(np.float32)
    ref = np.zeros((view.output_size, 16), dtype=np.float32)
    for i, (c, v) in enumerate(zip(view.cols, view.vals)):
        ref[int(view.row_ids[i])] = v.astype(np.float16).astype(np.float32) @ dense_in[c]
    np.testing.assert_allclose(ref.ravel(), g["export_split24_reference"], rtol=1e-6, atol=1e-6)


def test_optimized_planner_invariants():
    """Production planner: lossless, all tiles legal, no residual, shared structure."""

    from oracle import tkbn
    rng = np.random.default_rng(3)
    traj = rng.uniform(-0.5, 0.5, size=(200, 2)).astype(np.float32)
    omega = (traj.T * 2 * np.pi).astype(np.float32)
    cols, re, im = tkbn.kb_interpolation_f32(omega, (16, 16), (32, 32))
    coords = np.mod(traj.astype(np.float64) * 32 + 16, 32)
    real = layout.sparse_rows_from_ell("R", cols, re, 32 * 32, coords)
    imag = layout.sparse_rows_from_ell("I", cols, im, 32 * 32, coords)
    for view_r, view_i in ((real, imag), (layout.transpose_sparse(real, 32), layout.transpose_sparse(imag, 32))):
        structure = layout.SparseRows(view_r.name, view_r.output_size, view_r.n_cols, view_r.row_ids,
                                      view_r.coords, view_r.cols, tuple(np.ones(len(c)) for c in view_r.cols))
        tiles = layout.optimized_split24_pack(structure, "morton", 20260823)
        assert all(layout.validate_24(t.values)[0] for t in tiles)
        assert layout.sparse_checksum(structure) == layout.packed_checksum(tiles)
        packs = layout.pack_components(view_r, view_i)
        for comp in ("real", "imag"):
            arrays = packs[comp]
            dense = layout.expand_tiles(arrays["tile_a_comp.bin"], arrays["tile_pair_meta.bin"])
            assert dense.shape[0] == len(tiles)
        for key in ("group_offsets.bin", "group_row_ids.bin", "tile_col_ids.bin"):
            np.testing.assert_array_equal(packs["real"][key], packs["imag"][key])
