#!/usr/bin/env python3
"""Write the layout golden fixtures from the reference's numpy-only modules.

Run in the development container only (the reference tree is not present on
GPU machines):  python tests/golden/make_layout_golden.py /root/reference

Imported reference modules (numpy + stdlib only):
  src/nufft_layout_oracle.py, src/export_gpu_case.py,
  experiments/packing/prepare_dense_control.py,
  experiments/data_prep/generate_trajectories.py.
The fixtures are inputs and outputs only (arrays in ``layout_golden.npz``);
they pin the exact seeded tile layouts, compression and metadata packing,
which no invariant determines.  Cases follow the reference's own tests
(``tests/test_oracle.py:21-23, 32, 47-48`` and
``tests/test_export_gpu_case.py:19-33``).
"""

from __future__ import annotations

import argparse
import importlib.util
import sys
import tempfile
from argparse import Namespace
from contextlib import redirect_stdout
from io import StringIO
from pathlib import Path

import numpy as np


def load(root: Path, relative: str, name: str):
    spec = importlib.util.spec_from_file_location(name, root / relative)
    module = importlib.util.module_from_spec(spec)
    sys.modules[name] = module
    spec.loader.exec_module(module)
    return module


def tiles_to_arrays(prefix: str, tiles, out: dict) -> None:
    out[prefix + "_row_ids"] = np.stack([t.row_ids for t in tiles]).astype(np.int32)
    out[prefix + "_col_ids"] = np.stack([t.col_ids for t in tiles]).astype(np.int32)
    out[prefix + "_values"] = np.stack([t.values for t in tiles]).astype(np.float64)
    out[prefix + "_legal"] = np.asarray([t.legal_24 for t in tiles], dtype=np.bool_)


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("reference", type=Path)
    parser.add_argument("--output", type=Path, default=Path(__file__).with_name("layout_golden.npz"))
    args = parser.parse_args()
    root = args.reference.resolve()
    sys.path.insert(0, str(root))
    oracle = load(root, "src/nufft_layout_oracle.py", "src.nufft_layout_oracle")
    export = load(root, "src/export_gpu_case.py", "src.export_gpu_case")
    dense = load(root, "experiments/packing/prepare_dense_control.py", "prepare_dense_control")
    trajs = load(root, "experiments/data_prep/generate_trajectories.py", "generate_trajectories")

    out: dict[str, np.ndarray] = {}
    # H18: radial / golden-angle trajectories.
    out["radial_golden_8x16"] = trajs.radial(8, 16, True)
    out["radial_uniform_8x16"] = trajs.radial(8, 16, False)
    out["radial_golden_5x12"] = trajs.radial(5, 12, True)

    # Structural oracle case of tests/test_oracle.py.
    traj = oracle.generate_trajectory("radial", 24, 96)
    out["case_traj"] = traj
    out["case_spiral_traj"] = oracle.generate_trajectory("spiral", 24, 192)
    fwd = oracle.build_interpolation(traj, 24, 4)
    adj = oracle.transpose_sparse(fwd, 24)
    out["fwd_ptr"] = np.cumsum([0] + [len(c) for c in fwd.cols]).astype(np.int64)
    out["fwd_cols"] = np.concatenate(fwd.cols).astype(np.int32)
    out["fwd_vals"] = np.concatenate(fwd.vals).astype(np.float64)
    out["adj_row_ids"] = adj.row_ids.astype(np.int32)
    out["adj_ptr"] = np.cumsum([0] + [len(c) for c in adj.cols]).astype(np.int64)
    out["adj_cols"] = np.concatenate(adj.cols).astype(np.int32)
    out["adj_vals"] = np.concatenate(adj.vals).astype(np.float64)
    out["checksum_fwd"] = np.asarray(oracle.sparse_checksum(fwd))
    out["checksum_adj_t"] = np.asarray(oracle.canonical_operator_checksum(adj, False))
    for policy in oracle.ROW_GROUPING_POLICIES:
        out["order_fwd_" + policy] = np.asarray(oracle.row_order(fwd, policy, 32), dtype=np.int64)
        out["order_adj_" + policy] = np.asarray(oracle.row_order(adj, policy, 32), dtype=np.int64)
    tiles_to_arrays("pack_fwd_overlap_7", oracle.pack_view(fwd, "overlap", 32, 7).tiles, out)
    tiles_to_arrays("pack_adj_overlap_13", oracle.pack_view(adj, "overlap", 32, 13).tiles, out)
    tiles_to_arrays("pack_fwd_morton_11", oracle.pack_view(fwd, "morton", 32, 11).tiles, out)
    tiles_to_arrays("pack_adj_packing256_5", oracle.pack_view(adj, "packing256", 32, 5).tiles, out)

    # Export case of tests/test_export_gpu_case.py (csr_residual) and its split24 twin.
    for mode in ("csr_residual", "split24"):
        with tempfile.TemporaryDirectory() as directory:
            with redirect_stdout(StringIO()):
                export.export_case(Namespace(
                    trajectory="spiral", trajectory_file=None, periodic=False, view="G_T", grid=24,
                    width=4, samples=192, policy="overlap", illegal_mode=mode, seed=23,
                    output_dir=Path(directory)))
            d = Path(directory)
            out[f"export_{mode}_header"] = np.frombuffer((d / "header.bin").read_bytes(), dtype=np.uint8)
            for name, dtype in (("group_offsets", np.uint32), ("group_row_ids", np.int32),
                                ("tile_col_ids", np.int32), ("tile_a_comp", np.float16),
                                ("tile_pair_meta", np.uint32), ("residual_row_ptr", np.uint32),
                                ("residual_cols", np.int32), ("residual_vals", np.float16),
                                ("dense_input", np.float16), ("reference", np.float32)):
                out[f"export_{mode}_{name}"] = np.fromfile(d / f"{name}.bin", dtype=dtype)
            row_meta = dense.decode_row_metadata(
                out[f"export_{mode}_tile_pair_meta"].reshape(-1, 16))
            out[f"export_{mode}_row_meta"] = row_meta

    # compress_tile / pair_metadata on seeded legal tiles with FP16-underflowing entries.
    rng = np.random.default_rng(101)
    tiles = np.zeros((6, 16, 32))
    for t in range(6):
        for r in range(16):
            for q in range(8):
                k = rng.integers(0, 3)
                pos = rng.choice(4, size=k, replace=False)
                tiles[t, r, 4 * q + pos] = rng.standard_normal(k) * (1e-9 if t == 5 else 1.0)
    out["compress_in"] = tiles
    comps, metas = zip(*(export.compress_tile(tiles[t]) for t in range(6)))
    out["compress_comp"] = np.stack(comps)
    out["compress_meta"] = np.stack(metas)
    args.output.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(args.output, **out)
    print(f"wrote {args.output} ({args.output.stat().st_size} bytes, {len(out)} arrays)")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
